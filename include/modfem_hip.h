/*
 * modfem_hip.h -- C ABI of the MI355X-native ModFEM hot path
 * (element integration -> CRS assembly -> GMRES), libmodfem_hip.so.
 *
 * Plain C types only: int/long long/double pointers and sizes.  Every entry
 * point returns >= 0 on success and < 0 on error (-1 bad argument, -2 HIP
 * runtime failure, -3 unknown handle, -4 numerical failure); mf_last_error()
 * gives the message.  The reference instead calls exit(-1) on most errors
 * (e.g. sid_mkb/sis_mkb_intf.c:289-292).
 *
 * Objects are small integer handles, as in the reference's ID-indexed tables
 * (sid_mkb/sis_mkb_intf.c:95-97, lsd_mkb/lad_crs/las_crs_intf.c:65-66).
 * Paths are relative to the reference tree modfem2017/src/.
 *
 * Not re-entrant: like the reference's lar_/lsr_ layers, call from one host
 * thread.  All device work runs on one non-blocking HIP stream per process.
 */
#ifndef MODFEM_HIP_H
#define MODFEM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MF_PRISM 5 /* MMC_PRISM, include/mmh_intf.h */
#define MF_TETRA 7 /* MMC_TETRA */

/* coefficient storage models for mf_problem_set_coefficients */
#define MF_COEF_UNIFORM 0 /* all 21 values from coef_const                  */
#define MF_COEF_KAPPA 1   /* coef_el[nel]: diffusion a = kappa*I             */
#define MF_COEF_KAPPA_B 2 /* coef_el[4*nel] SoA: kappa, bx, by, bz            */
#define MF_COEF_FULL 3    /* coef_el[21*nel] SoA: a[9] b[3] t[3] c q[3] s m  */
#define MF_COEF_KAPPA_S 4 /* coef_el[2*nel] SoA: kappa, source s              */
/* source / Dirichlet data models */
#define MF_SRC_CONST 0
#define MF_SRC_DIFF_IN_CUBE 1 /* pdd_conv_diff/main/pds_conv_diff_util.c, pdr_exact_sol */
#define MF_SRC_DIFF_IN_BOX 2

/* ---------------------------------------------------------------- library */
/* select the HIP device and create the library stream (no reference
   counterpart: the reference's device set-up is tmr_init_... in
   tmd_opencl/tms_ocl_intf.c:105) */
int mf_init(int device);
const char* mf_last_error(void);
int mf_version(void);

/* ------------------------------------------------------------------- mesh
 * Input the approximation module reads through mmr_get_next_elem_all,
 * mmr_el_node_coor, mmr_fa_neig and mmr_fa_bc (include/mmh_intf.h;
 * mmd_prism/mms_prism_intf.c:385, mms_prism_datstr.c:365,993).  Indices are
 * 0-based, element nodes padded to 6 per element, boundary faces (bc>0)
 * listed in face-ID order with their owning element and local side. */
int mf_mesh_create(int nr_nodes, const double* node_coor, int nr_elems,
                   const int* el_type, const int* el_nodes, int nr_bfaces,
                   const int* bf_elem, const int* bf_side, const int* bf_bc);
/* jk ("gradmesh") half-structured prism format: mmr_import_mesh_grad
   (mmd_prism/mms_prism_input_gradmesh.c:29) */
int mf_mesh_read_jk(const char* filename);
/* NASTRAN tetrahedral mesh, free field format (utr_io_read_nas,
   utd_util/uts_io_intf.cpp:812-1020): GRID nodes, CTETRA elements in file
   order (vertices sorted ascending, then oriented positively), CTRIAX /
   CTRIA3 triangles with bc > 0 as boundary faces in file order; prisms and
   fixed-field files are rejected */
int mf_mesh_read_nas(const char* filename);
/* ModFEM prism mesh dump (".dat" / ".dmp"): mmr_read_mesh (mmd_prism/
   mms_prism_io_dump.c:26-270) with the element-local node order of
   mmr_el_node_coor (mmd_prism/mms_prism_datstr.c:365-510); meshes of active
   prisms only (no refinement history) */
int mf_mesh_read_dat(const char* filename);
int mf_mesh_write_jk(int mesh, const char* filename);
/* seeded synthetic boxes (no reference counterpart: workload generators) */
int mf_mesh_box_prism(int nx, int ny, int nz, double lx, double ly, double lz,
                      unsigned long long seed, double jitter, int bc);
int mf_mesh_box_tetra(int nx, int ny, int nz, double lx, double ly, double lz,
                      unsigned long long seed, double jitter, int bc);
/* z-layers of hex cells split into 2 prisms or 6 tetrahedra each (a layer
   is tetrahedral with probability tet_frac), conforming on the triangles
   where the layers meet: the mixed prism/tet meshes of DG problems */
int mf_mesh_box_hybrid(int nx, int ny, int nz, double lx, double ly, double lz,
                       unsigned long long seed, double jitter, int bc, double tet_frac);
int mf_mesh_sizes(int mesh, int* nr_nodes, int* nr_elems, int* nr_bfaces);
int mf_mesh_get(int mesh, double* node_coor, int* el_type, int* el_nodes,
                int* bf_elem, int* bf_side, int* bf_bc);
int mf_mesh_free(int mesh);
/* host-only integer set-up of a mesh (DOF list, CRS pattern, tables) without
   device work; same maps as mf_problem_create */
int mf_mesh_setup(int mesh, int* n, long long* nnz);
int mf_mesh_setup_get(int mesh, int* dof_node, int* row_ptr, int* col_ind);
int mf_mesh_setup_table(int mesh, int int_ent, int* table);
/* device-internal numbering chosen by the set-up (no reference counterpart:
   the reference keeps first-appearance numbering, and so does every mf_
   boundary): device row i holds DOF row_perm[i], device element slot e holds
   element el_perm[e] */
int mf_mesh_setup_order(int mesh, int* row_perm, int* el_perm);
/* field dumps in R's standard format (host only, no GPU needed).
   mf_field_write: apr_write_field (apd_std_lin/aps_std_lin_intf.c:532-650)
   for one equation and one solution vector -- "1 1", the number of DOF
   entities, then "id value " per node with a DOF in ascending node id,
   %.12lg, |value| < accuracy written as 0; x in DOF order.
   mf_field_read: the APC_READ branch of apr_create_field (:300-420); the
   first solution vector into x (DOF order); returns the entity count. */
int mf_field_write(int mesh, const double* x, double accuracy, const char* filename);
int mf_field_read(int mesh, const char* filename, double* x);

/* --------------------------------------------------------------- problem
 * mf_problem_create replaces sir_create's integer set-up for a std-linear
 * field with block size 1 (sid_mkb/sis_mkb_intf.c:396-1692):
 * utr_get_list_int_ent / utr_get_list_dof_ent (utd_util/uts_ls_intf.c:102,182),
 * DOF structures and sorted neighbour lists, lar_allocate_SM_and_LV_crs
 * (lsd_mkb/lad_crs/las_crs_intf.c:72) and the assembly tables
 * (las_crs_intf.c:354).  The matrix lives in device memory. */
int mf_problem_create(int mesh);
/* the same set-up computed on the device (the same replaced routines; the
   reference has no device set-up): DOF numbering, CRS pattern, assembly
   tables, incidence lists and the device ordering come from radix sorts on
   the GPU and are bitwise those of mf_problem_create; only the maps that
   host-side calls read are copied back.  Meshes of one element type with
   at most 2^31 element-matrix entries. */
int mf_problem_create_device(int mesh);
/* solver-only problem from a given CRS matrix (strictly ascending columns
   with the diagonal present), e.g. for a caller that assembles itself */
int mf_problem_create_crs(int n, const int* row_ptr, const int* col_ind,
                          const double* val, const double* rhs);
int mf_problem_sizes(int pid, int* n, long long* nnz, int* nr_elems,
                     int* nr_bfaces, int* max_row);
int mf_problem_get_dofs(int pid, int* dof_node);
int mf_problem_get_pattern(int pid, int* row_ptr, int* col_ind);
/* the reference-format table of one integration entity (elements first,
   then boundary faces): (k+1)*k ints, lar_fill_assembly_table_int_ent_crs
   (las_crs_intf.c:354-432) */
int mf_problem_get_table(int pid, int int_ent, int* table);

/* coefficients: coeff_ind[23] is the reference's coefficient indicator
   vector (pdr_select_el_coeff_vect, include/pdh_intf.h:486); bf_g[nbf]
   optional Dirichlet values for MF_SRC_CONST */
int mf_problem_set_coefficients(int pid, const int* coeff_ind, int coef_model,
                                const double* coef_const, const double* coef_el,
                                int src_model, const double* bf_g);

/* batched integrate + assemble on the device: replaces the create/assemble
   loop of sir_solve (sis_mkb_intf.c:1843-2073) with apr_num_int_el
   (apd_std_lin/aps_std_lin_intf.c:1537), pdr_conv_diff_comp_fa_stiff_mat_app
   (pdd_conv_diff/weak_formulation/approx_std/pds_conv_diff_weakform_std.c:253)
   and lar_assemble_SM_and_LV_with_table_crs (las_crs_intf.c:440); the
   OpenCL counterpart is tmr_perform_create_assemble_accel
   (utd_util/uts_accel_intf.c:70-103).  precision: 64 or 32 (FP32
   shape-function products with FP64 geometry and FP64 global sums). */
int mf_integrate_assemble(int pid, int precision);
/* mf_integrate_assemble without waiting for the device: the work is
   enqueued on the library stream, so every later library call on the
   problem (a solve, a batch cycle, a read-back) sees the new system */
int mf_integrate_assemble_async(int pid, int precision);
/* per boundary face (face-ID order of the mesh): type 1 Dirichlet penalty
   K += 1e7 phi_j phi_i dA, F += 1e7 g phi_i dA; type 2 Neumann F += g phi_i dA
   with constant flux g, or for the exact-solution source models the exact
   normal flux grad u . n at every face point (pdr_bc_neum_coeff)
   (pds_conv_diff_weakform_std.c:480-522; the type is what pdr_get_bc_type
   returns for the face's bc flag) */
int mf_problem_set_face_bc(int pid, const int* bc_type, const double* g);
/* element / face matrices without assembly, column-major stiff[u*k+w] then
   rhs[k], 42 doubles per entity (apr_num_int_el output layout,
   aps_std_lin_intf.c:1558-1577) */
int mf_element_matrices(int pid, int first, int count, int precision, double* out);
/* assembly strategy of mf_integrate_assemble: 0 (default) one lane per
   matrix row walking its elements in ascending order, deterministic and in
   the reference's summation order; 1 one lane group per element with FP64
   atomics (order-dependent rounding); 2 two passes: element rows written
   once to a device buffer, then gathered per matrix row in ascending element
   order (deterministic).  Modes 0 and 2 hold a row's sums in LDS for rows of
   up to 32 entries; a mesh with longer rows (the reference allows up to 200
   neighbours, sih_intf.h:57-63) is assembled by mode 1 */
int mf_set_assembly_mode(int pid, int mode);
int mf_face_matrices(int pid, int first, int count, double* out);
int mf_get_system(int pid, double* val, double* rhs);
int mf_set_system(int pid, const double* val, const double* rhs);

/* ---------------------------------------------------------------- solver */
/* y = A x (host arrays; lar_compute_residual_crs with Use_rhs=0 gives -y,
   las_crs_intf.c:788) */
int mf_spmv(int pid, const double* x, double* y);
/* preconditioner fill: 1 Jacobi 1/a_ii, lar_fill_preconditioner_crs
   (las_crs_intf.c:707); 2 ILU(0), the block ILU(0) of
   lar_fill_preconditioner_block (lad_block/las_block_intf.c:1409-2143) for
   1x1 blocks */
int mf_fill_precon(int pid, int precon);
/* ILU(0) elimination order: 0 the reference's DOF order (the reference's
   factors; as many levels as its dependency chains), 1 multicolour (greedy
   colouring of the matrix graph; a different ILU(0), few levels) */
int mf_set_ilu_order(int pid, int order);
/* precon 3: block Gauss-Seidel smoothing (BLOCK_GS, lar_perform_BJ_or_GS_
   iterations_crs, las_crs_intf.c:857-1013, in compreres' smoothing branch,
   lss_mkb_core.c:2001-2054) with nsweeps sweeps per application, rows
   updated in a multicolour order (the reference's order is the DOF order
   within each thread's chunk) */
int mf_set_gs(int pid, int nsweeps);
/* the Gauss-Seidel order: position of every reference row, and the colours */
int mf_gs_get(int pid, int* elim_pos, int* ncolors);
/* the ILU(0) factors in reference CRS positions (strict L, strict U, 1/u_ii
   on the diagonal), the elimination position of every reference row, and
   levels[3] = {forward levels, backward levels, colours} */
int mf_ilu_get(int pid, double* lu, int* elim_pos, int* levels);
/* restarted left-preconditioned GMRES(krylbas), lsr_mkb_core_gmres
   (lsd_mkb/lsd_mkb_core/lss_mkb_core.c:2317); precon 0 none, 1 Jacobi,
   2 ILU(0) (compreres' BLOCK_ILU branch, lss_mkb_core.c:1947-1997),
   3 multicolour block Gauss-Seidel (mf_set_gs).
   Returns 1 on convergence, 0 otherwise, <0 on error.  x: host array of n
   (input guess unless ini_zero, output solution), or NULL to start from zero
   and keep the solution in device memory only (read it with
   mf_get_solution). */
int mf_gmres(int pid, int precon, int krylbas, int max_iter, double tol,
             int ini_zero, double* x, int* iters, double* res_rel, double* resid);
/* Solver control file (mkb.dat), lsr_mkb_core_init single-level branch
   (lsd_mkb/lsd_mkb_core/lss_mkb_core.c:148-705).  params[13] receives
   {solver, nr_level, krylbas, precon, nr_prec, block_type, storage_type,
    ilu_k, max_iter, conv_type, monitor, nr_pre_smooth, nr_post_smooth};
   arguments > 0 override the file as in the reference; an empty file name
   gives the reference defaults (GMRES(50), block GS, 100 iterations,
   REL_RES_INI 1e-6). */
int mf_mkb_read(const char* filename, int max_iter, int error_type, double error_tol,
                int monitoring, int* params, double* conv_meas);
/* GMRES driven by mkb parameters, lsr_mkb_core_solve (lss_mkb_core.c:2200-2245):
   conv_type 0 REL_RES_INI, 1 ABS_RES, 2 REL_RES_RHS (reset to 0 by the
   reference, leaving the 1e-12 floors).  Available here: GMRES with no
   preconditioner or one Jacobi sweep; other settings return -1. */
int mf_solve_mkb(int pid, const int* params, double conv_meas, int ini_zero, double* x,
                 int* iters, double* res_rel, double* resid);
/* Mixed-precision solve (no reference counterpart; the reference is FP64
   throughout): FP64 iterative refinement r = b - A x, x += d, around
   GMRES(krylbas) whose operator reads FP32 matrix values while the Krylov
   basis, Gram-Schmidt, Givens rotations, corrections and residuals stay FP64.
   Stops when ||M^-1 (b - A x)|| <= tol ||M^-1 b|| in FP64 (M = I, D or the
   ILU(0) factors for precon 0, 1, 2): the measure of mf_gmres from a zero
   guess.  iters = total inner
   iterations, refinements = outer steps; x as in mf_gmres (NULL keeps the
   solution on the device). */
int mf_gmres_mixed(int pid, int precon, int krylbas, int max_iter, double tol, double* x,
                   int* iters, double* res_rel, double* resid, int* refinements);
/* mf_gmres with precon 0 (none) or 1 (Jacobi) that allocates no device
   memory: no counterpart in the reference; it is for callers that cannot
   allocate (code under graph capture, tight serving loops, harnesses that
   forbid hipMalloc).  Every work array of the solve (start and solution
   vectors, Jacobi inverses, Krylov basis, Hessenberg data, partial sums,
   status) lives in a fixed 128 MiB device array per GPU, set up by mf_init;
   a solve that needs more (about (krylbas + 4) * 8 * n bytes plus 2 MiB)
   returns an error.  The problem's matrix and load are its inputs, x its
   start (unless ini_zero) and output, in the DOF order of
   mf_problem_get_dofs.  The array keeps the last call's contents; calls that
   overlap on one GPU race.  Iterates equal mf_gmres's. */
int mf_gmres_noalloc(int pid, int precon, int krylbas, int max_iter, double tol, int ini_zero,
                     double* x, int* iters, double* res_rel, double* resid);
/* host copy of the last mf_gmres solution (DOF order of mf_problem_get_dofs) */
int mf_get_solution(int pid, double* x);
/* true relative residual ||b - A x|| / ||b|| in FP64 on the device */
int mf_true_residual(int pid, const double* x, double* rel);

/* ---------------------------------------------------------- measurement
 * kernel times measured with HIP events on the library stream */
int mf_bench_integrate(int pid, int precision, int reps, double* ms_elems,
                       double* ms_faces);
int mf_bench_spmv(int pid, int reps, double* ms);
/* ms per SELL-64 operator application y = A x with prec 0 (FP64 values and
   vectors), 1 (FP32 values, FP64 vectors) or 2 (FP32 values and vectors),
   16-bit columns in the slices spanning fewer than 65536 rows (as the
   solvers use them); prec + 10: 32-bit columns everywhere; ch in
   {8, 12, 16, 24} matrix entries per row in flight per lane */
int mf_bench_spmv_ex(int pid, int prec, int ch, int reps, double* ms);
int mf_last_solve_ms(int pid, double* ms);
/* plain device-to-device stream copy of nbytes (rounded down to 16 bytes):
   the achievable-bandwidth yardstick and the known-size kernel that scales
   the HBM byte counters (no reference counterpart) */
int mf_bench_copy(long long nbytes, int reps, double* ms);
/* per-launch HIP-event timing of the two Arnoldi kernels inside mf_gmres:
   out[6] = {launches, ms, algorithmic bytes} for the operator+dots kernel,
   then for the update+norm kernel (accumulated since the last call); enable
   switches the timing on or off for later solves */
int mf_gmres_kernel_stats(int enable, double* out);

int mf_problem_free(int pid);

/* ------------------------------------------------------------- batches
 * Many independent problems solved together (no reference counterpart: the
 * reference solves one system at a time).  Members keep their own device
 * matrices, vectors and Krylov bases; every kernel launch of a restart cycle
 * serves all members, and the restart logic of lsr_mkb_core_gmres runs on
 * the device, so a cycle can be replayed as one hipGraph (use_graph).
 * Members may join or leave between cycles (continuous batching); solutions
 * stay in device memory (mf_get_solution).  A member must not be freed, and
 * should not be re-assembled, while it is in a batch. */
int mf_batch_create(int krylbas);
/* start a member: zero initial guess, GMRES(krylbas) with precon 0/1/2 to
   the REL_RES_INI tolerance tol within max_iter iterations.  precon 2 is
   ILU(0) in the multicolour elimination order (mf_set_ilu_order(pid, 1)
   first); its substitutions run one launch per colour level for all ILU
   members together. */
int mf_batch_add(int bid, int pid, int precon, int max_iter, double tol);
/* mf_batch_add with the inner precision: 64 as mf_batch_add; 32 makes a
   mixed-precision member (as mf_gmres_mixed): FP64 refinement steps
   between cycles around inner GMRES solves whose matrix values, Krylov
   basis and Jacobi inverses / ILU(0) factors are FP32 (sums, Hessenberg
   data and the solution FP64).  The member ends when ||D^-1 (b - A x)||,
   evaluated in FP64 (D the Jacobi diagonal; the plain norm for precon 0),
   is at most tol times its value at x = 0. */
int mf_batch_add_ex(int bid, int pid, int precon, int max_iter, double tol, int inner_precision);
int mf_batch_remove(int bid, int pid);
/* a DG problem (one degree for every element, assembled) as a batch
   member: GMRES on the block-Jacobi-preconditioned system Dinv A x =
   Dinv b (Dinv: inverses of the diagonal blocks), which is the reference's
   left-preconditioned block-Jacobi GMRES of DG systems, held as a SELL-64
   matrix formed on the device; inner_precision 64 or 32 as mf_batch_add_ex.
   The solution is read with mf_dg_get_solution. */
int mf_batch_add_dg(int bid, int did, int max_iter, double tol, int inner_precision);
int mf_batch_remove_dg(int bid, int did);
int mf_batch_status_dg(int bid, int did, int* done, int* converged, int* iters,
                       double* res_rel, int* refinements);
/* integrate + assemble every member (mf_integrate_assemble of each) */
int mf_batch_integrate_assemble(int bid, int precision);
/* run up to max_cycles restart cycles; *active = members still iterating */
int mf_batch_run(int bid, int max_cycles, int use_graph, int* active);
/* give the batch a stream of its own (default priority): its cycles then
   run there, after the work enqueued on the library stream before each
   run, so that several batches can run at once */
int mf_batch_set_stream(int bid);
/* enqueue one restart cycle and return; mf_batch_wait completes it
   (status, membership and tile calls are refused while a cycle is in
   flight, and the members must not be touched until then) */
int mf_batch_run_async(int bid, int use_graph);
int mf_batch_wait(int bid, int* active);
int mf_batch_status(int bid, int pid, int* done, int* converged, int* iters,
                    double* res_rel);
/* mf_batch_status with the refinement steps of a mixed-precision member
   (0 otherwise); for such a member iters counts its inner iterations and
   res_rel is the FP64 measure of the last refinement step */
int mf_batch_status_ex(int bid, int pid, int* done, int* converged, int* iters,
                       double* res_rel, int* refinements);
/* Batch tiles (no reference counterpart).  A tile is the number of 64-row
 * slices one block of a member streams in every batched kernel.  spb in
 * [1,64] fixes it; spb = 0 (the default) lets the library tune it: when the
 * membership has changed, the next mf_batch_run times one restart cycle per
 * candidate (mf_batch_tile_candidates) and keeps the fastest.  Trial cycles
 * are undone (x and the GMRES state are restored), so tuning changes no
 * iterate.  The choice is cached for the process per batch shape (total
 * rows in half-octaves, member count in octaves, graph or eager launches);
 * a later batch of that shape, or within two steps of it on either count,
 * takes the cached choice without trials. */
int mf_batch_set_tiles(int bid, int spb);
/* tile in use, whether it was tuned, and the trial cycle times in ms of the
   candidates (trial_ms: room for the returned count); returns the count */
int mf_batch_tiles(int bid, int* spb, int* tuned, double* trial_ms);
/* candidate tiles in trial order (spb: room for the returned count) */
int mf_batch_tile_candidates(int* spb);
/* forget every cached tile choice */
int mf_batch_tile_cache_clear(void);
int mf_batch_free(int bid);

/* ------------------------------------------------------------ DG prisms
 * Discontinuous Galerkin approximation on prism meshes (apd_dg_prism):
 * tensor-product monomial base of degree pdeg = 100*pz + px, 1 <= px, pz <= 3
 * (apr_shape_fun_3D, apd_dg_prism/aps_dg_util.c:408; 6 / 18 / 40 shape
 * functions for p = 1 / 2 / 3, apr_get_el_pdeg_numshap,
 * aps_dg_prism_intf.c:725), R's pz = 0 degrees pdeg 1 and 2 (3 / 6 functions
 * constant in z, one Gauss point in z, penalty degree min(px, pz) = 0,
 * aps_dg_util.c:495-600, pds_conv_diff_weakform_dg.c:516-527), or the
 * complete base APC_BASE_COMPLETE_DG of degree p, given as MF_DG_COMPLETE + p
 * (every monomial of total degree <= p: 4 / 10 / 20 functions, the rules of
 * the tensor degree (p, p), penalty degree p, aps_dg_util.c:566-870,
 * 1446-1451).  Element DOF blocks follow the element order of the mesh. */
#define MF_DG_COMPLETE 1000
/* one degree for every element */
int mf_dg_create(int mesh, int pdeg);
/* a degree per element (apr_get_ent_pdeg, as an adapted field has them)
   from one family: 101, 202 or 303 each; 1 or 2 each; or MF_DG_COMPLETE + p
   each (any one degree for all).  A face between
   different degrees integrates with the rule of the neighbour with more
   points and penalises with the smaller degree
   (pds_conv_diff_weakform_dg.c:508-531); the coupling blocks are
   rectangular.  Element blocks sit at the offsets of mf_dg_layout */
int mf_dg_create_mixed(int mesh, const int* pdeg);
/* nshape: the block size of a one-degree problem, 0 for a mixed one */
int mf_dg_sizes(int did, int* nel, int* nshape, int* ngauss);
/* number of unknowns, offsets of every element's unknowns (nel + 1) and of
   its six block-row slots in the block storage (6 nel, then the total:
   own block ns_e x ns_e, then the coupling across each side, ns_e x ns_nb
   when interior, ns_e x ns_e otherwise; column-major; a row's slots are
   contiguous, rows of a mixed problem stored degree by degree); any pointer
   may be NULL (no reference counterpart) */
int mf_dg_layout(int did, long long* ndof, long long* xoff, long long* soff);
/* coefficients as pdr_select_el_coeff_vect / pdr_el_coeff provide them to
   apr_num_int_el (same ind[23], coefficient slots and storage models as
   mf_problem_set_coefficients); the DG form has no t (tx, ty, tz) terms
   (aps_dg_prism_intf.c:1653) and takes the 2D diffusion / convection / q
   branches when azz / bz / qz are absent */
int mf_dg_set_coefficients(int did, const int* ind, int coef_model, const double* coef_const,
                           const double* coef_el, int src_model);
/* element matrices of elements first..first+count-1 into host arrays:
   stiff[count][ns*ns] column-major (stiff[j*ns + i], j trial, i test) and
   rhs[count][ns], as apr_num_int_el (aps_dg_prism_intf.c:1239-1728) writes
   them for nreq = 1 */
int mf_dg_element_matrices(int did, int first, int count, double* stiff, double* rhs);
/* the same for every element into device-resident buffers, then copied out
   on request (mf_dg_get_element_matrices) */
int mf_dg_integrate(int did);
int mf_dg_get_element_matrices(int did, int first, int count, double* stiff, double* rhs);
/* NIPG face terms (pdr_conv_diff_comp_fa_stiff_mat_app, pdd_conv_diff/
   weak_formulation/approx_dg/pds_conv_diff_weakform_dg.c:232-1400): penalty
   min(px, pz) * penalty_const / sqrt(face area) (PENALTY_CONSTANT of the
   problem file, default 1), steady form (implicitness 1).  Boundary faces
   take the mesh's bc flags as pdr_get_bc_type does (1..9 Dirichlet, 10..19
   Neumann; unflagged faces Dirichlet); their data is the exact solution
   (its normal derivative) when a source model selects one, else the value
   given per (element, side) by mf_dg_set_boundary_values (default 0) */
int mf_dg_set_penalty(int did, double penalty_const);
/* kernel choice for the block rows (no reference counterpart; the reference
   integrates DG elements on the CPU, aps_dg_prism_intf.c:1239): own and
   coupling blocks of at most max_ns shape functions (default 6, p = 1) are
   assembled by vector-ALU kernels holding several elements per wave, larger
   ones by FP64 MFMA kernels, one element (side) per wave; 0 puts every
   block on MFMA.  Both forms sum the same terms to rounding. */
int mf_dg_set_valu_blocks(int did, int max_ns);
int mf_dg_set_boundary_values(int did, const double* val);
/* element-centric face table, 5 entries per element: neighbour element or
   -1 (Dirichlet) / -2 (Neumann), orientation code, sqrt(face area) */
int mf_dg_get_faces(int did, int* nb, int* code, double* hf);
/* the block system of lad_block storage (lsd_mkb/lad_block/lah_block.h:
   22-83, las_block_intf.c:362): per element a row of six ns x ns blocks,
   column-major (slot 0 the element's own block with its volume and face
   terms, slot 1 + s the coupling to the neighbour across side s) and its
   load vector, built on the device */
int mf_dg_assemble(int did);
int mf_dg_get_block_row(int did, int e, double* blocks, double* rhs);
/* block Jacobi: inverses of the own blocks (lar_fill_preconditioner_block,
   las_block_intf.c:1057, which factors them with LAPACK) */
int mf_dg_fill_precon(int did);
/* the inverse of element e's own block, column-major */
int mf_dg_get_precon_block(int did, int e, double* inv);
/* block Gauss-Seidel smoothing for precon 2 (BLOCK_GS, lar_perform_BJ_or_GS_
   iterations for block storage, las_block_intf.c:2269-2460): nsweeps sweeps
   per application, elements updated colour by colour (a greedy colouring of
   the element graph; the reference sweeps the blocks in their order) */
int mf_dg_set_gs(int did, int nsweeps, int* ncolors);
/* p-multigrid V-cycle as precon 4 of mf_dg_gmres (the V-cycle of
   lsr_mkb_core_vcycle, lsd_mkb_core/lss_mkb_core.c:1288-1470, over levels of
   decreasing isotropic degree on the same mesh where the reference walks
   mesh levels): for a uniform degree 202 or 303, levels of degree p-1 .. 1
   with the Galerkin operators P^T A P (sub-blocks of the finer blocks: the
   tensor bases are nested, P injects coefficients), nu1 / nu2 multicolour
   block Gauss-Seidel sweeps before / after the coarse correction, `coarse`
   block-ILU(0) corrections on degree 1.  Returns the number of levels */
int mf_dg_set_multigrid(int did, int nu1, int nu2, int coarse);
/* handle of level l (0: degree 1) below the problem's own, with its Galerkin
   blocks current (read them with mf_dg_get_block_row) */
int mf_dg_multigrid_level(int did, int l);
/* block ILU(0) in element order (lar_fill_preconditioner for block storage,
   MULTI_ILU / BLOCK_ILU, las_block_intf.c:1409-2143): for every element i
   and its lower face neighbours k in ascending order a_ik <- a_ik inv(D_k),
   D_i <- D_i - a_ik a_ki, a_ij <- a_ij - a_ik a_kj for neighbours j > k of
   both; then D_i is inverted (LU with partial pivoting, inverse from the
   factors).  Rows run level by level of the dependency schedule;
   nlevels[0..1] (may be null) receive the forward and backward level
   counts.  Needs a copy of the block rows beside the matrix. */
int mf_dg_fill_ilu(int did, int* nlevels);
/* the factored block row of element e in the layout of mf_dg_get_block_row:
   slot 0 the inverse of U's diagonal block, the other slots the L (lower
   neighbour) or U (upper neighbour) blocks */
int mf_dg_get_ilu_row(int did, int e, double* blocks);
/* v = (LU)^-1 b (lar_perform_rhsub_block, las_block_intf.c:3085-3476),
   host vectors of mf_dg_layout's ndof entries */
int mf_dg_ilu_apply(int did, const double* b, double* v);
/* GMRES (lsr_mkb_core_gmres, lss_mkb_core.c:2317) on the block system with
   precon 0 (none), 1 (block Jacobi, left), 2 (block Gauss-Seidel
   smoothing, the smoothing branch of compreres, lss_mkb_core.c:2001-2054)
   3 (block ILU(0), the BLOCK_ILU branch: (LU)^-1 (b - A x), filled on
   first use) or 4 (one p-multigrid V-cycle as smoothing, after
   mf_dg_set_multigrid); zero initial guess, stopping on the preconditioned residual relative to its
   initial value; x may be null (the solution stays on the device,
   mf_dg_get_solution) */
int mf_dg_gmres(int did, int precon, int krylbas, int max_iter, double tol, double* x,
                int* iters, double* res_rel, double* resid);
int mf_dg_get_solution(int did, double* x);
/* HIP-event times on the library stream, averaged over reps: out[0] ms per
   assembly of the block system (own blocks + coupling blocks), out[1] ms per
   block-Jacobi-preconditioned operator application (no reference
   counterpart; measurement only) */
int mf_dg_bench(int did, int reps, double* out);
int mf_dg_free(int did);

/* storage-seam inspection (mf_lar_get_sizes, mf_lar_get_system,
   mf_lar_get_block_precon) is declared in modfem_lar.h */

#ifdef __cplusplus
}
#endif
#endif
