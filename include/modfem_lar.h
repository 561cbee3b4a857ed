/*
 * modfem_lar.h -- the storage/operations seam of the reference's MKB solver
 * with its own names and signatures, exported by libmodfem_hip.so for block
 * size 1 (CRS) and block size 4 (BCRS, column-major 4x4 blocks).  Each
 * declaration replaces the reference function of the same name in
 * lsd_mkb/lah_intf.h (line numbers below, relative to modfem2017/src/); the
 * dispatcher lsd_mkb/las_intf.c and the modules lsd_mkb/lad_crs/las_crs_intf.c
 * and lsd_mkb/lad_bcrs/las_bcrs_intf.c are what a ModFEM build drops when
 * linking this library instead.
 *
 * Matrices and vectors live in device memory.  Vector arguments (X, B, V)
 * may be host or device pointers.  Matrix ids are indices into a LIFO table
 * (lar_free_SM_and_LV must free the last allocated matrix first), as in
 * las_crs_intf.c:765-769.
 */
#ifndef MODFEM_LAR_H
#define MODFEM_LAR_H

#ifdef __cplusplus
extern "C" {
#endif

/* lah_intf.h:90 -- block pattern from sorted neighbour-block lists; returns
   the matrix id (<0 on error).  LAC_STORAGE_CRS (2) with Block_size 1,
   LAC_STORAGE_BCRS (3) with Block_size 4 and Nrdof_glob = 4 Nrblocks, or
   LAC_STORAGE_UNDEFINED (-1), which picks by Block_size (las_intf.c:90-97);
   LAC_STORAGE_BLOCK (1) with a constant block size of 1 or 4 is held as CRS
   or BCRS (same operations, sums in column order rather than lad_block's
   diagonal-first order); variable block sizes and other storage types are
   refused */
int lar_allocate_SM_and_LV(int Storage_type, int solver_type, int Nrdof_glob,
                           int Max_SM_size, int Nrblocks, int Block_size,
                           int* Nrdofbl, int* Posglob, int* Nroffbl,
                           int** L_offbl);
/* lah_intf.h:108 */
int lar_initialize_SM_and_LV(int Matrix_id, int Scope);
/* lah_intf.h:118 */
double lar_get_storage(int Matrix_id);
/* lah_intf.h:127 */
int lar_fill_assembly_table_int_ent(int Matrix_id, int Nr_dof_bl, int* L_bl_id,
                                    int* L_bl_nrdof, int* Assembly_table_int_ent);
/* lah_intf.h:143 */
int lar_assemble_SM_and_LV_with_table(int Matrix_id, int Scope, int Nr_dof_bl,
                                      int* Assembly_table_int_ent,
                                      double* Stiff_mat, double* Rhs_vect,
                                      char* Rewr_dofs);
/* lah_intf.h:164 */
int lar_assemble_SM_and_LV(int Matrix_id, int Scope, int Nr_dof_bl,
                           int* L_bl_id, int* L_bl_nrdof, double* Stiff_mat,
                           double* Rhs_vect, char* Rewr_dofs);
/* lah_intf.h:184 */
int lar_free_SM_and_LV(int Matrix_id);
/* lah_intf.h:192 -- CRS: NO_PRECON (0), BLOCK_JACOBI (1) and BLOCK_GS (2);
   BCRS also block ILU(k): MULTI_ILU (4) / BLOCK_ILU (5) with
   Second_arg = k on a structurally symmetric pattern; k > 0 grows the
   pattern by k rings of neighbours (las_bcrs_intf.c:751-1016), with the
   reference's list limit of (int)(0.5 (k+2)^3 max_nrneig) blocks per row
   (an error beyond it, where the reference stops).
   Returns the number of diagonal blocks */
int lar_allocate_preconditioner(int Matrix_id, int Precon, int Second_arg);
/* lah_intf.h:203 -- CRS: 1/a_ii; BCRS: LU with partial pivoting of every
   diagonal block (LAPACK dgetrf's algorithm), or the block ILU(k) factors
   (las_bcrs_intf.c:1175-1246, 2181-2640: L blocks a_ki inv(a_ii), U blocks,
   inverted diagonal blocks); <0 for a singular block */
int lar_fill_preconditioner(int Matrix_id);
/* lah_intf.h:210 */
int lar_free_preconditioner(int Matrix_id);
/* lah_intf.h:219 -- V = (Use_rhs ? B (or stored rhs) : 0) - A X */
void lar_compute_residual(int Matrix_id, int Use_rhs, int Ini_zero, int Ndof,
                          double* X, double* B, double* V);
/* lah_intf.h:252 -- CRS BLOCK_JACOBI: V = D^-1 (B - (A-D) V), Nr_prec
   sweeps; CRS BLOCK_GS: Nr_prec forward Gauss-Seidel sweeps in the row
   chunks of n/T + 1 rows that the reference's OpenMP loop gives T threads
   (las_crs_intf.c:905-998), bitwise its result on T threads (T from
   mf_lar_set_gs_threads, default 1).  BCRS (las_bcrs_intf.c:2938-3300): block Jacobi or block
   Gauss-Seidel in the reference's row order, bitwise its single-thread
   result; even Nr_prec: Nr_prec/2 forward+backward pairs, odd: Nr_prec/2+1
   forward sweeps */
void lar_perform_BJ_or_GS_iterations(int Matrix_id, int Use_rhs, int Ini_zero,
                                     int Nr_prec, int Ndof, double* V, double* B);
/* lah_intf.h:269 -- BCRS with ILU(k): V = (LU)^-1 B (B NULL: the stored
   rhs), forward reduction and back substitution in the reference's order
   (las_bcrs_intf.c:3319-3510), bitwise its result.  Empty for CRS storage in
   the reference (las_crs_intf.c:1019-1057): records an error there */
void lar_perform_rhsub(int Matrix_id, int Ndof, double* V, double* B);

/* Restarted GMRES(krylbas) on a BCRS matrix, or a CRS matrix with BLOCK_GS,
   with the preconditioner
   allocated and filled on it (lsr_mkb_core_gmres, lsd_mkb/lsd_mkb_core/
   lss_mkb_core.c:2317, over the storage operations as lsr_mkb_core_compreres
   applies them, :1915-1990): none, block Jacobi / block Gauss-Seidel with
   nr_prec sweeps per application (smoothing), MULTI_ILU (one ILU correction
   as smoothing) or BLOCK_ILU ((LU)^-1 of the residual).  The rhs is the
   stored one; x (host) is the initial guess unless ini_zero and receives the
   solution.  tol: relative to the first preconditioned residual norm.
   Returns 1 on convergence, 0 otherwise, <0 on error */
int mf_lar_gmres(int Matrix_id, int krylbas, int max_iter, double tol,
                 int nr_prec, int ini_zero, double* x, int* iters,
                 double* res_rel, double* resid);

/* Order of the BCRS block Gauss-Seidel sweeps (no reference counterpart):
   0 the reference's row order (default, bitwise the reference), 1 colour
   classes of a greedy colouring, swept in place one class per launch
   (classes ascending forward, descending backward; the same iteration
   count rule).  Returns the number of colours (1 for order 0) */
int mf_lar_set_gs_order(int Matrix_id, int order);

/* Thread count T whose row chunks the CRS BLOCK_GS sweeps reproduce (no
   reference counterpart: the reference takes it from its OpenMP runtime,
   las_crs_intf.c:910-911); 1 <= T <= 2^20, default 1 (forward Gauss-Seidel in
   DOF order).  One wave walks each chunk's rows in order, so small T is a
   compatibility mode, not a fast smoother */
int mf_lar_set_gs_threads(int Matrix_id, int Nthreads);

/* HIP-event time (ms, averaged over reps) of one BCRS operation on
   device-resident vectors: what 0 residual, 1 one forward sweep of the block
   Jacobi / Gauss-Seidel preconditioner, 2 the ILU(k) factorisation, 3 one
   ILU(k) solve (no reference counterpart; measurement only) */
int mf_lar_bench(int Matrix_id, int what, int reps, double* ms);

/* Inspection (no reference counterpart; lar_get_SM_and_LV_crs_from_crs /
   _from_bcrs, las_crs_intf.c:1063, las_bcrs_intf.c:3812, are the nearest):
   storage type, block size, rows of blocks and stored entries/blocks */
int mf_lar_get_sizes(int Matrix_id, int* storage, int* block_size,
                     int* nrblocks, long long* nr_stored);
/* row pointers, column indices, values (BCRS: 16 per block, column-major)
   and rhs; any pointer may be NULL */
int mf_lar_get_system(int Matrix_id, int* row_ptr, int* col_ind, double* val,
                      double* rhs);
/* BCRS: LU factors (16 per block row, column-major) and 0-based pivots
   (4 per block row) of the filled diagonal-block preconditioner; with ILU(k)
   the factors of every block of its pattern (16 each, mf_lar_get_ilu_pattern)
   and no pivots */
int mf_lar_get_block_precon(int Matrix_id, double* lu, int* piv);
/* BCRS block ILU(k): the block pattern its factors live on (k > 0: the grown
   pattern; k = 0: the matrix's), row_ptr [nrblocks+1] and col_ind
   [*nr_stored], either may be NULL; returns k */
int mf_lar_get_ilu_pattern(int Matrix_id, int* row_ptr, int* col_ind, long long* nr_stored);

#ifdef __cplusplus
}
#endif
#endif
