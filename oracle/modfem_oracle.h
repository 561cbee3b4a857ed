/*
 * modfem_oracle.h -- CPU restatement of ModFEM's std-linear
 * integrate -> assemble -> MKB GMRES path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or as the timed CPU baseline.  The product (modfem_amd, libmodfem_hip.so)
 * never links, loads or calls it.
 *
 * Every function names the reference routine (file:line, relative to
 * modfem2017/src/) whose algorithm, data layout and summation order it
 * follows.  Index conventions: element/node/face indices are 0-based here;
 * the reference uses 1-based IDs in the same order.
 */
#ifndef MODFEM_ORACLE_H
#define MODFEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_PRISM 5 /* matches MMC_PRISM */
#define ORC_TETRA 7 /* matches MMC_TETRA */

/* coefficient-vector indicator slots (pds_conv_diff_weakform.c:87-276) */
enum {
  ORC_IND_M = 1, ORC_IND_AXX = 2, ORC_IND_AXY = 3, ORC_IND_AXZ = 4,
  ORC_IND_AYX = 5, ORC_IND_AYY = 6, ORC_IND_AYZ = 7, ORC_IND_AZX = 8,
  ORC_IND_AZY = 9, ORC_IND_AZZ = 10, ORC_IND_BX = 11, ORC_IND_BY = 12,
  ORC_IND_BZ = 13, ORC_IND_TX = 14, ORC_IND_TY = 15, ORC_IND_TZ = 16,
  ORC_IND_C = 17, ORC_IND_L = 18, ORC_IND_QX = 19, ORC_IND_QY = 20,
  ORC_IND_QZ = 21, ORC_IND_S = 22
};

/* source/Dirichlet-data models */
#define ORC_SRC_CONST 0        /* s and g constant per element / face      */
#define ORC_SRC_DIFF_IN_CUBE 1 /* u=exp(-(x2+y2+z2)),     s=-lapl u, g=u   */
#define ORC_SRC_DIFF_IN_BOX 2  /* u=exp(-(x2+y2+z2)/100), s=-lapl u, g=u   */
#define ORC_SRC_LINEAR 9       /* u=0.3+x-2y+0.5z, s=0 (oracle pinning only) */

/* per-element PDE coefficients, constant over the element:
   a[9] = axx axy axz ayx ayy ayz azx azy azz, b[3], t[3], c, q[3], s, m */
#define ORC_NCOEF 21
#define ORC_C_A 0
#define ORC_C_B 9
#define ORC_C_T 12
#define ORC_C_C 15
#define ORC_C_Q 16
#define ORC_C_S 19
#define ORC_C_M 20

/* quadrature (aps_std_util.c:1335-1436, aps_gauss_util_std.c:126-) */
int orc_quadr_3d(int el_type, double *xg, double *wg);
int orc_quadr_2d(int fa_is_quad, double *xg, double *wg);

/* shape functions + element calc (aps_std_util.c:463-820, 826-1226) */
double orc_elem_calc_3d(int el_type, int control, const double *eta,
                        const double *node_coor, double *phi, double *dphix,
                        double *dphiy, double *dphiz, double *xcoor,
                        double *vec_nor);
double orc_mat3_inv(const double *mat, double *mat_inv);

/* element integral (aps_std_lin_intf.c:1537-2870), nreq=1 */
int orc_num_int_el(int el_type, const double *node_coor, const int *ind,
                   const double *coef, int src_model, double *stiff,
                   double *rhs);
/* Dirichlet penalty face (pds_conv_diff_weakform_std.c:253-629) */
int orc_face_penalty(int el_type, int side, const double *node_coor,
                     double g_const, int src_model, double *stiff,
                     double *rhs);

/* Neumann face with constant flux g (pds_conv_diff_weakform_std.c:513-522) */
int orc_face_neumann(int el_type, int side, const double *node_coor, double g,
                     double *rhs);
int orc_face_neumann_exact(int el_type, int side, const double *nc, int src_model,
                           double *rhs);

/* lists and DOF structures (uts_ls_intf.c:102-398, sis_mkb_intf.c:396-1142)
   el_nodes: nel x 6 (tets use the first 4).  Returns number of DOFs.
   dof_node[n]: node index of DOF n (first-appearance order).
   node_dof[nno]: inverse map (-1 when the node carries no DOF). */
int orc_dof_list(int nel, const int *el_type, const int *el_nodes, int nno,
                 int *dof_node, int *node_dof);
/* CRS pattern (las_crs_intf.c:72-294 fed by sis_mkb_intf.c:762-1142).
   row_ptr[n+1]; col_ind sized by orc_crs_count (first call with col_ind=NULL
   returns nnz). */
long orc_crs_pattern(int nel, const int *el_type, const int *el_nodes,
                     const int *node_dof, int n, int *row_ptr, int *col_ind);
/* assembly tables (las_crs_intf.c:354-432): for every int-ent (elements,
   then boundary faces) (k+1)*k ints; table_ptr[nint+1] offsets. */
long orc_assembly_tables(int nel, const int *el_type, const int *el_nodes,
                         int nbf, const int *bf_elem, const int *node_dof,
                         const int *row_ptr, const int *col_ind,
                         long *table_ptr, int *table);

/* test hook: the factor the hybrid mesh module applies to face coordinates
   on tetrahedron side 3 (0.75 as the reference; 1.0 plants the undistorted
   map) */
void orc_set_tet_side3_scale(double s);

/* test hook: 1 plants the downwind neighbour in the DG interior upwind terms
   (0 restores the reference's upwind side) */
void orc_dg_set_plant_downwind(int on);

/* L2 norm and H1 seminorm of u - u_exact (pdr_error_test_exact) */
int orc_std_errors(int nel, const int *el_type, const int *el_nodes,
                   const double *coor, const double *u, int src_model,
                   double *l2h1);

/* integrate + assemble sequentially in int-ent order (sis_mkb_intf.c:1843-1866,
   las_crs_intf.c:440-513).  val[nnz], rhs[n] zeroed inside. */
int orc_integrate_assemble(int nel, const int *el_type, const int *el_nodes,
                           const double *node_coor_all, const int *ind,
                           const double *coef_el, int src_model, int nbf,
                           const int *bf_elem, const int *bf_side,
                           const double *bf_g, const long *table_ptr,
                           const int *table, double *val, double *rhs);

/* linear algebra (las_crs_intf.c:707-1013) */
void orc_fill_jacobi(int n, const int *row_ptr, const int *col_ind,
                     const double *val, int *diag_ptr, double *diag_precon);
void orc_compute_residual(int n, const int *row_ptr, const int *col_ind,
                          const double *val, int use_rhs, const double *b,
                          const double *x, double *v);
void orc_bj_iteration(int n, const int *row_ptr, const int *col_ind,
                      const double *val, const int *diag_ptr,
                      const double *diag_precon, int use_rhs,
                      const double *b, double *v);
/* BLOCK_GS sweeps as R runs them on nthreads OpenMP threads
   (las_crs_intf.c:905-998) */
void orc_crs_gs_chunked(int n, const int *row_ptr, const int *col_ind,
                        const double *val, const int *diag_ptr,
                        const double *diag_precon, int use_rhs, const double *b,
                        int nthreads, int nr_prec, double *v);

/* restarted left-preconditioned GMRES with CGS and Givens
   (lss_mkb_core.c:2317-2668), Jacobi preconditioner applied as one BJ sweep
   through compreres (lss_mkb_core.c:1861-2097).
   ini_zero: x is taken as zero on entry.  Returns 1 on convergence. */
int orc_gmres(int n, const int *row_ptr, const int *col_ind, const double *val,
              const double *b, int precon, int krylbas, int max_iter,
              double tol_rel, int ini_zero, double *x, int *iters,
              double *res_rel_out, double *res_out);

/* the same with an absolute tolerance on the preconditioned residual
   (ABS_RES; lss_mkb_core.c:2220-2225 passes one of the two, the other 0) */
int orc_gmres_tol(int n, const int *row_ptr, const int *col_ind, const double *val,
                  const double *b, int precon, int krylbas, int max_iter,
                  double tol_rel, double tol_abs, int ini_zero, double *x,
                  int *iters, double *res_rel_out, double *res_out);

/* ILU(0) on the CRS pattern (Saad Alg. 10.4, IKJ; the reference's block
   ILU(0) with 1x1 blocks) and its forward/backward substitution; precon 2
   in orc_gmres / orc_gmres_tol selects it */
int orc_ilu0(int n, const int *rp, const int *ci, const double *val, double *lu);
void orc_ilu_solve(int n, const int *rp, const int *ci, const double *lu, const double *in,
                   double *out);

/* Gauss-Seidel sweeps in a row order (NULL: natural, the reference's GS
   with one thread); orc_set_gs selects order and sweeps for precon 3 */
void orc_gs_sweeps(int n, const int *row_ptr, const int *col_ind, const double *val,
                   const int *diag_ptr, const double *diag_precon, const int *order,
                   int nsweeps, int use_rhs, const double *b, double *v);
void orc_set_gs(const int *order, int nsweeps);

/* DG prism approximation (apd_dg_prism): tensor base pdeg = 100 pz + px
   (px 1..3, pz 1..3, or pz = 0 with px 1, 2), complete base of degree p as
   ORC_DG_COMPLETE + p (p 1..3): shape functions, number of them, the
   penalty degree of a face (pdeg2 <= 0: boundary), prism quadrature,
   element integral (nreq = 1; column-major stiff[j*ns + i], i = test
   function; arrays sized for ns <= 128) */
#define ORC_DG_COMPLETE 1000
int orc_dg_shape(int pdeg, const double *eta, double *phi, double *dx, double *dy,
                 double *dz);
int orc_dg_ns(int pdeg);
int orc_dg_penalty_degree(int pdeg1, int pdeg2);
int orc_dg_quadr(int pdeg, double *xg, double *wg);
int orc_dg_num_int_el(int pdeg, const double *node_coor, const int *ind, const double *coef,
                      int src_model, double *stiff, double *rhs);

/* DG faces: quadrature (quad != 0: side faces), the second neighbour's
   local coordinates of a first-neighbour point (g1, g2: global node numbers
   of the two elements), face area, one face's matrices (bc 0 interior,
   1 Dirichlet, 2 Neumann; stiff (2 ns)^2 or ns^2, column-major), and the
   element's squared L2 / H1-seminorm errors against the exact solution */
int orc_dg_face_quadr(int pdeg, int quad, double *xg, double *wg);
void orc_dg_face_eta(int side, const double *xf, double *eta);
double orc_dg_face_ds(int side, const double *node_coor, const double *eta, double *normal,
                      double *x);
int orc_dg_map_point(int side1, const int *g1, int side2, const int *g2, const double *eta,
                     double *eta2);
double orc_dg_face_area(const double *node_coor, int side);
double orc_dg_eval(int pdeg, const double *nc, const double *dofs, const double *eta, double *x);
int orc_dg_face_int2(int pdeg1, int pdeg2, int bc, const double *nc1, const int *g1, int side1,
                     const double *coef1, const double *nc2, const int *g2, int side2,
                     const double *coef2, double penalty, double hf, int src_model, double bval,
                     double *stiff, double *rhs);
int orc_dg_face_int(int pdeg, int bc, const double *nc1, const int *g1, int side1,
                    const double *coef1, const double *nc2, const int *g2, int side2,
                    const double *coef2, double penalty, double hf, int src_model, double bval,
                    double *stiff, double *rhs);
int orc_dg_errors(int pdeg, const double *node_coor, const double *dofs, int src_model,
                  double *l2sq, double *h1sq);

double orc_dg_assemble_sample(int pdeg, int nel, const double *nodes, const int *gids, int nfaces,
                              const int *faces, const int *ind, const double *coef,
                              double penalty, int src_model);

/* BCRS storage with 4x4 column-major blocks (lad_bcrs): pattern, element
   assembly, residual, LAPACK-style 4x4 LU (dgetf2 / dgetrs), block Jacobi /
   Gauss-Seidel iterations in the one-thread order */
int orc_bcrs_pattern(int nb, const int *nb_ptr, const int *nb_ind, int *row_ptr, int *col_ind);
void orc_bcrs_assemble(const int *row_ptr, const int *col_ind, int k, const int *l_bl,
                       const double *stiff, const double *rhs_e, double *val, double *rhs);
void orc_bcrs_residual(int nb, const int *row_ptr, const int *col_ind, const double *val,
                       int use_rhs, const double *x, const double *b, double *v);
int orc_lu4(double *a, int *ipiv);
void orc_lu4_solve(const double *lu, const int *ipiv, double *b);
void orc_bcrs_bj_gs(int nb, const int *row_ptr, const int *col_ind, const int *diag_ptr,
                    const double *val, const double *lu, const int *ipiv, int gs, int use_rhs,
                    int nr_prec, const double *b, double *v);
void orc_inv4(double *a, const int *ipiv);
int orc_lun(int n, double *a, int *ipiv);
void orc_invn(int n, double *a, const int *ipiv);
int orc_block_ilu0(int nb, const int *ns, const int *nptr, const int *nbr,
                   const long long *doff, const long long *ooff, double *val);
void orc_block_ilu_solve(int nb, const int *ns, const long long *xoff, const int *nptr,
                         const int *nbr, const long long *doff, const long long *ooff,
                         const double *val, const double *b, double *v);
int orc_bcrs_ilu0(int nb, const int *rp, const int *ci, const int *diag, double *ilu);
long orc_bcrs_iluk_pattern(int nb, const int *rp, const int *ci, int k, int cap, int *rp2,
                           int *ci2);
void orc_bcrs_rhsub(int nb, const int *rp, const int *ci, const int *diag, const double *ilu,
                    const double *b, double *v);

/* BLAS level-1/2 restatements (netlib reference algorithms) */
void orc_drotg(double *a, double *b, double *c, double *s);

/* threads used by the OpenMP parts (0 = all) */
void orc_set_threads(int nthreads);

#ifdef __cplusplus
}
#endif
/* DG on tetrahedra (complete base of degree p; no reference counterpart) */
int orc_dgt_ns(int p);
int orc_dgt_shape(int p, const double *eta, double *phi, double *dx, double *dy, double *dz);
int orc_dgt_exponents(int p, int *ex);
int orc_dgt_quadr(int p, double *xg, double *wg);
int orc_dgt_num_int_el(int p, const double *node_coor, const int *ind, const double *coef,
                       int src_model, double *stiff, double *rhs);
void orc_dgt_face_eta(int side, const double *xf, double *eta);
double orc_dgt_face_ds(int side, const double *node_coor, const double *eta, double *normal,
                       double *x);
double orc_dgt_face_area(const double *node_coor, int side);
int orc_dgx_map_point(int t1, int s1, const int *g1, const double *eta, int t2, int s2,
                      const int *g2, double *eta2);
int orc_dgx_face_int(int t1, int pdeg1, int t2, int pdeg2, int bc, const double *nc1,
                     const int *g1, int side1, const double *coef1, const double *nc2,
                     const int *g2, int side2, const double *coef2, double penalty, double hf,
                     int src_model, double bval, double *stiff, double *rhs);
double orc_dgt_eval(int p, const double *nc, const double *dofs, const double *eta, double *x);

#endif
