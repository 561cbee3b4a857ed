/*
 * modfem_lar.h -- the storage/operations seam of the reference's MKB solver
 * with its own names and signatures, exported by libmodfem_hip.so for block
 * size 1 (CRS).  Each declaration replaces the reference function of the same
 * name in lsd_mkb/lah_intf.h (line numbers below, relative to
 * modfem2017/src/); the dispatcher lsd_mkb/las_intf.c and the CRS module
 * lsd_mkb/lad_crs/las_crs_intf.c are what a ModFEM build drops when linking
 * this library instead.
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

/* lah_intf.h:90 -- CRS pattern from sorted neighbour-block lists; returns
   the matrix id (<0 on error; only LAC_STORAGE_CRS / _UNDEFINED with
   Block_size 1 are accepted) */
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
/* lah_intf.h:192 -- NO_PRECON (0) and BLOCK_JACOBI (1) */
int lar_allocate_preconditioner(int Matrix_id, int Precon, int Second_arg);
/* lah_intf.h:203 */
int lar_fill_preconditioner(int Matrix_id);
/* lah_intf.h:210 */
int lar_free_preconditioner(int Matrix_id);
/* lah_intf.h:219 -- V = (Use_rhs ? B (or stored rhs) : 0) - A X */
void lar_compute_residual(int Matrix_id, int Use_rhs, int Ini_zero, int Ndof,
                          double* X, double* B, double* V);
/* lah_intf.h:252 -- BLOCK_JACOBI: V = D^-1 (B - (A-D) V), Nr_prec sweeps */
void lar_perform_BJ_or_GS_iterations(int Matrix_id, int Use_rhs, int Ini_zero,
                                     int Nr_prec, int Ndof, double* V, double* B);
/* lah_intf.h:269 -- empty for CRS storage in the reference
   (las_crs_intf.c:1019-1057); records an error */
void lar_perform_rhsub(int Matrix_id, int Ndof, double* V, double* B);

#ifdef __cplusplus
}
#endif
#endif
