/*
 * modfem_lsr.h -- the solver-module seam of the reference's MKB solver with
 * its own names and signatures, exported by libmodfem_hip.so.  Each
 * declaration replaces the reference function of the same name in
 * lsd_mkb/lsh_mkb_intf.h (line numbers below, relative to modfem2017/src/);
 * lsd_mkb/lss_mkb_intf.c, lsd_mkb/lsd_mkb_core/lss_mkb_core.c, the dispatcher
 * lsd_mkb/las_intf.c and the modules lad_crs / lad_bcrs are what a ModFEM
 * build drops when linking this library instead.  The problem drivers and
 * sid_mkb (sir_create / sir_solve) stay unchanged.
 *
 * One solver instance owns one matrix of the storage seam (modfem_lar.h) and
 * solves with restarted GMRES on the device.  Single-level solvers only
 * (Level_id 0); ids count from 0 and are destroyed in reverse order.  Return
 * codes are < 0 on error (mf_last_error gives the message) where the
 * reference exits.
 */
#ifndef MODFEM_LSR_H
#define MODFEM_LSR_H

#ifdef __cplusplus
extern "C" {
#endif

/* lsh_mkb_intf.h:165 -- control parameters from Filename (mkb.dat format;
   Max_iter, Error_type, Error_tolerance, Monitoring_level >= 0 override it);
   *Max_num_levels_p returns 1, *Storage_type the file's storage; Parallel
   must be 0.  Returns the solver id */
int lsr_mkb_init(int Solver_type, int Parallel, int* Max_num_levels_p, int* Storage_type,
                 char* Filename, int Max_iter, int Error_type, double Error_tolerance,
                 int Monitoring_level);
/* lsh_mkb_intf.h:185 -- after lsr_mkb_free_matrix; last created first */
int lsr_mkb_destroy(int Solver_id);
/* lsh_mkb_intf.h:193 -- lar_allocate_SM_and_LV with the enforced storage
   type, else the control file's (CRS for block size 1, BCRS for 4) */
int lsr_mkb_create_matrix(int Solver_id, int Level_id, int Storage_type, int Nrblocks,
                          int Nrdof_glob, int Block_size, int Max_sm_size, int* Nrdofbl,
                          int* Posglob, int* Nroffbl, int** L_offbl);
/* lsh_mkb_intf.h:213 -- Comp_type 0 (LSC_SOLVE) clears values and rhs,
   1 (LSC_RESOLVE) the rhs */
int lsr_mkb_clear_matrix(int Solver_id, int Level_id, int Comp_type);
/* lsh_mkb_intf.h:227 */
int lsr_mkb_fill_assembly_table_int_ent(int Solver_id, int Level_id, int Nr_dof_bl, int* L_bl_id,
                                        int* L_bl_nrdof, int* Assembly_table_int_ent);
/* lsh_mkb_intf.h:244 */
int lsr_mkb_assemble_local_stiff_mat_with_table(int Solver_id, int Level_id, int Comp_type,
                                                int Nr_dof_bl, int* Assembly_table_int_ent,
                                                double* Stiff_mat, double* Rhs_vect,
                                                char* Rewr_dofs);
/* lsh_mkb_intf.h:266 */
int lsr_mkb_assemble_local_sm(int Solver_id, int Level_id, int Comp_type, int Nr_dof_bl,
                              int* L_bl_id, int* L_bl_nrdof, double* Stiff_mat, double* Rhs_vect,
                              char* Rewr_dofs);
/* lsh_mkb_intf.h:288 */
int lsr_mkb_free_matrix(int Solver_id);
/* lsh_mkb_intf.h:296 -- the control file's preconditioner (CRS: none or
   block Jacobi; BCRS: also block Gauss-Seidel and block ILU(k)); returns the
   number of diagonal blocks */
int lsr_mkb_create_precon(int Solver_id, int Level_id);
/* lsh_mkb_intf.h:306 */
int lsr_mkb_fill_precon(int Solver_id, int Level_id);
/* lsh_mkb_intf.h:317 -- restarted GMRES(krylbas of the file) on the device;
   *Nr_iter > 0 and *Toler > 0 override the file; X (host or device) is the
   initial guess unless Ini_zero and receives the solution; B != NULL
   replaces the stored rhs.  Returns 1 converged, 0 not; *Nr_iter, *Toler
   (final measure of the file's kind) and *Conv_rate (mean reduction per
   iteration) come back */
int lsr_mkb_solve(int Solver_id, int Comp_type, int Ini_zero, double* X, double* B, int* Nr_iter,
                  double* Toler, int Monitor, double* Conv_rate);

#ifdef __cplusplus
}
#endif
#endif
