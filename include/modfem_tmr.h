/*
 * modfem_tmr.h -- the accelerator seam of the reference's create-assemble
 * path with its own names and signatures, exported by libmodfem_hip.so.  The
 * declarations replace tmd_opencl/tms_ocl_num_int.c (the OpenCL integration
 * and assembly) as utd_util/uts_accel_intf.c:70-103 and tmd_opencl/
 * tmh_ocl.h:266-292 declare it; utr_create_assemble_stiff_mat
 * (utd_util/uts_ls_intf.c, GPU_ASSEMBLING build), the streaming packers and
 * the CRS module that registers its arrays stay unchanged.
 *
 * Streamed data per call (one colour of elements): Gauss_dat [Ngauss][4]
 * (x, y, z, weight), Shape_fun_dat [Ngauss][Num_shap][4] (phi and its three
 * reference derivatives), El_geo_dat [Nr_elems][3 Num_geo_dofs] (vertex
 * coordinates in element order), El_pde_dat in one of the reference's two
 * packings: one source value per integration point (Laplace: sizes 0, 0, 1)
 * or 20 coefficients per element (axx..azz, bx..bz, tx..tz, c, qx..qz, s:
 * sizes 0, 20, 0).  Std-linear prisms and tetrahedra.  Return codes are < 0
 * on error (mf_last_error gives the message).
 */
#ifndef MODFEM_TMR_H
#define MODFEM_TMR_H

#ifdef __cplusplus
extern "C" {
#endif

/* uts_accel_intf.c:70 -- opens the device (MF_DEVICE, default 0) if needed */
int tmr_prepare_create_assemble_accel(int Problem_id, int Monitor);
/* tmh_ocl.h:285 -- registers the solver's CRS arrays (host memory): the
   device-assembled values are added into crs_val and rhs by
   tmr_prepare_final_crs */
int tmr_get_crs_structure(int Nrdof_glob, int nnz, int* crs_col_ind, int* crs_row_ptr,
                          double* crs_val, double* rhs);
/* tmh_ocl.h:268 */
int tmr_perform_creation_crs(int Problem_id, int Nr_int_ent, int* Asse_pos_first_dof_int_ent,
                             int* Assembly_table, int Max_dofs_int_ent);
/* uts_accel_intf.c:76 -- integrates the Nr_elems elements of one colour
   (integration entities First_elem_in_color_index ...) and adds them through
   their assembly-table rows (Comp_type 1 matrix, 2 rhs, 3 both) */
int tmr_perform_create_assemble_accel(int Problem_id, int Level_id, int Comp_type, int Nr_elems,
                                      int First_elem_in_color_index, int* L_elem_id,
                                      int* Asse_pos_first_dof_int_ent, int* Assembly_table,
                                      int Ngauss, double* Gauss_dat_host, int Num_shap,
                                      double* Shape_fun_dat_host, int Num_geo_dofs,
                                      double* El_geo_dat_host, int Num_dofs,
                                      int Size_global_pde_data, int Size_per_element_pde_data,
                                      int Size_per_int_point_pde_data, double* El_pde_dat_host,
                                      int Max_dofs_int_ent);
/* uts_ls_intf.c:2767 (tms_ocl_num_int.c) -- adds the device values into the
   registered CRS */
int tmr_prepare_final_crs(void);
/* tmh_ocl.h:280 */
int tmr_cleanup_crs(void);

#ifdef __cplusplus
}
#endif
#endif
