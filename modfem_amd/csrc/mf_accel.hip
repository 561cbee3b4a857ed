// Accelerator seam of the reference's create-assemble path with its own names
// and signatures (utd_util/uts_accel_intf.c:70-103, tmd_opencl/tmh_ocl.h:
// 266-292, implemented by tmd_opencl/tms_ocl_num_int.c): the driver
// utr_create_assemble_stiff_mat (utd_util/uts_ls_intf.c:3480-3560, built with
// GPU_ASSEMBLING) registers the solver's CRS arrays (tmr_get_crs_structure,
// called from lar_allocate_SM_and_LV_crs), uploads the pattern and assembly
// tables (tmr_perform_creation_crs), streams each colour's elements as the
// reference packs them (utr_prepare_geo_dat_for_all_color_streaming,
// uts_accel_intf.c:518-700; pdr_prepare_pde_coeff_streaming,
// pdd_conv_diff/linear_solver_interface/opencl/std_lin/
// pds_conv_diff_ls_intf_opencl.c) into tmr_perform_create_assemble_accel,
// and finally adds the device-assembled values into the solver's CRS
// (tmr_prepare_final_crs) and releases the device copies (tmr_cleanup_crs).
//
// Streamed data, per call (one colour of Nr_elems elements):
//   Gauss_dat      [Ngauss][4]          x, y, z, weight on the reference element
//   Shape_fun_dat  [Ngauss][Num_shap][4] phi, dphi/dxi, dphi/deta, dphi/dzeta
//   El_geo_dat     [Nr_elems][3 Num_geo_dofs]  vertex coordinates, element order
//   El_pde_dat     the two packings the reference produces:
//     (0, 0, 1)  LAPLACE: a = I, one source value s per integration point
//     (0, 20, 0) TEST_SCALAR: per element axx..azz, bx..bz, tx..tz, c, qx..qz, s
// Integration follows apr_num_int_el's terms for these coefficients
// (aps_std_lin_intf.c:1824-2862): K(u,w) += [(A grad u) . grad w + (b . grad u)
// w + (t . grad w) u + c u w] |det J| w_g, F(w) += [q . grad w + s w] |det J|
// w_g (tetrahedra: det J / 6 as apr_elem_calc_3D_tetra), then
// val[table[w + u k]] += K(u, w), rhs[table[k^2 + w]] += F(w).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <vector>

#include "mf_device.h"

using namespace mf;

namespace {

constexpr int kCompSm = 1, kCompRhs = 2, kCompBoth = 3;  // APC_COMP_* (aph_intf.h:141-144)

struct AccelCrs {
  int n = 0;
  long long nnz = 0;
  int* h_col = nullptr;
  int* h_row = nullptr;
  double* h_val = nullptr;
  double* h_rhs = nullptr;
  double* d_val = nullptr;
  double* d_rhs = nullptr;
  int nint = 0;
  double* d_data = nullptr;  // gauss, shape, geo, pde data of one call
  size_t data_cap = 0;
  int* d_table = nullptr;    // the call's assembly-table rows
  size_t table_cap = 0;
  bool registered = false, created = false;
};
AccelCrs g_acc;

// one thread per element; NS shape functions = vertices (std-linear)
template <int NS>
__global__ void __launch_bounds__(128) k_accel_elems(
    int nel, int ng, const double* __restrict__ gauss, const double* __restrict__ shape,
    const double* __restrict__ geo, int layout, const double* __restrict__ pde,
    const int* __restrict__ table, int comp, double* val, double* rhs) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nel) return;
  double x[NS][3];
#pragma unroll
  for (int i = 0; i < NS; i++)
#pragma unroll
    for (int d = 0; d < 3; d++) x[i][d] = geo[(size_t)e * 3 * NS + 3 * i + d];
  double a[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, b[3] = {0, 0, 0}, t[3] = {0, 0, 0}, q[3] = {0, 0, 0};
  double c = 0.0, s = 0.0;
  if (layout == 2) {
    const double* p = pde + (size_t)e * 20;
    for (int k = 0; k < 9; k++) a[k] = p[k];
    for (int k = 0; k < 3; k++) { b[k] = p[9 + k]; t[k] = p[12 + k]; q[k] = p[16 + k]; }
    c = p[15];
    s = p[19];
  }
  double K[NS][NS], F[NS];
#pragma unroll
  for (int u = 0; u < NS; u++) {
    F[u] = 0.0;
#pragma unroll
    for (int w = 0; w < NS; w++) K[u][w] = 0.0;
  }
  for (int g = 0; g < ng; g++) {
    const double* sh = shape + (size_t)g * NS * 4;
    // J[3r + c] = d x_r / d xi_c (apr_elem_calc_3D)
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < NS; i++)
#pragma unroll
      for (int r = 0; r < 3; r++) {
        J[3 * r + 0] += x[i][r] * sh[4 * i + 1];
        J[3 * r + 1] += x[i][r] * sh[4 * i + 2];
        J[3 * r + 2] += x[i][r] * sh[4 * i + 3];
      }
    // cofactor inverse (ut_mat3_inv)
    const double det = J[0] * (J[4] * J[8] - J[7] * J[5]) + J[3] * (J[7] * J[2] - J[1] * J[8]) +
                       J[6] * (J[1] * J[5] - J[4] * J[2]);
    const double r = 1.0 / det;
    double Ji[9];
    Ji[0] = (J[4] * J[8] - J[7] * J[5]) * r;
    Ji[3] = (-J[3] * J[8] + J[6] * J[5]) * r;
    Ji[6] = (J[3] * J[7] - J[6] * J[4]) * r;
    Ji[1] = (-J[1] * J[8] + J[7] * J[2]) * r;
    Ji[4] = (J[0] * J[8] - J[6] * J[2]) * r;
    Ji[7] = (-J[0] * J[7] + J[6] * J[1]) * r;
    Ji[2] = (J[1] * J[5] - J[4] * J[2]) * r;
    Ji[5] = (-J[0] * J[5] + J[3] * J[2]) * r;
    Ji[8] = (J[0] * J[4] - J[3] * J[1]) * r;
    const double vol = fabs(NS == 4 ? det * 0.16666666666666666666666666666667 : det) *
                       gauss[4 * g + 3];
    double gx[NS], gy[NS], gz[NS], ph[NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
      const double dx = sh[4 * i + 1], dy = sh[4 * i + 2], dz = sh[4 * i + 3];
      ph[i] = sh[4 * i];
      gx[i] = dx * Ji[0] + dy * Ji[3] + dz * Ji[6];
      gy[i] = dx * Ji[1] + dy * Ji[4] + dz * Ji[7];
      gz[i] = dx * Ji[2] + dy * Ji[5] + dz * Ji[8];
    }
    const double sg = layout == 1 ? pde[(size_t)e * ng + g] : s;
#pragma unroll
    for (int u = 0; u < NS; u++)
#pragma unroll
      for (int w = 0; w < NS; w++)
        K[u][w] += ((a[0] * gx[u] * gx[w] + a[1] * gy[u] * gx[w] + a[2] * gz[u] * gx[w] +
                     a[3] * gx[u] * gy[w] + a[4] * gy[u] * gy[w] + a[5] * gz[u] * gy[w] +
                     a[6] * gx[u] * gz[w] + a[7] * gy[u] * gz[w] + a[8] * gz[u] * gz[w]) +
                    (b[0] * gx[u] + b[1] * gy[u] + b[2] * gz[u]) * ph[w] +
                    (t[0] * gx[w] + t[1] * gy[w] + t[2] * gz[w]) * ph[u] + c * ph[u] * ph[w]) *
                   vol;
#pragma unroll
    for (int w = 0; w < NS; w++) F[w] += (q[0] * gx[w] + q[1] * gy[w] + q[2] * gz[w] + sg * ph[w]) * vol;
  }
  const int* tb = table + (size_t)e * (NS * NS + NS);
  if (comp == kCompSm || comp == kCompBoth)
#pragma unroll
    for (int u = 0; u < NS; u++)
#pragma unroll
      for (int w = 0; w < NS; w++) unsafeAtomicAdd(&val[tb[w + u * NS]], K[u][w]);
  if (comp == kCompRhs || comp == kCompBoth)
#pragma unroll
    for (int w = 0; w < NS; w++) unsafeAtomicAdd(&rhs[tb[NS * NS + w]], F[w]);
}

}  // namespace

extern "C" {

// tmr_prepare_create_assemble_accel (uts_accel_intf.c:70): opens the device
// (MF_DEVICE, default 0) when the library has not been initialised
int tmr_prepare_create_assemble_accel(int Problem_id, int Monitor) {
  (void)Problem_id; (void)Monitor;
  if (!stream()) {
    const char* dev = std::getenv("MF_DEVICE");
    if (device_init(dev ? std::atoi(dev) : 0)) return -2;
  }
  return 0;
}

// tmr_get_crs_structure (tmh_ocl.h:285): the solver's CRS arrays (host), to
// which tmr_prepare_final_crs adds the device-assembled values
int tmr_get_crs_structure(int Nrdof_glob, int nnz, int* crs_col_ind, int* crs_row_ptr,
                          double* crs_val, double* rhs) {
  if (Nrdof_glob < 1 || nnz < 1 || !crs_col_ind || !crs_row_ptr || !crs_val || !rhs) {
    set_error("tmr_get_crs_structure: invalid arguments");
    return -1;
  }
  if (tmr_prepare_create_assemble_accel(0, 0)) return -2;
  g_acc.n = Nrdof_glob;
  g_acc.nnz = nnz;
  g_acc.h_col = crs_col_ind;
  g_acc.h_row = crs_row_ptr;
  g_acc.h_val = crs_val;
  g_acc.h_rhs = rhs;
  if (g_acc.d_val) (void)hipFree(g_acc.d_val);
  if (g_acc.d_rhs) (void)hipFree(g_acc.d_rhs);
  g_acc.d_val = g_acc.d_rhs = nullptr;
  MF_HIP(hipMalloc(&g_acc.d_val, (size_t)nnz * sizeof(double)));
  MF_HIP(hipMalloc(&g_acc.d_rhs, (size_t)Nrdof_glob * sizeof(double)));
  MF_HIP(hipMemset(g_acc.d_val, 0, (size_t)nnz * sizeof(double)));
  MF_HIP(hipMemset(g_acc.d_rhs, 0, (size_t)Nrdof_glob * sizeof(double)));
  g_acc.registered = true;
  return 0;
}

// tmr_perform_creation_crs (tmh_ocl.h:268): checks that the solver's CRS is
// registered; every assembling call reads its own elements' table rows
int tmr_perform_creation_crs(int Problem_id, int Nr_int_ent, int* Asse_pos_first_dof_int_ent,
                             int* Assembly_table, int Max_dofs_int_ent) {
  (void)Problem_id;
  if (!g_acc.registered) {
    set_error("tmr_perform_creation_crs: no CRS registered (tmr_get_crs_structure)");
    return -1;
  }
  if (Nr_int_ent < 1 || !Asse_pos_first_dof_int_ent || !Assembly_table || Max_dofs_int_ent < 1) {
    set_error("tmr_perform_creation_crs: invalid arguments");
    return -1;
  }
  g_acc.nint = Nr_int_ent;
  g_acc.created = true;
  return 0;
}

// tmr_perform_create_assemble_accel (uts_accel_intf.c:76-103): integrate the
// Nr_elems elements of one colour (integration entities First_elem_in_color_
// index ...) from the streamed data and add them into the device copy of the
// registered CRS through their assembly-table rows
int tmr_perform_create_assemble_accel(int Problem_id, int Level_id, int Comp_type, int Nr_elems,
                                      int First_elem_in_color_index, int* L_elem_id,
                                      int* Asse_pos_first_dof_int_ent, int* Assembly_table,
                                      int Ngauss, double* Gauss_dat_host, int Num_shap,
                                      double* Shape_fun_dat_host, int Num_geo_dofs,
                                      double* El_geo_dat_host, int Num_dofs,
                                      int Size_global_pde_data, int Size_per_element_pde_data,
                                      int Size_per_int_point_pde_data, double* El_pde_dat_host,
                                      int Max_dofs_int_ent) {
  (void)Problem_id; (void)Level_id; (void)L_elem_id; (void)Max_dofs_int_ent;
  if (!g_acc.created) {
    set_error("tmr_perform_create_assemble_accel: call tmr_get_crs_structure and "
              "tmr_perform_creation_crs first");
    return -1;
  }
  if (Nr_elems < 0 || Ngauss < 1 || Ngauss > 64 || !Gauss_dat_host || !Shape_fun_dat_host ||
      !El_geo_dat_host || !Asse_pos_first_dof_int_ent || !Assembly_table ||
      First_elem_in_color_index < 0 || First_elem_in_color_index + Nr_elems > g_acc.nint) {
    set_error("tmr_perform_create_assemble_accel: invalid arguments");
    return -1;
  }
  if (Comp_type != kCompSm && Comp_type != kCompRhs && Comp_type != kCompBoth) return 0;
  if (Nr_elems == 0) return 0;
  const int k = Num_shap;
  if (!((k == 6 || k == 4) && Num_geo_dofs == k && Num_dofs == k)) {
    set_error("tmr_perform_create_assemble_accel: std-linear prisms (6) and tetrahedra (4) with "
              "one unknown per vertex only");
    return -1;
  }
  int layout = 0;
  if (Size_global_pde_data == 0 && Size_per_element_pde_data == 0 &&
      Size_per_int_point_pde_data == 1)
    layout = 1;  // LAPLACE packing
  else if (Size_global_pde_data == 0 && Size_per_element_pde_data == 20 &&
           Size_per_int_point_pde_data == 0)
    layout = 2;  // TEST_SCALAR packing
  if (!layout || !El_pde_dat_host) {
    set_error("tmr_perform_create_assemble_accel: PDE data must be one value per integration "
              "point (Laplace) or 20 coefficients per element (scalar test)");
    return -1;
  }
  // the call's table rows, each k^2 + k entries
  const size_t rowlen = (size_t)k * k + k;
  std::vector<int> tab((size_t)Nr_elems * rowlen);
  for (int i = 0; i < Nr_elems; i++) {
    const int* src = Assembly_table + Asse_pos_first_dof_int_ent[First_elem_in_color_index + i];
    for (size_t j = 0; j < rowlen; j++) {
      const int v = src[j];
      const bool bad = j < (size_t)k * k ? (v < 0 || v >= g_acc.nnz) : (v < 0 || v >= g_acc.n);
      if (bad) {
        set_error("tmr_perform_create_assemble_accel: assembly-table entry out of range");
        return -1;
      }
      tab[(size_t)i * rowlen + j] = v;
    }
  }
  const size_t n_gauss = 4 * (size_t)Ngauss, n_shape = 4 * (size_t)k * Ngauss;
  const size_t n_geo = (size_t)Nr_elems * 3 * k;
  const size_t n_pde = layout == 1 ? (size_t)Nr_elems * Ngauss : (size_t)Nr_elems * 20;
  const size_t need = n_gauss + n_shape + n_geo + n_pde;
  if (need > g_acc.data_cap) {
    if (g_acc.d_data) (void)hipFree(g_acc.d_data);
    g_acc.d_data = nullptr;
    MF_HIP(hipMalloc(&g_acc.d_data, need * sizeof(double)));
    g_acc.data_cap = need;
  }
  if (tab.size() > g_acc.table_cap) {
    if (g_acc.d_table) (void)hipFree(g_acc.d_table);
    g_acc.d_table = nullptr;
    MF_HIP(hipMalloc(&g_acc.d_table, tab.size() * sizeof(int)));
    g_acc.table_cap = tab.size();
  }
  double* dg = g_acc.d_data;
  double* ds = dg + n_gauss;
  double* dgeo = ds + n_shape;
  double* dp = dgeo + n_geo;
  hipStream_t st = stream();
  MF_HIP(hipMemcpyAsync(dg, Gauss_dat_host, n_gauss * sizeof(double), hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(ds, Shape_fun_dat_host, n_shape * sizeof(double), hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(dgeo, El_geo_dat_host, n_geo * sizeof(double), hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(dp, El_pde_dat_host, n_pde * sizeof(double), hipMemcpyHostToDevice, st));
  MF_HIP(hipMemcpyAsync(g_acc.d_table, tab.data(), tab.size() * sizeof(int), hipMemcpyHostToDevice,
                        st));
  const dim3 grid((Nr_elems + 127) / 128), block(128);
  if (k == 6)
    hipLaunchKernelGGL(k_accel_elems<6>, grid, block, 0, st, Nr_elems, Ngauss, dg, ds, dgeo,
                       layout, dp, g_acc.d_table, Comp_type, g_acc.d_val, g_acc.d_rhs);
  else
    hipLaunchKernelGGL(k_accel_elems<4>, grid, block, 0, st, Nr_elems, Ngauss, dg, ds, dgeo,
                       layout, dp, g_acc.d_table, Comp_type, g_acc.d_val, g_acc.d_rhs);
  MF_HIP(hipGetLastError());
  MF_HIP(hipStreamSynchronize(st));
  return 0;
}

// tmr_prepare_final_crs (uts_ls_intf.c:2767; tms_ocl_num_int.c): the
// device-assembled values and rhs added into the registered CRS arrays; the
// device copies start again from zero
int tmr_prepare_final_crs(void) {
  if (!g_acc.registered) {
    set_error("tmr_prepare_final_crs: no CRS registered");
    return -1;
  }
  std::vector<double> v((size_t)g_acc.nnz), r((size_t)g_acc.n);
  MF_HIP(hipMemcpy(v.data(), g_acc.d_val, v.size() * sizeof(double), hipMemcpyDeviceToHost));
  MF_HIP(hipMemcpy(r.data(), g_acc.d_rhs, r.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < v.size(); i++) g_acc.h_val[i] += v[i];
  for (size_t i = 0; i < r.size(); i++) g_acc.h_rhs[i] += r[i];
  MF_HIP(hipMemset(g_acc.d_val, 0, v.size() * sizeof(double)));
  MF_HIP(hipMemset(g_acc.d_rhs, 0, r.size() * sizeof(double)));
  return 0;
}

// tmr_cleanup_crs (tmh_ocl.h:280): releases the device copies
int tmr_cleanup_crs(void) {
  void* ptrs[] = {g_acc.d_val, g_acc.d_rhs, g_acc.d_data, g_acc.d_table};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  g_acc = AccelCrs();
  return 0;
}

}  // extern "C"
