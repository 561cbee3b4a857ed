// HIP kernels for gfx950: element integration + assembly and Dirichlet
// penalty faces (the solver kernels are in mf_solver.hip).
//
// Reference behaviour followed (paths relative to modfem2017/src/):
//   element integral  apd_std_lin/aps_std_lin_intf.c:1537-2870 (nreq=1)
//   geometry          apd_std_lin/aps_std_util.c:463-1226, 2770-2822
//   quadrature        apd_std_lin/aps_std_util.c:1335-1530,
//                     apd_std_lin/aps_gauss_util_std.c:152-299 (literals)
//   penalty faces     pdd_conv_diff/weak_formulation/approx_std/
//                     pds_conv_diff_weakform_std.c:253-629
//   assembly          lsd_mkb/lad_crs/las_crs_intf.c:440-513
//
// Design: a group of k lanes integrates one element (k = 6 for prisms, 4 for
// tetrahedra): geometry (Jacobian, determinant, inverse) is evaluated in
// FP64, one Gauss point per lane, and staged in LDS; each lane then
// accumulates one row of the element matrix and scatters it with hardware
// FP64 atomics through compact per-element row-offset tables (1 byte per
// entry).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "mf_device.h"

namespace mf {

static hipStream_t g_stream = nullptr;
static int g_device = -1;

int device_init(int device) {
  if (g_device == device && g_stream) return 0;
  int ndev = 0;
  MF_HIP(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) {
    set_error("mf_init: device index out of range (no HIP device?)");
    return -1;
  }
  MF_HIP(hipSetDevice(device));
  MF_HIP(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  g_device = device;
  return 0;
}

hipStream_t stream() { return g_stream; }

// ---------------------------------------------------------------------
// quadrature literals (aps_gauss_util_std.c) -- see the oracle for the
// derivation of the point order
// ---------------------------------------------------------------------
constexpr double G2 = 0.577350269189626;
constexpr double T1 = 0.166666666666667, T2 = 0.666666666666667;
constexpr double TW = 0.333333333333333;
constexpr double TA = 0.585410196624969, TB = 0.138196601125011;

struct PrismQ {
  static constexpr int NG = 6;
  __device__ static constexpr double x(int g) { return (g % 3 == 2) ? T2 : T1; }
  __device__ static constexpr double y(int g) { return (g % 3 == 1) ? T2 : T1; }
  __device__ static constexpr double z(int g) { return g < 3 ? -G2 : G2; }
  __device__ static constexpr double w(int) { return 0.5 * TW * 1.0; }
};
struct TetQ {
  static constexpr int NG = 4;
  // points (TA,TB,TB) (TB,TB,TA) (TB,TB,TB) (TB,TA,TB)
  __device__ static constexpr double x(int g) { return g == 0 ? TA : TB; }
  __device__ static constexpr double y(int g) { return g == 3 ? TA : TB; }
  __device__ static constexpr double z(int g) { return g == 1 ? TA : TB; }
  __device__ static constexpr double w(int) { return 0.25; }
};

// reference shape functions and local derivatives
template <int NS>
__device__ __forceinline__ void ref_shape(double e0, double e1, double e2,
                                          double* phi, double* dx, double* dy,
                                          double* dz) {
  if constexpr (NS == 6) {
    phi[0] = (-e2 + 1) / 2. * (1 - e0 - e1);
    phi[1] = (-e2 + 1) / 2. * e0;
    phi[2] = (-e2 + 1) / 2. * e1;
    phi[3] = (e2 + 1) / 2. * (1 - e0 - e1);
    phi[4] = (e2 + 1) / 2. * e0;
    phi[5] = (e2 + 1) / 2. * e1;
    dx[0] = (e2 - 1) / 2.; dx[1] = (-e2 + 1) / 2.; dx[2] = 0.0;
    dx[3] = (-e2 - 1) / 2.; dx[4] = (e2 + 1) / 2.; dx[5] = 0.0;
    dy[0] = (e2 - 1) / 2.; dy[1] = 0.0; dy[2] = (-e2 + 1) / 2.;
    dy[3] = (-e2 - 1) / 2.; dy[4] = 0.0; dy[5] = (e2 + 1) / 2.;
    dz[0] = -0.5 * (1 - e0 - e1); dz[1] = -0.5 * e0; dz[2] = -0.5 * e1;
    dz[3] = 0.5 * (1 - e0 - e1); dz[4] = 0.5 * e0; dz[5] = 0.5 * e1;
  } else {
    phi[0] = 1.0 - e0 - e1 - e2; phi[1] = e0; phi[2] = e1; phi[3] = e2;
    dx[0] = -1.0; dx[1] = 1.0; dx[2] = 0.0; dx[3] = 0.0;
    dy[0] = -1.0; dy[1] = 0.0; dy[2] = 1.0; dy[3] = 0.0;
    dz[0] = -1.0; dz[1] = 0.0; dz[2] = 0.0; dz[3] = 1.0;
  }
}

// Jacobian dxdeta[3r+c] = dx_r/deta_c, cofactor inverse (ut_mat3_inv),
// returns the determinant (scaled by 1/6 for tetrahedra)
template <int NS>
__device__ __forceinline__ double jacobian(const double (*xn)[3], const double* dx,
                                           const double* dy, const double* dz,
                                           double* J, double* Ji) {
  if constexpr (NS == 6) {
#pragma unroll
    for (int k = 0; k < 9; k++) J[k] = 0.0;
#pragma unroll
    for (int i = 0; i < 6; i++) {
      J[0] += xn[i][0] * dx[i]; J[1] += xn[i][0] * dy[i]; J[2] += xn[i][0] * dz[i];
      J[3] += xn[i][1] * dx[i]; J[4] += xn[i][1] * dy[i]; J[5] += xn[i][1] * dz[i];
      J[6] += xn[i][2] * dx[i]; J[7] += xn[i][2] * dy[i]; J[8] += xn[i][2] * dz[i];
    }
  } else {
    J[0] = xn[1][0] - xn[0][0]; J[3] = xn[1][1] - xn[0][1]; J[6] = xn[1][2] - xn[0][2];
    J[1] = xn[2][0] - xn[0][0]; J[4] = xn[2][1] - xn[0][1]; J[7] = xn[2][2] - xn[0][2];
    J[2] = xn[3][0] - xn[0][0]; J[5] = xn[3][1] - xn[0][1]; J[8] = xn[3][2] - xn[0][2];
  }
  const double* m = J;
  double rjac = m[0] * (m[4] * m[8] - m[7] * m[5]) + m[3] * (m[7] * m[2] - m[1] * m[8]) +
                m[6] * (m[1] * m[5] - m[4] * m[2]);
  double r = 1.0 / rjac;
  Ji[0] = (m[4] * m[8] - m[7] * m[5]) * r;
  Ji[3] = (-m[3] * m[8] + m[6] * m[5]) * r;
  Ji[6] = (m[3] * m[7] - m[6] * m[4]) * r;
  Ji[1] = (m[7] * m[2] - m[1] * m[8]) * r;
  Ji[4] = (m[0] * m[8] - m[6] * m[2]) * r;
  Ji[7] = (-m[0] * m[7] + m[6] * m[1]) * r;
  Ji[2] = (m[1] * m[5] - m[4] * m[2]) * r;
  Ji[5] = (-m[0] * m[5] + m[3] * m[2]) * r;
  Ji[8] = (m[0] * m[4] - m[3] * m[1]) * r;
  if constexpr (NS == 4) rjac *= 0.16666666666666666666666666666667;
  return rjac;
}

struct IntArgs {
  int nel;
  const double* coor;      // 3*n, DOF order
  const int* el_dof;       // [6][nel]
  const uint8_t* el_off;   // [36][nel]
  const long long* sl_ptr;  // SELL-64 slice offsets
  double* val;
  double* rhs;
  int coef_model, src_model;
  TermMask mask;
  double cc[NCOEF];        // uniform coefficients
  const double* coef;      // per-element SoA
  int write_elements;      // 1: [count][42] column-major; 2: row-major buffers below
  double* emat;            // [count][42] when write_elements == 1
  void* erow;              // [nel][6][6] rows (double or float) when == 2
  void* ef;                // [nel][6] load vector entries
  int e0, count;
  const int* elist;        // optional element list (write_elements == 1)
};

struct Coefs {
  double a[9], b[3], t[3], c, q[3], s, m;
};

__device__ __forceinline__ void load_coefs(const IntArgs& A, int e, Coefs& C) {
#pragma unroll
  for (int k = 0; k < 9; k++) C.a[k] = A.cc[C_A + k];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    C.b[k] = A.cc[C_B + k]; C.t[k] = A.cc[C_T + k]; C.q[k] = A.cc[C_Q + k];
  }
  C.c = A.cc[C_C]; C.s = A.cc[C_S]; C.m = A.cc[C_M];
  const size_t N = A.nel;
  if (A.coef_model == CM_KAPPA || A.coef_model == CM_KAPPA_B || A.coef_model == CM_KAPPA_S) {
    double kap = A.coef[e];
    C.a[0] = kap; C.a[4] = kap; C.a[8] = kap;
    C.a[1] = C.a[2] = C.a[3] = C.a[5] = C.a[6] = C.a[7] = 0.0;
    if (A.coef_model == CM_KAPPA_B) {
      C.b[0] = A.coef[N + e]; C.b[1] = A.coef[2 * N + e]; C.b[2] = A.coef[3 * N + e];
    }
    if (A.coef_model == CM_KAPPA_S) C.s = A.coef[N + e];
  } else if (A.coef_model == CM_FULL) {
#pragma unroll
    for (int k = 0; k < 9; k++) C.a[k] = A.coef[(C_A + k) * N + e];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      C.b[k] = A.coef[(C_B + k) * N + e];
      C.t[k] = A.coef[(C_T + k) * N + e];
      C.q[k] = A.coef[(C_Q + k) * N + e];
    }
    C.c = A.coef[C_C * N + e]; C.s = A.coef[C_S * N + e]; C.m = A.coef[C_M * N + e];
  }
}

// Cooperative element kernel: a group of NS lanes integrates one element
// (prism: 6 lanes, 10 elements per 64-lane wavefront; tetra: 4 lanes, 16).
//   phase 0  lane i loads node i (DOF index + FP64 coordinates) into LDS;
//   phase 1  lane g evaluates Gauss point g: Jacobian, determinant and
//            inverse in FP64, physical shape gradients, |det|*w and the
//            source value -> LDS;
//   phase 2  lane w accumulates row w of the element matrix (test function
//            w against every trial function u) over all Gauss points in the
//            arithmetic type R, plus load-vector entry w;
//   phase 3  lane w scatters its row with FP64 hardware atomics through the
//            compact row-offset table.
template <class Q>
__device__ __forceinline__ void gauss_point(int g, double* e) {
  // runtime index into the unrolled literal set
  double x = 0, y = 0, z = 0;
#pragma unroll
  for (int k = 0; k < Q::NG; k++)
    if (k == g) { x = Q::x(k); y = Q::y(k); z = Q::z(k); }
  e[0] = x; e[1] = y; e[2] = z;
}

template <int NS>
__device__ __forceinline__ double ref_phi(int i, double e0, double e1, double e2) {
  double phi[NS], dx[NS], dy[NS], dz[NS];
  ref_shape<NS>(e0, e1, e2, phi, dx, dy, dz);
  double r = 0;
#pragma unroll
  for (int k = 0; k < NS; k++)
    if (k == i) r = phi[k];
  return r;
}

constexpr int INT_BLOCK = 128;

template <int NS, class Q, int TS, typename R>
__device__ __forceinline__ void integrate_block(const IntArgs& A, int bid) {
  constexpr int NG = Q::NG;
  constexpr int EPW = 64 / NS;                 // elements per wavefront
  constexpr int EPB = EPW * (INT_BLOCK / 64);  // elements per block
  constexpr int GPD = 4 * NS + 2;              // doubles per Gauss point record
  __shared__ double s_x[EPB][NS][3];
  __shared__ double s_gp[EPB][NG][GPD];
  __shared__ double s_cf[EPB][NCOEF];
  __shared__ int s_dof[EPB][NS];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int le = wave * EPW + lane / NS;       // element slot in block
  const int sub = lane % NS;
  const bool act_lane = (lane / NS) < EPW;
  const int t = bid * EPB + le;
  const bool active = act_lane && t < A.count;
  const int e = active ? (A.elist ? A.elist[t] : A.e0 + t) : 0;
  const size_t N = A.nel;
  // phase 0
  if (active) {
    int d = A.el_dof[sub * N + e];
    s_dof[le][sub] = d;
    s_x[le][sub][0] = A.coor[3 * (size_t)d];
    s_x[le][sub][1] = A.coor[3 * (size_t)d + 1];
    s_x[le][sub][2] = A.coor[3 * (size_t)d + 2];
    if (sub == 0) {
      Coefs C;
      load_coefs(A, e, C);
#pragma unroll
      for (int k = 0; k < 9; k++) s_cf[le][C_A + k] = C.a[k];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        s_cf[le][C_B + k] = C.b[k]; s_cf[le][C_T + k] = C.t[k]; s_cf[le][C_Q + k] = C.q[k];
      }
      s_cf[le][C_C] = C.c; s_cf[le][C_S] = C.s; s_cf[le][C_M] = C.m;
    }
  }
  __syncthreads();
  // phase 1: lane sub handles Gauss points sub, sub+NS, ...
  if (active) {
    for (int g = sub; g < NG; g += NS) {
      double eta[3];
      gauss_point<Q>(g, eta);
      double phi[NS], dx[NS], dy[NS], dz[NS], J[9], Ji[9];
      ref_shape<NS>(eta[0], eta[1], eta[2], phi, dx, dy, dz);
      double det = jacobian<NS>(s_x[le], dx, dy, dz, J, Ji);
      double* rec = s_gp[le][g];
#pragma unroll
      for (int i = 0; i < NS; i++) {
        rec[3 * i] = dx[i] * Ji[0] + dy[i] * Ji[3] + dz[i] * Ji[6];
        rec[3 * i + 1] = dx[i] * Ji[1] + dy[i] * Ji[4] + dz[i] * Ji[7];
        rec[3 * i + 2] = dx[i] * Ji[2] + dy[i] * Ji[5] + dz[i] * Ji[8];
      }
      double wg = 0;
#pragma unroll
      for (int k = 0; k < NG; k++)
        if (k == g) wg = Q::w(k);
      rec[3 * NS] = fabs(det) * wg;
#pragma unroll
      for (int i = 0; i < NS; i++) rec[3 * NS + 2 + i] = phi[i];
      double sval = s_cf[le][C_S];
      if (A.src_model != SRC_CONST) {
        double xc[3] = {0, 0, 0};
        if constexpr (NS == 6) {
#pragma unroll
          for (int i = 0; i < 6; i++) {
            xc[0] += s_x[le][i][0] * phi[i]; xc[1] += s_x[le][i][1] * phi[i];
            xc[2] += s_x[le][i][2] * phi[i];
          }
        } else {
#pragma unroll
          for (int d = 0; d < 3; d++)
            xc[d] = s_x[le][0][d] + eta[0] * J[3 * d] + eta[1] * J[3 * d + 1] + eta[2] * J[3 * d + 2];
        }
        double u, lapl;
        exact_sol(A.src_model, xc, &u, &lapl);
        sval = -lapl;
      }
      rec[3 * NS + 1] = sval;
    }
  }
  __syncthreads();
  if (!active) return;
  // phase 2: row w = sub
  const int w = sub;
  const TermMask M = A.mask;
  const double* cf = s_cf[le];
  R a[9], b[3], tt[3], q[3];
#pragma unroll
  for (int k = 0; k < 9; k++) a[k] = (R)cf[C_A + k];
#pragma unroll
  for (int k = 0; k < 3; k++) { b[k] = (R)cf[C_B + k]; tt[k] = (R)cf[C_T + k]; q[k] = (R)cf[C_Q + k]; }
  const R cm = (R)((M.reac ? cf[C_C] : 0.0) + (M.m ? cf[C_M] : 0.0));
  R K[NS], F = 0;
#pragma unroll
  for (int u = 0; u < NS; u++) K[u] = 0;
#pragma unroll 1
  for (int g = 0; g < NG; g++) {
    const double* rec = s_gp[le][g];
    const R v = (R)rec[3 * NS];
    const R gwx = (R)rec[3 * w], gwy = (R)rec[3 * w + 1], gwz = (R)rec[3 * w + 2];
    const R phw = (R)rec[3 * NS + 2 + w];
    // A^T grad_w: k_diff(u) = grad_u . (A^T grad_w) = sum_rc a_rc dc_u dr_w
    R aw0, aw1, aw2;
    if (M.diff_full) {
      aw0 = a[0] * gwx + a[3] * gwy + a[6] * gwz;
      aw1 = a[1] * gwx + a[4] * gwy + a[7] * gwz;
      aw2 = a[2] * gwx + a[5] * gwy + a[8] * gwz;
    } else if (M.diff_diag) {
      aw0 = a[0] * gwx; aw1 = a[4] * gwy; aw2 = a[8] * gwz;
    } else {
      aw0 = aw1 = aw2 = 0;
    }
    R tw = 0;
    if (TS >= TS_ALL && M.tconv) tw = tt[0] * gwx + tt[1] * gwy + tt[2] * gwz;
#pragma unroll
    for (int u = 0; u < NS; u++) {
      const R gux = (R)rec[3 * u], guy = (R)rec[3 * u + 1], guz = (R)rec[3 * u + 2];
      R k = gux * aw0 + guy * aw1 + guz * aw2;
      if (TS >= TS_DIFF_CONV && M.conv) k += (b[0] * gux + b[1] * guy + b[2] * guz) * phw;
      if (TS >= TS_ALL) {
        const R phu = (R)rec[3 * NS + 2 + u];
        k += phu * tw + cm * phu * phw;
      }
      K[u] += k * v;
    }
    R f = 0;
    if (TS >= TS_ALL && M.q) f += q[0] * gwx + q[1] * gwy + q[2] * gwz;
    if (M.s) f += (R)rec[3 * NS + 1] * phw;
    F += f * v;
  }
  if (A.write_elements == 1) {
    double* o = A.emat + (size_t)t * 42;
#pragma unroll
    for (int u = 0; u < NS; u++) o[u * NS + w] = (double)K[u];
    o[NS * NS + w] = (double)F;
    return;
  }
  if (A.write_elements == 2) {
    R* o = reinterpret_cast<R*>(A.erow) + (size_t)e * 36 + w * 6;
#pragma unroll
    for (int u = 0; u < NS; u++) o[u] = K[u];
    reinterpret_cast<R*>(A.ef)[(size_t)e * 6 + w] = F;
    return;
  }
  // phase 3
  const int dw = s_dof[le][w];
  const long long base = A.sl_ptr[dw >> 6] + (dw & 63);
#pragma unroll
  for (int u = 0; u < NS; u++) {
    const int off = A.el_off[(size_t)(w * 6 + u) * N + e];
    unsafeAtomicAdd(&A.val[base + 64LL * off], (double)K[u]);
  }
  unsafeAtomicAdd(&A.rhs[dw], (double)F);
}

template <int NS, class Q, int TS, typename R>
__global__ void __launch_bounds__(INT_BLOCK) k_integrate(IntArgs A) {
  integrate_block<NS, Q, TS, R>(A, blockIdx.x);
}

// Face point (c0, c1) of the reference triangle -> tetrahedron coordinates,
// as the hybrid mesh module maps it (mmr_fa_elem_coor, mmd_t4_prism/
// mms_prism_t4_intf.cpp:2004-2037): sides 0-2 lie in the coordinate planes;
// side 3 (vertices 1, 2, 3) takes the face coordinates scaled by 0.75 first,
// which moves the points toward vertex 1 on that face.  The tetrahedral
// results of the reference (its L2/H1 error figures on the NASTRAN cube)
// are made with this map, so it is kept as the reference has it.
__device__ __forceinline__ void tet_face_eta(int side, double c0, double c1, double* eta) {
  switch (side) {
    case 0: eta[0] = c1; eta[1] = c0; eta[2] = 0.0; break;
    case 1: eta[0] = c0; eta[1] = 0.0; eta[2] = c1; break;
    case 2: eta[0] = 0.0; eta[1] = c1; eta[2] = c0; break;
    default:
      c0 *= 0.75;
      c1 *= 0.75;
      eta[0] = 1.0 - c0 - c1; eta[1] = c0; eta[2] = c1;
      break;
  }
}

// ---------------------------------------------------------------------
// Dirichlet penalty faces: K += 1e7 phi_j phi_i dA, F += 1e7 g phi_i dA
// on the owning element's DOFs.
// ---------------------------------------------------------------------
struct FaceArgs {
  int nel, nbf;
  const double* coor;
  const int* el_dof;
  const uint8_t* el_off;
  const int* bf_elem;
  const int* bf_side;
  const double* bf_g;
  const int* bf_type;       // per face: 2 Neumann flux, else Dirichlet penalty
  const long long* sl_ptr;  // SELL-64 slice offsets
  double* val;
  double* rhs;
  int src_model;
  int write_faces;
  double* fmat;
  int f0, count;
};

template <int NS>
__global__ void __launch_bounds__(256) k_faces(FaceArgs A) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= A.count) return;
  const int f = A.f0 + t;
  const int e = A.bf_elem[f], side = A.bf_side[f];
  const size_t N = A.nel;
  int dof[NS];
  double xn[NS][3];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    dof[i] = A.el_dof[i * N + e];
    xn[i][0] = A.coor[3 * (size_t)dof[i]];
    xn[i][1] = A.coor[3 * (size_t)dof[i] + 1];
    xn[i][2] = A.coor[3 * (size_t)dof[i] + 2];
  }
  const double gconst = A.bf_g ? A.bf_g[f] : 0.0;
  const bool neu = A.bf_type && A.bf_type[f] == BC_NEUMANN;
  double K[NS][NS], F[NS];
#pragma unroll
  for (int i = 0; i < NS; i++) {
    F[i] = 0;
#pragma unroll
    for (int j = 0; j < NS; j++) K[i][j] = 0;
  }
  const bool quad = (NS == 6 && side >= 2);
  const int ng = quad ? 4 : 3;
  for (int g = 0; g < ng; g++) {
    double c0, c1, wgt;
    if (quad) {
      c0 = (g / 2 == 0) ? -G2 : G2;
      c1 = (g % 2 == 0) ? -G2 : G2;
      wgt = 0.5 * 1.0 * 1.0;
    } else {
      c0 = (g == 2) ? T2 : T1;
      c1 = (g == 1) ? T2 : T1;
      wgt = 0.5 * TW;
    }
    double eta[3];
    if constexpr (NS == 6) {
      switch (side) {
        case 0: eta[0] = c1; eta[1] = c0; eta[2] = -1.0; break;
        case 1: eta[0] = c0; eta[1] = c1; eta[2] = 1.0; break;
        case 2: eta[0] = 0.5 * (c0 + 1.0); eta[1] = 0.0; eta[2] = c1; break;
        case 3: eta[0] = 0.5 * (-c0 + 1.0); eta[1] = 1.0 - eta[0]; eta[2] = c1; break;
        default: eta[0] = 0.0; eta[1] = 0.5 * (-c0 + 1.0); eta[2] = c1; break;
      }
    } else {
      tet_face_eta(side, c0, c1, eta);
    }
    double phi[NS], dx[NS], dy[NS], dz[NS], J[9], Ji[9];
    ref_shape<NS>(eta[0], eta[1], eta[2], phi, dx, dy, dz);
    double ds[3];
    if constexpr (NS == 6) {
      jacobian<NS>(xn, dx, dy, dz, J, Ji);
      switch (side) {
        case 0:
          ds[0] = -J[3] * J[7] + J[6] * J[4]; ds[1] = -J[6] * J[1] + J[0] * J[7];
          ds[2] = -J[0] * J[4] + J[3] * J[1]; break;
        case 1:
          ds[0] = J[3] * J[7] - J[6] * J[4]; ds[1] = J[6] * J[1] - J[0] * J[7];
          ds[2] = J[0] * J[4] - J[3] * J[1]; break;
        case 2:
          ds[0] = J[3] * J[8] - J[6] * J[5]; ds[1] = J[6] * J[2] - J[0] * J[8];
          ds[2] = J[0] * J[5] - J[3] * J[2]; break;
        case 3:
          ds[0] = (J[4] - J[3]) * J[8] - (J[7] - J[6]) * J[5];
          ds[1] = (J[7] - J[6]) * J[2] - (J[1] - J[0]) * J[8];
          ds[2] = (J[1] - J[0]) * J[5] - (J[4] - J[3]) * J[2]; break;
        default:
          ds[0] = J[5] * J[7] - J[8] * J[4]; ds[1] = J[8] * J[1] - J[2] * J[7];
          ds[2] = J[2] * J[4] - J[5] * J[1]; break;
      }
    } else {
      const int fn[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};  // outward (v1 x v2)
      const double *p0 = xn[fn[side][0]], *p1 = xn[fn[side][1]], *p2 = xn[fn[side][2]];
      double v1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      double v2[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
      ds[0] = v1[1] * v2[2] - v1[2] * v2[1];
      ds[1] = v1[2] * v2[0] - v1[0] * v2[2];
      ds[2] = v1[0] * v2[1] - v1[1] * v2[0];
    }
    double area = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]) * wgt;
    double gv = gconst;
    if (neu) {  // flux only (pds_conv_diff_weakform_std.c:513-522)
      if (A.src_model != SRC_CONST) {  // the exact normal flux
        double xc[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < NS; i++) {
          xc[0] += xn[i][0] * phi[i]; xc[1] += xn[i][1] * phi[i]; xc[2] += xn[i][2] * phi[i];
        }
        gv = exact_flux(A.src_model, xc, ds);
      }
#pragma unroll
      for (int i = 0; i < NS; i++) F[i] += gv * phi[i] * area;
      continue;
    }
    if (A.src_model != SRC_CONST) {
      double xc[3] = {0, 0, 0};
      if constexpr (NS == 6) {
        for (int i = 0; i < 6; i++) {
          xc[0] += xn[i][0] * phi[i]; xc[1] += xn[i][1] * phi[i]; xc[2] += xn[i][2] * phi[i];
        }
      } else {
        for (int i = 0; i < 4; i++) {
          xc[0] += xn[i][0] * phi[i]; xc[1] += xn[i][1] * phi[i]; xc[2] += xn[i][2] * phi[i];
        }
      }
      double lapl;
      exact_sol(A.src_model, xc, &gv, &lapl);
    }
    const double pen = 1.0e7;
#pragma unroll
    for (int j = 0; j < NS; j++)
#pragma unroll
      for (int i = 0; i < NS; i++) K[i][j] += pen * phi[j] * phi[i] * area;
#pragma unroll
    for (int i = 0; i < NS; i++) F[i] += pen * gv * phi[i] * area;
  }
  if (A.write_faces) {
    double* o = A.fmat + (size_t)t * 42;
    for (int j = 0; j < NS; j++)
      for (int i = 0; i < NS; i++) o[j * NS + i] = K[i][j];
    for (int i = 0; i < NS; i++) o[NS * NS + i] = F[i];
    return;
  }
#pragma unroll
  for (int w = 0; w < NS; w++) {
    const long long base = A.sl_ptr[dof[w] >> 6] + (dof[w] & 63);
#pragma unroll
    for (int u = 0; u < NS; u++) {
      const int off = A.el_off[(size_t)(w * 6 + u) * N + e];
      unsafeAtomicAdd(&A.val[base + 64LL * off], K[w][u]);
    }
    unsafeAtomicAdd(&A.rhs[dof[w]], F[w]);
  }
}

__global__ void k_zero(double* p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) p[i] = 0.0;
}

int launch_zero(double* p, long long n) {
  if (n <= 0) return 0;
  int blocks = (int)std::min<long long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, g_stream, p, n);
  MF_HIP(hipGetLastError());
  return 0;
}

static void fill_int_args(Problem& P, IntArgs& A) {
  A.nel = P.nel;
  A.coor = P.d_coor;
  A.el_dof = P.d_el_dof;
  A.el_off = P.d_el_off;
  A.sl_ptr = P.d_sl_ptr;
  A.val = P.d_val;
  A.rhs = P.d_rhs;
  A.coef_model = P.coef_model;
  A.src_model = P.src_model;
  A.mask = P.mask;
  for (int k = 0; k < NCOEF; k++) A.cc[k] = P.coef_const[k];
  A.coef = P.d_coef;
  A.elist = nullptr;
}

template <int NS, class Q, typename R>
static void launch_int_ts(int ts, int count, IntArgs& A) {
  constexpr int EPB = (64 / NS) * (INT_BLOCK / 64);
  dim3 g((count + EPB - 1) / EPB);
  if (ts == TS_DIFF)
    hipLaunchKernelGGL((k_integrate<NS, Q, TS_DIFF, R>), g, dim3(INT_BLOCK), 0, g_stream, A);
  else if (ts == TS_DIFF_CONV)
    hipLaunchKernelGGL((k_integrate<NS, Q, TS_DIFF_CONV, R>), g, dim3(INT_BLOCK), 0, g_stream, A);
  else
    hipLaunchKernelGGL((k_integrate<NS, Q, TS_ALL, R>), g, dim3(INT_BLOCK), 0, g_stream, A);
}

int launch_integrate(Problem& P, int precision, int write_elements, double* d_emat,
                     int e0, int ecount, const int* d_elist) {
  if (ecount <= 0) return 0;
  IntArgs A;
  fill_int_args(P, A);
  A.write_elements = write_elements;
  A.emat = d_emat;
  A.e0 = e0;
  A.count = ecount;
  A.elist = d_elist;
  const int ts = P.term_set;
  if (P.el_type == kTetra) {
    if (precision == 32) launch_int_ts<4, TetQ, float>(ts, ecount, A);
    else launch_int_ts<4, TetQ, double>(ts, ecount, A);
  } else {
    if (precision == 32) launch_int_ts<6, PrismQ, float>(ts, ecount, A);
    else launch_int_ts<6, PrismQ, double>(ts, ecount, A);
  }
  MF_HIP(hipGetLastError());
  return 0;
}

int launch_faces(Problem& P, int write_faces, double* d_fmat, int f0, int fcount) {
  if (fcount <= 0) return 0;
  FaceArgs A;
  A.nel = P.nel; A.nbf = P.nbf; A.coor = P.d_coor; A.el_dof = P.d_el_dof;
  A.el_off = P.d_el_off; A.bf_elem = P.d_bf_elem; A.bf_side = P.d_bf_side;
  A.bf_g = P.d_bf_g; A.bf_type = P.d_bf_type; A.sl_ptr = P.d_sl_ptr; A.val = P.d_val; A.rhs = P.d_rhs;
  A.src_model = P.src_model; A.write_faces = write_faces; A.fmat = d_fmat;
  A.f0 = f0; A.count = fcount;
  dim3 g((fcount + 255) / 256);
  if (P.el_type == kTetra)
    hipLaunchKernelGGL(k_faces<4>, g, dim3(256), 0, g_stream, A);
  else
    hipLaunchKernelGGL(k_faces<6>, g, dim3(256), 0, g_stream, A);
  MF_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------
// Row-wise integrate + assemble (default assembly path): one lane owns one
// matrix row (DOF r) and walks the integration entities that contain r in
// ascending order -- elements, then boundary faces -- exactly the order in
// which the reference's sequential create/assemble loop adds into that row
// (sid_mkb/sis_mkb_intf.c:1843-1866).  For each entity the lane evaluates
// the element-matrix row of r (FP64 geometry at every Gauss point) and adds
// it into register accumulators indexed by the row position from the
// compact table.  A wavefront owns one 64-row SELL slice and writes it once
// with coalesced stores: no atomics, no zero fill, bitwise reproducible.
// ---------------------------------------------------------------------
struct RowArgs {
  IntArgs I;              // element data + coefficients (val/rhs unused)
  int n;
  const int* inc_ptr;
  const int* inc_el;
  const uint8_t* inc_loc;
  const int* finc_ptr;
  const int* finc_f;
  const uint8_t* finc_loc;
  const int* bf_elem;
  const int* bf_side;
  const double* bf_g;
  const int* bf_type;     // per face: 2 Neumann flux, else Dirichlet penalty
  double* val;            // SELL-64
  double* rhs;
};

template <int NS>
__device__ __forceinline__ void gather_nodes(const IntArgs& I, int e, double (*xn)[3]) {
  const size_t N = I.nel;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const int d = I.el_dof[i * N + e];
    xn[i][0] = I.coor[3 * (size_t)d];
    xn[i][1] = I.coor[3 * (size_t)d + 1];
    xn[i][2] = I.coor[3 * (size_t)d + 2];
  }
}

// reference shape function i and its local derivatives at eta (i is a
// compile-time index after unrolling, eta runtime)
template <int NS>
__device__ __forceinline__ void ref_one(int i, const double* eta, double& phi, double& dx,
                                        double& dy, double& dz) {
  double p[NS], x[NS], y[NS], z[NS];
  ref_shape<NS>(eta[0], eta[1], eta[2], p, x, y, z);
  phi = p[0]; dx = x[0]; dy = y[0]; dz = z[0];
#pragma unroll
  for (int k = 1; k < NS; k++)
    if (k == i) { phi = p[k]; dx = x[k]; dy = y[k]; dz = z[k]; }
}

// row w of element e's matrix and load vector, Gauss points outer, kept to a
// small live set: node coordinates, the row accumulators and one Gauss
// point's Jacobian inverse
template <int NS, class Q, int TS, typename R>
__device__ __forceinline__ void element_row(const IntArgs& I, int e, int w, R* Krow, R& Fw) {
  double xn[NS][3];
  gather_nodes<NS>(I, e, xn);
  Coefs C;
  load_coefs(I, e, C);
  const TermMask M = I.mask;
  const R cm = (R)((M.reac ? C.c : 0.0) + (M.m ? C.m : 0.0));
#pragma unroll
  for (int u = 0; u < NS; u++) Krow[u] = 0;
  Fw = 0;
#pragma unroll 1
  for (int g = 0; g < Q::NG; g++) {
    double eta[3];
    gauss_point<Q>(g, eta);
    double phi[NS], dx[NS], dy[NS], dz[NS], J[9], Ji[9];
    ref_shape<NS>(eta[0], eta[1], eta[2], phi, dx, dy, dz);
    const double det = jacobian<NS>(xn, dx, dy, dz, J, Ji);
    double wg = 0;
#pragma unroll
    for (int k = 0; k < Q::NG; k++)
      if (k == g) wg = Q::w(k);
    const R v = (R)(fabs(det) * wg);
    double pw, dxw, dyw, dzw;
    pw = phi[0]; dxw = dx[0]; dyw = dy[0]; dzw = dz[0];
#pragma unroll
    for (int k = 1; k < NS; k++)
      if (k == w) { pw = phi[k]; dxw = dx[k]; dyw = dy[k]; dzw = dz[k]; }
    const R gwx = (R)(dxw * Ji[0] + dyw * Ji[3] + dzw * Ji[6]);
    const R gwy = (R)(dxw * Ji[1] + dyw * Ji[4] + dzw * Ji[7]);
    const R gwz = (R)(dxw * Ji[2] + dyw * Ji[5] + dzw * Ji[8]);
    const R phw = (R)pw;
    R aw0, aw1, aw2;
    if (M.diff_full) {
      aw0 = (R)C.a[0] * gwx + (R)C.a[3] * gwy + (R)C.a[6] * gwz;
      aw1 = (R)C.a[1] * gwx + (R)C.a[4] * gwy + (R)C.a[7] * gwz;
      aw2 = (R)C.a[2] * gwx + (R)C.a[5] * gwy + (R)C.a[8] * gwz;
    } else if (M.diff_diag) {
      aw0 = (R)C.a[0] * gwx; aw1 = (R)C.a[4] * gwy; aw2 = (R)C.a[8] * gwz;
    } else {
      aw0 = aw1 = aw2 = 0;
    }
    R tw = 0;
    if (TS >= TS_ALL && M.tconv) tw = (R)C.t[0] * gwx + (R)C.t[1] * gwy + (R)C.t[2] * gwz;
    const R bphw0 = (R)C.b[0] * phw, bphw1 = (R)C.b[1] * phw, bphw2 = (R)C.b[2] * phw;
#pragma unroll
    for (int u = 0; u < NS; u++) {
      const R gux = (R)(dx[u] * Ji[0] + dy[u] * Ji[3] + dz[u] * Ji[6]);
      const R guy = (R)(dx[u] * Ji[1] + dy[u] * Ji[4] + dz[u] * Ji[7]);
      const R guz = (R)(dx[u] * Ji[2] + dy[u] * Ji[5] + dz[u] * Ji[8]);
      R k = gux * aw0 + guy * aw1 + guz * aw2;
      if (TS >= TS_DIFF_CONV && M.conv) k += gux * bphw0 + guy * bphw1 + guz * bphw2;
      if (TS >= TS_ALL) k += (R)phi[u] * (tw + cm * phw);
      Krow[u] += k * v;
    }
    double sval = C.s;
    if (I.src_model != SRC_CONST) {
      double xc[3] = {0, 0, 0};
      if constexpr (NS == 6) {
#pragma unroll
        for (int i = 0; i < 6; i++) {
          xc[0] += xn[i][0] * phi[i]; xc[1] += xn[i][1] * phi[i]; xc[2] += xn[i][2] * phi[i];
        }
      } else {
#pragma unroll
        for (int d = 0; d < 3; d++)
          xc[d] = xn[0][d] + eta[0] * J[3 * d] + eta[1] * J[3 * d + 1] + eta[2] * J[3 * d + 2];
      }
      double uu, lapl;
      exact_sol(I.src_model, xc, &uu, &lapl);
      sval = -lapl;
    }
    R f = 0;
    if (TS >= TS_ALL && M.q) f += (R)C.q[0] * gwx + (R)C.q[1] * gwy + (R)C.q[2] * gwz;
    if (M.s) f += (R)sval * phw;
    Fw += f * v;
  }
}

// penalty-face row w of the owning element's local matrix
template <int NS>
__device__ __forceinline__ void face_row(const RowArgs& A, int f, int w, double* Krow, double& Fw) {
  const int e = A.bf_elem[f], side = A.bf_side[f];
  double xn[NS][3];
  gather_nodes<NS>(A.I, e, xn);
  const double gconst = A.bf_g ? A.bf_g[f] : 0.0;
  const bool neu = A.bf_type && A.bf_type[f] == BC_NEUMANN;
#pragma unroll
  for (int u = 0; u < NS; u++) Krow[u] = 0;
  Fw = 0;
  const bool quad = (NS == 6 && side >= 2);
  const int ng = quad ? 4 : 3;
  for (int g = 0; g < ng; g++) {
    double c0, c1, wgt;
    if (quad) {
      c0 = (g / 2 == 0) ? -G2 : G2;
      c1 = (g % 2 == 0) ? -G2 : G2;
      wgt = 0.5 * 1.0 * 1.0;
    } else {
      c0 = (g == 2) ? T2 : T1;
      c1 = (g == 1) ? T2 : T1;
      wgt = 0.5 * TW;
    }
    double eta[3];
    if constexpr (NS == 6) {
      switch (side) {
        case 0: eta[0] = c1; eta[1] = c0; eta[2] = -1.0; break;
        case 1: eta[0] = c0; eta[1] = c1; eta[2] = 1.0; break;
        case 2: eta[0] = 0.5 * (c0 + 1.0); eta[1] = 0.0; eta[2] = c1; break;
        case 3: eta[0] = 0.5 * (-c0 + 1.0); eta[1] = 1.0 - eta[0]; eta[2] = c1; break;
        default: eta[0] = 0.0; eta[1] = 0.5 * (-c0 + 1.0); eta[2] = c1; break;
      }
    } else {
      tet_face_eta(side, c0, c1, eta);
    }
    double phi[NS], dx[NS], dy[NS], dz[NS], J[9], Ji[9];
    ref_shape<NS>(eta[0], eta[1], eta[2], phi, dx, dy, dz);
    double ds[3];
    if constexpr (NS == 6) {
      jacobian<NS>(xn, dx, dy, dz, J, Ji);
      switch (side) {
        case 0:
          ds[0] = -J[3] * J[7] + J[6] * J[4]; ds[1] = -J[6] * J[1] + J[0] * J[7];
          ds[2] = -J[0] * J[4] + J[3] * J[1]; break;
        case 1:
          ds[0] = J[3] * J[7] - J[6] * J[4]; ds[1] = J[6] * J[1] - J[0] * J[7];
          ds[2] = J[0] * J[4] - J[3] * J[1]; break;
        case 2:
          ds[0] = J[3] * J[8] - J[6] * J[5]; ds[1] = J[6] * J[2] - J[0] * J[8];
          ds[2] = J[0] * J[5] - J[3] * J[2]; break;
        case 3:
          ds[0] = (J[4] - J[3]) * J[8] - (J[7] - J[6]) * J[5];
          ds[1] = (J[7] - J[6]) * J[2] - (J[1] - J[0]) * J[8];
          ds[2] = (J[1] - J[0]) * J[5] - (J[4] - J[3]) * J[2]; break;
        default:
          ds[0] = J[5] * J[7] - J[8] * J[4]; ds[1] = J[8] * J[1] - J[2] * J[7];
          ds[2] = J[2] * J[4] - J[5] * J[1]; break;
      }
    } else {
      const int fn[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};  // outward (v1 x v2)
      const double *p0 = xn[fn[side][0]], *p1 = xn[fn[side][1]], *p2 = xn[fn[side][2]];
      double v1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
      double v2[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
      ds[0] = v1[1] * v2[2] - v1[2] * v2[1];
      ds[1] = v1[2] * v2[0] - v1[0] * v2[2];
      ds[2] = v1[0] * v2[1] - v1[1] * v2[0];
    }
    const double area = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]) * wgt;
    double gv = gconst;
    if (A.I.src_model != SRC_CONST && !neu) {
      double xc[3] = {0, 0, 0};
#pragma unroll
      for (int i = 0; i < NS; i++) {
        xc[0] += xn[i][0] * phi[i]; xc[1] += xn[i][1] * phi[i]; xc[2] += xn[i][2] * phi[i];
      }
      double lapl;
      exact_sol(A.I.src_model, xc, &gv, &lapl);
    }
    double phw = 0;
#pragma unroll
    for (int i = 0; i < NS; i++)
      if (i == w) phw = phi[i];
    if (neu) {
      if (A.I.src_model != SRC_CONST) {  // the exact normal flux
        double xc[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < NS; i++) {
          xc[0] += xn[i][0] * phi[i]; xc[1] += xn[i][1] * phi[i]; xc[2] += xn[i][2] * phi[i];
        }
        gv = exact_flux(A.I.src_model, xc, ds);
      }
      Fw += gv * phw * area;
      continue;
    }
    const double pen = 1.0e7;
#pragma unroll
    for (int u = 0; u < NS; u++) Krow[u] += pen * phi[u] * phw * area;
    Fw += pen * gv * phw * area;
  }
}

constexpr int ROW_WAVES = 4;  // wavefronts (SELL slices) per block

// boundary faces of row r, added after all elements as the reference
// assembles them (face rows of the owning elements, ascending face order)
template <int NS>
__device__ __forceinline__ double face_rows_into(const RowArgs& A, int r, double (*acc)[64],
                                                 int lane) {
  const size_t N = A.I.nel;
  double F = 0.0;
  for (int p = A.finc_ptr[r]; p < A.finc_ptr[r + 1]; p++) {
    const int f = A.finc_f[p], w = A.finc_loc[p];
    const int e = A.bf_elem[f];
    double Kd[NS], Fd;
    face_row<NS>(A, f, w, Kd, Fd);
#pragma unroll
    for (int u = 0; u < NS; u++) acc[A.I.el_off[(size_t)(w * 6 + u) * N + e]][lane] += Kd[u];
    F += Fd;
  }
  return F;
}

// elements, then boundary faces: one lane per row, accumulators in LDS
// (k-major per wave, so the 64 lanes of a wave touch 64 consecutive words)
template <int NS, class Q, int TS, int RMAX, typename R>
__device__ __forceinline__ void assemble_rows_block(const RowArgs& A, int bid) {
  __shared__ double s_acc[ROW_WAVES][RMAX][64];
  const int nsl = (A.n + 63) >> 6;
  const int wv = threadIdx.x >> 6;
  const int sl = bid * ROW_WAVES + wv;
  if (sl >= nsl) return;  // wave-uniform, no block barriers below
  const int lane = threadIdx.x & 63;
  const int r = sl * 64 + lane;
  double(*acc)[64] = s_acc[wv];
#pragma unroll
  for (int k = 0; k < RMAX; k++) acc[k][lane] = 0.0;
  double F = 0.0;
  if (r < A.n) {
    const size_t N = A.I.nel;
    for (int p = A.inc_ptr[r]; p < A.inc_ptr[r + 1]; p++) {
      const int e = A.inc_el[p], w = A.inc_loc[p];
      R Kr[NS], Fr;
      element_row<NS, Q, TS, R>(A.I, e, w, Kr, Fr);
#pragma unroll
      for (int u = 0; u < NS; u++) {
        const int o = A.I.el_off[(size_t)(w * 6 + u) * N + e];
        acc[o][lane] += (double)Kr[u];
      }
      F += (double)Fr;
    }
    F += face_rows_into<NS>(A, r, acc, lane);
  }
  const long long base = A.I.sl_ptr[sl];
  const int len = (int)((A.I.sl_ptr[sl + 1] - base) >> 6);
  for (int k = 0; k < len; k++) A.val[base + 64LL * k + lane] = acc[k][lane];
  if (r < A.n) A.rhs[r] = F;
}

template <int NS, class Q, int TS, int RMAX, typename R>
__global__ void __launch_bounds__(64 * ROW_WAVES) k_assemble_rows(RowArgs A) {
  assemble_rows_block<NS, Q, TS, RMAX, R>(A, blockIdx.x);
}

template <int NS, class Q, int RMAX, typename R>
static void launch_rows_ts(int ts, dim3 g, RowArgs& A) {
  const dim3 b(64 * ROW_WAVES);
  if (ts == TS_DIFF)
    hipLaunchKernelGGL((k_assemble_rows<NS, Q, TS_DIFF, RMAX, R>), g, b, 0, g_stream, A);
  else if (ts == TS_DIFF_CONV)
    hipLaunchKernelGGL((k_assemble_rows<NS, Q, TS_DIFF_CONV, RMAX, R>), g, b, 0, g_stream, A);
  else
    hipLaunchKernelGGL((k_assemble_rows<NS, Q, TS_ALL, RMAX, R>), g, b, 0, g_stream, A);
}

int launch_assemble_rows(Problem& P, int precision) {
  RowArgs A;
  fill_int_args(P, A.I);
  A.n = P.n;
  A.inc_ptr = P.d_inc_ptr; A.inc_el = P.d_inc_el; A.inc_loc = P.d_inc_loc;
  A.finc_ptr = P.d_finc_ptr; A.finc_f = P.d_finc_f; A.finc_loc = P.d_finc_loc;
  A.bf_elem = P.d_bf_elem; A.bf_side = P.d_bf_side; A.bf_g = P.d_bf_g; A.bf_type = P.d_bf_type;
  A.val = P.d_val; A.rhs = P.d_rhs;
  const int nsl = (P.n + 63) / 64;
  dim3 g((nsl + ROW_WAVES - 1) / ROW_WAVES);
  const int mr = P.setup.max_row;
  const int ts = P.term_set;
  if (mr > 32) {
    set_error("row-wise assembly supports rows of at most 32 entries; use the atomic path");
    return -1;
  }
  if (P.el_type == kTetra) {
    if (precision == 32) {
      if (mr <= 16) launch_rows_ts<4, TetQ, 16, float>(ts, g, A);
      else launch_rows_ts<4, TetQ, 32, float>(ts, g, A);
    } else {
      if (mr <= 16) launch_rows_ts<4, TetQ, 16, double>(ts, g, A);
      else launch_rows_ts<4, TetQ, 32, double>(ts, g, A);
    }
  } else {
    if (precision == 32) {
      if (mr <= 24) launch_rows_ts<6, PrismQ, 24, float>(ts, g, A);
      else launch_rows_ts<6, PrismQ, 32, float>(ts, g, A);
    } else {
      if (mr <= 24) launch_rows_ts<6, PrismQ, 24, double>(ts, g, A);
      else launch_rows_ts<6, PrismQ, 32, double>(ts, g, A);
    }
  }
  MF_HIP(hipGetLastError());
  return 0;
}

// Two-pass assembly, pass 2: one lane per matrix row gathers the rows its
// DOF owns in the element-row buffers (ascending element order), adds them
// into LDS accumulators through the per-element offset table and writes the
// SELL slice once, after the boundary faces of the row.  T is the buffer
// type (FP64, or FP32 for the mixed-precision form; sums are FP64 either way).
template <int NS, int RMAX, typename T>
__device__ __forceinline__ void gather_rows_block(const RowArgs& A, int bid, const T* erow,
                                                  const T* ef, const uint8_t* offT) {
  constexpr int ns = NS;
  __shared__ double s_acc[ROW_WAVES][RMAX][64];
  const int nsl = (A.n + 63) >> 6;
  const int wv = threadIdx.x >> 6;
  const int sl = bid * ROW_WAVES + wv;
  if (sl >= nsl) return;
  const int lane = threadIdx.x & 63;
  const int r = sl * 64 + lane;
  double(*acc)[64] = s_acc[wv];
#pragma unroll
  for (int k = 0; k < RMAX; k++) acc[k][lane] = 0.0;
  double F = 0.0;
  if (r < A.n) {
    for (int p = A.inc_ptr[r]; p < A.inc_ptr[r + 1]; p++) {
      const size_t q = (size_t)A.inc_el[p] * 36 + A.inc_loc[p] * 6;
      for (int u = 0; u < ns; u++) acc[offT[q + u]][lane] += (double)erow[q + u];
      F += (double)ef[(size_t)A.inc_el[p] * 6 + A.inc_loc[p]];
    }
    F += face_rows_into<NS>(A, r, acc, lane);
  }
  const long long base = A.I.sl_ptr[sl];
  const int len = (int)((A.I.sl_ptr[sl + 1] - base) >> 6);
  for (int k = 0; k < len; k++) A.val[base + 64LL * k + lane] = acc[k][lane];
  if (r < A.n) A.rhs[r] = F;
}

template <int NS, int RMAX, typename T>
__global__ void __launch_bounds__(64 * ROW_WAVES) k_gather_rows(RowArgs A, const T* erow,
                                                               const T* ef, const uint8_t* offT) {
  gather_rows_block<NS, RMAX, T>(A, blockIdx.x, erow, ef, offT);
}

int launch_assemble_two_pass(Problem& P, int precision) {
  const size_t esz = precision == 32 ? sizeof(float) : sizeof(double);
  if (!P.d_erow || P.erow_bytes < (size_t)P.nel * 42 * esz) {
    if (P.d_erow) (void)hipFree(P.d_erow);
    MF_HIP(hipMalloc(&P.d_erow, (size_t)P.nel * 42 * esz));
    P.erow_bytes = (size_t)P.nel * 42 * esz;
  }
  // pass 1
  IntArgs I;
  fill_int_args(P, I);
  I.write_elements = 2;
  I.erow = P.d_erow;
  I.ef = (char*)P.d_erow + (size_t)P.nel * 36 * esz;
  I.emat = nullptr;
  I.e0 = 0;
  I.count = P.nel;
  const int ts = P.term_set;
  if (P.el_type == kTetra) {
    if (precision == 32) launch_int_ts<4, TetQ, float>(ts, P.nel, I);
    else launch_int_ts<4, TetQ, double>(ts, P.nel, I);
  } else {
    if (precision == 32) launch_int_ts<6, PrismQ, float>(ts, P.nel, I);
    else launch_int_ts<6, PrismQ, double>(ts, P.nel, I);
  }
  MF_HIP(hipGetLastError());
  // pass 2
  RowArgs A;
  A.I = I;
  A.n = P.n;
  A.inc_ptr = P.d_inc_ptr; A.inc_el = P.d_inc_el; A.inc_loc = P.d_inc_loc;
  A.finc_ptr = P.d_finc_ptr; A.finc_f = P.d_finc_f; A.finc_loc = P.d_finc_loc;
  A.bf_elem = P.d_bf_elem; A.bf_side = P.d_bf_side; A.bf_g = P.d_bf_g; A.bf_type = P.d_bf_type;
  A.val = P.d_val; A.rhs = P.d_rhs;
  const int nsl = (P.n + 63) / 64;
  dim3 g((nsl + ROW_WAVES - 1) / ROW_WAVES), b(64 * ROW_WAVES);
  const int mr = P.setup.max_row;
  if (mr > 32) {
    set_error("two-pass assembly supports rows of at most 32 entries; use the atomic path");
    return -1;
  }
  const bool tet = P.el_type == kTetra;
  if (precision == 32) {
    const float* er = (const float*)P.d_erow;
    const float* ef = er + (size_t)P.nel * 36;
    if (tet)
      hipLaunchKernelGGL((k_gather_rows<4, 32, float>), g, b, 0, g_stream, A, er, ef, P.d_el_offT);
    else if (mr <= 24)
      hipLaunchKernelGGL((k_gather_rows<6, 24, float>), g, b, 0, g_stream, A, er, ef, P.d_el_offT);
    else
      hipLaunchKernelGGL((k_gather_rows<6, 32, float>), g, b, 0, g_stream, A, er, ef, P.d_el_offT);
  } else {
    const double* er = (const double*)P.d_erow;
    const double* ef = er + (size_t)P.nel * 36;
    if (tet)
      hipLaunchKernelGGL((k_gather_rows<4, 32, double>), g, b, 0, g_stream, A, er, ef, P.d_el_offT);
    else if (mr <= 24)
      hipLaunchKernelGGL((k_gather_rows<6, 24, double>), g, b, 0, g_stream, A, er, ef, P.d_el_offT);
    else
      hipLaunchKernelGGL((k_gather_rows<6, 32, double>), g, b, 0, g_stream, A, er, ef, P.d_el_offT);
  }
  MF_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------
// Assembly of many problems with one launch per kernel (no reference
// counterpart: the reference assembles one system at a time).  Problems of
// one element type, term set and row bound share a launch: a block table
// maps every block to (problem, block of that problem) and the problems'
// launch arguments sit in a device array that the blocks read.  Each block
// runs exactly the code of the single-problem kernels, so every problem's
// matrix is bitwise the one its own launches give.
// ---------------------------------------------------------------------
struct GatherArgs {
  RowArgs A;
  const void* erow;
  const void* ef;
  const uint8_t* offT;
};

template <int NS, class Q, int TS, typename R>
__global__ void __launch_bounds__(INT_BLOCK) kb_integrate(const IntArgs* args, const int2* tab) {
  const int2 tb = tab[blockIdx.x];
  integrate_block<NS, Q, TS, R>(args[tb.x], tb.y);
}

template <int NS, int RMAX, typename T>
__global__ void __launch_bounds__(64 * ROW_WAVES) kb_gather_rows(const GatherArgs* args,
                                                                const int2* tab) {
  const int2 tb = tab[blockIdx.x];
  const GatherArgs& G = args[tb.x];
  gather_rows_block<NS, RMAX, T>(G.A, tb.y, (const T*)G.erow, (const T*)G.ef, G.offT);
}

template <int NS, class Q, int TS, int RMAX, typename R>
__global__ void __launch_bounds__(64 * ROW_WAVES) kb_assemble_rows(const RowArgs* args,
                                                                  const int2* tab) {
  const int2 tb = tab[blockIdx.x];
  assemble_rows_block<NS, Q, TS, RMAX, R>(args[tb.x], tb.y);
}

namespace {

void fill_row_args(Problem& P, RowArgs& A) {
  fill_int_args(P, A.I);
  A.n = P.n;
  A.inc_ptr = P.d_inc_ptr; A.inc_el = P.d_inc_el; A.inc_loc = P.d_inc_loc;
  A.finc_ptr = P.d_finc_ptr; A.finc_f = P.d_finc_f; A.finc_loc = P.d_finc_loc;
  A.bf_elem = P.d_bf_elem; A.bf_side = P.d_bf_side; A.bf_g = P.d_bf_g; A.bf_type = P.d_bf_type;
  A.val = P.d_val; A.rhs = P.d_rhs;
}

// one launch of a batched kernel: its template key, the argument records
// (as bytes) and the block table
struct ManyLaunch {
  int kind;   // 0 row-wise, 1 element rows (pass 1), 2 row gathers (pass 2)
  int ns, ts, rmax;
  std::vector<char> args;
  std::vector<int2> tab;
  size_t args_off = 0, tab_off = 0;
};

// device copy of every launch's records and tables; the host image stays
// alive until the next call, which first waits for the stream
struct ManyBuf {
  void* d = nullptr;
  size_t cap = 0;
  std::vector<char> h;
};
ManyBuf g_many;

template <class T>
void add_record(ManyLaunch& L, const T& rec, int blocks) {
  const int idx = (int)(L.args.size() / sizeof(T));
  L.args.insert(L.args.end(), (const char*)&rec, (const char*)&rec + sizeof(T));
  for (int b = 0; b < blocks; b++) L.tab.push_back(make_int2(idx, b));
}

ManyLaunch& launch_for(std::vector<ManyLaunch>& ls, int kind, int ns, int ts, int rmax) {
  for (auto& L : ls)
    if (L.kind == kind && L.ns == ns && L.ts == ts && L.rmax == rmax) return L;
  ls.push_back(ManyLaunch{kind, ns, ts, rmax, {}, {}});
  return ls.back();
}

template <int NS, class Q, typename R>
void kb_int_ts(int ts, dim3 g, const IntArgs* a, const int2* t) {
  if (ts == TS_DIFF)
    hipLaunchKernelGGL((kb_integrate<NS, Q, TS_DIFF, R>), g, dim3(INT_BLOCK), 0, g_stream, a, t);
  else if (ts == TS_DIFF_CONV)
    hipLaunchKernelGGL((kb_integrate<NS, Q, TS_DIFF_CONV, R>), g, dim3(INT_BLOCK), 0, g_stream, a, t);
  else
    hipLaunchKernelGGL((kb_integrate<NS, Q, TS_ALL, R>), g, dim3(INT_BLOCK), 0, g_stream, a, t);
}

template <int NS, class Q, int RMAX, typename R>
void kb_rows_ts(int ts, dim3 g, const RowArgs* a, const int2* t) {
  const dim3 b(64 * ROW_WAVES);
  if (ts == TS_DIFF)
    hipLaunchKernelGGL((kb_assemble_rows<NS, Q, TS_DIFF, RMAX, R>), g, b, 0, g_stream, a, t);
  else if (ts == TS_DIFF_CONV)
    hipLaunchKernelGGL((kb_assemble_rows<NS, Q, TS_DIFF_CONV, RMAX, R>), g, b, 0, g_stream, a, t);
  else
    hipLaunchKernelGGL((kb_assemble_rows<NS, Q, TS_ALL, RMAX, R>), g, b, 0, g_stream, a, t);
}

template <typename R>
void run_many(const ManyLaunch& L, const char* dbase) {
  const dim3 g((unsigned)L.tab.size());
  const int2* t = (const int2*)(dbase + L.tab_off);
  const dim3 b(64 * ROW_WAVES);
  if (L.kind == 1) {
    const IntArgs* a = (const IntArgs*)(dbase + L.args_off);
    if (L.ns == 4) kb_int_ts<4, TetQ, R>(L.ts, g, a, t);
    else kb_int_ts<6, PrismQ, R>(L.ts, g, a, t);
  } else if (L.kind == 2) {
    const GatherArgs* a = (const GatherArgs*)(dbase + L.args_off);
    if (L.ns == 4) hipLaunchKernelGGL((kb_gather_rows<4, 32, R>), g, b, 0, g_stream, a, t);
    else if (L.rmax == 24) hipLaunchKernelGGL((kb_gather_rows<6, 24, R>), g, b, 0, g_stream, a, t);
    else hipLaunchKernelGGL((kb_gather_rows<6, 32, R>), g, b, 0, g_stream, a, t);
  } else {
    const RowArgs* a = (const RowArgs*)(dbase + L.args_off);
    if (L.ns == 4) {
      if (L.rmax == 16) kb_rows_ts<4, TetQ, 16, R>(L.ts, g, a, t);
      else kb_rows_ts<4, TetQ, 32, R>(L.ts, g, a, t);
    } else {
      if (L.rmax == 24) kb_rows_ts<6, PrismQ, 24, R>(L.ts, g, a, t);
      else kb_rows_ts<6, PrismQ, 32, R>(L.ts, g, a, t);
    }
  }
}

}  // namespace

int launch_assemble_many(const std::vector<Problem*>& ps, int precision) {
  MF_HIP(hipStreamSynchronize(g_stream));
  const size_t esz = precision == 32 ? sizeof(float) : sizeof(double);
  std::vector<ManyLaunch> ls;
  ls.reserve(8);
  for (Problem* P : ps) {
    const int mr = P->setup.max_row;
    const bool tet = P->el_type == kTetra;
    const int ns = tet ? 4 : 6;
    if (P->nel <= 0) continue;
    if (mr > 32 || (P->assembly_mode != 0 && P->assembly_mode != 2)) {
      set_error("launch_assemble_many: row-wise or two-pass problems with rows of at most 32 entries");
      return -1;
    }
    const int nsl = (P->n + 63) / 64;
    const int row_blocks = (nsl + ROW_WAVES - 1) / ROW_WAVES;
    if (P->assembly_mode == 0) {
      const int rmax = tet ? (mr <= 16 ? 16 : 32) : (mr <= 24 ? 24 : 32);
      RowArgs A;
      fill_row_args(*P, A);
      add_record(launch_for(ls, 0, ns, P->term_set, rmax), A, row_blocks);
      continue;
    }
    if (!P->d_erow || P->erow_bytes < (size_t)P->nel * 42 * esz) {
      if (P->d_erow) (void)hipFree(P->d_erow);
      P->d_erow = nullptr;
      MF_HIP(hipMalloc(&P->d_erow, (size_t)P->nel * 42 * esz));
      P->erow_bytes = (size_t)P->nel * 42 * esz;
    }
    IntArgs I;
    fill_int_args(*P, I);
    I.write_elements = 2;
    I.erow = P->d_erow;
    I.ef = (char*)P->d_erow + (size_t)P->nel * 36 * esz;
    I.emat = nullptr;
    I.e0 = 0;
    I.count = P->nel;
    constexpr int EPB6 = (64 / 6) * (INT_BLOCK / 64), EPB4 = (64 / 4) * (INT_BLOCK / 64);
    const int epb = tet ? EPB4 : EPB6;
    add_record(launch_for(ls, 1, ns, P->term_set, 0), I, (P->nel + epb - 1) / epb);
    GatherArgs G;
    fill_row_args(*P, G.A);
    G.A.I = I;
    G.erow = I.erow;
    G.ef = I.ef;
    G.offT = P->d_el_offT;
    add_record(launch_for(ls, 2, ns, 0, tet ? 32 : (mr <= 24 ? 24 : 32)), G, row_blocks);
  }
  if (ls.empty()) return 0;
  // one host image: records (256-byte aligned) then tables
  size_t off = 0;
  for (auto& L : ls) {
    L.args_off = off;
    off = (off + L.args.size() + 255) & ~(size_t)255;
    L.tab_off = off;
    off = (off + L.tab.size() * sizeof(int2) + 255) & ~(size_t)255;
  }
  g_many.h.assign(off, 0);
  for (auto& L : ls) {
    std::memcpy(g_many.h.data() + L.args_off, L.args.data(), L.args.size());
    std::memcpy(g_many.h.data() + L.tab_off, L.tab.data(), L.tab.size() * sizeof(int2));
  }
  if (off > g_many.cap) {
    if (g_many.d) (void)hipFree(g_many.d);
    g_many.d = nullptr;
    g_many.cap = 0;
    MF_HIP(hipMalloc(&g_many.d, off));
    g_many.cap = off;
  }
  MF_HIP(hipMemcpyAsync(g_many.d, g_many.h.data(), off, hipMemcpyHostToDevice, g_stream));
  const char* dbase = (const char*)g_many.d;
  // element rows before their gathers
  for (int kind : {0, 1, 2})
    for (auto& L : ls)
      if (L.kind == kind) {
        if (precision == 32) run_many<float>(L, dbase);
        else run_many<double>(L, dbase);
      }
  MF_HIP(hipGetLastError());
  return 0;
}

}  // namespace mf
