// Discontinuous Galerkin prism approximation (apd_dg_prism) on the device:
// element integrals of the convection-diffusion-reaction weak form for the
// tensor-product monomial base of degree pdeg = 100 pz + px (1 <= px, pz <= 3;
// 6 / 18 / 40 shape functions for p = 1 / 2 / 3).
//
// Formulation.  With u_j = (phi_j, grad phi_j) at a Gauss point and the 4x4
// coefficient matrix M (M00 = m + c, M0k = b_k, M(1+c)(1+r) = a_rc), the
// element matrix of apr_num_int_el (aps_dg_prism_intf.c:1239-1728) is
//   K[j][i] = sum_g vol_g u_j^T M u_i.
// Writing u = T uhat, uhat = (phi, d/deta phi) the reference-element values
// (the same for every element) and T = diag(1, J^-T), gives
//   K = Uhat^T blockdiag(Mhat_g) Uhat,   Mhat_g = vol_g T^T M T,
// a product whose inner dimension is 4 G: one K-step of four of the FP64
// matrix instruction v_mfma_f64_16x16x4 is exactly one Gauss point.  The
// reference table Uhat lives in LDS; each wave builds the 16 + 4 numbers of
// Mhat_g and fhat_g for its element (one lane per Gauss point), then runs
// NT x NT 16x16 output tiles through the G steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <memory>
#include <vector>

#include "../../include/modfem_hip.h"
#include "mf_device.h"

namespace mf {

namespace {

// ---------------------------------------------------------------------
// quadrature (host, once per degree): apr_set_quadr_3D for the tensor base
// (aps_dg_util.c:1415-1496) -- degree 2 px on the triangle, 2 pz on the
// line, point (l2, l3, z), weight 1/2 w_tri w_line
// ---------------------------------------------------------------------
void gauss_legendre(int n, double* x, double* w) {
  if (n == 1) { x[0] = 0.0; w[0] = 2.0; return; }
  for (int i = 0; i < n; i++) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0, p1 = 0.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0;
      p1 = z;
      for (int k = 2; k <= n; k++) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1; p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-17) break;
    }
    double p0 = 1.0;
    p1 = z;
    for (int k = 2; k <= n; k++) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1; p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// symmetric triangle rules of degree 2, 4, 6 (Dunavant 1985: 3, 6, 12
// points), barycentric points, weights summing to one, listed orbit by
// orbit in the order the reference's table holds them (aps_gauss_util.c)
int triangle_rule(int order, std::vector<double>& l, std::vector<double>& w) {
  auto orb3 = [&](double a, double b, double ww) {
    const double p[3][3] = {{a, b, b}, {b, b, a}, {b, a, b}};
    for (auto& q : p) { l.insert(l.end(), q, q + 3); w.push_back(ww); }
  };
  if (order == 2) {
    orb3(0.666666666666667, 0.166666666666667, 0.333333333333333);
  } else if (order == 4) {
    orb3(0.108103018168070, 0.445948490915965, 0.223381589678011);
    orb3(0.816847572980459, 0.091576213509771, 0.109951743655322);
  } else if (order == 6) {
    orb3(0.501426509658179, 0.249286745170910, 0.116786275726379);
    orb3(0.873821971016996, 0.063089014491502, 0.050844906370207);
    const double a = 0.053145049844817, b = 0.310352451033784, c = 0.636502499121399;
    const double p[6][3] = {{a, b, c}, {b, c, a}, {c, a, b}, {c, b, a}, {a, c, b}, {b, a, c}};
    for (auto& q : p) { l.insert(l.end(), q, q + 3); w.push_back(0.082851075618374); }
  } else {
    return -1;
  }
  return (int)w.size();
}

// [ng][4]: eta0, eta1, eta2, weight
int volume_rule(int px, int pz, std::vector<double>& q) {
  std::vector<double> l, wt;
  if (triangle_rule(2 * px, l, wt) < 0) return -1;
  const int n2 = (int)wt.size(), nz = pz + 1;
  double xz[8], wz[8];
  gauss_legendre(nz, xz, wz);
  q.clear();
  for (int ki = 0; ki < nz; ki++)
    for (int kj = 0; kj < n2; kj++) {
      q.push_back(l[3 * kj + 1]);
      q.push_back(l[3 * kj + 2]);
      q.push_back(xz[ki]);
      q.push_back(0.5 * wt[kj] * wz[ki]);
    }
  return nz * n2;
}

constexpr int tri_points(int px) { return px == 1 ? 3 : px == 2 ? 6 : 12; }
constexpr int num_shape(int px, int pz) { return (px + 1) * (px + 2) / 2 * (pz + 1); }

}  // namespace

// ---------------------------------------------------------------------
// shape functions: the tensor-product monomial base of apr_shape_fun_3D
// (aps_dg_util.c:408-992) -- 1, e0, e1, z, z e0, z e1, then the 2D
// functions s_i (i >= 3) times z^0 and z^1, then every s_i times z^j,
// j >= 2; s_i is the complete 2D monomial base of apr_shape_fun_2D
// (1, e0, e1, e0 e1, e0^i e1^j for j <= 1 < i, then j >= 2)
// ---------------------------------------------------------------------
__device__ int dg_shape(int px, int pz, const double* eta, double* phi, double* dx,
                        double* dy, double* dz) {
  double s[4], ds[4], t[4], dt[4], u[4], du[4];
  auto mono = [](int p, double e, double* f, double* df) {
    f[0] = 1.0; df[0] = 0.0;
    if (p > 0) { f[1] = e; df[1] = 1.0; }
    for (int i = 2; i <= p; i++) { f[i] = e * f[i - 1]; df[i] = e * df[i - 1]; }
    for (int i = 2; i <= p; i++) df[i] *= i;
  };
  mono(px, eta[0], s, ds);
  mono(px, eta[1], t, dt);
  mono(pz, eta[2], u, du);
  double f[10], fx[10], fy[10];
  f[0] = 1.0; f[1] = eta[0]; f[2] = eta[1];
  fx[0] = 0.0; fx[1] = 1.0; fx[2] = 0.0;
  fy[0] = 0.0; fy[1] = 0.0; fy[2] = 1.0;
  int n2 = 3;
  if (px > 1) {
    f[3] = eta[0] * eta[1]; fx[3] = eta[1]; fy[3] = eta[0];
    n2 = 4;
    for (int j = 0; j < 2; j++)
      for (int i = 2; i <= px - j; i++) {
        f[n2] = s[i] * t[j]; fx[n2] = ds[i] * t[j]; fy[n2] = s[i] * dt[j]; n2++;
      }
    for (int j = 2; j <= px; j++)
      for (int i = 0; i <= px - j; i++) {
        f[n2] = s[i] * t[j]; fx[n2] = ds[i] * t[j]; fy[n2] = s[i] * dt[j]; n2++;
      }
  }
  phi[0] = 1.0; phi[1] = eta[0]; phi[2] = eta[1]; phi[3] = eta[2];
  phi[4] = eta[2] * eta[0]; phi[5] = eta[2] * eta[1];
  dx[0] = 0.0; dx[1] = 1.0; dx[2] = 0.0; dx[3] = 0.0; dx[4] = eta[2]; dx[5] = 0.0;
  dy[0] = 0.0; dy[1] = 0.0; dy[2] = 1.0; dy[3] = 0.0; dy[4] = 0.0; dy[5] = eta[2];
  dz[0] = 0.0; dz[1] = 0.0; dz[2] = 0.0; dz[3] = 1.0; dz[4] = eta[0]; dz[5] = eta[1];
  int k = 6;
  for (int j = 0; j < 2; j++)
    for (int i = 3; i < n2; i++) {
      phi[k] = f[i] * u[j]; dx[k] = fx[i] * u[j]; dy[k] = fy[i] * u[j]; dz[k] = f[i] * du[j];
      k++;
    }
  for (int j = 2; j <= pz; j++)
    for (int i = 0; i < n2; i++) {
      phi[k] = f[i] * u[j]; dx[k] = fx[i] * u[j]; dy[k] = fy[i] * u[j]; dz[k] = f[i] * du[j];
      k++;
    }
  return k;
}

// reference table U[g][q][i] (q: phi, d/deta0, d/deta1, d/deta2), i padded
// with zeros to nsp
__global__ void k_dg_table(int px, int pz, int ng, int nsp, const double* qv, double* U) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  double phi[40], dx[40], dy[40], dz[40];
  const int ns = dg_shape(px, pz, qv + 4 * g, phi, dx, dy, dz);
  double* Ug = U + (size_t)g * 4 * nsp;
  for (int i = 0; i < ns; i++) {
    Ug[i] = phi[i]; Ug[nsp + i] = dx[i]; Ug[2 * nsp + i] = dy[i]; Ug[3 * nsp + i] = dz[i];
  }
  for (int i = ns; i < nsp; i++) Ug[i] = Ug[nsp + i] = Ug[2 * nsp + i] = Ug[3 * nsp + i] = 0.0;
}

// which terms of the weak form are present (apr_num_int_el branches)
struct DgMask {
  int m, diff;  // diff: 0 none, 1 the xy block (ind[2] without ind[10]), 2 full
  int conv;     // 0 none, 1 bx, by (ind[11] without ind[13]), 2 full
  int reac, q, s;  // q: 0 none, 1 qx, qy, 2 full
};

struct DgElemArgs {
  int nel, e0, count;
  const double* nc;   // [18][nel] node coordinates, SoA
  const double* qv;   // [ng][4]
  const double* U;    // [ng][4][nsp]
  int coef_model, src_model;
  DgMask mask;
  double cc[NCOEF];
  const double* coef;  // per-element SoA per coef_model
  double* stiff;       // [count][ns*ns], column-major per element
  double* rhs;         // [count][ns]
};

__device__ __forceinline__ void dg_coefs(const DgElemArgs& A, int e, double* a, double* b,
                                         double& mc, double& q0, double& q1, double& q2,
                                         double& s) {
#pragma unroll
  for (int k = 0; k < 9; k++) a[k] = A.cc[C_A + k];
  b[0] = A.cc[C_B]; b[1] = A.cc[C_B + 1]; b[2] = A.cc[C_B + 2];
  double m = A.cc[C_M], c = A.cc[C_C];
  q0 = A.cc[C_Q]; q1 = A.cc[C_Q + 1]; q2 = A.cc[C_Q + 2];
  s = A.cc[C_S];
  const size_t N = A.nel;
  if (A.coef_model == CM_KAPPA || A.coef_model == CM_KAPPA_B || A.coef_model == CM_KAPPA_S) {
    const double kap = A.coef[e];
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = (k % 4 == 0) ? kap : 0.0;
    if (A.coef_model == CM_KAPPA_B) {
      b[0] = A.coef[N + e]; b[1] = A.coef[2 * N + e]; b[2] = A.coef[3 * N + e];
    }
    if (A.coef_model == CM_KAPPA_S) s = A.coef[N + e];
  } else if (A.coef_model == CM_FULL) {
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = A.coef[(C_A + k) * N + e];
    b[0] = A.coef[C_B * N + e]; b[1] = A.coef[(C_B + 1) * N + e]; b[2] = A.coef[(C_B + 2) * N + e];
    m = A.coef[C_M * N + e]; c = A.coef[C_C * N + e];
    q0 = A.coef[C_Q * N + e]; q1 = A.coef[(C_Q + 1) * N + e]; q2 = A.coef[(C_Q + 2) * N + e];
    s = A.coef[C_S * N + e];
  }
  const DgMask& K = A.mask;
  if (K.diff == 1) { a[2] = a[5] = a[6] = a[7] = a[8] = 0.0; }
  if (K.diff == 0) {
#pragma unroll
    for (int k = 0; k < 9; k++) a[k] = 0.0;
  }
  if (K.conv < 2) b[2] = 0.0;
  if (K.conv == 0) b[0] = b[1] = 0.0;
  mc = (K.m ? m : 0.0) + (K.reac ? c : 0.0);
  if (K.q < 2) q2 = 0.0;
  if (K.q == 0) q0 = q1 = 0.0;
  if (!K.s) s = 0.0;
}

// linear prism geometry at eta: Jacobian dx_r/deta_c (J[3r+c]), its
// inverse Ji[3c+r] = deta_c/dx_r by cofactors, and the point x
__device__ __forceinline__ double prism_geometry(const double* xn, const double* eta,
                                                 double* Ji, double* x) {
  const double e0 = eta[0], e1 = eta[1], z = eta[2];
  const double gp[6] = {(1.0 - e0 - e1) * (1.0 - z) / 2.0, e0 * (1.0 - z) / 2.0,
                        e1 * (1.0 - z) / 2.0, (1.0 - e0 - e1) * (1.0 + z) / 2.0,
                        e0 * (1.0 + z) / 2.0, e1 * (1.0 + z) / 2.0};
  const double gx[6] = {-(1.0 - z) / 2.0, (1.0 - z) / 2.0, 0.0,
                        -(1.0 + z) / 2.0, (1.0 + z) / 2.0, 0.0};
  const double gy[6] = {-(1.0 - z) / 2.0, 0.0, (1.0 - z) / 2.0,
                        -(1.0 + z) / 2.0, 0.0, (1.0 + z) / 2.0};
  const double gz[6] = {-(1.0 - e0 - e1) / 2.0, -e0 / 2.0, -e1 / 2.0,
                        (1.0 - e0 - e1) / 2.0, e0 / 2.0, e1 / 2.0};
  double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  x[0] = x[1] = x[2] = 0.0;
#pragma unroll
  for (int i = 0; i < 6; i++) {
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const double c = xn[3 * i + r];
      m[3 * r] += c * gx[i]; m[3 * r + 1] += c * gy[i]; m[3 * r + 2] += c * gz[i];
      x[r] += c * gp[i];
    }
  }
  const double det = m[0] * (m[4] * m[8] - m[7] * m[5]) + m[3] * (m[7] * m[2] - m[1] * m[8]) +
                     m[6] * (m[1] * m[5] - m[4] * m[2]);
  const double r = 1.0 / det;
  Ji[0] = (m[4] * m[8] - m[7] * m[5]) * r;
  Ji[3] = (-m[3] * m[8] + m[6] * m[5]) * r;
  Ji[6] = (m[3] * m[7] - m[6] * m[4]) * r;
  Ji[1] = (m[7] * m[2] - m[1] * m[8]) * r;
  Ji[4] = (m[0] * m[8] - m[6] * m[2]) * r;
  Ji[7] = (-m[0] * m[7] + m[6] * m[1]) * r;
  Ji[2] = (m[1] * m[5] - m[4] * m[2]) * r;
  Ji[5] = (-m[0] * m[5] + m[3] * m[2]) * r;
  Ji[8] = (m[0] * m[4] - m[3] * m[1]) * r;
  return det;
}

typedef double dbl4 __attribute__((ext_vector_type(4)));

// 8 waves per block for 20 or more shape functions, else 4
constexpr int dg_block_threads(int px, int pz) { return num_shape(px, pz) > 18 ? 512 : 256; }

template <int PX, int PZ>
struct DgShape {
  static constexpr int NS = num_shape(PX, PZ);
  static constexpr int NG = tri_points(PX) * (PZ + 1);
  static constexpr int NT = (NS + 15) / 16;
  static constexpr int NSP = NT * 16;
  static constexpr int WREC = NG * 20 + 24;  // per-wave LDS doubles: Mhat, fhat, nodes
  static constexpr int WPB = dg_block_threads(PX, PZ) / 64;
  static constexpr size_t lds_bytes() {
    return sizeof(double) * ((size_t)NG * 4 * NSP + (size_t)WPB * WREC);
  }
};

template <int PX, int PZ>
__global__ void __launch_bounds__(dg_block_threads(PX, PZ)) k_dg_elem(DgElemArgs A) {
  using S = DgShape<PX, PZ>;
  constexpr int NS = S::NS, NG = S::NG, NT = S::NT, NSP = S::NSP, WPB = S::WPB;
  extern __shared__ double lds[];
  double* sU = lds;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double* sM = lds + NG * 4 * NSP + wave * S::WREC;  // [NG][4][4]
  double* sF = sM + NG * 16;                         // [NG][4]
  double* sX = sF + NG * 4;                          // [6][3]
  for (int k = threadIdx.x; k < NG * 4 * NSP; k += 64 * WPB) sU[k] = A.U[k];
  const int p = lane >> 4, c16 = lane & 15;
  for (int base = blockIdx.x * WPB; base < A.count; base += gridDim.x * WPB) {
    const int el = base + wave;
    const bool valid = el < A.count;
    const int e = A.e0 + (valid ? el : 0);
    if (valid && lane < 18) sX[lane] = A.nc[(size_t)lane * A.nel + e];
    __syncthreads();
    if (valid && lane < NG) {
      double a[9], b[3], mc, q0, q1, q2, s;
      dg_coefs(A, e, a, b, mc, q0, q1, q2, s);
      const double* qg = A.qv + 4 * lane;
      double Ji[9], x[3];
      const double det = prism_geometry(sX, qg, Ji, x);
      const double vol = fabs(det) * qg[3];
      if (A.src_model != SRC_CONST && A.mask.s) {
        double uu, lapl;
        exact_sol(A.src_model, x, &uu, &lapl);
        s = -lapl;
      }
      double* Mg = sM + lane * 16;
      // Mhat = vol T^T M T with T = diag(1, Ji^T): (d phi/dx_c) = sum_k Ji[3k+c] d phi/deta_k
      Mg[0] = vol * mc;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        Mg[1 + k] = vol * (b[0] * Ji[3 * k] + b[1] * Ji[3 * k + 1] + b[2] * Ji[3 * k + 2]);
        Mg[4 * (1 + k)] = 0.0;
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        // row k: t_r = sum_c Ji[3k+c] a[3r+c]
        double t[3];
#pragma unroll
        for (int r = 0; r < 3; r++)
          t[r] = Ji[3 * k] * a[3 * r] + Ji[3 * k + 1] * a[3 * r + 1] + Ji[3 * k + 2] * a[3 * r + 2];
#pragma unroll
        for (int l = 0; l < 3; l++)
          Mg[4 * (1 + k) + 1 + l] =
              vol * (t[0] * Ji[3 * l] + t[1] * Ji[3 * l + 1] + t[2] * Ji[3 * l + 2]);
      }
      double* Fg = sF + lane * 4;
      Fg[0] = vol * s;
#pragma unroll
      for (int k = 0; k < 3; k++)
        Fg[1 + k] = vol * (q0 * Ji[3 * k] + q1 * Ji[3 * k + 1] + q2 * Ji[3 * k + 2]);
    }
    __syncthreads();
    if (valid) {
      dbl4 acc[NT][NT];
#pragma unroll
      for (int jt = 0; jt < NT; jt++)
#pragma unroll
        for (int it = 0; it < NT; it++) acc[jt][it] = dbl4{0.0, 0.0, 0.0, 0.0};
      double f = 0.0;
      for (int g = 0; g < NG; g++) {
        const double* Ug = sU + g * 4 * NSP;
        const double* Mp = sM + g * 16 + 4 * p;
        const double m0 = Mp[0], m1 = Mp[1], m2 = Mp[2], m3 = Mp[3];
        double av[NT], bv[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
          const int i = t * 16 + c16;
          av[t] = Ug[p * NSP + i];
          bv[t] = m0 * Ug[i] + m1 * Ug[NSP + i] + m2 * Ug[2 * NSP + i] + m3 * Ug[3 * NSP + i];
        }
#pragma unroll
        for (int jt = 0; jt < NT; jt++)
#pragma unroll
          for (int it = 0; it < NT; it++)
            acc[jt][it] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[jt], bv[it], acc[jt][it], 0, 0, 0);
        if (lane < NS) {
          const double* Fg = sF + g * 4;
          f += Fg[0] * Ug[lane] + Fg[1] * Ug[NSP + lane] + Fg[2] * Ug[2 * NSP + lane] +
               Fg[3] * Ug[3 * NSP + lane];
        }
      }
      // v_mfma_f64_16x16x4: A row m comes from lanes m, m+16, m+32, m+48
      // (k = lane / 16), B column n from lanes n + 16 k, and accumulator
      // register r of lane l holds D[4 r + l / 16][l % 16] (measured lane
      // layout); so lane holds K[j = 16 jt + 4 r + p][i = 16 it + c16].  The
      // element matrix is written column-major by trial function j:
      // stiff[j*NS + i]
      double* out = A.stiff + (size_t)el * NS * NS;
#pragma unroll
      for (int jt = 0; jt < NT; jt++)
#pragma unroll
        for (int it = 0; it < NT; it++)
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int j = 16 * jt + 4 * r + p, i = 16 * it + c16;
            if (j < NS && i < NS) out[j * NS + i] = acc[jt][it][r];
          }
      if (lane < NS) A.rhs[(size_t)el * NS + lane] = f;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------
// DG problem objects
// ---------------------------------------------------------------------
struct DgProblem {
  int mesh = 0, nel = 0, px = 1, pz = 1, ns = 0, ng = 0, nsp = 0;
  double* d_nc = nullptr;    // [18][nel]
  double* d_qv = nullptr;    // [ng][4]
  double* d_U = nullptr;     // [ng][4][nsp]
  DgMask mask{0, 2, 0, 0, 0, 1};
  int coef_model = CM_UNIFORM, src_model = SRC_CONST;
  double cc[NCOEF] = {0};
  double* d_coef = nullptr;
  double* d_stiff = nullptr;  // device-resident element matrices (mf_dg_integrate)
  double* d_rhs = nullptr;
  int ncu = 0;
};

namespace {
std::map<int, std::unique_ptr<DgProblem>> g_dg;
int g_next_dg = 1;

DgProblem* get_dg(int id) {
  auto it = g_dg.find(id);
  if (it == g_dg.end()) {
    set_error("unknown DG problem handle " + std::to_string(id));
    return nullptr;
  }
  return it->second.get();
}

template <int PX, int PZ>
int launch_elem_t(DgProblem& D, int e0, int count, double* d_stiff, double* d_rhs) {
  using S = DgShape<PX, PZ>;
  DgElemArgs A;
  A.nel = D.nel; A.e0 = e0; A.count = count;
  A.nc = D.d_nc; A.qv = D.d_qv; A.U = D.d_U;
  A.coef_model = D.coef_model; A.src_model = D.src_model; A.mask = D.mask;
  for (int k = 0; k < NCOEF; k++) A.cc[k] = D.cc[k];
  A.coef = D.d_coef; A.stiff = d_stiff; A.rhs = d_rhs;
  const size_t lds = S::lds_bytes();
  static bool attr = false;
  if (!attr) {
    MF_HIP(hipFuncSetAttribute((const void*)k_dg_elem<PX, PZ>,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    attr = true;
  }
  const int per_cu = std::max(1, (int)((160 * 1024) / lds));
  const int blocks = std::min((count + S::WPB - 1) / S::WPB, D.ncu * std::min(per_cu, 4));
  hipLaunchKernelGGL((k_dg_elem<PX, PZ>), dim3(blocks), dim3(64 * S::WPB), lds, stream(), A);
  MF_HIP(hipGetLastError());
  return 0;
}

int launch_elem(DgProblem& D, int e0, int count, double* d_stiff, double* d_rhs) {
  if (count <= 0) return 0;
#define MF_DG_CASE(X, Z) \
  if (D.px == X && D.pz == Z) return launch_elem_t<X, Z>(D, e0, count, d_stiff, d_rhs);
  MF_DG_CASE(1, 1) MF_DG_CASE(2, 1) MF_DG_CASE(3, 1)
  MF_DG_CASE(1, 2) MF_DG_CASE(2, 2) MF_DG_CASE(3, 2)
  MF_DG_CASE(1, 3) MF_DG_CASE(2, 3) MF_DG_CASE(3, 3)
#undef MF_DG_CASE
  set_error("DG degree out of range");
  return -1;
}

void free_dg(DgProblem& D) {
  void* ptrs[] = {D.d_nc, D.d_qv, D.d_U, D.d_coef, D.d_stiff, D.d_rhs};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}
}  // namespace

}  // namespace mf

using namespace mf;

extern "C" {

int mf_dg_create(int mesh, int pdeg) {
  Mesh* m = mesh_by_id(mesh);
  if (!m) return -3;
  if (!stream()) {
    set_error("mf_init has not been called");
    return -1;
  }
  const int px = pdeg % 100, pz = pdeg / 100;
  if (px < 1 || px > 3 || pz < 1 || pz > 3) {
    set_error("mf_dg_create: pdeg must be 100*pz + px with 1 <= px, pz <= 3");
    return -1;
  }
  for (int e = 0; e < m->nel; e++)
    if (m->type[e] != kPrism) {
      set_error("mf_dg_create: the DG approximation is defined on prisms only");
      return -1;
    }
  auto D = std::make_unique<DgProblem>();
  D->mesh = mesh; D->nel = m->nel; D->px = px; D->pz = pz;
  D->ns = num_shape(px, pz);
  D->nsp = (D->ns + 15) / 16 * 16;
  std::vector<double> qv;
  D->ng = volume_rule(px, pz, qv);
  int dev = 0;
  MF_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  MF_HIP(hipGetDeviceProperties(&prop, dev));
  D->ncu = prop.multiProcessorCount;
  const size_t nel = m->nel;
  std::vector<double> nc(18 * nel);
  for (size_t e = 0; e < nel; e++)
    for (int i = 0; i < 6; i++)
      for (int r = 0; r < 3; r++)
        nc[(size_t)(3 * i + r) * nel + e] = m->coor[3 * (size_t)m->nodes[6 * e + i] + r];
  MF_HIP(hipMalloc(&D->d_nc, nc.size() * sizeof(double)));
  MF_HIP(hipMalloc(&D->d_qv, qv.size() * sizeof(double)));
  MF_HIP(hipMalloc(&D->d_U, (size_t)D->ng * 4 * D->nsp * sizeof(double)));
  MF_HIP(hipMemcpyAsync(D->d_nc, nc.data(), nc.size() * sizeof(double), hipMemcpyHostToDevice,
                        stream()));
  MF_HIP(hipMemcpyAsync(D->d_qv, qv.data(), qv.size() * sizeof(double), hipMemcpyHostToDevice,
                        stream()));
  hipLaunchKernelGGL(k_dg_table, dim3(1), dim3(64), 0, stream(), px, pz, D->ng, D->nsp,
                     D->d_qv, D->d_U);
  MF_HIP(hipGetLastError());
  MF_HIP(hipStreamSynchronize(stream()));
  const int id = g_next_dg++;
  g_dg[id] = std::move(D);
  return id;
}

int mf_dg_sizes(int did, int* nel, int* ns, int* ng) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (nel) *nel = D->nel;
  if (ns) *ns = D->ns;
  if (ng) *ng = D->ng;
  return 0;
}

int mf_dg_set_coefficients(int did, const int* ind, int coef_model, const double* coef_const,
                           const double* coef_el, int src_model) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!ind || coef_model < CM_UNIFORM || coef_model > CM_KAPPA_S || src_model < SRC_CONST ||
      src_model > SRC_DIFF_IN_BOX || (coef_model != CM_UNIFORM && !coef_el)) {
    set_error("mf_dg_set_coefficients: invalid arguments");
    return -1;
  }
  if (ind[18]) {
    set_error("the obsolete l (time rhs) coefficient is not supported");
    return -1;
  }
  DgMask M;
  M.m = ind[1] != 0;
  M.diff = ind[10] ? 2 : (ind[2] ? 1 : 0);
  M.conv = ind[13] ? 2 : (ind[11] ? 1 : 0);
  M.reac = ind[17] != 0;
  M.q = ind[21] ? 2 : (ind[19] ? 1 : 0);
  M.s = ind[22] != 0;
  D->mask = M;
  for (int k = 0; k < NCOEF; k++) D->cc[k] = coef_const ? coef_const[k] : 0.0;
  D->coef_model = coef_model;
  D->src_model = src_model;
  if (D->d_coef) {
    (void)hipFree(D->d_coef);
    D->d_coef = nullptr;
  }
  const size_t per = coef_model == CM_KAPPA ? 1
                     : coef_model == CM_KAPPA_B ? 4
                     : coef_model == CM_KAPPA_S ? 2
                     : coef_model == CM_FULL ? NCOEF : 0;
  if (per) {
    const size_t bytes = per * D->nel * sizeof(double);
    MF_HIP(hipMalloc(&D->d_coef, bytes));
    MF_HIP(hipMemcpyAsync(D->d_coef, coef_el, bytes, hipMemcpyHostToDevice, stream()));
    MF_HIP(hipStreamSynchronize(stream()));
  }
  return 0;
}

int mf_dg_element_matrices(int did, int first, int count, double* stiff, double* rhs) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (first < 0 || count < 0 || first + (long long)count > D->nel || !stiff || !rhs) {
    set_error("mf_dg_element_matrices: invalid range or output");
    return -1;
  }
  const int ns = D->ns;
  const int chunk = std::min(count, 1 << 16);
  double *ds = nullptr, *dr = nullptr;
  MF_HIP(hipMalloc(&ds, (size_t)std::max(chunk, 1) * ns * ns * sizeof(double)));
  MF_HIP(hipMalloc(&dr, (size_t)std::max(chunk, 1) * ns * sizeof(double)));
  int rc = 0;
  for (int c0 = 0; c0 < count && rc == 0; c0 += chunk) {
    const int n = std::min(chunk, count - c0);
    rc = launch_elem(*D, first + c0, n, ds, dr);
    if (rc == 0 &&
        (hipMemcpyAsync(stiff + (size_t)c0 * ns * ns, ds, (size_t)n * ns * ns * sizeof(double),
                        hipMemcpyDeviceToHost, stream()) != hipSuccess ||
         hipMemcpyAsync(rhs + (size_t)c0 * ns, dr, (size_t)n * ns * sizeof(double),
                        hipMemcpyDeviceToHost, stream()) != hipSuccess ||
         hipStreamSynchronize(stream()) != hipSuccess)) {
      set_error("mf_dg_element_matrices: copy failed");
      rc = -2;
    }
  }
  (void)hipFree(ds);
  (void)hipFree(dr);
  return rc;
}

int mf_dg_integrate(int did) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  const size_t ns = D->ns;
  if (!D->d_stiff) {
    MF_HIP(hipMalloc(&D->d_stiff, (size_t)D->nel * ns * ns * sizeof(double)));
    MF_HIP(hipMalloc(&D->d_rhs, (size_t)D->nel * ns * sizeof(double)));
  }
  const int rc = launch_elem(*D, 0, D->nel, D->d_stiff, D->d_rhs);
  if (rc) return rc;
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_dg_get_element_matrices(int did, int first, int count, double* stiff, double* rhs) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->d_stiff || first < 0 || count < 0 || first + (long long)count > D->nel) {
    set_error("mf_dg_get_element_matrices: nothing integrated or invalid range");
    return -1;
  }
  const size_t ns = D->ns;
  if (stiff)
    MF_HIP(hipMemcpyAsync(stiff, D->d_stiff + first * ns * ns, count * ns * ns * sizeof(double),
                          hipMemcpyDeviceToHost, stream()));
  if (rhs)
    MF_HIP(hipMemcpyAsync(rhs, D->d_rhs + first * ns, count * ns * sizeof(double),
                          hipMemcpyDeviceToHost, stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_dg_free(int did) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  free_dg(*D);
  g_dg.erase(did);
  return 0;
}

}  // extern "C"
