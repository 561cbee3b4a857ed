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
#include <cstring>
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

// Tetrahedra (no reference counterpart: the reference's DG module is
// prism-only).  Complete polynomial base of degree p in the reference
// coordinates, xi^a eta^b zeta^c with a + b + c <= p, by total degree, then
// a descending, then b descending; reference vertices (0,0,0), (1,0,0),
// (0,1,0), (0,0,1); sides with the mesh module's vertex sets {0,1,2},
// {0,1,3}, {0,2,3}, {1,2,3}; the prism module's triangle rule of degree 2p
// on the sides.  Degree indices: 0..8 the prisms' tensor (px, pz), 9..11
// tets of degree 1..3, 12..14 prisms with the complete base of degree 1..3
// (APC_BASE_COMPLETE_DG: every monomial of total degree <= p, the volume and
// face rules of the tensor degree (p, p), aps_dg_util.c:1446-1451), 15..16
// the tensor degrees px = 1, 2 with pz = 0 (pdeg 1, 2: constant in z, one
// Gauss point in z, penalty degree min(px, pz) = 0).  Element kinds of the
// kernels' templates: 0 tensor prism, 1 tet, 2 complete-base prism.
constexpr int kNDeg = 17;
constexpr int kDgComplete = 1000;  // degree code of the complete base: kDgComplete + p
constexpr int tet_ns(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }
constexpr int tet_points(int p) { return (p + 2) * (p + 1) * (p + 1); }
constexpr int kTetFaceV[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

// collapsed Gauss-Legendre rule of the tet, [ng][4]: xi = u, eta = v (1 - u),
// zeta = w (1 - u)(1 - v), weight (1 - u)^2 (1 - v); p + 2 points in u and
// p + 1 in v and w, exact for every monomial of degree <= 2p
int tet_volume_rule(int p, std::vector<double>& q) {
  double tu[8], wu[8], tv[8], wv[8];
  const int nu = p + 2, nv = p + 1;
  gauss_legendre(nu, tu, wu);
  gauss_legendre(nv, tv, wv);
  q.clear();
  for (int iu = 0; iu < nu; iu++)
    for (int iv = 0; iv < nv; iv++)
      for (int iw = 0; iw < nv; iw++) {
        const double u = 0.5 * (tu[iu] + 1.0), v = 0.5 * (tv[iv] + 1.0), w = 0.5 * (tv[iw] + 1.0);
        q.push_back(u);
        q.push_back(v * (1.0 - u));
        q.push_back(w * (1.0 - u) * (1.0 - v));
        q.push_back(0.125 * wu[iu] * wv[iv] * wv[iw] * (1.0 - u) * (1.0 - u) * (1.0 - v));
      }
  return nu * nv * nv;
}

// exponents a | b << 2 | c << 4 of the tet base in base order
void tet_exponents(int p, int* ex) {
  int n = 0;
  for (int t = 0; t <= p; t++)
    for (int a = t; a >= 0; a--)
      for (int b = t - a; b >= 0; b--) ex[n++] = a | b << 2 | (t - a - b) << 4;
}

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
struct DgOp;

struct DgMg;
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
  // own and coupling blocks of at most this many shape functions are
  // assembled by the vector-ALU kernels, larger ones on MFMA
  int valu_ns = 6;
  // faces, element-centric: for element e and side s (index 5 e + s)
  int* d_fnb = nullptr;       // neighbour element, or -1 Dirichlet / -2 Neumann
  int* d_fcode = nullptr;     // neighbour side (bits 0-2) | orientation (see build_faces)
  double* d_fhf = nullptr;    // sqrt(face area)
  double* d_fval = nullptr;   // boundary data when no exact solution is selected
  int* d_slots = nullptr;     // interior (element, side) slots 5 e + s
  int nslots = 0;
  double* d_qt = nullptr;     // triangle-face rule [ngt][3]
  double* d_qq = nullptr;     // quad-face rule [ngq][3]
  int ngt = 0, ngq = 0;
  int ex[40] = {0};           // exponents of the shape functions, a | b << 2 | c << 4
  double penalty_const = 1.0;
  // block system: [nel][6][ns*ns] column-major blocks (slot 0 the element's
  // own block, slot 1 + s the coupling across side s), rhs [nel][ns]
  double* d_blk = nullptr;
  double* d_brhs = nullptr;
  double* d_dinv = nullptr;   // inverses of the slot-0 blocks
  bool assembled = false, dinv_valid = false;
  // multicolour block Gauss-Seidel: elements grouped by colour (no two
  // neighbours share a colour), sweeps per application
  std::vector<int> gs_color_ptr;
  int* d_gs_elems = nullptr;
  int gs_sweeps = 2;
  double* d_work = nullptr;   // n, GS work vector
  Problem* P = nullptr;       // GMRES workspace over the block system
  DgOp* op = nullptr;
  // per-element degrees (degree index d = 3 (pz - 1) + (px - 1)), elements
  // grouped by degree (grp_ptr[d] .. grp_ptr[d + 1] in d_grp).  A uniform
  // problem keeps the implicit layout (rows [nel][6][ns^2], vectors
  // [nel][ns]) and null offset arrays; a mixed one stores the offsets of
  // every element's unknowns, block-row slots (own block ns_e^2, then per
  // side ns_e x ns_nb for an interior side, ns_e^2 otherwise) and inverse
  bool mixed = false;
  std::vector<int> deg, grp_ptr, h_fnb;
  int* d_grp = nullptr;
  int* d_deg = nullptr;
  std::vector<long long> xoff, soff, voff;
  long long *d_xoff = nullptr, *d_soff = nullptr, *d_voff = nullptr;
  long long ndof = 0, blk_len = 0, inv_len = 0;
  // face rules of all nine degrees, concatenated in d_qt / d_qq
  int qt_off[kNDeg] = {0}, ngt9[kNDeg] = {0}, qq_off[kNDeg] = {0}, ngq9[kNDeg] = {0};
  int ex9[kNDeg][40] = {};
  // volume rules of the degrees present (mixed problems)
  double* d_qv9 = nullptr;
  int qv_off[kNDeg] = {0};
  // interior slots grouped by (own degree, neighbour degree): slot_ptr[kNDeg^2 + 1]
  std::vector<int> slot_ptr;
  // Gauss-Seidel colours split by degree: gs_cd_ptr[kNDeg c + d]
  std::vector<int> gs_cd_ptr;
  // block ILU(0) in element order: factors on the block-row layout (d_ilu),
  // inverses of U's diagonal blocks (d_ilu_inv, inverse layout), rows by
  // level of the forward (factor, L solve) and backward (U solve) schedules
  bool ilu_valid = false;
  double* d_ilu = nullptr;
  double* d_ilu_inv = nullptr;
  std::vector<int> ilu_fptr, ilu_bptr;
  int* d_ilu_frows = nullptr;
  int* d_ilu_brows = nullptr;
  // wave tasks of the solves per level (ilu_ftask[l] .. ilu_ftask[l + 1])
  std::vector<int> ilu_ftask, ilu_btask;
  int2* d_ilu_ftasks = nullptr;
  int2* d_ilu_btasks = nullptr;
  // batch member (uniform degree): the block-Jacobi-preconditioned system
  // M = Dinv A, b' = Dinv b as a SELL-64 problem, rows in DOF order
  Problem* bp = nullptr;
  bool bp_valid = false;
  bool sol_batch = false;     // the last solution came from a batch
  // p-multigrid levels below this problem's degree (mf_dg_set_multigrid)
  DgMg* mg = nullptr;
  bool mg_valid = false;
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

void free_dg_solver(DgProblem& D);
int dg_ilu_solve(DgProblem& D, double* v, const int* stop);

void free_dg_batch(DgProblem& D);
void free_dg_mg(DgProblem& D);

void free_dg(DgProblem& D) {
  free_dg_mg(D);
  free_dg_solver(D);
  free_dg_batch(D);
  void* ptrs[] = {D.d_nc, D.d_qv, D.d_U, D.d_coef, D.d_stiff, D.d_rhs, D.d_fnb, D.d_fcode,
                  D.d_fhf, D.d_fval, D.d_qt, D.d_qq, D.d_blk, D.d_brhs, D.d_dinv, D.d_slots,
                  D.d_gs_elems, D.d_work, D.d_grp, D.d_deg, D.d_xoff, D.d_soff, D.d_voff,
                  D.d_qv9, D.d_ilu, D.d_ilu_inv, D.d_ilu_frows, D.d_ilu_brows,
                  D.d_ilu_ftasks, D.d_ilu_btasks};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
}
}  // namespace

// ---------------------------------------------------------------------
// DG block system: faces, assembly of block rows, block Jacobi, operator
// ---------------------------------------------------------------------
namespace {

// prism face vertices, quads cyclic with the bottom edge first (face-node
// table mmd_prism/mms_prism_intf.c:98)
constexpr int kFaceV[5][4] = {{0, 1, 2, -1}, {3, 4, 5, -1}, {0, 1, 4, 3}, {1, 2, 5, 4},
                              {2, 0, 3, 5}};

// exponents (a, b, c) of e0^a e1^b z^c for every shape function, in the
// order apr_shape_fun_3D lists them
void shape_exponents(int px, int pz, int* ex) {
  std::vector<std::pair<int, int>> b2 = {{0, 0}, {1, 0}, {0, 1}};
  if (px > 1) {
    b2.push_back({1, 1});
    for (int j = 0; j < 2; j++)
      for (int i = 2; i <= px - j; i++) b2.push_back({i, j});
    for (int j = 2; j <= px; j++)
      for (int i = 0; i <= px - j; i++) b2.push_back({i, j});
  }
  const int first[6][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {0, 1, 1}};
  int k = 0;
  for (auto& f : first) ex[k++] = f[0] | f[1] << 2 | f[2] << 4;
  for (int j = 0; j < 2; j++)
    for (size_t i = 3; i < b2.size(); i++) ex[k++] = b2[i].first | b2[i].second << 2 | j << 4;
  for (int j = 2; j <= pz; j++)
    for (size_t i = 0; i < b2.size(); i++) ex[k++] = b2[i].first | b2[i].second << 2 | j << 4;
}

// exponents of the complete base of degree p in apr_shape_fun_3D's order
// (aps_dg_util.c:566-870): 1, e0, e1, z, z e0, z e1, then e0 e1, e0^2, e1^2,
// z^2 (p = 2) or e0 e1, e0 e1 z and the loops over the 2D functions
// e0^i e1^j times z^k (p >= 3)
void complete_exponents(int p, int* ex) {
  int n = 0;
  auto put = [&](int a, int b, int c) { ex[n++] = a | b << 2 | c << 4; };
  put(0, 0, 0); put(1, 0, 0); put(0, 1, 0); put(0, 0, 1);
  if (p == 1) return;
  put(1, 0, 1); put(0, 1, 1); put(1, 1, 0);
  if (p == 2) {
    put(2, 0, 0); put(0, 2, 0); put(0, 0, 2);
    return;
  }
  put(1, 1, 1);
  for (int j = 0; j < 2; j++)
    for (int i = 0; i < 2; i++)
      for (int k = 2; k <= p - i - j; k++) put(i, j, k);
  for (int j = 0; j < 2; j++)
    for (int i = 2; i <= p - j; i++)
      for (int k = 0; k <= p - i - j; k++) put(i, j, k);
  for (int j = 2; j <= p; j++)
    for (int i = 0; i <= p - j; i++)
      for (int k = 0; k <= p - i - j; k++) put(i, j, k);
}

// exponents of the pz = 0 tensor degrees (aps_dg_util.c:495-600): 1, e0, e1
// (, e0^2, e0 e1, e1^2)
void flat_exponents(int px, int* ex) {
  const int e[6][2] = {{0, 0}, {1, 0}, {0, 1}, {2, 0}, {1, 1}, {0, 2}};
  for (int k = 0; k < (px == 1 ? 3 : 6); k++) ex[k] = e[k][0] | e[k][1] << 2;
}

// face rules (apr_set_quadr_2D, aps_dg_util.c:1499-1555): [n][3] = (x0, x1, w)
void face_rules(int px, int pz, std::vector<double>& qt, std::vector<double>& qq) {
  std::vector<double> l, w;
  triangle_rule(2 * px, l, w);
  qt.clear();
  for (size_t k = 0; k < w.size(); k++) {
    qt.push_back(l[3 * k + 1]); qt.push_back(l[3 * k + 2]); qt.push_back(0.5 * w[k]);
  }
  double x1[8], w1[8], x2[8], w2[8];
  gauss_legendre(px + 1, x1, w1);
  gauss_legendre(pz + 1, x2, w2);
  qq.clear();
  for (int a = 0; a <= px; a++)
    for (int b = 0; b <= pz; b++) {
      qq.push_back(x1[a]); qq.push_back(x2[b]); qq.push_back(0.5 * w1[a] * w2[b]);
    }
}

double face_area(const Mesh& m, int e, int s) {
  const bool tet = m.type[e] == kTetra;
  const int nv = tet || s < 2 ? 3 : 4;
  double p[4][3], d[4][3];
  for (int k = 0; k < nv; k++)
    for (int r = 0; r < 3; r++)
      p[k][r] = m.coor[3 * (size_t)m.nodes[6 * (size_t)e + (tet ? kTetFaceV[s][k] : kFaceV[s][k])] + r];
  for (int k = 0; k < nv; k++)
    for (int r = 0; r < 3; r++) d[k][r] = p[(k + 1) % nv][r] - p[k][r];
  auto cross_len = [](const double* a, const double* b) {
    const double c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2],
                 c2 = a[0] * b[1] - a[1] * b[0];
    return std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  };
  const double l1 = cross_len(d[0], d[1]);
  return nv == 3 ? 0.5 * l1 : 0.5 * (l1 + cross_len(d[2], d[3]));
}

// element-centric face table.  Pairs are matched by their sorted vertex
// numbers; the orientation code lets an element carry a point of its own
// face into the neighbour: between prisms, triangles store, for each
// neighbour face vertex, the matching vertex of the own face (2 bits each),
// quads whether the bottom edge runs the other way; a face with a tet on
// either side sets bit 3 and stores, for each own face vertex, the
// neighbour's local vertex (3 bits each).  A tet's fifth side is -3.  The face size comes from the
// lower-numbered element's vertex cycle (both sides use the same value).
int build_faces(const Mesh& m, std::vector<int>& fnb, std::vector<int>& fcode,
                std::vector<double>& fhf) {
  const size_t nel = m.nel, nf = 5 * nel;
  fnb.assign(nf, -1);
  fcode.assign(nf, 0);
  fhf.assign(nf, 0.0);
  struct Key {
    int v[4];
    int es;
  };
  auto is_tet = [&m](size_t e) { return m.type[e] == kTetra; };
  // the element's local vertex of side s, position k
  auto fv = [&](size_t e, int s, int k) { return is_tet(e) ? kTetFaceV[s][k] : kFaceV[s][k]; };
  auto tri = [&](size_t e, int s) { return is_tet(e) || s < 2; };
  std::vector<Key> keys;
  keys.reserve(nf);
  for (size_t e = 0; e < nel; e++)
    for (int s = 0; s < (is_tet(e) ? 4 : 5); s++) {
      Key k;
      const int nv = tri(e, s) ? 3 : 4;
      for (int t = 0; t < 4; t++) k.v[t] = t < nv ? m.nodes[6 * e + fv(e, s, t)] : -1;
      std::sort(k.v, k.v + nv);
      k.es = (int)(5 * e + s);
      keys.push_back(k);
    }
  for (size_t e = 0; e < nel; e++)
    if (is_tet(e)) fnb[5 * e + 4] = -3;  // a tet has four sides
  std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
    for (int t = 0; t < 4; t++)
      if (a.v[t] != b.v[t]) return a.v[t] < b.v[t];
    return a.es < b.es;
  });
  std::map<std::pair<int, int>, int> bc;
  for (int f = 0; f < m.nbf; f++) bc[{m.bf_elem[f], m.bf_side[f]}] = m.bf_bc[f];
  const size_t nk = keys.size();
  size_t i = 0;
  while (i < nk) {
    size_t j = i + 1;
    while (j < nk && std::equal(keys[j].v, keys[j].v + 4, keys[i].v)) j++;
    if (j - i > 2) {
      set_error("DG faces: a face is shared by more than two elements");
      return -1;
    }
    const int e1 = keys[i].es / 5, s1 = keys[i].es % 5;
    if (j - i == 1) {
      auto it = bc.find({e1, s1});
      const int flag = it == bc.end() ? 1 : it->second;
      if (flag <= 0 || flag >= 20) {
        set_error("DG faces: boundary flags must select Dirichlet (1..9) or Neumann (10..19)");
        return -1;
      }
      fnb[keys[i].es] = flag < 10 ? -1 : -2;
      fhf[keys[i].es] = std::sqrt(face_area(m, e1, s1));
    } else {
      const int e2 = keys[i + 1].es / 5, s2 = keys[i + 1].es % 5;
      if (tri(e1, s1) != tri(e2, s2)) {
        set_error("DG faces: a triangle face meets a quad face");
        return -1;
      }
      const double hf = std::sqrt(face_area(m, std::min(e1, e2), e1 < e2 ? s1 : s2));
      const bool generic = is_tet(e1) || is_tet(e2);
      for (int side = 0; side < 2; side++) {
        const int ea = side ? e2 : e1, sa = side ? s2 : s1, eb = side ? e1 : e2,
                  sb = side ? s1 : s2;
        const int* ga = &m.nodes[6 * (size_t)ea];
        const int* gb = &m.nodes[6 * (size_t)eb];
        int code = sb;
        if (generic) {
          // for each own face vertex k the neighbour's local vertex (3 bits)
          code |= 8;
          for (int k = 0; k < 3; k++) {
            int v2 = -1;
            for (int t = 0; t < 3; t++)
              if (gb[fv(eb, sb, t)] == ga[fv(ea, sa, k)]) v2 = fv(eb, sb, t);
            if (v2 < 0) {
              set_error("DG faces: inconsistent triangle face");
              return -1;
            }
            code |= v2 << (4 + 3 * k);
          }
        } else if (sa < 2) {
          for (int kb = 0; kb < 3; kb++) {
            int match = -1;
            for (int ka = 0; ka < 3; ka++)
              if (gb[kFaceV[sb][kb]] == ga[kFaceV[sa][ka]]) match = ka;
            if (match < 0) {
              set_error("DG faces: inconsistent triangle face");
              return -1;
            }
            code |= match << (4 + 2 * kb);
          }
        } else {
          const int A = ga[kFaceV[sa][0]], B = ga[kFaceV[sa][1]];
          const int A2 = gb[kFaceV[sb][0]], B2 = gb[kFaceV[sb][1]];
          if (A2 == B && B2 == A)
            code |= 1 << 4;
          else if (!(A2 == A && B2 == B)) {
            set_error("DG faces: neighbouring prisms are not stacked the same way up");
            return -1;
          }
        }
        fnb[5 * (size_t)ea + sa] = eb;
        fcode[5 * (size_t)ea + sa] = code;
        fhf[5 * (size_t)ea + sa] = hf;
      }
    }
    i = j;
  }
  return 0;
}

__device__ __forceinline__ double sel4(int p, double a0, double a1, double a2, double a3) {
  return p == 0 ? a0 : p == 1 ? a1 : p == 2 ? a2 : a3;
}

__device__ __forceinline__ double pw(double x, int a) {
  return a == 0 ? 1.0 : a == 1 ? x : a == 2 ? x * x : x * x * x;
}

// shape functions t*16 + c16 (t < NT) and their reference derivatives at eta;
// powers and derivatives are selected from x^0..x^3 without branches
template <int NS, int NT>
__device__ __forceinline__ void mono_eval(const int* ex, int c16, const double* eta,
                                          double (&u)[4][NT]) {
  double P[3][4], dP[3][4];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double x = eta[d], x2 = x * x;
    P[d][0] = 1.0; P[d][1] = x; P[d][2] = x2; P[d][3] = x2 * x;
    dP[d][0] = 0.0; dP[d][1] = 1.0; dP[d][2] = 2.0 * x; dP[d][3] = 3.0 * x2;
  }
#pragma unroll
  for (int t = 0; t < NT; t++) {
    const int i = 16 * t + c16;
    const int code = (i < NS) ? ex[i < NS ? i : 0] : 0;
    const int a = code & 3, b = (code >> 2) & 3, c = (code >> 4) & 3;
    const double pa = sel4(a, P[0][0], P[0][1], P[0][2], P[0][3]);
    const double pb = sel4(b, P[1][0], P[1][1], P[1][2], P[1][3]);
    const double pc = sel4(c, P[2][0], P[2][1], P[2][2], P[2][3]);
    const double da = sel4(a, dP[0][0], dP[0][1], dP[0][2], dP[0][3]);
    const double db = sel4(b, dP[1][0], dP[1][1], dP[1][2], dP[1][3]);
    const double dc = sel4(c, dP[2][0], dP[2][1], dP[2][2], dP[2][3]);
    const double m = (i < NS) ? 1.0 : 0.0;
    u[0][t] = m * (pa * pb * pc);
    u[1][t] = m * (da * pb * pc);
    u[2][t] = m * (pa * db * pc);
    u[3][t] = m * (pa * pb * dc);
  }
}

// element-local coordinates of face point xf on side s (the first-neighbour
// branch of mmr_fa_elem_coor, mms_prism_datstr.c:1541-1700)
__device__ __forceinline__ void face_eta(int s, const double* xf, double* eta) {
  if (s == 0) { eta[0] = xf[1]; eta[1] = xf[0]; eta[2] = -1.0; }
  else if (s == 1) { eta[0] = xf[0]; eta[1] = xf[1]; eta[2] = 1.0; }
  else if (s == 2) { eta[0] = 0.5 * (xf[0] + 1.0); eta[1] = 0.0; eta[2] = xf[1]; }
  else if (s == 3) { eta[0] = 0.5 * (-xf[0] + 1.0); eta[1] = 1.0 - eta[0]; eta[2] = xf[1]; }
  else { eta[0] = 0.0; eta[1] = 0.5 * (-xf[0] + 1.0); eta[2] = xf[1]; }
}

// the same point in the neighbour across side s (code from build_faces)
__device__ __forceinline__ void neighbour_eta(int s, int code, const double* eta, double* eta2) {
  const int s2 = code & 7;
  if (s < 2) {
    const double lam[3] = {1.0 - eta[0] - eta[1], eta[0], eta[1]};
    const int m1 = (code >> 6) & 3, m2 = (code >> 8) & 3;
    eta2[0] = m1 == 0 ? lam[0] : m1 == 1 ? lam[1] : lam[2];
    eta2[1] = m2 == 0 ? lam[0] : m2 == 1 ? lam[1] : lam[2];
    eta2[2] = s2 == 0 ? -1.0 : 1.0;
  } else {
    double t = s == 2 ? eta[0] : s == 3 ? eta[1] : 1.0 - eta[1];
    if ((code >> 4) & 1) t = 1.0 - t;
    if (s2 == 2) { eta2[0] = t; eta2[1] = 0.0; }
    else if (s2 == 3) { eta2[0] = 1.0 - t; eta2[1] = t; }
    else { eta2[0] = 0.0; eta2[1] = 1.0 - t; }
    eta2[2] = eta[2];
  }
}

// Jacobian columns, inverse and point of the linear prism map at eta
__device__ __forceinline__ double prism_jac(const double* xn, const double* eta, double* J,
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
#pragma unroll
  for (int k = 0; k < 9; k++) J[k] = 0.0;
  x[0] = x[1] = x[2] = 0.0;
#pragma unroll
  for (int i = 0; i < 6; i++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const double c = xn[3 * i + r];
      J[3 * r] += c * gx[i]; J[3 * r + 1] += c * gy[i]; J[3 * r + 2] += c * gz[i];
      x[r] += c * gp[i];
    }
  const double* m = J;
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

// area element |ds| and unit outward normal of side s (apr_elem_calc_3D
// with Control = 3 + s, aps_dg_util.c:334-367)
__device__ __forceinline__ double face_ds(int s, const double* J, double* n) {
  double a[3], b[3];
  const double t0[3] = {J[0], J[3], J[6]}, t1[3] = {J[1], J[4], J[7]}, t2[3] = {J[2], J[5], J[8]};
  double sg = 1.0;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    if (s == 0) { a[r] = t0[r]; b[r] = t1[r]; }
    else if (s == 1) { a[r] = t0[r]; b[r] = t1[r]; }
    else if (s == 2) { a[r] = t0[r]; b[r] = t2[r]; }
    else if (s == 3) { a[r] = t1[r] - t0[r]; b[r] = t2[r]; }
    else { a[r] = t2[r]; b[r] = t1[r]; }
  }
  if (s == 0) sg = -1.0;
  double ds[3] = {sg * (a[1] * b[2] - a[2] * b[1]), sg * (a[2] * b[0] - a[0] * b[2]),
                  sg * (a[0] * b[1] - a[1] * b[0])};
  const double d = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
  n[0] = ds[0] / d; n[1] = ds[1] / d; n[2] = ds[2] / d;
  return d;
}

struct DgRowArgs {
  DgElemArgs E;        // element data, coefficients, volume rule of the launch's degree
  const int* fnb;
  const int* fcode;
  const double* fhf;
  const double* fval;
  // face rules of every degree index d = 3 (pz - 1) + (px - 1): triangle
  // rules [ngt[d]][3] at qt + 3 qt_off[d], quad rules likewise
  const double* qt;
  const double* qq;
  int qt_off[kNDeg], ngt[kNDeg], qq_off[kNDeg], ngq[kNDeg];
  double penalty_const;
  const int* deg;      // degree index of every element
  const int* elems;    // elements of the launch (own blocks) or null: all, in order
  // offsets of every element's unknowns and block-row slots (own block, then
  // the coupling across each side); null for a uniform problem, whose rows
  // are [nel][6][NS*NS] and vectors [nel][NS]
  const long long* xoff;
  const long long* soff;
  int ex[40];          // exponents of the launch's (element's) degree
  int exn[40];         // exponents of the neighbour's degree (coupling launches)
  double* blk;
  double* brhs;
};

__host__ __device__ constexpr int cmax(int a, int b) { return a > b ? a : b; }
__host__ __device__ constexpr int deg_index(int px, int pz) {
  return pz == 0 ? 14 + px : 3 * (pz - 1) + (px - 1);
}
__host__ __device__ constexpr int tet_deg_index(int p) { return 8 + p; }
__host__ __device__ constexpr int complete_deg_index(int p) { return 11 + p; }
// degree index of a launch's elements of kind 0 (prism (PX, PZ)), 1 (tet of
// degree PX) or 2 (complete-base prism of degree PX)
__host__ __device__ constexpr int deg_of(int px, int pz, int kind) {
  return kind == 1 ? tet_deg_index(px) : kind == 2 ? complete_deg_index(px) : deg_index(px, pz);
}
__host__ __device__ constexpr int ns_of_deg(int d) {
  return d >= 15 ? num_shape(d - 14, 0)
         : d >= 12 ? tet_ns(d - 11)
         : d >= 9 ? tet_ns(d - 8) : num_shape(d % 3 + 1, d / 3 + 1);
}
__host__ __device__ constexpr int pmin_of_deg(int d) {
  return d >= 15 ? 0
         : d >= 12 ? d - 11
         : d >= 9 ? d - 8 : d % 3 < d / 3 ? d % 3 + 1 : d / 3 + 1;
}
// the tensor degree (px, pz) whose volume and face rules degree index d uses
// (tets: the isotropic prism degree's triangle rule)
__host__ __device__ constexpr int rule_px(int d) {
  return d >= 15 ? d - 14 : d >= 12 ? d - 11 : d >= 9 ? d - 8 : d % 3 + 1;
}
__host__ __device__ constexpr int rule_pz(int d) {
  return d >= 15 ? 0 : d >= 12 ? d - 11 : d >= 9 ? d - 8 : d / 3 + 1;
}

// the rule of a face between element e (degree de) and its neighbour nb
// (degree dn): the one with more points, the second neighbour's (the higher
// element) on a tie (pds_conv_diff_weakform_dg.c:508-514)
__device__ __forceinline__ int face_rule_deg(const DgRowArgs& R, int tri, int e, int de, int nb,
                                             int dn) {
  const int ce = tri ? R.ngt[de] : R.ngq[de], cn = tri ? R.ngt[dn] : R.ngq[dn];
  return e < nb ? (ce > cn ? de : dn) : (cn > ce ? dn : de);
}

// own side of a face point: local coordinates, area element times weight,
// outward normal, global point, Ji a^T n (the flux a grad(.) . n in
// reference derivatives) and b . n of element e
struct FacePt {
  double eta[3], area, n[3], x[3], gh[3], bn;
};

__device__ __forceinline__ void face_point(const DgElemArgs& A, const double* xe, int e, int s,
                                           const double* rule, FacePt& F) {
  double J[9], Ji[9];
  face_eta(s, rule, F.eta);
  prism_jac(xe, F.eta, J, Ji, F.x);
  F.area = face_ds(s, J, F.n) * rule[2];
  double a[9], b[3], mc, q0, q1, q2, sv;
  dg_coefs(A, e, a, b, mc, q0, q1, q2, sv);
  double An[3];
#pragma unroll
  for (int c = 0; c < 3; c++) An[c] = a[c] * F.n[0] + a[3 + c] * F.n[1] + a[6 + c] * F.n[2];
#pragma unroll
  for (int k = 0; k < 3; k++) F.gh[k] = Ji[3 * k] * An[0] + Ji[3 * k + 1] * An[1] + Ji[3 * k + 2] * An[2];
  F.bn = b[0] * F.n[0] + b[1] * F.n[1] + b[2] * F.n[2];
}

__device__ __forceinline__ double neighbour_bn(const DgElemArgs& A, int nb, const double* n) {
  double a[9], b[3], mc, q0, q1, q2, sv;
  dg_coefs(A, nb, a, b, mc, q0, q1, q2, sv);
  return b[0] * n[0] + b[1] * n[1] + b[2] * n[2];
}

// reference vertex coordinates: tets (0,0,0), (1,0,0), (0,1,0), (0,0,1);
// prisms the triangle times [-1, 1]
__device__ __forceinline__ void tet_vertex(int v, double* X) {
  X[0] = v == 1 ? 1.0 : 0.0;
  X[1] = v == 2 ? 1.0 : 0.0;
  X[2] = v == 3 ? 1.0 : 0.0;
}
__device__ __forceinline__ void prism_vertex(int v, double* X) {
  const int t = v % 3;
  X[0] = t == 1 ? 1.0 : 0.0;
  X[1] = t == 2 ? 1.0 : 0.0;
  X[2] = v < 3 ? -1.0 : 1.0;
}
// local vertex k of tet side s ({0,1,2}, {0,1,3}, {0,2,3}, {1,2,3})
__device__ __forceinline__ int tet_side_vertex(int s, int k) {
  return k == 0 ? (s == 3 ? 1 : 0) : k == 1 ? (s == 0 || s == 1 ? 1 : s == 2 ? 2 : 2)
                                            : (s == 0 ? 2 : 3);
}

// affine tet map at eta: J[3r + c] = x_{c+1, r} - x_{0, r}, its inverse and
// the point; returns det J
__device__ __forceinline__ double tet_jac(const double* xn, const double* eta, double* J,
                                          double* Ji, double* x) {
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int c = 0; c < 3; c++) J[3 * r + c] = xn[3 * (c + 1) + r] - xn[r];
#pragma unroll
  for (int r = 0; r < 3; r++)
    x[r] = xn[r] + J[3 * r] * eta[0] + J[3 * r + 1] * eta[1] + J[3 * r + 2] * eta[2];
  const double* m = J;
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

// own side of a face point of a tet: side s's point (r0, r1) of the
// reference triangle at X_a + r0 (X_b - X_a) + r1 (X_c - X_a); area element
// |(x_b - x_a) x (x_c - x_a)| times the weight, normal away from the vertex
// opposite the side
__device__ __forceinline__ void face_point_tet(const DgElemArgs& A, const double* xe, int e,
                                               int s, const double* rule, FacePt& F) {
  const int a = tet_side_vertex(s, 0), b = tet_side_vertex(s, 1), c = tet_side_vertex(s, 2);
  const int o = 6 - a - b - c;
  double Xa[3], Xb[3], Xc[3], J[9], Ji[9];
  tet_vertex(a, Xa);
  tet_vertex(b, Xb);
  tet_vertex(c, Xc);
#pragma unroll
  for (int d = 0; d < 3; d++) F.eta[d] = Xa[d] + rule[0] * (Xb[d] - Xa[d]) + rule[1] * (Xc[d] - Xa[d]);
  tet_jac(xe, F.eta, J, Ji, F.x);
  double u[3], v[3], w[3], ds[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    u[d] = xe[3 * b + d] - xe[3 * a + d];
    v[d] = xe[3 * c + d] - xe[3 * a + d];
    w[d] = xe[3 * o + d] - xe[3 * a + d];
  }
  ds[0] = u[1] * v[2] - u[2] * v[1];
  ds[1] = u[2] * v[0] - u[0] * v[2];
  ds[2] = u[0] * v[1] - u[1] * v[0];
  if (ds[0] * w[0] + ds[1] * w[1] + ds[2] * w[2] > 0) {
    ds[0] = -ds[0]; ds[1] = -ds[1]; ds[2] = -ds[2];
  }
  const double det = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
  F.n[0] = ds[0] / det; F.n[1] = ds[1] / det; F.n[2] = ds[2] / det;
  F.area = det * rule[2];
  double aa[9], b3[3], mc, q0, q1, q2, sv;
  dg_coefs(A, e, aa, b3, mc, q0, q1, q2, sv);
  double An[3];
#pragma unroll
  for (int cc = 0; cc < 3; cc++) An[cc] = aa[cc] * F.n[0] + aa[3 + cc] * F.n[1] + aa[6 + cc] * F.n[2];
#pragma unroll
  for (int k = 0; k < 3; k++) F.gh[k] = Ji[3 * k] * An[0] + Ji[3 * k + 1] * An[1] + Ji[3 * k + 2] * An[2];
  F.bn = b3[0] * F.n[0] + b3[1] * F.n[1] + b3[2] * F.n[2];
}

// the neighbour's local coordinates of a face point: prism pairs by their
// orientation code, faces with a tet through the shared vertices (code bit
// 3: the neighbour's local vertex of each own face vertex in 3-bit fields)
template <bool TE, bool TN>
__device__ __forceinline__ void neighbour_eta_x(int s, int code, const double* eta, double* eta2) {
  if (!TE && !TN) {
    neighbour_eta(s, code, eta, eta2);
    return;
  }
  double lam[3];
  if (TE) {
    const double bary[4] = {1.0 - eta[0] - eta[1] - eta[2], eta[0], eta[1], eta[2]};
#pragma unroll
    for (int k = 0; k < 3; k++) lam[k] = bary[tet_side_vertex(s, k)];
  } else {
    lam[0] = 1.0 - eta[0] - eta[1]; lam[1] = eta[0]; lam[2] = eta[1];
  }
  eta2[0] = eta2[1] = eta2[2] = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const int v2 = (code >> (4 + 3 * k)) & 7;
    double X[3];
    if (TN) tet_vertex(v2, X);
    else prism_vertex(v2, X);
#pragma unroll
    for (int d = 0; d < 3; d++) eta2[d] += lam[k] * X[d];
  }
  if (!TN) eta2[2] = (code & 7) == 0 ? -1.0 : 1.0;
}

template <int PX, int PZ>
struct DgRowShape {
  static constexpr int NGT = tri_points(PX), NGQ = (PX + 1) * (PZ + 1);
  static constexpr int NPTS = DgShape<PX, PZ>::NG + 2 * NGT + 3 * NGQ;  // volume + 5 faces
};

// powers x^0..x^3 of the three local coordinates of a point (the monomial
// base needs no more), stored with a point's record
__device__ __forceinline__ void put_powers(const double* eta, double* P) {
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double x = eta[d], x2 = x * x;
    P[4 * d] = 1.0; P[4 * d + 1] = x; P[4 * d + 2] = x2; P[4 * d + 3] = x2 * x;
  }
}

// per-lane constants for the lane's shape functions i = 16 t + c16: exponent
// offsets into a power table and the exponents as factors of the derivative
template <int NS, int NT>
struct LaneShapes {
  int oa[NT], ob[NT], oc[NT], da[NT], db[NT], dc[NT];
  double fa[NT], fb[NT], fc[NT], live[NT];
  __device__ __forceinline__ void init(const int* ex, int c16) {
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const int i = 16 * t + c16;
      const int code = i < NS ? ex[i < NS ? i : 0] : 0;
      const int a = code & 3, b = (code >> 2) & 3, c = (code >> 4) & 3;
      oa[t] = a; ob[t] = 4 + b; oc[t] = 8 + c;
      da[t] = a > 0 ? a - 1 : 0; db[t] = 4 + (b > 0 ? b - 1 : 0); dc[t] = 8 + (c > 0 ? c - 1 : 0);
      fa[t] = a; fb[t] = b; fc[t] = c;
      live[t] = i < NS ? 1.0 : 0.0;
    }
  }
  // phi and its three reference derivatives from a power table
  __device__ __forceinline__ void eval(const double* P, double (&u)[4][NT]) const {
#pragma unroll
    for (int t = 0; t < NT; t++) {
      const double pa = P[oa[t]], pb = P[ob[t]], pc = P[oc[t]];
      const double qa = fa[t] * P[da[t]], qb = fb[t] * P[db[t]], qc = fc[t] * P[dc[t]];
      u[0][t] = live[t] * (pa * pb * pc);
      u[1][t] = live[t] * (qa * pb * pc);
      u[2][t] = live[t] * (pa * qb * pc);
      u[3][t] = live[t] * (pa * pb * qc);
    }
  }
};

// Record of point k of element e for its own block (volume points first,
// then the points of sides 0..4 under each side's rule): the powers of the
// local coordinates (Rg[0..11]), the 4x4 matrix M between (phi, d/deta phi)
// of trial function (row) and test function (column) (Rg[12..27]) and the
// load-vector row (Rg[28..31]).  Mixed problems pass each side's rule,
// first point and penalty / h_F in fq, fst and fpen.
// sizes of a launch's elements: a tet of degree PX (TET) or a prism (PX, PZ)
template <int PX, int PZ, int TET>
struct DgEl {
  static constexpr int NS = TET ? tet_ns(PX) : num_shape(PX, PZ);  // complete: tet_ns too
  static constexpr int NG = TET == 1 ? tet_points(PX) : tri_points(PX) * (PZ + 1);
  static constexpr int NT = (NS + 15) / 16;
  static constexpr int NSIDES = TET == 1 ? 4 : 5;
  static constexpr int DE = deg_of(PX, PZ, TET);
};

template <int PX, int PZ, bool MIXED, int TET = 0>
__device__ __forceinline__ void diag_record(const DgRowArgs& R, int e, int k, const double* sX,
                                            const int* fq, const int* fst, const double* fpen,
                                            double* Rg) {
  using Q = DgRowShape<PX, PZ>;
  constexpr int NG = DgEl<PX, PZ, TET>::NG, DE = DgEl<PX, PZ, TET>::DE;
  const DgElemArgs& A = R.E;
  double* M = Rg + 12;
#pragma unroll
  for (int t = 0; t < 20; t++) M[t] = 0.0;
  if (k < NG) {
    const double* qg = A.qv + 4 * k;
    double a[9], b[3], mc, q0, q1, q2, sv, J[9], Ji[9], x[3];
    dg_coefs(A, e, a, b, mc, q0, q1, q2, sv);
    const double vol = fabs(TET == 1 ? tet_jac(sX, qg, J, Ji, x) : prism_jac(sX, qg, J, Ji, x)) * qg[3];
    if (A.src_model != SRC_CONST && A.mask.s) {
      double uu, lapl;
      exact_sol(A.src_model, x, &uu, &lapl);
      sv = -lapl;
    }
    put_powers(qg, Rg);
    M[0] = vol * mc;
#pragma unroll
    for (int kk = 0; kk < 3; kk++) {
      M[1 + kk] = vol * (b[0] * Ji[3 * kk] + b[1] * Ji[3 * kk + 1] + b[2] * Ji[3 * kk + 2]);
      double t[3];
#pragma unroll
      for (int r = 0; r < 3; r++)
        t[r] = Ji[3 * kk] * a[3 * r] + Ji[3 * kk + 1] * a[3 * r + 1] + Ji[3 * kk + 2] * a[3 * r + 2];
#pragma unroll
      for (int l = 0; l < 3; l++)
        M[4 * (1 + kk) + 1 + l] = vol * (t[0] * Ji[3 * l] + t[1] * Ji[3 * l + 1] + t[2] * Ji[3 * l + 2]);
    }
    M[16] = vol * sv;
#pragma unroll
    for (int kk = 0; kk < 3; kk++)
      M[17 + kk] = vol * (q0 * Ji[3 * kk] + q1 * Ji[3 * kk + 1] + q2 * Ji[3 * kk + 2]);
  } else {
    int sd = 0;
    const double* rule;
    if (MIXED) {
      while (k >= fst[sd + 1]) sd++;
      const int q = fq[sd];
      rule = (TET == 1 || sd < 2 ? R.qt + 3 * R.qt_off[q] : R.qq + 3 * R.qq_off[q]) + 3 * (k - fst[sd]);
    } else {
      int fq0 = k - NG;
      if (fq0 < 2 * Q::NGT) {
        sd = fq0 / Q::NGT;
        rule = R.qt + 3 * (R.qt_off[DE] + fq0 - sd * Q::NGT);
      } else {
        fq0 -= 2 * Q::NGT;
        sd = 2 + fq0 / Q::NGQ;
        rule = R.qq + 3 * (R.qq_off[DE] + fq0 - (sd - 2) * Q::NGQ);
      }
    }
    const int fi = 5 * e + sd, nb = R.fnb[fi];
    const double pen = MIXED ? fpen[sd] : pmin_of_deg(DE) * R.penalty_const / R.fhf[fi];
    FacePt F;
    if constexpr (TET == 1) face_point_tet(A, sX, e, sd, rule, F);
    else face_point(A, sX, e, sd, rule, F);
    put_powers(F.eta, Rg);
    if (nb >= 0) {
      // interior: 0.5 (phi_j g_i - g_j phi_i) + (pen + upwind) phi_j phi_i
      const double nv = 0.5 * (F.bn + neighbour_bn(A, nb, F.n));
      M[0] = F.area * (pen + (nv > 0 ? F.bn : 0.0));
#pragma unroll
      for (int kk = 0; kk < 3; kk++) {
        M[1 + kk] = 0.5 * F.area * F.gh[kk];
        M[4 * (1 + kk)] = -0.5 * F.area * F.gh[kk];
      }
    } else if (nb == -1) {
      // Dirichlet (weakform_dg.c:1036-1120)
      double fd = R.fval[fi];
      if (A.src_model != SRC_CONST) {
        double lapl;
        exact_sol(A.src_model, F.x, &fd, &lapl);
      }
      M[0] = F.area * (pen + (F.bn >= 0 ? F.bn : 0.0));
#pragma unroll
      for (int kk = 0; kk < 3; kk++) {
        M[1 + kk] = F.area * F.gh[kk];
        M[4 * (1 + kk)] = -F.area * F.gh[kk];
        M[17 + kk] = F.area * fd * F.gh[kk];
      }
      M[16] = F.area * (pen * fd - (F.bn < 0 ? F.bn * fd : 0.0));
    } else {
      // Neumann (weakform_dg.c:1148-1185): the exact normal derivative
      double gd = R.fval[fi];
      if (A.src_model != SRC_CONST) {
        const double a1 = A.src_model == SRC_DIFF_IN_BOX ? 100.0 : 1.0;
        const double* x = F.x;
        const double ex = exp((-x[0] * x[0] - x[1] * x[1] - x[2] * x[2]) / a1);
        gd = (-2.0 * x[0] * F.n[0] - 2.0 * x[1] * F.n[1] - 2.0 * x[2] * F.n[2]) * ex / a1;
      }
      M[0] = F.area * (F.bn > 0 ? F.bn : 0.0);
      M[16] = F.area * gd;
    }
  }
}

// Own blocks and load vectors.  One wave per element.  Points come in
// chunks of up to 64, one lane each: the volume points, then the points of
// the five faces with the element's own terms of each face; a lane writes its
// point's record (powers of the local coordinates, the 4x4 matrix between
// (phi, d/deta phi) of trial and test function, the load-vector row) to
// LDS, and the wave runs the chunk through v_mfma_f64_16x16x4, one K-step
// per point.
template <int PX, int PZ, bool MIXED, int TET = 0>
__global__ void __launch_bounds__(64) k_dg_diag(DgRowArgs R) {
  using S = DgEl<PX, PZ, TET>;
  using Q = DgRowShape<PX, PZ>;
  // records of 32 doubles at an odd stride: the lanes writing their own
  // records fall on different LDS banks
  constexpr int NS = S::NS, NG = S::NG, NT = S::NT, DE = S::DE, REC = 33, NSD = S::NSIDES;
  __shared__ double rec[64 * REC];
  __shared__ double sX[18];
  __shared__ int sex[40];
  __shared__ int fq[5], fst[6];
  __shared__ double fpen[5];
  const DgElemArgs& A = R.E;
  const int lane = threadIdx.x;
  const int e = R.elems ? R.elems[blockIdx.x] : blockIdx.x;
  if (lane < 40) sex[lane] = R.ex[lane];
  if (lane < 18) sX[lane] = A.nc[(size_t)lane * A.nel + e];
  if (MIXED && lane < NSD) {
    // each face's rule and penalty / h_F (weakform_dg.c:508-531)
    const int sd = lane, nb = R.fnb[5 * e + sd];
    const bool tri = TET == 1 || sd < 2;
    int q = DE, pm = pmin_of_deg(DE);
    if (nb >= 0) {
      const int dn = R.deg[nb];
      q = face_rule_deg(R, tri, e, DE, nb, dn);
      pm = min(pm, pmin_of_deg(dn));
    }
    fq[sd] = q;
    fst[sd + 1] = tri ? R.ngt[q] : R.ngq[q];
    fpen[sd] = pm * R.penalty_const / R.fhf[5 * e + sd];
  }
  __syncthreads();
  if (MIXED && lane == 0) {
    fst[0] = NG;
    for (int sd = 0; sd < NSD; sd++) fst[sd + 1] += fst[sd];
  }
  __syncthreads();
  const int NPTS = MIXED ? fst[NSD] : Q::NPTS;
  const int p = lane >> 4, c16 = lane & 15;
  LaneShapes<NS, NT> L;
  L.init(sex, c16);
  const double w0 = p == 0 ? 1.0 : 0.0, w1 = p == 1 ? 1.0 : 0.0, w2 = p == 2 ? 1.0 : 0.0,
               w3 = p == 3 ? 1.0 : 0.0;
  dbl4 acc[NT][NT];
  double f[NT];
#pragma unroll
  for (int t = 0; t < NT; t++) f[t] = 0.0;
#pragma unroll
  for (int jt = 0; jt < NT; jt++)
#pragma unroll
    for (int it = 0; it < NT; it++) acc[jt][it] = dbl4{0.0, 0.0, 0.0, 0.0};
  for (int c0 = 0; c0 < NPTS; c0 += 64) {
    const int k = c0 + lane;
    if (k < NPTS) {
      diag_record<PX, PZ, MIXED, TET>(R, e, k, sX, fq, fst, fpen, rec + REC * lane);
    }
    __syncthreads();
    const int nk = NPTS - c0 < 64 ? NPTS - c0 : 64;
    for (int g = 0; g < nk; g++) {
      const double* Rg = rec + REC * g;
      double u[4][NT];
      L.eval(Rg, u);
      const double* M = Rg + 12 + 4 * p;
      const double m0 = M[0], m1 = M[1], m2 = M[2], m3 = M[3];
      const double f0 = Rg[28], f1 = Rg[29], f2 = Rg[30], f3 = Rg[31];
      double av[NT], bv[NT];
#pragma unroll
      for (int t = 0; t < NT; t++) {
        av[t] = w0 * u[0][t] + w1 * u[1][t] + w2 * u[2][t] + w3 * u[3][t];
        bv[t] = m0 * u[0][t] + m1 * u[1][t] + m2 * u[2][t] + m3 * u[3][t];
        f[t] += f0 * u[0][t] + f1 * u[1][t] + f2 * u[2][t] + f3 * u[3][t];
      }
#pragma unroll
      for (int jt = 0; jt < NT; jt++)
#pragma unroll
        for (int it = 0; it < NT; it++)
          acc[jt][it] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[jt], bv[it], acc[jt][it], 0, 0, 0);
    }
    __syncthreads();
  }
  double* out = R.blk + (R.soff ? R.soff[6 * (size_t)e] : (size_t)e * 6 * NS * NS);
  double* rout = R.brhs + (R.xoff ? R.xoff[e] : (size_t)e * NS);
#pragma unroll
  for (int jt = 0; jt < NT; jt++)
#pragma unroll
    for (int it = 0; it < NT; it++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int j = 16 * jt + 4 * r + p, i = 16 * it + c16;
        if (j < NS && i < NS) out[j * NS + i] = acc[jt][it][r];
      }
  if (p == 0) {
#pragma unroll
    for (int t = 0; t < NT; t++)
      if (16 * t + c16 < NS) rout[16 * t + c16] = f[t];
  }
}

// Record of face point k of the interior side s of element e (neighbour nb
// across it, orientation code, the face's rule q): the powers of both
// elements' local coordinates (Rg[0..11] own, Rg[12..23] neighbour) and the
// 4x4 matrix M between (psi, d/deta psi) of the neighbour's trial function
// (row) and (phi, d/deta phi) of the element's test function (column)
// (Rg[24..39]).  sX holds both elements' node coordinates.
template <int PXE, int PZE, int PXN, int PZN, int TE = 0, int TN = 0>
__device__ __forceinline__ void couple_record(const DgRowArgs& R, int e, int s, int nb, int code,
                                              int q, int kp, const double* sX, double* Rg) {
  constexpr int DE = deg_of(PXE, PZE, TE), DN = deg_of(PXN, PZN, TN);
  const DgElemArgs& A = R.E;
  const int slot = 5 * e + s;
  const double* rule = (TE == 1 || TN == 1 || s < 2 ? R.qt + 3 * R.qt_off[q] : R.qq + 3 * R.qq_off[q]) + 3 * kp;
  FacePt F;
  if constexpr (TE == 1) face_point_tet(A, sX, e, s, rule, F);
  else face_point(A, sX, e, s, rule, F);
  double eta2[3], J2[9], Ji2[9], x2[3];
  neighbour_eta_x<TE == 1, TN == 1>(s, code, F.eta, eta2);
  if constexpr (TN == 1) tet_jac(sX + 18, eta2, J2, Ji2, x2);
  else prism_jac(sX + 18, eta2, J2, Ji2, x2);
  double a[9], b[3], mc, q0, q1, q2, sv;
  dg_coefs(A, nb, a, b, mc, q0, q1, q2, sv);
  double An[3], hh[3];
#pragma unroll
  for (int c = 0; c < 3; c++) An[c] = a[c] * F.n[0] + a[3 + c] * F.n[1] + a[6 + c] * F.n[2];
#pragma unroll
  for (int k = 0; k < 3; k++) hh[k] = Ji2[3 * k] * An[0] + Ji2[3 * k + 1] * An[1] + Ji2[3 * k + 2] * An[2];
  const double bnn = b[0] * F.n[0] + b[1] * F.n[1] + b[2] * F.n[2];
  const double nv = 0.5 * (F.bn + bnn);
  const int pm = min(pmin_of_deg(DE), pmin_of_deg(DN));
  const double pen = pm * R.penalty_const / R.fhf[slot];
  put_powers(F.eta, Rg);
  put_powers(eta2, Rg + 12);
  double* M = Rg + 24;
#pragma unroll
  for (int t = 0; t < 16; t++) M[t] = 0.0;
  M[0] = F.area * (-pen + (nv < 0 ? F.bn : 0.0));
#pragma unroll
  for (int k = 0; k < 3; k++) {
    M[1 + k] = -0.5 * F.area * F.gh[k];
    M[4 * (1 + k)] = -0.5 * F.area * hh[k];
  }
}

// Coupling blocks: one wave per interior (element, side): the block of the
// element's test functions against the neighbour's trial functions,
// -0.5 (psi_j g_i + h_j phi_i) + (-pen + upwind) psi_j phi_i over the face
// (the (1,2) block of pdr_conv_diff_comp_fa_stiff_mat_app seen from either
// side), ns_e x ns_nb, column-major.  Records hold the powers of both
// elements' local coordinates.
template <int PXE, int PZE, int PXN, int PZN, int TET_E = 0, int TET_N = 0>
__global__ void __launch_bounds__(64) k_dg_couple(DgRowArgs R, const int* slots) {
  constexpr int NSE = DgEl<PXE, PZE, TET_E>::NS, NTE = (NSE + 15) / 16;
  constexpr int NSN = DgEl<PXN, PZN, TET_N>::NS, NTN = (NSN + 15) / 16;
  constexpr int DE = deg_of(PXE, PZE, TET_E), DN = deg_of(PXN, PZN, TET_N), REC = 41;  // 40 + bank pad
  constexpr bool TRI = TET_E == 1 || TET_N == 1;  // a face with a tet is a triangle
  constexpr int QE = TRI ? 0 : (PXE + 1) * (PZE + 1), QN = TRI ? 0 : (PXN + 1) * (PZN + 1);
  constexpr int TE = tri_points(PXE), TN = tri_points(PXN);
  constexpr int NPMAX = cmax(cmax(QE, QN), cmax(TE, TN));
  __shared__ double rec[NPMAX * REC];
  __shared__ double sX[36];
  __shared__ int sex[40], sexn[40];
  const DgElemArgs& A = R.E;
  const int lane = threadIdx.x;
  const int slot = slots[blockIdx.x];
  const int e = slot / 5, s = slot % 5;
  const int nb = R.fnb[slot], code = R.fcode[slot];
  const int q = face_rule_deg(R, TRI || s < 2, e, DE, nb, DN);
  const int npf = TRI || s < 2 ? R.ngt[q] : R.ngq[q];
  if (lane < 40) {
    sex[lane] = R.ex[lane];
    sexn[lane] = R.exn[lane];
  }
  if (lane < 18) sX[lane] = A.nc[(size_t)lane * A.nel + e];
  else if (lane < 36) sX[lane] = A.nc[(size_t)(lane - 18) * A.nel + nb];
  __syncthreads();
  if (lane < npf)
    couple_record<PXE, PZE, PXN, PZN, TET_E, TET_N>(R, e, s, nb, code, q, lane, sX, rec + REC * lane);
  __syncthreads();
  const int p = lane >> 4, c16 = lane & 15;
  LaneShapes<NSE, NTE> L;
  L.init(sex, c16);
  LaneShapes<NSN, NTN> LN;
  LN.init(sexn, c16);
  const double w0 = p == 0 ? 1.0 : 0.0, w1 = p == 1 ? 1.0 : 0.0, w2 = p == 2 ? 1.0 : 0.0,
               w3 = p == 3 ? 1.0 : 0.0;
  dbl4 acc[NTN][NTE];
#pragma unroll
  for (int jt = 0; jt < NTN; jt++)
#pragma unroll
    for (int it = 0; it < NTE; it++) acc[jt][it] = dbl4{0.0, 0.0, 0.0, 0.0};
  for (int g = 0; g < npf; g++) {
    const double* Rg = rec + REC * g;
    double u[4][NTE], w[4][NTN];
    L.eval(Rg, u);
    LN.eval(Rg + 12, w);
    const double* M = Rg + 24 + 4 * p;
    const double m0 = M[0], m1 = M[1], m2 = M[2], m3 = M[3];
    double av[NTN], bv[NTE];
#pragma unroll
    for (int t = 0; t < NTN; t++) av[t] = w0 * w[0][t] + w1 * w[1][t] + w2 * w[2][t] + w3 * w[3][t];
#pragma unroll
    for (int t = 0; t < NTE; t++) bv[t] = m0 * u[0][t] + m1 * u[1][t] + m2 * u[2][t] + m3 * u[3][t];
#pragma unroll
    for (int jt = 0; jt < NTN; jt++)
#pragma unroll
      for (int it = 0; it < NTE; it++)
        acc[jt][it] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[jt], bv[it], acc[jt][it], 0, 0, 0);
  }
  double* out = R.blk + (R.soff ? R.soff[6 * (size_t)e + 1 + s] : ((size_t)e * 6 + 1 + s) * NSE * NSE);
#pragma unroll
  for (int jt = 0; jt < NTN; jt++)
#pragma unroll
    for (int it = 0; it < NTE; it++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int j = 16 * jt + 4 * r + p, i = 16 * it + c16;
        if (j < NSN && i < NSE) out[j * NSE + i] = acc[jt][it][r];
      }
}

// Small blocks on the vector ALUs.  With few shape functions a 16x16 MFMA
// tile holds mostly padding and a wave per element leaves most lanes idle
// while the records are built, so for these degrees a wave holds 64/NS
// elements (or interior sides) and lane (element, i) owns test function i,
// i.e. column i of the block (rows j over the trial functions) in registers.
// Points go in rounds: the element's lanes write one record each (the same
// records as the MFMA kernels), then the values (phi, d/deta phi) of their
// own function at every point of the round to LDS, and every lane contracts
// M with its test function and the trial functions' values.

// largest block size (shape functions) the vector-ALU kernels are built for
constexpr int kValuMaxNs = 18;

// values of the shape function with exponent code `code` (and its local
// derivatives) from a record's power table
__device__ __forceinline__ void shape_from_powers(const double* P, int code, double* u) {
  const int a = code & 3, b = (code >> 2) & 3, c = (code >> 4) & 3;
  const double pa = P[a], pb = P[4 + b], pc = P[8 + c];
  const double qa = a * P[a > 0 ? a - 1 : 0], qb = b * P[4 + (b > 0 ? b - 1 : 0)],
               qc = c * P[8 + (c > 0 ? c - 1 : 0)];
  u[0] = pa * pb * pc;
  u[1] = qa * pb * pc;
  u[2] = pa * qb * pc;
  u[3] = pa * pb * qc;
}

template <int PX, int PZ, bool MIXED, int TET = 0>
__global__ void __launch_bounds__(64) k_dg_diag_v(DgRowArgs R, int cnt) {
  using S = DgEl<PX, PZ, TET>;
  using Q = DgRowShape<PX, PZ>;
  constexpr int NS = S::NS, NG = S::NG, DE = S::DE, REC = 33, NSD = S::NSIDES;  // 32 + bank pad
  constexpr int EPW = 64 / NS;
  // values of the element's functions at the round's points, element-major
  // with an odd stride between elements (bank spread of the broadcast reads)
  constexpr int UVS = NS * NS * 4 + 1;
  __shared__ double rec[EPW][NS][REC];
  __shared__ double uv[EPW * UVS];
  __shared__ double sX[EPW][18];
  __shared__ int sex[NS];
  __shared__ int fq[EPW][5], fst[EPW][6];
  __shared__ double fpen[EPW][5];
  __shared__ int snp[EPW];
  const DgElemArgs& A = R.E;
  const int lane = threadIdx.x, le = lane / NS, i = lane % NS;
  const int t = blockIdx.x * EPW + le;
  const bool act = le < EPW && t < cnt;
  const int e = act ? (R.elems ? R.elems[t] : t) : 0;
  if (lane < NS) sex[lane] = R.ex[lane];
  if (act)
    for (int c = i; c < 18; c += NS) sX[le][c] = A.nc[(size_t)c * A.nel + e];
  // each face's rule and penalty / h_F (weakform_dg.c:508-531); an element
  // may have fewer lanes (NS) than sides
  for (int sd = i; MIXED && act && sd < NSD; sd += NS) {
    const int nb = R.fnb[5 * e + sd];
    const bool tri = TET == 1 || sd < 2;
    int q = DE, pm = pmin_of_deg(DE);
    if (nb >= 0) {
      const int dn = R.deg[nb];
      q = face_rule_deg(R, tri, e, DE, nb, dn);
      pm = min(pm, pmin_of_deg(dn));
    }
    fq[le][sd] = q;
    fst[le][sd + 1] = tri ? R.ngt[q] : R.ngq[q];
    fpen[le][sd] = pm * R.penalty_const / R.fhf[5 * e + sd];
  }
  __syncthreads();
  if (lane < EPW) {
    int np = 0;
    if (blockIdx.x * EPW + lane < cnt) {
      if (MIXED) {
        fst[lane][0] = NG;
        for (int sd = 0; sd < NSD; sd++) fst[lane][sd + 1] += fst[lane][sd];
        np = fst[lane][NSD];
      } else {
        np = Q::NPTS;
      }
    }
    snp[lane] = np;
  }
  __syncthreads();
  int nmax = 0;
#pragma unroll
  for (int k = 0; k < EPW; k++) nmax = max(nmax, snp[k]);
  const int npts = act ? snp[le] : 0;
  const int code = sex[i];
  double K[NS];
#pragma unroll
  for (int j = 0; j < NS; j++) K[j] = 0.0;
  double f = 0.0;
  for (int c0 = 0; c0 < nmax; c0 += NS) {
    const int k = c0 + i;
    if (act && k < npts)
      diag_record<PX, PZ, MIXED, TET>(R, e, k, sX[le], fq[le], fst[le], fpen[le], rec[le][i]);
    __syncthreads();
    const int nk = min(NS, npts - c0);
    double* uve = uv + le * UVS;
#pragma unroll 1
    for (int g = 0; g < nk; g++) shape_from_powers(rec[le][g], code, uve + (g * NS + i) * 4);
    __syncthreads();
#pragma unroll 1
    for (int g = 0; g < nk; g++) {
      const double* M = rec[le][g] + 12;
      const double* u = uve + (g * NS + i) * 4;
      const double u0 = u[0], u1 = u[1], u2 = u[2], u3 = u[3];
      double tv[4];
#pragma unroll
      for (int a = 0; a < 4; a++)
        tv[a] = M[4 * a] * u0 + M[4 * a + 1] * u1 + M[4 * a + 2] * u2 + M[4 * a + 3] * u3;
      f += M[16] * u0 + M[17] * u1 + M[18] * u2 + M[19] * u3;
#pragma unroll
      for (int j = 0; j < NS; j++) {
        const double* w = uve + (g * NS + j) * 4;
        K[j] += w[0] * tv[0] + w[1] * tv[1] + w[2] * tv[2] + w[3] * tv[3];
      }
    }
    __syncthreads();
  }
  if (!act) return;
  double* out = R.blk + (R.soff ? R.soff[6 * (size_t)e] : (size_t)e * 6 * NS * NS);
#pragma unroll
  for (int j = 0; j < NS; j++) out[j * NS + i] = K[j];
  R.brhs[(R.xoff ? R.xoff[e] : (size_t)e * NS) + i] = f;
}

template <int PXE, int PZE, int PXN, int PZN, int TET_E = 0, int TET_N = 0>
__global__ void __launch_bounds__(64) k_dg_couple_v(DgRowArgs R, const int* slots, int cnt) {
  constexpr int NSE = DgEl<PXE, PZE, TET_E>::NS, NSN = DgEl<PXN, PZN, TET_N>::NS;
  constexpr int DE = deg_of(PXE, PZE, TET_E), DN = deg_of(PXN, PZN, TET_N), REC = 41;  // 40 + bank pad
  constexpr bool TRI = TET_E == 1 || TET_N == 1;  // a face with a tet is a triangle
  constexpr int QE = TRI ? 0 : (PXE + 1) * (PZE + 1), QN = TRI ? 0 : (PXN + 1) * (PZN + 1);
  constexpr int TE = tri_points(PXE), TN = tri_points(PXN);
  constexpr int NPMAX = cmax(cmax(QE, QN), cmax(TE, TN));
  constexpr int FPW = 64 / NSE;  // interior sides per wave
  constexpr int WVS = NPMAX * NSN * 4 + 1;  // the neighbour's trial functions, per side
  __shared__ double rec[FPW][NPMAX][REC];
  __shared__ double wv[FPW * WVS];
  __shared__ double sX[FPW][36];
  __shared__ int sex[NSE], sexn[NSN];
  const DgElemArgs& A = R.E;
  const int lane = threadIdx.x, lf = lane / NSE, i = lane % NSE;
  const int t = blockIdx.x * FPW + lf;
  const bool act = lf < FPW && t < cnt;
  const int slot = act ? slots[t] : 0;
  const int e = slot / 5, s = slot % 5;
  const int nb = act ? R.fnb[slot] : 0, code = act ? R.fcode[slot] : 0;
  const int q = act ? face_rule_deg(R, TRI || s < 2, e, DE, nb, DN) : 0;
  const int npf = act ? (TRI || s < 2 ? R.ngt[q] : R.ngq[q]) : 0;
  if (lane < NSE) sex[lane] = R.ex[lane];
  if (lane < NSN) sexn[lane] = R.exn[lane];
  if (act) {
    for (int c = i; c < 18; c += NSE) sX[lf][c] = A.nc[(size_t)c * A.nel + e];
    for (int c = i; c < 18; c += NSE) sX[lf][18 + c] = A.nc[(size_t)c * A.nel + nb];
  }
  __syncthreads();
  // the side's points over its lanes, in a fixed number of rounds
  constexpr int RR = (NPMAX + NSE - 1) / NSE;
#pragma unroll
  for (int r = 0; r < RR; r++) {
    const int k = i + r * NSE;
    if (k < npf)
      couple_record<PXE, PZE, PXN, PZN, TET_E, TET_N>(R, e, s, nb, code, q, k, sX[lf], rec[lf][k]);
  }
  __syncthreads();
  double* wvf = wv + lf * WVS;
#pragma unroll 1
  for (int k = 0; k < npf; k++)
    for (int j = i; j < NSN; j += NSE) shape_from_powers(rec[lf][k] + 12, sexn[j], wvf + (k * NSN + j) * 4);
  __syncthreads();
  if (!act) return;
  const int cd = sex[i];
  double K[NSN];
#pragma unroll
  for (int j = 0; j < NSN; j++) K[j] = 0.0;
#pragma unroll 1
  for (int k = 0; k < npf; k++) {
    double u[4];
    shape_from_powers(rec[lf][k], cd, u);
    const double* M = rec[lf][k] + 24;
    double tv[4];
#pragma unroll
    for (int a = 0; a < 4; a++)
      tv[a] = M[4 * a] * u[0] + M[4 * a + 1] * u[1] + M[4 * a + 2] * u[2] + M[4 * a + 3] * u[3];
#pragma unroll
    for (int j = 0; j < NSN; j++) {
      const double* w = wvf + (k * NSN + j) * 4;
      K[j] += w[0] * tv[0] + w[1] * tv[1] + w[2] * tv[2] + w[3] * tv[3];
    }
  }
  double* out = R.blk + (R.soff ? R.soff[6 * (size_t)e + 1 + s] : ((size_t)e * 6 + 1 + s) * NSE * NSE);
#pragma unroll
  for (int j = 0; j < NSN; j++) out[j * NSE + i] = K[j];
}

// inverse of each element's own block by Gauss-Jordan elimination with
// partial pivoting (the reference factors these blocks with LAPACK dgetrf for
// its block Jacobi, lad_block/las_block_intf.c:1057).  One wave per element;
// lane r keeps row r of [A | I] in registers.  At step c the unused row with
// the largest |a_rc| (lowest lane on ties) is the pivot: it is scaled and
// broadcast through LDS, every other row subtracts its multiple, and the
// eliminated column is shifted out so that the step's column is always
// element 0 of the left part.  The pivot row of step c is row c of the
// inverse.  A zero pivot sets *bad.
template <int NS>
__global__ void __launch_bounds__(64) k_dg_inv(const int* elems, const long long* soff,
                                               const long long* voff, const double* blk,
                                               double* dinv, int* bad) {
  __shared__ double prow[2 * NS];
  const int lane = threadIdx.x, e = elems ? elems[blockIdx.x] : blockIdx.x;
  const bool row = lane < NS;
  const double* Ae = blk + (soff ? soff[6 * (size_t)e] : (size_t)e * 6 * NS * NS);
  double L[NS], Rr[NS];
#pragma unroll
  for (int j = 0; j < NS; j++) {
    L[j] = row ? Ae[(size_t)j * NS + lane] : 0.0;
    Rr[j] = (lane == j) ? 1.0 : 0.0;
  }
  bool used = !row;
  int mycol = -1;
  for (int c = 0; c < NS; c++) {
    double v = used ? -1.0 : fabs(L[0]);
    int r = lane;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double v2 = __shfl_xor(v, o);
      const int r2 = __shfl_xor(r, o);
      if (v2 > v || (v2 == v && r2 < r)) { v = v2; r = r2; }
    }
    if (!(v > 0.0)) {
      if (lane == 0) bad[0] = 1;
      return;
    }
    if (lane == r) {
      const double inv = 1.0 / L[0];
#pragma unroll
      for (int j = 1; j < NS; j++) prow[j] = L[j] * inv;
#pragma unroll
      for (int j = 0; j < NS; j++) prow[NS + j] = Rr[j] * inv;
      used = true;
      mycol = c;
    }
    __syncthreads();
    if (lane == r) {
#pragma unroll
      for (int j = 1; j < NS; j++) L[j] = prow[j];
#pragma unroll
      for (int j = 0; j < NS; j++) Rr[j] = prow[NS + j];
    } else {
      const double fac = L[0];
#pragma unroll
      for (int j = 1; j < NS; j++) L[j] -= fac * prow[j];
#pragma unroll
      for (int j = 0; j < NS; j++) Rr[j] -= fac * prow[NS + j];
    }
    // drop the eliminated column
#pragma unroll
    for (int j = 0; j + 1 < NS; j++) L[j] = L[j + 1];
    L[NS - 1] = 0.0;
    __syncthreads();
  }
  if (row) {
    double* Di = dinv + (voff ? voff[e] : (size_t)e * NS * NS);
#pragma unroll
    for (int j = 0; j < NS; j++) Di[(size_t)j * NS + mycol] = Rr[j];
  }
}

// acc + (or -) sum_k B[k ld + i] x[k] over N columns, unrolled
template <int N, int SIGN>
__device__ __forceinline__ double colsum(double acc, const double* B, int ld, int i,
                                         const double* x) {
#pragma unroll
  for (int k = 0; k < N; k++) {
    if (SIGN > 0) acc += B[(size_t)k * ld + i] * x[k];
    else acc -= B[(size_t)k * ld + i] * x[k];
  }
  return acc;
}

// the neighbour columns of a mixed block row: the isotropic block sizes
// unrolled, any other size by a plain loop
template <int SIGN>
__device__ __forceinline__ double nb_colsum(double acc, const double* B, int ld, int i,
                                            const double* x, int dn) {
  switch (dn) {
    case 0: return colsum<6, SIGN>(acc, B, ld, i, x);
    case 4: return colsum<18, SIGN>(acc, B, ld, i, x);
    case 8: return colsum<40, SIGN>(acc, B, ld, i, x);
    case 9: return colsum<4, SIGN>(acc, B, ld, i, x);
    case 10: return colsum<10, SIGN>(acc, B, ld, i, x);
    case 11: return colsum<20, SIGN>(acc, B, ld, i, x);
  }
  const int n = ns_of_deg(dn);
  for (int k = 0; k < n; k++) {
    if (SIGN > 0) acc += B[(size_t)k * ld + i] * x[k];
    else acc -= B[(size_t)k * ld + i] * x[k];
  }
  return acc;
}

// Layout of a block system for the operator kernels: a uniform problem has
// rows [nel][6][NS*NS] and vectors [nel][NS] (offset arrays null); a mixed
// one has per-element offsets of unknowns (xoff), block-row slots (soff) and
// block inverses (voff), and the degree index of every element
struct DgLayout {
  const int* fnb;
  const int* deg;
  const long long* xoff;
  const long long* soff;
  const long long* voff;
};

// w = scale * Dinv (A x) for the block rows (mode 0, scale = sc[j] when sc,
// else 1; Dinv skipped when dinv is null), or r = Dinv (b - A x) (mode 1),
// for the cnt elements of one degree (elems, or 0..cnt-1 when null).
// 64 / NS elements per wave, lane = (element, row).
template <int NS, bool MIXED>
__global__ void __launch_bounds__(256) k_dg_op(int cnt, const int* elems, DgLayout Y,
                                               const double* blk, const double* dinv,
                                               const double* x, const double* b, const double* sc,
                                               int j, const int* stop, double* w, int mode) {
  if (stop && *stop) return;
  constexpr int EPW = 64 / NS;
  __shared__ double T[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int el = lane / NS, i = lane - el * NS;
  const double s = sc ? sc[j] : 1.0;
  for (int g0 = (blockIdx.x * 4 + wave) * EPW; g0 < cnt; g0 += gridDim.x * 4 * EPW) {
    const int gi = g0 + el;
    const bool act = el < EPW && gi < cnt;
    const int e = act ? (MIXED ? elems[gi] : gi) : 0;
    const size_t xe0 = MIXED ? Y.xoff[e] : (size_t)e * NS;
    double acc = 0.0;
    if (act) {
      const double* Ae = blk + (MIXED ? Y.soff[6 * (size_t)e] : (size_t)e * 6 * NS * NS);
      const double* xe = x + xe0;
      for (int k = 0; k < NS; k++) acc += Ae[(size_t)k * NS + i] * xe[k];
      for (int sd = 0; sd < 5; sd++) {
        const int nb = Y.fnb[5 * (size_t)e + sd];
        if (nb < 0) continue;
        if (MIXED) {
          acc = nb_colsum<1>(acc, blk + Y.soff[6 * (size_t)e + 1 + sd], NS, i, x + Y.xoff[nb],
                             Y.deg[nb]);
        } else {
          const double* B = Ae + (size_t)(1 + sd) * NS * NS;
          const double* xn = x + (size_t)nb * NS;
          for (int k = 0; k < NS; k++) acc += B[(size_t)k * NS + i] * xn[k];
        }
      }
      if (mode == 1) acc = b[xe0 + i] - acc;
    }
    if (!dinv) {
      if (act) w[xe0 + i] = (mode == 1 ? 1.0 : s) * acc;
      continue;
    }
    T[wave][lane] = acc;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    double r = 0.0;
    if (act) {
      const double* Di = dinv + (MIXED ? Y.voff[e] : (size_t)e * NS * NS);
      const double* t = &T[wave][el * NS];
      for (int k = 0; k < NS; k++) r += Di[(size_t)k * NS + i] * t[k];
      w[xe0 + i] = (mode == 1 ? 1.0 : s) * r;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

// one colour of a block Gauss-Seidel sweep (lar_perform_BJ_or_GS_iterations
// for block storage, las_block_intf.c:2269-2460, in colour order):
// x_e = Dinv_e (b_e - sum_s B_{e,s} x_nb(s)) for the colour's elements of
// one degree (b null: zero right-hand side)
template <int NS, bool MIXED>
__global__ void __launch_bounds__(256) k_dg_gs(int cnt, const int* elems, DgLayout Y,
                                               const double* blk, const double* dinv,
                                               const double* b, double* x, const int* stop) {
  if (stop && *stop) return;
  constexpr int EPW = 64 / NS;
  __shared__ double T[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int el = lane / NS, i = lane - el * NS;
  for (int g0 = (blockIdx.x * 4 + wave) * EPW; g0 < cnt; g0 += gridDim.x * 4 * EPW) {
    const int gi = g0 + el;
    const bool act = el < EPW && gi < cnt;
    const int e = act ? elems[gi] : 0;
    const size_t xe0 = MIXED ? Y.xoff[e] : (size_t)e * NS;
    double acc = 0.0;
    if (act) {
      acc = b ? b[xe0 + i] : 0.0;
      const double* Ae = blk + (MIXED ? Y.soff[6 * (size_t)e] : (size_t)e * 6 * NS * NS);
      for (int sd = 0; sd < 5; sd++) {
        const int nb = Y.fnb[5 * (size_t)e + sd];
        if (nb < 0) continue;
        if (MIXED) {
          acc = nb_colsum<-1>(acc, blk + Y.soff[6 * (size_t)e + 1 + sd], NS, i, x + Y.xoff[nb],
                              Y.deg[nb]);
        } else {
          const double* B = Ae + (size_t)(1 + sd) * NS * NS;
          const double* xn = x + (size_t)nb * NS;
          for (int k = 0; k < NS; k++) acc -= B[(size_t)k * NS + i] * xn[k];
        }
      }
    }
    T[wave][lane] = acc;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    double r = 0.0;
    if (act) {
      const double* Di = dinv + (MIXED ? Y.voff[e] : (size_t)e * NS * NS);
      const double* t = &T[wave][el * NS];
      for (int k = 0; k < NS; k++) r += Di[(size_t)k * NS + i] * t[k];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    if (act) x[xe0 + i] = r;
  }
}


// ---------------------------------------------------------------------
// Block ILU(0) of the block system in element order (lar_fill_preconditioner
// and lar_perform_rhsub for block storage, MULTI_ILU / BLOCK_ILU:
// lad_block/las_block_intf.c:1409-2143 and 3085-3476).  Row i, for its
// lower neighbours k in ascending order: a_ik <- a_ik inv(D_k),
// D_i <- D_i - a_ik a_ki, a_ij <- a_ij - a_ik a_kj for every neighbour
// j > k of i that is also a neighbour of k; then D_i is inverted (LU with
// partial pivoting and the inverse from the factors, LAPACK's dgetf2 and
// dgetri order, as R calls dgetrf + dgetri).  Rows of one level of the
// schedule depend only on rows of earlier levels.
// ---------------------------------------------------------------------
struct DgIluLay {
  const int* fnb;
  const int* fcode;
  const int* deg;
  int ns;
  const long long* xoff;
  const long long* soff;
  const long long* voff;
  __device__ int nse(int e) const { return deg ? ns_of_deg(deg[e]) : ns; }
  __device__ long long xo(int e) const { return xoff ? xoff[e] : (long long)e * ns; }
  __device__ long long so(int e, int s) const {
    return soff ? soff[6 * (size_t)e + s] : ((long long)e * 6 + s) * ns * ns;
  }
  __device__ long long vo(int e) const { return voff ? voff[e] : (long long)e * ns * ns; }
};

// sides of element i towards neighbours below (lower) or above i, ordered by
// neighbour index
__device__ __forceinline__ int dg_sorted_sides(const int* fnb, int i, bool lower, int* side) {
  int n = 0;
  for (int s = 0; s < 5; s++) {
    const int nb = fnb[5 * (size_t)i + s];
    if (nb < 0 || (lower ? nb >= i : nb <= i)) continue;
    int t = n++;
    while (t > 0 && nb < fnb[5 * (size_t)i + side[t - 1]]) {
      side[t] = side[t - 1];
      t--;
    }
    side[t] = s;
  }
  return n;
}

constexpr int kIluMax = 40;  // largest block (p = 3 in both directions)

// one row of the factorisation per 64-lane block
__global__ void __launch_bounds__(64) k_dg_ilu_row(int cnt, const int* rows, DgIluLay Y,
                                                   double* F, double* Finv, int* bad) {
  __shared__ double D[kIluMax * kIluMax];
  __shared__ double T[kIluMax * kIluMax];
  __shared__ int piv[kIluMax];
  const int lane = threadIdx.x;
  const int i = rows[blockIdx.x];
  const int ni = Y.nse(i);
  const double* Ai = F + Y.so(i, 0);
  for (int q = lane; q < ni * ni; q += 64) D[q] = Ai[q];
  int side[5];
  const int nl = dg_sorted_sides(Y.fnb, i, true, side);
  __syncthreads();
  for (int t = 0; t < nl; t++) {
    const int s = side[t], k = Y.fnb[5 * (size_t)i + s], nk = Y.nse(k);
    double* Aik = F + Y.so(i, 1 + s);
    const double* Dk = Finv + Y.vo(k);
    // a_ik inv(D_k), sums from 0 over the inner index
    for (int q = lane; q < ni * nk; q += 64) {
      const int r = q % ni, c = q / ni;
      double d = 0.0;
      for (int l = 0; l < nk; l++) d += Aik[r + (size_t)ni * l] * Dk[l + (size_t)nk * c];
      T[q] = d;
    }
    __syncthreads();
    for (int q = lane; q < ni * nk; q += 64) Aik[q] = T[q];
    // D_i -= a_ik a_ki (a_ki: row k's block across the side facing i)
    const int sk = Y.fcode[5 * (size_t)i + s] & 7;  // bits 0-2: the neighbour's side
    const double* Aki = F + Y.so(k, 1 + sk);
    for (int q = lane; q < ni * ni; q += 64) {
      const int r = q % ni, c = q / ni;
      double d = 0.0;
      for (int l = 0; l < nk; l++) d += T[r + ni * l] * Aki[l + (size_t)nk * c];
      D[q] -= d;
    }
    // a_ij -= a_ik a_kj for neighbours j > k of both
    for (int s2 = 0; s2 < 5; s2++) {
      const int j = Y.fnb[5 * (size_t)i + s2];
      if (j <= k) continue;
      int t2 = -1;
      for (int u = 0; u < 5; u++)
        if (Y.fnb[5 * (size_t)k + u] == j) t2 = u;
      if (t2 < 0) continue;
      const int nj = Y.nse(j);
      double* Aij = F + Y.so(i, 1 + s2);
      const double* Akj = F + Y.so(k, 1 + t2);
      for (int q = lane; q < ni * nj; q += 64) {
        const int r = q % ni, c = q / ni;
        double d = 0.0;
        for (int l = 0; l < nk; l++) d += T[r + ni * l] * Akj[l + (size_t)nk * c];
        Aij[q] -= d;
      }
    }
    __syncthreads();
  }
  // LU of D with partial pivoting (dgetf2): lane = row
  for (int j = 0; j < ni; j++) {
    double v = (lane >= j && lane < ni) ? fabs(D[lane + ni * j]) : -1.0;
    int r = lane;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const double v2 = __shfl_xor(v, o);
      const int r2 = __shfl_xor(r, o);
      if (v2 > v || (v2 == v && r2 < r)) { v = v2; r = r2; }
    }
    if (!(v > 0.0)) {
      if (lane == 0) bad[0] = 1;
      return;
    }
    if (lane == 0) piv[j] = r;
    if (r != j && lane < ni) {  // swap rows j and r, lane = column
      const double a = D[j + ni * lane];
      D[j + ni * lane] = D[r + ni * lane];
      D[r + ni * lane] = a;
    }
    __syncthreads();
    const double rp = 1.0 / D[j + ni * j];
    if (lane > j && lane < ni) D[lane + ni * j] *= rp;
    __syncthreads();
    if (lane > j && lane < ni) {
      const double lij = D[lane + ni * j];
      for (int c = j + 1; c < ni; c++) D[lane + ni * c] += lij * -D[j + ni * c];
    }
    __syncthreads();
  }
  // inv(U) column by column (dtrti2), lane = row
  for (int j = 0; j < ni; j++) {
    double x = lane < j ? D[lane + ni * j] : 0.0;
    const double ujj = 1.0 / D[j + ni * j];
    for (int jj = 0; jj < j; jj++) {
      const double t = __shfl(x, jj);
      if (t != 0.0) {
        if (lane < jj) x = x + t * D[lane + ni * jj];
        else if (lane == jj) x = x * D[jj + ni * jj];
      }
    }
    __syncthreads();
    if (lane < j) D[lane + ni * j] = -ujj * x;
    if (lane == 0) D[j + ni * j] = ujj;
    __syncthreads();
  }
  // inv(A) = inv(U) inv(L) (dgetri), columns from the last, lane = row
  for (int j = ni - 1; j >= 0; j--) {
    const double w = (lane > j && lane < ni) ? D[lane + ni * j] : 0.0;
    double y = (lane <= j && lane < ni) ? D[lane + ni * j] : 0.0;
    for (int jj = j + 1; jj < ni; jj++) {
      const double t = -1.0 * __shfl(w, jj);
      if (lane < ni) y = y + t * D[lane + ni * jj];
    }
    __syncthreads();
    if (lane < ni) D[lane + ni * j] = y;
    __syncthreads();
  }
  for (int j = ni - 2; j >= 0; j--) {  // column interchanges in reverse order
    const int jp = piv[j];
    if (jp != j && lane < ni) {
      const double a = D[lane + ni * j];
      D[lane + ni * j] = D[lane + ni * jp];
      D[lane + ni * jp] = a;
    }
    __syncthreads();
  }
  double* Di = Finv + Y.vo(i);
  for (int q = lane; q < ni * ni; q += 64) Di[q] = D[q];
}

// The solves run wave tasks: a task is up to 64 / ns rows of one level with
// the same block size ns (rows[first .. first + count)), lane = (row of the
// task, entry of the row).  task.y = count | ns << 8.

// forward (L) solve of one level, in place: t = v_i less the lower
// neighbours' terms in ascending neighbour order, one term at a time
__global__ void __launch_bounds__(256) k_dg_ilu_fwd(int ntask, const int2* tasks, const int* rows,
                                                    DgIluLay Y, const double* F, double* v,
                                                    const int* stop) {
  if (stop && *stop) return;
  const int g = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (g >= ntask) return;
  const int2 tk = tasks[g];
  const int ni = tk.y >> 8, el = lane / ni, r = lane - el * ni;
  if (el >= (tk.y & 255)) return;
  const int i = rows[tk.x + el];
  int side[5];
  const int nl = dg_sorted_sides(Y.fnb, i, true, side);
  const long long xi = Y.xo(i);
  double t = v[xi + r];
  for (int q = 0; q < nl; q++) {
    const int s = side[q], k = Y.fnb[5 * (size_t)i + s], nk = Y.nse(k);
    const double* A = F + Y.so(i, 1 + s) + r;
    const double* xk = v + Y.xo(k);
#pragma unroll 2
    for (int c = 0; c < nk; c++) t -= A[(size_t)ni * c] * xk[c];
  }
  v[xi + r] = t;
}

// backward (U) solve of one level, in place: t = v_i less the upper
// neighbours' terms, then v_i = inv(D_i) t summed from 0
__global__ void __launch_bounds__(256) k_dg_ilu_bwd(int ntask, const int2* tasks, const int* rows,
                                                    DgIluLay Y, const double* F,
                                                    const double* Finv, double* v,
                                                    const int* stop) {
  if (stop && *stop) return;
  __shared__ double Tw[4][64];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int g = blockIdx.x * 4 + wave;
  if (g >= ntask) return;
  const int2 tk = tasks[g];
  const int ni = tk.y >> 8, el = lane / ni, r = lane - el * ni;
  const bool act = el < (tk.y & 255);
  const int i = act ? rows[tk.x + el] : 0;
  const long long xi = act ? Y.xo(i) : 0;
  double t = 0.0;
  if (act) {
    int side[5];
    const int nu = dg_sorted_sides(Y.fnb, i, false, side);
    t = v[xi + r];
    for (int q = 0; q < nu; q++) {
      const int s = side[q], j = Y.fnb[5 * (size_t)i + s], nj = Y.nse(j);
      const double* A = F + Y.so(i, 1 + s) + r;
      const double* xj = v + Y.xo(j);
#pragma unroll 2
      for (int c = 0; c < nj; c++) t -= A[(size_t)ni * c] * xj[c];
    }
  }
  Tw[wave][lane] = t;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (act) {
    const double* Di = Finv + Y.vo(i) + r;
    const double* tw = &Tw[wave][el * ni];
    double y = 0.0;
#pragma unroll 2
    for (int c = 0; c < ni; c++) y += Di[(size_t)ni * c] * tw[c];
    v[xi + r] = y;
  }
}

// w = s_j v and y = w (start of a GS-smoothed operator application)
__global__ void k_dg_scale2(long long n, const double* v, const double* sc, int j, double* w,
                            double* y, const int* stop) {
  if (stop && *stop) return;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double t = sc[j] * v[k];
  w[k] = t;
  y[k] = t;
}

// w -= y
__global__ void k_dg_sub(long long n, const double* y, double* w, const int* stop) {
  if (stop && *stop) return;
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) w[k] -= y[k];
}

}  // namespace

// call F::template run<NS>() for the block size of degree index d
template <class F>
static int with_ns(int d, F& f) {
  switch (ns_of_deg(d)) {
    case 3: return f.template run<3>(d);
    case 4: return f.template run<4>(d);
    case 6: return f.template run<6>(d);
    case 9: return f.template run<9>(d);
    case 10: return f.template run<10>(d);
    case 12: return f.template run<12>(d);
    case 18: return f.template run<18>(d);
    case 20: return f.template run<20>(d);
    case 24: return f.template run<24>(d);
    case 30: return f.template run<30>(d);
    case 40: return f.template run<40>(d);
  }
  set_error("DG: unsupported block size");
  return -1;
}

// GMRES operator over the DG block system (precon 0: none, 1: block Jacobi,
// 2: block Gauss-Seidel smoothing with gs_sweeps multicolour sweeps, the
// smoothing branch of compreres, lss_mkb_core.c:2001-2054)
struct DgOp : ExtOp {
  DgProblem* D = nullptr;
  int precon = 1;
  DgLayout layout() const {
    return DgLayout{D->d_fnb, D->mixed ? D->d_deg : nullptr, D->mixed ? D->d_xoff : nullptr,
                    D->mixed ? D->d_soff : nullptr, D->mixed ? D->d_voff : nullptr};
  }
  int grid(int cnt, int ns) const {
    const int epw = 64 / ns;
    const long long waves = (cnt + epw - 1) / epw;
    return (int)std::max<long long>(1, std::min<long long>((waves + 3) / 4, (long long)D->ncu * 16));
  }
  // arguments of the current operator or sweep launch, read by run<NS>
  const double *x_ = nullptr, *b_ = nullptr, *sc_ = nullptr;
  const int* stop_ = nullptr;
  double* w_ = nullptr;
  int j_ = 0, mode_ = 0, color_ = -1;
  template <int NS>
  int run(int d) {
    if (color_ < 0) {  // operator over the elements of degree d
      const int c0 = D->grp_ptr[d], cnt = D->grp_ptr[d + 1] - c0;
      if (D->mixed)
        hipLaunchKernelGGL((k_dg_op<NS, true>), dim3(grid(cnt, NS)), dim3(256), 0, stream(), cnt,
                           (const int*)D->d_grp + c0, layout(), D->d_blk,
                           precon == 1 ? D->d_dinv : nullptr, x_, b_, sc_, j_, stop_, w_, mode_);
      else
        hipLaunchKernelGGL((k_dg_op<NS, false>), dim3(grid(cnt, NS)), dim3(256), 0, stream(), cnt,
                           nullptr, layout(), D->d_blk, precon == 1 ? D->d_dinv : nullptr, x_, b_,
                           sc_, j_, stop_, w_, mode_);
    } else {  // the elements of degree d in colour color_
      const size_t k = kNDeg * (size_t)color_ + d;
      const int c0 = D->gs_cd_ptr[k], cnt = D->gs_cd_ptr[k + 1] - c0;
      if (D->mixed)
        hipLaunchKernelGGL((k_dg_gs<NS, true>), dim3(grid(cnt, NS)), dim3(256), 0, stream(), cnt,
                           (const int*)D->d_gs_elems + c0, layout(), D->d_blk, D->d_dinv, b_, w_,
                           stop_);
      else
        hipLaunchKernelGGL((k_dg_gs<NS, false>), dim3(grid(cnt, NS)), dim3(256), 0, stream(), cnt,
                           (const int*)D->d_gs_elems + c0, layout(), D->d_blk, D->d_dinv, b_, w_,
                           stop_);
    }
    MF_HIP(hipGetLastError());
    return 0;
  }
  int op(const double* x, const double* b, const double* sc, int j, const int* stop, double* w,
         int mode) {
    x_ = x; b_ = b; sc_ = sc; j_ = j; stop_ = stop; w_ = w; mode_ = mode; color_ = -1;
    for (int d = 0; d < kNDeg; d++)
      if (D->grp_ptr[d + 1] > D->grp_ptr[d] && with_ns(d, *this)) return -2;
    return 0;
  }
  int sweeps(const double* b, double* y, const int* stop) {
    const int ncol = (int)D->gs_color_ptr.size() - 1;
    b_ = b; w_ = y; stop_ = stop;
    for (int k = 0; k < D->gs_sweeps; k++)
      for (int c = 0; c < ncol; c++)
        for (int d = 0; d < kNDeg; d++) {
          const size_t q = kNDeg * (size_t)c + d;
          if (D->gs_cd_ptr[q + 1] == D->gs_cd_ptr[q]) continue;
          color_ = c;
          if (with_ns(d, *this)) return -2;
        }
    color_ = -1;
    return 0;
  }
  int run_mg(const double* x, const double* b, const double* sc, int j, const int* stop,
             double* w, int mode);
  int run_all(const double* x, const double* b, const double* sc, int j, const int* stop,
              double* w, int mode) {
    if (precon == 4) return run_mg(x, b, sc, j, stop, w, mode);
    if (precon == 3) {
      // BLOCK_ILU branch of compreres: w = (LU)^-1 (b - A x), or (LU)^-1 s_j A v
      if (op(x, b, sc, j, stop, w, mode)) return -2;
      return dg_ilu_solve(*D, w, stop);
    }
    if (precon != 2) return op(x, b, sc, j, stop, w, mode);
    const long long n = D->ndof;
    const int g = (int)((n + 255) / 256);
    if (mode == 1) {
      // r = GS_k(x; b) - x
      MF_HIP(hipMemcpyAsync(w, x, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
      if (sweeps(b, w, stop)) return -2;
      hipLaunchKernelGGL(k_dg_sub, dim3(g), dim3(256), 0, stream(), n, x, w, stop);
      MF_HIP(hipGetLastError());
      return 0;
    }
    // w = X - GS_k(X; 0), X = s_j v
    hipLaunchKernelGGL(k_dg_scale2, dim3(g), dim3(256), 0, stream(), n, x, sc, j, w, D->d_work,
                       stop);
    if (sweeps(nullptr, D->d_work, stop)) return -2;
    hipLaunchKernelGGL(k_dg_sub, dim3(g), dim3(256), 0, stream(), n, D->d_work, w, stop);
    MF_HIP(hipGetLastError());
    return 0;
  }
  int residual(const double* b, const double* x, double* r) override {
    return run_all(x, b, nullptr, 0, nullptr, r, 1);
  }
  int apply(const double* v, const double* scale, int j, double* w, const int* stop) override {
    return run_all(v, nullptr, scale, j, stop, w, 0);
  }
};

namespace {

void free_dg_solver(DgProblem& D) {
  if (D.P) {
    Problem& P = *D.P;
    void* ptrs[] = {P.d_V, P.d_H, P.d_partials, P.d_stat, P.d_x};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
    if (P.h_stat) (void)hipHostFree(P.h_stat);
    delete D.P;
    D.P = nullptr;
  }
  delete D.op;
  D.op = nullptr;
}

// the launch arguments common to the own-block and coupling kernels
void row_args(const DgProblem& D, int d, DgRowArgs& R) {
  DgElemArgs& A = R.E;
  A.nel = D.nel; A.e0 = 0; A.count = D.nel;
  A.nc = D.d_nc; A.U = D.d_U;
  A.qv = D.mixed ? D.d_qv9 + 4 * (size_t)D.qv_off[d] : D.d_qv;
  A.coef_model = D.coef_model; A.src_model = D.src_model; A.mask = D.mask;
  for (int k = 0; k < NCOEF; k++) A.cc[k] = D.cc[k];
  A.coef = D.d_coef; A.stiff = nullptr; A.rhs = nullptr;
  R.fnb = D.d_fnb; R.fcode = D.d_fcode; R.fhf = D.d_fhf; R.fval = D.d_fval;
  R.qt = D.d_qt; R.qq = D.d_qq;
  for (int k = 0; k < kNDeg; k++) {
    R.qt_off[k] = D.qt_off[k]; R.ngt[k] = D.ngt9[k];
    R.qq_off[k] = D.qq_off[k]; R.ngq[k] = D.ngq9[k];
  }
  R.penalty_const = D.penalty_const;
  R.deg = D.d_deg;
  R.elems = D.mixed ? D.d_grp + D.grp_ptr[d] : nullptr;
  R.xoff = D.mixed ? D.d_xoff : nullptr;
  R.soff = D.mixed ? D.d_soff : nullptr;
  for (int k = 0; k < 40; k++) R.ex[k] = D.ex9[d][k];
  R.blk = D.d_blk; R.brhs = D.d_brhs;
}

template <int PX, int PZ, int TET = 0>
int launch_diag_t(DgProblem& D) {
  constexpr int d = deg_of(PX, PZ, TET);
  const int cnt = D.grp_ptr[d + 1] - D.grp_ptr[d];
  if (cnt == 0) return 0;
  DgRowArgs R;
  row_args(D, d, R);
  constexpr int NS = DgEl<PX, PZ, TET>::NS;
  if constexpr (NS <= kValuMaxNs) {
    if (NS <= D.valu_ns) {
      const dim3 g((cnt + 64 / NS - 1) / (64 / NS));
      if (D.mixed)
        hipLaunchKernelGGL((k_dg_diag_v<PX, PZ, true, TET>), g, dim3(64), 0, stream(), R, cnt);
      else
        hipLaunchKernelGGL((k_dg_diag_v<PX, PZ, false>), g, dim3(64), 0, stream(), R, cnt);
      MF_HIP(hipGetLastError());
      return 0;
    }
  }
  if (D.mixed)
    hipLaunchKernelGGL((k_dg_diag<PX, PZ, true, TET>), dim3(cnt), dim3(64), 0, stream(), R);
  else
    hipLaunchKernelGGL((k_dg_diag<PX, PZ, false>), dim3(cnt), dim3(64), 0, stream(), R);
  MF_HIP(hipGetLastError());
  return 0;
}

template <int PXE, int PZE, int PXN, int PZN, int TET_E = 0, int TET_N = 0>
int launch_couple_t(DgProblem& D) {
  constexpr int de = deg_of(PXE, PZE, TET_E), dn = deg_of(PXN, PZN, TET_N);
  const int s0 = D.slot_ptr[kNDeg * de + dn], cnt = D.slot_ptr[kNDeg * de + dn + 1] - s0;
  if (cnt == 0) return 0;
  DgRowArgs R;
  row_args(D, de, R);
  for (int k = 0; k < 40; k++) R.exn[k] = D.ex9[dn][k];
  constexpr int NSE = DgEl<PXE, PZE, TET_E>::NS, NSN = DgEl<PXN, PZN, TET_N>::NS;
  if constexpr (NSE <= kValuMaxNs && NSN <= kValuMaxNs) {
    if (NSE <= D.valu_ns && NSN <= D.valu_ns) {
      const dim3 g((cnt + 64 / NSE - 1) / (64 / NSE));
      hipLaunchKernelGGL((k_dg_couple_v<PXE, PZE, PXN, PZN, TET_E, TET_N>), g, dim3(64), 0,
                         stream(), R, (const int*)D.d_slots + s0, cnt);
      MF_HIP(hipGetLastError());
      return 0;
    }
  }
  hipLaunchKernelGGL((k_dg_couple<PXE, PZE, PXN, PZN, TET_E, TET_N>), dim3(cnt), dim3(64), 0,
                     stream(), R, (const int*)D.d_slots + s0);
  MF_HIP(hipGetLastError());
  return 0;
}

int launch_rows(DgProblem& D) {
  int rc = 0;
#define MF_DG_CASE(X, Z) \
  if (rc == 0) rc = launch_diag_t<X, Z>(D);
  MF_DG_CASE(1, 1) MF_DG_CASE(2, 1) MF_DG_CASE(3, 1)
  MF_DG_CASE(1, 2) MF_DG_CASE(2, 2) MF_DG_CASE(3, 2)
  MF_DG_CASE(1, 3) MF_DG_CASE(2, 3) MF_DG_CASE(3, 3)
#undef MF_DG_CASE
  // tets of degree 1..3, complete-base prisms of degree 1..3, pz = 0 prisms
  if (rc == 0) rc = launch_diag_t<1, 1, 1>(D);
  if (rc == 0) rc = launch_diag_t<2, 2, 1>(D);
  if (rc == 0) rc = launch_diag_t<3, 3, 1>(D);
  if (rc == 0) rc = launch_diag_t<1, 1, 2>(D);
  if (rc == 0) rc = launch_diag_t<2, 2, 2>(D);
  if (rc == 0) rc = launch_diag_t<3, 3, 2>(D);
  if (rc == 0) rc = launch_diag_t<1, 0>(D);
  if (rc == 0) rc = launch_diag_t<2, 0>(D);
  // couplings between equal degrees, and between the isotropic degrees of
  // mixed problems
#define MF_DG_PAIR(XE, ZE, XN, ZN) \
  if (rc == 0) rc = launch_couple_t<XE, ZE, XN, ZN>(D);
  MF_DG_PAIR(1, 1, 1, 1) MF_DG_PAIR(2, 1, 2, 1) MF_DG_PAIR(3, 1, 3, 1)
  MF_DG_PAIR(1, 2, 1, 2) MF_DG_PAIR(2, 2, 2, 2) MF_DG_PAIR(3, 2, 3, 2)
  MF_DG_PAIR(1, 3, 1, 3) MF_DG_PAIR(2, 3, 2, 3) MF_DG_PAIR(3, 3, 3, 3)
  MF_DG_PAIR(1, 1, 2, 2) MF_DG_PAIR(1, 1, 3, 3) MF_DG_PAIR(2, 2, 1, 1)
  MF_DG_PAIR(2, 2, 3, 3) MF_DG_PAIR(3, 3, 1, 1) MF_DG_PAIR(3, 3, 2, 2)
#undef MF_DG_PAIR
  // faces with a tet (triangles): tet-tet, tet-prism and prism-tet pairs of
  // degrees 1..3 (isotropic prisms)
#define MF_DG_TPAIR(PE, PN)                                                      \
  if (rc == 0) rc = launch_couple_t<PE, PE, PN, PN, 1, 1>(D);                    \
  if (rc == 0) rc = launch_couple_t<PE, PE, PN, PN, 1, 0>(D);                    \
  if (rc == 0) rc = launch_couple_t<PE, PE, PN, PN, 0, 1>(D);                    \
  if (rc == 0) rc = launch_couple_t<PE, PE, PN, PN, 2, 2>(D);
  MF_DG_TPAIR(1, 1) MF_DG_TPAIR(1, 2) MF_DG_TPAIR(1, 3)
  MF_DG_TPAIR(2, 1) MF_DG_TPAIR(2, 2) MF_DG_TPAIR(2, 3)
  MF_DG_TPAIR(3, 1) MF_DG_TPAIR(3, 2) MF_DG_TPAIR(3, 3)
#undef MF_DG_TPAIR
  // pz = 0 prisms among themselves
  if (rc == 0) rc = launch_couple_t<1, 0, 1, 0>(D);
  if (rc == 0) rc = launch_couple_t<1, 0, 2, 0>(D);
  if (rc == 0) rc = launch_couple_t<2, 0, 1, 0>(D);
  if (rc == 0) rc = launch_couple_t<2, 0, 2, 0>(D);
  return rc;
}

template <int NS>
int launch_inv_t(DgProblem& D, int d, int* d_bad) {
  const int cnt = D.grp_ptr[d + 1] - D.grp_ptr[d];
  if (cnt == 0) return 0;
  hipLaunchKernelGGL((k_dg_inv<NS>), dim3(cnt), dim3(64), 0, stream(),
                     D.mixed ? (const int*)D.d_grp + D.grp_ptr[d] : nullptr,
                     (const long long*)D.d_soff, (const long long*)D.d_voff, D.d_blk, D.d_dinv,
                     d_bad);
  MF_HIP(hipGetLastError());
  return 0;
}



// greedy colouring of the element graph (elements sharing a face get
// different colours), elements listed colour by colour
int build_gs_colors(DgProblem& D) {
  if (D.d_gs_elems) return 0;
  const size_t nf = 5 * (size_t)D.nel;
  std::vector<int> fnb(nf);
  MF_HIP(hipMemcpy(fnb.data(), D.d_fnb, nf * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> col(D.nel, -1);
  int ncol = 0;
  for (int e = 0; e < D.nel; e++) {
    unsigned used = 0;
    for (int s = 0; s < 5; s++) {
      const int nb = fnb[5 * (size_t)e + s];
      if (nb >= 0 && col[nb] >= 0 && col[nb] < 32) used |= 1u << col[nb];
    }
    int c = 0;
    while (c < 32 && (used >> c & 1u)) c++;
    col[e] = c;
    ncol = std::max(ncol, c + 1);
  }
  D.gs_color_ptr.assign(ncol + 1, 0);
  for (int e = 0; e < D.nel; e++) D.gs_color_ptr[col[e] + 1]++;
  for (int c = 0; c < ncol; c++) D.gs_color_ptr[c + 1] += D.gs_color_ptr[c];
  // within a colour, elements grouped by degree (ascending element order in each group)
  D.gs_cd_ptr.assign(kNDeg * (size_t)ncol + 1, 0);
  for (int e = 0; e < D.nel; e++) D.gs_cd_ptr[kNDeg * (size_t)col[e] + D.deg[e] + 1]++;
  for (size_t k = 0; k + 1 < D.gs_cd_ptr.size(); k++) D.gs_cd_ptr[k + 1] += D.gs_cd_ptr[k];
  std::vector<int> pos(D.gs_cd_ptr.begin(), D.gs_cd_ptr.end() - 1), elems(D.nel);
  for (int e = 0; e < D.nel; e++) elems[pos[kNDeg * (size_t)col[e] + D.deg[e]]++] = e;
  MF_HIP(hipMalloc(&D.d_gs_elems, (size_t)D.nel * sizeof(int)));
  MF_HIP(hipMemcpy(D.d_gs_elems, elems.data(), (size_t)D.nel * sizeof(int), hipMemcpyHostToDevice));
  if (!D.d_work) MF_HIP(hipMalloc(&D.d_work, (size_t)D.ndof * sizeof(double)));
  return 0;
}

struct InvLaunch {
  DgProblem* D;
  int* d_bad;
  template <int NS>
  int run(int d) { return launch_inv_t<NS>(*D, d, d_bad); }
};

int fill_block_jacobi(DgProblem& D) {
  if (!D.d_dinv) MF_HIP(hipMalloc(&D.d_dinv, (size_t)D.inv_len * sizeof(double)));
  int* d_bad = nullptr;
  MF_HIP(hipMalloc(&d_bad, sizeof(int)));
  MF_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), stream()));
  int rc = 0;
  InvLaunch L{&D, d_bad};
  for (int d = 0; d < kNDeg && rc == 0; d++)
    if (D.grp_ptr[d + 1] > D.grp_ptr[d]) rc = with_ns(d, L);
  int bad = 0;
  if (rc == 0) {
    MF_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, stream()));
    MF_HIP(hipStreamSynchronize(stream()));
  }
  (void)hipFree(d_bad);
  if (rc) return rc;
  if (bad) {
    set_error("DG block Jacobi: a diagonal block is singular");
    return -1;
  }
  D.dinv_valid = true;
  D.bp_valid = false;  // the batch form reads the inverses
  return 0;
}

DgIluLay ilu_layout(const DgProblem& D) {
  return DgIluLay{D.d_fnb, D.d_fcode, D.mixed ? D.d_deg : nullptr, D.mixed ? 0 : D.ns,
                  D.mixed ? D.d_xoff : nullptr, D.mixed ? D.d_soff : nullptr,
                  D.mixed ? D.d_voff : nullptr};
}

// level schedules of the ILU: the forward level of element i is one more than
// the largest level of its lower neighbours (0 without any), the backward
// level likewise over its upper neighbours; rows listed level by level
int build_ilu_levels(DgProblem& D) {
  if (D.d_ilu_frows) return 0;
  const int nel = D.nel;
  std::vector<int> lf(nel, 0), lb(nel, 0);
  int nlf = 0, nlb = 0;
  for (int e = 0; e < nel; e++) {
    int l = 0;
    for (int s = 0; s < 5; s++) {
      const int nb = D.h_fnb[5 * (size_t)e + s];
      if (nb >= 0 && nb < e) l = std::max(l, lf[nb] + 1);
    }
    lf[e] = l;
    nlf = std::max(nlf, l + 1);
  }
  for (int e = nel - 1; e >= 0; e--) {
    int l = 0;
    for (int s = 0; s < 5; s++) {
      const int nb = D.h_fnb[5 * (size_t)e + s];
      if (nb > e) l = std::max(l, lb[nb] + 1);
    }
    lb[e] = l;
    nlb = std::max(nlb, l + 1);
  }
  // rows level by level, by block size within a level; wave tasks of up to
  // 64 / ns rows of one size
  auto nse = [&D](int e) { return D.mixed ? ns_of_deg(D.deg[e]) : D.ns; };
  auto group = [&](const std::vector<int>& lev, int nl, std::vector<int>& ptr,
                   std::vector<int>& rows, std::vector<int>& tptr, std::vector<int2>& tasks) {
    ptr.assign(nl + 1, 0);
    for (int e = 0; e < nel; e++) ptr[lev[e] + 1]++;
    for (int l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
    std::vector<int> pos(ptr.begin(), ptr.end() - 1);
    rows.resize(nel);
    for (int e = 0; e < nel; e++) rows[pos[lev[e]]++] = e;
    tptr.assign(1, 0);
    tasks.clear();
    for (int l = 0; l < nl; l++) {
      std::stable_sort(rows.begin() + ptr[l], rows.begin() + ptr[l + 1],
                       [&](int a, int b) { return nse(a) < nse(b); });
      int q = ptr[l];
      while (q < ptr[l + 1]) {
        const int ns = nse(rows[q]), epw = 64 / ns;
        int c = 0;
        while (c < epw && q + c < ptr[l + 1] && nse(rows[q + c]) == ns) c++;
        tasks.push_back(make_int2(q, c | ns << 8));
        q += c;
      }
      tptr.push_back((int)tasks.size());
    }
  };
  std::vector<int> fr, br;
  std::vector<int2> ft, bt;
  group(lf, nlf, D.ilu_fptr, fr, D.ilu_ftask, ft);
  group(lb, nlb, D.ilu_bptr, br, D.ilu_btask, bt);
  MF_HIP(hipMalloc(&D.d_ilu_frows, (size_t)nel * sizeof(int)));
  MF_HIP(hipMalloc(&D.d_ilu_brows, (size_t)nel * sizeof(int)));
  MF_HIP(hipMalloc(&D.d_ilu_ftasks, ft.size() * sizeof(int2)));
  MF_HIP(hipMalloc(&D.d_ilu_btasks, bt.size() * sizeof(int2)));
  MF_HIP(hipMemcpy(D.d_ilu_frows, fr.data(), (size_t)nel * sizeof(int), hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(D.d_ilu_brows, br.data(), (size_t)nel * sizeof(int), hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(D.d_ilu_ftasks, ft.data(), ft.size() * sizeof(int2), hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(D.d_ilu_btasks, bt.data(), bt.size() * sizeof(int2), hipMemcpyHostToDevice));
  return 0;
}

int fill_block_ilu(DgProblem& D) {
  if (build_ilu_levels(D)) return -2;
  if (!D.d_ilu) MF_HIP(hipMalloc(&D.d_ilu, (size_t)D.blk_len * sizeof(double)));
  if (!D.d_ilu_inv) MF_HIP(hipMalloc(&D.d_ilu_inv, (size_t)D.inv_len * sizeof(double)));
  hipStream_t s = stream();
  MF_HIP(hipMemcpyAsync(D.d_ilu, D.d_blk, (size_t)D.blk_len * sizeof(double),
                        hipMemcpyDeviceToDevice, s));
  int* d_bad = nullptr;
  MF_HIP(hipMalloc(&d_bad, sizeof(int)));
  MF_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), s));
  const DgIluLay Y = ilu_layout(D);
  for (size_t l = 0; l + 1 < D.ilu_fptr.size(); l++) {
    const int c0 = D.ilu_fptr[l], cnt = D.ilu_fptr[l + 1] - c0;
    hipLaunchKernelGGL(k_dg_ilu_row, dim3(cnt), dim3(64), 0, s, cnt,
                       (const int*)D.d_ilu_frows + c0, Y, D.d_ilu, D.d_ilu_inv, d_bad);
  }
  int bad = 0;
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_bad);
  MF_HIP(e);
  if (bad) {
    set_error("DG block ILU: a diagonal block is singular");
    return -1;
  }
  D.ilu_valid = true;
  return 0;
}

// v = (LU)^-1 v in place: the forward levels, then the backward levels
int dg_ilu_solve(DgProblem& D, double* v, const int* stop) {
  hipStream_t s = stream();
  const DgIluLay Y = ilu_layout(D);
  for (size_t l = 0; l + 1 < D.ilu_ftask.size(); l++) {
    const int t0 = D.ilu_ftask[l], nt = D.ilu_ftask[l + 1] - t0;
    hipLaunchKernelGGL(k_dg_ilu_fwd, dim3((nt + 3) / 4), dim3(256), 0, s, nt,
                       (const int2*)D.d_ilu_ftasks + t0, (const int*)D.d_ilu_frows, Y,
                       (const double*)D.d_ilu, v, stop);
  }
  for (size_t l = 0; l + 1 < D.ilu_btask.size(); l++) {
    const int t0 = D.ilu_btask[l], nt = D.ilu_btask[l + 1] - t0;
    hipLaunchKernelGGL(k_dg_ilu_bwd, dim3((nt + 3) / 4), dim3(256), 0, s, nt,
                       (const int2*)D.d_ilu_btasks + t0, (const int*)D.d_ilu_brows, Y,
                       (const double*)D.d_ilu, (const double*)D.d_ilu_inv, v, stop);
  }
  MF_HIP(hipGetLastError());
  return 0;
}

}  // namespace

// ---------------------------------------------------------------------
// p-multigrid V-cycle (the V-cycle of lsr_mkb_core_vcycle,
// lss_mkb_core.c:1288-1470, over levels of decreasing isotropic degree on
// the same mesh instead of the reference's mesh levels): pre-smoothing by
// block Gauss-Seidel sweeps, residual, restriction to the next lower
// degree, the V-cycle there for the correction from zero, prolongation and
// addition, post-smoothing; on degree 1 block-ILU(0) corrections
// (lsr_mkb_core_solve_coarse's role).  The tensor bases are nested (the
// functions of (p-1, p-1) are functions of (p, p)), so prolongation is the
// injection of coefficients, restriction its transpose (the coarse test
// functions' residual entries) and the coarse operators the Galerkin
// products P^T A P: sub-blocks of the finer level's blocks.
struct DgMg {
  std::vector<int> ids;          // handles of levels 0 (degree 1) .. L-2
  std::vector<DgProblem*> lev;   // levels 0 .. L-1 (L-1: the fine problem)
  std::vector<DgOp> ops;         // residual / smoothing of every level
  std::vector<int*> d_map;       // level l's functions in level l+1's base
  std::vector<double*> b, e, t;  // per level: rhs, correction, residual
  double* zero = nullptr;        // a zero rhs of the fine size
  int nu1 = 2, nu2 = 2, coarse = 2;
  ~DgMg() {
    for (int* p : d_map) if (p) (void)hipFree(p);
    for (auto* v : {&b, &e, &t})
      for (double* p : *v) if (p) (void)hipFree(p);
    if (zero) (void)hipFree(zero);
  }
};

namespace {

__global__ void k_mg_gather_blocks(long long cnt, int nc, int nf, const int* map, const double* bf,
                                   double* bc) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt) return;
  const long long nn = (long long)nc * nc, blk = k / nn;
  const int r = (int)(k % nn), j = r / nc, i = r % nc;
  bc[k] = bf[blk * nf * nf + (long long)map[j] * nf + map[i]];
}
__global__ void k_mg_restrict(long long cnt, int nc, int nf, const int* map, const double* rf,
                              double* bc, double* ec, const int* stop) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt || (stop && *stop)) return;
  bc[k] = rf[(k / nc) * nf + map[k % nc]];
  ec[k] = 0.0;
}
__global__ void k_mg_prolong_add(long long cnt, int nc, int nf, const int* map, const double* ec,
                                 double* yf, const int* stop) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= cnt || (stop && *stop)) return;
  yf[(k / nc) * nf + map[k % nc]] += ec[k];
}
__global__ void k_mg_add(long long n, const double* t, double* y, const int* stop) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n || (stop && *stop)) return;
  y[k] += t[k];
}

unsigned blocks_of(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// the Galerkin blocks, block inverses, colourings and coarse ILU factors of
// every level below the fine one
int mg_refresh(DgProblem& D) {
  DgMg& M = *D.mg;
  const int L = (int)M.lev.size();
  for (int l = L - 2; l >= 0; l--) {
    DgProblem& C = *M.lev[l];
    const DgProblem& F = *M.lev[l + 1];
    const long long cnt = (long long)C.nel * 6 * C.ns * C.ns;
    if (!C.d_blk) MF_HIP(hipMalloc(&C.d_blk, cnt * sizeof(double)));
    if (!C.d_brhs) MF_HIP(hipMalloc(&C.d_brhs, (size_t)C.ndof * sizeof(double)));
    hipLaunchKernelGGL(k_mg_gather_blocks, dim3(blocks_of(cnt)), dim3(256), 0, stream(), cnt, C.ns,
                       F.ns, M.d_map[l], F.d_blk, C.d_blk);
    MF_HIP(hipGetLastError());
    C.assembled = true;
    C.dinv_valid = C.ilu_valid = false;
  }
  for (int l = 0; l < L; l++) {
    DgProblem& C = *M.lev[l];
    if (l == 0) {
      if (!C.ilu_valid && fill_block_ilu(C)) return -2;
    } else {
      if (!C.dinv_valid && fill_block_jacobi(C)) return -2;
      if (build_gs_colors(C)) return -2;
    }
  }
  D.mg_valid = true;
  return 0;
}

// one V-cycle from level l on y (in/out) with rhs b (null: zero)
int mg_vcycle(DgMg& M, int l, const double* b, double* y, const int* stop) {
  DgProblem& C = *M.lev[l];
  DgOp& O = M.ops[l];
  const long long n = C.ndof;
  if (!b) b = M.zero;
  if (l == 0) {
    for (int k = 0; k < M.coarse; k++) {
      if (O.op(y, b, nullptr, 0, stop, M.t[0], 1)) return -2;
      if (dg_ilu_solve(C, M.t[0], stop)) return -2;
      hipLaunchKernelGGL(k_mg_add, dim3(blocks_of(n)), dim3(256), 0, stream(), n, M.t[0], y, stop);
    }
    MF_HIP(hipGetLastError());
    return 0;
  }
  C.gs_sweeps = M.nu1;
  if (M.nu1 > 0 && O.sweeps(b, y, stop)) return -2;
  if (O.op(y, b, nullptr, 0, stop, M.t[l], 1)) return -2;
  DgProblem& Q = *M.lev[l - 1];
  const long long cnt = (long long)Q.nel * Q.ns;
  hipLaunchKernelGGL(k_mg_restrict, dim3(blocks_of(cnt)), dim3(256), 0, stream(), cnt, Q.ns, C.ns,
                     M.d_map[l - 1], M.t[l], M.b[l - 1], M.e[l - 1], stop);
  MF_HIP(hipGetLastError());
  if (mg_vcycle(M, l - 1, M.b[l - 1], M.e[l - 1], stop)) return -2;
  hipLaunchKernelGGL(k_mg_prolong_add, dim3(blocks_of(cnt)), dim3(256), 0, stream(), cnt, Q.ns,
                     C.ns, M.d_map[l - 1], M.e[l - 1], y, stop);
  MF_HIP(hipGetLastError());
  C.gs_sweeps = M.nu2;
  if (M.nu2 > 0 && O.sweeps(b, y, stop)) return -2;
  return 0;
}

int DgOp::run_mg(const double* x, const double* b, const double* sc, int j, const int* stop,
                 double* w, int mode) {
  DgMg& M = *D->mg;
  const int L = (int)M.lev.size();
  const long long n = D->ndof;
  const int g = (int)((n + 255) / 256);
  if (mode == 1) {
    // r = MG(x; b) - x
    MF_HIP(hipMemcpyAsync(w, x, n * sizeof(double), hipMemcpyDeviceToDevice, stream()));
    if (mg_vcycle(M, L - 1, b, w, stop)) return -2;
    hipLaunchKernelGGL(k_dg_sub, dim3(g), dim3(256), 0, stream(), n, x, w, stop);
    MF_HIP(hipGetLastError());
    return 0;
  }
  // w = X - MG(X; 0), X = s_j v
  hipLaunchKernelGGL(k_dg_scale2, dim3(g), dim3(256), 0, stream(), n, x, sc, j, w, D->d_work, stop);
  if (mg_vcycle(M, L - 1, nullptr, D->d_work, stop)) return -2;
  hipLaunchKernelGGL(k_dg_sub, dim3(g), dim3(256), 0, stream(), n, D->d_work, w, stop);
  MF_HIP(hipGetLastError());
  return 0;
}

namespace {
void free_dg_mg(DgProblem& D) {
  if (!D.mg) return;
  for (int id : D.mg->ids) {
    auto it = g_dg.find(id);
    if (it == g_dg.end()) continue;
    free_dg(*it->second);
    g_dg.erase(it);
  }
  delete D.mg;
  D.mg = nullptr;
  D.mg_valid = false;
}
}  // namespace

}  // namespace mf

using namespace mf;

extern "C" {

}  // extern "C"

namespace {

// a DG problem on a prism, tet or hybrid mesh with the given degree (pdeg =
// 100 pz + px; a tet takes the complete base of degree px = pz) of every
// element; mixed (variable block sizes) when they differ or tets are present
int dg_create(int mesh, const std::vector<int>& pdeg) {
  Mesh* m = mesh_by_id(mesh);
  if (!m) return -3;
  if (!stream()) {
    set_error("mf_init has not been called");
    return -1;
  }
  bool any_tet = false, any_new = false;
  for (int e = 0; e < m->nel; e++) {
    any_new = any_new || pdeg[e] > kDgComplete || pdeg[e] < 100;
    if (m->type[e] == kTetra) {
      any_tet = true;
      if (pdeg[e] > kDgComplete || pdeg[e] % 100 != pdeg[e] / 100) {
        set_error("mf_dg_create: a tetrahedron takes an isotropic degree (101, 202 or 303)");
        return -1;
      }
    } else if (m->type[e] != kPrism) {
      set_error("mf_dg_create: DG elements are prisms or tetrahedra");
      return -1;
    }
  }
  const size_t nel = m->nel;
  // the complete base and the pz = 0 degrees always take the per-degree
  // (mixed) layout and rules
  bool mixed = any_tet || any_new;
  for (size_t e = 1; e < nel; e++) mixed = mixed || pdeg[e] != pdeg[0];
  auto D = std::make_unique<DgProblem>();
  D->mesh = mesh; D->nel = m->nel; D->mixed = mixed;
  D->deg.resize(nel);
  for (size_t e = 0; e < nel; e++)
    D->deg[e] = m->type[e] == kTetra     ? tet_deg_index(pdeg[e] % 100)
                : pdeg[e] > kDgComplete ? complete_deg_index(pdeg[e] - kDgComplete)
                                        : deg_index(pdeg[e] % 100, pdeg[e] / 100);
  if (!mixed) {
    D->px = pdeg[0] % 100; D->pz = pdeg[0] / 100;
    D->ns = num_shape(D->px, D->pz);
  } else {
    D->px = D->pz = 0;  // no single degree: element-matrix entry points are off
    D->ns = 0;
  }
  int dev = 0;
  MF_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  MF_HIP(hipGetDeviceProperties(&prop, dev));
  D->ncu = prop.multiProcessorCount;
  std::vector<double> nc(18 * nel, 0.0);
  for (size_t e = 0; e < nel; e++)
    for (int i = 0; i < (m->type[e] == kTetra ? 4 : 6); i++)
      for (int r = 0; r < 3; r++)
        nc[(size_t)(3 * i + r) * nel + e] = m->coor[3 * (size_t)m->nodes[6 * e + i] + r];
  MF_HIP(hipMalloc(&D->d_nc, nc.size() * sizeof(double)));
  MF_HIP(hipMemcpyAsync(D->d_nc, nc.data(), nc.size() * sizeof(double), hipMemcpyHostToDevice,
                        stream()));
  // degree groups, offsets of unknowns / block-row slots / inverses
  D->grp_ptr.assign(kNDeg + 1, 0);
  for (size_t e = 0; e < nel; e++) D->grp_ptr[D->deg[e] + 1]++;
  for (int d = 0; d < kNDeg; d++) D->grp_ptr[d + 1] += D->grp_ptr[d];
  std::vector<int> grp(nel), pos(D->grp_ptr.begin(), D->grp_ptr.end() - 1);
  for (size_t e = 0; e < nel; e++) grp[pos[D->deg[e]]++] = (int)e;
  // faces
  std::vector<int> fnb, fcode;
  std::vector<double> fhf;
  if (build_faces(*m, fnb, fcode, fhf)) {
    free_dg(*D);
    return -1;
  }
  const size_t nf = 5 * nel;
  D->h_fnb = fnb;
  if (mixed) {
    // unknowns in element order; block rows and inverses stored degree
    // group by group, so that a launch over one degree streams its blocks
    D->xoff.assign(nel + 1, 0);
    D->voff.assign(nel + 1, 0);
    D->soff.assign(6 * nel + 1, 0);
    for (size_t e = 0; e < nel; e++) D->xoff[e + 1] = D->xoff[e] + ns_of_deg(D->deg[e]);
    long long bpos = 0, vpos = 0;
    for (size_t g = 0; g < nel; g++) {
      const int e = grp[g];
      const long long ns = ns_of_deg(D->deg[e]);
      D->voff[e] = vpos;
      vpos += ns * ns;
      D->soff[6 * (size_t)e] = bpos;
      bpos += ns * ns;
      for (int sd = 0; sd < 5; sd++) {
        const int nb = fnb[5 * (size_t)e + sd];
        D->soff[6 * (size_t)e + 1 + sd] = bpos;
        bpos += ns * (nb >= 0 ? ns_of_deg(D->deg[nb]) : ns);
      }
    }
    D->ndof = D->xoff[nel];
    D->blk_len = bpos;
    D->inv_len = vpos;
    D->soff[6 * nel] = bpos;
    D->voff[nel] = vpos;
    MF_HIP(hipMalloc(&D->d_xoff, D->xoff.size() * sizeof(long long)));
    MF_HIP(hipMalloc(&D->d_soff, D->soff.size() * sizeof(long long)));
    MF_HIP(hipMalloc(&D->d_voff, D->voff.size() * sizeof(long long)));
    MF_HIP(hipMalloc(&D->d_grp, nel * sizeof(int)));
    MF_HIP(hipMemcpyAsync(D->d_xoff, D->xoff.data(), D->xoff.size() * sizeof(long long),
                          hipMemcpyHostToDevice, stream()));
    MF_HIP(hipMemcpyAsync(D->d_soff, D->soff.data(), D->soff.size() * sizeof(long long),
                          hipMemcpyHostToDevice, stream()));
    MF_HIP(hipMemcpyAsync(D->d_voff, D->voff.data(), D->voff.size() * sizeof(long long),
                          hipMemcpyHostToDevice, stream()));
    MF_HIP(hipMemcpyAsync(D->d_grp, grp.data(), nel * sizeof(int), hipMemcpyHostToDevice,
                          stream()));
  } else {
    const long long ns = D->ns;
    D->ndof = (long long)nel * ns;
    D->blk_len = (long long)nel * 6 * ns * ns;
    D->inv_len = (long long)nel * ns * ns;
  }
  MF_HIP(hipMalloc(&D->d_deg, nel * sizeof(int)));
  MF_HIP(hipMemcpyAsync(D->d_deg, D->deg.data(), nel * sizeof(int), hipMemcpyHostToDevice,
                        stream()));
  // volume rules: the table of the single degree, or every degree present
  if (!mixed) {
    std::vector<double> qv;
    D->ng = volume_rule(D->px, D->pz, qv);
    D->nsp = (D->ns + 15) / 16 * 16;
    MF_HIP(hipMalloc(&D->d_qv, qv.size() * sizeof(double)));
    MF_HIP(hipMalloc(&D->d_U, (size_t)D->ng * 4 * D->nsp * sizeof(double)));
    MF_HIP(hipMemcpyAsync(D->d_qv, qv.data(), qv.size() * sizeof(double), hipMemcpyHostToDevice,
                          stream()));
    hipLaunchKernelGGL(k_dg_table, dim3(1), dim3(64), 0, stream(), D->px, D->pz, D->ng, D->nsp,
                       D->d_qv, D->d_U);
    MF_HIP(hipGetLastError());
  } else {
    std::vector<double> all, qv;
    for (int d = 0; d < kNDeg; d++) {
      if (D->grp_ptr[d + 1] == D->grp_ptr[d]) continue;
      D->qv_off[d] = (int)all.size() / 4;
      if (d >= 9 && d < 12) tet_volume_rule(d - 8, qv);
      else volume_rule(rule_px(d), rule_pz(d), qv);
      all.insert(all.end(), qv.begin(), qv.end());
    }
    MF_HIP(hipMalloc(&D->d_qv9, all.size() * sizeof(double)));
    MF_HIP(hipMemcpyAsync(D->d_qv9, all.data(), all.size() * sizeof(double),
                          hipMemcpyHostToDevice, stream()));
  }
  // face rules and shape-function exponents of all nine degrees
  std::vector<double> qt_all, qq_all, qt, qq;
  for (int d = 0; d < 9; d++) {
    face_rules(d % 3 + 1, d / 3 + 1, qt, qq);
    D->qt_off[d] = (int)qt_all.size() / 3;
    D->ngt9[d] = (int)qt.size() / 3;
    D->qq_off[d] = (int)qq_all.size() / 3;
    D->ngq9[d] = (int)qq.size() / 3;
    qt_all.insert(qt_all.end(), qt.begin(), qt.end());
    qq_all.insert(qq_all.end(), qq.begin(), qq.end());
    shape_exponents(d % 3 + 1, d / 3 + 1, D->ex9[d]);
  }
  // tets: the triangle rule of the isotropic prism degree, no quads
  for (int p = 1; p <= 3; p++) {
    const int d = tet_deg_index(p), dp = deg_index(p, p);
    D->qt_off[d] = D->qt_off[dp];
    D->ngt9[d] = D->ngt9[dp];
    D->qq_off[d] = 0;
    D->ngq9[d] = 0;
    tet_exponents(p, D->ex9[d]);
  }
  // complete-base prisms: the rules of the tensor degree (p, p)
  for (int p = 1; p <= 3; p++) {
    const int d = complete_deg_index(p), dp = deg_index(p, p);
    D->qt_off[d] = D->qt_off[dp];
    D->ngt9[d] = D->ngt9[dp];
    D->qq_off[d] = D->qq_off[dp];
    D->ngq9[d] = D->ngq9[dp];
    complete_exponents(p, D->ex9[d]);
  }
  // pz = 0: one Gauss point in z on the quads
  for (int px = 1; px <= 2; px++) {
    const int d = deg_index(px, 0);
    face_rules(px, 0, qt, qq);
    D->qt_off[d] = (int)qt_all.size() / 3;
    D->ngt9[d] = (int)qt.size() / 3;
    D->qq_off[d] = (int)qq_all.size() / 3;
    D->ngq9[d] = (int)qq.size() / 3;
    qt_all.insert(qt_all.end(), qt.begin(), qt.end());
    qq_all.insert(qq_all.end(), qq.begin(), qq.end());
    flat_exponents(px, D->ex9[d]);
  }
  if (!mixed) {
    const int d = D->deg[0];
    D->ngt = D->ngt9[d];
    D->ngq = D->ngq9[d];
    for (int k = 0; k < 40; k++) D->ex[k] = D->ex9[d][k];
  }
  MF_HIP(hipMalloc(&D->d_fnb, nf * sizeof(int)));
  MF_HIP(hipMalloc(&D->d_fcode, nf * sizeof(int)));
  MF_HIP(hipMalloc(&D->d_fhf, nf * sizeof(double)));
  MF_HIP(hipMalloc(&D->d_fval, nf * sizeof(double)));
  MF_HIP(hipMalloc(&D->d_qt, qt_all.size() * sizeof(double)));
  MF_HIP(hipMalloc(&D->d_qq, qq_all.size() * sizeof(double)));
  MF_HIP(hipMemcpyAsync(D->d_fnb, fnb.data(), nf * sizeof(int), hipMemcpyHostToDevice, stream()));
  MF_HIP(hipMemcpyAsync(D->d_fcode, fcode.data(), nf * sizeof(int), hipMemcpyHostToDevice,
                        stream()));
  MF_HIP(hipMemcpyAsync(D->d_fhf, fhf.data(), nf * sizeof(double), hipMemcpyHostToDevice,
                        stream()));
  MF_HIP(hipMemsetAsync(D->d_fval, 0, nf * sizeof(double), stream()));
  MF_HIP(hipMemcpyAsync(D->d_qt, qt_all.data(), qt_all.size() * sizeof(double),
                        hipMemcpyHostToDevice, stream()));
  MF_HIP(hipMemcpyAsync(D->d_qq, qq_all.data(), qq_all.size() * sizeof(double),
                        hipMemcpyHostToDevice, stream()));
  // interior slots grouped by (own degree, neighbour degree), ascending within a pair
  constexpr int NP = kNDeg * kNDeg;
  D->slot_ptr.assign(NP + 1, 0);
  for (size_t k = 0; k < nf; k++)
    if (fnb[k] >= 0) D->slot_ptr[kNDeg * D->deg[k / 5] + D->deg[fnb[k]] + 1]++;
  for (int q = 0; q < NP; q++) D->slot_ptr[q + 1] += D->slot_ptr[q];
  std::vector<int> slots(D->slot_ptr[NP]), spos(D->slot_ptr.begin(), D->slot_ptr.end() - 1);
  for (size_t k = 0; k < nf; k++)
    if (fnb[k] >= 0) slots[spos[kNDeg * D->deg[k / 5] + D->deg[fnb[k]]]++] = (int)k;
  D->nslots = (int)slots.size();
  if (D->nslots) {
    MF_HIP(hipMalloc(&D->d_slots, slots.size() * sizeof(int)));
    MF_HIP(hipMemcpyAsync(D->d_slots, slots.data(), slots.size() * sizeof(int),
                          hipMemcpyHostToDevice, stream()));
  }
  MF_HIP(hipStreamSynchronize(stream()));
  const int id = g_next_dg++;
  g_dg[id] = std::move(D);
  return id;
}

}  // namespace

namespace mf {
namespace {

// ---------------------------------------------------------------------
// DG problems as batch members.  Left-preconditioned GMRES with block
// Jacobi (the reference's precon of DG systems, lss_mkb_core.c:1861-2097 for
// block storage) is plain GMRES on M x = b' with M = Dinv A and b' = Dinv b:
// the block rows of the system scaled by the inverses of their diagonal
// blocks.  M is stored as a scalar SELL-64 matrix in DOF order (row
// (e, i): the own block's columns, then those of the neighbour across each
// interior side in side order), so a DG problem joins a batch like any
// other SELL problem, preconditioner folded in; its measure ||b' - M x|| is
// the block-Jacobi-preconditioned residual of the DG solve.
// ---------------------------------------------------------------------
// M's row (e, i) and b'_(e, i), one thread per row (uniform degree)
__global__ void k_dg_to_sell(long long n, int ns, const int* fnb, const double* blk,
                             const double* dinv, const double* b, const long long* sl_ptr,
                             double* val, double* bp) {
  const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const long long e = r / ns;
  const int i = (int)(r - e * ns);
  const size_t nn = (size_t)ns * ns;
  const double* Di = dinv + e * nn;  // column-major: (i, k) at k ns + i
  const long long base = sl_ptr[r >> 6] + (r & 63);
  int q = 0;
  for (int slot = 0; slot < 6; slot++) {
    if (slot > 0 && fnb[5 * e + slot - 1] < 0) continue;
    const double* Bk = blk + (e * 6 + slot) * nn;
    for (int j = 0; j < ns; j++, q++) {
      double m = 0.0;
      for (int k = 0; k < ns; k++) m += Di[(size_t)k * ns + i] * Bk[(size_t)j * ns + k];
      val[base + 64LL * q] = m;
    }
  }
  double t = 0.0;
  for (int k = 0; k < ns; k++) t += Di[(size_t)k * ns + i] * b[e * ns + k];
  bp[r] = t;
}

// structure of M (host): slice lengths and columns; padding entries point
// at the own row with value zero
int dg_batch_structure(DgProblem& D) {
  const long long n = D.ndof;
  const int ns = D.ns;
  const int nsl = (int)((n + 63) / 64);
  std::vector<long long> sl_ptr(nsl + 1, 0);
  auto row_len = [&](long long r) {
    const long long e = r / ns;
    int k = 1;
    for (int sd = 0; sd < 5; sd++) k += D.h_fnb[5 * e + sd] >= 0;
    return k * ns;
  };
  long long nnz = 0;
  for (int sl = 0; sl < nsl; sl++) {
    int len = 0;
    for (long long r = 64LL * sl; r < std::min(n, 64LL * sl + 64); r++) {
      len = std::max(len, row_len(r));
      nnz += row_len(r);
    }
    sl_ptr[sl + 1] = sl_ptr[sl] + 64LL * len;
  }
  std::vector<int> col(sl_ptr[nsl]);
  for (int sl = 0; sl < nsl; sl++) {
    const int len = (int)((sl_ptr[sl + 1] - sl_ptr[sl]) / 64);
    for (int lane = 0; lane < 64; lane++) {
      const long long r = 64LL * sl + lane;
      const long long own = r < n ? r : n - 1;
      for (int q = 0; q < len; q++) col[sl_ptr[sl] + 64LL * q + lane] = (int)own;
      if (r >= n) continue;
      const long long e = r / ns;
      int q = 0;
      for (int slot = 0; slot < 6; slot++) {
        const long long nb = slot == 0 ? e : D.h_fnb[5 * e + slot - 1];
        if (nb < 0) continue;
        for (int j = 0; j < ns; j++, q++) col[sl_ptr[sl] + 64LL * q + lane] = (int)(nb * ns + j);
      }
    }
  }
  Problem* P = new Problem();
  D.bp = P;
  P->n = (int)n;
  P->el_type = kPrism;
  P->sell = true;
  P->nnz = nnz;
  P->nnz_sell = sl_ptr[nsl];
  MF_HIP(hipMalloc(&P->d_sl_ptr, sl_ptr.size() * sizeof(long long)));
  MF_HIP(hipMemcpy(P->d_sl_ptr, sl_ptr.data(), sl_ptr.size() * sizeof(long long),
                   hipMemcpyHostToDevice));
  MF_HIP(hipMalloc(&P->d_scol, col.size() * sizeof(int)));
  MF_HIP(hipMemcpy(P->d_scol, col.data(), col.size() * sizeof(int), hipMemcpyHostToDevice));
  MF_HIP(hipMalloc(&P->d_val, (size_t)P->nnz_sell * sizeof(double)));
  MF_HIP(hipMemset(P->d_val, 0, (size_t)P->nnz_sell * sizeof(double)));  // padding stays zero
  MF_HIP(hipMalloc(&P->d_rhs, (size_t)n * sizeof(double)));
  return 0;
}

void free_dg_batch(DgProblem& D) {
  if (!D.bp) return;
  Problem& P = *D.bp;
  void* ptrs[] = {P.d_sl_ptr, P.d_scol, P.d_val, P.d_rhs, P.d_val32, P.d_dinv, P.d_dinv32,
                  P.d_scol16, P.d_sbase, P.d_V, P.d_H, P.d_partials, P.d_stat, P.d_x,
                  P.d_work, P.d_work2};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  if (P.h_stat) (void)hipHostFree(P.h_stat);
  delete D.bp;
  D.bp = nullptr;
}

}  // namespace

Problem* dg_batch_problem(int did) {
  DgProblem* D = get_dg(did);
  if (!D) return nullptr;
  if (D->mixed) {
    set_error("batch: DG members need one degree for all elements");
    return nullptr;
  }
  if (!D->assembled) {
    set_error("batch: the DG system is not assembled");
    return nullptr;
  }
  if (!D->dinv_valid && fill_block_jacobi(*D)) return nullptr;
  if (!D->bp && dg_batch_structure(*D)) return nullptr;
  if (!D->bp_valid) {
    Problem& P = *D->bp;
    const long long n = D->ndof;
    hipLaunchKernelGGL(k_dg_to_sell, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(), n,
                       D->ns, D->d_fnb, D->d_blk, D->d_dinv, D->d_brhs, P.d_sl_ptr, P.d_val,
                       P.d_rhs);
    if (hipGetLastError() != hipSuccess) {
      set_error("batch: the DG member's matrix could not be formed");
      return nullptr;
    }
    P.dinv_valid = false;
    P.val32_valid = false;
    D->bp_valid = true;
  }
  D->sol_batch = true;
  return D->bp;
}

}  // namespace mf

namespace {
// a degree code of the C ABI: tensor 100 pz + px (px, pz 1..3, or pz = 0 with
// px 1, 2), or the complete base kDgComplete + p (p 1..3)
bool dg_degree_ok(int pdeg) {
  if (pdeg > kDgComplete) return pdeg - kDgComplete >= 1 && pdeg - kDgComplete <= 3;
  const int px = pdeg % 100, pz = pdeg / 100;
  return (px >= 1 && px <= 3 && pz >= 1 && pz <= 3) || pdeg == 1 || pdeg == 2;
}
}  // namespace

extern "C" {

int mf_dg_create(int mesh, int pdeg) {
  Mesh* m = mesh_by_id(mesh);
  if (!m) return -3;
  if (!dg_degree_ok(pdeg)) {
    set_error("mf_dg_create: pdeg must be 100*pz + px with 1 <= px, pz <= 3, 1 or 2 (pz = 0), "
              "or MF_DG_COMPLETE + p with 1 <= p <= 3");
    return -1;
  }
  return dg_create(mesh, std::vector<int>(m->nel, pdeg));
}

int mf_dg_create_mixed(int mesh, const int* pdeg) {
  Mesh* m = mesh_by_id(mesh);
  if (!m) return -3;
  if (!pdeg) {
    set_error("mf_dg_create_mixed: no degrees");
    return -1;
  }
  std::vector<int> v(pdeg, pdeg + m->nel);
  bool same = true;
  for (int e = 0; e < m->nel; e++) same = same && v[e] == v[0];
  // one family per mesh: isotropic tensor degrees, pz = 0 degrees, or the
  // complete base
  auto family = [](int c) {
    if (c > kDgComplete) return dg_degree_ok(c) ? 2 : -1;
    if (c == 1 || c == 2) return 1;
    return (c % 100 == c / 100 && c % 100 >= 1 && c % 100 <= 3) ? 0 : -1;
  };
  for (int e = 0; e < m->nel; e++) {
    const bool ok = same ? dg_degree_ok(v[e]) : family(v[e]) >= 0 && family(v[e]) == family(v[0]);
    if (!ok) {
      set_error("mf_dg_create_mixed: degrees per element from one family: 101, 202, 303; or "
                "1, 2 (pz = 0); or MF_DG_COMPLETE + p, 1 <= p <= 3 (any one degree when all are "
                "equal)");
      return -1;
    }
  }
  return dg_create(mesh, v);
}

int mf_dg_layout(int did, long long* ndof, long long* xoff, long long* soff) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (ndof) *ndof = D->ndof;
  const size_t nel = D->nel;
  if (xoff) {
    if (D->mixed) memcpy(xoff, D->xoff.data(), (nel + 1) * sizeof(long long));
    else for (size_t e = 0; e <= nel; e++) xoff[e] = (long long)e * D->ns;
  }
  if (soff) {
    if (D->mixed) memcpy(soff, D->soff.data(), (6 * nel + 1) * sizeof(long long));
    else for (size_t k = 0; k <= 6 * nel; k++) soff[k] = (long long)k * D->ns * D->ns;
  }
  return 0;
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
  D->assembled = D->dinv_valid = D->ilu_valid = D->bp_valid = false;
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
  if (D->mixed) {
    set_error("mf_dg_element_matrices: one degree per problem (mixed degrees: mf_dg_get_block_row)");
    return -1;
  }
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
  if (D->mixed) {
    set_error("mf_dg_integrate: one degree per problem (mixed degrees: mf_dg_assemble)");
    return -1;
  }
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

int mf_dg_set_valu_blocks(int did, int max_ns) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (max_ns < 0) {
    set_error("mf_dg_set_valu_blocks: max_ns >= 0 required");
    return -1;
  }
  D->valu_ns = max_ns;
  return 0;
}

int mf_dg_set_penalty(int did, double penalty_const) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!(penalty_const >= 0.0)) {
    set_error("mf_dg_set_penalty: the penalty constant must be >= 0");
    return -1;
  }
  D->penalty_const = penalty_const;
  D->assembled = D->dinv_valid = D->ilu_valid = D->bp_valid = false;
  return 0;
}

int mf_dg_set_boundary_values(int did, const double* val) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  const size_t nf = 5 * (size_t)D->nel;
  if (val)
    MF_HIP(hipMemcpyAsync(D->d_fval, val, nf * sizeof(double), hipMemcpyHostToDevice, stream()));
  else
    MF_HIP(hipMemsetAsync(D->d_fval, 0, nf * sizeof(double), stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  D->assembled = D->dinv_valid = D->ilu_valid = D->bp_valid = false;
  return 0;
}

int mf_dg_get_faces(int did, int* nb, int* code, double* hf) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  const size_t nf = 5 * (size_t)D->nel;
  if (nb) MF_HIP(hipMemcpyAsync(nb, D->d_fnb, nf * sizeof(int), hipMemcpyDeviceToHost, stream()));
  if (code)
    MF_HIP(hipMemcpyAsync(code, D->d_fcode, nf * sizeof(int), hipMemcpyDeviceToHost, stream()));
  if (hf) MF_HIP(hipMemcpyAsync(hf, D->d_fhf, nf * sizeof(double), hipMemcpyDeviceToHost, stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_dg_assemble(int did) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->d_blk) {
    MF_HIP(hipMalloc(&D->d_blk, (size_t)D->blk_len * sizeof(double)));
    MF_HIP(hipMalloc(&D->d_brhs, (size_t)D->ndof * sizeof(double)));
    // boundary slots are never written: keep them zero
    MF_HIP(hipMemsetAsync(D->d_blk, 0, (size_t)D->blk_len * sizeof(double), stream()));
  }
  const int rc = launch_rows(*D);
  if (rc) return rc;
  MF_HIP(hipStreamSynchronize(stream()));
  D->assembled = true;
  D->dinv_valid = D->ilu_valid = D->bp_valid = D->mg_valid = false;
  return 0;
}

int mf_dg_get_block_row(int did, int e, double* blocks, double* rhs) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->assembled || e < 0 || e >= D->nel) {
    set_error("mf_dg_get_block_row: not assembled or element out of range");
    return -1;
  }
  const size_t ns = D->mixed ? ns_of_deg(D->deg[e]) : D->ns;
  const size_t b0 = D->mixed ? D->soff[6 * (size_t)e] : (size_t)e * 6 * ns * ns;
  size_t b1 = b0 + 6 * ns * ns;
  if (D->mixed) {  // the row's six slots are contiguous; the last one ends the row
    const int nb = D->h_fnb[5 * (size_t)e + 4];
    b1 = D->soff[6 * (size_t)e + 5] + ns * (nb >= 0 ? ns_of_deg(D->deg[nb]) : ns);
  }
  const size_t x0 = D->mixed ? D->xoff[e] : (size_t)e * ns;
  if (blocks)
    MF_HIP(hipMemcpyAsync(blocks, D->d_blk + b0, (b1 - b0) * sizeof(double), hipMemcpyDeviceToHost,
                          stream()));
  if (rhs)
    MF_HIP(hipMemcpyAsync(rhs, D->d_brhs + x0, ns * sizeof(double), hipMemcpyDeviceToHost,
                          stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_dg_fill_precon(int did) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->assembled) {
    set_error("mf_dg_fill_precon: the system is not assembled");
    return -1;
  }
  return fill_block_jacobi(*D);
}

int mf_dg_get_precon_block(int did, int e, double* inv) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->dinv_valid || e < 0 || e >= D->nel || !inv) {
    set_error("mf_dg_get_precon_block: no block Jacobi filled or element out of range");
    return -1;
  }
  const size_t ns = D->mixed ? ns_of_deg(D->deg[e]) : D->ns, nn = ns * ns;
  const size_t v0 = D->mixed ? D->voff[e] : (size_t)e * nn;
  MF_HIP(hipMemcpyAsync(inv, D->d_dinv + v0, nn * sizeof(double), hipMemcpyDeviceToHost, stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_dg_fill_ilu(int did, int* nlevels) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->assembled) {
    set_error("mf_dg_fill_ilu: the system is not assembled");
    return -1;
  }
  const int rc = fill_block_ilu(*D);
  if (rc) return rc;
  if (nlevels) {
    nlevels[0] = (int)D->ilu_fptr.size() - 1;
    nlevels[1] = (int)D->ilu_bptr.size() - 1;
  }
  return 0;
}

int mf_dg_get_ilu_row(int did, int e, double* blocks) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->ilu_valid || e < 0 || e >= D->nel || !blocks) {
    set_error("mf_dg_get_ilu_row: no block ILU filled or element out of range");
    return -1;
  }
  const size_t ns = D->mixed ? ns_of_deg(D->deg[e]) : D->ns, nn = ns * ns;
  const size_t b0 = D->mixed ? D->soff[6 * (size_t)e] : (size_t)e * 6 * nn;
  size_t b1 = b0 + 6 * nn;
  if (D->mixed) {
    const int nb = D->h_fnb[5 * (size_t)e + 4];
    b1 = D->soff[6 * (size_t)e + 5] + ns * (nb >= 0 ? ns_of_deg(D->deg[nb]) : ns);
  }
  const size_t v0 = D->mixed ? D->voff[e] : (size_t)e * nn;
  MF_HIP(hipMemcpyAsync(blocks, D->d_ilu_inv + v0, nn * sizeof(double), hipMemcpyDeviceToHost,
                        stream()));
  MF_HIP(hipMemcpyAsync(blocks + nn, D->d_ilu + b0 + nn, (b1 - b0 - nn) * sizeof(double),
                        hipMemcpyDeviceToHost, stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_dg_ilu_apply(int did, const double* b, double* v) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->ilu_valid || !b || !v) {
    set_error("mf_dg_ilu_apply: no block ILU filled");
    return -1;
  }
  double* d_v = nullptr;
  const size_t bytes = (size_t)D->ndof * sizeof(double);
  MF_HIP(hipMalloc(&d_v, bytes));
  hipError_t e = hipMemcpyAsync(d_v, b, bytes, hipMemcpyHostToDevice, stream());
  int rc = e == hipSuccess ? dg_ilu_solve(*D, d_v, nullptr) : 0;
  if (e == hipSuccess && rc == 0) e = hipMemcpyAsync(v, d_v, bytes, hipMemcpyDeviceToHost, stream());
  if (e == hipSuccess && rc == 0) e = hipStreamSynchronize(stream());
  (void)hipFree(d_v);
  MF_HIP(e);
  return rc;
}

int mf_dg_gmres(int did, int precon, int krylbas, int max_iter, double tol, double* x,
                int* iters, double* res_rel, double* resid) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->assembled) {
    set_error("mf_dg_gmres: the system is not assembled");
    return -1;
  }
  if (precon < 0 || precon > 4) {
    set_error("mf_dg_gmres: precon must be 0 (none), 1 (block Jacobi), 2 (block Gauss-Seidel), "
              "3 (block ILU(0)) or 4 (p-multigrid V-cycle)");
    return -1;
  }
  if (precon == 4) {
    if (!D->mg) {
      set_error("mf_dg_gmres: precon 4 needs mf_dg_set_multigrid first");
      return -1;
    }
    if (!D->mg_valid && mg_refresh(*D)) return -2;
  }
  if ((precon == 1 || precon == 2) && !D->dinv_valid && fill_block_jacobi(*D)) return -2;
  if (precon == 2 && build_gs_colors(*D)) return -2;
  if (precon == 3 && !D->ilu_valid) {
    const int rc = fill_block_ilu(*D);
    if (rc) return rc;
  }
  if (!D->P) {
    D->P = new Problem();
    D->P->n = (int)D->ndof;
    D->P->el_type = kPrism;
    D->op = new DgOp();
    D->op->D = D;
    D->P->ext = D->op;
  }
  D->op->precon = precon;
  D->P->d_rhs = D->d_brhs;
  D->sol_batch = false;
  if (!D->P->d_x) MF_HIP(hipMalloc(&D->P->d_x, (size_t)D->P->n * sizeof(double)));
  return gmres_solve(*D->P, precon, krylbas, max_iter, tol, 0.0, 1, x, iters, res_rel, resid);
}

int mf_dg_set_gs(int did, int nsweeps, int* ncolors) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (nsweeps < 1) {
    set_error("mf_dg_set_gs: nsweeps must be >= 1");
    return -1;
  }
  D->gs_sweeps = nsweeps;
  if (build_gs_colors(*D)) return -2;
  if (ncolors) *ncolors = (int)D->gs_color_ptr.size() - 1;
  return 0;
}

// p-multigrid V-cycle as the GMRES preconditioner (precon 4) of a uniform
// isotropic tensor degree p >= 2: levels p-1 .. 1 on the same mesh, nu1 / nu2
// block Gauss-Seidel sweeps before / after the coarse correction, `coarse`
// block-ILU(0) corrections on degree 1.  Returns the number of levels
int mf_dg_set_multigrid(int did, int nu1, int nu2, int coarse) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (D->mixed || D->px != D->pz || D->px < 2 || nu1 < 0 || nu2 < 0 || nu1 + nu2 < 1 ||
      coarse < 1) {
    set_error("mf_dg_set_multigrid: needs a uniform degree 202 or 303, nu1, nu2 >= 0 with "
              "nu1 + nu2 >= 1, and coarse >= 1");
    return -1;
  }
  if (!D->mg) {
    auto M = std::make_unique<DgMg>();
    const int p = D->px;
    for (int q = 1; q < p; q++) {
      const int id = dg_create(D->mesh, std::vector<int>(D->nel, 101 * q));
      if (id < 0) {
        for (int h : M->ids) {
          auto it = g_dg.find(h);
          if (it != g_dg.end()) {
            free_dg(*it->second);
            g_dg.erase(it);
          }
        }
        return id;
      }
      M->ids.push_back(id);
      M->lev.push_back(get_dg(id));
    }
    M->lev.push_back(D);
    const int L = (int)M->lev.size();
    M->ops.resize(L);
    for (int l = 0; l < L; l++) {
      M->ops[l].D = M->lev[l];
      M->ops[l].precon = 0;
    }
    M->d_map.assign(L - 1, nullptr);
    M->b.assign(L, nullptr);
    M->e.assign(L, nullptr);
    M->t.assign(L, nullptr);
    D->mg = M.release();  // owned from here: freed with the problem
    DgMg& G = *D->mg;
    for (int l = 0; l + 1 < L; l++) {
      int exc[40], exf[40];
      shape_exponents(l + 1, l + 1, exc);
      shape_exponents(l + 2, l + 2, exf);
      const int nc = num_shape(l + 1, l + 1), nf = num_shape(l + 2, l + 2);
      std::vector<int> map(nc, -1);
      for (int i = 0; i < nc; i++)
        for (int j = 0; j < nf; j++)
          if (exf[j] == exc[i]) map[i] = j;
      MF_HIP(hipMalloc(&G.d_map[l], nc * sizeof(int)));
      MF_HIP(hipMemcpy(G.d_map[l], map.data(), nc * sizeof(int), hipMemcpyHostToDevice));
    }
    for (int l = 0; l < L; l++) {
      const size_t bytes = std::max<long long>(G.lev[l]->ndof, 1) * sizeof(double);
      MF_HIP(hipMalloc(&G.t[l], bytes));
      if (l + 1 < L) {
        MF_HIP(hipMalloc(&G.b[l], bytes));
        MF_HIP(hipMalloc(&G.e[l], bytes));
      }
    }
    MF_HIP(hipMalloc(&G.zero, std::max<long long>(D->ndof, 1) * sizeof(double)));
    MF_HIP(hipMemset(G.zero, 0, std::max<long long>(D->ndof, 1) * sizeof(double)));
  }
  D->mg->nu1 = nu1;
  D->mg->nu2 = nu2;
  D->mg->coarse = coarse;
  D->mg_valid = false;
  return (int)D->mg->lev.size();
}

// handle of p-multigrid level l (0: degree 1) below the problem's own,
// its Galerkin blocks current (inspection)
int mf_dg_multigrid_level(int did, int l) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->mg || l < 0 || l >= (int)D->mg->ids.size() || !D->assembled) {
    set_error("mf_dg_multigrid_level: needs an assembled problem with multigrid levels and "
              "0 <= l < levels - 1");
    return -1;
  }
  if (!D->mg_valid && mg_refresh(*D)) return -2;
  return D->mg->ids[l];
}

int mf_dg_get_solution(int did, double* x) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  const Problem* S = D->sol_batch ? D->bp : D->P;  // the last solve's workspace
  if (!S || !S->d_x) {
    set_error("mf_dg_get_solution: no solve has run");
    return -1;
  }
  MF_HIP(hipMemcpyAsync(x, S->d_x, (size_t)S->n * sizeof(double), hipMemcpyDeviceToHost,
                        stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_dg_bench(int did, int reps, double* out) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (!D->assembled || reps < 1 || !out) {
    set_error("mf_dg_bench: assemble first; reps >= 1");
    return -1;
  }
  if (!D->dinv_valid && fill_block_jacobi(*D)) return -2;
  hipStream_t s = stream();
  hipEvent_t ev[4];
  for (auto& e : ev) MF_HIP(hipEventCreate(&e));
  double t_asm = 0, t_op = 0;
  const size_t n = (size_t)D->ndof;
  double *x = nullptr, *w = nullptr;
  MF_HIP(hipMalloc(&x, n * sizeof(double)));
  MF_HIP(hipMalloc(&w, n * sizeof(double)));
  MF_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
  DgOp op;
  op.D = D;
  op.precon = 1;
  int rc = 0;
  for (int r = 0; r < reps && rc == 0; r++) {
    float ms = 0;
    MF_HIP(hipEventRecord(ev[0], s));
    rc = launch_rows(*D);
    MF_HIP(hipEventRecord(ev[1], s));
    if (rc == 0) rc = op.run_all(x, nullptr, nullptr, 0, nullptr, w, 0);
    MF_HIP(hipEventRecord(ev[2], s));
    MF_HIP(hipEventSynchronize(ev[2]));
    MF_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    t_asm += ms;
    MF_HIP(hipEventElapsedTime(&ms, ev[1], ev[2]));
    t_op += ms;
  }
  (void)hipFree(x);
  (void)hipFree(w);
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc) return rc;
  out[0] = t_asm / reps;
  out[1] = t_op / reps;
  return 0;
}

int mf_dg_free(int did) {
  DgProblem* D = get_dg(did);
  if (!D) return -3;
  if (D->bp && D->bp->batch_refs > 0) {
    set_error("mf_dg_free: the problem is a member of a batch (mf_batch_remove_dg first)");
    return -1;
  }
  free_dg(*D);
  g_dg.erase(did);
  return 0;
}

}  // extern "C"
