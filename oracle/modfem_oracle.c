/*
 * modfem_oracle.c -- CPU restatement of ModFEM's std-linear
 * integrate -> assemble -> MKB GMRES path (test infrastructure only; see the
 * header).  Paths cited as file:line are relative to modfem2017/src/.
 *
 * Parity status of each part (what pins it; tests/test_oracle_*.py):
 *   quadrature, shape functions, Jacobians  -- compared with the reference's
 *       own apr_set_quadr_3D / apr_elem_calc_3D compiled from its sources
 *       (oracle/_ref) and with closed forms (partition of unity, volumes);
 *   element integral                          -- closed-form P1 tet stiffness,
 *       zero row sums, patch test, reference Gauss-point data;
 *   penalty faces                             -- closed-form face mass;
 *   DOF list, CRS pattern, assembly tables,
 *   assembly, Jacobi, BJ sweep, residual      -- the reference's own
 *       las_crs_intf.c compiled from its sources (oracle/_ref), bit-exact;
 *   GMRES                                     -- dense LU solution on small
 *       systems, Krylov invariance, the reference's field dump
 *       (ctest/conv_diff/Laplace_std_in_box_opencl/comparison/field_box.dmp).
 *   drotg/drot/dtrsv/dgemv                    -- the BLAS the reference links
 *       is not in the tree (parity unpinned at that boundary); restated from
 *       the netlib reference algorithms and checked by their identities.
 */
#include "modfem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* ---------------------------------------------------------------------
 * Quadrature.  Literals are the 15-digit values of
 * apd_std_lin/aps_gauss_util_std.c:152-299; the point ordering follows
 * apr_set_quadr_3D (aps_std_util.c:1400-1434: z outer, triangle inner,
 * triangle coordinates taken from barycentric slots 1 and 2).
 * The tetrahedral rule reads the 4-coordinate table with stride 3 exactly as
 * aps_std_util.c:1377-1383 does, which yields the four points below.
 * ------------------------------------------------------------------- */
static const double G2 = 0.577350269189626;
static const double T1 = 0.166666666666667, T2 = 0.666666666666667;
static const double TW = 0.333333333333333;
static const double TA = 0.585410196624969, TB = 0.138196601125011;

int orc_quadr_3d(int el_type, double *xg, double *wg) {
  if (el_type == ORC_TETRA) {
    const double p[4][3] = {{TA, TB, TB}, {TB, TB, TA}, {TB, TB, TB}, {TB, TA, TB}};
    for (int k = 0; k < 4; k++) {
      xg[3 * k] = p[k][0]; xg[3 * k + 1] = p[k][1]; xg[3 * k + 2] = p[k][2];
      wg[k] = 0.25;
    }
    return 4;
  }
  const double tri[3][2] = {{T1, T1}, {T1, T2}, {T2, T1}};
  const double zg[2] = {-G2, G2};
  for (int ki = 0; ki < 2; ki++)
    for (int kj = 0; kj < 3; kj++) {
      int g = ki * 3 + kj;
      xg[3 * g] = tri[kj][0];
      xg[3 * g + 1] = tri[kj][1];
      xg[3 * g + 2] = zg[ki];
      wg[g] = 0.5 * TW * 1.0;
    }
  return 6;
}

/* apr_set_quadr_2D (aps_std_util.c:1441-1530): triangles use the 3-point
   rule with weights 0.5*w, quads 2x2 Gauss on [-1,1]^2 with weights 0.5*w*w */
int orc_quadr_2d(int fa_is_quad, double *xg, double *wg) {
  if (fa_is_quad) {
    const double g[2] = {-G2, G2};
    for (int ki = 0; ki < 2; ki++)
      for (int kj = 0; kj < 2; kj++) {
        xg[2 * (ki * 2 + kj)] = g[ki];
        xg[2 * (ki * 2 + kj) + 1] = g[kj];
        wg[ki * 2 + kj] = 0.5 * 1.0 * 1.0;
      }
    return 4;
  }
  const double tri[3][2] = {{T1, T1}, {T1, T2}, {T2, T1}};
  for (int k = 0; k < 3; k++) {
    xg[2 * k] = tri[k][0];
    xg[2 * k + 1] = tri[k][1];
    wg[k] = 0.5 * TW;
  }
  return 3;
}

/* ut_mat3_inv (aps_std_util.c:2770-2822): cofactor inverse, returns det */
double orc_mat3_inv(const double *m, double *mi) {
  double rjac = m[0] * (m[4] * m[8] - m[7] * m[5]) +
                m[3] * (m[7] * m[2] - m[1] * m[8]) +
                m[6] * (m[1] * m[5] - m[4] * m[2]);
  double r = 1.0 / rjac;
  double s0, s1, s2;
  s0 = m[4] * m[8] - m[7] * m[5];
  s1 = -m[3] * m[8] + m[6] * m[5];
  s2 = m[3] * m[7] - m[6] * m[4];
  mi[0] = s0 * r; mi[3] = s1 * r; mi[6] = s2 * r;
  s0 = m[7] * m[2] - m[1] * m[8];
  s1 = m[0] * m[8] - m[6] * m[2];
  s2 = -m[0] * m[7] + m[6] * m[1];
  mi[1] = s0 * r; mi[4] = s1 * r; mi[7] = s2 * r;
  s0 = m[1] * m[5] - m[4] * m[2];
  s1 = -m[0] * m[5] + m[3] * m[2];
  s2 = m[0] * m[4] - m[3] * m[1];
  mi[2] = s0 * r; mi[5] = s1 * r; mi[8] = s2 * r;
  return rjac;
}

static double vec3_len(const double *v) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* Element calculations.  Prism: apr_shape_fun_3D_std_lin_prism
   (aps_std_util.c:463-511) and apr_elem_calc_3D_prism (:520-820); tetra:
   apr_shape_fun_3D_std_lin_tetra (:826-867), apr_elem_calc_3D_tetra
   (:869-1226).  control==2: volume (returns |det|); control>=3: face
   control-3 (returns the area element, vec_nor optional). */
double orc_elem_calc_3d(int el_type, int control, const double *eta,
                        const double *nc, double *phi, double *dphix,
                        double *dphiy, double *dphiz, double *xcoor,
                        double *vec_nor) {
  double J[9], Ji[9], det;
  int ns;
  if (el_type == ORC_TETRA) {
    ns = 4;
    phi[0] = 1.0 - eta[0] - eta[1] - eta[2];
    phi[1] = eta[0]; phi[2] = eta[1]; phi[3] = eta[2];
    dphix[0] = -1.0; dphix[1] = 1.0; dphix[2] = 0.0; dphix[3] = 0.0;
    dphiy[0] = -1.0; dphiy[1] = 0.0; dphiy[2] = 1.0; dphiy[3] = 0.0;
    dphiz[0] = -1.0; dphiz[1] = 0.0; dphiz[2] = 0.0; dphiz[3] = 1.0;
    J[0] = nc[3] - nc[0]; J[3] = nc[4] - nc[1]; J[6] = nc[5] - nc[2];
    J[1] = nc[6] - nc[0]; J[4] = nc[7] - nc[1]; J[7] = nc[8] - nc[2];
    J[2] = nc[9] - nc[0]; J[5] = nc[10] - nc[1]; J[8] = nc[11] - nc[2];
    if (xcoor) {
      xcoor[0] = nc[0] + eta[0] * J[0] + eta[1] * J[1] + eta[2] * J[2];
      xcoor[1] = nc[1] + eta[0] * J[3] + eta[1] * J[4] + eta[2] * J[5];
      xcoor[2] = nc[2] + eta[0] * J[6] + eta[1] * J[7] + eta[2] * J[8];
    }
    det = orc_mat3_inv(J, Ji) * 0.16666666666666666666666666666667;
  } else {
    ns = 6;
    const double e0 = eta[0], e1 = eta[1], e2 = eta[2];
    phi[0] = (-e2 + 1) / 2. * (1 - e0 - e1);
    phi[1] = (-e2 + 1) / 2. * e0;
    phi[2] = (-e2 + 1) / 2. * e1;
    phi[3] = (e2 + 1) / 2. * (1 - e0 - e1);
    phi[4] = (e2 + 1) / 2. * e0;
    phi[5] = (e2 + 1) / 2. * e1;
    dphix[0] = (e2 - 1) / 2.; dphix[1] = (-e2 + 1) / 2.; dphix[2] = 0.0;
    dphix[3] = (-e2 - 1) / 2.; dphix[4] = (e2 + 1) / 2.; dphix[5] = 0.0;
    dphiy[0] = (e2 - 1) / 2.; dphiy[1] = 0.0; dphiy[2] = (-e2 + 1) / 2.;
    dphiy[3] = (-e2 - 1) / 2.; dphiy[4] = 0.0; dphiy[5] = (e2 + 1) / 2.;
    dphiz[0] = -0.5 * (1 - e0 - e1); dphiz[1] = -0.5 * e0; dphiz[2] = -0.5 * e1;
    dphiz[3] = 0.5 * (1 - e0 - e1); dphiz[4] = 0.5 * e0; dphiz[5] = 0.5 * e1;
    double gp[6], gx[6], gy[6], gz[6];
    gp[0] = (1.0 - e0 - e1) * (1.0 - e2) / 2.0;
    gp[1] = e0 * (1.0 - e2) / 2.0;
    gp[2] = e1 * (1.0 - e2) / 2.0;
    gp[3] = (1.0 - e0 - e1) * (1.0 + e2) / 2.0;
    gp[4] = e0 * (1.0 + e2) / 2.0;
    gp[5] = e1 * (1.0 + e2) / 2.0;
    if (xcoor) {
      xcoor[0] = 0; xcoor[1] = 0; xcoor[2] = 0;
      for (int i = 0; i < 6; i++) {
        xcoor[0] += nc[3 * i] * gp[i];
        xcoor[1] += nc[3 * i + 1] * gp[i];
        xcoor[2] += nc[3 * i + 2] * gp[i];
      }
    }
    gx[0] = -(1.0 - e2) / 2.0; gx[1] = (1.0 - e2) / 2.0; gx[2] = 0.0;
    gx[3] = -(1.0 + e2) / 2.0; gx[4] = (1.0 + e2) / 2.0; gx[5] = 0.0;
    gy[0] = -(1.0 - e2) / 2.0; gy[1] = 0.0; gy[2] = (1.0 - e2) / 2.0;
    gy[3] = -(1.0 + e2) / 2.0; gy[4] = 0.0; gy[5] = (1.0 + e2) / 2.0;
    gz[0] = -(1.0 - e0 - e1) / 2.0; gz[1] = -e0 / 2.0; gz[2] = -e1 / 2.0;
    gz[3] = (1.0 - e0 - e1) / 2.0; gz[4] = e0 / 2.0; gz[5] = e1 / 2.0;
    for (int k = 0; k < 9; k++) J[k] = 0.0;
    for (int i = 0; i < 6; i++) {
      J[0] += nc[3 * i] * gx[i];
      J[1] += nc[3 * i] * gy[i];
      J[2] += nc[3 * i] * gz[i];
      J[3] += nc[3 * i + 1] * gx[i];
      J[4] += nc[3 * i + 1] * gy[i];
      J[5] += nc[3 * i + 1] * gz[i];
      J[6] += nc[3 * i + 2] * gx[i];
      J[7] += nc[3 * i + 2] * gy[i];
      J[8] += nc[3 * i + 2] * gz[i];
    }
    det = orc_mat3_inv(J, Ji);
  }
  for (int i = 0; i < ns; i++) {
    double daux = dphix[i] * Ji[0] + dphiy[i] * Ji[3] + dphiz[i] * Ji[6];
    double faux = dphix[i] * Ji[1] + dphiy[i] * Ji[4] + dphiz[i] * Ji[7];
    dphiz[i] = dphix[i] * Ji[2] + dphiy[i] * Ji[5] + dphiz[i] * Ji[8];
    dphix[i] = daux;
    dphiy[i] = faux;
  }
  if (control == 2) return fabs(det);
  double ds[3];
  if (el_type == ORC_TETRA) {
    /* MMC_FACE_NODES_FOR_TETRA of the hybrid mesh module
       (mmd_t4_prism/mms_prism_t4_intf.cpp:218): v1 x v2 points outward */
    static const int fn[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
    int f = control - 3;
    const double *p0 = nc + 3 * fn[f][0], *p1 = nc + 3 * fn[f][1],
                 *p2 = nc + 3 * fn[f][2];
    double v1[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double v2[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    ds[0] = v1[1] * v2[2] - v1[2] * v2[1];
    ds[1] = v1[2] * v2[0] - v1[0] * v2[2];
    ds[2] = v1[0] * v2[1] - v1[1] * v2[0];
  } else {
    switch (control) {
      case 3:
        ds[0] = -J[3] * J[7] + J[6] * J[4];
        ds[1] = -J[6] * J[1] + J[0] * J[7];
        ds[2] = -J[0] * J[4] + J[3] * J[1];
        break;
      case 4:
        ds[0] = J[3] * J[7] - J[6] * J[4];
        ds[1] = J[6] * J[1] - J[0] * J[7];
        ds[2] = J[0] * J[4] - J[3] * J[1];
        break;
      case 5:
        ds[0] = J[3] * J[8] - J[6] * J[5];
        ds[1] = J[6] * J[2] - J[0] * J[8];
        ds[2] = J[0] * J[5] - J[3] * J[2];
        break;
      case 6:
        ds[0] = (J[4] - J[3]) * J[8] - (J[7] - J[6]) * J[5];
        ds[1] = (J[7] - J[6]) * J[2] - (J[1] - J[0]) * J[8];
        ds[2] = (J[1] - J[0]) * J[5] - (J[4] - J[3]) * J[2];
        break;
      default:
        ds[0] = J[5] * J[7] - J[8] * J[4];
        ds[1] = J[8] * J[1] - J[2] * J[7];
        ds[2] = J[2] * J[4] - J[5] * J[1];
        break;
    }
  }
  det = vec3_len(ds);
  if (vec_nor) {
    vec_nor[0] = ds[0] / det; vec_nor[1] = ds[1] / det; vec_nor[2] = ds[2] / det;
  }
  return det;
}

/* exact solutions of the DIFF_IN_CUBE / DIFF_IN_BOX test problems
   (pdd_conv_diff/main/pds_conv_diff_util.c, pdr_exact_sol) */
static void exact_sol(int model, const double *x, double *u, double *lapl) {
  if (model == ORC_SRC_LINEAR) { /* pinning case only: u = 0.3 + x - 2y + 0.5z */
    if (u) *u = 0.3 + x[0] - 2.0 * x[1] + 0.5 * x[2];
    if (lapl) *lapl = 0.0;
    return;
  }
  double a1 = (model == ORC_SRC_DIFF_IN_BOX) ? 100.0 : 1.0;
  double X = x[0], Y = x[1], Z = x[2];
  double e = exp((-X * X - Y * Y - Z * Z) / a1);
  if (u) *u = e;
  if (lapl)
    *lapl = -(2.0 / a1) *
            ((1 - 2 * X * X / a1) + (1 - 2 * Y * Y / a1) + (1 - 2 * Z * Z / a1)) * e;
}

/* gradient of the exact solution (pdr_exact_sol, pds_conv_diff_util.c:
   1195-1260: u_x = -2 x exp(-|x|^2 / a1) / a1, likewise y, z) */
static void exact_grad(int model, const double *x, double *gr) {
  if (model == ORC_SRC_LINEAR) {
    gr[0] = 1.0; gr[1] = -2.0; gr[2] = 0.5;
    return;
  }
  double a1 = (model == ORC_SRC_DIFF_IN_BOX) ? 100.0 : 1.0;
  double X = x[0], Y = x[1], Z = x[2];
  gr[0] = -2.0 * X * exp((-X * X - Y * Y - Z * Z) / a1) / a1;
  gr[1] = -2.0 * Y * exp((-X * X - Y * Y - Z * Z) / a1) / a1;
  gr[2] = -2.0 * Z * exp((-X * X - Y * Y - Z * Z) / a1) / a1;
}

/* apr_num_int_el, nreq==1 (aps_std_lin_intf.c:1537-2870).  Term order and
   branch selection follow the reference: mass (ind[1]); diffusion full 9-term
   branch when ind[10] && ind[2] (:2112-2136), diagonal 3-term branch when
   ind[10] && !ind[2] (:2209-2224); convection b (ind[13], :2346-2360); t
   terms (ind[16]); reaction c (ind[17], :2612-2625); rhs q (ind[21]) and s
   (ind[22]) (:2672-2810).  Stiff is column-major stiff[u*ns+w]. */
int orc_num_int_el(int el_type, const double *nc, const int *ind,
                   const double *cf, int src_model, double *stiff,
                   double *rhs) {
  double xg[18], wg[6], phi[6], dx[6], dy[6], dz[6], xcoor[3];
  int ng = orc_quadr_3d(el_type, xg, wg);
  int ns = (el_type == ORC_TETRA) ? 4 : 6;
  const double *a = cf + ORC_C_A, *b = cf + ORC_C_B, *t = cf + ORC_C_T;
  const double *q = cf + ORC_C_Q;
  for (int ki = 0; ki < ng; ki++) {
    double determ = orc_elem_calc_3d(el_type, 2, &xg[3 * ki], nc, phi, dx, dy,
                                     dz, xcoor, NULL);
    double vol = determ * wg[ki];
    double sval = cf[ORC_C_S];
    if (src_model != ORC_SRC_CONST) {
      double lapl;
      exact_sol(src_model, xcoor, NULL, &lapl);
      sval = -lapl;
    }
    int kk;
    if (stiff) {
      if (ind[ORC_IND_M]) {
        kk = 0;
        for (int u = 0; u < ns; u++) {
          for (int w = 0; w < ns; w++)
            stiff[kk + w] += cf[ORC_C_M] * phi[u] * phi[w] * vol;
          kk += ns;
        }
      }
      if (ind[ORC_IND_AZZ]) {
        if (ind[ORC_IND_AXX]) {
          kk = 0;
          for (int u = 0; u < ns; u++) {
            for (int w = 0; w < ns; w++)
              stiff[kk + w] += (a[0] * dx[u] * dx[w] + a[1] * dy[u] * dx[w] +
                                a[2] * dz[u] * dx[w] + a[3] * dx[u] * dy[w] +
                                a[4] * dy[u] * dy[w] + a[5] * dz[u] * dy[w] +
                                a[6] * dx[u] * dz[w] + a[7] * dy[u] * dz[w] +
                                a[8] * dz[u] * dz[w]) * vol;
            kk += ns;
          }
        } else {
          kk = 0;
          for (int u = 0; u < ns; u++) {
            for (int w = 0; w < ns; w++)
              stiff[kk + w] += (a[0] * dx[u] * dx[w] + a[4] * dy[u] * dy[w] +
                                a[8] * dz[u] * dz[w]) * vol;
            kk += ns;
          }
        }
      }
      if (ind[ORC_IND_BZ]) {
        kk = 0;
        for (int u = 0; u < ns; u++) {
          for (int w = 0; w < ns; w++)
            stiff[kk + w] += (b[0] * dx[u] * phi[w] + b[1] * dy[u] * phi[w] +
                              b[2] * dz[u] * phi[w]) * vol;
          kk += ns;
        }
      }
      if (ind[ORC_IND_TZ]) {
        kk = 0;
        for (int u = 0; u < ns; u++) {
          for (int w = 0; w < ns; w++)
            stiff[kk + w] += (t[0] * phi[u] * dx[w] + t[1] * phi[u] * dy[w] +
                              t[2] * phi[u] * dz[w]) * vol;
          kk += ns;
        }
      }
      if (ind[ORC_IND_C]) {
        kk = 0;
        for (int u = 0; u < ns; u++) {
          for (int w = 0; w < ns; w++)
            stiff[kk + w] += cf[ORC_C_C] * phi[u] * phi[w] * vol;
          kk += ns;
        }
      }
    }
    if (rhs) {
      if (ind[ORC_IND_QZ]) {
        for (int w = 0; w < ns; w++)
          rhs[w] += (q[0] * dx[w] + q[1] * dy[w] + q[2] * dz[w]) * vol;
      }
      if (ind[ORC_IND_S]) {
        for (int w = 0; w < ns; w++) rhs[w] += sval * phi[w] * vol;
      }
    }
  }
  return ns;
}

/* face-local -> element-local coordinates for the first (owning) neighbour,
   mmr_fa_elem_coor (prisms: mmd_prism/mms_prism_datstr.c:1562-1700;
   tetrahedra: the hybrid mesh module, mmd_t4_prism/mms_prism_t4_intf.cpp:
   2004-2037, whose side 3 scales the face coordinates by 0.75 before
   mapping them onto the face opposite vertex 0). */
static double tet_side3_scale = 0.75;

/* test hook: plants another side-3 scale (1.0 is the undistorted map) so the
   tests can show which map the reference's figures need */
void orc_set_tet_side3_scale(double s) { tet_side3_scale = s; }

static void face_to_elem(int el_type, int side, const double *c, double *eta) {
  if (el_type == ORC_TETRA) {
    double c0 = c[0], c1 = c[1];
    switch (side) {
      case 0: eta[0] = c1; eta[1] = c0; eta[2] = 0.0; break;
      case 1: eta[0] = c0; eta[1] = 0.0; eta[2] = c1; break;
      case 2: eta[0] = 0.0; eta[1] = c1; eta[2] = c0; break;
      default:
        c0 *= tet_side3_scale;
        c1 *= tet_side3_scale;
        eta[0] = 1.0 - c0 - c1; eta[1] = c0; eta[2] = c1;
        break;
    }
    return;
  }
  switch (side) {
    case 0: eta[0] = c[1]; eta[1] = c[0]; eta[2] = -1.0; break;
    case 1: eta[0] = c[0]; eta[1] = c[1]; eta[2] = 1.0; break;
    case 2: eta[0] = 0.5 * (c[0] + 1.0); eta[1] = 0.0; eta[2] = c[1]; break;
    case 3:
      eta[0] = 0.5 * (-c[0] + 1.0); eta[1] = 1.0 - eta[0]; eta[2] = c[1];
      break;
    default: eta[0] = 0.0; eta[1] = 0.5 * (-c[0] + 1.0); eta[2] = c[1]; break;
  }
}

/* Dirichlet penalty on a boundary face (pds_conv_diff_weakform_std.c:
   253-629): K += 1e7 phi_j phi_i dA, F += 1e7 g phi_i dA on the owning
   element's full local matrix (column-major). */
int orc_face_penalty(int el_type, int side, const double *nc, double g_const,
                     int src_model, double *stiff, double *rhs) {
  int quad = (el_type == ORC_PRISM && side >= 2);
  double xg[8], wg[4], eta[3], phi[6], dx[6], dy[6], dz[6], xcoor[3], vn[3];
  int ng = orc_quadr_2d(quad, xg, wg);
  int ns = (el_type == ORC_TETRA) ? 4 : 6;
  const double penalty = 1.0e7;
  for (int ki = 0; ki < ng; ki++) {
    face_to_elem(el_type, side, &xg[2 * ki], eta);
    double determ = orc_elem_calc_3d(el_type, 3 + side, eta, nc, phi, dx, dy,
                                     dz, xcoor, vn);
    double area = determ * wg[ki];
    double g = g_const;
    if (src_model != ORC_SRC_CONST) exact_sol(src_model, xcoor, &g, NULL);
    int kk = 0;
    for (int j = 0; j < ns; j++) {
      for (int i = 0; i < ns; i++)
        stiff[kk + i] += penalty * phi[j] * phi[i] * area;
      kk += ns;
    }
    for (int i = 0; i < ns; i++) rhs[i] += penalty * g * phi[i] * area;
  }
  return ns;
}

/* Neumann face (pds_conv_diff_weakform_std.c:513-522): F += g phi_i dA with a
   constant flux g, no stiffness contribution. */
int orc_face_neumann(int el_type, int side, const double *nc, double g, double *rhs) {
  int quad = (el_type == ORC_PRISM && side >= 2);
  double xg[8], wg[4], eta[3], phi[6], dx[6], dy[6], dz[6], xcoor[3], vn[3];
  int ng = orc_quadr_2d(quad, xg, wg);
  int ns = (el_type == ORC_TETRA) ? 4 : 6;
  for (int ki = 0; ki < ng; ki++) {
    face_to_elem(el_type, side, &xg[2 * ki], eta);
    double determ = orc_elem_calc_3d(el_type, 3 + side, eta, nc, phi, dx, dy,
                                     dz, xcoor, vn);
    double area = determ * wg[ki];
    for (int i = 0; i < ns; i++) rhs[i] += g * phi[i] * area;
  }
  return ns;
}

/* Neumann face of an exact-solution problem (DIFF_IN_CUBE / DIFF_IN_BOX):
   at every face point the data is the exact normal flux g = grad u . n with
   the face's unit normal from apr_elem_calc_3D (pdr_bc_neum_coeff,
   pds_conv_diff_weakform.c; face loop pds_conv_diff_weakform_std.c:
   420-522), rhs_i += g phi_i dA */
int orc_face_neumann_exact(int el_type, int side, const double *nc, int src_model,
                           double *rhs) {
  int quad = (el_type == ORC_PRISM && side >= 2);
  double xg[8], wg[4], eta[3], phi[6], dx[6], dy[6], dz[6], xcoor[3], vn[3], gr[3];
  int ng = orc_quadr_2d(quad, xg, wg);
  int ns = (el_type == ORC_TETRA) ? 4 : 6;
  for (int ki = 0; ki < ng; ki++) {
    face_to_elem(el_type, side, &xg[2 * ki], eta);
    double determ = orc_elem_calc_3d(el_type, 3 + side, eta, nc, phi, dx, dy,
                                     dz, xcoor, vn);
    exact_grad(src_model, xcoor, gr);
    double g = gr[0] * vn[0] + gr[1] * vn[1] + gr[2] * vn[2];
    double area = determ * wg[ki];
    for (int i = 0; i < ns; i++) rhs[i] += g * phi[i] * area;
  }
  return ns;
}

/* ---------------------------------------------------------------------
 * Lists and DOF structures
 * ------------------------------------------------------------------- */
static int nodes_of(int t) { return t == ORC_TETRA ? 4 : 6; }

/* utr_get_list_dof_ent, std-linear branch (uts_ls_intf.c:229-398): DOF
   entities are vertices in first-appearance order over element int-ents in
   list order, element-local node order within each element. */
int orc_dof_list(int nel, const int *el_type, const int *el_nodes, int nno,
                 int *dof_node, int *node_dof) {
  for (int i = 0; i < nno; i++) node_dof[i] = -1;
  int n = 0;
  for (int e = 0; e < nel; e++) {
    int k = nodes_of(el_type[e]);
    for (int i = 0; i < k; i++) {
      int v = el_nodes[6 * e + i];
      if (node_dof[v] < 0) {
        node_dof[v] = n;
        dof_node[n] = v;
        n++;
      }
    }
  }
  return n;
}

static int cmp_int(const void *a, const void *b) {
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* neighbour block sets (sis_mkb_intf.c:762-984; boundary faces reuse the
   owning element's DOF list, so elements suffice), sorted ascending
   (:1107-1142), then the CRS layout of lar_allocate_SM_and_LV_crs
   (las_crs_intf.c:184-250): strictly ascending columns with the diagonal
   inserted in place. */
long orc_crs_pattern(int nel, const int *el_type, const int *el_nodes,
                     const int *node_dof, int n, int *row_ptr, int *col_ind) {
  /* count candidate (row, col) pairs per row */
  int *cnt = (int *)calloc((size_t)n + 1, sizeof(int));
  for (int e = 0; e < nel; e++) {
    int k = nodes_of(el_type[e]);
    for (int i = 0; i < k; i++) cnt[node_dof[el_nodes[6 * e + i]]] += k;
  }
  long *off = (long *)malloc(((size_t)n + 1) * sizeof(long));
  off[0] = 0;
  for (int i = 0; i < n; i++) off[i + 1] = off[i] + cnt[i];
  int *buf = (int *)malloc((size_t)off[n] * sizeof(int) + 1);
  for (int i = 0; i < n; i++) cnt[i] = 0;
  for (int e = 0; e < nel; e++) {
    int k = nodes_of(el_type[e]);
    for (int i = 0; i < k; i++) {
      int r = node_dof[el_nodes[6 * e + i]];
      for (int j = 0; j < k; j++) buf[off[r] + cnt[r]++] = node_dof[el_nodes[6 * e + j]];
    }
  }
  long nnz = 0;
  for (int r = 0; r < n; r++) {
    int *p = buf + off[r];
    int m = cnt[r];
    qsort(p, (size_t)m, sizeof(int), cmp_int);
    int u = 0;
    for (int i = 0; i < m; i++)
      if (u == 0 || p[i] != p[u - 1]) p[u++] = p[i];
    /* u distinct columns including the diagonal itself */
    if (row_ptr) row_ptr[r] = (int)nnz;
    if (col_ind) memcpy(col_ind + nnz, p, (size_t)u * sizeof(int));
    nnz += u;
  }
  if (row_ptr) row_ptr[n] = (int)nnz;
  free(buf); free(off); free(cnt);
  return nnz;
}

/* lar_fill_assembly_table_int_ent_crs (las_crs_intf.c:354-432) for every
   int-ent in list order (elements, then boundary faces which reuse the
   owning element's DOF list): table[j+i*k] = CRS position of
   (row L[j], col L[i]); table[k*k+j] = L[j]. */
long orc_assembly_tables(int nel, const int *el_type, const int *el_nodes,
                         int nbf, const int *bf_elem, const int *node_dof,
                         const int *row_ptr, const int *col_ind,
                         long *table_ptr, int *table) {
  long pos = 0;
  int nint = nel + nbf;
  for (int ie = 0; ie < nint; ie++) {
    int e = ie < nel ? ie : bf_elem[ie - nel];
    int k = nodes_of(el_type[e]);
    if (table_ptr) table_ptr[ie] = pos;
    if (table) {
      int L[6];
      for (int i = 0; i < k; i++) L[i] = node_dof[el_nodes[6 * e + i]];
      for (int i = 0; i < k; i++) {
        int col = L[i];
        for (int j = 0; j < k; j++) {
          int row = L[j];
          int found = -1;
          for (int c = row_ptr[row]; c < row_ptr[row + 1]; c++)
            if (col_ind[c] == col) { found = c; break; }
          table[pos + j + i * k] = found;
        }
      }
      for (int j = 0; j < k; j++) table[pos + k * k + j] = L[j];
    }
    pos += (long)(k + 1) * k;
  }
  if (table_ptr) table_ptr[nint] = pos;
  return pos;
}

static void gather_coor(const double *all, const int *nodes, int k, double *nc) {
  for (int i = 0; i < k; i++)
    for (int d = 0; d < 3; d++) nc[3 * i + d] = all[3 * nodes[i] + d];
}

/* pdr_error_test_exact (pdd_conv_diff/main/pds_conv_diff_util.c:749-960)
   for the std-linear approximation: over the elements in ID order, at the
   element's volume rule (apr_set_quadr_3D), u_h and grad u_h from the nodal
   values (apr_elem_calc_3D control 2) against the exact solution of the
   source model; err_L2^2 += (u - u_h)^2 |det| w, err_H1^2 += |grad u -
   grad u_h|^2 |det| w, per element first, then summed.  u holds one value
   per mesh node; returns the square roots in l2[0], l2[1]. */
int orc_std_errors(int nel, const int *el_type, const int *el_nodes,
                   const double *coor, const double *u, int src_model,
                   double *l2h1) {
  double l2 = 0.0, h1 = 0.0;
  for (int e = 0; e < nel; e++) {
    int k = nodes_of(el_type[e]);
    double nc[18], xg[3 * 6], wg[6], phi[6], dx[6], dy[6], dz[6], xc[3];
    gather_coor(coor, el_nodes + 6 * e, k, nc);
    int ng = orc_quadr_3d(el_type[e], xg, wg);
    double el2 = 0.0, eh1 = 0.0;
    for (int ki = 0; ki < ng; ki++) {
      double det = orc_elem_calc_3d(el_type[e], 2, &xg[3 * ki], nc, phi, dx, dy, dz, xc, NULL);
      double vol = det * wg[ki];
      double uh = 0.0, ux = 0.0, uy = 0.0, uz = 0.0;
      for (int i = 0; i < k; i++) {
        double d = u[el_nodes[6 * e + i]];
        uh += d * phi[i];
        ux += d * dx[i];
        uy += d * dy[i];
        uz += d * dz[i];
      }
      double f, g[3];
      exact_sol(src_model, xc, &f, NULL);
      exact_grad(src_model, xc, g);
      el2 += (f - uh) * (f - uh) * vol;
      eh1 += ((g[0] - ux) * (g[0] - ux) + (g[1] - uy) * (g[1] - uy) +
              (g[2] - uz) * (g[2] - uz)) * vol;
    }
    l2 += el2;
    h1 += eh1;
  }
  l2h1[0] = sqrt(l2);
  l2h1[1] = sqrt(h1);
  return nel;
}

/* sequential create+assemble (sis_mkb_intf.c:1843-1866): element int-ents in
   ID order then faces; each entity's matrix is zeroed (pds_conv_diff_
   weakform_std.c:161-173), integrated, and added through its table
   (las_crs_intf.c:440-513).  Element integrals run in parallel into
   per-element buffers; the scatter stays sequential so sums keep the
   reference order. */
int orc_integrate_assemble(int nel, const int *el_type, const int *el_nodes,
                           const double *coor, const int *ind,
                           const double *coef_el, int src_model, int nbf,
                           const int *bf_elem, const int *bf_side,
                           const double *bf_g, const long *table_ptr,
                           const int *table, double *val, double *rhs) {
  int n_rows = 0;
  long nnz = 0;
  (void)n_rows; (void)nnz;
  int nint = nel + nbf;
  double *em = (double *)malloc((size_t)nint * 42 * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int ie = 0; ie < nint; ie++) {
    double *sm = em + (size_t)ie * 42;
    for (int i = 0; i < 42; i++) sm[i] = 0.0;
    double nc[18];
    if (ie < nel) {
      int k = nodes_of(el_type[ie]);
      gather_coor(coor, el_nodes + 6 * ie, k, nc);
      orc_num_int_el(el_type[ie], nc, ind, coef_el + (size_t)ie * ORC_NCOEF,
                     src_model, sm, sm + k * k);
    } else {
      int f = ie - nel, e = bf_elem[f];
      int k = nodes_of(el_type[e]);
      gather_coor(coor, el_nodes + 6 * e, k, nc);
      orc_face_penalty(el_type[e], bf_side[f], nc, bf_g ? bf_g[f] : 0.0,
                       src_model, sm, sm + k * k);
    }
  }
  for (int ie = 0; ie < nint; ie++) {
    int e = ie < nel ? ie : bf_elem[ie - nel];
    int k = nodes_of(el_type[e]);
    const double *sm = em + (size_t)ie * 42;
    const int *tb = table + table_ptr[ie];
    for (int i = 0; i < k; i++)
      for (int j = 0; j < k; j++) val[tb[j + i * k]] += sm[j + i * k];
    for (int j = 0; j < k; j++) rhs[tb[k * k + j]] += sm[k * k + j];
  }
  free(em);
  return nint;
}

/* ---------------------------------------------------------------------
 * Storage operations (lad_crs)
 * ------------------------------------------------------------------- */
/* lar_allocate_preconditioner_crs + lar_fill_preconditioner_crs
   (las_crs_intf.c:658-737) */
void orc_fill_jacobi(int n, const int *row_ptr, const int *col_ind,
                     const double *val, int *diag_ptr, double *diag_precon) {
  for (int r = 0; r < n; r++)
    for (int c = row_ptr[r]; c < row_ptr[r + 1]; c++)
      if (col_ind[c] == r) diag_ptr[r] = c;
  for (int r = 0; r < n; r++) diag_precon[r] = 1.0 / val[diag_ptr[r]];
}

/* lar_compute_residual_crs (las_crs_intf.c:788-849): v = 0; v -= A x; v += b */
void orc_compute_residual(int n, const int *row_ptr, const int *col_ind,
                          const double *val, int use_rhs, const double *b,
                          const double *x, double *v) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int k = row_ptr[i]; k < row_ptr[i + 1]; k++)
      s -= val[k] * x[col_ind[k]];
    if (use_rhs) s += b[i];
    v[i] = s;
  }
}

/* lar_perform_BJ_or_GS_iterations_crs, BLOCK_JACOBI, one sweep
   (las_crs_intf.c:857-1013): v_i <- d_i * (b_i - sum_{j != i} a_ij v_j),
   entries left of the diagonal first, then right of it. */
void orc_bj_iteration(int n, const int *row_ptr, const int *col_ind,
                      const double *val, const int *diag_ptr,
                      const double *diag_precon, int use_rhs,
                      const double *b, double *v) {
  double *vt = (double *)malloc((size_t)n * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n; r++) {
    double vloc = 0.0;
    for (int k = row_ptr[r]; k < diag_ptr[r]; k++) vloc -= val[k] * v[col_ind[k]];
    for (int k = diag_ptr[r] + 1; k < row_ptr[r + 1]; k++)
      vloc -= val[k] * v[col_ind[k]];
    if (use_rhs) vloc += b[r];
    vloc *= diag_precon[r];
    vt[r] = vloc;
  }
  memcpy(v, vt, (size_t)n * sizeof(double));
  free(vt);
}

/* lar_perform_BJ_or_GS_iterations_crs, BLOCK_GS, as it runs on nthreads
   OpenMP threads (las_crs_intf.c:905-998): the rows are cut into chunks of
   portion = n/nthreads + 1, chunk t holding rows [t*portion, (t+1)*portion);
   within a chunk the rows are updated in order and an entry left of the
   diagonal (CRS position, k < diag_ptr) whose column lies in
   [chunk start, row) takes the chunk's new value, every other entry the value
   of the sweep's start; the chunks' new values replace v after the sweep.
   Columns that R's per-thread copy does not cover (more than half bandwidth
   + 10 below the chunk) read the sweep's start value here; R leaves them
   uninitialised, so matrices whose DOF half bandwidth exceeds R's block half
   bandwidth differ there.  nthreads = 1 is textbook forward Gauss-Seidel. */
void orc_crs_gs_chunked(int n, const int *row_ptr, const int *col_ind,
                        const double *val, const int *diag_ptr,
                        const double *diag_precon, int use_rhs, const double *b,
                        int nthreads, int nr_prec, double *v) {
  double *vold = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  const int portion = n / nthreads + 1;
  for (int it = 0; it < nr_prec; it++) {
    memcpy(vold, v, (size_t)n * sizeof(double));
    for (int t = 0; t < nthreads; t++) {
      const int first = t * portion;
      int last = (t + 1) * portion;
      if (last > n) last = n;
      for (int r = first; r < last; r++) {
        double vloc = 0.0;
        for (int k = row_ptr[r]; k < diag_ptr[r]; k++) {
          const int c = col_ind[k];
          vloc -= val[k] * ((c >= first && c < r) ? v[c] : vold[c]);
        }
        for (int k = diag_ptr[r] + 1; k < row_ptr[r + 1]; k++)
          vloc -= val[k] * vold[col_ind[k]];
        if (use_rhs) vloc += b[r];
        vloc *= diag_precon[r];
        v[r] = vloc;
      }
    }
  }
  free(vold);
}

/* ---------------------------------------------------------------------
 * BLAS restatements (netlib reference algorithms)
 * ------------------------------------------------------------------- */
void orc_drotg(double *a, double *b, double *c, double *s) {
  double roe = *b;
  if (fabs(*a) > fabs(*b)) roe = *a;
  double scale = fabs(*a) + fabs(*b);
  double r, z;
  if (scale == 0.0) {
    *c = 1.0; *s = 0.0; r = 0.0; z = 0.0;
  } else {
    double ra = *a / scale, rb = *b / scale;
    r = scale * sqrt(ra * ra + rb * rb);
    r = (roe >= 0 ? 1.0 : -1.0) * r;
    *c = *a / r;
    *s = *b / r;
    z = 1.0;
    if (fabs(*a) > fabs(*b)) z = *s;
    if (fabs(*b) >= fabs(*a) && *c != 0.0) z = 1.0 / *c;
  }
  *a = r;
  *b = z;
}

static void drot1(double *x, double *y, double c, double s) {
  double t = c * (*x) + s * (*y);
  *y = c * (*y) - s * (*x);
  *x = t;
}

static double dot(int n, const double *x, const double *y) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}

static double nrm2(int n, const double *x) { return sqrt(dot(n, x, x)); }

/* lsr_mkb_core_compreres for NO_PRECON / BLOCK_JACOBI with one sweep
   (lss_mkb_core.c:1861-2097):
     control 1: V = X + M^-1 (b - A X); V -= X
     control 0: V = X; one BJ sweep with no rhs; V -= X; V *= -1 */
typedef struct {
  int n;
  const int *rp, *ci;
  const double *val, *b;
  int precon;
  int *diag_ptr;
  double *dinv;
  double *lu, *tmp; /* precon 2: ILU(0) factors on the pattern, work vector */
} orc_op;

/* ILU(0) on the CRS pattern in row order, IKJ form (Saad, Iterative Methods
   for Sparse Linear Systems, 2nd ed., Alg. 10.4) as the reference's block
   ILU(0) does it for 1x1 blocks (lad_block/las_block_intf.c:1830-2143): for
   each row i and each k < i of its pattern in ascending order,
   a_ik = a_ik * d_k with d_k = 1/u_kk, then a_ij -= a_ik a_kj for every
   j > k present in both rows i and k (the diagonal included); when row i is
   done its diagonal is replaced by d_i = 1/u_ii.  lu holds the strict lower
   part of the unit-lower L, the strict upper part of U and d on the
   diagonal. */
int orc_ilu0(int n, const int *rp, const int *ci, const double *val, double *lu) {
  memcpy(lu, val, (size_t)rp[n] * sizeof(double));
  int *dg = (int *)malloc((size_t)n * sizeof(int));
  for (int i = 0; i < n; i++) {
    dg[i] = -1;
    for (int p = rp[i]; p < rp[i + 1]; p++)
      if (ci[p] == i) dg[i] = p;
    if (dg[i] < 0) { free(dg); return -1; }
  }
  for (int i = 0; i < n; i++) {
    for (int p = rp[i]; p < dg[i]; p++) {
      const int k = ci[p];
      const double aik = lu[p] * lu[dg[k]];
      lu[p] = aik;
      int r = p + 1, q = dg[k] + 1;
      while (r < rp[i + 1] && q < rp[k + 1]) {
        if (ci[r] == ci[q]) { lu[r] -= aik * lu[q]; r++; q++; }
        else if (ci[r] < ci[q]) r++;
        else q++;
      }
    }
    lu[dg[i]] = 1.0 / lu[dg[i]];
  }
  free(dg);
  return 0;
}

/* out = U^-1 L^-1 in with the factors of orc_ilu0: forward substitution in
   row order, then backward with the stored reciprocal diagonal (the
   reference's lar_perform_rhsub for 1x1 blocks, las_block_intf.c:3082-3490) */
void orc_ilu_solve(int n, const int *rp, const int *ci, const double *lu, const double *in,
                   double *out) {
  for (int i = 0; i < n; i++) {
    double s = in[i];
    for (int p = rp[i]; p < rp[i + 1] && ci[p] < i; p++) s -= lu[p] * out[ci[p]];
    out[i] = s;
  }
  for (int i = n - 1; i >= 0; i--) {
    double s = out[i], d = 0.0;
    for (int p = rp[i]; p < rp[i + 1]; p++) {
      if (ci[p] > i) s -= lu[p] * out[ci[p]];
      else if (ci[p] == i) d = lu[p];
    }
    out[i] = s * d;
  }
}

/* Gauss-Seidel sweeps in a given row order (NULL: natural order): every
   row takes the newest values of the rows before it in the order and the
   old values of the rows after it, v_r = d_r (b_r - sum_{j != r} a_rj v_j)
   with the sum taken as the reference does (vloc -= a v over the row in
   column order, then + b, then * d; las_crs_intf.c:905-998).  In natural
   order this is the reference's GS with one thread; the device runs it in a
   multicolour order, which the order argument reproduces. */
static const int *g_gs_order = NULL;
static int g_gs_sweeps = 1;

void orc_set_gs(const int *order, int nsweeps) {
  g_gs_order = order;
  g_gs_sweeps = nsweeps;
}

void orc_gs_sweeps(int n, const int *row_ptr, const int *col_ind, const double *val,
                   const int *diag_ptr, const double *diag_precon, const int *order,
                   int nsweeps, int use_rhs, const double *b, double *v) {
  for (int it = 0; it < nsweeps; it++)
    for (int q = 0; q < n; q++) {
      const int r = order ? order[q] : q;
      double vloc = 0.0;
      for (int k = row_ptr[r]; k < row_ptr[r + 1]; k++)
        if (k != diag_ptr[r]) vloc -= val[k] * v[col_ind[k]];
      if (use_rhs) vloc += b[r];
      v[r] = vloc * diag_precon[r];
    }
}

static void compreres(const orc_op *op, int control, const double *x, double *v) {
  int n = op->n;
  if (op->precon == 0) {
    orc_compute_residual(n, op->rp, op->ci, op->val, control, op->b, x, v);
  } else if (op->precon == 2) {
    /* BLOCK_ILU branch (lss_mkb_core.c:1947-1997): V = M^-1 (B - A X) */
    orc_compute_residual(n, op->rp, op->ci, op->val, control, op->b, x, op->tmp);
    orc_ilu_solve(n, op->rp, op->ci, op->lu, op->tmp, v);
  } else if (op->precon == 3) {
    /* smoothing branch with GS sweeps (lss_mkb_core.c:2001-2054) */
    memcpy(v, x, (size_t)n * sizeof(double));
    orc_gs_sweeps(n, op->rp, op->ci, op->val, op->diag_ptr, op->dinv, g_gs_order, g_gs_sweeps,
                  control, op->b, v);
    for (int i = 0; i < n; i++) v[i] += -1.0 * x[i];
  } else {
    memcpy(v, x, (size_t)n * sizeof(double));
    orc_bj_iteration(n, op->rp, op->ci, op->val, op->diag_ptr, op->dinv,
                     control, op->b, v);
    for (int i = 0; i < n; i++) v[i] += -1.0 * x[i];
  }
  if (control == 0)
    for (int i = 0; i < n; i++) v[i] *= -1.0;
}

int orc_gmres(int n, const int *row_ptr, const int *col_ind, const double *val,
              const double *b, int precon, int krylbas, int max_iter,
              double tol_rel, int ini_zero, double *x, int *iters,
              double *res_rel_out, double *res_out) {
  return orc_gmres_tol(n, row_ptr, col_ind, val, b, precon, krylbas, max_iter, tol_rel, 0.0,
                       ini_zero, x, iters, res_rel_out, res_out);
}

int orc_gmres_tol(int n, const int *row_ptr, const int *col_ind, const double *val,
                  const double *b, int precon, int krylbas, int max_iter,
                  double tol_rel, double tol_abs, int ini_zero, double *x, int *iters,
                  double *res_rel_out, double *res_out) {
  const double SMALL = 1.0e-15;
  const double MAX_CONVRATE = 0.999;
  orc_op op = {n, row_ptr, col_ind, val, b, precon, NULL, NULL, NULL, NULL};
  if (precon == 2) {
    op.lu = (double *)malloc((size_t)row_ptr[n] * sizeof(double));
    op.tmp = (double *)malloc((size_t)n * sizeof(double));
    if (orc_ilu0(n, row_ptr, col_ind, val, op.lu)) { free(op.lu); free(op.tmp); return -1; }
  } else if (precon) {
    op.diag_ptr = (int *)malloc((size_t)n * sizeof(int));
    op.dinv = (double *)malloc((size_t)n * sizeof(double));
    orc_fill_jacobi(n, row_ptr, col_ind, val, op.diag_ptr, op.dinv);
  }
  int m = krylbas;
  double *v = (double *)calloc((size_t)n * (m + 1), sizeof(double));
  double *h = (double *)calloc((size_t)(m + 1) * (m + 1), sizeof(double));
  double *giv = (double *)calloc((size_t)2 * (m + 1), sizeof(double));
  /* both tolerances floored at 1e-12 (:2377-2378) */
  double tol_res = tol_abs < 1.0e-12 ? 1.0e-12 : tol_abs;
  if (tol_rel < 1.0e-12) tol_rel = 1.0e-12;
  int restrt = max_iter / m + 1, restit = 0, gmiter = 0, conv = 0;
  double beta0 = 0, beta1 = 0, rnrm2, bnrm2, convrate = 0, faux = 0;
  double resid = 0, res_rel = 0;
  if (ini_zero) memset(x, 0, (size_t)n * sizeof(double));
  do {
    compreres(&op, 1, x, v);
    rnrm2 = nrm2(n, v);
    bnrm2 = 0.0;
    if (ini_zero && gmiter == 0) bnrm2 = rnrm2;
    if (bnrm2 < SMALL * SMALL) bnrm2 = 1.0;
    if (restit == 0) {
      convrate = 0.0; faux = 0.0; beta0 = rnrm2; beta1 = beta0;
    } else if (rnrm2 < SMALL * SMALL || gmiter == 0 || beta0 < SMALL * SMALL ||
               beta1 < SMALL * SMALL) {
      convrate = 0.0; faux = 0.0; beta1 = 0.0;
    } else {
      convrate = exp(log(rnrm2 / beta0) / gmiter);
      faux = exp(log(rnrm2 / beta1) / m);
      beta1 = rnrm2;
    }
    resid = rnrm2;
    res_rel = beta0 != 0.0 ? rnrm2 / beta0 : 0.0;
    if (rnrm2 < tol_res || res_rel < tol_rel || rnrm2 < SMALL * SMALL) {
      conv = 1;
      break;
    }
    if (convrate > MAX_CONVRATE || faux > MAX_CONVRATE) break;
    double daux = 1.0 / rnrm2;
    for (int i = 0; i < n; i++) v[i] *= daux;
    h[0] = rnrm2;
    for (int i = 1; i <= m; i++) h[i] = 0.0;
    int done = 0;
    for (int ik = 1; ik <= m; ik++) {
      gmiter++;
      double *w = v + (size_t)ik * n;
      compreres(&op, 0, v + (size_t)(ik - 1) * n, w);
      int k = (m + 1) * ik;
      for (int i = 1; i <= ik; i++) h[k++] = dot(n, v + (size_t)(i - 1) * n, w);
      k = (m + 1) * ik;
      for (int i = 1; i <= ik; i++) {
        double a = -h[k++];
        const double *vi = v + (size_t)(i - 1) * n;
        for (int j = 0; j < n; j++) w[j] += a * vi[j];
      }
      k = (m + 1) * ik + ik;
      h[k] = nrm2(n, w);
      daux = 1.0 / h[k];
      for (int j = 0; j < n; j++) w[j] *= daux;
      double *col = h + (size_t)ik * (m + 1);
      for (int i = 1; i < ik; i++)
        drot1(&col[i - 1], &col[i], giv[2 * (i - 1)], giv[2 * (i - 1) + 1]);
      double ra = col[ik - 1], rb = col[ik];
      orc_drotg(&ra, &rb, &giv[2 * (ik - 1)], &giv[2 * (ik - 1) + 1]);
      drot1(&col[ik - 1], &col[ik], giv[2 * (ik - 1)], giv[2 * (ik - 1) + 1]);
      drot1(&h[ik - 1], &h[ik], giv[2 * (ik - 1)], giv[2 * (ik - 1) + 1]);
      resid = fabs(h[ik]);
      res_rel = beta0 != 0.0 ? resid / beta0 : 0.0;
      if (ik == m || gmiter == max_iter || resid < tol_res || res_rel < tol_rel) {
        /* dtrsv("U","N","N") on h[1..ik][..] then x += V y (dgemv) */
        for (int i = ik - 1; i >= 0; i--) {
          double s = h[i];
          for (int j = i + 1; j < ik; j++) s -= h[(size_t)(j + 1) * (m + 1) + i] * h[j];
          h[i] = s / h[(size_t)(i + 1) * (m + 1) + i];
        }
        for (int j = 0; j < ik; j++) {
          const double *vj = v + (size_t)j * n;
          double y = h[j];
          for (int r = 0; r < n; r++) x[r] += vj[r] * y;
        }
        if (resid < tol_res || res_rel < tol_rel) { conv = 1; done = 1; }
        else if (gmiter == max_iter) done = 1;
        break;
      }
    }
    if (done) break;
    restit++;
  } while (restit <= restrt);
  if (iters) *iters = gmiter;
  if (res_rel_out) *res_rel_out = res_rel;
  if (res_out) *res_out = resid;
  free(v); free(h); free(giv);
  free(op.diag_ptr); free(op.dinv); free(op.lu); free(op.tmp);
  return conv;
}

/* ---------------------------------------------------------------------
 * DG prism approximation (apd_dg_prism), tensor-product base
 * ------------------------------------------------------------------- */

/* 1D monomials eta^i, i <= p, and their derivatives
   (aps_dg_util.c:1371-1410: derivative built as eta*d_{i-1} then times i) */
static void dg_mono1d(int p, double eta, double *f, double *df) {
  f[0] = 1.0;
  if (p > 0) f[1] = eta;
  for (int i = 2; i <= p; i++) f[i] = eta * f[i - 1];
  if (!df) return;
  df[0] = 0.0;
  if (p > 0) df[1] = 1.0;
  for (int i = 2; i <= p; i++) df[i] = eta * df[i - 1];
  for (int i = 2; i <= p; i++) df[i] *= i;
}

/* complete 2D base of degree p on the triangle in the reference's order
   (aps_dg_util.c:997-1365): 1, e0, e1, e0 e1, then e0^i e1^j for
   j = 0,1 (i >= 2, i + j <= p) and j >= 2 (i + j <= p) */
static int dg_base2d(int p, const double *eta, double *f, double *fx, double *fy) {
  f[0] = 1.0; f[1] = eta[0]; f[2] = eta[1];
  fx[0] = 0.0; fx[1] = 1.0; fx[2] = 0.0;
  fy[0] = 0.0; fy[1] = 0.0; fy[2] = 1.0;
  if (p == 1) return 3;
  f[3] = eta[0] * eta[1];
  fx[3] = eta[1];
  fy[3] = eta[0];
  double s[16], ds[16], t[16], dt[16];
  dg_mono1d(p, eta[0], s, ds);
  dg_mono1d(p, eta[1], t, dt);
  int k = 4;
  for (int j = 0; j < 2; j++)
    for (int i = 2; i <= p - j; i++) {
      f[k] = s[i] * t[j]; fx[k] = ds[i] * t[j]; fy[k] = s[i] * dt[j]; k++;
    }
  for (int j = 2; j <= p; j++)
    for (int i = 0; i <= p - j; i++) {
      f[k] = s[i] * t[j]; fx[k] = ds[i] * t[j]; fy[k] = s[i] * dt[j]; k++;
    }
  return k;
}

/* prism bases of the DG module (apr_shape_fun_3D, aps_dg_util.c:408-992).
   Tensor base, pdeg = 100*pz + px:
     pz = 0 (pdeg 1, 2; :495-600): 1, e0, e1 (, e0^2, e0 e1, e1^2), constant
     in z;
     px, pz >= 1: 1, e0, e1, z, z e0, z e1, then the 2D functions i >= 3
     times z^0 and z^1, then all 2D functions times z^j, j >= 2.
   Complete base of degree p, encoded ORC_DG_COMPLETE + p (:566-870):
     p = 1: 1, e0, e1, z; p = 2: 1, e0, e1, z, z e0, z e1, e0 e1, e0^2, e1^2,
     z^2; p >= 3: 1, e0, e1, z, z e0, z e1, e0 e1, e0 e1 z, then the 2D
     complete functions times z^k in the loops of :822-870 (total degree
     <= p). */
int orc_dg_shape(int pdeg, const double *eta, double *phi, double *dx, double *dy,
                 double *dz) {
  const int complete = pdeg > ORC_DG_COMPLETE;
  const int px = complete ? pdeg - ORC_DG_COMPLETE : pdeg % 100;
  const int pz = complete ? px : pdeg / 100;
  if (px < 1 || px > 3 || pz < 0 || (pz == 0 && px > 2)) return -1;
  if (pz == 0) {
    phi[0] = 1.0; phi[1] = eta[0]; phi[2] = eta[1];
    dx[0] = 0.0; dx[1] = 1.0; dx[2] = 0.0;
    dy[0] = 0.0; dy[1] = 0.0; dy[2] = 1.0;
    dz[0] = 0.0; dz[1] = 0.0; dz[2] = 0.0;
    if (px == 1) return 3;
    phi[3] = eta[0] * eta[0]; phi[4] = eta[0] * eta[1]; phi[5] = eta[1] * eta[1];
    dx[3] = 2 * eta[0]; dx[4] = eta[1]; dx[5] = 0.0;
    dy[3] = 0.0; dy[4] = eta[0]; dy[5] = 2 * eta[1];
    dz[3] = 0.0; dz[4] = 0.0; dz[5] = 0.0;
    return 6;
  }
  phi[0] = 1.0; phi[1] = eta[0]; phi[2] = eta[1]; phi[3] = eta[2];
  dx[0] = 0.0; dx[1] = 1.0; dx[2] = 0.0; dx[3] = 0.0;
  dy[0] = 0.0; dy[1] = 0.0; dy[2] = 1.0; dy[3] = 0.0;
  dz[0] = 0.0; dz[1] = 0.0; dz[2] = 0.0; dz[3] = 1.0;
  if (complete && px == 1) return 4;
  phi[4] = eta[2] * eta[0]; phi[5] = eta[2] * eta[1];
  dx[4] = eta[2]; dx[5] = 0.0;
  dy[4] = 0.0; dy[5] = eta[2];
  dz[4] = eta[0]; dz[5] = eta[1];
  if (!complete && px == 1 && pz == 1) return 6;
  double f[64], fx[64], fy[64], u[16], du[16];
  if (complete) {
    phi[6] = eta[1] * eta[0];
    dx[6] = eta[1]; dy[6] = eta[0]; dz[6] = 0.0;
    if (px == 2) {
      phi[7] = eta[0] * eta[0]; phi[8] = eta[1] * eta[1]; phi[9] = eta[2] * eta[2];
      dx[7] = 2.0 * eta[0]; dx[8] = 0.0; dx[9] = 0.0;
      dy[7] = 0.0; dy[8] = 2.0 * eta[1]; dy[9] = 0.0;
      dz[7] = 0.0; dz[8] = 0.0; dz[9] = 2.0 * eta[2];
      return 10;
    }
    phi[7] = eta[2] * eta[1] * eta[0];
    dx[7] = eta[2] * eta[1]; dy[7] = eta[2] * eta[0]; dz[7] = eta[1] * eta[0];
    dg_base2d(px, eta, f, fx, fy);
    dg_mono1d(px, eta[2], u, du);
    const int p = px;
    int kk = 8, ii = 0;
    for (int j = 0; j < 2; j++)
      for (int i = 0; i < 2; i++) {
        for (int k = 2; k <= p - i - j; k++) {
          phi[kk] = f[ii] * u[k]; dx[kk] = fx[ii] * u[k]; dy[kk] = fy[ii] * u[k];
          dz[kk] = f[ii] * du[k]; kk++;
        }
        ii++;
      }
    for (int j = 0; j < 2; j++)
      for (int i = 2; i <= p - j; i++) {
        for (int k = 0; k <= p - i - j; k++) {
          phi[kk] = f[ii] * u[k]; dx[kk] = fx[ii] * u[k]; dy[kk] = fy[ii] * u[k];
          dz[kk] = f[ii] * du[k]; kk++;
        }
        ii++;
      }
    for (int j = 2; j <= p; j++)
      for (int i = 0; i <= p - j; i++) {
        for (int k = 0; k <= p - i - j; k++) {
          phi[kk] = f[ii] * u[k]; dx[kk] = fx[ii] * u[k]; dy[kk] = fy[ii] * u[k];
          dz[kk] = f[ii] * du[k]; kk++;
        }
        ii++;
      }
    return kk;
  }
  int n2 = dg_base2d(px, eta, f, fx, fy);
  dg_mono1d(pz, eta[2], u, du);
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

/* number of shape functions (apr_get_el_pdeg_numshap,
   aps_dg_prism_intf.c:725-780) */
int orc_dg_ns(int pdeg) {
  if (pdeg > ORC_DG_COMPLETE) {
    const int p = pdeg - ORC_DG_COMPLETE;
    return (p + 1) * (p + 2) * (p + 3) / 6;
  }
  const int px = pdeg % 100, pz = pdeg / 100;
  return (pz + 1) * (px + 1) * (px + 2) / 2;
}

/* penalty degree of a face (pds_conv_diff_weakform_dg.c:516-527): complete
   base min(p1, p2) (boundary: p1); tensor base min over the neighbours of
   min(px, pz), which is zero for the pz = 0 degrees; pdeg2 <= 0: boundary */
int orc_dg_penalty_degree(int pdeg1, int pdeg2) {
  if (pdeg1 > ORC_DG_COMPLETE) {
    const int p1 = pdeg1 - ORC_DG_COMPLETE;
    if (pdeg2 <= 0) return p1;
    const int p2 = pdeg2 - ORC_DG_COMPLETE;
    return p1 < p2 ? p1 : p2;
  }
  int d = pdeg1 / 100 < pdeg1 % 100 ? pdeg1 / 100 : pdeg1 % 100;
  if (pdeg2 > 0) {
    const int e = pdeg2 / 100 < pdeg2 % 100 ? pdeg2 / 100 : pdeg2 % 100;
    if (e < d) d = e;
  }
  return d;
}

/* the tensor degree whose volume and face rules a complete-base degree
   uses (aps_dg_util.c:1446-1451, 1531-1536: orders 2p in x and z) */
static int dg_rule_pdeg(int pdeg) {
  return pdeg > ORC_DG_COMPLETE ? 101 * (pdeg - ORC_DG_COMPLETE) : pdeg;
}

/* Gauss-Legendre points and weights on [-1,1] (Newton on P_n) */
static void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; i++) {
    double z = cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5)), dp = 0.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; k++) {
        double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1; p1 = p2;
      }
      if (n == 1) { p1 = z; p0 = 1.0; }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-17) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; k++) {
      double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1; p1 = p2;
    }
    if (n == 1) { p1 = z; p0 = 1.0; }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[n - 1 - i] = z;
    w[n - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
  if (n == 1) { x[0] = 0.0; w[0] = 2.0; }
}

/* symmetric triangle rules of degree 1, 2, 4, 6 (Dunavant, Int. J. Numer.
   Meth. Engng. 21 (1985) 1129-1148; 1, 3, 6, 12 points), barycentric
   points (l1, l2, l3), weights summing to 1; the reference selects them by
   degree (aps_gauss_util.c:27-150) */
static int dunavant(int order, double *l, double *w) {
  int n = 0;
#define ORB3(a, b, ww)                                                    \
  do {                                                                    \
    const double A_ = (a), B_ = (b);                                      \
    l[3 * n] = A_; l[3 * n + 1] = B_; l[3 * n + 2] = B_; w[n++] = (ww);   \
    l[3 * n] = B_; l[3 * n + 1] = B_; l[3 * n + 2] = A_; w[n++] = (ww);   \
    l[3 * n] = B_; l[3 * n + 1] = A_; l[3 * n + 2] = B_; w[n++] = (ww);   \
  } while (0)
  if (order <= 1) {
    l[0] = l[1] = l[2] = 1.0 / 3.0; w[0] = 1.0; return 1;
  } else if (order == 2) {
    ORB3(0.666666666666667, 0.166666666666667, 0.333333333333333);
  } else if (order == 4) {
    ORB3(0.108103018168070, 0.445948490915965, 0.223381589678011);
    ORB3(0.816847572980459, 0.091576213509771, 0.109951743655322);
  } else if (order == 6) {
    ORB3(0.501426509658179, 0.249286745170910, 0.116786275726379);
    ORB3(0.873821971016996, 0.063089014491502, 0.050844906370207);
    const double a = 0.053145049844817, b = 0.310352451033784, c = 0.636502499121399;
    const double p6[6][3] = {{a, b, c}, {b, c, a}, {c, a, b}, {c, b, a}, {a, c, b}, {b, a, c}};
    for (int q = 0; q < 6; q++) {
      l[3 * n] = p6[q][0]; l[3 * n + 1] = p6[q][1]; l[3 * n + 2] = p6[q][2];
      w[n++] = 0.082851075618374;
    }
  } else {
    return -1;
  }
#undef ORB3
  return n;
}

/* prism quadrature of the DG module (apr_set_quadr_3D, aps_dg_util.c:
   1415-1496): degree 2 px on the triangle x degree 2 pz on the line; point
   (l2, l3, z), weight 1/2 w_tri w_line */
int orc_dg_quadr(int pdeg, double *xg, double *wg) {
  pdeg = dg_rule_pdeg(pdeg);
  int ox = 2 * (pdeg % 100), oz = 2 * (pdeg / 100);
  if (ox == 0) ox = 1;
  if (oz == 0) oz = 1;
  double l[3 * 16], wt[16], xz[12], wz[12];
  int n2 = dunavant(ox, l, wt);
  if (n2 < 0) return -1;
  int nz = oz / 2 + 1;
  gauss_legendre(nz, xz, wz);
  for (int ki = 0; ki < nz; ki++)
    for (int kj = 0; kj < n2; kj++) {
      const int g = ki * n2 + kj;
      xg[3 * g] = l[3 * kj + 1];
      xg[3 * g + 1] = l[3 * kj + 2];
      xg[3 * g + 2] = xz[ki];
      wg[g] = 0.5 * wt[kj] * wz[ki];
    }
  return nz * n2;
}

/* DG element integral (apr_num_int_el, aps_dg_prism_intf.c:1239-1728),
   nreq = 1, column-major stiff[j*ns + i] (i: test function):
     m phi_j phi_i + sum_rc a_rc d_c phi_j d_r phi_i (full branch when
     ind[10], diagonal-free 2D branch when only ind[2]) + sum_k b_k phi_j
     d_k phi_i + c phi_j phi_i;  rhs: q . grad phi_i + s phi_i.
   The convection term carries the derivative on the test function (the
   reference's DG form), unlike the std-linear module. */
int orc_dg_num_int_el(int pdeg, const double *nc, const int *ind, const double *coef,
                      int src_model, double *stiff, double *rhs) {
  double xg[3 * 64 * 12], wg[64 * 12];
  const int ng = orc_dg_quadr(pdeg, xg, wg);
  if (ng < 0) return -1;
  double phi[128], dx[128], dy[128], dz[128];
  int ns = 0;
  const double *a = coef + ORC_C_A, *b = coef + ORC_C_B, *q = coef + ORC_C_Q;
  for (int ki = 0; ki < ng; ki++) {
    const double *eta = &xg[3 * ki];
    ns = orc_dg_shape(pdeg, eta, phi, dx, dy, dz);
    double gp[6], gdx[6], gdy[6], gdz[6];
    gp[0] = (1.0 - eta[0] - eta[1]) * (1.0 - eta[2]) / 2.0;
    gp[1] = eta[0] * (1.0 - eta[2]) / 2.0;
    gp[2] = eta[1] * (1.0 - eta[2]) / 2.0;
    gp[3] = (1.0 - eta[0] - eta[1]) * (1.0 + eta[2]) / 2.0;
    gp[4] = eta[0] * (1.0 + eta[2]) / 2.0;
    gp[5] = eta[1] * (1.0 + eta[2]) / 2.0;
    gdx[0] = -(1.0 - eta[2]) / 2.0; gdx[1] = (1.0 - eta[2]) / 2.0; gdx[2] = 0.0;
    gdx[3] = -(1.0 + eta[2]) / 2.0; gdx[4] = (1.0 + eta[2]) / 2.0; gdx[5] = 0.0;
    gdy[0] = -(1.0 - eta[2]) / 2.0; gdy[1] = 0.0; gdy[2] = (1.0 - eta[2]) / 2.0;
    gdy[3] = -(1.0 + eta[2]) / 2.0; gdy[4] = 0.0; gdy[5] = (1.0 + eta[2]) / 2.0;
    gdz[0] = -(1.0 - eta[0] - eta[1]) / 2.0; gdz[1] = -eta[0] / 2.0; gdz[2] = -eta[1] / 2.0;
    gdz[3] = (1.0 - eta[0] - eta[1]) / 2.0; gdz[4] = eta[0] / 2.0; gdz[5] = eta[1] / 2.0;
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Ji[9], xc[3] = {0, 0, 0};
    for (int i = 0; i < 6; i++) {
      J[0] += nc[3 * i] * gdx[i]; J[1] += nc[3 * i] * gdy[i]; J[2] += nc[3 * i] * gdz[i];
      J[3] += nc[3 * i + 1] * gdx[i]; J[4] += nc[3 * i + 1] * gdy[i]; J[5] += nc[3 * i + 1] * gdz[i];
      J[6] += nc[3 * i + 2] * gdx[i]; J[7] += nc[3 * i + 2] * gdy[i]; J[8] += nc[3 * i + 2] * gdz[i];
      xc[0] += nc[3 * i] * gp[i]; xc[1] += nc[3 * i + 1] * gp[i]; xc[2] += nc[3 * i + 2] * gp[i];
    }
    const double det = orc_mat3_inv(J, Ji);
    for (int i = 0; i < ns; i++) {
      const double ax = dx[i] * Ji[0] + dy[i] * Ji[3] + dz[i] * Ji[6];
      const double ay = dx[i] * Ji[1] + dy[i] * Ji[4] + dz[i] * Ji[7];
      const double az = dx[i] * Ji[2] + dy[i] * Ji[5] + dz[i] * Ji[8];
      dx[i] = ax; dy[i] = ay; dz[i] = az;
    }
    const double vol = fabs(det) * wg[ki];
    double s = coef[ORC_C_S];
    if (src_model != ORC_SRC_CONST) {
      double u, lapl;
      exact_sol(src_model, xc, &u, &lapl);
      s = -lapl;
    }
    for (int j = 0; j < ns; j++)
      for (int i = 0; i < ns; i++) {
        double k = 0.0;
        if (ind[ORC_IND_M]) k += coef[ORC_C_M] * phi[j] * phi[i] * vol;
        if (ind[ORC_IND_AZZ])
          k += (a[0] * dx[j] * dx[i] + a[1] * dy[j] * dx[i] + a[2] * dz[j] * dx[i] +
                a[3] * dx[j] * dy[i] + a[4] * dy[j] * dy[i] + a[5] * dz[j] * dy[i] +
                a[6] * dx[j] * dz[i] + a[7] * dy[j] * dz[i] + a[8] * dz[j] * dz[i]) * vol;
        else if (ind[ORC_IND_AXX])
          k += (a[0] * dx[j] * dx[i] + a[1] * dy[j] * dx[i] + a[3] * dx[j] * dy[i] +
                a[4] * dy[j] * dy[i]) * vol;
        if (ind[ORC_IND_BZ])
          k += (b[0] * phi[j] * dx[i] + b[1] * phi[j] * dy[i] + b[2] * phi[j] * dz[i]) * vol;
        else if (ind[ORC_IND_BX])
          k += (b[0] * phi[j] * dx[i] + b[1] * phi[j] * dy[i]) * vol;
        if (ind[ORC_IND_C]) k += coef[ORC_C_C] * phi[j] * phi[i] * vol;
        stiff[j * ns + i] += k;
      }
    for (int i = 0; i < ns; i++) {
      double f = 0.0;
      if (ind[ORC_IND_QZ]) f += (q[0] * dx[i] + q[1] * dy[i] + q[2] * dz[i]) * vol;
      else if (ind[ORC_IND_QX]) f += (q[0] * dx[i] + q[1] * dy[i]) * vol;
      if (ind[ORC_IND_S]) f += s * phi[i] * vol;
      rhs[i] += f;
    }
  }
  return ns;
}

/* ---------------------------------------------------------------------
 * DG prism faces (pdr_conv_diff_comp_fa_stiff_mat_app, pdd_conv_diff/
 * weak_formulation/approx_dg/pds_conv_diff_weakform_dg.c:232-1400)
 * ------------------------------------------------------------------- */

/* prism face vertices, quads in cyclic order (bottom edge first), from the
   face-node table mmd_prism/mms_prism_intf.c:98 */
static const int DG_FACE_V[5][4] = {{0, 1, 2, -1}, {3, 4, 5, -1}, {0, 1, 4, 3}, {1, 2, 5, 4},
                                    {2, 0, 3, 5}};

/* linear prism map at eta: J[3r+c] = dx_r/deta_c, point x; returns det */
static double dg_geom(const double *nc, const double *eta, double *J, double *x) {
  const double e0 = eta[0], e1 = eta[1], z = eta[2];
  double gp[6], gx[6], gy[6], gz[6];
  gp[0] = (1.0 - e0 - e1) * (1.0 - z) / 2.0; gp[1] = e0 * (1.0 - z) / 2.0;
  gp[2] = e1 * (1.0 - z) / 2.0; gp[3] = (1.0 - e0 - e1) * (1.0 + z) / 2.0;
  gp[4] = e0 * (1.0 + z) / 2.0; gp[5] = e1 * (1.0 + z) / 2.0;
  gx[0] = -(1.0 - z) / 2.0; gx[1] = (1.0 - z) / 2.0; gx[2] = 0.0;
  gx[3] = -(1.0 + z) / 2.0; gx[4] = (1.0 + z) / 2.0; gx[5] = 0.0;
  gy[0] = -(1.0 - z) / 2.0; gy[1] = 0.0; gy[2] = (1.0 - z) / 2.0;
  gy[3] = -(1.0 + z) / 2.0; gy[4] = 0.0; gy[5] = (1.0 + z) / 2.0;
  gz[0] = -(1.0 - e0 - e1) / 2.0; gz[1] = -e0 / 2.0; gz[2] = -e1 / 2.0;
  gz[3] = (1.0 - e0 - e1) / 2.0; gz[4] = e0 / 2.0; gz[5] = e1 / 2.0;
  for (int k = 0; k < 9; k++) J[k] = 0.0;
  x[0] = x[1] = x[2] = 0.0;
  for (int i = 0; i < 6; i++)
    for (int r = 0; r < 3; r++) {
      J[3 * r] += nc[3 * i + r] * gx[i];
      J[3 * r + 1] += nc[3 * i + r] * gy[i];
      J[3 * r + 2] += nc[3 * i + r] * gz[i];
      x[r] += nc[3 * i + r] * gp[i];
    }
  return J[0] * (J[4] * J[8] - J[7] * J[5]) + J[3] * (J[7] * J[2] - J[1] * J[8]) +
         J[6] * (J[1] * J[5] - J[4] * J[2]);
}

/* element-local coordinates of a face quadrature point seen from the face's
   first neighbour (mmr_fa_elem_coor, mmd_prism/mms_prism_datstr.c:1541-1700):
   triangles take (x0, x1) in [0,1]^2, quads (x0, x1) in [-1,1]^2 */
void orc_dg_face_eta(int side, const double *xf, double *eta) {
  switch (side) {
    case 0: eta[0] = xf[1]; eta[1] = xf[0]; eta[2] = -1.0; break;
    case 1: eta[0] = xf[0]; eta[1] = xf[1]; eta[2] = 1.0; break;
    case 2: eta[0] = 0.5 * (xf[0] + 1.0); eta[1] = 0.0; eta[2] = xf[1]; break;
    case 3: eta[0] = 0.5 * (-xf[0] + 1.0); eta[1] = 1.0 - eta[0]; eta[2] = xf[1]; break;
    default: eta[0] = 0.0; eta[1] = 0.5 * (-xf[0] + 1.0); eta[2] = xf[1]; break;
  }
}

/* area element and outward unit normal of an element side at eta
   (apr_elem_calc_3D with Control = 3 + side, aps_dg_util.c:334-367) */
double orc_dg_face_ds(int side, const double *nc, const double *eta, double *n, double *x) {
  double J[9], ds[3];
  dg_geom(nc, eta, J, x);
  const double t0[3] = {J[0], J[3], J[6]}, t1[3] = {J[1], J[4], J[7]}, t2[3] = {J[2], J[5], J[8]};
  const double *a, *b;
  double d[3];
  int neg = 0;
  switch (side) {
    case 0: a = t0; b = t1; neg = 1; break;
    case 1: a = t0; b = t1; break;
    case 2: a = t0; b = t2; break;
    case 3:
      d[0] = t1[0] - t0[0]; d[1] = t1[1] - t0[1]; d[2] = t1[2] - t0[2];
      a = d; b = t2; break;
    default: a = t2; b = t1; break;
  }
  ds[0] = a[1] * b[2] - a[2] * b[1];
  ds[1] = a[2] * b[0] - a[0] * b[2];
  ds[2] = a[0] * b[1] - a[1] * b[0];
  if (neg) { ds[0] = -ds[0]; ds[1] = -ds[1]; ds[2] = -ds[2]; }
  const double det = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
  n[0] = ds[0] / det; n[1] = ds[1] / det; n[2] = ds[2] / det;
  return det;
}

/* face quadrature (apr_set_quadr_2D, aps_dg_util.c:1499-1555): quads are
   Gauss-Legendre of degree 2 px (first coordinate) x 2 pz on [-1,1]^2 with
   weights 1/2 wx wz, triangles the degree-2 px rule with weights 1/2 w */
int orc_dg_face_quadr(int pdeg, int quad, double *xg, double *wg) {
  pdeg = dg_rule_pdeg(pdeg);
  const int px = pdeg % 100, pz = pdeg / 100;
  if (quad) {
    double x1[8], w1[8], x2[8], w2[8];
    const int n1 = px + 1, n2 = pz + 1;
    gauss_legendre(n1, x1, w1);
    gauss_legendre(n2, x2, w2);
    for (int ki = 0; ki < n1; ki++)
      for (int kj = 0; kj < n2; kj++) {
        xg[2 * (ki * n2 + kj)] = x1[ki];
        xg[2 * (ki * n2 + kj) + 1] = x2[kj];
        wg[ki * n2 + kj] = 0.5 * w1[ki] * w2[kj];
      }
    return n1 * n2;
  }
  double l[3 * 16], w[16];
  const int n = dunavant(2 * px, l, w);
  for (int k = 0; k < n; k++) {
    xg[2 * k] = l[3 * k + 1];
    xg[2 * k + 1] = l[3 * k + 2];
    wg[k] = 0.5 * w[k];
  }
  return n;
}

/* local coordinates in the second neighbour of a point given in the first
   one: the point's barycentric (triangle) or edge/height (quad) coordinates
   on the shared face carried over through the shared vertices (g1, g2 are
   the global node numbers of the two elements' nodes) */
int orc_dg_map_point(int side1, const int *g1, int side2, const int *g2, const double *eta,
                     double *eta2) {
  if (side1 < 2) {
    const double lam[3] = {1.0 - eta[0] - eta[1], eta[0], eta[1]};
    double mu[3];
    for (int k2 = 0; k2 < 3; k2++) {
      int found = 0;
      for (int k1 = 0; k1 < 3; k1++)
        if (g2[DG_FACE_V[side2][k2]] == g1[DG_FACE_V[side1][k1]]) {
          mu[k2] = lam[k1];
          found = 1;
        }
      if (!found) return -1;
    }
    eta2[0] = mu[1]; eta2[1] = mu[2]; eta2[2] = side2 == 0 ? -1.0 : 1.0;
    return 0;
  }
  /* t runs along the bottom edge from face vertex 0 to face vertex 1 */
  double t = side1 == 2 ? eta[0] : side1 == 3 ? eta[1] : 1.0 - eta[1];
  const int a1 = g1[DG_FACE_V[side1][0]], b1 = g1[DG_FACE_V[side1][1]];
  const int a2 = g2[DG_FACE_V[side2][0]], b2 = g2[DG_FACE_V[side2][1]];
  if (a2 == a1 && b2 == b1) { /* same direction */
  } else if (a2 == b1 && b2 == a1) {
    t = 1.0 - t;
  } else {
    return -1;
  }
  eta2[2] = eta[2];
  if (side2 == 2) { eta2[0] = t; eta2[1] = 0.0; }
  else if (side2 == 3) { eta2[0] = 1.0 - t; eta2[1] = t; }
  else { eta2[0] = 0.0; eta2[1] = 1.0 - t; }
  return 0;
}

/* face area (mmr_fa_area, mms_prism_datstr.c:1732-1790): triangles 1/2|a x b|,
   quads 1/2(|a x b| + |c x d|) around the face's vertex cycle */
double orc_dg_face_area(const double *nc, int side) {
  const int nv = side < 2 ? 3 : 4;
  double p[4][3], e[4][3];
  for (int k = 0; k < nv; k++)
    for (int r = 0; r < 3; r++) p[k][r] = nc[3 * DG_FACE_V[side][k] + r];
  for (int k = 0; k < nv; k++)
    for (int r = 0; r < 3; r++) e[k][r] = p[(k + 1) % nv][r] - p[k][r];
  double c1[3], c2[3];
  c1[0] = e[0][1] * e[1][2] - e[0][2] * e[1][1];
  c1[1] = e[0][2] * e[1][0] - e[0][0] * e[1][2];
  c1[2] = e[0][0] * e[1][1] - e[0][1] * e[1][0];
  const double l1 = sqrt(c1[0] * c1[0] + c1[1] * c1[1] + c1[2] * c1[2]);
  if (nv == 3) return 0.5 * l1;
  c2[0] = e[2][1] * e[3][2] - e[2][2] * e[3][1];
  c2[1] = e[2][2] * e[3][0] - e[2][0] * e[3][2];
  c2[2] = e[2][0] * e[3][1] - e[2][1] * e[3][0];
  return 0.5 * (l1 + sqrt(c2[0] * c2[0] + c2[1] * c2[1] + c2[2] * c2[2]));
}

/* shape functions with physical gradients at eta */
static int dg_phys(int pdeg, const double *nc, const double *eta, double *phi, double *dx,
                   double *dy, double *dz) {
  const int ns = orc_dg_shape(pdeg, eta, phi, dx, dy, dz);
  double J[9], Ji[9], x[3];
  dg_geom(nc, eta, J, x);
  orc_mat3_inv(J, Ji);
  for (int i = 0; i < ns; i++) {
    const double ax = dx[i] * Ji[0] + dy[i] * Ji[3] + dz[i] * Ji[6];
    const double ay = dx[i] * Ji[1] + dy[i] * Ji[4] + dz[i] * Ji[7];
    const double az = dx[i] * Ji[2] + dy[i] * Ji[5] + dz[i] * Ji[8];
    dx[i] = ax; dy[i] = ay; dz[i] = az;
  }
  return ns;
}

static int dg_plant_downwind = 0;

/* test hook: 1 plants the downwind neighbour in the interior upwind terms */
void orc_dg_set_plant_downwind(int on) { dg_plant_downwind = on; }

/* one face's stiffness matrix and load vector for nreq = 1, steady
   (implicitness 1), as pdr_conv_diff_comp_fa_stiff_mat_app writes them
   (pds_conv_diff_weakform_dg.c:232-1400): ndofs_fa = ns1 (boundary) or
   ns1 + ns2 (interior), column-major stiff[j*ndofs_fa + i] with the first
   neighbour's functions first.  The neighbours may have different degrees:
   the face rule is the one of the neighbour with more points, the second
   one's on a tie (:508-514); the caller passes penalty = min over the
   neighbours of min(px, pz), times the penalty constant (:516-531).
   bc: 0 interior, 1 Dirichlet, 2 Neumann.  Normal fluxes An = a^T n and
   Bn = b . n of each neighbour as pdr_fa_coeff forms them
   (pds_conv_diff_weakform.c:577-740, Qn = 0); hf = sqrt(face area);
   Dirichlet data f: the exact solution when src_model selects one
   (pdr_bc_diri_coeff), else bval; Neumann data g: the exact normal
   derivative when src_model selects one, else bval. */
int orc_dg_face_int2(int pdeg1, int pdeg2, int bc, const double *nc1, const int *g1, int side1,
                     const double *coef1, const double *nc2, const int *g2, int side2,
                     const double *coef2, double penalty, double hf, int src_model, double bval,
                     double *stiff, double *rhs) {
  double xg[2 * 64], wg[64];
  int ng = orc_dg_face_quadr(pdeg1, side1 >= 2, xg, wg);
  if (bc == 0) {
    double xg2[2 * 64], wg2[64];
    const int ng2 = orc_dg_face_quadr(pdeg2, side1 >= 2, xg2, wg2);
    if (!(ng > ng2)) {
      ng = ng2;
      memcpy(xg, xg2, sizeof(double) * 2 * ng2);
      memcpy(wg, wg2, sizeof(double) * ng2);
    }
  }
  double p1[64], x1[64], y1[64], z1[64], p2[64], x2[64], y2[64], z2[64];
  int ns1 = 0, ns2 = 0, nf = 0;
  for (int ki = 0; ki < ng; ki++) {
    double eta[3], eta2[3], n[3], xc[3];
    orc_dg_face_eta(side1, &xg[2 * ki], eta);
    const double area = orc_dg_face_ds(side1, nc1, eta, n, xc) * wg[ki];
    ns1 = dg_phys(pdeg1, nc1, eta, p1, x1, y1, z1);
    const double *a = coef1 + ORC_C_A, *b = coef1 + ORC_C_B;
    double an1[3], an2[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++) an1[c] = a[c] * n[0] + a[3 + c] * n[1] + a[6 + c] * n[2];
    const double bn1 = b[0] * n[0] + b[1] * n[1] + b[2] * n[2];
    double bn2 = 0.0, nv = bn1;
    if (bc == 0) {
      if (orc_dg_map_point(side1, g1, side2, g2, eta, eta2)) return -1;
      ns2 = dg_phys(pdeg2, nc2, eta2, p2, x2, y2, z2);
      const double *a2 = coef2 + ORC_C_A, *b2 = coef2 + ORC_C_B;
      for (int c = 0; c < 3; c++) an2[c] = a2[c] * n[0] + a2[3 + c] * n[1] + a2[6 + c] * n[2];
      bn2 = b2[0] * n[0] + b2[1] * n[1] + b2[2] * n[2];
      nv = 0.5 * (bn1 + bn2);
    }
    nf = bc == 0 ? ns1 + ns2 : ns1;
    if (bc == 0) {
      double gg1[64], gg2[64];
      for (int i = 0; i < ns1; i++) gg1[i] = an1[0] * x1[i] + an1[1] * y1[i] + an1[2] * z1[i];
      for (int i = 0; i < ns2; i++) gg2[i] = an2[0] * x2[i] + an2[1] * y2[i] + an2[2] * z2[i];
      /* (1,1): test 1, trial 1 */
      for (int j = 0; j < ns1; j++)
        for (int i = 0; i < ns1; i++)
          stiff[j * nf + i] += ((p1[j] * gg1[i] - p1[i] * gg1[j]) * 0.5 +
                                penalty * p1[j] * p1[i] / hf) * area;
      /* (1,2): test 1, trial 2 */
      for (int j = 0; j < ns2; j++)
        for (int i = 0; i < ns1; i++)
          stiff[(ns1 + j) * nf + i] += ((-p2[j] * gg1[i] - p1[i] * gg2[j]) * 0.5 -
                                        penalty * p2[j] * p1[i] / hf) * area;
      /* (2,1): test 2, trial 1 */
      for (int j = 0; j < ns1; j++)
        for (int i = 0; i < ns2; i++)
          stiff[j * nf + ns1 + i] += ((p1[j] * gg2[i] + p2[i] * gg1[j]) * 0.5 -
                                      penalty * p1[j] * p2[i] / hf) * area;
      /* (2,2) */
      for (int j = 0; j < ns2; j++)
        for (int i = 0; i < ns2; i++)
          stiff[(ns1 + j) * nf + ns1 + i] += ((-p2[j] * gg2[i] + p2[i] * gg2[j]) * 0.5 +
                                              penalty * p2[j] * p2[i] / hf) * area;
      /* upwinding on the averaged normal velocity (a test hook can plant
         the downwind side instead) */
      const double nvs = dg_plant_downwind ? -nv : nv;
      if (nvs > 0) {
        for (int j = 0; j < ns1; j++) {
          for (int i = 0; i < ns1; i++) stiff[j * nf + i] += area * bn1 * p1[i] * p1[j];
          for (int i = 0; i < ns2; i++) stiff[j * nf + ns1 + i] -= area * bn2 * p2[i] * p1[j];
        }
      } else if (nvs < 0) {
        for (int j = 0; j < ns2; j++) {
          for (int i = 0; i < ns1; i++) stiff[(ns1 + j) * nf + i] += area * bn1 * p1[i] * p2[j];
          for (int i = 0; i < ns2; i++) stiff[(ns1 + j) * nf + ns1 + i] -= area * bn2 * p2[i] * p2[j];
        }
      }
    } else if (bc == 1) {
      double f = bval;
      if (src_model != ORC_SRC_CONST) exact_sol(src_model, xc, &f, NULL);
      for (int j = 0; j < ns1; j++)
        for (int i = 0; i < ns1; i++) {
          const double gi = an1[0] * x1[i] + an1[1] * y1[i] + an1[2] * z1[i];
          const double gj = an1[0] * x1[j] + an1[1] * y1[j] + an1[2] * z1[j];
          stiff[j * nf + i] += (gi * p1[j] - gj * p1[i] + penalty * p1[j] * p1[i] / hf) * area;
        }
      for (int i = 0; i < ns1; i++) {
        const double gi = an1[0] * x1[i] + an1[1] * y1[i] + an1[2] * z1[i];
        rhs[i] += (f * gi + penalty * f * p1[i] / hf) * area;
      }
      if (nv < 0) {
        for (int i = 0; i < ns1; i++) rhs[i] += -bn1 * f * p1[i] * area;
      } else {
        for (int j = 0; j < ns1; j++)
          for (int i = 0; i < ns1; i++) stiff[j * nf + i] += area * bn1 * p1[i] * p1[j];
      }
    } else {
      /* Neumann data (pdr_bc_neum_coeff, pds_conv_diff_weakform.c:1100-1140):
         the exact solution's normal derivative when src_model selects one */
      double g = bval;
      if (src_model == ORC_SRC_LINEAR) {
        g = n[0] - 2.0 * n[1] + 0.5 * n[2];
      } else if (src_model != ORC_SRC_CONST) {
        const double a1 = src_model == ORC_SRC_DIFF_IN_BOX ? 100.0 : 1.0;
        const double e = exp((-xc[0] * xc[0] - xc[1] * xc[1] - xc[2] * xc[2]) / a1);
        g = (-2.0 * xc[0] * n[0] - 2.0 * xc[1] * n[1] - 2.0 * xc[2] * n[2]) * e / a1;
      }
      if (nv > 0)
        for (int j = 0; j < ns1; j++)
          for (int i = 0; i < ns1; i++) stiff[j * nf + i] += area * bn1 * p1[i] * p1[j];
      for (int i = 0; i < ns1; i++) rhs[i] += g * p1[i] * area;
    }
  }
  return nf;
}

/* both neighbours of degree pdeg */
int orc_dg_face_int(int pdeg, int bc, const double *nc1, const int *g1, int side1,
                    const double *coef1, const double *nc2, const int *g2, int side2,
                    const double *coef2, double penalty, double hf, int src_model, double bval,
                    double *stiff, double *rhs) {
  return orc_dg_face_int2(pdeg, pdeg, bc, nc1, g1, side1, coef1, nc2, g2, side2, coef2, penalty,
                          hf, src_model, bval, stiff, rhs);
}

/* value of a DG field at a local point of one element and the point's
   physical coordinates (test helper) */
double orc_dg_eval(int pdeg, const double *nc, const double *dofs, const double *eta, double *x) {
  double phi[128], dx[128], dy[128], dz[128], J[9];
  const int ns = orc_dg_shape(pdeg, eta, phi, dx, dy, dz);
  dg_geom(nc, eta, J, x);
  double u = 0.0;
  for (int i = 0; i < ns; i++) u += dofs[i] * phi[i];
  return u;
}

/* squared L2 and H1-seminorm errors of a DG solution against the exact
   solution over one element (pdr_error_test_exact,
   pdd_conv_diff/main/pds_conv_diff_util.c:749-965, element volume rule) */
int orc_dg_errors(int pdeg, const double *nc, const double *dofs, int src_model, double *l2sq,
                  double *h1sq) {
  double xg[3 * 64 * 12], wg[64 * 12];
  const int ng = orc_dg_quadr(pdeg, xg, wg);
  double phi[128], dx[128], dy[128], dz[128];
  *l2sq = 0.0; *h1sq = 0.0;
  for (int ki = 0; ki < ng; ki++) {
    double J[9], x[3];
    const double det = fabs(dg_geom(nc, &xg[3 * ki], J, x));
    const int ns = dg_phys(pdeg, nc, &xg[3 * ki], phi, dx, dy, dz);
    double u = 0, ux = 0, uy = 0, uz = 0;
    for (int i = 0; i < ns; i++) {
      u += dofs[i] * phi[i]; ux += dofs[i] * dx[i]; uy += dofs[i] * dy[i]; uz += dofs[i] * dz[i];
    }
    const double a1 = src_model == ORC_SRC_DIFF_IN_BOX ? 100.0 : 1.0;
    const double e = exp((-x[0] * x[0] - x[1] * x[1] - x[2] * x[2]) / a1);
    const double f = e, fx = -2.0 * x[0] * e / a1, fy = -2.0 * x[1] * e / a1,
                 fz = -2.0 * x[2] * e / a1;
    const double vol = det * wg[ki];
    *l2sq += (f - u) * (f - u) * vol;
    *h1sq += ((fx - ux) * (fx - ux) + (fy - uy) * (fy - uy) + (fz - uz) * (fz - uz)) * vol;
  }
  return ng;
}

/* CPU timing driver for the DG path: element integrals of elements
   0..nel-1 and the face integrals of the listed faces (e1, s1, e2, s2 per
   face, e2 < 0 for a boundary face with type s2 = -1 Dirichlet / -2
   Neumann), every coefficient uniform; returns a checksum of the matrices */
double orc_dg_assemble_sample(int pdeg, int nel, const double *nodes, const int *gids, int nfaces,
                              const int *faces, const int *ind, const double *coef,
                              double penalty, int src_model) {
  double st[80 * 80], rhs[80], sum = 0.0;
  int ns = 0;
  for (int e = 0; e < nel; e++) {
    memset(st, 0, sizeof(st));
    memset(rhs, 0, sizeof(rhs));
    ns = orc_dg_num_int_el(pdeg, nodes + 18 * (size_t)e, ind, coef, src_model, st, rhs);
    sum += st[0] + rhs[0];
  }
  for (int f = 0; f < nfaces; f++) {
    const int *fc = faces + 4 * (size_t)f;
    const int e1 = fc[0], s1 = fc[1], e2 = fc[2], s2 = fc[3];
    const double *nc1 = nodes + 18 * (size_t)e1;
    const double hf = sqrt(orc_dg_face_area(nc1, s1));
    memset(st, 0, sizeof(st));
    memset(rhs, 0, sizeof(rhs));
    if (e2 >= 0)
      orc_dg_face_int(pdeg, 0, nc1, gids + 6 * (size_t)e1, s1, coef, nodes + 18 * (size_t)e2,
                      gids + 6 * (size_t)e2, s2, coef, penalty, hf, src_model, 0.0, st, rhs);
    else
      orc_dg_face_int(pdeg, s2 == -1 ? 1 : 2, nc1, gids + 6 * (size_t)e1, s1, coef, NULL, NULL, 0,
                      NULL, penalty, hf, src_model, 0.0, st, rhs);
    sum += st[0] + rhs[0];
  }
  (void)ns;
  return sum;
}

/* ---------------------------------------------------------------------
 * BCRS storage with 4x4 blocks (lsd_mkb/lad_bcrs/las_bcrs_intf.c)
 * ------------------------------------------------------------------- */

/* pattern (lar_allocate_SM_and_LV_bcrs, las_bcrs_intf.c:109-277): block
   rows in order, the sorted neighbour blocks with the diagonal inserted in
   place; neighbour lists flattened as nb_ptr / nb_ind.  Returns the number
   of stored blocks (col_ind may be NULL to count) */
int orc_bcrs_pattern(int nb, const int *nb_ptr, const int *nb_ind, int *row_ptr, int *col_ind) {
  int idx = 0;
  for (int r = 0; r < nb; r++) {
    if (row_ptr) row_ptr[r] = idx;
    const int cnt = nb_ptr[r + 1] - nb_ptr[r];
    int flag = 0;
    for (int i = 0; i < cnt; i++) {
      const int c = nb_ind[nb_ptr[r] + i];
      if (!flag && c > r) {
        if (col_ind) col_ind[idx] = r;
        idx++;
        flag = 1;
      }
      if (col_ind) col_ind[idx] = c;
      idx++;
    }
    if (!flag && cnt > 0) {
      if (col_ind) col_ind[idx] = r;
      idx++;
    }
  }
  if (row_ptr) row_ptr[nb] = idx;
  return idx;
}

/* assembly of one element (lar_assemble_SM_and_LV_bcrs, :530-623): block
   (row l_bl[r], col l_bl[c]) += the 4x4 sub-block of the column-major
   element matrix (4 k rows) at rows 4r.., columns 4c.. */
void orc_bcrs_assemble(const int *row_ptr, const int *col_ind, int k, const int *l_bl,
                       const double *stiff, const double *rhs_e, double *val, double *rhs) {
  const int nrdof = 4 * k;
  for (int r = 0; r < k; r++) {
    const int rg = l_bl[r];
    for (int c = 0; c < k; c++) {
      const int cg = l_bl[c];
      int ind = -1;
      for (int s = row_ptr[rg]; s < row_ptr[rg + 1]; s++)
        if (col_ind[s] == cg) { ind = s * 16; break; }
      if (ind < 0) continue;
      for (int kk = 0; kk < 4; kk++)
        for (int l = 0; l < 4; l++) val[ind + kk * 4 + l] += stiff[(c * nrdof + r) * 4 + kk * nrdof + l];
    }
    for (int l = 0; l < 4; l++) rhs[rg * 4 + l] += rhs_e[r * 4 + l];
  }
}

/* V = (use_rhs ? b : 0) - A x (lar_compute_residual_bcrs, :2787-2930):
   V = 0, then for each stored block in row order and l = 0..3, k = 0..3:
   V[4r+k] -= a_kl x[4c+l]; then V += b */
void orc_bcrs_residual(int nb, const int *row_ptr, const int *col_ind, const double *val,
                       int use_rhs, const double *x, const double *b, double *v) {
  for (int i = 0; i < 4 * nb; i++) v[i] = 0.0;
  if (x)
    for (int r = 0; r < nb; r++)
      for (int s = row_ptr[r]; s < row_ptr[r + 1]; s++) {
        const int c = col_ind[s];
        for (int l = 0; l < 4; l++)
          for (int k = 0; k < 4; k++) v[4 * r + k] -= val[16 * s + k + 4 * l] * x[4 * c + l];
      }
  if (use_rhs)
    for (int i = 0; i < 4 * nb; i++) v[i] += b[i];
}

/* LU factorisation of a column-major n x n block with partial pivoting, the
   unblocked right-looking algorithm of LAPACK's dgetf2 (pivot = first
   largest |a| in the column, rows swapped whole, column scaled by the
   reciprocal pivot, rank-1 update); ipiv 0-based.  Returns 0, or j+1 for a
   zero pivot in column j.  dgetrf runs this unblocked path for blocks
   smaller than LAPACK's block size (64), which covers the 4x4 BCRS blocks
   and DG blocks of up to 40 shape functions */
int orc_lun(int n, double *a, int *ipiv) {
  int info = 0;
  for (int j = 0; j < n; j++) {
    int jp = j;
    double amax = fabs(a[j + n * j]);
    for (int i = j + 1; i < n; i++)
      if (fabs(a[i + n * j]) > amax) { amax = fabs(a[i + n * j]); jp = i; }
    ipiv[j] = jp;
    if (a[jp + n * j] != 0.0) {
      if (jp != j)
        for (int c = 0; c < n; c++) {
          const double t = a[j + n * c]; a[j + n * c] = a[jp + n * c]; a[jp + n * c] = t;
        }
      const double r = 1.0 / a[j + n * j];
      for (int i = j + 1; i < n; i++) a[i + n * j] *= r;
    } else if (!info) {
      info = j + 1;
    }
    for (int c = j + 1; c < n; c++) {
      const double t = -a[j + n * c];
      for (int i = j + 1; i < n; i++) a[i + n * c] += a[i + n * j] * t;
    }
  }
  return info;
}

int orc_lu4(double *a, int *ipiv) { return orc_lun(4, a, ipiv); }

/* solve with the factors of orc_lu4 as LAPACK's dgetrs ('N'): row
   interchanges, unit-lower forward and upper backward column sweeps
   (dtrsm, skipping zero entries, dividing by the diagonal) */
void orc_lu4_solve(const double *lu, const int *ipiv, double *b) {
  for (int j = 0; j < 4; j++)
    if (ipiv[j] != j) { const double t = b[j]; b[j] = b[ipiv[j]]; b[ipiv[j]] = t; }
  for (int k = 0; k < 4; k++)
    if (b[k] != 0.0)
      for (int i = k + 1; i < 4; i++) b[i] -= b[k] * lu[i + 4 * k];
  for (int k = 3; k >= 0; k--)
    if (b[k] != 0.0) {
      b[k] /= lu[k + 4 * k];
      for (int i = 0; i < k; i++) b[i] -= b[k] * lu[i + 4 * k];
    }
}

/* block Jacobi / block Gauss-Seidel iterations (lar_perform_BJ_or_GS_
   iterations_bcrs, :2938-3300) as one thread runs them: Nr_prec even gives
   Nr_prec/2 loops of a forward and a backward sweep, odd gives Nr_prec/2+1
   forward sweeps; a sweep forms vloc = -(lower blocks) - (upper blocks)
   (+ b), solves with the diagonal block's factors and stores into a copy
   that is written back after the sweep.  Jacobi reads the old V everywhere;
   Gauss-Seidel reads the copy (already updated rows) for the blocks on the
   side the sweep has passed.  lu/ipiv: factors of the diagonal blocks */
void orc_bcrs_bj_gs(int nb, const int *row_ptr, const int *col_ind, const int *diag_ptr,
                    const double *val, const double *lu, const int *ipiv, int gs, int use_rhs,
                    int nr_prec, const double *b, double *v) {
  const int loops = nr_prec % 2 == 0 ? nr_prec / 2 : nr_prec / 2 + 1;
  double *tmp = (double *)malloc(sizeof(double) * 4 * (size_t)nb);
  for (int it = 0; it < loops; it++) {
    for (int dir = 0; dir < (nr_prec % 2 == 0 ? 2 : 1); dir++) {
      memcpy(tmp, v, sizeof(double) * 4 * (size_t)nb);
      for (int q = 0; q < nb; q++) {
        const int r = dir == 0 ? q : nb - 1 - q;
        double vloc[4] = {0, 0, 0, 0};
        /* blocks the sweep has passed: lower for forward, upper for backward */
        const int p0 = dir == 0 ? row_ptr[r] : diag_ptr[r] + 1;
        const int p1 = dir == 0 ? diag_ptr[r] : row_ptr[r + 1];
        const int o0 = dir == 0 ? diag_ptr[r] + 1 : row_ptr[r];
        const int o1 = dir == 0 ? row_ptr[r + 1] : diag_ptr[r];
        for (int s = p0; s < p1; s++) {
          const double *xs = gs ? tmp : v;
          const int c = col_ind[s];
          for (int l = 0; l < 4; l++)
            for (int k = 0; k < 4; k++) vloc[k] -= val[16 * s + k + 4 * l] * xs[4 * c + l];
        }
        for (int s = o0; s < o1; s++) {
          const int c = col_ind[s];
          for (int k = 0; k < 4; k++)
            for (int l = 0; l < 4; l++) vloc[k] -= val[16 * s + k + 4 * l] * v[4 * c + l];
        }
        if (use_rhs)
          for (int k = 0; k < 4; k++) vloc[k] = 1.0 * b[4 * r + k] + vloc[k];
        orc_lu4_solve(lu + 16 * (size_t)r, ipiv + 4 * (size_t)r, vloc);
        for (int k = 0; k < 4; k++) tmp[4 * r + k] = vloc[k];
      }
      memcpy(v, tmp, sizeof(double) * 4 * (size_t)nb);
    }
  }
  free(tmp);
}

/* inverse of a column-major n x n block from its orc_lun factors, the
   unblocked algorithm of LAPACK's dgetri: inv(U) column by column (dtrti2:
   a_jj = 1/a_jj, then the column above it by the inverted leading triangle
   (dtrmv, skipping zero entries) scaled by -a_jj), then for j = n-1..0 the
   column of L moved out and inv(A)(:, j) -= inv(A)(:, j+1:) L(j+1:, j)
   (dgemv), then the column interchanges in reverse order.  n <= 64 */
void orc_invn(int n, double *a, const int *ipiv) {
  for (int j = 0; j < n; j++) {
    a[j + n * j] = 1.0 / a[j + n * j];
    const double ajj = -a[j + n * j];
    double *x = a + n * j;
    for (int jj = 0; jj < j; jj++)
      if (x[jj] != 0.0) {
        const double t = x[jj];
        for (int ii = 0; ii < jj; ii++) x[ii] = x[ii] + t * a[ii + n * jj];
        x[jj] = x[jj] * a[jj + n * jj];
      }
    for (int ii = 0; ii < j; ii++) x[ii] = ajj * x[ii];
  }
  double work[64];
  for (int j = n - 1; j >= 0; j--) {
    for (int i = j + 1; i < n; i++) {
      work[i] = a[i + n * j];
      a[i + n * j] = 0.0;
    }
    for (int jj = j + 1; jj < n; jj++) {
      const double t = -1.0 * work[jj];
      for (int i = 0; i < n; i++) a[i + n * j] = a[i + n * j] + t * a[i + n * jj];
    }
  }
  for (int j = n - 2; j >= 0; j--) {
    const int jp = ipiv[j];
    if (jp != j)
      for (int i = 0; i < n; i++) {
        const double t = a[i + n * j]; a[i + n * j] = a[i + n * jp]; a[i + n * jp] = t;
      }
  }
}

void orc_inv4(double *a, const int *ipiv) { orc_invn(4, a, ipiv); }

/* block ILU(0) on the BCRS pattern (lar_fill_preconditioner_bcrs, the
   active variant 2, las_bcrs_intf.c:2181-2640), one thread: for every pivot
   row i in order, W = inv(a_ii) (dgetrf + dgetri); for every block a_ik of
   row i's upper part, the mirrored block a_ki of row k becomes a_ki W
   (sums from 0 over the inner index), and every a_kj present in row k with j
   in row i's upper part gets a_kj - a_ki a_ij (dgemm with alpha -1, beta 1:
   for each column, inner index outer, rows inner); finally a_ii = W.  The
   pattern must be structurally symmetric.  ilu: in the values, out the
   factors.  Returns 0, or -1 for a missing mirrored block, or j+1 for a
   zero pivot in a diagonal block */
int orc_bcrs_ilu0(int nb, const int *rp, const int *ci, const int *diag, double *ilu) {
  for (int i = 0; i < nb; i++) {
    if (rp[i] >= rp[i + 1]) continue;
    double *aii = ilu + 16 * (size_t)diag[i];
    double w[16];
    int piv[4];
    memcpy(w, aii, sizeof w);
    const int info = orc_lu4(w, piv);
    if (info) return info;
    orc_inv4(w, piv);
    for (int s = diag[i] + 1; s < rp[i + 1]; s++) {
      const int k = ci[s];
      int ski = -1;
      for (int t = rp[k]; t < rp[k + 1]; t++)
        if (ci[t] == i) { ski = t; break; }
      if (ski < 0) return -1;
      double *aki = ilu + 16 * (size_t)ski;
      double prod[16];
      for (int r = 0; r < 4; r++)
        for (int c = 0; c < 4; c++) {
          double d = 0.0;
          for (int l = 0; l < 4; l++) d += aki[r + 4 * l] * w[l + 4 * c];
          prod[r + 4 * c] = d;
        }
      memcpy(aki, prod, sizeof prod);
      for (int sij = diag[i] + 1; sij < rp[i + 1]; sij++) {
        const int j = ci[sij];
        int skj = -1;
        for (int t = ski + 1; t < rp[k + 1]; t++)
          if (ci[t] == j) { skj = t; break; }
        if (skj < 0) continue;
        const double *aij = ilu + 16 * (size_t)sij;
        double *akj = ilu + 16 * (size_t)skj;
        for (int c = 0; c < 4; c++)
          for (int l = 0; l < 4; l++) {
            const double t = -1.0 * aij[l + 4 * c];
            for (int r = 0; r < 4; r++) akj[r + 4 * c] = akj[r + 4 * c] + t * aki[r + 4 * l];
          }
      }
    }
    memcpy(aii, w, sizeof w);
  }
  return 0;
}

/* pattern of block ILU(k), k > 0 (lar_allocate_preconditioner_bcrs,
   las_bcrs_intf.c:751-1016): row i starts from its own stored blocks; each of
   k rings appends, for every column added by the previous ring (in list
   order), the columns of that column's original row not yet listed; the
   row's list is then sorted.  A row holds at most cap columns (R's list
   length, (int)(0.5 (k+2)^3 max_nrneig)); R stops when a row needs more, and
   this returns -1 - row then.  rp2 [nb+1] is always filled; ci2 (may be
   NULL) receives the columns.  Returns the number of blocks. */
long orc_bcrs_iluk_pattern(int nb, const int *rp, const int *ci, int k, int cap, int *rp2,
                           int *ci2) {
  int *list = (int *)malloc(sizeof(int) * (cap > 0 ? cap : 1));
  long total = 0;
  rp2[0] = 0;
  for (int i = 0; i < nb; i++) {
    int n = 0;
    for (int s = rp[i]; s < rp[i + 1]; s++) {
      if (n == cap) { free(list); return -1 - i; }
      list[n++] = ci[s];
    }
    int first = 0, last = n;
    for (int ring = 0; ring < k; ring++) {
      for (int q = first; q < last; q++) {
        const int r = list[q];
        for (int s = rp[r]; s < rp[r + 1]; s++) {
          const int c = ci[s];
          int found = 0;
          for (int t = 0; t < n && !found; t++) found = list[t] == c;
          if (found) continue;
          if (n == cap) { free(list); return -1 - i; }
          list[n++] = c;
        }
      }
      first = last;
      last = n;
    }
    for (int a = 1; a < n; a++) {  /* insertion sort */
      const int v = list[a];
      int b = a;
      while (b > 0 && list[b - 1] > v) { list[b] = list[b - 1]; b--; }
      list[b] = v;
    }
    if (ci2) memcpy(ci2 + total, list, sizeof(int) * n);
    total += n;
    rp2[i + 1] = (int)total;
  }
  free(list);
  return total;
}

/* v = (LU)^-1 b with the orc_bcrs_ilu0 factors (lar_perform_rhsub_bcrs,
   variant 0, las_bcrs_intf.c:3349-3510): forward, rows in order (rows with
   no stored blocks skipped), v_i = b_i - sum over the lower blocks of
   L_ij v_j (block by block, column outer, row inner); backward, rows in
   reverse, t = v_i - sum over the upper blocks of U_ij v_j, then v_i =
   inv(a_ii) t summed from 0 (column outer) */
void orc_bcrs_rhsub(int nb, const int *rp, const int *ci, const int *diag, const double *ilu,
                    const double *b, double *v) {
  for (int i = 0; i < nb; i++) {
    if (rp[i] >= rp[i + 1]) continue;
    double t[4];
    for (int r = 0; r < 4; r++) t[r] = b[4 * i + r];
    for (int s = rp[i]; s < diag[i]; s++) {
      const double *a = ilu + 16 * (size_t)s;
      const double *x = v + 4 * (size_t)ci[s];
      for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) t[r] -= a[r + 4 * c] * x[c];
    }
    for (int r = 0; r < 4; r++) v[4 * i + r] = t[r];
  }
  for (int i = nb - 1; i >= 0; i--) {
    double t[4];
    for (int r = 0; r < 4; r++) t[r] = v[4 * i + r];
    for (int s = diag[i] + 1; s < rp[i + 1]; s++) {
      const double *a = ilu + 16 * (size_t)s;
      const double *x = v + 4 * (size_t)ci[s];
      for (int c = 0; c < 4; c++)
        for (int r = 0; r < 4; r++) t[r] -= a[r + 4 * c] * x[c];
    }
    const double *d = ilu + 16 * (size_t)diag[i];
    for (int r = 0; r < 4; r++) v[4 * i + r] = 0.0;
    for (int c = 0; c < 4; c++)
      for (int r = 0; r < 4; r++) v[4 * i + r] += d[r + 4 * c] * t[c];
  }
}

/* ------------------------------------------------------------------
   Block ILU(0) of a matrix of variable-size blocks (the DG block system),
   lar_fill_preconditioner for block storage, MULTI_ILU / BLOCK_ILU branch
   (lad_block/las_block_intf.c:1409-2143), one thread.  Row i has its
   diagonal block at val + doff[i] (ns[i] x ns[i], column-major) and, for
   p in nptr[i]..nptr[i+1], the block a_ij, j = nbr[p], at val + ooff[p]
   (ns[i] x ns[j], column-major with leading dimension ns[i]).  For every
   row i in order, its lower neighbours k in ascending order (R sorts
   Llowerneig, :1490-1560):
     a_ik <- a_ik D_k, D_k the already inverted diagonal block (sums from 0
             over the inner index, :1880-1900);
     D_i  <- D_i - a_ik a_ki (:1930-1962);
     a_ij <- a_ij - a_ik a_kj for every neighbour j > k of i that is also a
             neighbour of k (:1975-2040);
   then D_i <- inv(D_i): 1/d for a 1x1 block, else dgetrf + dgetri (:2085-
   2110).  On exit the lower blocks hold L (without its unit diagonal), the
   upper blocks U and the diagonal blocks the inverses of U's diagonal.
   Returns 0, -1 when a_ki is missing (the pattern must be symmetric), or
   i+1 for a singular diagonal block in row i.  Blocks of at most 64. */
static int nbr_pos(const int *nptr, const int *nbr, int row, int col) {
  for (int p = nptr[row]; p < nptr[row + 1]; p++)
    if (nbr[p] == col) return p;
  return -1;
}

static int lower_sorted(const int *nptr, const int *nbr, int i, int *kk, int *pp) {
  int nl = 0;
  for (int p = nptr[i]; p < nptr[i + 1]; p++) {
    const int k = nbr[p];
    if (k >= i) continue;
    int t = nl++;
    while (t > 0 && k < kk[t - 1]) { kk[t] = kk[t - 1]; pp[t] = pp[t - 1]; t--; }
    kk[t] = k; pp[t] = p;
  }
  return nl;
}

static int upper_sorted(const int *nptr, const int *nbr, int i, int *jj, int *pp) {
  int nu = 0;
  for (int p = nptr[i]; p < nptr[i + 1]; p++) {
    const int j = nbr[p];
    if (j <= i) continue;
    int t = nu++;
    while (t > 0 && j < jj[t - 1]) { jj[t] = jj[t - 1]; pp[t] = pp[t - 1]; t--; }
    jj[t] = j; pp[t] = p;
  }
  return nu;
}

int orc_block_ilu0(int nb, const int *ns, const int *nptr, const int *nbr,
                   const long long *doff, const long long *ooff, double *val) {
  double work[64 * 64];
  int piv[64];
  for (int i = 0; i < nb; i++) {
    const int ni = ns[i], deg = nptr[i + 1] - nptr[i];
    double *di = val + doff[i];
    int *kk = (int *)malloc(sizeof(int) * (deg + 1)), *pp = (int *)malloc(sizeof(int) * (deg + 1));
    const int nl = lower_sorted(nptr, nbr, i, kk, pp);
    for (int t = 0; t < nl; t++) {
      const int k = kk[t], nk = ns[k];
      double *aik = val + ooff[pp[t]];
      const double *dk = val + doff[k];
      for (int r = 0; r < ni; r++)
        for (int c = 0; c < nk; c++) {
          double d = 0.0;
          for (int l = 0; l < nk; l++) d += aik[r + ni * l] * dk[l + nk * c];
          work[r + ni * c] = d;
        }
      memcpy(aik, work, sizeof(double) * ni * nk);
      const int q = nbr_pos(nptr, nbr, k, i);
      if (q < 0) { free(kk); free(pp); return -1; }
      const double *aki = val + ooff[q];
      for (int r = 0; r < ni; r++)
        for (int c = 0; c < ni; c++) {
          double d = 0.0;
          for (int l = 0; l < nk; l++) d += aik[r + ni * l] * aki[l + nk * c];
          di[r + ni * c] -= d;
        }
      for (int p2 = nptr[i]; p2 < nptr[i + 1]; p2++) {
        const int j = nbr[p2];
        if (j <= k) continue;
        const int q2 = nbr_pos(nptr, nbr, k, j);
        if (q2 < 0) continue;
        const int nj = ns[j];
        double *aij = val + ooff[p2];
        const double *akj = val + ooff[q2];
        for (int r = 0; r < ni; r++)
          for (int c = 0; c < nj; c++) {
            double d = 0.0;
            for (int l = 0; l < nk; l++) d += aik[r + ni * l] * akj[l + nk * c];
            aij[r + ni * c] -= d;
          }
      }
    }
    free(kk); free(pp);
    if (ni == 1) {
      if (di[0] == 0.0) return i + 1;
      di[0] = 1.0 / di[0];
    } else {
      if (orc_lun(ni, di, piv)) return i + 1;
      orc_invn(ni, di, piv);
    }
  }
  return 0;
}

/* v = (LU)^-1 b with the orc_block_ilu0 factors (lar_perform_rhsub_block,
   las_block_intf.c:3085-3476): forward, rows in order, t = b_i, then for
   the lower neighbours in ascending order t_r -= L_ik(r, c) v_k(c) one term
   at a time (row outer, column inner), v_i = t; backward, rows in reverse,
   t = v_i less the upper neighbours' terms in the same way, then v_i =
   inv(D_i) t summed from 0 (column inner).  xoff[i]: first unknown of row
   i.  b and v may be the same array. */
void orc_block_ilu_solve(int nb, const int *ns, const long long *xoff, const int *nptr,
                         const int *nbr, const long long *doff, const long long *ooff,
                         const double *val, const double *b, double *v) {
  double t[64];
  for (int i = 0; i < nb; i++) {
    const int ni = ns[i], deg = nptr[i + 1] - nptr[i];
    int *kk = (int *)malloc(sizeof(int) * (deg + 1)), *pp = (int *)malloc(sizeof(int) * (deg + 1));
    const int nl = lower_sorted(nptr, nbr, i, kk, pp);
    for (int r = 0; r < ni; r++) t[r] = b[xoff[i] + r];
    for (int q = 0; q < nl; q++) {
      const int k = kk[q], nk = ns[k];
      const double *a = val + ooff[pp[q]];
      for (int r = 0; r < ni; r++)
        for (int c = 0; c < nk; c++) t[r] -= a[r + ni * c] * v[xoff[k] + c];
    }
    for (int r = 0; r < ni; r++) v[xoff[i] + r] = t[r];
    free(kk); free(pp);
  }
  for (int i = nb - 1; i >= 0; i--) {
    const int ni = ns[i], deg = nptr[i + 1] - nptr[i];
    int *jj = (int *)malloc(sizeof(int) * (deg + 1)), *pp = (int *)malloc(sizeof(int) * (deg + 1));
    const int nu = upper_sorted(nptr, nbr, i, jj, pp);
    for (int r = 0; r < ni; r++) t[r] = v[xoff[i] + r];
    for (int q = 0; q < nu; q++) {
      const int j = jj[q], nj = ns[j];
      const double *a = val + ooff[pp[q]];
      for (int r = 0; r < ni; r++)
        for (int c = 0; c < nj; c++) t[r] -= a[r + ni * c] * v[xoff[j] + c];
    }
    const double *d = val + doff[i];
    for (int r = 0; r < ni; r++) {
      double y = 0.0;
      for (int c = 0; c < ni; c++) y += d[r + ni * c] * t[c];
      v[xoff[i] + r] = y;
    }
    free(jj); free(pp);
  }
}

/* ---------------------------------------------------------------------
 * DG on tetrahedra.  No reference counterpart: the reference's DG module is
 * prism-only (apd_dg_prism/aph_dg_prism.h:1-3).  Built from the mathematics
 * and the prism module's weak form:
 *  - base: the complete polynomials of degree p in the reference
 *    coordinates, xi^a eta^b zeta^c with a + b + c <= p, by total degree,
 *    then a descending, then b descending; ns = (p+1)(p+2)(p+3)/6;
 *  - reference vertices (0,0,0), (1,0,0), (0,1,0), (0,0,1) and the affine
 *    map x = x0 + J xi, J = [x1 - x0, x2 - x0, x3 - x0];
 *  - volume rule: collapsed Gauss-Legendre (xi = u, eta = v (1 - u),
 *    zeta = w (1 - u)(1 - v), weight (1 - u)^2 (1 - v)), p + 2 points in u
 *    and p + 1 in v and w: exact for every monomial of degree <= 2 p;
 *  - sides s = 0..3 with the mesh module's vertex sets {0,1,2}, {0,1,3},
 *    {0,2,3}, {1,2,3}; face point (c0, c1) of the reference triangle at
 *    X_a + c0 (X_b - X_a) + c1 (X_c - X_a); the prism module's triangle rule
 *    of degree 2 p; the normal points away from the opposite vertex;
 *  - the prism module's NIPG face terms, penalty and upwinding
 *    (pds_conv_diff_weakform_dg.c:232-1400) with the tet's p in place of
 *    min(px, pz).  Faces between a tet and a prism are triangles; the
 *    neighbour's local coordinates of a face point follow from the shared
 *    vertices, as orc_dg_map_point does for prisms.
 * ------------------------------------------------------------------- */
static const int DGT_FACE_V[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
static const double DGT_X[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
/* reference vertices of the prism (aps_dg_prism: triangle x [-1, 1]) */
static const double DGP_X[6][3] = {{0, 0, -1}, {1, 0, -1}, {0, 1, -1},
                                   {0, 0, 1},  {1, 0, 1},  {0, 1, 1}};

int orc_dgt_ns(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

int orc_dgt_shape(int p, const double *eta, double *phi, double *dx, double *dy, double *dz) {
  if (p < 0 || p > 3) return -1;
  double P[3][4], D[3][4];
  for (int d = 0; d < 3; d++) {
    const double x = eta[d], x2 = x * x;
    P[d][0] = 1.0; P[d][1] = x; P[d][2] = x2; P[d][3] = x2 * x;
    D[d][0] = 0.0; D[d][1] = 1.0; D[d][2] = 2.0 * x; D[d][3] = 3.0 * x2;
  }
  int n = 0;
  for (int t = 0; t <= p; t++)
    for (int a = t; a >= 0; a--)
      for (int b = t - a; b >= 0; b--) {
        const int c = t - a - b;
        phi[n] = P[0][a] * P[1][b] * P[2][c];
        dx[n] = D[0][a] * P[1][b] * P[2][c];
        dy[n] = P[0][a] * D[1][b] * P[2][c];
        dz[n] = P[0][a] * P[1][b] * D[2][c];
        n++;
      }
  return n;
}

/* exponents a | b << 2 | c << 4 of the base, in base order */
int orc_dgt_exponents(int p, int *ex) {
  int n = 0;
  for (int t = 0; t <= p; t++)
    for (int a = t; a >= 0; a--)
      for (int b = t - a; b >= 0; b--) ex[n++] = a | b << 2 | (t - a - b) << 4;
  return n;
}

int orc_dgt_quadr(int p, double *xg, double *wg) {
  if (p < 1 || p > 3) return -1;
  double tu[8], wu[8], tv[8], wv[8];
  const int nu = p + 2, nv = p + 1;
  gauss_legendre(nu, tu, wu);
  gauss_legendre(nv, tv, wv);
  int g = 0;
  for (int iu = 0; iu < nu; iu++)
    for (int iv = 0; iv < nv; iv++)
      for (int iw = 0; iw < nv; iw++) {
        const double u = 0.5 * (tu[iu] + 1.0), v = 0.5 * (tv[iv] + 1.0), w = 0.5 * (tv[iw] + 1.0);
        xg[3 * g] = u;
        xg[3 * g + 1] = v * (1.0 - u);
        xg[3 * g + 2] = w * (1.0 - u) * (1.0 - v);
        wg[g] = 0.125 * wu[iu] * wv[iv] * wv[iw] * (1.0 - u) * (1.0 - u) * (1.0 - v);
        g++;
      }
  return g;
}

/* affine map of a tet: J[3r+c] = dx_r/dxi_c, point x at eta; returns det */
static double dgt_geom(const double *nc, const double *eta, double *J, double *x) {
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) J[3 * r + c] = nc[3 * (c + 1) + r] - nc[r];
  for (int r = 0; r < 3; r++)
    x[r] = nc[r] + J[3 * r] * eta[0] + J[3 * r + 1] * eta[1] + J[3 * r + 2] * eta[2];
  return J[0] * (J[4] * J[8] - J[7] * J[5]) + J[3] * (J[7] * J[2] - J[1] * J[8]) +
         J[6] * (J[1] * J[5] - J[4] * J[2]);
}

static int dgt_phys(int p, const double *nc, const double *eta, double *phi, double *dx,
                    double *dy, double *dz) {
  const int ns = orc_dgt_shape(p, eta, phi, dx, dy, dz);
  double J[9], Ji[9], x[3];
  dgt_geom(nc, eta, J, x);
  orc_mat3_inv(J, Ji);
  for (int i = 0; i < ns; i++) {
    const double ax = dx[i] * Ji[0] + dy[i] * Ji[3] + dz[i] * Ji[6];
    const double ay = dx[i] * Ji[1] + dy[i] * Ji[4] + dz[i] * Ji[7];
    const double az = dx[i] * Ji[2] + dy[i] * Ji[5] + dz[i] * Ji[8];
    dx[i] = ax; dy[i] = ay; dz[i] = az;
  }
  return ns;
}

/* element integral of a tet, the terms and column-major layout of
   orc_dg_num_int_el */
int orc_dgt_num_int_el(int p, const double *nc, const int *ind, const double *coef,
                       int src_model, double *stiff, double *rhs) {
  double xg[3 * 128], wg[128];
  const int ng = orc_dgt_quadr(p, xg, wg);
  if (ng < 0) return -1;
  double phi[32], dx[32], dy[32], dz[32];
  int ns = 0;
  const double *a = coef + ORC_C_A, *b = coef + ORC_C_B, *q = coef + ORC_C_Q;
  for (int ki = 0; ki < ng; ki++) {
    const double *eta = &xg[3 * ki];
    ns = orc_dgt_shape(p, eta, phi, dx, dy, dz);
    double J[9], Ji[9], xc[3];
    dgt_geom(nc, eta, J, xc);
    const double det = orc_mat3_inv(J, Ji);
    for (int i = 0; i < ns; i++) {
      const double ax = dx[i] * Ji[0] + dy[i] * Ji[3] + dz[i] * Ji[6];
      const double ay = dx[i] * Ji[1] + dy[i] * Ji[4] + dz[i] * Ji[7];
      const double az = dx[i] * Ji[2] + dy[i] * Ji[5] + dz[i] * Ji[8];
      dx[i] = ax; dy[i] = ay; dz[i] = az;
    }
    const double vol = fabs(det) * wg[ki];
    double s = coef[ORC_C_S];
    if (src_model != ORC_SRC_CONST) {
      double u, lapl;
      exact_sol(src_model, xc, &u, &lapl);
      s = -lapl;
    }
    for (int j = 0; j < ns; j++)
      for (int i = 0; i < ns; i++) {
        double k = 0.0;
        if (ind[ORC_IND_M]) k += coef[ORC_C_M] * phi[j] * phi[i] * vol;
        if (ind[ORC_IND_AZZ])
          k += (a[0] * dx[j] * dx[i] + a[1] * dy[j] * dx[i] + a[2] * dz[j] * dx[i] +
                a[3] * dx[j] * dy[i] + a[4] * dy[j] * dy[i] + a[5] * dz[j] * dy[i] +
                a[6] * dx[j] * dz[i] + a[7] * dy[j] * dz[i] + a[8] * dz[j] * dz[i]) * vol;
        else if (ind[ORC_IND_AXX])
          k += (a[0] * dx[j] * dx[i] + a[1] * dy[j] * dx[i] + a[3] * dx[j] * dy[i] +
                a[4] * dy[j] * dy[i]) * vol;
        if (ind[ORC_IND_BZ])
          k += (b[0] * phi[j] * dx[i] + b[1] * phi[j] * dy[i] + b[2] * phi[j] * dz[i]) * vol;
        else if (ind[ORC_IND_BX])
          k += (b[0] * phi[j] * dx[i] + b[1] * phi[j] * dy[i]) * vol;
        if (ind[ORC_IND_C]) k += coef[ORC_C_C] * phi[j] * phi[i] * vol;
        stiff[j * ns + i] += k;
      }
    for (int i = 0; i < ns; i++) {
      double f = 0.0;
      if (ind[ORC_IND_QZ]) f += (q[0] * dx[i] + q[1] * dy[i] + q[2] * dz[i]) * vol;
      else if (ind[ORC_IND_QX]) f += (q[0] * dx[i] + q[1] * dy[i]) * vol;
      if (ind[ORC_IND_S]) f += s * phi[i] * vol;
      rhs[i] += f;
    }
  }
  return ns;
}

void orc_dgt_face_eta(int side, const double *xf, double *eta) {
  const double *A = DGT_X[DGT_FACE_V[side][0]], *B = DGT_X[DGT_FACE_V[side][1]],
               *C = DGT_X[DGT_FACE_V[side][2]];
  for (int d = 0; d < 3; d++) eta[d] = A[d] + xf[0] * (B[d] - A[d]) + xf[1] * (C[d] - A[d]);
}

/* area element |(x_b - x_a) x (x_c - x_a)| (twice the face area; the face
   rule's weights sum to 1/2) and the normal pointing away from the vertex
   opposite the side; x = the point at eta */
double orc_dgt_face_ds(int side, const double *nc, const double *eta, double *n, double *x) {
  double J[9];
  dgt_geom(nc, eta, J, x);
  const int a = DGT_FACE_V[side][0], b = DGT_FACE_V[side][1], c = DGT_FACE_V[side][2];
  const int o = 6 - a - b - c;
  double u[3], v[3], ds[3], w[3];
  for (int d = 0; d < 3; d++) {
    u[d] = nc[3 * b + d] - nc[3 * a + d];
    v[d] = nc[3 * c + d] - nc[3 * a + d];
    w[d] = nc[3 * o + d] - nc[3 * a + d];
  }
  ds[0] = u[1] * v[2] - u[2] * v[1];
  ds[1] = u[2] * v[0] - u[0] * v[2];
  ds[2] = u[0] * v[1] - u[1] * v[0];
  if (ds[0] * w[0] + ds[1] * w[1] + ds[2] * w[2] > 0) {
    ds[0] = -ds[0]; ds[1] = -ds[1]; ds[2] = -ds[2];
  }
  const double det = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
  n[0] = ds[0] / det; n[1] = ds[1] / det; n[2] = ds[2] / det;
  return det;
}

double orc_dgt_face_area(const double *nc, int side) {
  double n[3], x[3], eta[3];
  const double xf[2] = {0.0, 0.0};
  orc_dgt_face_eta(side, xf, eta);
  return 0.5 * orc_dgt_face_ds(side, nc, eta, n, x);
}

/* local coordinates in the neighbour (type t2: 0 prism, 1 tet, side s2,
   global node numbers g2) of a triangle-face point given by the local
   coordinates eta of element (t1, s1, g1): the point's barycentric
   coordinates on the face carried over through the shared vertices */
int orc_dgx_map_point(int t1, int s1, const int *g1, const double *eta, int t2, int s2,
                      const int *g2, double *eta2) {
  const int *f1 = t1 ? DGT_FACE_V[s1] : DG_FACE_V[s1];
  const int *f2 = t2 ? DGT_FACE_V[s2] : DG_FACE_V[s2];
  double lam[3];
  if (t1) {
    const double bary[4] = {1.0 - eta[0] - eta[1] - eta[2], eta[0], eta[1], eta[2]};
    for (int k = 0; k < 3; k++) lam[k] = bary[f1[k]];
  } else {
    if (s1 >= 2) return -1;
    lam[0] = 1.0 - eta[0] - eta[1]; lam[1] = eta[0]; lam[2] = eta[1];
  }
  eta2[0] = eta2[1] = eta2[2] = 0.0;
  for (int k = 0; k < 3; k++) {
    int v2 = -1;
    for (int m = 0; m < 3; m++)
      if (g2[f2[m]] == g1[f1[k]]) v2 = f2[m];
    if (v2 < 0) return -1;
    const double *X = t2 ? DGT_X[v2] : DGP_X[v2];
    for (int d = 0; d < 3; d++) eta2[d] += lam[k] * X[d];
  }
  if (!t2) {
    if (s2 >= 2) return -1;
    eta2[2] = s2 == 0 ? -1.0 : 1.0;
  }
  return 0;
}

/* one face with at least one tet (a triangle), the terms and layout of
   orc_dg_face_int2: t = 0 prism (pdeg as there), 1 tet (pdeg = p) */
int orc_dgx_face_int(int t1, int pdeg1, int t2, int pdeg2, int bc, const double *nc1,
                     const int *g1, int side1, const double *coef1, const double *nc2,
                     const int *g2, int side2, const double *coef2, double penalty, double hf,
                     int src_model, double bval, double *stiff, double *rhs) {
  if (!t1 && (bc != 0 ? 1 : !t2))
    return orc_dg_face_int2(pdeg1, pdeg2, bc, nc1, g1, side1, coef1, nc2, g2, side2, coef2,
                            penalty, hf, src_model, bval, stiff, rhs);
  double xg[2 * 64], wg[64];
  int ng = orc_dg_face_quadr(t1 ? pdeg1 : pdeg1 % 100, 0, xg, wg);
  if (bc == 0) {
    double xg2[2 * 64], wg2[64];
    const int ng2 = orc_dg_face_quadr(t2 ? pdeg2 : pdeg2 % 100, 0, xg2, wg2);
    if (!(ng > ng2)) {
      ng = ng2;
      memcpy(xg, xg2, sizeof(double) * 2 * ng2);
      memcpy(wg, wg2, sizeof(double) * ng2);
    }
  }
  double p1[64], x1[64], y1[64], z1[64], p2[64], x2[64], y2[64], z2[64];
  int ns1 = 0, ns2 = 0, nf = 0;
  for (int ki = 0; ki < ng; ki++) {
    double eta[3], eta2[3], n[3], xc[3], area;
    if (t1) {
      orc_dgt_face_eta(side1, &xg[2 * ki], eta);
      area = orc_dgt_face_ds(side1, nc1, eta, n, xc) * wg[ki];
      ns1 = dgt_phys(pdeg1, nc1, eta, p1, x1, y1, z1);
    } else {
      orc_dg_face_eta(side1, &xg[2 * ki], eta);
      area = orc_dg_face_ds(side1, nc1, eta, n, xc) * wg[ki];
      ns1 = dg_phys(pdeg1, nc1, eta, p1, x1, y1, z1);
    }
    const double *a = coef1 + ORC_C_A, *b = coef1 + ORC_C_B;
    double an1[3], an2[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++) an1[c] = a[c] * n[0] + a[3 + c] * n[1] + a[6 + c] * n[2];
    const double bn1 = b[0] * n[0] + b[1] * n[1] + b[2] * n[2];
    double bn2 = 0.0, nv = bn1;
    if (bc == 0) {
      if (orc_dgx_map_point(t1, side1, g1, eta, t2, side2, g2, eta2)) return -1;
      ns2 = t2 ? dgt_phys(pdeg2, nc2, eta2, p2, x2, y2, z2) : dg_phys(pdeg2, nc2, eta2, p2, x2, y2, z2);
      const double *a2 = coef2 + ORC_C_A, *b2 = coef2 + ORC_C_B;
      for (int c = 0; c < 3; c++) an2[c] = a2[c] * n[0] + a2[3 + c] * n[1] + a2[6 + c] * n[2];
      bn2 = b2[0] * n[0] + b2[1] * n[1] + b2[2] * n[2];
      nv = 0.5 * (bn1 + bn2);
    }
    nf = bc == 0 ? ns1 + ns2 : ns1;
    if (bc == 0) {
      double gg1[64], gg2[64];
      for (int i = 0; i < ns1; i++) gg1[i] = an1[0] * x1[i] + an1[1] * y1[i] + an1[2] * z1[i];
      for (int i = 0; i < ns2; i++) gg2[i] = an2[0] * x2[i] + an2[1] * y2[i] + an2[2] * z2[i];
      for (int j = 0; j < ns1; j++)
        for (int i = 0; i < ns1; i++)
          stiff[j * nf + i] += ((p1[j] * gg1[i] - p1[i] * gg1[j]) * 0.5 +
                                penalty * p1[j] * p1[i] / hf) * area;
      for (int j = 0; j < ns2; j++)
        for (int i = 0; i < ns1; i++)
          stiff[(ns1 + j) * nf + i] += ((-p2[j] * gg1[i] - p1[i] * gg2[j]) * 0.5 -
                                        penalty * p2[j] * p1[i] / hf) * area;
      for (int j = 0; j < ns1; j++)
        for (int i = 0; i < ns2; i++)
          stiff[j * nf + ns1 + i] += ((p1[j] * gg2[i] + p2[i] * gg1[j]) * 0.5 -
                                      penalty * p1[j] * p2[i] / hf) * area;
      for (int j = 0; j < ns2; j++)
        for (int i = 0; i < ns2; i++)
          stiff[(ns1 + j) * nf + ns1 + i] += ((-p2[j] * gg2[i] + p2[i] * gg2[j]) * 0.5 +
                                              penalty * p2[j] * p2[i] / hf) * area;
      if (nv > 0) {
        for (int j = 0; j < ns1; j++) {
          for (int i = 0; i < ns1; i++) stiff[j * nf + i] += area * bn1 * p1[i] * p1[j];
          for (int i = 0; i < ns2; i++) stiff[j * nf + ns1 + i] -= area * bn2 * p2[i] * p1[j];
        }
      } else if (nv < 0) {
        for (int j = 0; j < ns2; j++) {
          for (int i = 0; i < ns1; i++) stiff[(ns1 + j) * nf + i] += area * bn1 * p1[i] * p2[j];
          for (int i = 0; i < ns2; i++) stiff[(ns1 + j) * nf + ns1 + i] -= area * bn2 * p2[i] * p2[j];
        }
      }
    } else if (bc == 1) {
      double f = bval;
      if (src_model != ORC_SRC_CONST) exact_sol(src_model, xc, &f, NULL);
      for (int j = 0; j < ns1; j++)
        for (int i = 0; i < ns1; i++) {
          const double gi = an1[0] * x1[i] + an1[1] * y1[i] + an1[2] * z1[i];
          const double gj = an1[0] * x1[j] + an1[1] * y1[j] + an1[2] * z1[j];
          stiff[j * nf + i] += (gi * p1[j] - gj * p1[i] + penalty * p1[j] * p1[i] / hf) * area;
        }
      for (int i = 0; i < ns1; i++) {
        const double gi = an1[0] * x1[i] + an1[1] * y1[i] + an1[2] * z1[i];
        rhs[i] += (f * gi + penalty * f * p1[i] / hf) * area;
      }
      if (nv < 0) {
        for (int i = 0; i < ns1; i++) rhs[i] += -bn1 * f * p1[i] * area;
      } else {
        for (int j = 0; j < ns1; j++)
          for (int i = 0; i < ns1; i++) stiff[j * nf + i] += area * bn1 * p1[i] * p1[j];
      }
    } else {
      double g = bval;
      if (src_model == ORC_SRC_LINEAR) {
        g = n[0] - 2.0 * n[1] + 0.5 * n[2];
      } else if (src_model != ORC_SRC_CONST) {
        const double a1 = src_model == ORC_SRC_DIFF_IN_BOX ? 100.0 : 1.0;
        const double e = exp((-xc[0] * xc[0] - xc[1] * xc[1] - xc[2] * xc[2]) / a1);
        g = (-2.0 * xc[0] * n[0] - 2.0 * xc[1] * n[1] - 2.0 * xc[2] * n[2]) * e / a1;
      }
      if (nv > 0)
        for (int j = 0; j < ns1; j++)
          for (int i = 0; i < ns1; i++) stiff[j * nf + i] += area * bn1 * p1[i] * p1[j];
      for (int i = 0; i < ns1; i++) rhs[i] += g * p1[i] * area;
    }
  }
  return nf;
}

/* value of a tet DG field at a local point and the point's coordinates */
double orc_dgt_eval(int p, const double *nc, const double *dofs, const double *eta, double *x) {
  double phi[32], dx[32], dy[32], dz[32], J[9];
  const int ns = orc_dgt_shape(p, eta, phi, dx, dy, dz);
  dgt_geom(nc, eta, J, x);
  double u = 0.0;
  for (int i = 0; i < ns; i++) u += dofs[i] * phi[i];
  return u;
}
