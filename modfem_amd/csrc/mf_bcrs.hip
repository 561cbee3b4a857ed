// BCRS storage with 4x4 blocks (lsd_mkb/lad_bcrs/las_bcrs_intf.c) on the
// device, for the lar_* storage seam: residual, diagonal-block LU, and block
// Jacobi / block Gauss-Seidel sweeps.  Blocks are column-major (entry (row k,
// column l) at 16 s + k + 4 l), rows of blocks in CRS order with the diagonal
// in place.  Every kernel performs the reference's operations in the
// reference's order without fused multiply-adds, so results are bitwise those
// of the reference's single-thread code.  Gauss-Seidel sweeps follow the
// reference's sequential row order through level scheduling: a row's level is
// one more than the highest level among the rows it reads new values from.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "mf_device.h"
#include "mf_solver_dev.h"

namespace mf {

namespace {

// V = (b ? b : 0) - A x (lar_compute_residual_bcrs, las_bcrs_intf.c:2787-2930),
// one thread per unknown; the blocks of a row are taken two at a time with
// all loads issued before the (sequential) subtractions
__global__ void k_bcrs_residual(int nb, const int* rp, const int* ci, const double* val,
                                const double* x, const double* b, double* v, const int* stop) {
  // contiguous block rows per XCD: neighbouring rows share their x blocks in
  // that XCD's L2
  const long long i = (long long)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= 4LL * nb || (stop && *stop)) return;
  const int r = (int)(i >> 2), k = (int)(i & 3);
  double acc = 0.0;
  if (x) {
    const int s1 = rp[r + 1];
    int s = rp[r];
    for (; s + 2 <= s1; s += 2) {
      const int c0 = ci[s], c1 = ci[s + 1];
      double a0[4], a1[4], x0[4], x1[4];
      for (int l = 0; l < 4; l++) {
        a0[l] = val[16 * (size_t)s + k + 4 * l];
        a1[l] = val[16 * (size_t)(s + 1) + k + 4 * l];
        x0[l] = x[4 * (size_t)c0 + l];
        x1[l] = x[4 * (size_t)c1 + l];
      }
      for (int l = 0; l < 4; l++) acc = __dsub_rn(acc, __dmul_rn(a0[l], x0[l]));
      for (int l = 0; l < 4; l++) acc = __dsub_rn(acc, __dmul_rn(a1[l], x1[l]));
    }
    if (s < s1) {
      const int c = ci[s];
      for (int l = 0; l < 4; l++)
        acc = __dsub_rn(acc, __dmul_rn(val[16 * (size_t)s + k + 4 * l], x[4 * (size_t)c + l]));
    }
  }
  if (b) acc = __dadd_rn(acc, b[i]);
  v[i] = acc;
}

// LU of every diagonal block with partial pivoting, the unblocked algorithm of
// LAPACK dgetf2 (lar_fill_preconditioner_bcrs calls dgetrf, :1119-1148)
__global__ void k_bcrs_lu4(int nb, const int* diag, const double* val, double* lu, int* piv,
                           int* bad) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nb) return;
  double a[16];
  for (int t = 0; t < 16; t++) a[t] = val[16 * (size_t)diag[r] + t];
  int ip[4];
  for (int j = 0; j < 4; j++) {
    int jp = j;
    double amax = fabs(a[j + 4 * j]);
    for (int i = j + 1; i < 4; i++)
      if (fabs(a[i + 4 * j]) > amax) { amax = fabs(a[i + 4 * j]); jp = i; }
    ip[j] = jp;
    if (a[jp + 4 * j] != 0.0) {
      if (jp != j)
        for (int c = 0; c < 4; c++) {
          const double t = a[j + 4 * c];
          a[j + 4 * c] = a[jp + 4 * c];
          a[jp + 4 * c] = t;
        }
      const double rec = __ddiv_rn(1.0, a[j + 4 * j]);
      for (int i = j + 1; i < 4; i++) a[i + 4 * j] = __dmul_rn(a[i + 4 * j], rec);
    } else {
      bad[0] = 1;
    }
    for (int c = j + 1; c < 4; c++) {
      const double t = -a[j + 4 * c];
      for (int i = j + 1; i < 4; i++) a[i + 4 * c] = __dadd_rn(a[i + 4 * c], __dmul_rn(a[i + 4 * j], t));
    }
  }
  for (int t = 0; t < 16; t++) lu[16 * (size_t)r + t] = a[t];
  for (int t = 0; t < 4; t++) piv[4 * (size_t)r + t] = ip[t];
}

// solve with the factors as LAPACK dgetrs ('N') does
__device__ __forceinline__ void lu4_solve(const double* lu, const int* ip, double* b) {
  for (int j = 0; j < 4; j++)
    if (ip[j] != j) {
      const double t = b[j];
      b[j] = b[ip[j]];
      b[ip[j]] = t;
    }
  for (int k = 0; k < 4; k++)
    if (b[k] != 0.0)
      for (int i = k + 1; i < 4; i++) b[i] = __dsub_rn(b[i], __dmul_rn(b[k], lu[i + 4 * k]));
  for (int k = 3; k >= 0; k--)
    if (b[k] != 0.0) {
      b[k] = __ddiv_rn(b[k], lu[k + 4 * k]);
      for (int i = 0; i < k; i++) b[i] = __dsub_rn(b[i], __dmul_rn(b[k], lu[i + 4 * k]));
    }
}

// one row of a sweep (lar_perform_BJ_or_GS_iterations_bcrs, :2938-3300):
// vloc = -(blocks the sweep has passed, read from `passed`) - (the other
// off-diagonal blocks, read from v) (+ b), then the diagonal solve
__device__ __forceinline__ void bcrs_row(int r, int dir, const int* rp, const int* ci,
                                         const int* diag, const double* val, const double* lu,
                                         const int* piv, const double* b, const double* v,
                                         const double* passed, double* out) {
  double vloc[4] = {0.0, 0.0, 0.0, 0.0};
  const int p0 = dir == 0 ? rp[r] : diag[r] + 1, p1 = dir == 0 ? diag[r] : rp[r + 1];
  const int o0 = dir == 0 ? diag[r] + 1 : rp[r], o1 = dir == 0 ? rp[r + 1] : diag[r];
  for (int s = p0; s < p1; s++) {
    const int c = ci[s];
    for (int l = 0; l < 4; l++)
      for (int k = 0; k < 4; k++)
        vloc[k] = __dsub_rn(vloc[k], __dmul_rn(val[16 * (size_t)s + k + 4 * l], passed[4 * (size_t)c + l]));
  }
  for (int s = o0; s < o1; s++) {
    const int c = ci[s];
    for (int k = 0; k < 4; k++)
      for (int l = 0; l < 4; l++)
        vloc[k] = __dsub_rn(vloc[k], __dmul_rn(val[16 * (size_t)s + k + 4 * l], v[4 * (size_t)c + l]));
  }
  if (b)
    for (int k = 0; k < 4; k++) vloc[k] = __dadd_rn(vloc[k], b[4 * (size_t)r + k]);
  lu4_solve(lu + 16 * (size_t)r, piv + 4 * (size_t)r, vloc);
  for (int k = 0; k < 4; k++) out[4 * (size_t)r + k] = vloc[k];
}

// subtract the products of blocks [s0, s1) of component k in order (two
// blocks per step, loads first)
__device__ __forceinline__ double chain_sub(double acc, int s0, int s1, int k, const int* ci,
                                            const double* val, const double* y) {
  int s = s0;
  for (; s + 2 <= s1; s += 2) {
    const int c0 = ci[s], c1 = ci[s + 1];
    double a0[4], a1[4], y0[4], y1[4];
    for (int l = 0; l < 4; l++) {
      a0[l] = val[16 * (size_t)s + k + 4 * l];
      a1[l] = val[16 * (size_t)(s + 1) + k + 4 * l];
      y0[l] = y[4 * (size_t)c0 + l];
      y1[l] = y[4 * (size_t)c1 + l];
    }
    for (int l = 0; l < 4; l++) acc = __dsub_rn(acc, __dmul_rn(a0[l], y0[l]));
    for (int l = 0; l < 4; l++) acc = __dsub_rn(acc, __dmul_rn(a1[l], y1[l]));
  }
  if (s < s1) {
    const int c = ci[s];
    for (int l = 0; l < 4; l++)
      acc = __dsub_rn(acc, __dmul_rn(val[16 * (size_t)s + k + 4 * l], y[4 * (size_t)c + l]));
  }
  return acc;
}

// block Jacobi sweep: every row reads the old v; one thread per unknown, the
// four threads of a row exchange their sums and each solves with the
// diagonal block's factors, keeping its own component.  With a row list
// (rows of one colour class, none sharing a block) the same kernel sweeps
// the class in place (out = v).
__global__ void k_bcrs_bj(int nrows, const int* rows, int dir, const int* rp, const int* ci,
                          const int* diag, const double* val, const double* lu, const int* piv,
                          const double* b, const double* v, double* out, const int* stop) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 4LL * nrows || (stop && *stop)) return;
  const int r = rows ? rows[t >> 2] : (int)(t >> 2), k = (int)(t & 3);
  const long long i = 4LL * r + k;
  const int d = diag[r];
  double acc = 0.0;
  if (dir == 0) {
    acc = chain_sub(acc, rp[r], d, k, ci, val, v);
    acc = chain_sub(acc, d + 1, rp[r + 1], k, ci, val, v);
  } else {
    acc = chain_sub(acc, d + 1, rp[r + 1], k, ci, val, v);
    acc = chain_sub(acc, rp[r], d, k, ci, val, v);
  }
  if (b) acc = __dadd_rn(acc, b[i]);
  const int base = (int)(threadIdx.x & ~3u);
  double y[4];
  for (int q = 0; q < 4; q++) y[q] = __shfl(acc, base + q, 64);
  lu4_solve(lu + 16 * (size_t)r, piv + 4 * (size_t)r, y);
  out[i] = y[k];
}

// block Gauss-Seidel, the rows of one level: passed blocks read the sweep's
// output (rows of earlier levels are final there)
__global__ void k_bcrs_gs_level(int cnt, const int* rows, int dir, const int* rp, const int* ci,
                                const int* diag, const double* val, const double* lu,
                                const int* piv, const double* b, const double* v, double* out,
                                const int* stop) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt || (stop && *stop)) return;
  bcrs_row(rows[t], dir, rp, ci, diag, val, lu, piv, b, v, out, out);
}

// inverse of a 4x4 block from its lu4 factors, the unblocked algorithm of
// LAPACK dgetri: inv(U) column by column (dtrti2 with dtrmv, which skips zero
// entries), then inv(A)(:, j) -= inv(A)(:, j+1:) L(j+1:, j) for j = 3..0
// (dgemv), then the column interchanges in reverse order
__device__ __forceinline__ void inv4(double* a, const int* ip) {
  for (int j = 0; j < 4; j++) {
    a[j + 4 * j] = __ddiv_rn(1.0, a[j + 4 * j]);
    const double ajj = -a[j + 4 * j];
    for (int jj = 0; jj < j; jj++)
      if (a[jj + 4 * j] != 0.0) {
        const double t = a[jj + 4 * j];
        for (int ii = 0; ii < jj; ii++) a[ii + 4 * j] = __dadd_rn(a[ii + 4 * j], __dmul_rn(t, a[ii + 4 * jj]));
        a[jj + 4 * j] = __dmul_rn(a[jj + 4 * j], a[jj + 4 * jj]);
      }
    for (int ii = 0; ii < j; ii++) a[ii + 4 * j] = __dmul_rn(ajj, a[ii + 4 * j]);
  }
  double work[4];
  for (int j = 3; j >= 0; j--) {
    for (int i = j + 1; i < 4; i++) {
      work[i] = a[i + 4 * j];
      a[i + 4 * j] = 0.0;
    }
    for (int jj = j + 1; jj < 4; jj++) {
      const double t = -work[jj];
      for (int i = 0; i < 4; i++) a[i + 4 * j] = __dadd_rn(a[i + 4 * j], __dmul_rn(t, a[i + 4 * jj]));
    }
  }
  for (int j = 2; j >= 0; j--) {
    const int jp = ip[j];
    if (jp != j)
      for (int i = 0; i < 4; i++) {
        const double t = a[i + 4 * j];
        a[i + 4 * j] = a[i + 4 * jp];
        a[i + 4 * jp] = t;
      }
  }
}

// LU with partial pivoting of a 4x4 block in registers (dgetf2); false for a
// zero pivot
__device__ __forceinline__ bool lu4(double* a, int* ip) {
  bool ok = true;
  for (int j = 0; j < 4; j++) {
    int jp = j;
    double amax = fabs(a[j + 4 * j]);
    for (int i = j + 1; i < 4; i++)
      if (fabs(a[i + 4 * j]) > amax) { amax = fabs(a[i + 4 * j]); jp = i; }
    ip[j] = jp;
    if (a[jp + 4 * j] != 0.0) {
      if (jp != j)
        for (int c = 0; c < 4; c++) {
          const double t = a[j + 4 * c];
          a[j + 4 * c] = a[jp + 4 * c];
          a[jp + 4 * c] = t;
        }
      const double rec = __ddiv_rn(1.0, a[j + 4 * j]);
      for (int i = j + 1; i < 4; i++) a[i + 4 * j] = __dmul_rn(a[i + 4 * j], rec);
    } else {
      ok = false;
    }
    for (int c = j + 1; c < 4; c++) {
      const double t = -a[j + 4 * c];
      for (int i = j + 1; i < 4; i++) a[i + 4 * c] = __dadd_rn(a[i + 4 * c], __dmul_rn(a[i + 4 * j], t));
    }
  }
  return ok;
}

// block ILU(0), the rows of one level (a row's level exceeds those of the
// rows in its lower part).  Row k in the order of the reference's
// right-looking variant (las_bcrs_intf.c:2181-2640), which applies to every
// block of row k the pivots i < k in increasing order: for each lower block
// a_ki, a_ki = a_ki inv(a_ii) (sums from 0), then a_kj -= a_ki a_ij for the
// blocks of row i's upper part present in row k (dgemm order: column, inner
// index, row); finally the diagonal block is replaced by its inverse
__global__ void k_bcrs_ilu_level(int cnt, const int* rows, const int* rp, const int* ci,
                                 const int* diag, double* ilu, int* bad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt) return;
  const int k = rows[t];
  const int dk = diag[k], end = rp[k + 1];
  for (int s = rp[k]; s < dk; s++) {
    const int i = ci[s];
    double* aki = ilu + 16 * (size_t)s;
    const double* w = ilu + 16 * (size_t)diag[i];
    double a[16], p[16];
    for (int q = 0; q < 16; q++) a[q] = aki[q];
    for (int r = 0; r < 4; r++)
      for (int c = 0; c < 4; c++) {
        double d = 0.0;
        for (int l = 0; l < 4; l++) d = __dadd_rn(d, __dmul_rn(a[r + 4 * l], w[l + 4 * c]));
        p[r + 4 * c] = d;
      }
    for (int q = 0; q < 16; q++) aki[q] = p[q];
    int skj = s + 1;
    for (int sij = diag[i] + 1; sij < rp[i + 1]; sij++) {
      const int j = ci[sij];
      while (skj < end && ci[skj] < j) skj++;
      if (skj >= end) break;
      if (ci[skj] != j) continue;
      const double* aij = ilu + 16 * (size_t)sij;
      double* akj = ilu + 16 * (size_t)skj;
      double c4[16];
      for (int q = 0; q < 16; q++) c4[q] = akj[q];
      for (int c = 0; c < 4; c++)
        for (int l = 0; l < 4; l++) {
          const double tt = -aij[l + 4 * c];
          for (int r = 0; r < 4; r++) c4[r + 4 * c] = __dadd_rn(c4[r + 4 * c], __dmul_rn(tt, p[r + 4 * l]));
        }
      for (int q = 0; q < 16; q++) akj[q] = c4[q];
    }
  }
  double d[16];
  int ip[4];
  double* akk = ilu + 16 * (size_t)dk;
  for (int q = 0; q < 16; q++) d[q] = akk[q];
  if (!lu4(d, ip)) {
    bad[0] = 1;
    return;
  }
  inv4(d, ip);
  for (int q = 0; q < 16; q++) akk[q] = d[q];
}

// forward reduction of the ILU solve (lar_perform_rhsub_bcrs, variant 0,
// las_bcrs_intf.c:3349-3430), the rows of one level: v_i = b_i - L_i. v
__global__ void k_bcrs_fwd_level(int cnt, const int* rows, const int* rp, const int* ci,
                                 const int* diag, const double* ilu, const double* b, double* v,
                                 const int* stop) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt || (stop && *stop)) return;
  const int i = rows[t];
  double x[4];
  for (int r = 0; r < 4; r++) x[r] = b[4 * (size_t)i + r];
  for (int s = rp[i]; s < diag[i]; s++) {
    const double* a = ilu + 16 * (size_t)s;
    const double* y = v + 4 * (size_t)ci[s];
    for (int c = 0; c < 4; c++)
      for (int r = 0; r < 4; r++) x[r] = __dsub_rn(x[r], __dmul_rn(a[r + 4 * c], y[c]));
  }
  for (int r = 0; r < 4; r++) v[4 * (size_t)i + r] = x[r];
}

// back substitution (:3432-3510), the rows of one level:
// v_i = inv(a_ii) (v_i - U_i. v)
__global__ void k_bcrs_bwd_level(int cnt, const int* rows, const int* rp, const int* ci,
                                 const int* diag, const double* ilu, double* v, const int* stop) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cnt || (stop && *stop)) return;
  const int i = rows[t];
  double x[4];
  for (int r = 0; r < 4; r++) x[r] = v[4 * (size_t)i + r];
  for (int s = diag[i] + 1; s < rp[i + 1]; s++) {
    const double* a = ilu + 16 * (size_t)s;
    const double* y = v + 4 * (size_t)ci[s];
    for (int c = 0; c < 4; c++)
      for (int r = 0; r < 4; r++) x[r] = __dsub_rn(x[r], __dmul_rn(a[r + 4 * c], y[c]));
  }
  const double* d = ilu + 16 * (size_t)diag[i];
  double o[4] = {0.0, 0.0, 0.0, 0.0};
  for (int c = 0; c < 4; c++)
    for (int r = 0; r < 4; r++) o[r] = __dadd_rn(o[r], __dmul_rn(d[r + 4 * c], x[c]));
  for (int r = 0; r < 4; r++) v[4 * (size_t)i + r] = o[r];
}

// One wave per row for the sweeps of the level schedules, when no row holds
// more than kWaveBlocks blocks: the wave forms every product a_kl y_l of the
// row in parallel (coalesced block loads) into LDS, then lane k subtracts the
// products of component k in the sequential order (blocks ascending, l
// ascending) -- the same roundings in the same order as one thread per row.
//   mode 0: Gauss-Seidel row in direction dir (passed side from `passed`,
//           other side from v, + b, then the diagonal block's LU solve) -> out
//   mode 1: ILU forward row, y = b - L y (lower blocks from v, in place)
//   mode 2: ILU backward row, y = inv(a_ii) (y - U y) (upper blocks from v)
constexpr int kWaveBlocks = 128;

__global__ void __launch_bounds__(64) k_bcrs_rows_wave(
    int cnt, const int* rows, int mode, int dir, const int* rp, const int* ci, const int* diag,
    const double* val, const double* lu, const int* piv, const double* b, const double* v,
    const double* passed, double* out, const int* stop) {
  __shared__ double prod[kWaveBlocks * 16];
  __shared__ double acc[4];
  const int t = blockIdx.x, lane = threadIdx.x;
  if (t >= cnt || (stop && *stop)) return;  // uniform over the workgroup
  const int r = rows[t];
  int p0, p1, o0 = 0, o1 = 0;
  if (mode == 0) {
    p0 = dir == 0 ? rp[r] : diag[r] + 1;
    p1 = dir == 0 ? diag[r] : rp[r + 1];
    o0 = dir == 0 ? diag[r] + 1 : rp[r];
    o1 = dir == 0 ? rp[r + 1] : diag[r];
  } else if (mode == 1) {
    p0 = rp[r];
    p1 = diag[r];
  } else {
    p0 = diag[r] + 1;
    p1 = rp[r + 1];
  }
  const int np = p1 - p0, nbk = np + (o1 - o0);
  for (int idx = lane; idx < 16 * nbk; idx += 64) {
    const int q = idx >> 4, e = idx & 15;
    const int s = q < np ? p0 + q : o0 + (q - np);
    const double* src = q < np ? passed : v;
    prod[idx] = __dmul_rn(val[16 * (size_t)s + e], src[4 * (size_t)ci[s] + (e >> 2)]);
  }
  __syncthreads();
  if (lane < 4) {
    const int k = lane;
    double x = mode == 2 ? v[4 * (size_t)r + k] : (mode == 1 ? b[4 * (size_t)r + k] : 0.0);
    for (int q = 0; q < nbk; q++)
      for (int l = 0; l < 4; l++) x = __dsub_rn(x, prod[16 * q + k + 4 * l]);
    if (mode == 0 && b) x = __dadd_rn(x, b[4 * (size_t)r + k]);
    acc[k] = x;
  }
  __syncthreads();
  if (lane == 0) {
    double x[4] = {acc[0], acc[1], acc[2], acc[3]};
    if (mode == 0) {
      lu4_solve(lu + 16 * (size_t)r, piv + 4 * (size_t)r, x);
      for (int k = 0; k < 4; k++) out[4 * (size_t)r + k] = x[k];
    } else if (mode == 1) {
      for (int k = 0; k < 4; k++) out[4 * (size_t)r + k] = x[k];
    } else {
      const double* d = lu + 16 * (size_t)diag[r];
      double o[4] = {0.0, 0.0, 0.0, 0.0};
      for (int c = 0; c < 4; c++)
        for (int k = 0; k < 4; k++) o[k] = __dadd_rn(o[k], __dmul_rn(d[k + 4 * c], x[c]));
      for (int k = 0; k < 4; k++) out[4 * (size_t)r + k] = o[k];
    }
  }
}

// Block ILU(0) of the rows of one level, one wave per row: the row's blocks
// in LDS; for each lower block a_ki in order, lanes 0..15 form a_ki inv(a_ii)
// (one entry each, sums from 0), then the wave updates the blocks a_kj of the
// row present in row i's upper part, one entry per lane (dgemm order: inner
// index ascending); finally lane 0 inverts the diagonal block.  The same
// operations in the same order per entry as k_bcrs_ilu_level.
__global__ void __launch_bounds__(64) k_bcrs_ilu_wave(int cnt, const int* rows, const int* rp,
                                                       const int* ci, const int* diag, double* ilu,
                                                       int* bad) {
  __shared__ double A[kWaveBlocks * 16];
  __shared__ int C[kWaveBlocks];
  __shared__ int Pos[kWaveBlocks];
  const int t = blockIdx.x, lane = threadIdx.x;
  if (t >= cnt) return;
  const int k = rows[t];
  const int r0 = rp[k], nbk = rp[k + 1] - r0, dk = diag[k] - r0;
  double* row = ilu + 16 * (size_t)r0;
  for (int idx = lane; idx < 16 * nbk; idx += 64) A[idx] = row[idx];
  for (int q = lane; q < nbk; q += 64) C[q] = ci[r0 + q];
  __syncthreads();
  for (int q = 0; q < dk; q++) {
    const int i = C[q];
    const double* w = ilu + 16 * (size_t)diag[i];
    double pv = 0.0;
    if (lane < 16) {
      const int rr = lane & 3, cc = lane >> 2;
      for (int l = 0; l < 4; l++) pv = __dadd_rn(pv, __dmul_rn(A[16 * q + rr + 4 * l], w[l + 4 * cc]));
    }
    __syncthreads();
    if (lane < 16) A[16 * q + lane] = pv;
    __syncthreads();
    const int u0 = diag[i] + 1, nu = rp[i + 1] - u0;
    // position in row k of every block of row i's upper part (-1: absent),
    // by bisection over the sorted columns after q
    for (int jj = lane; jj < nu; jj += 64) {
      const int j = ci[u0 + jj];
      int lo = q + 1, hi = nbk;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (C[mid] < j) lo = mid + 1;
        else hi = mid;
      }
      Pos[jj] = (lo < nbk && C[lo] == j) ? lo : -1;
    }
    __syncthreads();
    for (int idx = lane; idx < 16 * nu; idx += 64) {
      const int jj = idx >> 4, e = idx & 15, rr = e & 3, cc = e >> 2;
      const int pos = Pos[jj];
      if (pos < 0) continue;
      const double* aij = ilu + 16 * (size_t)(u0 + jj);
      double c4 = A[16 * pos + e];
      for (int l = 0; l < 4; l++)
        c4 = __dadd_rn(c4, __dmul_rn(-aij[l + 4 * cc], A[16 * q + rr + 4 * l]));
      A[16 * pos + e] = c4;
    }
    __syncthreads();
  }
  if (lane == 0) {
    double d[16];
    int ip[4];
    for (int z = 0; z < 16; z++) d[z] = A[16 * dk + z];
    if (!lu4(d, ip)) bad[0] = 1;
    inv4(d, ip);
    for (int z = 0; z < 16; z++) A[16 * dk + z] = d[z];
  }
  __syncthreads();
  for (int idx = lane; idx < 16 * nbk; idx += 64) row[idx] = A[idx];
}

// vector steps of the preconditioned residual (lsr_mkb_core_compreres,
// lss_mkb_core.c:1915-1990): y = s_j x (s null: copy), y = a - b (negated
// when neg), y = -y, y += z
__global__ void k_bcrs_vscale(long long n, const double* x, const double* s, int j, double* y,
                              const int* stop) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (stop && *stop)) return;
  y[i] = s ? __dmul_rn(s[j], x[i]) : x[i];
}
__global__ void k_bcrs_vsub(long long n, const double* a, const double* b, double* y, int neg,
                            const int* stop) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (stop && *stop)) return;
  const double d = __dsub_rn(a[i], b[i]);
  y[i] = neg ? -d : d;
}
__global__ void k_bcrs_vneg(long long n, double* y, const int* stop) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (stop && *stop)) return;
  y[i] = -y[i];
}
__global__ void k_bcrs_vadd(long long n, const double* z, double* y, const int* stop) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (stop && *stop)) return;
  y[i] = __dadd_rn(y[i], z[i]);
}

}  // namespace

int bcrs_levels(int nb, const std::vector<int>& rp, const std::vector<int>& ci,
                const std::vector<int>& diag, int dir, std::vector<int>& lev_ptr,
                std::vector<int>& rows) {
  std::vector<int> lev(nb, 0);
  int nlev = 0;
  for (int q = 0; q < nb; q++) {
    const int r = dir == 0 ? q : nb - 1 - q;
    const int p0 = dir == 0 ? rp[r] : diag[r] + 1, p1 = dir == 0 ? diag[r] : rp[r + 1];
    int l = 0;
    for (int s = p0; s < p1; s++) l = std::max(l, lev[ci[s]] + 1);
    lev[r] = l;
    nlev = std::max(nlev, l + 1);
  }
  lev_ptr.assign(nlev + 1, 0);
  for (int r = 0; r < nb; r++) lev_ptr[lev[r] + 1]++;
  for (int l = 0; l < nlev; l++) lev_ptr[l + 1] += lev_ptr[l];
  std::vector<int> pos(lev_ptr.begin(), lev_ptr.end() - 1);
  rows.assign(nb, 0);
  for (int r = 0; r < nb; r++) rows[pos[lev[r]]++] = r;
  return nlev;
}

int bcrs_colors(int nb, const std::vector<int>& rp, const std::vector<int>& ci,
                std::vector<int>& col_ptr, std::vector<int>& rows) {
  std::vector<int> color(nb, -1), mark;
  int ncol = 0;
  for (int r = 0; r < nb; r++) {
    mark.assign(ncol + 1, 0);
    for (int s = rp[r]; s < rp[r + 1]; s++)
      if (color[ci[s]] >= 0) mark[color[ci[s]]] = 1;
    int c = 0;
    while (c < ncol && mark[c]) c++;
    color[r] = c;
    ncol = std::max(ncol, c + 1);
  }
  col_ptr.assign(ncol + 1, 0);
  for (int r = 0; r < nb; r++) col_ptr[color[r] + 1]++;
  for (int c = 0; c < ncol; c++) col_ptr[c + 1] += col_ptr[c];
  std::vector<int> pos(col_ptr.begin(), col_ptr.end() - 1);
  rows.assign(nb, 0);
  for (int r = 0; r < nb; r++) rows[pos[color[r]]++] = r;
  return ncol;
}

int bcrs_residual(const BcrsDev& M, const double* x, const double* b, double* v,
                  const int* stop) {
  const long long n = 4LL * M.nb;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_bcrs_residual, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream(),
                     M.nb, M.d_rp, M.d_ci, M.d_val, x, b, v, stop);
  MF_HIP(hipGetLastError());
  return 0;
}

int bcrs_fill_lu(BcrsDev& M) {
  if (M.nb == 0) return 0;
  int* d_bad = nullptr;
  MF_HIP(hipMalloc(&d_bad, sizeof(int)));
  MF_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), stream()));
  hipLaunchKernelGGL(k_bcrs_lu4, dim3((M.nb + 255) / 256), dim3(256), 0, stream(), M.nb, M.d_diag,
                     M.d_val, M.d_lu, M.d_piv, d_bad);
  int bad = 0;
  MF_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  (void)hipFree(d_bad);
  if (bad) {
    set_error("lar_fill_preconditioner: a diagonal block is singular");
    return -1;
  }
  return 0;
}

int bcrs_sweeps(BcrsDev& M, int gs, int nr_prec, const double* b, double* v, double* tmp,
                const int* stop, bool sync) {
  const int loops = nr_prec % 2 == 0 ? nr_prec / 2 : nr_prec / 2 + 1;
  const int dirs = nr_prec % 2 == 0 ? 2 : 1;
  const long long n = 4LL * M.nb;
  if (M.nb == 0) return 0;
  const unsigned gv = (unsigned)((n + 255) / 256);
  if (gs && M.gs_order == 1) {  // colour classes, in place
    const int ncol = (int)M.col_ptr.size() - 1;
    for (int it = 0; it < loops; it++)
      for (int dir = 0; dir < dirs; dir++)
        for (int q = 0; q < ncol; q++) {
          const int c = dir == 0 ? q : ncol - 1 - q;
          const int cnt = M.col_ptr[c + 1] - M.col_ptr[c];
          hipLaunchKernelGGL(k_bcrs_bj, dim3((4 * cnt + 255) / 256), dim3(256), 0, stream(), cnt,
                             M.d_col_rows + M.col_ptr[c], 0, M.d_rp, M.d_ci, M.d_diag, M.d_val,
                             M.d_lu, M.d_piv, b, v, v, stop);
        }
    MF_HIP(hipGetLastError());
    if (sync) MF_HIP(hipStreamSynchronize(stream()));
    return 0;
  }
  for (int it = 0; it < loops; it++)
    for (int dir = 0; dir < dirs; dir++) {
      if (!gs) {
        hipLaunchKernelGGL(k_bcrs_bj, dim3(gv), dim3(256), 0, stream(), M.nb, nullptr, dir,
                           M.d_rp, M.d_ci, M.d_diag, M.d_val, M.d_lu, M.d_piv, b, v, tmp, stop);
      } else {
        const std::vector<int>& lp = M.lev_ptr[dir];
        for (size_t l = 0; l + 1 < lp.size(); l++) {
          const int cnt = lp[l + 1] - lp[l];
          if (M.max_row <= kWaveBlocks)
            hipLaunchKernelGGL(k_bcrs_rows_wave, dim3(cnt), dim3(64), 0, stream(), cnt,
                               M.d_lev_rows[dir] + lp[l], 0, dir, M.d_rp, M.d_ci, M.d_diag,
                               M.d_val, M.d_lu, M.d_piv, b, v, tmp, tmp, stop);
          else
            hipLaunchKernelGGL(k_bcrs_gs_level, dim3((cnt + 63) / 64), dim3(64), 0, stream(), cnt,
                               M.d_lev_rows[dir] + lp[l], dir, M.d_rp, M.d_ci, M.d_diag, M.d_val,
                               M.d_lu, M.d_piv, b, v, tmp, stop);
        }
      }
      hipLaunchKernelGGL(k_bcrs_vscale, dim3(gv), dim3(256), 0, stream(), n, tmp, nullptr, 0, v,
                         stop);
      MF_HIP(hipGetLastError());
    }
  if (sync) MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

// ILU(k) pattern of lar_allocate_preconditioner_bcrs (las_bcrs_intf.c:751-
// 1016): row i starts from its stored blocks; each of k rings appends, for
// every column the previous ring added (in list order), the columns of that
// column's row not yet listed; the list is then sorted.  A row holds at most
// cap columns, as the reference's fixed-length list does (it stops the
// program when a row needs more; here the row is reported)
int bcrs_iluk_pattern(int nb, const std::vector<int>& rp, const std::vector<int>& ci, int k,
                      int cap, std::vector<int>& prp, std::vector<int>& pci, int* bad_row) {
  std::vector<int> seen(nb, -1), list;
  prp.assign(nb + 1, 0);
  pci.clear();
  for (int i = 0; i < nb; i++) {
    list.clear();
    auto put = [&](int c) {
      if (seen[c] == i) return true;
      if ((int)list.size() == cap) return false;
      seen[c] = i;
      list.push_back(c);
      return true;
    };
    bool ok = true;
    for (int s = rp[i]; s < rp[i + 1] && ok; s++) ok = put(ci[s]);
    size_t first = 0, last = list.size();
    for (int ring = 0; ring < k && ok; ring++) {
      for (size_t q = first; q < last && ok; q++) {
        const int r = list[q];
        for (int s = rp[r]; s < rp[r + 1] && ok; s++) ok = put(ci[s]);
      }
      first = last;
      last = list.size();
    }
    if (!ok) {
      if (bad_row) *bad_row = i;
      return -1;
    }
    std::sort(list.begin(), list.end());
    pci.insert(pci.end(), list.begin(), list.end());
    prp[i + 1] = (int)pci.size();
  }
  return 0;
}

// ILU values on the grown pattern: the matrix block, or zero for fill
__global__ void k_bcrs_ilu_gather(long long n, const int* src, const double* val, double* ilu) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int s = src[q >> 4];
  ilu[q] = s >= 0 ? val[16 * (long long)s + (q & 15)] : 0.0;
}

// the pattern the factors live on: the matrix's for ILU(0), the grown one
// for ILU(k > 0)
struct IluPattern {
  const int *rp, *ci, *diag;
  int max_row;
  const std::vector<int>* lev_ptr;
  int* const* lev_rows;
};
static IluPattern ilu_pattern(const BcrsDev& M) {
  if (M.ilu_k > 0)
    return {M.d_ilu_rp, M.d_ilu_ci, M.d_ilu_diag, M.ilu_max_row, M.ilu_lev_ptr, M.d_ilu_lev_rows};
  return {M.d_rp, M.d_ci, M.d_diag, M.max_row, M.lev_ptr, M.d_lev_rows};
}

int bcrs_ilu_factor(BcrsDev& M) {
  if (M.nb == 0) return 0;
  if (M.ilu_k > 0) {
    const long long n = M.ilu_nblk * 16;
    hipLaunchKernelGGL(k_bcrs_ilu_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       stream(), n, (const int*)M.d_ilu_src, (const double*)M.d_val, M.d_ilu);
  } else {
    MF_HIP(hipMemcpyAsync(M.d_ilu, M.d_val, (size_t)M.nblk * 16 * sizeof(double),
                          hipMemcpyDeviceToDevice, stream()));
  }
  const IluPattern Q = ilu_pattern(M);
  int* d_bad = nullptr;
  MF_HIP(hipMalloc(&d_bad, sizeof(int)));
  MF_HIP(hipMemsetAsync(d_bad, 0, sizeof(int), stream()));
  const std::vector<int>& lp = Q.lev_ptr[0];
  for (size_t l = 0; l + 1 < lp.size(); l++) {
    const int cnt = lp[l + 1] - lp[l];
    if (Q.max_row <= kWaveBlocks)
      hipLaunchKernelGGL(k_bcrs_ilu_wave, dim3(cnt), dim3(64), 0, stream(), cnt,
                         Q.lev_rows[0] + lp[l], Q.rp, Q.ci, Q.diag, M.d_ilu, d_bad);
    else
      hipLaunchKernelGGL(k_bcrs_ilu_level, dim3((cnt + 63) / 64), dim3(64), 0, stream(), cnt,
                         Q.lev_rows[0] + lp[l], Q.rp, Q.ci, Q.diag, M.d_ilu, d_bad);
  }
  MF_HIP(hipGetLastError());
  int bad = 0;
  MF_HIP(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  (void)hipFree(d_bad);
  if (bad) {
    set_error("lar_fill_preconditioner: a diagonal block of the ILU(0) factors is singular");
    return -1;
  }
  return 0;
}

int bcrs_ilu_solve(const BcrsDev& M, const double* b, double* v, const int* stop) {
  if (M.nb == 0) return 0;
  const IluPattern Q = ilu_pattern(M);
  const std::vector<int>& lf = Q.lev_ptr[0];
  for (size_t l = 0; l + 1 < lf.size(); l++) {
    const int cnt = lf[l + 1] - lf[l];
    if (Q.max_row <= kWaveBlocks)
      hipLaunchKernelGGL(k_bcrs_rows_wave, dim3(cnt), dim3(64), 0, stream(), cnt,
                         Q.lev_rows[0] + lf[l], 1, 0, Q.rp, Q.ci, Q.diag, M.d_ilu,
                         nullptr, nullptr, b, v, v, v, stop);
    else
      hipLaunchKernelGGL(k_bcrs_fwd_level, dim3((cnt + 63) / 64), dim3(64), 0, stream(), cnt,
                         Q.lev_rows[0] + lf[l], Q.rp, Q.ci, Q.diag, M.d_ilu, b, v, stop);
  }
  const std::vector<int>& lb = Q.lev_ptr[1];
  for (size_t l = 0; l + 1 < lb.size(); l++) {
    const int cnt = lb[l + 1] - lb[l];
    if (Q.max_row <= kWaveBlocks)
      hipLaunchKernelGGL(k_bcrs_rows_wave, dim3(cnt), dim3(64), 0, stream(), cnt,
                         Q.lev_rows[1] + lb[l], 2, 0, Q.rp, Q.ci, Q.diag, M.d_ilu,
                         M.d_ilu, nullptr, nullptr, v, v, v, stop);
    else
      hipLaunchKernelGGL(k_bcrs_bwd_level, dim3((cnt + 63) / 64), dim3(64), 0, stream(), cnt,
                         Q.lev_rows[1] + lb[l], Q.rp, Q.ci, Q.diag, M.d_ilu, v, stop);
  }
  MF_HIP(hipGetLastError());
  return 0;
}

// The preconditioned operator of GMRES on a BCRS matrix, as
// lsr_mkb_core_compreres (lss_mkb_core.c:1915-1990) forms it from the
// storage operations: none, V = b - A x; BLOCK_ILU, V = (LU)^-1 (b - A x);
// block Jacobi / Gauss-Seidel (nr_prec sweeps) and MULTI_ILU through
// smoothing, V = smooth(x; b) - x, with smooth one ILU correction
// x + (LU)^-1 (b - A x) for MULTI_ILU (lsr_mkb_core_smooth, :1621-1645).
// The Arnoldi product takes b = 0 and negates the result (dscal by -1):
// w = -V(X), X = s_j V_j.
struct BcrsOp : ExtOp {
  BcrsDev* B = nullptr;
  int kind = 0, nr_prec = 1;
  double *t1 = nullptr, *t2 = nullptr, *t3 = nullptr, *xs = nullptr;  // 4 nb each
  ~BcrsOp() override {
    for (double* p : {t1, t2, t3, xs})
      if (p) (void)hipFree(p);
  }
  long long n() const { return 4LL * B->nb; }
  unsigned gv() const { return (unsigned)((n() + 255) / 256); }
  int alloc() {
    for (double** p : {&t1, &t2, &t3, &xs})
      if (!*p) MF_HIP(hipMalloc(p, std::max<long long>(n(), 1) * sizeof(double)));
    return 0;
  }
  // out = (neg ? -1 : 1) M^-1 (b - A x), b null: zero
  int precon_residual(const double* b, const double* x, double* out, const int* stop, int neg) {
    hipStream_t s = stream();
    if (kind == 0 || kind == 5) {
      if (kind == 0) {
        if (bcrs_residual(*B, x, b, out, stop)) return -2;
      } else {
        if (bcrs_residual(*B, x, b, t1, stop)) return -2;
        if (bcrs_ilu_solve(*B, t1, out, stop)) return -2;
      }
      if (neg) hipLaunchKernelGGL(k_bcrs_vneg, dim3(gv()), dim3(256), 0, s, n(), out, stop);
      MF_HIP(hipGetLastError());
      return 0;
    }
    hipLaunchKernelGGL(k_bcrs_vscale, dim3(gv()), dim3(256), 0, s, n(), x, nullptr, 0, t2, stop);
    if (kind == 4) {
      if (bcrs_residual(*B, t2, b, t1, stop)) return -2;
      if (bcrs_ilu_solve(*B, t1, t3, stop)) return -2;
      hipLaunchKernelGGL(k_bcrs_vadd, dim3(gv()), dim3(256), 0, s, n(), t3, t2, stop);
    } else if (bcrs_sweeps(*B, kind == 2, nr_prec, b, t2, t1, stop, false)) {
      return -2;
    }
    hipLaunchKernelGGL(k_bcrs_vsub, dim3(gv()), dim3(256), 0, s, n(), t2, x, out, neg, stop);
    MF_HIP(hipGetLastError());
    return 0;
  }
  int residual(const double* b, const double* x, double* r) override {
    return precon_residual(b, x, r, nullptr, 0);
  }
  int apply(const double* v, const double* scale, int j, double* w, const int* stop) override {
    hipLaunchKernelGGL(k_bcrs_vscale, dim3(gv()), dim3(256), 0, stream(), n(), v, scale, j, xs,
                       stop);
    return precon_residual(nullptr, xs, w, stop, 1);
  }
};

ExtOp* bcrs_make_op(BcrsDev* B, int kind, int nr_prec) {
  BcrsOp* op = new BcrsOp();
  op->B = B;
  op->kind = kind;
  op->nr_prec = nr_prec;
  if (op->alloc()) {
    delete op;
    return nullptr;
  }
  return op;
}

}  // namespace mf
