// Krylov solver kernels for gfx950: CRS-stream operator, fused Arnoldi
// steps of restarted GMRES, and the reference-order row operations of the
// storage seam.
//
// Reference behaviour followed (paths relative to modfem2017/src/):
//   GMRES(m), CGS, Givens   lsd_mkb/lsd_mkb_core/lss_mkb_core.c:2317-2668
//   preconditioned operator lss_mkb_core.c:1861-2097 (compreres)
//   residual / BJ sweep     lsd_mkb/lad_crs/las_crs_intf.c:788-1013
//   Jacobi fill             las_crs_intf.c:707-737
//
// Layout: CRS values (FP64), columns and row pointers (int32); the Krylov
// basis V is (m+1) vectors with a padded stride ldv (multiple of 64 doubles)
// so that every vector starts 512-byte aligned.
//
// One Arnoldi step j is four launches on the library stream:
//   k_op_dots    w = s_j D^-1 A V_j written to V_{j+1}; per-block partial
//                dot products V_i . w for i <= j.  Rows are processed in
//                tiles: a tile's CRS entries are streamed coalesced into LDS
//                products, then one lane owns one row for the row sum, the
//                scaling and the dot products.
//   k_fin_dots   one block: h_i = s_i sum(partials), c_i = h_i s_i
//   k_update     w -= sum_i c_i V_i, per-block partial ||w||^2
//   k_fin_update one block: h_{j+1,j} = ||w||, Givens rotations, stopping test
// The basis vectors stay unnormalised (scale factors s_i = 1/||V_i||), so no
// separate normalisation pass is needed.  Every kernel of a step returns at
// once when the stopping flag is set, so a whole restart cycle is enqueued
// without host synchronisation.  Blocks are remapped so that each XCD
// streams one contiguous eighth of the rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <map>
#include <tuple>
#include <vector>

#include "mf_device.h"
#include "mf_solver_dev.h"

namespace mf {


static int g_ncu = 0;

static int num_cus() {
  if (!g_ncu) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 256;
    g_ncu = p.multiProcessorCount;
  }
  return g_ncu;
}

// ---------------------------------------------------------------------
// optional per-launch timing of the Arnoldi kernels (HIP events on the
// library stream) with the algorithmic bytes of every launch
// ---------------------------------------------------------------------
struct KStats {
  bool on = false;
  std::vector<hipEvent_t> ev;  // (start, stop) per recorded launch
  std::vector<int> kind;       // 0: op_dots, 1: update
  std::vector<double> bytes;
  std::vector<int> step;       // Arnoldi step index of the launch
  size_t used = 0;
  double t[2] = {0, 0}, b[2] = {0, 0};
  long long nl[2] = {0, 0};
};
static KStats g_ks;

void kstats_enable(int on) {
  g_ks.on = on != 0;
  g_ks.t[0] = g_ks.t[1] = g_ks.b[0] = g_ks.b[1] = 0;
  g_ks.nl[0] = g_ks.nl[1] = 0;
}

void kstats_get(double* out) {
  out[0] = (double)g_ks.nl[0]; out[1] = g_ks.t[0]; out[2] = g_ks.b[0];
  out[3] = (double)g_ks.nl[1]; out[4] = g_ks.t[1]; out[5] = g_ks.b[1];
}

static hipEvent_t ks_event() {
  if (g_ks.used == g_ks.ev.size()) {
    hipEvent_t e;
    (void)hipEventCreate(&e);
    g_ks.ev.push_back(e);
  }
  return g_ks.ev[g_ks.used++];
}

// accumulate the launches of Arnoldi steps < steps_done (after a sync)
static void ks_collect(int steps_done) {
  for (size_t q = 0; q < g_ks.kind.size() && 2 * q + 1 < g_ks.used; q++) {
    if (g_ks.step[q] >= steps_done) continue;  // returned at once (stopped)
    float ms = 0;
    (void)hipEventElapsedTime(&ms, g_ks.ev[2 * q], g_ks.ev[2 * q + 1]);
    int k = g_ks.kind[q];
    g_ks.t[k] += ms;
    g_ks.b[k] += g_ks.bytes[q];
    g_ks.nl[k]++;
  }
  g_ks.used = 0;
  g_ks.kind.clear();
  g_ks.bytes.clear();
  g_ks.step.clear();
}

// V_0 = D^-1 (b - A x) over slices [s0, s1); partial ||V_0||^2 to slot pb
// (of the stored, possibly FP32-rounded, V_0)
template <typename VT>
__device__ __forceinline__ void residual_body(const SolveArgs& A, int s0, int s1, int pb) {
  __shared__ double red[1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  VT* V = basis<VT>(A);
  double acc = 0.0;
  for (int sl = s0 + wv; sl < s1; sl += SW) {
    const double s = sell_row_op(A, A.x, sl, lane);
    const int r = sl * 64 + lane;
    if (r < A.n) {
      const VT v = (VT)jacobi(A, r, A.b[r] - s);
      V[r] = v;
      acc += (double)v * (double)v;
    }
  }
  block_sum_multi<1>(&acc, 1, red);
  if (threadIdx.x == 0) A.partials[pb] = red[0];
}

__device__ __forceinline__ void residual_any(const SolveArgs& A, int s0, int s1, int pb) {
  if (A.v32)
    residual_body<float>(A, s0, s1, pb);
  else
    residual_body<double>(A, s0, s1, pb);
}

__global__ void __launch_bounds__(SB) k_residual_sell(SolveArgs A) {
  const int nsl = (A.n + 63) >> 6;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int s0, s1;
  slice_range(nsl, lb, s0, s1);
  residual_any(A, s0, s1, lb);
}

// Arnoldi step j, part 1 (SELL-64).  Phase 1: w = s_j D^-1 A V_j for the
// block's slices, written to V_{j+1} and staged in LDS.  Phase 2: wavefront q
// forms V_i . w over the block's rows for i = q, q+SW, ..., streaming V_i with
// four independent accumulators and one wave-wide sum per i, so the number of
// basis vectors costs neither registers nor occupancy.
template <typename VT>
__device__ __forceinline__ void op_dots_body(const SolveArgs& A, int j, const SliceSet& S, int pb,
                                             int G, double* ws) {
  const int ns = S.count();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  VT* V = basis<VT>(A);
  const VT* x = V + (size_t)j * A.ldv;
  VT* w = V + (size_t)(j + 1) * A.ldv;
  const double sj = Scale(A)[j];
  for (int t = wv; t < ns; t += SW) {
    const int sl = S.slice(t);
    const double s = sell_row_op(A, x, sl, lane);
    const int r = sl * 64 + lane;
    double wr = 0.0;
    if (r < A.n) {
      const VT wq = (VT)jacobi(A, r, sj * s);
      w[r] = wq;
      wr = (double)wq;
    }
    ws[t * 64 + lane] = wr;
  }
  __syncthreads();
  for (int i = wv; i <= j; i += SW) {
    const VT* v = V + (size_t)i * A.ldv + lane;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int t = 0;
    // rows past n read zero padding of V (ldv covers whole slices) times w = 0
    for (; t + 3 < ns; t += 4) {
      a0 += (double)v[(size_t)S.slice(t) * 64] * ws[t * 64 + lane];
      a1 += (double)v[(size_t)S.slice(t + 1) * 64] * ws[(t + 1) * 64 + lane];
      a2 += (double)v[(size_t)S.slice(t + 2) * 64] * ws[(t + 2) * 64 + lane];
      a3 += (double)v[(size_t)S.slice(t + 3) * 64] * ws[(t + 3) * 64 + lane];
    }
    for (; t < ns; t++) a0 += (double)v[(size_t)S.slice(t) * 64] * ws[t * 64 + lane];
    const double d = wave_sum((a0 + a1) + (a2 + a3));
    if (lane == 0) A.partials[(size_t)i * G + pb] = d;
  }
}

__device__ __forceinline__ void op_dots_any(const SolveArgs& A, int j, const SliceSet& S, int pb,
                                            int G, double* ws) {
  if (A.v32)
    op_dots_body<float>(A, j, S, pb, G, ws);
  else
    op_dots_body<double>(A, j, S, pb, G, ws);
}

__global__ void __launch_bounds__(SB) k_op_dots_sell(SolveArgs A, int j) {
  if (A.stat->stop) return;
  extern __shared__ double ws[];  // w of the block's slices, 64 per slice
  const int nsl = (A.n + 63) >> 6;
  op_dots_any(A, j, block_slices(nsl), blockIdx.x, gridDim.x, ws);
}

// ILU-preconditioned Arnoldi step (precon 2), part 1a: w = s_j A V_j into
// V_{j+1}; the ILU solve then runs in place on w, and k_dots_sell forms the
// CGS dots with the preconditioned w
template <typename VT>
__device__ __forceinline__ void op_body(const SolveArgs& A, int j, const SliceSet& S) {
  const int ns = S.count();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  VT* V = basis<VT>(A);
  const VT* x = V + (size_t)j * A.ldv;
  VT* w = V + (size_t)(j + 1) * A.ldv;
  const double sj = Scale(A)[j];
  for (int t = wv; t < ns; t += SW) {
    const int sl = S.slice(t);
    const double sum = sell_row_op(A, x, sl, lane);
    const int r = sl * 64 + lane;
    if (r < A.n) w[r] = (VT)(sj * sum);
  }
}

__global__ void __launch_bounds__(SB) k_op_sell(SolveArgs A, int j) {
  if (A.stat->stop) return;
  const SliceSet S = block_slices((A.n + 63) >> 6);
  if (A.v32)
    op_body<float>(A, j, S);
  else
    op_body<double>(A, j, S);
}

template <typename VT>
__device__ __forceinline__ void dots_body(const SolveArgs& A, int j, const SliceSet& S, int pb,
                                          int G) {
  const int ns = S.count();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const VT* V = basis<VT>(A);
  const VT* w = V + (size_t)(j + 1) * A.ldv + lane;
  for (int i = wv; i <= j; i += SW) {
    const VT* v = V + (size_t)i * A.ldv + lane;
    double a0 = 0.0, a1 = 0.0;
    int t = 0;
    for (; t + 1 < ns; t += 2) {
      const size_t o0 = (size_t)S.slice(t) * 64, o1 = (size_t)S.slice(t + 1) * 64;
      a0 += (double)v[o0] * (double)w[o0];
      a1 += (double)v[o1] * (double)w[o1];
    }
    for (; t < ns; t++) {
      const size_t o = (size_t)S.slice(t) * 64;
      a0 += (double)v[o] * (double)w[o];
    }
    const double d = wave_sum(a0 + a1);
    if (lane == 0) A.partials[(size_t)i * G + pb] = d;
  }
}

__global__ void __launch_bounds__(SB) k_dots_sell(SolveArgs A, int j) {
  if (A.stat->stop) return;
  const SliceSet S = block_slices((A.n + 63) >> 6);
  if (A.v32)
    dots_body<float>(A, j, S, blockIdx.x, gridDim.x);
  else
    dots_body<double>(A, j, S, blockIdx.x, gridDim.x);
}

// GS-smoothed operator (precon 3, lss_mkb_core.c:2001-2054): with
// X = s_j V_j, w = X - GS(X; b = 0); the sweeps run in place on w
__global__ void k_gs_op_init(SolveArgs A, int j) {
  if (A.stat->stop) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  A.V[(size_t)(j + 1) * A.ldv + i] = Scale(A)[j] * A.V[(size_t)j * A.ldv + i];
}

__global__ void k_gs_op_finish(SolveArgs A, int j) {
  if (A.stat->stop) return;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  double* w = A.V + (size_t)(j + 1) * A.ldv;
  w[i] = Scale(A)[j] * A.V[(size_t)j * A.ldv + i] - w[i];
}

// restart: V_0 = GS(x; b) - x, starting from V_0 = x
__global__ void k_gs_restart_finish(SolveArgs A) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < A.n) A.V[i] += -1.0 * A.x[i];
}

// partial ||V_0||^2 over slices [s0, s1) to slot pb (ILU restart)
template <typename VT>
__device__ __forceinline__ void sumsq_body(const SolveArgs& A, int s0, int s1, int pb) {
  __shared__ double red[1];
  const VT* V = basis<VT>(A);
  double acc = 0.0;
  for (long long r = (long long)s0 * 64 + threadIdx.x; r < (long long)s1 * 64 && r < A.n; r += SB)
    acc += (double)V[r] * (double)V[r];
  block_sum_multi<1>(&acc, 1, red);
  if (threadIdx.x == 0) A.partials[pb] = red[0];
}

__global__ void __launch_bounds__(SB) k_sumsq_v0(SolveArgs A) {
  const int nsl = (A.n + 63) >> 6;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int s0, s1;
  slice_range(nsl, lb, s0, s1);
  if (A.v32)
    sumsq_body<float>(A, s0, s1, lb);
  else
    sumsq_body<double>(A, s0, s1, lb);
}

// y = A x for the operator benchmark: values FP64 or FP32 (VT), vector FP64
// or FP32 (XT), CH entries of a row in flight per lane
template <typename VT, typename XT, int CH>
__global__ void __launch_bounds__(SB) k_spmv_bench(SolveArgs A, const VT* val, const XT* x, XT* y) {
  const int nsl = (A.n + 63) >> 6;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int s0, s1;
  slice_range(nsl, lb, s0, s1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int sl = s0 + wv; sl < s1; sl += SW) {
    const double s = sell_row<VT, XT, CH>(A, val, x, sl, lane);
    const int r = sl * 64 + lane;
    if (r < A.n) y[r] = (XT)s;
  }
}

__global__ void __launch_bounds__(SB) k_spmv_sell(SolveArgs A, const double* x, double* y) {
  const int nsl = (A.n + 63) >> 6;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int s0, s1;
  slice_range(nsl, lb, s0, s1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int sl = s0 + wv; sl < s1; sl += SW) {
    const double s = sell_row(A, A.val, x, sl, lane);
    const int r = sl * 64 + lane;
    if (r < A.n) y[r] = s;
  }
}

// r = b - A x in FP64 and per-block partial ||D^-1 r||^2 (refinement
// residual of the mixed-precision solve)
__global__ void __launch_bounds__(SB) k_refine_residual(SolveArgs A, const double* b,
                                                        const double* x, double* r) {
  __shared__ double red[1];
  const int nsl = (A.n + 63) >> 6;
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int s0, s1;
  slice_range(nsl, lb, s0, s1);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double acc = 0.0;
  for (int sl = s0 + wv; sl < s1; sl += SW) {
    const double s = sell_row(A, A.val, x, sl, lane);
    const int i = sl * 64 + lane;
    if (i < A.n) {
      const double v = b[i] - s;
      r[i] = v;
      const double pv = A.dinv ? v * A.dinv[i] : v;
      acc += pv * pv;
    }
  }
  block_sum_multi<1>(&acc, 1, red);
  if (threadIdx.x == 0) A.partials[lb] = red[0];
}

__global__ void __launch_bounds__(SB) k_sum_partials(const double* partials, int G, double* out) {
  __shared__ double red[1];
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += SB) s += partials[b];
  block_sum_multi<1>(&s, 1, red);
  if (threadIdx.x == 0) *out = red[0];
}

__global__ void k_axpy(long long n, const double* d, double* x) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += d[i];
}

__global__ void k_round_f32(long long n, const double* a, float* b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_fill_jacobi_sell(int n, const long long* sl_ptr, const uint8_t* diag_k,
                                   const double* val, double* dinv) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  dinv[r] = 1.0 / val[sl_ptr[r >> 6] + 64LL * diag_k[r] + (r & 63)];
}

// ---------------------------------------------------------------------
// GMRES kernels
// ---------------------------------------------------------------------
// V_0 = D^-1 (b - A x); per-block partial ||V_0||^2
__global__ void __launch_bounds__(SB) k_residual(SolveArgs A) {
  __shared__ double prod[TILE_NNZ];
  __shared__ double red[1];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int t0, t1;
  tile_range(A, lb, t0, t1);
  double acc = 0.0;
  for (int t = t0; t < t1; t++) {
    const int r0 = A.tile_row[t], r1 = A.tile_row[t + 1];
    double s = tile_rows(A, A.x, r0, r1, prod);
    const int r = r0 + threadIdx.x;
    if (r < r1) {
      double v = A.b[r] - s;
      if (A.dinv) v *= A.dinv[r];
      A.V[r] = v;
      acc += v * v;
    }
  }
  block_sum_multi<1>(&acc, 1, red);
  if (threadIdx.x == 0) A.partials[lb] = red[0];
}

__global__ void __launch_bounds__(SB) k_fin_residual(SolveArgs A, int G) {
  __shared__ double red[1];
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += SB) s += A.partials[b];
  block_sum_multi<1>(&s, 1, red);
  if (threadIdx.x == 0) {
    A.stat->rnrm = sqrt(red[0]);
    if (A.hstat) *A.hstat = *A.stat;
  }
}

// Arnoldi step j, part 1
template <int NV>
__global__ void __launch_bounds__(SB) k_op_dots(SolveArgs A, int j) {
  if (A.stat->stop) return;
  __shared__ double prod[TILE_NNZ];
  __shared__ double red[NV];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int t0, t1;
  tile_range(A, lb, t0, t1);
  const double* x = A.V + (size_t)j * A.ldv;
  double* w = A.V + (size_t)(j + 1) * A.ldv;
  const double sj = Scale(A)[j];
  double acc[NV];
#pragma unroll
  for (int i = 0; i < NV; i++) acc[i] = 0.0;
  for (int t = t0; t < t1; t++) {
    const int r0 = A.tile_row[t], r1 = A.tile_row[t + 1];
    double s = tile_rows(A, x, r0, r1, prod);
    const int r = r0 + threadIdx.x;
    if (r < r1) {
      double wr = sj * s;
      if (A.dinv) wr *= A.dinv[r];
      w[r] = wr;
#pragma unroll
      for (int i = 0; i < NV; i++)
        if (i <= j) acc[i] += A.V[(size_t)i * A.ldv + r] * wr;
    }
  }
  block_sum_multi<NV>(acc, j + 1, red);
  if ((int)threadIdx.x <= j) A.partials[(size_t)threadIdx.x * gridDim.x + lb] = red[threadIdx.x];
}

// one block per dot product i <= j: h_i = s_i * sum_b partials[i][b]
__device__ __forceinline__ void fin_dots_body(const SolveArgs& A, int j, int i, int G) {
  __shared__ double red[1];
  const double* p = A.partials + (size_t)i * G;
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += SB) s += p[b];
  block_sum_multi<1>(&s, 1, red);
  if (threadIdx.x == 0) {
    const double si = Scale(A)[i];
    const double h = si * red[0];
    Hcol(A, j + 1)[i] = h;
    Coef(A)[i] = h * si;
  }
}

__global__ void __launch_bounds__(SB) k_fin_dots(SolveArgs A, int j, int G) {
  if (A.stat->stop) return;
  fin_dots_body(A, j, blockIdx.x, G);
}

// Arnoldi step j, part 2: w -= sum_i c_i V_i over row pairs (16-byte loads),
// per-block partial ||w||^2.  The vectors are read in groups of eight
// independent loads; the subtraction order is i = 0, 1, ..., j.
// rows are updated in groups of W consecutive rows per lane (16-byte
// loads): W = 2 for the FP64 basis, 4 for the FP32 one
template <typename VT> struct Grp;
template <> struct Grp<double> { using T = double2; static constexpr int W = 2; };
template <> struct Grp<float> { using T = float4; static constexpr int W = 4; };

// block r of G of a member updates the row groups [r g / G, (r+1) g / G)
// cs: the coefficients c_i, i <= j, in shared memory (visible to the block)
template <typename VT>
__device__ __forceinline__ void update_body(const SolveArgs& A, int j, int r, int G, int pb,
                                            const double* cs) {
  using GT = typename Grp<VT>::T;
  constexpr int W = Grp<VT>::W;
  __shared__ double red[1];
  const long long ng = (A.n + W - 1) / W;  // ldv is a multiple of W: whole groups stay inside
  const long long p0 = r * ng / G, p1 = (r + 1) * ng / G;
  const long long ldg = A.ldv / W;
  const GT* VG = reinterpret_cast<const GT*>(basis<VT>(A));
  GT* w = reinterpret_cast<GT*>(basis<VT>(A) + (size_t)(j + 1) * A.ldv);
  double acc = 0.0;
  for (long long p = p0 + threadIdx.x; p < p1; p += SB) {
    const GT w0 = w[p];
    double wv[W];
#pragma unroll
    for (int e = 0; e < W; e++) wv[e] = (double)reinterpret_cast<const VT*>(&w0)[e];
    int i = 0;
    for (; i + 7 <= j; i += 8) {
      GT v[8];
#pragma unroll
      for (int q = 0; q < 8; q++) v[q] = VG[(size_t)(i + q) * ldg + p];
#pragma unroll
      for (int q = 0; q < 8; q++)
#pragma unroll
        for (int e = 0; e < W; e++) wv[e] -= cs[i + q] * (double)reinterpret_cast<const VT*>(&v[q])[e];
    }
    for (; i <= j; i++) {
      const GT v = VG[(size_t)i * ldg + p];
#pragma unroll
      for (int e = 0; e < W; e++) wv[e] -= cs[i] * (double)reinterpret_cast<const VT*>(&v)[e];
    }
    GT wr;
#pragma unroll
    for (int e = 0; e < W; e++) {
      const VT q = W * p + e < A.n ? (VT)wv[e] : (VT)0;  // padding rows stay zero
      reinterpret_cast<VT*>(&wr)[e] = q;
      acc += (double)q * (double)q;
    }
    w[p] = wr;
  }
  block_sum_multi<1>(&acc, 1, red);
  if (threadIdx.x == 0) A.partials[pb] = red[0];
}

__device__ __forceinline__ void update_any(const SolveArgs& A, int j, int r, int G, int pb,
                                           const double* cs) {
  if (A.v32)
    update_body<float>(A, j, r, G, pb, cs);
  else
    update_body<double>(A, j, r, G, pb, cs);
}

__global__ void __launch_bounds__(SB) k_update(SolveArgs A, int j) {
  if (A.stat->stop) return;
  __shared__ double cs[MAX_NV];
  if ((int)threadIdx.x <= j) cs[threadIdx.x] = Coef(A)[threadIdx.x];
  __syncthreads();
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  update_any(A, j, lb, gridDim.x, lb, cs);
}

// h_{j+1,j} = ||w||, Givens rotations (drot/drotg as lss_mkb_core.c:2539-2560),
// stopping tests (:2562-2590)
__device__ __forceinline__ void fin_update_body(const SolveArgs& A, int j, int G) {
  __shared__ double red[1];
  __shared__ double col[MAX_NV + 1], giv[2 * MAX_NV], g2[2];
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += SB) s += A.partials[b];
  double* hc = Hcol(A, j + 1);
  double* gv = Giv(A);
  if ((int)threadIdx.x <= j) col[threadIdx.x] = hc[threadIdx.x];
  if ((int)threadIdx.x < 2 * j) giv[threadIdx.x] = gv[threadIdx.x];
  block_sum_multi<1>(&s, 1, red);
  if (threadIdx.x == 0) {
    const int m = A.m;
    const double hn = sqrt(red[0]);
    col[j + 1] = hn;
    Scale(A)[j + 1] = 1.0 / hn;
    for (int i = 1; i <= j; i++) rot(&col[i - 1], &col[i], giv[2 * (i - 1)], giv[2 * (i - 1) + 1]);
    double ra = col[j], rb = col[j + 1], c, sn;
    drotg_dev(&ra, &rb, &c, &sn);
    gv[2 * j] = c;
    gv[2 * j + 1] = sn;
    rot(&col[j], &col[j + 1], c, sn);
    double* g = Hcol(A, 0);
    g2[0] = g[j];
    g2[1] = g[j + 1];
    rot(&g2[0], &g2[1], c, sn);
    g[j] = g2[0];
    g[j + 1] = g2[1];
    GmresStat* st = A.stat;
    st->gmiter += 1;
    st->j_done = j + 1;
    st->resid = fabs(g2[1]);
    st->res_rel = st->beta0 != 0.0 ? st->resid / st->beta0 : 0.0;
    if (st->resid < st->tol_res || st->res_rel < st->tol_rel) {
      st->converged = 1;
      st->stop = 1;
    }
    if (j + 1 == m || st->gmiter == st->max_iter) st->stop = 1;
    if (A.hstat) *A.hstat = *st;
  }
  __syncthreads();
  if ((int)threadIdx.x <= j + 1) hc[threadIdx.x] = col[threadIdx.x];
}

__global__ void __launch_bounds__(SB) k_fin_update(SolveArgs A, int j, int G) {
  if (A.stat->stop) return;
  fin_update_body(A, j, G);
}

// x += sum_c (y_c s_c) V_c over rows [r0, r1), strided by `step`
template <typename VT>
__device__ __forceinline__ void xupdate_body(const SolveArgs& A, int ik, const double* cs,
                                             long long r0, long long r1, long long step) {
  const VT* V = basis<VT>(A);
  for (long long r = r0; r < r1; r += step) {
    double xr = A.x[r];
    for (int c = 0; c < ik; c++) xr += (double)V[(size_t)c * A.ldv + r] * cs[c];
    A.x[r] = xr;
  }
}

__global__ void __launch_bounds__(SB) k_xupdate(SolveArgs A, int ik) {
  __shared__ double cs[MAX_NV];
  if ((int)threadIdx.x < ik) cs[threadIdx.x] = Coef(A)[threadIdx.x];
  __syncthreads();
  const long long r0 = (long long)blockIdx.x * SB + threadIdx.x, step = (long long)gridDim.x * SB;
  if (A.v32)
    xupdate_body<float>(A, ik, cs, r0, A.n, step);
  else
    xupdate_body<double>(A, ik, cs, r0, A.n, step);
}

__global__ void k_fill_jacobi(int n, const int* rp, const int* col, const double* val,
                              double* dinv) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double d = 0.0;
  for (int k = rp[r]; k < rp[r + 1]; k++)
    if (col[k] == r) d = val[k];
  dinv[r] = 1.0 / d;
}

// plain y = A x through the same tiles
__global__ void __launch_bounds__(SB) k_spmv(SolveArgs A, const double* x, double* y) {
  __shared__ double prod[TILE_NNZ];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int t0, t1;
  tile_range(A, lb, t0, t1);
  for (int t = t0; t < t1; t++) {
    const int r0 = A.tile_row[t], r1 = A.tile_row[t + 1];
    double s = tile_rows(A, x, r0, r1, prod);
    const int r = r0 + threadIdx.x;
    if (r < r1) y[r] = s;
  }
}

// Reference-order row operations for the storage seam (lar_*):
//   mode 0: v = (use_rhs ? b : 0) - A x        lar_compute_residual_crs
//                                              (las_crs_intf.c:788-849)
//   mode 1: v = d .* ((use_rhs ? b : 0) - sum_{j!=i} a_ij x_j)
//                                              lar_perform_BJ_or_GS_iterations_crs,
//                                              BLOCK_JACOBI (las_crs_intf.c:857-1013)
// Row sums run in column order over LDS-staged products without FMA
// contraction, matching the reference's sequential "v -= a*x" exactly.
__global__ void __launch_bounds__(SB) k_rowop(SolveArgs A, int mode, int use_rhs,
                                              const double* x, const double* b,
                                              const int* diag_ptr, double* v) {
  __shared__ double prod[TILE_NNZ];
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  int t0, t1;
  tile_range(A, lb, t0, t1);
  for (int t = t0; t < t1; t++) {
    const int r0 = A.tile_row[t], r1 = A.tile_row[t + 1];
    const int k0 = A.row_ptr[r0], k1 = A.row_ptr[r1];
    if (x)
      for (int k = k0 + threadIdx.x; k < k1; k += SB) prod[k - k0] = __dmul_rn(A.val[k], x[A.col[k]]);
    __syncthreads();
    const int r = r0 + threadIdx.x;
    if (r < r1) {
      double s = 0.0;
      if (x) {
        const int a = A.row_ptr[r] - k0, z = A.row_ptr[r + 1] - k0;
        const int dk = (mode == 1) ? diag_ptr[r] - k0 : -1;
        for (int k = a; k < z; k++)
          if (k != dk) s = __dsub_rn(s, prod[k]);
      }
      if (use_rhs) s = __dadd_rn(s, b[r]);
      if (mode == 1) s = __dmul_rn(s, A.dinv[r]);
      v[r] = s;
    }
    __syncthreads();
  }
}

__global__ void k_diag_ptr(int n, const int* rp, const int* col, int* dp) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  int p = -1;
  for (int k = rp[r]; k < rp[r + 1]; k++)
    if (col[k] == r) p = k;
  dp[r] = p;
}

// ---------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------
static long long ldv_of(int n) { return ((long long)n + 63) / 64 * 64; }

static void fill_solve_args(Problem& P, SolveArgs& A, int precon, int m) {
  A.n = P.n; A.ntiles = P.ntiles; A.m = m; A.ldv = ldv_of(P.n);
  A.tile_row = P.d_tile_row; A.row_ptr = P.d_row_ptr; A.col = P.d_col; A.val = P.d_val;
  A.sl_ptr = P.d_sl_ptr; A.scol = P.d_scol;
  A.val32 = P.d_val32; A.op32 = 0;
  A.dinv = precon ? P.d_dinv : nullptr;
  A.b = P.d_rhs; A.x = P.d_x; A.V = P.d_V; A.H = P.d_H; A.partials = P.d_partials;
  A.stat = P.d_stat; A.hstat = P.h_stat;
  A.Vf = reinterpret_cast<float*>(P.d_V); A.dinv32 = nullptr; A.v32 = 0; A.ilu = 0;
  A.scol16 = P.d_scol16; A.sbase = P.d_sbase;
}

// grid sizes: one resident round over the CUs (LDS limits a tile kernel to 5
// blocks per CU), never more blocks than work items
static int grid_tiles(const Problem& P) {
  return std::max(1, std::min(P.ntiles, std::min(MAX_BLOCKS, num_cus() * 5)));
}
// at most MAX_SLICES slices per block: the op kernel stages their w in LDS
constexpr int MAX_SLICES = 32;
static int grid_slices(const Problem& P) {
  const int nsl = (P.n + 63) / 64;
  const int want = std::max(num_cus() * 8, (nsl + MAX_SLICES - 1) / MAX_SLICES);
  return std::max(1, std::min((nsl + SW - 1) / SW, want));
}
static size_t op_lds_bytes(int n, int G) {
  // a block holds <= SW * ceil(part / (nq*SW)) slices, part <= ceil(nsl/nx)
  // the XCD's share and nq >= floor(G/nx) its blocks
  const int nsl = (n + 63) / 64;
  const int nx = G < 8 ? G : 8;
  const int part = (nsl + nx - 1) / nx + 1, nq = G / nx;
  return (size_t)(SW * ((part + nq * SW - 1) / (nq * SW))) * 64 * sizeof(double);
}
static int grid_rows(const Problem& P) {
  long long pairs = (P.n + 1) / 2;
  return (int)std::max(1LL, std::min<long long>((pairs + SB - 1) / SB, std::min(MAX_BLOCKS, num_cus() * 8)));
}

int build_tiles(Problem& P) {
  const auto& rp = P.setup.row_ptr;
  P.tile_row.clear();
  P.tile_row.push_back(0);
  int r = 0;
  while (r < P.n) {
    int start = r;
    while (r < P.n && r - start < SB && rp[r + 1] - rp[start] <= TILE_NNZ) r++;
    if (r == start) {
      set_error("a CRS row exceeds the tile capacity");
      return -1;
    }
    P.tile_row.push_back(r);
  }
  P.ntiles = (int)P.tile_row.size() - 1;
  if (P.d_tile_row) (void)hipFree(P.d_tile_row);
  MF_HIP(hipMalloc(&P.d_tile_row, P.tile_row.size() * sizeof(int)));
  MF_HIP(hipMemcpy(P.d_tile_row, P.tile_row.data(), P.tile_row.size() * sizeof(int),
                   hipMemcpyHostToDevice));
  return 0;
}

int launch_fill_jacobi(Problem& P) {
  if (P.sell && !P.d_diag_k) {
    set_error("Jacobi: this system has no diagonal positions (no point Jacobi)");
    return -1;
  }
  if (!P.d_dinv) MF_HIP(hipMalloc(&P.d_dinv, (size_t)P.n * sizeof(double)));
  P.dinv_valid = true;
  if (P.sell)
    hipLaunchKernelGGL(k_fill_jacobi_sell, dim3((P.n + 255) / 256), dim3(256), 0, stream(), P.n,
                       P.d_sl_ptr, P.d_diag_k, P.d_val, P.d_dinv);
  else
    hipLaunchKernelGGL(k_fill_jacobi, dim3((P.n + 255) / 256), dim3(256), 0, stream(), P.n,
                       P.d_row_ptr, P.d_col, P.d_val, P.d_dinv);
  MF_HIP(hipGetLastError());
  return 0;
}

int launch_rowop(Problem& P, int mode, int use_rhs, const double* d_x, const double* d_b,
                 const int* d_diag_ptr, double* d_v) {
  SolveArgs A;
  fill_solve_args(P, A, mode == 1, 0);
  hipLaunchKernelGGL(k_rowop, dim3(grid_tiles(P)), dim3(SB), 0, stream(), A, mode, use_rhs,
                     d_x, d_b, d_diag_ptr, d_v);
  MF_HIP(hipGetLastError());
  return 0;
}

int launch_diag_ptr(Problem& P, int* d_dp) {
  hipLaunchKernelGGL(k_diag_ptr, dim3((P.n + 255) / 256), dim3(256), 0, stream(), P.n,
                     P.d_row_ptr, P.d_col, d_dp);
  MF_HIP(hipGetLastError());
  return 0;
}

// plain stream copy (bandwidth yardstick, counter calibration)
__global__ void __launch_bounds__(256) k_copy(const double2* __restrict__ a, double2* __restrict__ b,
                                              long long n2) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n2; i += stride) b[i] = a[i];
}

int bench_copy(long long nbytes, int reps, double* ms) {
  const long long n2 = nbytes / 16;
  if (n2 <= 0 || reps < 1) {
    set_error("mf_bench_copy: nbytes >= 16 and reps >= 1 required");
    return -1;
  }
  double2 *a = nullptr, *b = nullptr;
  MF_HIP(hipMalloc(&a, n2 * 16));
  MF_HIP(hipMalloc(&b, n2 * 16));
  MF_HIP(hipMemsetAsync(a, 0, n2 * 16, stream()));
  // 4 blocks per CU sweeping one contiguous window: the fastest plain copy
  // shape measured on MI355X (5.6 TB/s for 1 GiB)
  const int G = (int)std::min<long long>((n2 + 255) / 256, num_cus() * 4LL);
  hipEvent_t e0, e1;
  MF_HIP(hipEventCreate(&e0));
  MF_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_copy, dim3(G), dim3(256), 0, stream(), a, b, n2);
  MF_HIP(hipEventRecord(e0, stream()));
  for (int r = 0; r < reps; r++) hipLaunchKernelGGL(k_copy, dim3(G), dim3(256), 0, stream(), a, b, n2);
  MF_HIP(hipEventRecord(e1, stream()));
  MF_HIP(hipEventSynchronize(e1));
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return 0;
}

int launch_spmv(Problem& P, const double* d_x, double* d_y) {
  SolveArgs A;
  fill_solve_args(P, A, 0, 0);
  if (P.sell)
    hipLaunchKernelGGL(k_spmv_sell, dim3(grid_slices(P)), dim3(SB), 0, stream(), A, d_x, d_y);
  else
    hipLaunchKernelGGL(k_spmv, dim3(grid_tiles(P)), dim3(SB), 0, stream(), A, d_x, d_y);
  MF_HIP(hipGetLastError());
  return 0;
}

static int round_mixed(Problem& P, int precon);

// 16-bit columns of the SELL operator: one wavefront per slice finds the
// smallest and largest column of the slice; a slice spanning fewer than
// 65536 rows stores offsets from its smallest column
__global__ void k_narrow_columns(int nsl, const long long* sl_ptr, const int* scol,
                                 uint16_t* scol16, int* sbase) {
  const int sl = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (sl >= nsl) return;
  const int lane = threadIdx.x & 63;
  const long long b = sl_ptr[sl], e = sl_ptr[sl + 1];
  int lo = 0x7fffffff, hi = -1;
  for (long long q = b + lane; q < e; q += 64) {
    lo = min(lo, scol[q]);
    hi = max(hi, scol[q]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o, 64));
    hi = max(hi, __shfl_xor(hi, o, 64));
  }
  const bool narrow = hi >= lo && hi - lo < 65536;
  if (lane == 0) sbase[sl] = narrow ? lo : -1;
  if (narrow)
    for (long long q = b + lane; q < e; q += 64) scol16[q] = (uint16_t)(scol[q] - lo);
}

static int build_narrow_columns(Problem& P) {
  if (P.d_sbase) return 0;
  const int nsl = (P.n + 63) / 64;
  MF_HIP(hipMalloc(&P.d_scol16, (size_t)std::max<long long>(P.nnz_sell, 1) * sizeof(uint16_t)));
  MF_HIP(hipMalloc(&P.d_sbase, (size_t)std::max(nsl, 1) * sizeof(int)));
  hipLaunchKernelGGL(k_narrow_columns, dim3((nsl + 3) / 4), dim3(256), 0, stream(), nsl, P.d_sl_ptr,
                     P.d_scol, P.d_scol16, P.d_sbase);
  MF_HIP(hipGetLastError());
  return 0;
}


// operator benchmark variants: prec 0 FP64 values and vector, 1 FP32
// values and FP64 vector, 2 FP32 values and vector; ch entries in flight
template <typename VT, typename XT>
static void launch_spmv_bench(const Problem& P, const SolveArgs& A, const VT* val, const void* x,
                              void* y, int ch) {
  const dim3 g(grid_slices(P)), b(SB);
  const XT* xx = static_cast<const XT*>(x);
  XT* yy = static_cast<XT*>(y);
  hipStream_t s = stream();
  if (ch == 8)
    hipLaunchKernelGGL((k_spmv_bench<VT, XT, 8>), g, b, 0, s, A, val, xx, yy);
  else if (ch == 16)
    hipLaunchKernelGGL((k_spmv_bench<VT, XT, 16>), g, b, 0, s, A, val, xx, yy);
  else if (ch == 24)
    hipLaunchKernelGGL((k_spmv_bench<VT, XT, 24>), g, b, 0, s, A, val, xx, yy);
  else
    hipLaunchKernelGGL((k_spmv_bench<VT, XT, 12>), g, b, 0, s, A, val, xx, yy);
}

int bench_spmv_variant(Problem& P, int prec, int ch, int reps, double* ms) {
  const bool wide = prec >= 10;  // the 32-bit columns everywhere
  if (wide) prec -= 10;
  if (!P.sell || prec < 0 || prec > 2 || reps < 1) {
    set_error("mf_bench_spmv_ex: SELL problem, prec 0/1/2 (+10) and reps >= 1 required");
    return -1;
  }
  if (build_narrow_columns(P)) return -2;
  if (prec > 0 && round_mixed(P, 0)) return -2;
  SolveArgs A;
  fill_solve_args(P, A, 0, 0);
  if (wide) {
    A.scol16 = nullptr;
    A.sbase = nullptr;
  }
  void *x = nullptr, *y = nullptr;
  MF_HIP(hipMalloc(&x, (size_t)P.n * sizeof(double)));
  MF_HIP(hipMalloc(&y, (size_t)P.n * sizeof(double)));
  MF_HIP(hipMemsetAsync(x, 0, (size_t)P.n * sizeof(double), stream()));
  auto run = [&]() {
    if (prec == 0)
      launch_spmv_bench<double, double>(P, A, P.d_val, x, y, ch);
    else if (prec == 1)
      launch_spmv_bench<float, double>(P, A, P.d_val32, x, y, ch);
    else
      launch_spmv_bench<float, float>(P, A, P.d_val32, x, y, ch);
  };
  hipEvent_t e0, e1;
  MF_HIP(hipEventCreate(&e0));
  MF_HIP(hipEventCreate(&e1));
  run();
  MF_HIP(hipEventRecord(e0, stream()));
  for (int r = 0; r < reps; r++) run();
  MF_HIP(hipEventRecord(e1, stream()));
  MF_HIP(hipEventSynchronize(e1));
  MF_HIP(hipGetLastError());
  float t = 0;
  (void)hipEventElapsedTime(&t, e0, e1);
  *ms = t / reps;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(x);
  (void)hipFree(y);
  return 0;
}

static int ensure_workspace(Problem& P, int m) {
  const long long ldv = ldv_of(P.n);
  if (P.kryl_alloc >= m && P.d_V) return 0;
  if (P.d_V) (void)hipFree(P.d_V);
  if (P.d_H) (void)hipFree(P.d_H);
  MF_HIP(hipMalloc(&P.d_V, (size_t)(m + 1) * ldv * sizeof(double)));
  MF_HIP(hipMemset(P.d_V, 0, (size_t)(m + 1) * ldv * sizeof(double)));
  MF_HIP(hipMalloc(&P.d_H, (size_t)((m + 1) * (m + 1) + 4 * (m + 1)) * sizeof(double)));
  if (!P.d_partials) {
    const size_t G = std::max<size_t>(MAX_BLOCKS, (size_t)grid_slices(P));
    MF_HIP(hipMalloc(&P.d_partials, G * MAX_NV * sizeof(double)));
  }
  if (!P.d_stat) MF_HIP(hipMalloc(&P.d_stat, sizeof(GmresStat)));
  if (!P.h_stat) MF_HIP(hipHostMalloc(&P.h_stat, sizeof(GmresStat), hipHostMallocDefault));
  if (!P.d_x) MF_HIP(hipMalloc(&P.d_x, (size_t)P.n * sizeof(double)));
  P.kryl_alloc = m;
  return 0;
}

// The operator+dots kernels read the rows [n, ldv) of every basis vector
// (times w = 0) and need them zero.  The FP32 basis shares the storage of
// the FP64 one, so those rows of one view are data of the other: a solve
// clears them in the view it uses before it starts.
__global__ void k_zero_padding(long long ldv, int n, int nvec, int v32, double* V) {
  const long long pad = ldv - n;
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= pad * nvec) return;
  const long long at = (q / pad) * ldv + n + q % pad;
  if (v32)
    reinterpret_cast<float*>(V)[at] = 0.0f;
  else
    V[at] = 0.0;
}

static int zero_padding(const Problem& P, int m, int v32) {
  const long long ldv = ldv_of(P.n), tot = (ldv - P.n) * (m + 1);
  if (tot == 0) return 0;
  hipLaunchKernelGGL(k_zero_padding, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, stream(),
                     ldv, P.n, m + 1, v32, P.d_V);
  MF_HIP(hipGetLastError());
  return 0;
}

template <int NV>
static void launch_step(int G_op, int G_up, SolveArgs& A, int j) {
  hipStream_t s = stream();
  hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
  if (g_ks.on) {
    e0 = ks_event(); e1 = ks_event();
    (void)hipEventRecord(e0, s);
  }
  if (A.sl_ptr)
    hipLaunchKernelGGL(k_op_dots_sell, dim3(G_op), dim3(SB), op_lds_bytes(A.n, G_op), s, A, j);
  else
    hipLaunchKernelGGL(k_op_dots<NV>, dim3(G_op), dim3(SB), 0, s, A, j);
  if (g_ks.on) (void)hipEventRecord(e1, s);
  hipLaunchKernelGGL(k_fin_dots, dim3(j + 1), dim3(SB), 0, s, A, j, G_op);
  if (g_ks.on) {
    e2 = ks_event(); e3 = ks_event();
    (void)hipEventRecord(e2, s);
  }
  hipLaunchKernelGGL(k_update, dim3(G_up), dim3(SB), 0, s, A, j);
  if (g_ks.on) (void)hipEventRecord(e3, s);
  hipLaunchKernelGGL(k_fin_update, dim3(1), dim3(SB), 0, s, A, j, G_up);
  if (g_ks.on) {
    // algorithmic bytes: op = CRS values+columns, row pointers, input
    // vector, Jacobi diagonal, output, and V_0..V_{j-1} for the dots;
    // update = w read+write and V_0..V_j
    const double n = A.n;
    g_ks.kind.push_back(0);
    g_ks.step.push_back(j);
    g_ks.bytes.push_back(-1.0);  // set by the caller (needs nnz)
    g_ks.kind.push_back(1);
    g_ks.step.push_back(j);
    g_ks.bytes.push_back(8.0 * n * (j + 3));
  }
}

static void launch_arnoldi(int G_op, int G_up, SolveArgs& A, int j) {
  int nv = j + 1;
  if (nv <= 8) launch_step<8>(G_op, G_up, A, j);
  else if (nv <= 16) launch_step<16>(G_op, G_up, A, j);
  else if (nv <= 24) launch_step<24>(G_op, G_up, A, j);
  else if (nv <= 32) launch_step<32>(G_op, G_up, A, j);
  else if (nv <= 40) launch_step<40>(G_op, G_up, A, j);
  else if (nv <= 48) launch_step<48>(G_op, G_up, A, j);
  else launch_step<64>(G_op, G_up, A, j);
}

// Restarted left-preconditioned GMRES(m), classical Gram-Schmidt and Givens
// rotations, following lsr_mkb_core_gmres (lss_mkb_core.c:2317-2668): same
// restart count, tolerance floors, stopping tests on the preconditioned
// residual relative to the initial one, and convergence-rate abandonment.
// d_b / d_x: right-hand side and iterate in device memory; op32: the
// operator reads the FP32 copy of the values (Gram-Schmidt, Givens and the
// iterate stay FP64)
// noalloc: no events are created and the status is read back by a plain
// copy (the pinned mirror is absent); d_x already holds the start vector
// unless ini_zero
static int gmres_core(Problem& P, int precon, int krylbas, int max_iter, double tol,
                      double tol_abs, int ini_zero, double* h_x, const double* d_b, double* d_x,
                      int op32, int* iters, double* res_rel_out, double* resid_out,
                      bool noalloc = false) {
  const int m = krylbas;
  if (m < 1 || m >= MAX_NV) {
    set_error("GMRES: number of Krylov vectors must be in [1,63]");
    return -1;
  }
  if (ensure_workspace(P, m)) return -2;
  const bool ext = P.ext != nullptr;
  const bool ilu = !ext && precon == 2, gs = !ext && precon == 3;
  if (ext) {
    if (op32) {
      set_error("GMRES: no FP32 operator for this system");
      return -1;
    }
  } else if (gs) {
    if (!P.sell) {
      set_error("GMRES: Gauss-Seidel needs SELL storage");
      return -1;
    }
    if ((!P.d_dinv || !P.dinv_valid) && launch_fill_jacobi(P)) return -2;
    if (!P.gs && gs_build(P)) return -2;
    if (!P.gs->valid && gs_refresh(P)) return -2;
  } else if (ilu) {
    if (!P.sell) {
      set_error("GMRES: ILU(0) needs SELL storage");
      return -1;
    }
    if (!P.ilu && ilu_build(P, P.ilu_order)) return -2;
    if (!P.ilu->valid && ilu_factor(P)) return -2;
  } else if (precon && (!P.d_dinv || !P.dinv_valid)) {
    if (launch_fill_jacobi(P)) return -2;
  }
  // 16-bit columns of the operator (made once; a no-allocation solve uses
  // them only if an earlier solve made them)
  if (P.sell && !noalloc && build_narrow_columns(P)) return -2;
  SolveArgs A;
  fill_solve_args(P, A, (ilu || gs || ext) ? 0 : precon, m);
  A.b = d_b;
  A.x = d_x;
  A.op32 = op32;
  const int gn = (P.n + 255) / 256;
  const int* d_stop = reinterpret_cast<const int*>(reinterpret_cast<const char*>(P.d_stat) +
                                                   offsetof(GmresStat, stop));
  if (op32 && (!P.sell || !P.d_val32)) {
    set_error("GMRES: the FP32 operator needs SELL storage and its FP32 values");
    return -1;
  }
  if (op32) {
    // inner solve of the mixed-precision refinement: FP32 basis and FP32
    // preconditioner data
    if (gs) {
      set_error("GMRES: no FP32 inner solve with Gauss-Seidel");
      return -1;
    }
    if (ilu && !P.ilu->valid32) {
      set_error("GMRES: the FP32 ILU(0) factors are not current");
      return -1;
    }
    if (!ilu && precon && !P.d_dinv32) {
      set_error("GMRES: the FP32 Jacobi inverses are missing");
      return -1;
    }
    A.v32 = 1;
    A.dinv32 = (!ilu && precon) ? P.d_dinv32 : nullptr;
    A.dinv = nullptr;
  }
  if (P.sell && zero_padding(P, m, A.v32)) return -2;
  const int G_op = (P.sell || ext) ? grid_slices(P) : grid_tiles(P), G_up = grid_rows(P);
  const double* d_scale = P.d_H + (size_t)(m + 1) * (m + 1) + 2 * (m + 1);
  if (!P.sell) A.sl_ptr = nullptr;
  hipStream_t s = stream();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!noalloc) {
    MF_HIP(hipEventCreate(&e0));
    MF_HIP(hipEventCreate(&e1));
    MF_HIP(hipEventRecord(e0, s));
  }
  if (ini_zero || (!h_x && !noalloc)) {
    MF_HIP(hipMemsetAsync(d_x, 0, (size_t)P.n * sizeof(double), s));
  } else if (h_x) {
    MF_HIP(hipMemcpyAsync(d_x, h_x, (size_t)P.n * sizeof(double), hipMemcpyHostToDevice, s));
  }
  // status written by the finishing kernels: pinned mirror, or a copy
  auto read_stat = [&](GmresStat& out) -> int {
    if (P.h_stat) {
      out = *P.h_stat;
      return 0;
    }
    MF_HIP(hipMemcpyAsync(&out, P.d_stat, sizeof(GmresStat), hipMemcpyDeviceToHost, s));
    MF_HIP(hipStreamSynchronize(s));
    return 0;
  };
  const double SMALL = 1.0e-15, MAX_CONVRATE = 0.999;
  // both tolerances floored at 1e-12 (lss_mkb_core.c:2377-2378)
  const double tol_rel = tol < 1e-12 ? 1e-12 : tol;
  GmresStat st{};
  st.tol_rel = tol_rel;
  st.tol_res = tol_abs < 1e-12 ? 1e-12 : tol_abs;
  st.max_iter = max_iter;
  int restrt = max_iter / m + 1, restit = 0, gmiter = 0, conv = 0;
  double beta0 = 0, beta1 = 0, convrate = 0, faux = 0, resid = 0, res_rel = 0;
  // Hessenberg columns, Givens pairs, scales and update coefficients (host
  // copy of the device array), sized for the largest basis on the stack
  double H[(MAX_NV + 1) * (MAX_NV + 1) + 4 * (MAX_NV + 1)];
  const size_t Hn = (size_t)(m + 1) * (m + 1) + 4 * (m + 1);
  const double op_bytes_base =
      (op32 ? 8.0 : 12.0) * (double)P.nnz + 4.0 * (P.n + 1) + 24.0 * (double)P.n;
  do {
    st.gmiter = gmiter;
    st.stop = 0;
    st.j_done = 0;
    MF_HIP(hipMemcpyAsync(P.d_stat, &st, sizeof(GmresStat), hipMemcpyHostToDevice, s));
    if (ext) {
      if (P.ext->residual(A.b, A.x, A.V)) return -2;
      hipLaunchKernelGGL(k_sumsq_v0, dim3(G_op), dim3(SB), 0, s, A);
    } else if (gs) {
      // V_0 = GS_k(x; b) - x (the smoothing branch, lss_mkb_core.c:2001-2054)
      MF_HIP(hipMemcpyAsync(A.V, A.x, (size_t)P.n * sizeof(double), hipMemcpyDeviceToDevice, s));
      if (gs_sweeps(P, P.gs_sweeps, A.b, A.V, nullptr)) return -2;
      hipLaunchKernelGGL(k_gs_restart_finish, dim3(gn), dim3(256), 0, s, A);
      hipLaunchKernelGGL(k_sumsq_v0, dim3(G_op), dim3(SB), 0, s, A);
    } else if (P.sell)
      hipLaunchKernelGGL(k_residual_sell, dim3(G_op), dim3(SB), 0, s, A);
    else
      hipLaunchKernelGGL(k_residual, dim3(G_op), dim3(SB), 0, s, A);
    if (ilu) {
      // V_0 = (LU)^-1 (b - A x), lss_mkb_core.c:1953-1977
      if (A.v32 ? ilu_apply32(P, A.Vf, A.Vf, nullptr) : ilu_apply(P, A.V, A.V, nullptr)) return -2;
      hipLaunchKernelGGL(k_sumsq_v0, dim3(G_op), dim3(SB), 0, s, A);
    }
    hipLaunchKernelGGL(k_fin_residual, dim3(1), dim3(SB), 0, s, A, G_op);
    MF_HIP(hipGetLastError());
    MF_HIP(hipStreamSynchronize(s));
    GmresStat rs{};
    if (read_stat(rs)) return -2;
    double rnrm2 = rs.rnrm;
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
    if (!(rnrm2 == rnrm2)) {
      set_error("GMRES: residual is not a number (singular matrix or zero diagonal?)");
      if (e0) (void)hipEventDestroy(e0);
      if (e1) (void)hipEventDestroy(e1);
      return -4;
    }
    if (rnrm2 < st.tol_res || res_rel < tol_rel || rnrm2 < SMALL * SMALL) {
      conv = 1;
      break;
    }
    if (convrate > MAX_CONVRATE || faux > MAX_CONVRATE) break;
    std::fill(H, H + Hn, 0.0);
    H[0] = rnrm2;
    double* scale = H + (size_t)(m + 1) * (m + 1) + 2 * (m + 1);
    scale[0] = 1.0 / rnrm2;
    st.beta0 = beta0;
    st.rnrm = rnrm2;
    MF_HIP(hipMemcpyAsync(P.d_H, H, Hn * sizeof(double), hipMemcpyHostToDevice, s));
    MF_HIP(hipMemcpyAsync(P.d_stat, &st, sizeof(GmresStat), hipMemcpyHostToDevice, s));
    if (ilu || gs || ext) {
      for (int j = 0; j < m; j++) {
        double* w = A.V + (size_t)(j + 1) * A.ldv;
        if (ext) {
          if (P.ext->apply(A.V + (size_t)j * A.ldv, d_scale, j, w, d_stop)) return -2;
        } else if (gs) {
          hipLaunchKernelGGL(k_gs_op_init, dim3(gn), dim3(256), 0, s, A, j);
          if (gs_sweeps(P, P.gs_sweeps, nullptr, w, d_stop)) return -2;
          hipLaunchKernelGGL(k_gs_op_finish, dim3(gn), dim3(256), 0, s, A, j);
        } else {
          hipLaunchKernelGGL(k_op_sell, dim3(G_op), dim3(SB), 0, s, A, j);
          float* wf = A.Vf + (size_t)(j + 1) * A.ldv;
          if (A.v32 ? ilu_apply32(P, wf, wf, d_stop) : ilu_apply(P, w, w, d_stop)) return -2;
        }
        hipLaunchKernelGGL(k_dots_sell, dim3(G_op), dim3(SB), 0, s, A, j);
        hipLaunchKernelGGL(k_fin_dots, dim3(j + 1), dim3(SB), 0, s, A, j, G_op);
        hipLaunchKernelGGL(k_update, dim3(G_up), dim3(SB), 0, s, A, j);
        hipLaunchKernelGGL(k_fin_update, dim3(1), dim3(SB), 0, s, A, j, G_up);
      }
    } else {
      for (int j = 0; j < m; j++) launch_arnoldi(G_op, G_up, A, j);
    }
    MF_HIP(hipGetLastError());
    MF_HIP(hipMemcpyAsync(H, P.d_H, Hn * sizeof(double), hipMemcpyDeviceToHost, s));
    MF_HIP(hipStreamSynchronize(s));
    GmresStat hs{};
    if (read_stat(hs)) return -2;
    if (g_ks.on) {
      for (size_t q = 0; q < g_ks.kind.size(); q++)
        if (g_ks.kind[q] == 0) g_ks.bytes[q] = op_bytes_base + 8.0 * P.n * g_ks.step[q];
      ks_collect(hs.j_done);
    }
    const int ik = hs.j_done;
    gmiter = hs.gmiter;
    resid = hs.resid;
    res_rel = hs.res_rel;
    // y = R^-1 g (dtrsv "U","N","N"), then x += V diag(s) y (dgemv)
    double y[MAX_NV];
    for (int i = ik - 1; i >= 0; i--) {
      double sum = H[i];
      for (int c = i + 1; c < ik; c++) sum -= H[(size_t)(c + 1) * (m + 1) + i] * y[c];
      y[i] = sum / H[(size_t)(i + 1) * (m + 1) + i];
    }
    const double* sc = H + (size_t)(m + 1) * (m + 1) + 2 * (m + 1);
    double coef[MAX_NV];
    for (int c = 0; c < ik; c++) coef[c] = y[c] * sc[c];
    double* d_coef = P.d_H + (size_t)(m + 1) * (m + 1) + 3 * (m + 1);
    if (ik > 0) {
      MF_HIP(hipMemcpyAsync(d_coef, coef, ik * sizeof(double), hipMemcpyHostToDevice, s));
      int gx = (int)std::min<long long>((P.n + SB - 1) / SB, 4096);
      hipLaunchKernelGGL(k_xupdate, dim3(gx), dim3(SB), 0, s, A, ik);
      MF_HIP(hipGetLastError());
    }
    MF_HIP(hipStreamSynchronize(s));
    if (hs.converged) {
      conv = 1;
      break;
    }
    if (gmiter == max_iter) break;
    restit++;
  } while (restit <= restrt);
  if (!noalloc) {
    MF_HIP(hipEventRecord(e1, s));
    MF_HIP(hipEventSynchronize(e1));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    P.t_solve_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (h_x) MF_HIP(hipMemcpy(h_x, d_x, (size_t)P.n * sizeof(double), hipMemcpyDeviceToHost));
  if (iters) *iters = gmiter;
  if (res_rel_out) *res_rel_out = res_rel;
  if (resid_out) *resid_out = resid;
  return conv;
}

int gmres_solve(Problem& P, int precon, int krylbas, int max_iter, double tol, double tol_abs,
                int ini_zero, double* h_x, int* iters, double* res_rel_out, double* resid_out) {
  if (!P.d_x) MF_HIP(hipMalloc(&P.d_x, (size_t)P.n * sizeof(double)));
  return gmres_core(P, precon, krylbas, max_iter, tol, tol_abs, ini_zero, h_x, P.d_rhs, P.d_x, 0,
                    iters, res_rel_out, resid_out);
}

// FP32 copies for mixed-precision inner solves: matrix values, and the
// Jacobi inverses (precon 1) or the ILU(0) factors (precon 2), refreshed
// when the matrix changed since they were made; the 16-bit columns once
static int round_mixed(Problem& P, int precon) {
  hipStream_t s = stream();
  if (build_narrow_columns(P)) return -2;
  if (precon == 1 && (!P.d_dinv || !P.dinv_valid)) {
    if (launch_fill_jacobi(P)) return -2;
    P.val32_valid = false;
  }
  if (!P.val32_valid) {
    if (!P.d_val32) MF_HIP(hipMalloc(&P.d_val32, (size_t)P.nnz_sell * sizeof(float)));
    hipLaunchKernelGGL(k_round_f32, dim3((unsigned)((P.nnz_sell + 255) / 256)), dim3(256), 0, s,
                       P.nnz_sell, P.d_val, P.d_val32);
    if (P.d_dinv && P.dinv_valid) {
      if (!P.d_dinv32) MF_HIP(hipMalloc(&P.d_dinv32, (size_t)P.n * sizeof(float)));
      hipLaunchKernelGGL(k_round_f32, dim3((P.n + 255) / 256), dim3(256), 0, s, (long long)P.n,
                         P.d_dinv, P.d_dinv32);
    }
    MF_HIP(hipGetLastError());
    P.val32_valid = true;
  }
  if (precon == 1 && !P.d_dinv32) {
    MF_HIP(hipMalloc(&P.d_dinv32, (size_t)P.n * sizeof(float)));
    hipLaunchKernelGGL(k_round_f32, dim3((P.n + 255) / 256), dim3(256), 0, s, (long long)P.n,
                       P.d_dinv, P.d_dinv32);
  }
  if (precon == 2 && !P.ilu->valid32 && ilu_round_f32(P)) return -2;
  return 0;
}

// Mixed-precision solve: FP64 iterative refinement around GMRES(m) whose
// operator reads FP32 matrix values (Gram-Schmidt, Givens rotations, Krylov
// basis, corrections and residuals FP64).  Step k: r_k = b - A x_k and
// rho_k = ||D^-1 r_k|| in FP64; stop when rho_k <= tol * rho_0 (the same
// preconditioned relative measure as the plain solve from x_0 = 0);
// otherwise solve A d = r_k with the FP32 operator to the relative tolerance
// max(tol rho_0 / rho_k, 1e-6) and set x_{k+1} = x_k + d.
int gmres_mixed(Problem& P, int precon, int krylbas, int max_iter, double tol, double* h_x,
                int* iters, double* res_rel_out, double* resid_out, int* refinements) {
  if (!P.sell) {
    set_error("mixed-precision GMRES needs a SELL problem");
    return -1;
  }
  if (krylbas < 1 || krylbas >= MAX_NV || max_iter < 1) {
    set_error("GMRES: number of Krylov vectors must be in [1,63], max_iter >= 1");
    return -1;
  }
  if (ensure_workspace(P, krylbas)) return -2;
  const bool ilu = precon == 2;
  if (ilu) {
    if (!P.ilu && ilu_build(P, P.ilu_order)) return -2;
    if (!P.ilu->valid && ilu_factor(P)) return -2;
  } else if (precon && (!P.d_dinv || !P.dinv_valid) && launch_fill_jacobi(P)) {
    return -2;
  }
  hipStream_t s = stream();
  const long long n = P.n;
  if (round_mixed(P, precon)) return -2;
  if (!P.d_work) MF_HIP(hipMalloc(&P.d_work, (size_t)n * sizeof(double)));
  if (!P.d_work2) MF_HIP(hipMalloc(&P.d_work2, (size_t)(n + 1) * sizeof(double)));
  double *d_r = P.d_work, *d_d = P.d_work2, *d_sum = P.d_work2 + n;
  MF_HIP(hipMemsetAsync(P.d_x, 0, (size_t)n * sizeof(double), s));
  SolveArgs A;
  fill_solve_args(P, A, ilu ? 0 : precon, krylbas);
  const int G = grid_slices(P);
  const int gx = (int)((n + 255) / 256);
  int total = 0, conv = 0, k = 0, rc = 0;
  // ILU: the measure is ||(LU)^-1 r||, formed in d_d before each inner solve
  SolveArgs An = A;
  An.V = d_d;
  double rho0 = 0.0, rho = 0.0;
  hipEvent_t e0, e1;
  MF_HIP(hipEventCreate(&e0));
  MF_HIP(hipEventCreate(&e1));
  MF_HIP(hipEventRecord(e0, s));
  for (k = 0; k < 30; k++) {
    hipLaunchKernelGGL(k_refine_residual, dim3(G), dim3(SB), 0, s, A, P.d_rhs, P.d_x, d_r);
    if (ilu) {
      if (ilu_apply(P, d_r, d_d, nullptr)) return -2;
      hipLaunchKernelGGL(k_sumsq_v0, dim3(G), dim3(SB), 0, s, An);
    }
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(SB), 0, s, P.d_partials, G, d_sum);
    double h_sum = 0.0;
    MF_HIP(hipMemcpyAsync(&h_sum, d_sum, sizeof(double), hipMemcpyDeviceToHost, s));
    MF_HIP(hipStreamSynchronize(s));
    rho = std::sqrt(h_sum);
    if (k == 0) rho0 = rho;
    if (!(rho == rho)) {
      set_error("mixed GMRES: residual is not a number");
      rc = -4;
      break;
    }
    if (rho <= tol * rho0 || rho < 1e-300 || total >= max_iter) {
      conv = rho <= tol * rho0 || rho < 1e-300;
      break;
    }
    const double tol_in = std::max(tol * rho0 / rho, 1e-6);
    int it = 0;
    double rr = 0, rs = 0;
    rc = gmres_core(P, precon, krylbas, max_iter - total, tol_in, 0.0, 1, nullptr, d_r, d_d, 1,
                    &it, &rr, &rs);
    if (rc < 0) break;
    total += it;
    hipLaunchKernelGGL(k_axpy, dim3(gx), dim3(256), 0, s, n, d_d, P.d_x);
    MF_HIP(hipGetLastError());
  }
  MF_HIP(hipEventRecord(e1, s));
  MF_HIP(hipEventSynchronize(e1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  P.t_solve_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc < 0) return rc;
  if (h_x) MF_HIP(hipMemcpy(h_x, P.d_x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
  if (iters) *iters = total;
  if (res_rel_out) *res_rel_out = rho0 > 0 ? rho / rho0 : 0.0;
  if (resid_out) *resid_out = rho;
  if (refinements) *refinements = k;
  return conv;
}


// ---------------------------------------------------------------------
// GMRES without device allocation (mf_gmres_noalloc), for callers that may
// not allocate: every work array of the solve (start and solution vectors,
// Jacobi inverses, Krylov basis, Hessenberg data, dot-product partials,
// status) lives in one file-scope device array of fixed size.  Each device
// the library runs on holds its own copy from module load to process end;
// its address is looked up once per device at mf_init.  The array keeps the
// last call's contents, so everything a solve reads before writing is set
// by the call itself; calls that overlap on one device race.
// ---------------------------------------------------------------------
constexpr size_t kNoallocScratchBytes = 128ull << 20;  // 128 MiB per device
constexpr int kNoallocMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];
static double* g_scratch_ptr[kNoallocMaxDevices] = {};

int noalloc_scratch_init(int device) {
  if (device < 0 || device >= kNoallocMaxDevices) {
    set_error("mf_init: device index beyond the no-allocation table");
    return -1;
  }
  if (g_scratch_ptr[device]) return 0;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(std::string("mf_init: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    return -2;
  }
  g_scratch_ptr[device] = static_cast<double*>(p);
  return 0;
}

__global__ void k_gather_perm(int n, const int* __restrict__ perm, const double* __restrict__ src,
                              double* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__global__ void k_scatter_perm(int n, const int* __restrict__ perm, const double* __restrict__ src,
                               double* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[perm[i]] = src[i];
}

// doubles of scratch the solve of P with m Krylov vectors needs (each
// region starts on a 256-byte boundary)
static size_t noalloc_doubles(const Problem& P, int m) {
  auto up = [](size_t d) { return (d + 31) / 32 * 32; };
  const size_t ldv = (size_t)ldv_of(P.n);
  const size_t G = std::max<size_t>(MAX_BLOCKS, (size_t)grid_slices(P));
  return up(ldv) * 3                                   // reference x, device x, Jacobi
         + up(((size_t)P.n + 1) / 2)                   // row permutation (int)
         + up(G * MAX_NV)                              // dot-product partials
         + up((sizeof(GmresStat) + 7) / 8)             // status
         + up((size_t)(m + 1) * (m + 1) + 4 * (m + 1)) // Hessenberg data
         + (size_t)(m + 1) * ldv;                      // Krylov basis
}

int gmres_noalloc(Problem& P, int precon, int krylbas, int max_iter, double tol, int ini_zero,
                  double* h_x, int* iters, double* res_rel, double* resid) {
  const int m = krylbas;
  if (precon != 0 && precon != 1) {
    set_error("mf_gmres_noalloc: precon must be 0 (none) or 1 (Jacobi)");
    return -1;
  }
  if (m < 1 || m >= MAX_NV || max_iter < 1 || tol < 0) {
    set_error("mf_gmres_noalloc: krylbas in [1,63], max_iter >= 1 and tol >= 0 required");
    return -1;
  }
  if (!P.sell || P.ext) {
    set_error("mf_gmres_noalloc: needs a problem assembled in SELL storage");
    return -1;
  }
  int dev = -1;
  MF_HIP(hipGetDevice(&dev));
  double* base = (dev >= 0 && dev < kNoallocMaxDevices) ? g_scratch_ptr[dev] : nullptr;
  if (!base) {
    set_error("mf_gmres_noalloc: call mf_init on this device first");
    return -1;
  }
  const size_t need = noalloc_doubles(P, m);
  if (need * sizeof(double) > kNoallocScratchBytes) {
    set_error("mf_gmres_noalloc: the solve needs " + std::to_string(need * sizeof(double) >> 20) +
              " MiB of work space, more than the 128 MiB the no-allocation path holds");
    return -1;
  }
  auto up = [](size_t d) { return (d + 31) / 32 * 32; };
  const size_t ldv = (size_t)ldv_of(P.n);
  const size_t G = std::max<size_t>(MAX_BLOCKS, (size_t)grid_slices(P));
  double* q = base;
  double* xr = q; q += up(ldv);
  double* xd = q; q += up(ldv);
  double* dinv = q; q += up(ldv);
  int* perm = reinterpret_cast<int*>(q); q += up(((size_t)P.n + 1) / 2);
  double* part = q; q += up(G * MAX_NV);
  GmresStat* stat = reinterpret_cast<GmresStat*>(q); q += up((sizeof(GmresStat) + 7) / 8);
  double* Hd = q; q += up((size_t)(m + 1) * (m + 1) + 4 * (m + 1));
  double* V = q;
  hipStream_t s = stream();
  const int gn = (P.n + 255) / 256;
  MF_HIP(hipMemsetAsync(V, 0, (size_t)(m + 1) * ldv * sizeof(double), s));
  MF_HIP(hipMemsetAsync(stat, 0, sizeof(GmresStat), s));
  if (!ini_zero && h_x) {
    // start vector in reference order -> device order
    MF_HIP(hipMemcpyAsync(xr, h_x, (size_t)P.n * sizeof(double), hipMemcpyHostToDevice, s));
    MF_HIP(hipMemcpyAsync(perm, P.setup.perm.data(), (size_t)P.n * sizeof(int),
                          hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_gather_perm, dim3(gn), dim3(256), 0, s, P.n, perm, xr, xd);
  } else {
    MF_HIP(hipMemsetAsync(xd, 0, (size_t)P.n * sizeof(double), s));
  }
  // the solve's work arrays point into the scratch array for this call
  // only; the problem's own pointers come back on every exit path
  struct Swap {
    Problem& P;
    double *V, *H, *partials, *x, *dinv;
    GmresStat *stat, *hstat;
    int kryl;
    bool dinv_valid, ks_on;
    explicit Swap(Problem& p)
        : P(p), V(p.d_V), H(p.d_H), partials(p.d_partials), x(p.d_x), dinv(p.d_dinv),
          stat(p.d_stat), hstat(p.h_stat), kryl(p.kryl_alloc), dinv_valid(p.dinv_valid),
          ks_on(g_ks.on) {}
    ~Swap() {
      P.d_V = V; P.d_H = H; P.d_partials = partials; P.d_x = x; P.d_dinv = dinv;
      P.d_stat = stat; P.h_stat = hstat; P.kryl_alloc = kryl; P.dinv_valid = dinv_valid;
      g_ks.on = ks_on;
    }
  } swap(P);
  P.d_V = V; P.d_H = Hd; P.d_partials = part; P.d_x = xd; P.d_dinv = dinv;
  P.d_stat = stat; P.h_stat = nullptr; P.kryl_alloc = m; P.dinv_valid = false;
  g_ks.on = false;
  // ini_zero 0 with no host vector: the iterate already sits in xd
  const int rc = gmres_core(P, precon, m, max_iter, tol, 0.0, 0, nullptr, P.d_rhs, xd, 0, iters,
                            res_rel, resid, true);
  if (rc >= 0 && h_x) {
    MF_HIP(hipMemcpyAsync(perm, P.setup.perm.data(), (size_t)P.n * sizeof(int),
                          hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_scatter_perm, dim3(gn), dim3(256), 0, s, P.n, perm, xd, xr);
    MF_HIP(hipMemcpyAsync(h_x, xr, (size_t)P.n * sizeof(double), hipMemcpyDeviceToHost, s));
    MF_HIP(hipStreamSynchronize(s));
  }
  return rc;
}

// ---------------------------------------------------------------------
// Batched GMRES.  All members advance one restart cycle per round of
// launches: residual, restart test, m Arnoldi steps (operator+dots, finish
// dots, update, finish update), y = R^-1 g, x update, bookkeeping.  The
// restart logic of gmres_core runs on the device (kb_restart, kb_advance),
// so a cycle needs no host decision and can be replayed as one hipGraph.
// Members that have stopped leave every later launch at once.
//
// Members preconditioned by multicolour ILU(0) add, after the residual and
// after each operator application, one launch per substitution level that
// covers the rows of that level in every such member (the number of levels
// is the number of colours, not the number of members), followed by the
// norm or the Gram-Schmidt dots of the preconditioned vector.
//
// Mixed-precision members run FP32-valued inner solves (FP32 values, basis,
// Jacobi inverses or ILU factors; FP64 sums and Hessenberg data) on the
// correction equation A d = r.  Their FP64 refinement steps x += d,
// r = b - A x, rho = ||D^-1 r|| run on the device before and after every
// cycle (three launches each, no host decision), restarting the inner solve
// to max(tol rho_0 / rho, 1e-5) until rho <= tol rho_0.
// ---------------------------------------------------------------------
struct BatchArgs {
  const SolveArgs* args;
  const int* blk_prob;
  const int* blk_rank;
  const int* nblk;
};

__global__ void __launch_bounds__(SB) kb_residual(BatchArgs B) {
  const int p = B.blk_prob[blockIdx.x], r = B.blk_rank[blockIdx.x], G = B.nblk[p];
  const SolveArgs A = B.args[p];
  if (A.stat->done) return;
  const int nsl = (A.n + 63) >> 6;
  residual_any(A, (int)((long long)r * nsl / G), (int)((long long)(r + 1) * nsl / G), r);
}

// restart test of lsr_mkb_core_gmres (lss_mkb_core.c:2404-2470) per member
__global__ void __launch_bounds__(SB) kb_restart(BatchArgs B) {
  const SolveArgs A = B.args[blockIdx.x];
  GmresStat* st = A.stat;
  if (st->done) return;
  __shared__ double red[1];
  __shared__ int go;
  const int G = B.nblk[blockIdx.x];
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += SB) s += A.partials[b];
  block_sum_multi<1>(&s, 1, red);
  if (threadIdx.x == 0) {
    const double SMALL = 1.0e-15, MAX_CONVRATE = 0.999;
    const double rnrm2 = sqrt(red[0]);
    double convrate = 0.0, faux = 0.0;
    st->stop = 0;
    st->j_done = 0;
    if (st->restit == 0) {
      st->beta0 = rnrm2;
      st->beta1 = rnrm2;
    } else if (rnrm2 < SMALL * SMALL || st->gmiter == 0 || st->beta0 < SMALL * SMALL ||
               st->beta1 < SMALL * SMALL) {
      st->beta1 = 0.0;
    } else {
      convrate = exp(log(rnrm2 / st->beta0) / st->gmiter);
      faux = exp(log(rnrm2 / st->beta1) / st->m);
      st->beta1 = rnrm2;
    }
    st->resid = rnrm2;
    st->res_rel = st->beta0 != 0.0 ? rnrm2 / st->beta0 : 0.0;
    st->rnrm = rnrm2;
    go = 0;
    if (!(rnrm2 == rnrm2)) {
      st->nan = 1;
      st->done = 1;
      st->stop = 1;
    } else if (rnrm2 < st->tol_res || st->res_rel < st->tol_rel || rnrm2 < SMALL * SMALL) {
      st->converged = 1;
      st->done = 1;
      st->stop = 1;
    } else if (convrate > MAX_CONVRATE || faux > MAX_CONVRATE) {
      st->done = 1;
      st->stop = 1;
    } else {
      go = 1;
    }
  }
  __syncthreads();
  if (!go) return;
  const int nh = (A.m + 1) * (A.m + 1) + 4 * (A.m + 1);
  for (int k = threadIdx.x; k < nh; k += SB) A.H[k] = 0.0;
  __syncthreads();
  if (threadIdx.x == 0) {
    A.H[0] = st->rnrm;
    Scale(A)[0] = 1.0 / st->rnrm;
  }
}

__device__ __forceinline__ SliceSet member_slices(int n, int r, int G) {
  const int nsl = (n + 63) >> 6;
  SliceSet S;
  S.base = (int)((long long)r * nsl / G);
  S.end = (int)((long long)(r + 1) * nsl / G);
  S.step = SW;  // contiguous
  return S;
}

// ILU members: partial ||(LU)^-1 r||^2 of the preconditioned V_0
__global__ void __launch_bounds__(SB) kb_sumsq(BatchArgs B) {
  const int p = B.blk_prob[blockIdx.x], r = B.blk_rank[blockIdx.x], G = B.nblk[p];
  const SolveArgs A = B.args[p];
  if (!A.ilu || A.stat->done) return;
  const SliceSet S = member_slices(A.n, r, G);
  if (A.v32)
    sumsq_body<float>(A, S.base, S.end, r);
  else
    sumsq_body<double>(A, S.base, S.end, r);
}

// one substitution level of every ILU member: block b covers up to SW
// slices [s0[b], s1[b]) of member mem[b]; j < 0 preconditions V_0 (restart),
// j >= 0 the new vector V_{j+1} of Arnoldi step j
template <bool BWD>
__global__ void __launch_bounds__(SB) kb_ilu_level(BatchArgs B, const IluMember* L, const int* mem,
                                                   const int* s0, const int* s1, int off, int j) {
  const int b = off + blockIdx.x;
  const int k = mem[b];
  const SolveArgs& A = B.args[k];
  if (j < 0 ? A.stat->done : A.stat->stop) return;
  const int sl = s0[b] + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (sl >= s1[b]) return;
  const IluMember& M = L[k];
  const size_t v = (size_t)(j < 0 ? 0 : j + 1) * A.ldv;
  if (A.v32) {
    if (BWD)
      ilu_bwd_row(M.b32, sl, lane, A.Vf + v);
    else
      ilu_fwd_row(M.f32, sl, lane, A.Vf + v);
  } else {
    if (BWD)
      ilu_bwd_row(M.b, sl, lane, A.V + v);
    else
      ilu_fwd_row(M.f, sl, lane, A.V + v);
  }
}

__global__ void __launch_bounds__(SB) kb_op_dots(BatchArgs B, int j) {
  const int p = B.blk_prob[blockIdx.x], r = B.blk_rank[blockIdx.x], G = B.nblk[p];
  const SolveArgs A = B.args[p];
  if (A.stat->stop) return;
  extern __shared__ double ws[];
  const SliceSet S = member_slices(A.n, r, G);
  if (!A.ilu)
    op_dots_any(A, j, S, r, G, ws);
  else if (A.v32)
    op_body<float>(A, j, S);  // dots after the substitution (kb_dots)
  else
    op_body<double>(A, j, S);
}

// ILU members: Gram-Schmidt dots with the preconditioned w
__global__ void __launch_bounds__(SB) kb_dots(BatchArgs B, int j) {
  const int p = B.blk_prob[blockIdx.x], r = B.blk_rank[blockIdx.x], G = B.nblk[p];
  const SolveArgs A = B.args[p];
  if (!A.ilu || A.stat->stop) return;
  const SliceSet S = member_slices(A.n, r, G);
  if (A.v32)
    dots_body<float>(A, j, S, r, G);
  else
    dots_body<double>(A, j, S, r, G);
}

__global__ void __launch_bounds__(SB) kb_fin_dots(BatchArgs B, int j) {
  const SolveArgs A = B.args[blockIdx.y];
  if (A.stat->stop) return;
  fin_dots_body(A, j, blockIdx.x, B.nblk[blockIdx.y]);
}

__global__ void __launch_bounds__(SB) kb_update(BatchArgs B, int j) {
  const int p = B.blk_prob[blockIdx.x], r = B.blk_rank[blockIdx.x], G = B.nblk[p];
  const SolveArgs A = B.args[p];
  if (A.stat->stop) return;
  __shared__ double cs[MAX_NV];
  if ((int)threadIdx.x <= j) cs[threadIdx.x] = Coef(A)[threadIdx.x];
  __syncthreads();
  update_any(A, j, r, G, r, cs);
}

__global__ void __launch_bounds__(SB) kb_fin_update(BatchArgs B, int j) {
  const SolveArgs A = B.args[blockIdx.x];
  if (A.stat->stop) return;
  fin_update_body(A, j, B.nblk[blockIdx.x]);
}

// y = R^-1 g (dtrsv "U","N","N", lss_mkb_core.c:2591-2593), coefficients y_c s_c
__global__ void __launch_bounds__(64) kb_ysolve(BatchArgs B) {
  const SolveArgs A = B.args[blockIdx.x];
  GmresStat* st = A.stat;
  if (st->done || threadIdx.x != 0) return;
  __shared__ double y[MAX_NV];
  const int m = A.m, ik = st->j_done;
  for (int i = ik - 1; i >= 0; i--) {
    double sum = A.H[i];
    for (int c = i + 1; c < ik; c++) sum -= A.H[(size_t)(c + 1) * (m + 1) + i] * y[c];
    y[i] = sum / A.H[(size_t)(i + 1) * (m + 1) + i];
  }
  for (int c = 0; c < ik; c++) Coef(A)[c] = y[c] * Scale(A)[c];
  st->ik_pending = ik;
}

// x += sum_c coef_c V_c over the block's rows (dgemv, lss_mkb_core.c:2602)
__global__ void __launch_bounds__(SB) kb_xupdate(BatchArgs B) {
  const int p = B.blk_prob[blockIdx.x], r = B.blk_rank[blockIdx.x], G = B.nblk[p];
  const SolveArgs A = B.args[p];
  const int ik = A.stat->ik_pending;
  if (ik == 0) return;
  __shared__ double cs[MAX_NV];
  if ((int)threadIdx.x < ik) cs[threadIdx.x] = Coef(A)[threadIdx.x];
  __syncthreads();
  const long long r0 = (long long)r * A.n / G, r1 = (long long)(r + 1) * A.n / G;
  if (A.v32)
    xupdate_body<float>(A, ik, cs, r0 + threadIdx.x, r1, SB);
  else
    xupdate_body<double>(A, ik, cs, r0 + threadIdx.x, r1, SB);
}

// end of a restart (lss_mkb_core.c:2606-2668): converged, out of
// iterations, or the next restart
__global__ void kb_advance(BatchArgs B, int nprob) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nprob) return;
  GmresStat* st = B.args[p].stat;
  if (st->ik_pending == 0) return;
  st->ik_pending = 0;
  if (st->converged) {
    st->done = 1;
  } else if (st->gmiter == st->max_iter) {
    st->done = 1;
  } else {
    st->restit += 1;
    if (st->restit > st->restrt) st->done = 1;
  }
  if (st->done) st->stop = 1;
}

// Refinement steps of mixed-precision members, on the device.  A step is
// due for a member that has not ended when its first step has not been
// taken (phase 1) or its inner solve has ended: x += d (not on the first
// step), d = 0, r = b - A x and rho = ||D^-1 r|| in FP64 (D the Jacobi
// diagonal; the plain norm for precon 0).  The member then ends (rho <= tol
// rho_0, out of iterations or refinement steps, or not a number) or starts
// a new inner solve to the relative tolerance max(tol rho_0 / rho, 1e-5).
constexpr double kInnerTolFloor = 1e-5;
constexpr int kMaxRefinements = 30;

__device__ __forceinline__ bool refine_due(const MixedState& M, const GmresStat& st) {
  return M.mixed && !M.fin && (M.phase == 1 || st.done);
}

__global__ void __launch_bounds__(SB) kb_refine_axpy(BatchArgs B, const RefineArgs* R,
                                                     const MixedState* mix) {
  const int p = B.blk_prob[blockIdx.x], r = B.blk_rank[blockIdx.x], G = B.nblk[p];
  const SolveArgs& A = B.args[p];
  if (!refine_due(mix[p], *A.stat)) return;
  const bool add = mix[p].phase != 1;
  double* x = R[p].x;
  const long long r0 = (long long)r * A.n / G, r1 = (long long)(r + 1) * A.n / G;
  for (long long i = r0 + threadIdx.x; i < r1; i += SB) {
    if (add) x[i] += A.x[i];
    A.x[i] = 0.0;
  }
}

// r = b - A x with the FP64 values, partial ||D^-1 r||^2 per block
__global__ void __launch_bounds__(SB) kb_refine_resid(BatchArgs B, const RefineArgs* R,
                                                      const MixedState* mix) {
  const int p = B.blk_prob[blockIdx.x], r = B.blk_rank[blockIdx.x], G = B.nblk[p];
  const SolveArgs& A = B.args[p];
  if (!refine_due(mix[p], *A.stat)) return;
  __shared__ double red[1];
  const RefineArgs& Q = R[p];
  double* res = const_cast<double*>(A.b);
  const int nsl = (A.n + 63) >> 6;
  const int s0 = (int)((long long)r * nsl / G), s1 = (int)((long long)(r + 1) * nsl / G);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double acc = 0.0;
  for (int sl = s0 + wv; sl < s1; sl += SW) {
    const double sum = sell_row(A, A.val, Q.x, sl, lane);
    const int i = sl * 64 + lane;
    if (i < A.n) {
      const double v = Q.rhs[i] - sum;
      res[i] = v;
      const double pv = Q.dinv ? v * Q.dinv[i] : v;
      acc += pv * pv;
    }
  }
  block_sum_multi<1>(&acc, 1, red);
  if (threadIdx.x == 0) A.partials[r] = red[0];
}

// one block per member: rho, the test, and the next inner solve
__global__ void __launch_bounds__(SB) kb_refine_fin(BatchArgs B, MixedState* mix, int m) {
  const int p = blockIdx.x;
  const SolveArgs& A = B.args[p];
  if (!refine_due(mix[p], *A.stat)) return;
  __shared__ double red[1];
  const int G = B.nblk[p];
  double s = 0.0;
  for (int b = threadIdx.x; b < G; b += SB) s += A.partials[b];
  block_sum_multi<1>(&s, 1, red);
  if (threadIdx.x != 0) return;
  MixedState& M = mix[p];
  GmresStat* st = A.stat;
  const double rho = sqrt(red[0]);
  if (M.phase == 1)
    M.rho0 = rho;
  else
    M.total += st->gmiter;
  M.rho = rho;
  M.phase = 0;
  if (!(rho == rho) || st->nan) {
    M.fin = M.nan = 1;
  } else if (rho <= M.tol * M.rho0 || rho < 1e-300) {
    M.fin = M.conv = 1;
  } else if (M.total >= M.max_iter || M.outer >= kMaxRefinements) {
    M.fin = 1;
  }
  if (M.fin) {
    st->done = st->stop = 1;
    return;
  }
  // a new inner solve from d = 0 (tolerances floored as lss_mkb_core.c:2377-2378)
  const int mi = M.max_iter - M.total;
  const double tr = fmax(M.tol * M.rho0 / rho, kInnerTolFloor);
  GmresStat z = {};
  z.tol_rel = tr < 1e-12 ? 1e-12 : tr;
  z.tol_res = 1e-12;
  z.max_iter = mi;
  z.m = m;
  z.restrt = mi / m + 1;
  *st = z;
  M.outer++;
}

static BatchArgs batch_args(const Batch& B) {
  BatchArgs a;
  a.args = reinterpret_cast<const SolveArgs*>(B.d_args);
  a.blk_prob = B.d_blk_prob;
  a.blk_rank = B.d_blk_rank;
  a.nblk = B.d_nblk;
  return a;
}

static void batch_drop_graph(Batch& B) {
  if (B.exec) (void)hipGraphExecDestroy(B.exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  B.exec = nullptr;
  B.graph = nullptr;
}

static void batch_drop_ilu_tables(Batch& B) {
  void* ptrs[] = {B.d_ilu_mem, B.d_ilu_s0, B.d_ilu_s1};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  B.d_ilu_mem = B.d_ilu_s0 = B.d_ilu_s1 = nullptr;
  B.ilu_off_f.clear();
  B.ilu_off_b.clear();
}

static void batch_drop_tables(Batch& B) {
  batch_drop_graph(B);
  batch_drop_ilu_tables(B);
  void* ptrs[] = {B.d_args, B.d_blk_prob, B.d_blk_rank, B.d_nblk, B.d_stat, B.d_ilu, B.d_mix,
                  B.d_ref};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  void* hptrs[] = {B.h_pinned, B.h_mix_pinned};
  for (void* q : hptrs)
    if (q) (void)hipHostFree(q);
  B.d_args = nullptr; B.d_blk_prob = B.d_blk_rank = B.d_nblk = nullptr;
  B.d_stat = nullptr; B.h_pinned = nullptr;
  B.d_ilu = nullptr; B.d_mix = nullptr; B.h_mix_pinned = nullptr;
  B.d_ref = nullptr;
  B.cap = 0;
  B.blk_cap = 0;
}

// initial state of a member: tolerances floored as lss_mkb_core.c:2377-2378
static GmresStat member_init(const BatchMember& M, int m) {
  GmresStat st{};
  st.tol_rel = M.tol < 1e-12 ? 1e-12 : M.tol;
  st.tol_res = 1e-12;
  st.max_iter = M.max_iter;
  st.m = m;
  st.restrt = M.max_iter / m + 1;
  return st;
}

// the data a member's solve reads, made current: Jacobi inverses (FP64;
// also the refinement measure of mixed members), ILU(0) factors, FP32 copies
static int prepare_member(const BatchMember& M) {
  Problem& P = *M.P;
  // the Jacobi inverses: the preconditioner of precon 1, and the refinement
  // measure of mixed-precision ILU(0) members (precon 0 measures the plain
  // residual and needs none)
  if ((M.precon == 1 || (M.mixed && M.precon == 2)) && (!P.d_dinv || !P.dinv_valid)) {
    if (launch_fill_jacobi(P)) return -2;
    P.val32_valid = false;  // the FP32 inverses follow the FP64 ones
  }
  if (M.precon == 2 && !P.ilu->valid) {
    if (ilu_factor(P)) return -2;
  }
  if (M.mixed && round_mixed(P, M.precon)) return -2;
  return 0;
}

int batch_add(Batch& B, Problem* P, int pid, int precon, int max_iter, double tol, bool mixed) {
  if (!P->sell) {
    set_error("batch: member needs SELL storage");
    return -1;
  }
  if (max_iter < 1 || tol < 0 || precon < 0 || precon > 2) {
    set_error("batch: precon 0/1/2, max_iter >= 1 and tol >= 0 required");
    return -1;
  }
  if (mixed && tol <= 0) {
    set_error("batch: a mixed-precision member needs tol > 0");
    return -1;
  }
  if (precon == 2 && P->ilu_order != 1) {
    set_error("batch: ILU(0) members need the multicolour elimination order (mf_set_ilu_order 1)");
    return -1;
  }
  for (auto& M : B.mem)
    if (M.pid == pid) {
      set_error("batch: problem is already a member");
      return -1;
    }
  if (ensure_workspace(*P, B.m)) return -2;
  if (build_narrow_columns(*P)) return -2;
  if (zero_padding(*P, B.m, mixed ? 1 : 0)) return -2;
  if (precon == 2 && !P->ilu && ilu_build(*P, 1)) return -2;
  BatchMember M;
  M.P = P; M.pid = pid; M.precon = precon; M.max_iter = max_iter; M.tol = tol;
  M.mixed = mixed;
  if (prepare_member(M)) return -2;
  hipStream_t s = stream();
  if (!P->d_x) MF_HIP(hipMalloc(&P->d_x, (size_t)P->n * sizeof(double)));
  MF_HIP(hipMemsetAsync(P->d_x, 0, (size_t)P->n * sizeof(double), s));
  GmresStat st = member_init(M, B.m);
  if (mixed) {
    if (!P->d_work) MF_HIP(hipMalloc(&P->d_work, (size_t)P->n * sizeof(double)));
    if (!P->d_work2) MF_HIP(hipMalloc(&P->d_work2, (size_t)(P->n + 1) * sizeof(double)));
    MF_HIP(hipMemsetAsync(P->d_work2, 0, (size_t)P->n * sizeof(double), s));
    st.done = st.stop = 1;  // until the first refinement step starts the inner solve
  }
  MixedState ms{};
  ms.mixed = mixed ? 1 : 0;
  ms.phase = mixed ? 1 : 0;
  ms.max_iter = max_iter;
  ms.tol = tol;
  P->batch_refs++;
  B.mem.push_back(M);
  B.h_stat.push_back(st);
  B.h_mix.push_back(ms);
  B.dirty = B.retile = true;
  return (int)B.mem.size() - 1;
}

int batch_remove(Batch& B, int pid) {
  for (size_t k = 0; k < B.mem.size(); k++)
    if (B.mem[k].pid == pid) {
      B.mem[k].P->batch_refs--;
      B.mem.erase(B.mem.begin() + k);
      B.h_stat.erase(B.h_stat.begin() + k);
      B.h_mix.erase(B.h_mix.begin() + k);
      B.dirty = B.retile = true;
      return 0;
    }
  set_error("batch: problem is not a member");
  return -1;
}

static int batch_alloc(Batch& B, int np) {
  if (np <= B.cap) return 0;
  batch_drop_tables(B);
  const int cap = std::max(np, 16);
  MF_HIP(hipMalloc(&B.d_args, (size_t)cap * sizeof(SolveArgs)));
  MF_HIP(hipMalloc(&B.d_nblk, (size_t)cap * sizeof(int)));
  MF_HIP(hipMalloc(&B.d_stat, (size_t)cap * sizeof(GmresStat)));
  MF_HIP(hipMalloc(&B.d_ilu, (size_t)cap * sizeof(IluMember)));
  MF_HIP(hipMalloc(&B.d_mix, (size_t)cap * sizeof(MixedState)));
  MF_HIP(hipHostMalloc(&B.h_pinned, (size_t)cap * sizeof(GmresStat), hipHostMallocDefault));
  MF_HIP(hipHostMalloc(&B.h_mix_pinned, (size_t)cap * sizeof(MixedState), hipHostMallocDefault));
  MF_HIP(hipMalloc(&B.d_ref, (size_t)cap * sizeof(RefineArgs)));
  B.cap = cap;
  B.dirty = true;
  return 0;
}

// substitution level tables of the ILU members: level l of the forward
// (backward) sweep of every member that has it, SW slices per block
static int batch_build_ilu(Batch& B) {
  batch_drop_ilu_tables(B);
  const int np = (int)B.mem.size();
  std::vector<int> bm, b0, b1;
  for (int sweep = 0; sweep < 2; sweep++) {
    std::vector<int>& off = sweep ? B.ilu_off_b : B.ilu_off_f;
    int nlev = 0;
    for (const auto& M : B.mem)
      if (M.precon == 2) nlev = std::max(nlev, sweep ? M.P->ilu->nlev_u : M.P->ilu->nlev_l);
    off.assign(1, (int)bm.size());
    for (int l = 0; l < nlev; l++) {
      for (int k = 0; k < np; k++) {
        const BatchMember& M = B.mem[k];
        if (M.precon != 2) continue;
        const IluData& L = *M.P->ilu;
        if (l >= (sweep ? L.nlev_u : L.nlev_l)) continue;
        const std::vector<int>& ls = sweep ? L.bwd.lev_sl : L.fwd.lev_sl;
        for (int q = ls[l]; q < ls[l + 1]; q += SW) {
          bm.push_back(k);
          b0.push_back(q);
          b1.push_back(std::min(q + SW, ls[l + 1]));
        }
      }
      off.push_back((int)bm.size());
    }
  }
  if (bm.empty()) return 0;
  hipStream_t s = stream();
  MF_HIP(hipMalloc(&B.d_ilu_mem, bm.size() * sizeof(int)));
  MF_HIP(hipMalloc(&B.d_ilu_s0, b0.size() * sizeof(int)));
  MF_HIP(hipMalloc(&B.d_ilu_s1, b1.size() * sizeof(int)));
  MF_HIP(hipMemcpyAsync(B.d_ilu_mem, bm.data(), bm.size() * sizeof(int), hipMemcpyHostToDevice, s));
  MF_HIP(hipMemcpyAsync(B.d_ilu_s0, b0.data(), b0.size() * sizeof(int), hipMemcpyHostToDevice, s));
  MF_HIP(hipMemcpyAsync(B.d_ilu_s1, b1.data(), b1.size() * sizeof(int), hipMemcpyHostToDevice, s));
  MF_HIP(hipStreamSynchronize(s));
  return 0;
}

// the vectors a member's inner iteration reads and updates
static void member_args(const BatchMember& M, SolveArgs& A, int m) {
  Problem& P = *M.P;
  fill_solve_args(P, A, (M.precon == 1 && !M.mixed) ? 1 : 0, m);
  A.ilu = M.precon == 2;
  if (M.mixed) {
    A.op32 = 1;
    A.v32 = 1;
    A.dinv32 = M.precon == 1 ? P.d_dinv32 : nullptr;
    A.b = P.d_work;   // r of the current refinement step
    A.x = P.d_work2;  // its correction d
  } else {
    A.b = P.d_rhs;
    A.x = P.d_x;
  }
}

static int batch_build(Batch& B) {
  const int np = (int)B.mem.size();
  batch_drop_graph(B);
  if (batch_alloc(B, np)) return -2;
  std::vector<SolveArgs> args(np);
  std::vector<IluMember> ilus(np);
  std::vector<RefineArgs> refs(np);
  std::vector<int> nblk(np), bp, br;
  int max_sl = 1;
  B.has_ilu = false;
  B.has_mixed = false;
  for (int k = 0; k < np; k++) {
    Problem& P = *B.mem[k].P;
    SolveArgs& A = args[k];
    member_args(B.mem[k], A, B.m);
    A.stat = B.d_stat + k;
    A.hstat = nullptr;
    refs[k] = RefineArgs{P.d_x, P.d_rhs, B.mem[k].precon > 0 ? P.d_dinv : nullptr};
    B.has_mixed = B.has_mixed || B.mem[k].mixed;
    IluMember& L = ilus[k];
    L = IluMember{};
    if (B.mem[k].precon == 2) {
      B.has_ilu = true;
      L.f = sweep_args(P.ilu->fwd);
      L.b = sweep_args(P.ilu->bwd);
      L.f32 = sweep_args32(P.ilu->fwd);
      L.b32 = sweep_args32(P.ilu->bwd);
    }
    const int nsl = (P.n + 63) / 64;
    const int G = std::max(1, std::min(MAX_BLOCKS, (nsl + B.spb_used - 1) / B.spb_used));
    nblk[k] = G;
    max_sl = std::max(max_sl, (nsl + G - 1) / G + 1);
    for (int r = 0; r < G; r++) {
      bp.push_back(k);
      br.push_back(r);
    }
  }
  B.nblk = (int)bp.size();
  B.lds = (size_t)max_sl * 64 * sizeof(double);
  if (B.lds > 64 * 1024) {
    set_error("batch: a member is too large for the batched operator (> 33M rows)");
    return -1;
  }
  if ((int)bp.size() > B.blk_cap) {
    if (B.d_blk_prob) (void)hipFree(B.d_blk_prob);
    if (B.d_blk_rank) (void)hipFree(B.d_blk_rank);
    B.d_blk_prob = B.d_blk_rank = nullptr;
    B.blk_cap = 0;
    const int cap = (int)bp.size() * 2;
    MF_HIP(hipMalloc(&B.d_blk_prob, (size_t)cap * sizeof(int)));
    MF_HIP(hipMalloc(&B.d_blk_rank, (size_t)cap * sizeof(int)));
    B.blk_cap = cap;
  }
  if (np > 0) {
    hipStream_t s = stream();
    MF_HIP(hipMemcpyAsync(B.d_blk_prob, bp.data(), bp.size() * sizeof(int), hipMemcpyHostToDevice, s));
    MF_HIP(hipMemcpyAsync(B.d_blk_rank, br.data(), br.size() * sizeof(int), hipMemcpyHostToDevice, s));
    MF_HIP(hipMemcpyAsync(B.d_args, args.data(), np * sizeof(SolveArgs), hipMemcpyHostToDevice, s));
    MF_HIP(hipMemcpyAsync(B.d_nblk, nblk.data(), np * sizeof(int), hipMemcpyHostToDevice, s));
    MF_HIP(hipMemcpyAsync(B.d_stat, B.h_stat.data(), np * sizeof(GmresStat), hipMemcpyHostToDevice, s));
    MF_HIP(hipMemcpyAsync(B.d_ilu, ilus.data(), np * sizeof(IluMember), hipMemcpyHostToDevice, s));
    MF_HIP(hipMemcpyAsync(B.d_ref, refs.data(), np * sizeof(RefineArgs), hipMemcpyHostToDevice, s));
    MF_HIP(hipMemcpyAsync(B.d_mix, B.h_mix.data(), np * sizeof(MixedState), hipMemcpyHostToDevice, s));
    MF_HIP(hipStreamSynchronize(s));
  }
  if (batch_build_ilu(B)) return -2;
  B.dirty = false;
  return 0;
}

static void batch_ilu_levels(Batch& B, const BatchArgs& a, int j, hipStream_t s) {
  const IluMember* L = reinterpret_cast<const IluMember*>(B.d_ilu);
  for (size_t l = 0; l + 1 < B.ilu_off_f.size(); l++) {
    const int nb = B.ilu_off_f[l + 1] - B.ilu_off_f[l];
    if (nb > 0)
      hipLaunchKernelGGL(kb_ilu_level<false>, dim3(nb), dim3(SB), 0, s, a, L, B.d_ilu_mem,
                         B.d_ilu_s0, B.d_ilu_s1, B.ilu_off_f[l], j);
  }
  for (size_t l = 0; l + 1 < B.ilu_off_b.size(); l++) {
    const int nb = B.ilu_off_b[l + 1] - B.ilu_off_b[l];
    if (nb > 0)
      hipLaunchKernelGGL(kb_ilu_level<true>, dim3(nb), dim3(SB), 0, s, a, L, B.d_ilu_mem,
                         B.d_ilu_s0, B.d_ilu_s1, B.ilu_off_b[l], j);
  }
}

static void batch_refine_steps(Batch& B, hipStream_t s) {
  const BatchArgs a = batch_args(B);
  const RefineArgs* R = reinterpret_cast<const RefineArgs*>(B.d_ref);
  hipLaunchKernelGGL(kb_refine_axpy, dim3(B.nblk), dim3(SB), 0, s, a, R, B.d_mix);
  hipLaunchKernelGGL(kb_refine_resid, dim3(B.nblk), dim3(SB), 0, s, a, R, B.d_mix);
  hipLaunchKernelGGL(kb_refine_fin, dim3((int)B.mem.size()), dim3(SB), 0, s, a, B.d_mix, B.m);
}

// one restart cycle; with refine, the refinement steps of mixed-precision
// members due before it (first steps) and after it (ended inner solves)
static void batch_enqueue_cycle(Batch& B, hipStream_t s, bool refine = true) {
  const BatchArgs a = batch_args(B);
  const int np = (int)B.mem.size();
  if (refine && B.has_mixed) batch_refine_steps(B, s);
  hipLaunchKernelGGL(kb_residual, dim3(B.nblk), dim3(SB), 0, s, a);
  if (B.has_ilu) {
    batch_ilu_levels(B, a, -1, s);
    hipLaunchKernelGGL(kb_sumsq, dim3(B.nblk), dim3(SB), 0, s, a);
  }
  hipLaunchKernelGGL(kb_restart, dim3(np), dim3(SB), 0, s, a);
  for (int j = 0; j < B.m; j++) {
    hipLaunchKernelGGL(kb_op_dots, dim3(B.nblk), dim3(SB), B.lds, s, a, j);
    if (B.has_ilu) {
      batch_ilu_levels(B, a, j, s);
      hipLaunchKernelGGL(kb_dots, dim3(B.nblk), dim3(SB), 0, s, a, j);
    }
    hipLaunchKernelGGL(kb_fin_dots, dim3(j + 1, np), dim3(SB), 0, s, a, j);
    hipLaunchKernelGGL(kb_update, dim3(B.nblk), dim3(SB), 0, s, a, j);
    hipLaunchKernelGGL(kb_fin_update, dim3(np), dim3(SB), 0, s, a, j);
  }
  hipLaunchKernelGGL(kb_ysolve, dim3(np), dim3(64), 0, s, a);
  hipLaunchKernelGGL(kb_xupdate, dim3(B.nblk), dim3(SB), 0, s, a);
  hipLaunchKernelGGL(kb_advance, dim3((np + 255) / 256), dim3(256), 0, s, a, np);
  if (refine && B.has_mixed) batch_refine_steps(B, s);
}

// ---------------------------------------------------------------------
// Batch tile tuner.  The tile of a batch is the number of 64-row slices one
// block of a member streams: small tiles give many blocks (more parallel
// dot-product partials, more work per finishing pass), large tiles fewer,
// fuller blocks.  Which is fastest depends on how the rows are spread over
// the members, so the tuner times one restart cycle per candidate on the
// batch itself and keeps the fastest.  A trial cycle is undone: every
// member's x is restored from a copy and the GMRES state from the host copy
// taken before the trial, so tuning does not change any iterate.  Choices
// are cached per shape: (rows in half-octaves, members in octaves, graph or
// eager launches), and a shape next to a cached one takes its choice.
// ---------------------------------------------------------------------
struct TileChoice {
  int spb;
  float ms[kBatchTileNum];
};
static std::map<std::tuple<int, int, int>, TileChoice> g_tile_cache;

void batch_tile_cache_clear() { g_tile_cache.clear(); }

int batch_set_tiles(Batch& B, int spb) {
  if (spb < 0 || spb > 64) {
    set_error("batch tiles: slices per block must be in [1,64], or 0 to tune");
    return -1;
  }
  B.spb = spb;
  B.dirty = B.retile = true;
  return 0;
}

static std::tuple<int, int, int> tile_key(const Batch& B, int use_graph) {
  long long rows = 0;
  for (const auto& M : B.mem) rows += M.P->n;
  const int half_oct = (int)std::lround(2.0 * std::log2((double)std::max(1LL, rows)));
  const int oct = (int)std::lround(std::log2((double)std::max<size_t>(1, B.mem.size())));
  return std::make_tuple(half_oct, oct, use_graph ? 1 : 0);
}

// a continuously batched server changes shape every few cycles; a choice
// cached within two half-octaves of rows and two octaves of members (same
// launch mode) is taken instead of new trials
static std::map<std::tuple<int, int, int>, TileChoice>::iterator tile_near(
    const std::tuple<int, int, int>& key) {
  auto best = g_tile_cache.end();
  int bd = 1 << 30;
  for (auto it = g_tile_cache.begin(); it != g_tile_cache.end(); ++it) {
    const auto& k = it->first;
    if (std::get<2>(k) != std::get<2>(key)) continue;
    const int dh = std::abs(std::get<0>(k) - std::get<0>(key));
    const int dm = std::abs(std::get<1>(k) - std::get<1>(key));
    if (dh > 2 || dm > 2) continue;
    if (dh + dm < bd) {
      bd = dh + dm;
      best = it;
    }
  }
  return best;
}

static int batch_capture(Batch& B, hipStream_t s, bool refine = true) {
  MF_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  batch_enqueue_cycle(B, s, refine);
  MF_HIP(hipStreamEndCapture(s, &B.graph));
  MF_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
  return 0;
}

// the iterate a member's cycles update: x, or a mixed member's correction
static double* member_iterate(const BatchMember& M) { return M.mixed ? M.P->d_work2 : M.P->d_x; }

static int batch_tune(Batch& B, int use_graph, TileChoice& out) {
  hipStream_t s = stream();
  const int np = (int)B.mem.size();
  std::vector<long long> off(np + 1, 0);
  for (int k = 0; k < np; k++) off[k + 1] = off[k] + B.mem[k].P->n;
  double* snap = nullptr;
  MF_HIP(hipMalloc(&snap, (size_t)std::max(1LL, off[np]) * sizeof(double)));
  for (int k = 0; k < np; k++)
    MF_HIP(hipMemcpyAsync(snap + off[k], member_iterate(B.mem[k]),
                          (size_t)B.mem[k].P->n * sizeof(double), hipMemcpyDeviceToDevice, s));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  MF_HIP(hipEventCreate(&e0));
  MF_HIP(hipEventCreate(&e1));
  int rc = 0;
  float best = 0.0f;
  out.spb = kBatchTileCands[0];
  // the first candidate runs twice; its first (cold) cycle is not counted
  for (int t = -1; t < kBatchTileNum && rc == 0; t++) {
    B.spb_used = kBatchTileCands[t < 0 ? 0 : t];
    rc = batch_build(B);  // also uploads the pre-trial GMRES state
    if (rc) break;
    // trial cycles run the inner solves only (no refinement steps)
    if (use_graph) {
      if (!B.exec && (rc = batch_capture(B, s, false))) break;
      MF_HIP(hipGraphUpload(B.exec, s));
    }
    MF_HIP(hipEventRecord(e0, s));
    if (use_graph)
      MF_HIP(hipGraphLaunch(B.exec, s));
    else
      batch_enqueue_cycle(B, s, false);
    MF_HIP(hipEventRecord(e1, s));
    MF_HIP(hipGetLastError());
    for (int k = 0; k < np; k++)
      MF_HIP(hipMemcpyAsync(member_iterate(B.mem[k]), snap + off[k],
                            (size_t)B.mem[k].P->n * sizeof(double), hipMemcpyDeviceToDevice, s));
    MF_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    MF_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (t < 0) continue;
    out.ms[t] = ms;
    if (t == 0 || ms < best) {
      best = ms;
      out.spb = B.spb_used;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(snap);
  B.dirty = true;
  return rc;
}


// a member still being solved: an FP64 member whose GMRES has not ended, or
// a mixed member that has not met its refinement test
static bool member_active(const MixedState& M, const GmresStat& st) {
  return M.mixed ? !M.fin : !st.done;
}

// device state to the host copies
static int batch_read_state(Batch& B, hipStream_t s) {
  const int np = (int)B.mem.size();
  MF_HIP(hipMemcpyAsync(B.h_pinned, B.d_stat, np * sizeof(GmresStat), hipMemcpyDeviceToHost, s));
  if (B.has_mixed)
    MF_HIP(hipMemcpyAsync(B.h_mix_pinned, B.d_mix, np * sizeof(MixedState), hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  std::copy(B.h_pinned, B.h_pinned + np, B.h_stat.begin());
  if (B.has_mixed) std::copy(B.h_mix_pinned, B.h_mix_pinned + np, B.h_mix.begin());
  return 0;
}

static hipStream_t batch_stream(const Batch& B) { return B.own ? B.own : stream(); }

// Before cycles: preconditioner data of members re-assembled since they
// joined, the tile choice, the tables; then the batch's own stream (if it
// has one) waits for the work enqueued on the library stream (assembly of
// new members, their preconditioner data).  act: members still active.
static int batch_prologue(Batch& B, int max_cycles, int use_graph, int& act) {
  const int np = (int)B.mem.size();
  if (B.m < 1 || B.m >= MAX_NV) {
    set_error("batch: number of Krylov vectors must be in [1,63]");
    return -1;
  }
  for (auto& M : B.mem)
    if (prepare_member(M)) return -2;
  act = 0;
  for (int k = 0; k < np; k++) act += member_active(B.h_mix[k], B.h_stat[k]);
  if (B.retile) {
    B.retile = false;
    B.spb_tuned = false;
    const int prev = B.spb_used;
    if (B.spb > 0) {
      B.spb_used = B.spb;
    } else if (act > 0 && max_cycles > 0) {
      const auto key = tile_key(B, use_graph);
      auto it = g_tile_cache.find(key);
      if (it == g_tile_cache.end()) it = tile_near(key);
      if (it == g_tile_cache.end()) {
        // the trial cycles time the inner solves: members whose first
        // refinement step is due take it before
        bool fresh = false;
        for (const auto& ms : B.h_mix) fresh = fresh || (ms.mixed && ms.phase == 1);
        if (fresh) {
          if (B.dirty && batch_build(B)) return -2;
          batch_refine_steps(B, stream());
          MF_HIP(hipGetLastError());
          if (batch_read_state(B, stream())) return -2;
        }
        TileChoice tc{};
        const int rc = batch_tune(B, use_graph, tc);
        if (rc) return rc;
        it = g_tile_cache.emplace(key, tc).first;
      }
      B.spb_used = it->second.spb;
      B.spb_tuned = true;
      std::copy(it->second.ms, it->second.ms + kBatchTileNum, B.tile_ms);
    }
    if (B.spb_used != prev) B.dirty = true;
  }
  if (B.dirty) {
    const int rc = batch_build(B);
    if (rc) return rc;
  }
  if (B.own) {
    MF_HIP(hipEventRecord(B.ev_ready, stream()));
    MF_HIP(hipStreamWaitEvent(B.own, B.ev_ready, 0));
  }
  return 0;
}

static int batch_launch_cycle(Batch& B, int use_graph, hipStream_t s) {
  if (use_graph) {
    if (!B.exec && batch_capture(B, s)) return -2;
    MF_HIP(hipGraphLaunch(B.exec, s));
  } else {
    batch_enqueue_cycle(B, s);
  }
  MF_HIP(hipGetLastError());
  return 0;
}

int batch_run(Batch& B, int max_cycles, int use_graph, int* active) {
  if (B.pending) {
    set_error("batch: a cycle is in flight (mf_batch_wait first)");
    return -1;
  }
  const int np = (int)B.mem.size();
  int act = 0;
  int rc = batch_prologue(B, max_cycles, use_graph, act);
  if (rc) return rc;
  const hipStream_t s = batch_stream(B);
  for (int c = 0; c < max_cycles && act > 0; c++) {
    if ((rc = batch_launch_cycle(B, use_graph, s))) return rc;
    if (batch_read_state(B, s)) return -2;
    B.cycles++;
    act = 0;
    for (int k = 0; k < np; k++) act += member_active(B.h_mix[k], B.h_stat[k]);
  }
  if (active) *active = act;
  return 0;
}

// one cycle enqueued with the copy of the state to the host; batch_wait
// completes it
int batch_run_async(Batch& B, int use_graph) {
  if (B.pending) {
    set_error("batch: a cycle is in flight (mf_batch_wait first)");
    return -1;
  }
  const int np = (int)B.mem.size();
  int act = 0;
  int rc = batch_prologue(B, 1, use_graph, act);
  if (rc || act == 0) return rc;
  const hipStream_t s = batch_stream(B);
  if ((rc = batch_launch_cycle(B, use_graph, s))) return rc;
  MF_HIP(hipMemcpyAsync(B.h_pinned, B.d_stat, np * sizeof(GmresStat), hipMemcpyDeviceToHost, s));
  if (B.has_mixed)
    MF_HIP(hipMemcpyAsync(B.h_mix_pinned, B.d_mix, np * sizeof(MixedState), hipMemcpyDeviceToHost, s));
  MF_HIP(hipEventRecord(B.ev_done, s));
  B.pending = true;
  return 0;
}

int batch_wait(Batch& B, int* active) {
  const int np = (int)B.mem.size();
  if (B.pending) {
    MF_HIP(hipEventSynchronize(B.ev_done));
    std::copy(B.h_pinned, B.h_pinned + np, B.h_stat.begin());
    if (B.has_mixed) std::copy(B.h_mix_pinned, B.h_mix_pinned + np, B.h_mix.begin());
    B.cycles++;
    B.pending = false;
  }
  int act = 0;
  for (int k = 0; k < np; k++) act += member_active(B.h_mix[k], B.h_stat[k]);
  if (active) *active = act;
  return 0;
}

int batch_own_stream(Batch& B) {
  if (B.own) return 0;
  MF_HIP(hipStreamCreateWithFlags(&B.own, hipStreamNonBlocking));
  MF_HIP(hipEventCreateWithFlags(&B.ev_ready, hipEventDisableTiming));
  MF_HIP(hipEventCreateWithFlags(&B.ev_done, hipEventDisableTiming));
  return 0;
}

void batch_free(Batch& B) {
  if (B.own) (void)hipStreamSynchronize(B.own);
  (void)hipStreamSynchronize(stream());
  B.pending = false;
  batch_drop_tables(B);
  if (B.d_blk_prob) (void)hipFree(B.d_blk_prob);
  if (B.d_blk_rank) (void)hipFree(B.d_blk_rank);
  B.d_blk_prob = B.d_blk_rank = nullptr;
  for (auto& M : B.mem) M.P->batch_refs--;
  B.mem.clear();
  B.h_stat.clear();
  B.h_mix.clear();
  if (B.ev_ready) (void)hipEventDestroy(B.ev_ready);
  if (B.ev_done) (void)hipEventDestroy(B.ev_done);
  if (B.own) (void)hipStreamDestroy(B.own);
  B.ev_ready = B.ev_done = nullptr;
  B.own = nullptr;
}

}  // namespace mf
