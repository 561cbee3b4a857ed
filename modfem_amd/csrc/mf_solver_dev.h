// Device helpers shared by the single-problem and the batched Krylov
// kernels: launch constants, reductions, the solver argument block, SELL-64
// row sums and the Givens rotation (drotg semantics of netlib BLAS, as
// called by lss_mkb_core.c:2539-2560).
#pragma once

#include <hip/hip_runtime.h>

#include "mf_device.h"

namespace mf {

constexpr int SB = 256;         // threads per solver block
constexpr int TILE_NNZ = 4096;  // CRS entries staged per tile (32 KB LDS)
constexpr int MAX_BLOCKS = 4096;
constexpr int MAX_NV = 64;

// ---------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------
// blocks are dispatched round-robin over the 8 XCDs; give XCD x the x-th
// contiguous range of logical blocks so its L2 sees neighbouring rows
__device__ __forceinline__ int xcd_block(int bid, int nb) {
  int xcd = bid & 7, k = bid >> 3;
  int per = nb >> 3, rem = nb & 7;
  return xcd * per + min(xcd, rem) + k;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// sum NV per-thread values over the block; results in red[0..nact-1]
template <int NV>
__device__ __forceinline__ void block_sum_multi(const double* acc, int nact, double* red) {
  __shared__ double wsum[SB / 64][NV];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NV; i++) {
    if (i < nact) {
      double v = wave_sum(acc[i]);
      if (lane == 0) wsum[wid][i] = v;
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < nact) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < SB / 64; w++) s += wsum[w][threadIdx.x];
    red[threadIdx.x] = s;
  }
  __syncthreads();
}

struct SolveArgs {
  int n, ntiles, m;
  long long ldv;        // stride of the Krylov vectors
  const int* tile_row;  // CRS-tile path (storage seam)
  const int* row_ptr;
  const int* col;
  const double* val;    // CRS values (tile path) or SELL-64 values
  const float* val32;   // SELL-64 values rounded to FP32 (mixed-precision inner solves)
  int op32;             // 1: the operator kernels read val32
  // 16-bit columns relative to a per-slice base where the slice's columns
  // span less than 65536 rows (null, or sbase[sl] < 0: use scol)
  const uint16_t* scol16;
  const int* sbase;
  const long long* sl_ptr;  // SELL-64 path
  const int* scol;
  const double* dinv;   // null: no preconditioner
  const double* b;
  double* x;
  double* V;            // (m+1) x ldv
  double* H;            // (m+1)^2 | givens 2(m+1) | scale (m+1) | coef (m+1)
  double* partials;
  GmresStat* stat;
  GmresStat* hstat;     // pinned host mirror
  // mixed-precision inner solves: the Krylov basis and the Jacobi inverses
  // are FP32 (Vf is the storage of V read with a stride of ldv floats);
  // sums, dot products, Hessenberg data and the iterate stay FP64
  float* Vf;
  const float* dinv32;
  int v32;
  int ilu;              // batched solves: the member applies ILU(0) factors
};

template <typename VT> __device__ __forceinline__ VT* basis(const SolveArgs& A);
template <> __device__ __forceinline__ double* basis<double>(const SolveArgs& A) { return A.V; }
template <> __device__ __forceinline__ float* basis<float>(const SolveArgs& A) { return A.Vf; }

// v scaled by the Jacobi inverse of row r, when there is one
__device__ __forceinline__ double jacobi(const SolveArgs& A, long long r, double v) {
  if (A.dinv) return v * A.dinv[r];
  if (A.dinv32) return v * (double)A.dinv32[r];
  return v;
}

__device__ __forceinline__ double* Hcol(const SolveArgs& A, int c) { return A.H + (size_t)(A.m + 1) * c; }
__device__ __forceinline__ double* Giv(const SolveArgs& A) { return A.H + (size_t)(A.m + 1) * (A.m + 1); }
__device__ __forceinline__ double* Scale(const SolveArgs& A) { return Giv(A) + 2 * (A.m + 1); }
__device__ __forceinline__ double* Coef(const SolveArgs& A) { return Scale(A) + (A.m + 1); }

// tile range of a logical block
__device__ __forceinline__ void tile_range(const SolveArgs& A, int lb, int& t0, int& t1) {
  t0 = (int)((long long)lb * A.ntiles / gridDim.x);
  t1 = (int)((long long)(lb + 1) * A.ntiles / gridDim.x);
}

// stream one tile's CRS entries into LDS products and return the row sum of
// the calling lane's row (valid when r0 + threadIdx.x < r1)
__device__ __forceinline__ double tile_rows(const SolveArgs& A, const double* x, int r0, int r1,
                                            double* prod) {
  const int k0 = A.row_ptr[r0], k1 = A.row_ptr[r1];
#pragma unroll 4
  for (int k = k0 + threadIdx.x; k < k1; k += SB) prod[k - k0] = A.val[k] * x[A.col[k]];
  __syncthreads();
  double s = 0.0;
  const int r = r0 + threadIdx.x;
  if (r < r1) {
    const int a = A.row_ptr[r] - k0, z = A.row_ptr[r + 1] - k0;
    for (int k = a; k < z; k++) s += prod[k];
  }
  __syncthreads();
  return s;
}

// SELL-64: wavefront = slice of 64 rows, lane = row; entries column-major
// inside the slice, so every load of values and columns is one contiguous
// 512/256-byte access.  Logical block lb owns a contiguous range of slices
// (one contiguous eighth of the matrix per XCD); its 4 wavefronts take the
// range's slices round-robin.
constexpr int SW = SB / 64;  // wavefronts per block

__device__ __forceinline__ void slice_range(int nsl, int lb, int& s0, int& s1) {
  s0 = (int)((long long)lb * nsl / gridDim.x);
  s1 = (int)((long long)(lb + 1) * nsl / gridDim.x);
}

// The slices of one operator block, in order t = 0, 1, ...: XCD x owns the
// x-th contiguous part of the slices and its blocks take steps of SW
// consecutive slices cyclically, so each XCD sweeps one contiguous window of
// the matrix at a time (measured 3% faster than a contiguous range per block,
// which leaves ~2k streams open at scattered addresses).
struct SliceSet {
  int base, step, end;
  __device__ __forceinline__ int slice(int t) const { return base + (t / SW) * step + (t % SW); }
  __device__ __forceinline__ int count() const {
    if (end <= base) return 0;
    const int full = (end - base) / step, rem = (end - base) % step;
    return full * SW + min(rem, SW);
  }
};

__device__ __forceinline__ SliceSet block_slices(int nsl) {
  SliceSet S;
  const int G = gridDim.x;
  const int nx = G < 8 ? G : 8;
  const int x = blockIdx.x % nx, q = blockIdx.x / nx;
  const int nq = G / nx + (x < G % nx ? 1 : 0);
  const int r0 = (int)((long long)x * nsl / nx), r1 = (int)((long long)(x + 1) * nsl / nx);
  S.base = r0 + q * SW;
  S.step = nq * SW;
  S.end = r1;
  return S;
}

// row sum of the calling lane's row: the slice's column indices and values
// are loaded first (CH entries in flight per lane), then the vector gathers
// are issued together, so a chunk costs about two memory round trips instead
// of one per entry.  Every load is unconditional (indices past the slice's
// length are clamped to its last entry and their products dropped by a
// select), so the compiler keeps the loads of a chunk in flight together.
template <typename VT, typename XT, int CH = 12>
__device__ __forceinline__ double sell_row(const SolveArgs& A, const VT* val, const XT* x,
                                           int sl, int lane) {
  const long long base = A.sl_ptr[sl] + lane;
  const int len = (int)((A.sl_ptr[sl + 1] - A.sl_ptr[sl]) >> 6);
  // 16-bit columns where the slice has them (a wave-uniform test)
  const int cb = A.sbase ? A.sbase[sl] : -1;
  double s = 0.0;
  for (int k0 = 0; k0 < len; k0 += CH) {
    int c[CH];
    VT v[CH];
    if (cb >= 0) {
#pragma unroll
      for (int k = 0; k < CH; k++) {
        const long long q = base + 64LL * min(k0 + k, len - 1);
        c[k] = cb + (int)A.scol16[q];
        v[k] = val[q];
      }
    } else {
#pragma unroll
      for (int k = 0; k < CH; k++) {
        const long long q = base + 64LL * min(k0 + k, len - 1);
        c[k] = A.scol[q];
        v[k] = val[q];
      }
    }
    XT xv[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) xv[k] = x[c[k]];
#pragma unroll
    for (int k = 0; k < CH; k++) {
      const double t = (double)v[k] * (double)xv[k];
      s += k0 + k < len ? t : 0.0;
    }
  }
  return s;
}

// row sum with the FP64 or the FP32 copy of the values (accumulated in FP64)
template <typename XT>
__device__ __forceinline__ double sell_row_op(const SolveArgs& A, const XT* x, int sl, int lane) {
  return A.op32 ? sell_row(A, A.val32, x, sl, lane) : sell_row(A, A.val, x, sl, lane);
}

// ILU(0) substitution over SELL-64 packed levels: a wavefront owns a slice
// of 64 rows of one level, lane = row, and the factor entries of the slice
// are read column-major (coalesced).  Row ids and columns are device rows;
// the substitutions run in place on the vector.  FT: factor type, VT:
// vector type; the sums are FP64.
template <typename FT>
struct IluSweepArgs {
  const long long* ptr;  // slice start in the packed entries
  const int* col;        // device row of each entry's column
  const FT* val;
  const int* row;        // device row of each packed row (-1: padding)
  const int* len;        // entries of each packed row
  const FT* diag;        // backward sweep: 1/u_ii per packed row
};

// host: the packed factors of a sweep, FP64 or their FP32 copies
IluSweepArgs<double> sweep_args(const IluData::Sweep& W);
IluSweepArgs<float> sweep_args32(const IluData::Sweep& W);

// sum_k val_k v[col_k] over the L entries of one packed row, loads issued
// in groups of 8 so that a group costs about two memory round trips
template <typename FT, typename VT>
__device__ __forceinline__ double packed_row_sum(const int* col, const FT* val, long long base,
                                                 int L, const VT* v) {
  constexpr int CH = 8;
  double s = 0.0;
  for (int k0 = 0; k0 < L; k0 += CH) {
    int c[CH];
    double w[CH], x[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
      const bool ok = k0 + k < L;
      const long long q = base + 64LL * (ok ? k0 + k : k0);
      c[k] = col[q];
      w[k] = ok ? (double)val[q] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < CH; k++) x[k] = (double)v[c[k]];
#pragma unroll
    for (int k = 0; k < CH; k++) s += w[k] * x[k];
  }
  return s;
}

// forward: v_i <- v_i - sum_{k before i} l_ik v_k (the v_k are already
// substituted: they belong to earlier levels)
template <typename FT, typename VT>
__device__ __forceinline__ void ilu_fwd_row(const IluSweepArgs<FT>& W, int sl, int lane, VT* v) {
  const int pr = sl * 64 + lane;
  const int i = W.row[pr];
  if (i < 0) return;
  v[i] = (VT)((double)v[i] - packed_row_sum(W.col, W.val, W.ptr[sl] + lane, W.len[pr], v));
}

// backward: v_i <- (v_i - sum_{k after i} u_ik v_k) / u_ii
template <typename FT, typename VT>
__device__ __forceinline__ void ilu_bwd_row(const IluSweepArgs<FT>& W, int sl, int lane, VT* v) {
  const int pr = sl * 64 + lane;
  const int i = W.row[pr];
  if (i < 0) return;
  v[i] = (VT)(((double)v[i] - packed_row_sum(W.col, W.val, W.ptr[sl] + lane, W.len[pr], v)) *
              (double)W.diag[pr]);
}

// FP64 data of a mixed-precision batch member's refinement steps
struct RefineArgs {
  double* x;            // the solution
  const double* rhs;
  const double* dinv;   // Jacobi inverses of the measure (null: plain norm)
};

// factors of one batch member, FP64 and FP32 copies
struct IluMember {
  IluSweepArgs<double> f, b;
  IluSweepArgs<float> f32, b32;
};

__device__ inline void drotg_dev(double* a, double* b, double* c, double* s) {
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

__device__ __forceinline__ void rot(double* x, double* y, double c, double s) {
  double t = c * (*x) + s * (*y);
  *y = c * (*y) - s * (*x);
  *x = t;
}


}  // namespace mf
