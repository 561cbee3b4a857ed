// Multicolour Gauss-Seidel sweeps on the device, the BLOCK_GS smoother of
// the reference's MKB solver for 1x1 blocks (lsd_mkb/lad_crs/
// las_crs_intf.c:857-1013, used through lsr_mkb_core_smooth,
// lss_mkb_core.c:1549-1773):
//   v_r = d_r (b_r - sum_{j != r} a_rj v_j),  d_r = 1/a_rr,
// each row taking the newest values of the rows updated before it.  The
// reference updates rows in DOF order within a thread's chunk (its result
// depends on the thread count); here the rows of one colour of a greedy
// colouring are independent and are updated together, colour by colour,
// which is GS in the multicolour order.  Values are copied from the SELL
// matrix into per-colour SELL-64 slices so the sweeps read them coalesced.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "mf_device.h"

namespace mf {

__global__ void __launch_bounds__(256) k_gs_color(int s0, int s1, const long long* ptr,
                                                  const int* col, const double* val,
                                                  const int* row, const int* len,
                                                  const double* dinv, const double* b,
                                                  double* v, const int* stop) {
  if (stop && *stop) return;
  const int sl = s0 + blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (sl >= s1) return;
  const int pr = sl * 64 + lane;
  const int i = row[pr];
  if (i < 0) return;
  const long long base = ptr[sl] + lane;
  const int L = len[pr];
  constexpr int CH = 8;
  double vloc = 0.0;
  for (int k0 = 0; k0 < L; k0 += CH) {
    int c[CH];
    double a[CH], x[CH];
#pragma unroll
    for (int k = 0; k < CH; k++) {
      const bool ok = k0 + k < L;
      const long long q = base + 64LL * (ok ? k0 + k : k0);
      c[k] = col[q];
      a[k] = ok ? val[q] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < CH; k++) x[k] = v[c[k]];
#pragma unroll
    for (int k = 0; k < CH; k++) vloc -= a[k] * x[k];
  }
  if (b) vloc += b[i];
  v[i] = vloc * dinv[i];
}

__global__ void k_gs_pack(long long nent, const long long* src, const double* sell, double* val) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nent) val[q] = src[q] >= 0 ? sell[src[q]] : 0.0;
}

template <typename T>
static int up(T** d, const std::vector<T>& h) {
  MF_HIP(hipMalloc(d, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) MF_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

void gs_free(Problem& P) {
  if (!P.gs) return;
  GsData& G = *P.gs;
  void* ptrs[] = {G.d_ptr, G.d_col, G.d_src, G.d_val, G.d_row, G.d_len};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  delete P.gs;
  P.gs = nullptr;
}

int gs_build(Problem& P) {
  if (!P.sell || P.setup.perm.empty()) {
    set_error("Gauss-Seidel: needs a SELL problem with its set-up");
    return -1;
  }
  const Setup& S = P.setup;
  const int n = S.n;
  std::vector<int> color;
  const int nc = greedy_colors(S, color);
  std::vector<int> byc(n);
  for (int i = 0; i < n; i++) byc[i] = i;
  std::stable_sort(byc.begin(), byc.end(), [&](int a, int b) { return color[a] < color[b]; });
  gs_free(P);
  P.gs = new GsData();
  GsData& G = *P.gs;
  G.n = n;
  G.ncolors = nc;
  G.elim_pos.assign(n, 0);
  for (int k = 0; k < n; k++) G.elim_pos[byc[k]] = k;
  std::vector<long long> ptr(1, 0), src;
  std::vector<int> prow, plen, col;
  G.col_sl.assign(nc + 1, 0);
  int k = 0;
  for (int c = 0; c < nc; c++) {
    G.col_sl[c] = (int)ptr.size() - 1;
    int end = k;
    while (end < n && color[byc[end]] == c) end++;
    for (int a = k; a < end; a += 64) {
      int mx = 0;
      for (int q = 0; q < 64; q++) {
        const int i = a + q < end ? byc[a + q] : -1;
        const int len = i < 0 ? 0 : S.row_ptr[S.perm[i] + 1] - S.row_ptr[S.perm[i]] - 1;
        prow.push_back(i);
        plen.push_back(len);
        mx = std::max(mx, len);
      }
      ptr.push_back(ptr.back() + 64LL * mx);
    }
    k = end;
  }
  G.col_sl[nc] = (int)ptr.size() - 1;
  const long long nent = ptr.back();
  col.assign(nent, 0);
  src.assign(nent, -1);
  const int nsl = (int)ptr.size() - 1;
  for (int sl = 0; sl < nsl; sl++)
    for (int q = 0; q < 64; q++) {
      const int i = prow[(size_t)sl * 64 + q];
      if (i < 0) continue;
      const int r = S.perm[i];
      const long long base = S.sl_ptr[i >> 6] + (i & 63);
      int e2 = 0;
      for (int e = S.row_ptr[r]; e < S.row_ptr[r + 1]; e++) {
        if (S.col_ind[e] == r) continue;  // diagonal
        col[ptr[sl] + 64LL * e2 + q] = S.iperm[S.col_ind[e]];
        src[ptr[sl] + 64LL * e2 + q] = base + 64LL * (e - S.row_ptr[r]);
        e2++;
      }
    }
  G.nent = nent;
  G.nrows = (int)prow.size();
  if (up(&G.d_ptr, ptr) || up(&G.d_col, col) || up(&G.d_src, src) || up(&G.d_row, prow) ||
      up(&G.d_len, plen))
    return -2;
  MF_HIP(hipMalloc(&G.d_val, std::max<long long>(nent, 1) * sizeof(double)));
  G.valid = false;
  return 0;
}

int gs_refresh(Problem& P) {
  GsData& G = *P.gs;
  if (G.nent > 0)
    hipLaunchKernelGGL(k_gs_pack, dim3((unsigned)((G.nent + 255) / 256)), dim3(256), 0, stream(),
                       G.nent, G.d_src, P.d_val, G.d_val);
  MF_HIP(hipGetLastError());
  G.valid = true;
  return 0;
}

int gs_sweeps(Problem& P, int nsweeps, const double* b, double* v, const int* stop) {
  GsData& G = *P.gs;
  for (int it = 0; it < nsweeps; it++)
    for (int c = 0; c < G.ncolors; c++) {
      const int s0 = G.col_sl[c], s1 = G.col_sl[c + 1];
      hipLaunchKernelGGL(k_gs_color, dim3((s1 - s0 + 3) / 4), dim3(256), 0, stream(), s0, s1,
                         G.d_ptr, G.d_col, G.d_val, G.d_row, G.d_len, P.d_dinv, b, v, stop);
    }
  MF_HIP(hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------
// BLOCK_GS of the CRS storage as the reference computes it on T OpenMP
// threads (lar_perform_BJ_or_GS_iterations_crs, las_crs_intf.c:905-998).
// The rows are cut into chunks of portion = n/T + 1; within a chunk the rows
// are updated in order, and an entry left of the diagonal (CRS position
// k < diag_ptr) whose column lies in [chunk start, row) takes the chunk's new
// value, every other entry the sweep's start value.  The result depends on T
// and equals the reference's run on T threads bit for bit (T = 1: forward
// Gauss-Seidel in DOF order).
// One wave per chunk walks its rows: the lanes form the products a_k x_k, the
// row sum is taken in CRS position order exactly as "vloc -= a*x" runs, then
// "+= b", "*= d".  The chunk's newest kGsRing values are kept in LDS; older
// ones are read back from vnew (written by this workgroup before a barrier).
namespace {

constexpr int kGsRing = 4096;

__global__ void __launch_bounds__(64) k_crs_gs_chunks(int n, int portion, const int* __restrict__ rp,
                                                      const int* __restrict__ col,
                                                      const double* __restrict__ val,
                                                      const int* __restrict__ dp,
                                                      const double* __restrict__ dinv,
                                                      const double* b, const double* vold,
                                                      double* vnew, const int* stop) {
  if (stop && *stop) return;
  __shared__ double ring[kGsRing];
  __shared__ double prod[64];
  const int lane = threadIdx.x;
  const long long f = (long long)blockIdx.x * portion;
  if (f >= n) return;
  const int first = (int)f;
  const int last = (int)std::min<long long>(n, f + portion);
  for (int r = first; r < last; r++) {
    const int k0 = rp[r], k1 = rp[r + 1], dk = dp[r];
    double s = 0.0;
    for (int base = k0; base < k1; base += 64) {
      const int k = base + lane;
      double p = 0.0;
      if (k < k1 && k != dk) {
        const int c = col[k];
        double x;
        if (k < dk && c >= first && c < r)
          x = (r - c < kGsRing) ? ring[c & (kGsRing - 1)] : vnew[c];
        else
          x = vold[c];
        p = __dmul_rn(val[k], x);
      }
      prod[lane] = p;
      __syncthreads();
      const int cnt = std::min(64, k1 - base);
      for (int j = 0; j < cnt; j++)
        if (base + j != dk) s = __dsub_rn(s, prod[j]);
      __syncthreads();
    }
    if (b) s = __dadd_rn(s, b[r]);
    s = __dmul_rn(s, dinv[r]);
    if (lane == 0) {
      vnew[r] = s;
      ring[r & (kGsRing - 1)] = s;
    }
    __syncthreads();
  }
}

__global__ void k_gs_vcopy(long long n, const double* x, const double* s, int j, double* y,
                           const int* stop) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (stop && *stop)) return;
  y[i] = s ? __dmul_rn(s[j], x[i]) : x[i];
}

// y = (neg ? -1 : 1) (a - c)
__global__ void k_gs_vsub(long long n, const double* a, const double* c, double* y, int neg,
                          const int* stop) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || (stop && *stop)) return;
  const double d = __dsub_rn(a[i], c[i]);
  y[i] = neg ? -d : d;
}

}  // namespace

// nr_prec sweeps over v (device), b null: no rhs; tmp: n doubles
int crs_gs_chunked(Problem& P, int nthreads, int nr_prec, const double* b, double* v, double* tmp,
                   const int* stop) {
  if (P.n == 0 || nr_prec < 1) return 0;
  if (nthreads < 1 || nthreads > kMaxGsChunks) {
    set_error("CRS Gauss-Seidel: the thread count must be in [1, " +
              std::to_string(kMaxGsChunks) + "]");
    return -1;
  }
  if (!P.d_dinv || !P.d_diag_ptr) {
    set_error("CRS Gauss-Seidel: the preconditioner is not filled");
    return -1;
  }
  const int portion = P.n / nthreads + 1;
  const int grid = (int)std::min<long long>(nthreads, ((long long)P.n + portion - 1) / portion);
  const unsigned gv = (unsigned)((P.n + 255) / 256);
  for (int it = 0; it < nr_prec; it++) {
    hipLaunchKernelGGL(k_crs_gs_chunks, dim3(grid), dim3(64), 0, stream(), P.n, portion,
                       P.d_row_ptr, P.d_col, P.d_val, P.d_diag_ptr, P.d_dinv, b, v, tmp, stop);
    hipLaunchKernelGGL(k_gs_vcopy, dim3(gv), dim3(256), 0, stream(), (long long)P.n, tmp, nullptr,
                       0, v, stop);
  }
  MF_HIP(hipGetLastError());
  return 0;
}

namespace {

// GMRES operator of a CRS matrix preconditioned by chunked Gauss-Seidel, as
// lsr_mkb_core_compreres applies a smoother (lss_mkb_core.c:1861-2097):
// V = x; nr_prec sweeps on V with (Control = 1) or without the rhs; V -= x;
// V = -V when Control = 0
struct CrsGsOp : ExtOp {
  Problem* P = nullptr;
  int nthreads = 1, nr_prec = 1;
  double *t1 = nullptr, *t2 = nullptr, *xs = nullptr;
  ~CrsGsOp() override {
    for (double* p : {t1, t2, xs})
      if (p) (void)hipFree(p);
  }
  unsigned gv() const { return (unsigned)((P->n + 255) / 256); }
  int precon_residual(const double* b, const double* x, double* out, const int* stop, int neg) {
    hipLaunchKernelGGL(k_gs_vcopy, dim3(gv()), dim3(256), 0, stream(), (long long)P->n, x,
                       nullptr, 0, t2, stop);
    if (crs_gs_chunked(*P, nthreads, nr_prec, b, t2, t1, stop)) return -2;
    hipLaunchKernelGGL(k_gs_vsub, dim3(gv()), dim3(256), 0, stream(), (long long)P->n, t2, x, out,
                       neg, stop);
    MF_HIP(hipGetLastError());
    return 0;
  }
  int residual(const double* b, const double* x, double* r) override {
    return precon_residual(b, x, r, nullptr, 0);
  }
  int apply(const double* v, const double* scale, int j, double* w, const int* stop) override {
    hipLaunchKernelGGL(k_gs_vcopy, dim3(gv()), dim3(256), 0, stream(), (long long)P->n, v, scale,
                       j, xs, stop);
    return precon_residual(nullptr, xs, w, stop, 1);
  }
};

}  // namespace

ExtOp* crs_gs_make_op(Problem* P, int nthreads, int nr_prec) {
  CrsGsOp* op = new CrsGsOp();
  op->P = P;
  op->nthreads = nthreads;
  op->nr_prec = nr_prec;
  const size_t bytes = (size_t)std::max(P->n, 1) * sizeof(double);
  if (hipMalloc(&op->t1, bytes) != hipSuccess || hipMalloc(&op->t2, bytes) != hipSuccess ||
      hipMalloc(&op->xs, bytes) != hipSuccess) {
    set_error("CRS Gauss-Seidel operator: out of device memory");
    delete op;
    return nullptr;
  }
  return op;
}

}  // namespace mf
