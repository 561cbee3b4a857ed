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

}  // namespace mf
