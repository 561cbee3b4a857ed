// ILU(0) preconditioner on the device, as the reference's block ILU(0) does
// it for 1x1 blocks (lsd_mkb/lad_block/las_block_intf.c:1409-2143 factor,
// :3082-3490 forward/backward substitution; used by compreres' BLOCK_ILU
// branch, lsd_mkb/lsd_mkb_core/lss_mkb_core.c:1947-1997):
//   for each row i, lower k ascending: a_ik *= d_k (d_k = 1/u_kk), then
//   a_ij -= a_ik a_kj for j > k in both rows; finally d_i = 1/u_ii.
// The elimination order is either the reference's DOF order (factors equal
// the reference's) or a multicolour order (greedy colouring of the matrix
// graph), which has as many forward levels as colours.  Rows of one level
// depend only on rows of earlier levels, so each level is one launch with a
// lane per row; the factorization uses the forward levels, the backward
// substitution its own levels.  The substitutions read copies of the
// factors packed per level in SELL-64 slices, so their loads coalesce, and
// run in place on the vector in device-row numbering: a row's gathers then
// touch its mesh neighbours, which the device order keeps close in memory
// (rows of one colour are spread over the whole mesh).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "mf_device.h"
#include "mf_solver_dev.h"

namespace mf {

__global__ void k_ilu_load(long long nnz, const long long* src, const double* val, double* lu) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) lu[e] = val[src[e]];
}

__global__ void k_ilu_factor_level(const int* rows, int count, const int* rp, const int* ci,
                                   const int* dg, double* lu) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int i = rows[t];
  const int e_end = rp[i + 1];
  for (int p = rp[i]; p < dg[i]; p++) {
    const int k = ci[p];
    const double aik = lu[p] * lu[dg[k]];
    lu[p] = aik;
    int r = p + 1, q = dg[k] + 1;
    const int q_end = rp[k + 1];
    while (r < e_end && q < q_end) {
      const int cr = ci[r], cq = ci[q];
      if (cr == cq) {
        lu[r] -= aik * lu[q];
        r++;
        q++;
      } else if (cr < cq) {
        r++;
      } else {
        q++;
      }
    }
  }
  lu[dg[i]] = 1.0 / lu[dg[i]];
}

// substitution over SELL-packed levels: a wavefront owns a slice of 64
// rows of the level, lane = row (ilu_fwd_row / ilu_bwd_row)
template <bool BWD, typename FT, typename VT>
__global__ void __launch_bounds__(256) k_ilu_level_sell(int s0, int s1, IluSweepArgs<FT> W, VT* v,
                                                        const int* stop) {
  if (stop && *stop) return;
  const int sl = s0 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (sl >= s1) return;
  if (BWD)
    ilu_bwd_row(W, sl, threadIdx.x & 63, v);
  else
    ilu_fwd_row(W, sl, threadIdx.x & 63, v);
}

__global__ void k_ilu_pack(long long nent, const int* src, const double* lu, double* val) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nent) val[q] = src[q] >= 0 ? lu[src[q]] : 0.0;
}

__global__ void k_ilu_pack_diag(int nrows, const int* dpos, const double* lu, double* diag) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nrows) diag[q] = dpos[q] >= 0 ? lu[dpos[q]] : 0.0;
}

static void free_sweep(IluData::Sweep& W) {
  void* ptrs[] = {W.d_ptr, W.d_col, W.d_src, W.d_val, W.d_row, W.d_len, W.d_diag, W.d_dpos,
                  W.d_val32, W.d_diag32};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  W = IluData::Sweep();
}

void ilu_free(Problem& P) {
  if (!P.ilu) return;
  IluData& L = *P.ilu;
  void* ptrs[] = {L.d_rp, L.d_ci, L.d_dg, L.d_lev_l, L.d_lev_u, L.d_src, L.d_lu};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  free_sweep(L.fwd);
  free_sweep(L.bwd);
  delete P.ilu;
  P.ilu = nullptr;
}

template <typename T>
static int up(T** d, const std::vector<T>& h) {
  MF_HIP(hipMalloc(d, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty()) MF_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// rows grouped by level; lev[] per row
static void group_levels(const std::vector<int>& lev, int nlev, std::vector<int>& ptr,
                         std::vector<int>& rows) {
  ptr.assign(nlev + 1, 0);
  for (int l : lev) ptr[l + 1]++;
  for (int k = 0; k < nlev; k++) ptr[k + 1] += ptr[k];
  rows.resize(lev.size());
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (int i = 0; i < (int)lev.size(); i++) rows[fill[lev[i]]++] = i;
}

// greedy colouring of the (symmetric) pattern, rows visited in device order
int greedy_colors(const Setup& S, std::vector<int>& color) {
  const int n = S.n;
  int ncolors = 0;
  color.assign(n, -1);
  std::vector<int> mark;
  for (int i = 0; i < n; i++) {
    const int r = S.perm[i];
    mark.assign(ncolors + 1, 0);
    for (int e = S.row_ptr[r]; e < S.row_ptr[r + 1]; e++) {
      const int c = color[S.iperm[S.col_ind[e]]];
      if (c >= 0) mark[c] = 1;
    }
    int c = 0;
    while (c < ncolors && mark[c]) c++;
    color[i] = c;
    ncolors = std::max(ncolors, c + 1);
  }
  return ncolors;
}

// pack the rows of every level into SELL-64 slices: lower (fwd) or upper
// (bwd) entries of each row, in ascending elimination order of the columns.
// The packed row ids and column indices are device rows (ord maps an
// elimination position to its device row).
static int pack_sweep(IluData::Sweep& W, int nlev, const std::vector<int>& lev_ptr,
                      const std::vector<int>& rows, const std::vector<int>& rp,
                      const std::vector<int>& ci, const std::vector<int>& dg,
                      const std::vector<int>& ord, bool upper) {
  std::vector<long long> ptr(1, 0);
  std::vector<int> prow, plen, pdiag;
  W.lev_sl.assign(nlev + 1, 0);
  for (int l = 0; l < nlev; l++) {
    W.lev_sl[l] = (int)ptr.size() - 1;
    for (int a = lev_ptr[l]; a < lev_ptr[l + 1]; a += 64) {
      int mx = 0;
      for (int q = 0; q < 64; q++) {
        const int k = a + q < lev_ptr[l + 1] ? rows[a + q] : -1;
        const int len = k < 0 ? 0 : (upper ? rp[k + 1] - dg[k] - 1 : dg[k] - rp[k]);
        prow.push_back(k);
        plen.push_back(len);
        pdiag.push_back(k < 0 ? -1 : dg[k]);
        mx = std::max(mx, len);
      }
      ptr.push_back(ptr.back() + 64LL * mx);
    }
  }
  W.lev_sl[nlev] = (int)ptr.size() - 1;
  const long long nent = ptr.back();
  std::vector<int> col(nent, 0), src(nent, -1);
  const int nsl = (int)ptr.size() - 1;
  for (int sl = 0; sl < nsl; sl++)
    for (int q = 0; q < 64; q++) {
      const int k = prow[(size_t)sl * 64 + q];
      if (k < 0) continue;
      const int first = upper ? dg[k] + 1 : rp[k];
      for (int e = 0; e < plen[(size_t)sl * 64 + q]; e++) {
        col[ptr[sl] + 64LL * e + q] = ci[first + e];
        src[ptr[sl] + 64LL * e + q] = first + e;
      }
    }
  W.nent = nent;
  W.nrows = (int)prow.size();
  for (auto& c : col) c = ord[c];
  for (auto& k : prow)
    if (k >= 0) k = ord[k];
  if (up(&W.d_ptr, ptr) || up(&W.d_col, col) || up(&W.d_src, src) || up(&W.d_row, prow) ||
      up(&W.d_len, plen) || (upper && up(&W.d_dpos, pdiag)))
    return -2;
  MF_HIP(hipMalloc(&W.d_val, std::max<long long>(nent, 1) * sizeof(double)));
  if (upper) MF_HIP(hipMalloc(&W.d_diag, std::max(W.nrows, 1) * sizeof(double)));
  return 0;
}

int ilu_build(Problem& P, int order) {
  if (!P.sell || P.setup.perm.empty()) {
    set_error("ILU(0): needs a SELL problem with its set-up");
    return -1;
  }
  if (order != 0 && order != 1) {
    set_error("ILU(0): order must be 0 (reference DOF order) or 1 (multicolour)");
    return -1;
  }
  const Setup& S = P.setup;
  const int n = S.n;
  // elimination order: ord[k] = device row of ILU position k, pos[device row]
  std::vector<int> ord(n), pos(n);
  int ncolors = 0;
  if (order == 0) {
    for (int k = 0; k < n; k++) ord[k] = S.iperm[k];  // reference row k
  } else {
    std::vector<int> color;
    ncolors = greedy_colors(S, color);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return color[a] < color[b]; });
  }
  for (int k = 0; k < n; k++) pos[ord[k]] = k;
  // pattern in ILU positions, columns ascending, with the SELL position of
  // every entry
  std::vector<int> rp(n + 1, 0), ci(S.nnz), dg(n, -1);
  std::vector<long long> src(S.nnz), refp(S.nnz);
  for (int k = 0; k < n; k++) {
    const int r = S.perm[ord[k]];
    rp[k + 1] = rp[k] + (S.row_ptr[r + 1] - S.row_ptr[r]);
  }
  std::vector<std::pair<int, int>> buf;  // (column position, reference entry)
  for (int k = 0; k < n; k++) {
    const int i = ord[k], r = S.perm[i];
    const long long base = S.sl_ptr[i >> 6] + (i & 63);
    buf.clear();
    for (int e = S.row_ptr[r]; e < S.row_ptr[r + 1]; e++)
      buf.push_back({pos[S.iperm[S.col_ind[e]]], e});
    std::sort(buf.begin(), buf.end());
    for (size_t q = 0; q < buf.size(); q++) {
      const int e = buf[q].second;
      ci[rp[k] + q] = buf[q].first;
      src[rp[k] + q] = base + 64LL * (e - S.row_ptr[r]);
      refp[rp[k] + q] = e;
      if (buf[q].first == k) dg[k] = rp[k] + (int)q;
    }
    if (dg[k] < 0) {
      set_error("ILU(0): a row has no diagonal entry");
      return -1;
    }
  }
  // forward levels (lower dependencies) and backward levels (upper)
  std::vector<int> levl(n, 0), levu(n, 0);
  int nl = 0, nu = 0;
  for (int k = 0; k < n; k++) {
    int l = 0;
    for (int p = rp[k]; p < dg[k]; p++) l = std::max(l, levl[ci[p]] + 1);
    levl[k] = l;
    nl = std::max(nl, l + 1);
  }
  for (int k = n - 1; k >= 0; k--) {
    int l = 0;
    for (int p = dg[k] + 1; p < rp[k + 1]; p++) l = std::max(l, levu[ci[p]] + 1);
    levu[k] = l;
    nu = std::max(nu, l + 1);
  }
  std::vector<int> rows_l, rows_u;
  ilu_free(P);
  P.ilu = new IluData();
  IluData& L = *P.ilu;
  L.order = order;
  L.n = n;
  L.nnz = S.nnz;
  L.ncolors = ncolors;
  L.nlev_l = nl;
  L.nlev_u = nu;
  L.ref_pos = refp;
  L.pos_of_dev = pos;
  group_levels(levl, nl, L.lev_ptr_l, rows_l);
  group_levels(levu, nu, L.lev_ptr_u, rows_u);
  if (up(&L.d_rp, rp) || up(&L.d_ci, ci) || up(&L.d_dg, dg) || up(&L.d_lev_l, rows_l) ||
      up(&L.d_lev_u, rows_u) || up(&L.d_src, src))
    return -2;
  MF_HIP(hipMalloc(&L.d_lu, std::max<long long>(S.nnz, 1) * sizeof(double)));
  if (pack_sweep(L.fwd, nl, L.lev_ptr_l, rows_l, rp, ci, dg, ord, false) ||
      pack_sweep(L.bwd, nu, L.lev_ptr_u, rows_u, rp, ci, dg, ord, true))
    return -2;
  L.valid = false;
  return 0;
}

int ilu_factor(Problem& P) {
  if (!P.ilu) {
    set_error("ILU(0): structure not built");
    return -1;
  }
  IluData& L = *P.ilu;
  hipStream_t s = stream();
  hipLaunchKernelGGL(k_ilu_load, dim3((unsigned)((L.nnz + 255) / 256)), dim3(256), 0, s, L.nnz,
                     L.d_src, P.d_val, L.d_lu);
  for (int l = 0; l < L.nlev_l; l++) {
    const int c = L.lev_ptr_l[l + 1] - L.lev_ptr_l[l];
    hipLaunchKernelGGL(k_ilu_factor_level, dim3((c + 255) / 256), dim3(256), 0, s,
                       L.d_lev_l + L.lev_ptr_l[l], c, L.d_rp, L.d_ci, L.d_dg, L.d_lu);
  }
  for (IluData::Sweep* W : {&L.fwd, &L.bwd})
    if (W->nent > 0)
      hipLaunchKernelGGL(k_ilu_pack, dim3((unsigned)((W->nent + 255) / 256)), dim3(256), 0, s,
                         W->nent, W->d_src, L.d_lu, W->d_val);
  hipLaunchKernelGGL(k_ilu_pack_diag, dim3((L.bwd.nrows + 255) / 256), dim3(256), 0, s,
                     L.bwd.nrows, L.bwd.d_dpos, L.d_lu, L.bwd.d_diag);
  MF_HIP(hipGetLastError());
  L.valid = true;
  L.valid32 = false;
  return 0;
}

__global__ void k_ilu_to_f32(long long n, const double* a, float* b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

int ilu_round_f32(Problem& P) {
  if (!P.ilu || !P.ilu->valid) {
    set_error("ILU(0): no factors to round");
    return -1;
  }
  IluData& L = *P.ilu;
  hipStream_t s = stream();
  for (IluData::Sweep* W : {&L.fwd, &L.bwd}) {
    if (!W->d_val32) MF_HIP(hipMalloc(&W->d_val32, std::max<long long>(W->nent, 1) * sizeof(float)));
    if (W->nent > 0)
      hipLaunchKernelGGL(k_ilu_to_f32, dim3((unsigned)((W->nent + 255) / 256)), dim3(256), 0, s,
                         W->nent, W->d_val, W->d_val32);
  }
  if (!L.bwd.d_diag32) MF_HIP(hipMalloc(&L.bwd.d_diag32, std::max(L.bwd.nrows, 1) * sizeof(float)));
  hipLaunchKernelGGL(k_ilu_to_f32, dim3((L.bwd.nrows + 255) / 256), dim3(256), 0, s,
                     (long long)L.bwd.nrows, L.bwd.d_diag, L.bwd.d_diag32);
  MF_HIP(hipGetLastError());
  L.valid32 = true;
  return 0;
}

IluSweepArgs<double> sweep_args(const IluData::Sweep& W) {
  return IluSweepArgs<double>{W.d_ptr, W.d_col, W.d_val, W.d_row, W.d_len, W.d_diag};
}
IluSweepArgs<float> sweep_args32(const IluData::Sweep& W) {
  return IluSweepArgs<float>{W.d_ptr, W.d_col, W.d_val32, W.d_row, W.d_len, W.d_diag32};
}

// out = (LU)^-1 in: in is copied to out (unless they are the same vector),
// then both substitutions run in place on out
template <typename FT, typename VT>
static int ilu_sweeps(const Problem& P, IluSweepArgs<FT> F, IluSweepArgs<FT> U, const VT* in,
                      VT* out, const int* stop) {
  const IluData& L = *P.ilu;
  hipStream_t s = stream();
  if (in != out)
    MF_HIP(hipMemcpyAsync(out, in, (size_t)P.n * sizeof(VT), hipMemcpyDeviceToDevice, s));
  for (int l = 0; l < L.nlev_l; l++) {
    const int s0 = L.fwd.lev_sl[l], s1 = L.fwd.lev_sl[l + 1];
    hipLaunchKernelGGL((k_ilu_level_sell<false, FT, VT>), dim3((s1 - s0 + 3) / 4), dim3(256), 0, s,
                       s0, s1, F, out, stop);
  }
  for (int l = 0; l < L.nlev_u; l++) {
    const int s0 = L.bwd.lev_sl[l], s1 = L.bwd.lev_sl[l + 1];
    hipLaunchKernelGGL((k_ilu_level_sell<true, FT, VT>), dim3((s1 - s0 + 3) / 4), dim3(256), 0, s,
                       s0, s1, U, out, stop);
  }
  MF_HIP(hipGetLastError());
  return 0;
}

int ilu_apply(Problem& P, const double* in, double* out, const int* stop) {
  IluData& L = *P.ilu;
  return ilu_sweeps(P, sweep_args(L.fwd), sweep_args(L.bwd), in, out, stop);
}

int ilu_apply32(Problem& P, const float* in, float* out, const int* stop) {
  IluData& L = *P.ilu;
  if (!L.valid32) {
    set_error("ILU(0): FP32 factors are not current");
    return -1;
  }
  return ilu_sweeps(P, sweep_args32(L.fwd), sweep_args32(L.bwd), in, out, stop);
}

}  // namespace mf
