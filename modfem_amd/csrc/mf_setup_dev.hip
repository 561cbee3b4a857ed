// Set-up of a std-linear problem on the device (no reference counterpart:
// the reference builds its DOF list, neighbour sets and assembly tables on
// the host: utd_util/uts_ls_intf.c:229-398, sid_mkb/sis_mkb_intf.c:904-1142,
// lsd_mkb/lad_crs/las_crs_intf.c:72-432).  Every map is bitwise the host
// set-up's (mf_setup.cpp).  The orders the host builds with sequential
// scans come from radix sorts of keys that encode the same order:
//   DOF list      nodes by their first position 6 e + i over the elements
//                 (first appearance, uts_ls_intf.c:258-308);
//   incidence     (DOF, element) pairs, i.e. elements ascending per DOF;
//   CRS rows      unique (row, column) pairs of every element's DOFs, i.e.
//                 sorted unions of the neighbour DOFs (sis_mkb_intf.c:904-1142);
//   device order  elements by the Morton key of their centroid, ties by
//                 element (the sort is stable), then DOFs by first position
//                 over the sorted elements;
// and the tables follow by binary search in the rows, as on the host.  The
// host keeps the maps its own calls read (pattern, permutations, element
// tables), downloaded once.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <vector>

#include "mf_device.h"

namespace mf {
namespace {

constexpr int TPB = 256;

unsigned grid_of(long long n) { return (unsigned)std::max<long long>(1, (n + TPB - 1) / TPB); }

// device allocation released on every exit path
template <class T>
struct DevArr {
  T* p = nullptr;
  DevArr() = default;
  DevArr(const DevArr&) = delete;
  DevArr& operator=(const DevArr&) = delete;
  ~DevArr() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n) {
    MF_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return 0;
  }
  T* release() {
    T* q = p;
    p = nullptr;
    return q;
  }
};

// hipcub's two-call idiom with a scratch buffer that grows as needed
struct Scratch {
  void* p = nullptr;
  size_t n = 0;
  ~Scratch() {
    if (p) (void)hipFree(p);
  }
  int ensure(size_t bytes) {
    if (bytes <= n) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    MF_HIP(hipMalloc(&p, std::max<size_t>(bytes, 1)));
    n = bytes;
    return 0;
  }
};

template <class K, class V>
int sort_pairs(Scratch& S, const K* kin, K* kout, const V* vin, V* vout, long long n, int end_bit,
               hipStream_t s) {
  size_t bytes = 0;
  MF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0,
                                            end_bit, s));
  if (S.ensure(bytes)) return -2;
  MF_HIP(hipcub::DeviceRadixSort::SortPairs(S.p, bytes, kin, kout, vin, vout, (int)n, 0, end_bit,
                                            s));
  return 0;
}

int bits_of(unsigned long long v) {
  int b = 1;
  while (b < 64 && (v >> b)) b++;
  return b;
}

// ---------------------------------------------------------------- kernels
__global__ void k_fill_u32(unsigned* a, long long n, unsigned v) {
  const long long i = (long long)blockIdx.x * TPB + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void k_iota(int* a, long long n) {
  const long long i = (long long)blockIdx.x * TPB + threadIdx.x;
  if (i < n) a[i] = (int)i;
}

// first position 6 e + i at which node v appears
__global__ void k_first_node(int nel, int K, const int* nodes, unsigned* first) {
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= (long long)nel * K) return;
  const int e = (int)(t / K), i = (int)(t % K);
  atomicMin(&first[nodes[6 * (size_t)e + i]], (unsigned)(6 * e + i));
}

// nodes in first-appearance order -> DOF numbers; *n = DOFs
__global__ void k_number_dofs(int nno, const unsigned* key, const int* node, int* node_dof,
                              int* dof_node, int* n) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= nno) return;
  if (key[i] == 0xffffffffu) return;
  node_dof[node[i]] = i;
  dof_node[i] = node[i];
  if (i == nno - 1 || key[i + 1] == 0xffffffffu) *n = i + 1;
}

// el_dof[k][e] (SoA, -1 past the element's nodes) and the incidence pairs
__global__ void k_el_dof(int nel, int K, const int* nodes, const int* node_dof, int* el_dof,
                         unsigned long long* inc_key, int* inc_val) {
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= (long long)nel * 6) return;
  const int e = (int)(t / 6), k = (int)(t % 6);
  const int d = k < K ? node_dof[nodes[6 * (size_t)e + k]] : -1;
  el_dof[(size_t)k * nel + e] = d;
  if (k < K) {
    const size_t q = (size_t)e * K + k;
    inc_key[q] = ((unsigned long long)d << 32) | (unsigned)e;
    inc_val[q] = k;
  }
}

// (row, column) of every element matrix entry
__global__ void k_pairs(int nel, int K, const int* el_dof, unsigned long long* key) {
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= (long long)nel * K * K) return;
  const int e = (int)(t / (K * K)), w = (int)(t % (K * K)) / K, u = (int)(t % K);
  const unsigned long long r = (unsigned)el_dof[(size_t)w * nel + e];
  const unsigned long long c = (unsigned)el_dof[(size_t)u * nel + e];
  key[t] = (r << 32) | c;
}

// rows of sorted (row, .) keys: ptr[r] = first position of row r (rows
// without entries take the next row's start)
__global__ void k_row_starts(int n, const unsigned long long* key, long long m, int* ptr) {
  const int r = blockIdx.x * TPB + threadIdx.x;
  if (r > n) return;
  long long lo = 0, hi = m;
  const unsigned long long target = (unsigned long long)r << 32;
  while (lo < hi) {
    const long long mid = (lo + hi) >> 1;
    if (key[mid] < target) lo = mid + 1;
    else hi = mid;
  }
  ptr[r] = (int)lo;
}

__global__ void k_low_half(long long m, const unsigned long long* key, int* out) {
  const long long i = (long long)blockIdx.x * TPB + threadIdx.x;
  if (i < m) out[i] = (int)(key[i] & 0xffffffffu);
}

__global__ void k_row_max(int n, const int* ptr, int* mx) {
  const int r = blockIdx.x * TPB + threadIdx.x;
  if (r < n) atomicMax(mx, ptr[r + 1] - ptr[r]);
}

// compact element table: position of column dof(u) inside row dof(w)
__global__ void k_el_off(int nel, int K, const int* el_dof, const int* row_ptr,
                         const int* col_ind, uint8_t* el_off, int* bad) {
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= (long long)nel * 36) return;
  const int e = (int)(t % nel), q = (int)(t / nel), w = q / 6, u = q % 6;
  uint8_t o = 0;
  if (w < K && u < K) {
    const int r = el_dof[(size_t)w * nel + e], c = el_dof[(size_t)u * nel + e];
    int lo = row_ptr[r], hi = row_ptr[r + 1];
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (col_ind[mid] < c) lo = mid + 1;
      else hi = mid;
    }
    if (lo < row_ptr[r + 1] && col_ind[lo] == c) o = (uint8_t)(lo - row_ptr[r]);
    else atomicAdd(bad, 1);
  }
  el_off[t] = o;
}

// boundary-face incidence pairs through the owning element's DOFs
__global__ void k_face_pairs(int nbf, int nel, int K, const int* bf_elem, const int* el_dof,
                             unsigned long long* key, int* val) {
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= (long long)nbf * K) return;
  const int f = (int)(t / K), i = (int)(t % K);
  const int d = el_dof[(size_t)i * nel + bf_elem[f]];
  key[t] = ((unsigned long long)d << 32) | (unsigned)f;
  val[t] = i;
}

__global__ void k_centroids(int nel, int K, const int* nodes, const double* coor, double* cen) {
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= nel) return;
  for (int d = 0; d < 3; d++) {
    double c = 0;
    for (int i = 0; i < K; i++) c += coor[3 * (size_t)nodes[6 * (size_t)e + i] + d];
    c /= K;
    cen[(size_t)d * nel + e] = c;
  }
}

__device__ unsigned long long spread21(unsigned long long v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

struct Box {
  double lo[3], hi[3];
};

__global__ void k_morton(int nel, const double* cen, Box B, unsigned long long* code) {
  const int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= nel) return;
  unsigned long long c = 0;
  for (int d = 0; d < 3; d++) {
    const double span = B.hi[d] - B.lo[d];
    const double t = span > 0 ? (cen[(size_t)d * nel + e] - B.lo[d]) / span : 0.0;
    const unsigned long long q = (unsigned long long)(t * 2097151.0);
    c |= spread21(q) << d;
  }
  code[e] = c;
}

__global__ void k_inverse(int n, const int* perm, int* iperm) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < n) iperm[perm[i]] = i;
}

// first position 6 slot + k of every DOF over the elements in device order
__global__ void k_first_dof(int nel, int K, const int* eperm, const int* el_dof, unsigned* first) {
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= (long long)nel * K) return;
  const int slot = (int)(t / K), k = (int)(t % K);
  const int d = el_dof[(size_t)k * nel + eperm[slot]];
  atomicMin(&first[d], (unsigned)(6 * slot + k));
}

__global__ void k_slice_len(int n, const int* perm, const int* row_ptr, long long* len) {
  const int sl = blockIdx.x * TPB + threadIdx.x;
  const int ns = (n + 63) / 64;
  if (sl >= ns) return;
  int mx = 0;
  for (int i = sl * 64; i < min(n, sl * 64 + 64); i++)
    mx = max(mx, row_ptr[perm[i] + 1] - row_ptr[perm[i]]);
  len[sl] = 64LL * mx;
}

__global__ void k_fill_sell(int n, const int* perm, const int* iperm, const int* row_ptr,
                            const int* col_ind, const long long* sl_ptr, int* sl_col,
                            uint8_t* diag_k) {
  const int ns = (n + 63) / 64;
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= (long long)ns * 64) return;
  const int sl = (int)(t / 64), lane = (int)(t % 64);
  const long long base = sl_ptr[sl];
  const int len = (int)((sl_ptr[sl + 1] - base) / 64);
  const int i = sl * 64 + lane;
  const int r = i < n ? perm[i] : -1;
  const int rl = r >= 0 ? row_ptr[r + 1] - row_ptr[r] : 0;
  for (int k = 0; k < len; k++) {
    int c = i < n ? i : n - 1;
    if (k < rl) c = iperm[col_ind[row_ptr[r] + k]];
    sl_col[base + 64LL * k + lane] = c;
  }
  if (r >= 0)
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; k++)
      if (col_ind[k] == r) diag_k[i] = (uint8_t)(k - row_ptr[r]);
}

// device-order arrays of the problem
__global__ void k_dev_coor(int n, const int* perm, const int* dof_node, const double* coor,
                           double* cd) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const int v = dof_node[perm[i]];
  for (int k = 0; k < 3; k++) cd[3 * (size_t)i + k] = coor[3 * (size_t)v + k];
}

__global__ void k_dev_elems(int nel, const int* eperm, const int* iperm, const int* el_dof,
                            const uint8_t* el_off, int* ed, uint8_t* eo, uint8_t* offT) {
  const long long t = (long long)blockIdx.x * TPB + threadIdx.x;
  if (t >= (long long)nel * 36) return;
  const int slot = (int)(t % nel), q = (int)(t / nel);
  const int e = eperm[slot];
  const uint8_t o = el_off[(size_t)q * nel + e];
  eo[(size_t)q * nel + slot] = o;
  offT[(size_t)slot * 36 + q] = o;
  if (q < 6) {
    const int d = el_dof[(size_t)q * nel + e];
    ed[(size_t)q * nel + slot] = d < 0 ? -1 : iperm[d];
  }
}

__global__ void k_gather_i(long long n, const int* idx, const int* src, int* dst) {
  const long long i = (long long)blockIdx.x * TPB + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// lengths of the rows of a CSR list taken in device row order
__global__ void k_perm_len(int n, const int* perm, const int* ptr, int* len) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i < n) len[i] = ptr[perm[i] + 1] - ptr[perm[i]];
}

// a CSR list in device row order; entries mapped through emap (or kept)
__global__ void k_perm_rows(int n, const int* perm, const int* ptr, const int* ent,
                            const int* loc, const int* nptr, const int* emap, int* nent,
                            uint8_t* nloc) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  const int r = perm[i];
  for (int p = ptr[r], q = nptr[i]; p < ptr[r + 1]; p++, q++) {
    nent[q] = emap ? emap[ent[p]] : ent[p];
    nloc[q] = (uint8_t)loc[p];
  }
}

// copies in stream order: the library stream does not synchronise with the
// null stream, so a plain hipMemcpy could read before the kernels finish
template <class T>
int download(std::vector<T>& h, const T* d, size_t n, hipStream_t s) {
  h.resize(n);
  if (n) MF_HIP(hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  return 0;
}

int scan_exclusive_int(Scratch& S, const int* in, int* out, int n, hipStream_t s) {
  size_t bytes = 0;
  MF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
  if (S.ensure(bytes)) return -2;
  MF_HIP(hipcub::DeviceScan::ExclusiveSum(S.p, bytes, in, out, n, s));
  return 0;
}

}  // namespace

int problem_setup_device(const Mesh& M, Problem& P) {
  hipStream_t s = stream();
  const int nel = M.nel, nno = M.nno, nbf = M.nbf;
  const int K = nodes_of(M.type[0]);
  Setup& S = P.setup;
  S = Setup();
  Scratch scr;
  // inputs
  DevArr<int> nodes, bfe;
  DevArr<double> coor;
  if (nodes.alloc(6 * (size_t)nel) || coor.alloc(3 * (size_t)nno) || bfe.alloc(nbf)) return -2;
  MF_HIP(hipMemcpyAsync(nodes.p, M.nodes.data(), 6 * (size_t)nel * sizeof(int), hipMemcpyHostToDevice, s));
  MF_HIP(hipMemcpyAsync(coor.p, M.coor.data(), 3 * (size_t)nno * sizeof(double), hipMemcpyHostToDevice, s));
  if (nbf)
    MF_HIP(hipMemcpyAsync(bfe.p, M.bf_elem.data(), (size_t)nbf * sizeof(int), hipMemcpyHostToDevice, s));
  // DOF list: nodes by first appearance
  DevArr<unsigned> first, first_s;
  DevArr<int> nid, nid_s, node_dof, dof_node, cnt;
  if (first.alloc(nno) || first_s.alloc(nno) || nid.alloc(nno) || nid_s.alloc(nno) ||
      node_dof.alloc(nno) || dof_node.alloc(nno) || cnt.alloc(2))
    return -2;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(nno)), dim3(TPB), 0, s, first.p, (long long)nno, 0xffffffffu);
  hipLaunchKernelGGL(k_iota, dim3(grid_of(nno)), dim3(TPB), 0, s, nid.p, (long long)nno);
  hipLaunchKernelGGL(k_first_node, dim3(grid_of((long long)nel * K)), dim3(TPB), 0, s, nel, K,
                     nodes.p, first.p);
  if (sort_pairs(scr, first.p, first_s.p, nid.p, nid_s.p, nno, 32, s)) return -2;
  MF_HIP(hipMemsetAsync(node_dof.p, 0xff, (size_t)nno * sizeof(int), s));
  MF_HIP(hipMemsetAsync(cnt.p, 0, 2 * sizeof(int), s));
  hipLaunchKernelGGL(k_number_dofs, dim3(grid_of(nno)), dim3(TPB), 0, s, nno, first_s.p, nid_s.p,
                     node_dof.p, dof_node.p, cnt.p);
  int n = 0;
  MF_HIP(hipMemcpyAsync(&n, cnt.p, sizeof(int), hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  S.n = n;
  // element DOFs and the DOF -> element incidence (elements ascending)
  const long long ninc = (long long)nel * K;
  DevArr<int> el_dof, inc_v, inc_vs, inc_ptr, inc_el;
  DevArr<unsigned long long> inc_k, inc_ks;
  if (el_dof.alloc(6 * (size_t)nel) || inc_k.alloc(ninc) || inc_ks.alloc(ninc) ||
      inc_v.alloc(ninc) || inc_vs.alloc(ninc) || inc_ptr.alloc(n + 1) || inc_el.alloc(ninc))
    return -2;
  hipLaunchKernelGGL(k_el_dof, dim3(grid_of((long long)nel * 6)), dim3(TPB), 0, s, nel, K, nodes.p,
                     node_dof.p, el_dof.p, inc_k.p, inc_v.p);
  const int kb = 32 + bits_of((unsigned long long)std::max(n, 1));
  if (sort_pairs(scr, inc_k.p, inc_ks.p, inc_v.p, inc_vs.p, ninc, kb, s)) return -2;
  hipLaunchKernelGGL(k_row_starts, dim3(grid_of(n + 1)), dim3(TPB), 0, s, n, inc_ks.p, ninc, inc_ptr.p);
  hipLaunchKernelGGL(k_low_half, dim3(grid_of(ninc)), dim3(TPB), 0, s, ninc, inc_ks.p, inc_el.p);
  // CRS pattern: unique (row, column) pairs
  const long long npair = (long long)nel * K * K;
  DevArr<unsigned long long> pk, pks, pku;
  DevArr<int> row_ptr, col_ind, nsel, mx;
  if (pk.alloc(npair) || pks.alloc(npair) || pku.alloc(npair) || nsel.alloc(1) || mx.alloc(2) ||
      row_ptr.alloc(n + 1))
    return -2;
  hipLaunchKernelGGL(k_pairs, dim3(grid_of(npair)), dim3(TPB), 0, s, nel, K, el_dof.p, pk.p);
  {
    size_t bytes = 0;
    MF_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, pk.p, pks.p, (int)npair, 0, kb, s));
    if (scr.ensure(bytes)) return -2;
    MF_HIP(hipcub::DeviceRadixSort::SortKeys(scr.p, bytes, pk.p, pks.p, (int)npair, 0, kb, s));
    bytes = 0;
    MF_HIP(hipcub::DeviceSelect::Unique(nullptr, bytes, pks.p, pku.p, nsel.p, (int)npair, s));
    if (scr.ensure(bytes)) return -2;
    MF_HIP(hipcub::DeviceSelect::Unique(scr.p, bytes, pks.p, pku.p, nsel.p, (int)npair, s));
  }
  int nnz = 0;
  MF_HIP(hipMemcpyAsync(&nnz, nsel.p, sizeof(int), hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  S.nnz = nnz;
  if (col_ind.alloc(nnz)) return -2;
  hipLaunchKernelGGL(k_row_starts, dim3(grid_of(n + 1)), dim3(TPB), 0, s, n, pku.p, (long long)nnz,
                     row_ptr.p);
  hipLaunchKernelGGL(k_low_half, dim3(grid_of(nnz)), dim3(TPB), 0, s, (long long)nnz, pku.p, col_ind.p);
  MF_HIP(hipMemsetAsync(mx.p, 0, 2 * sizeof(int), s));
  hipLaunchKernelGGL(k_row_max, dim3(grid_of(n)), dim3(TPB), 0, s, n, row_ptr.p, mx.p);
  // element tables
  DevArr<uint8_t> el_off;
  if (el_off.alloc(36 * (size_t)nel)) return -2;
  hipLaunchKernelGGL(k_el_off, dim3(grid_of((long long)nel * 36)), dim3(TPB), 0, s, nel, K,
                     el_dof.p, row_ptr.p, col_ind.p, el_off.p, mx.p + 1);
  int hm[2];
  MF_HIP(hipMemcpyAsync(hm, mx.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  MF_HIP(hipStreamSynchronize(s));
  S.max_row = hm[0];
  if (hm[0] > 256) {
    set_error("a CRS row has more than 256 entries; compact tables need <= 256");
    return -1;
  }
  if (hm[1]) {
    set_error("assembly table lookup failed");
    return -1;
  }
  // boundary-face incidence (faces ascending per DOF)
  const long long nfp = (long long)nbf * K;
  DevArr<unsigned long long> fk, fks;
  DevArr<int> fv, fvs, finc_ptr, finc_f;
  if (fk.alloc(nfp) || fks.alloc(nfp) || fv.alloc(nfp) || fvs.alloc(nfp) || finc_ptr.alloc(n + 1) ||
      finc_f.alloc(nfp))
    return -2;
  if (nfp) {
    hipLaunchKernelGGL(k_face_pairs, dim3(grid_of(nfp)), dim3(TPB), 0, s, nbf, nel, K, bfe.p,
                       el_dof.p, fk.p, fv.p);
    if (sort_pairs(scr, fk.p, fks.p, fv.p, fvs.p, nfp, kb, s)) return -2;
    hipLaunchKernelGGL(k_low_half, dim3(grid_of(nfp)), dim3(TPB), 0, s, nfp, fks.p, finc_f.p);
  }
  hipLaunchKernelGGL(k_row_starts, dim3(grid_of(n + 1)), dim3(TPB), 0, s, n, fks.p, nfp, finc_ptr.p);
  // device order: elements along the Morton curve of their centroids
  DevArr<double> cen, red;
  DevArr<unsigned long long> code, code_s;
  DevArr<int> eid, eperm, ieperm;
  if (cen.alloc(3 * (size_t)nel) || red.alloc(6) || code.alloc(nel) || code_s.alloc(nel) ||
      eid.alloc(nel) || eperm.alloc(nel) || ieperm.alloc(nel))
    return -2;
  hipLaunchKernelGGL(k_centroids, dim3(grid_of(nel)), dim3(TPB), 0, s, nel, K, nodes.p, coor.p, cen.p);
  for (int d = 0; d < 3; d++) {
    size_t bytes = 0;
    MF_HIP(hipcub::DeviceReduce::Min(nullptr, bytes, cen.p + (size_t)d * nel, red.p + d, nel, s));
    if (scr.ensure(bytes)) return -2;
    MF_HIP(hipcub::DeviceReduce::Min(scr.p, bytes, cen.p + (size_t)d * nel, red.p + d, nel, s));
    bytes = 0;
    MF_HIP(hipcub::DeviceReduce::Max(nullptr, bytes, cen.p + (size_t)d * nel, red.p + 3 + d, nel, s));
    if (scr.ensure(bytes)) return -2;
    MF_HIP(hipcub::DeviceReduce::Max(scr.p, bytes, cen.p + (size_t)d * nel, red.p + 3 + d, nel, s));
  }
  Box B;
  {
    double h[6];
    MF_HIP(hipMemcpyAsync(h, red.p, 6 * sizeof(double), hipMemcpyDeviceToHost, s));
    MF_HIP(hipStreamSynchronize(s));
    for (int d = 0; d < 3; d++) {
      B.lo[d] = h[d];
      B.hi[d] = h[3 + d];
    }
  }
  hipLaunchKernelGGL(k_morton, dim3(grid_of(nel)), dim3(TPB), 0, s, nel, cen.p, B, code.p);
  hipLaunchKernelGGL(k_iota, dim3(grid_of(nel)), dim3(TPB), 0, s, eid.p, (long long)nel);
  if (sort_pairs(scr, code.p, code_s.p, eid.p, eperm.p, nel, 64, s)) return -2;
  hipLaunchKernelGGL(k_inverse, dim3(grid_of(nel)), dim3(TPB), 0, s, nel, eperm.p, ieperm.p);
  // DOFs by first appearance over the sorted elements
  DevArr<int> did, perm, iperm;
  if (did.alloc(n) || perm.alloc(n) || iperm.alloc(n)) return -2;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(n)), dim3(TPB), 0, s, first.p, (long long)n, 0xffffffffu);
  hipLaunchKernelGGL(k_first_dof, dim3(grid_of((long long)nel * K)), dim3(TPB), 0, s, nel, K,
                     eperm.p, el_dof.p, first.p);
  hipLaunchKernelGGL(k_iota, dim3(grid_of(n)), dim3(TPB), 0, s, did.p, (long long)n);
  if (sort_pairs(scr, first.p, first_s.p, did.p, perm.p, n, 32, s)) return -2;
  hipLaunchKernelGGL(k_inverse, dim3(grid_of(n)), dim3(TPB), 0, s, n, perm.p, iperm.p);
  // SELL-64 in device row order
  const int nsl = (n + 63) / 64;
  DevArr<long long> slen, sl_ptr;
  if (slen.alloc(nsl + 1) || sl_ptr.alloc(nsl + 1)) return -2;
  MF_HIP(hipMemsetAsync(slen.p, 0, (size_t)(nsl + 1) * sizeof(long long), s));
  hipLaunchKernelGGL(k_slice_len, dim3(grid_of(nsl)), dim3(TPB), 0, s, n, perm.p, row_ptr.p, slen.p);
  {
    size_t bytes = 0;
    MF_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, slen.p, sl_ptr.p, nsl + 1, s));
    if (scr.ensure(bytes)) return -2;
    MF_HIP(hipcub::DeviceScan::ExclusiveSum(scr.p, bytes, slen.p, sl_ptr.p, nsl + 1, s));
  }
  if (download(S.sl_ptr, sl_ptr.p, (size_t)nsl + 1, s)) return -2;
  const long long nnz_sell = S.sl_ptr.back();
  DevArr<int> scol;
  DevArr<uint8_t> diag_k;
  if (scol.alloc(nnz_sell) || diag_k.alloc(n)) return -2;
  MF_HIP(hipMemsetAsync(diag_k.p, 0, (size_t)n, s));
  hipLaunchKernelGGL(k_fill_sell, dim3(grid_of((long long)nsl * 64)), dim3(TPB), 0, s, n, perm.p,
                     iperm.p, row_ptr.p, col_ind.p, sl_ptr.p, scol.p, diag_k.p);
  // the problem's device arrays in device order
  DevArr<double> dcoor;
  DevArr<int> ded;
  DevArr<uint8_t> deo, doffT;
  if (dcoor.alloc(3 * (size_t)n) || ded.alloc(6 * (size_t)nel) || deo.alloc(36 * (size_t)nel) ||
      doffT.alloc(36 * (size_t)nel))
    return -2;
  hipLaunchKernelGGL(k_dev_coor, dim3(grid_of(n)), dim3(TPB), 0, s, n, perm.p, dof_node.p, coor.p, dcoor.p);
  hipLaunchKernelGGL(k_dev_elems, dim3(grid_of((long long)nel * 36)), dim3(TPB), 0, s, nel, eperm.p,
                     iperm.p, el_dof.p, el_off.p, ded.p, deo.p, doffT.p);
  DevArr<int> dbfe, dbfs;
  if (dbfe.alloc(nbf) || dbfs.alloc(nbf)) return -2;
  if (nbf) {
    hipLaunchKernelGGL(k_gather_i, dim3(grid_of(nbf)), dim3(TPB), 0, s, (long long)nbf, bfe.p,
                       ieperm.p, dbfe.p);
    MF_HIP(hipMemcpyAsync(dbfs.p, M.bf_side.data(), (size_t)nbf * sizeof(int), hipMemcpyHostToDevice, s));
  }
  // incidence lists in device row order
  DevArr<int> len, ip, fpp, iel, ff;
  DevArr<uint8_t> il, fl;
  if (len.alloc(n + 1) || ip.alloc(n + 1) || fpp.alloc(n + 1) || iel.alloc(ninc) || il.alloc(ninc) ||
      ff.alloc(nfp) || fl.alloc(nfp))
    return -2;
  MF_HIP(hipMemsetAsync(len.p, 0, (size_t)(n + 1) * sizeof(int), s));
  hipLaunchKernelGGL(k_perm_len, dim3(grid_of(n)), dim3(TPB), 0, s, n, perm.p, inc_ptr.p, len.p);
  if (scan_exclusive_int(scr, len.p, ip.p, n + 1, s)) return -2;
  hipLaunchKernelGGL(k_perm_rows, dim3(grid_of(n)), dim3(TPB), 0, s, n, perm.p, inc_ptr.p, inc_el.p,
                     inc_vs.p, ip.p, ieperm.p, iel.p, il.p);
  hipLaunchKernelGGL(k_perm_len, dim3(grid_of(n)), dim3(TPB), 0, s, n, perm.p, finc_ptr.p, len.p);
  if (scan_exclusive_int(scr, len.p, fpp.p, n + 1, s)) return -2;
  hipLaunchKernelGGL(k_perm_rows, dim3(grid_of(n)), dim3(TPB), 0, s, n, perm.p, finc_ptr.p,
                     finc_f.p, fvs.p, fpp.p, (const int*)nullptr, ff.p, fl.p);
  MF_HIP(hipGetLastError());
  // the maps the host reads later
  if (download(S.node_dof, node_dof.p, nno, s) || download(S.dof_node, dof_node.p, n, s) ||
      download(S.row_ptr, row_ptr.p, (size_t)n + 1, s) || download(S.col_ind, col_ind.p, nnz, s) ||
      download(S.el_dof, el_dof.p, 6 * (size_t)nel, s) ||
      download(S.el_off, el_off.p, 36 * (size_t)nel, s) || download(S.perm, perm.p, n, s) ||
      download(S.iperm, iperm.p, n, s) || download(S.eperm, eperm.p, nel, s) ||
      download(S.ieperm, ieperm.p, nel, s))
    return -2;
  // hand the device arrays to the problem
  P.n = n;
  P.nnz = nnz;
  P.nnz_sell = nnz_sell;
  P.sell = true;
  P.d_coor = dcoor.release();
  P.d_el_dof = ded.release();
  P.d_el_off = deo.release();
  P.d_el_offT = doffT.release();
  P.d_bf_elem = dbfe.release();
  P.d_bf_side = dbfs.release();
  P.d_sl_ptr = sl_ptr.release();
  P.d_scol = scol.release();
  P.d_diag_k = diag_k.release();
  P.d_inc_ptr = ip.release();
  P.d_inc_el = iel.release();
  P.d_inc_loc = il.release();
  P.d_finc_ptr = fpp.release();
  P.d_finc_f = ff.release();
  P.d_finc_loc = fl.release();
  MF_HIP(hipMalloc(&P.d_val, std::max<long long>(nnz_sell, 1) * sizeof(double)));
  MF_HIP(hipMalloc(&P.d_rhs, std::max(n, 1) * sizeof(double)));
  MF_HIP(hipMemsetAsync(P.d_val, 0, (size_t)nnz_sell * sizeof(double), s));
  MF_HIP(hipMemsetAsync(P.d_rhs, 0, (size_t)n * sizeof(double), s));
  MF_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace mf
