// Host setup: DOF numbering, sparsity and assembly tables.  Integer work,
// bit-exact with the reference path:
//   DOF entities in first-appearance order over element int-ents
//     (utd_util/uts_ls_intf.c:258-308); no renumbering
//     (sid_mkb/sis_mkb_intf.c:81-82), so block_id == list index and
//     posglob == block_id for block size 1 (sis_mkb_intf.c:1017-1033);
//   neighbour blocks = union over int-ents containing the DOF
//     (sis_mkb_intf.c:904-978), sorted ascending (:1107-1142);
//   CRS rows strictly ascending with the diagonal in place
//     (lsd_mkb/lad_crs/las_crs_intf.c:184-250);
//   assembly table entry (j + i*k) = position of (row L[j], col L[i])
//     (las_crs_intf.c:379-410), then k RHS rows (:413-425).
// Rows are built in parallel from a DOF -> element incidence list; the
// result does not depend on the thread count.
#include <algorithm>
#include <cstring>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "mf_internal.h"

namespace mf {

int build_setup(const Mesh& m, Setup& s) {
  s = Setup();
  s.node_dof.assign(m.nno, -1);
  s.dof_node.reserve(m.nno);
  for (int e = 0; e < m.nel; e++) {
    int k = nodes_of(m.type[e]);
    for (int i = 0; i < k; i++) {
      int v = m.nodes[6 * (size_t)e + i];
      if (v < 0 || v >= m.nno) {
        set_error("element references a node outside the mesh");
        return -1;
      }
      if (s.node_dof[v] < 0) {
        s.node_dof[v] = (int)s.dof_node.size();
        s.dof_node.push_back(v);
      }
    }
  }
  const int n = s.n = (int)s.dof_node.size();
  // element connectivity in DOF numbering
  s.el_dof.assign(6 * (size_t)m.nel, -1);
  for (int e = 0; e < m.nel; e++) {
    int k = nodes_of(m.type[e]);
    for (int i = 0; i < k; i++)
      s.el_dof[(size_t)i * m.nel + e] = s.node_dof[m.nodes[6 * (size_t)e + i]];
  }
  // DOF -> element incidence (CSR)
  std::vector<int> inc_ptr(n + 1, 0);
  for (int e = 0; e < m.nel; e++) {
    int k = nodes_of(m.type[e]);
    for (int i = 0; i < k; i++) inc_ptr[s.el_dof[(size_t)i * m.nel + e] + 1]++;
  }
  for (int r = 0; r < n; r++) inc_ptr[r + 1] += inc_ptr[r];
  std::vector<int> inc(inc_ptr[n]);
  {
    std::vector<int> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (int e = 0; e < m.nel; e++) {
      int k = nodes_of(m.type[e]);
      for (int i = 0; i < k; i++) inc[fill[s.el_dof[(size_t)i * m.nel + e]]++] = e;
    }
  }
  // rows: pass 1 counts, pass 2 fills
  std::vector<int> rowlen(n);
  auto row_cols = [&](int r, std::vector<int>& buf) {
    buf.clear();
    for (int p = inc_ptr[r]; p < inc_ptr[r + 1]; p++) {
      int e = inc[p];
      int k = nodes_of(m.type[e]);
      for (int i = 0; i < k; i++) buf.push_back(s.el_dof[(size_t)i * m.nel + e]);
    }
    std::sort(buf.begin(), buf.end());
    buf.erase(std::unique(buf.begin(), buf.end()), buf.end());
  };
#pragma omp parallel
  {
    std::vector<int> buf;
#pragma omp for schedule(static)
    for (int r = 0; r < n; r++) {
      row_cols(r, buf);
      rowlen[r] = (int)buf.size();
    }
  }
  s.row_ptr.assign(n + 1, 0);
  long long acc = 0;
  int mx = 0;
  for (int r = 0; r < n; r++) {
    s.row_ptr[r] = (int)acc;
    acc += rowlen[r];
    mx = std::max(mx, rowlen[r]);
  }
  if (acc > 2147483647LL) {
    set_error("CRS pattern exceeds 2^31 non-zeros");
    return -1;
  }
  s.row_ptr[n] = (int)acc;
  s.nnz = acc;
  s.max_row = mx;
  if (mx > 256) {
    set_error("a CRS row has more than 256 entries; compact tables need <= 256");
    return -1;
  }
  s.col_ind.resize(acc);
#pragma omp parallel
  {
    std::vector<int> buf;
#pragma omp for schedule(static)
    for (int r = 0; r < n; r++) {
      row_cols(r, buf);
      std::memcpy(&s.col_ind[s.row_ptr[r]], buf.data(), buf.size() * sizeof(int));
    }
  }
  // compact element tables: off[(w*6+u)*nel+e] = pos(row dof_w, col dof_u) - row_ptr
  auto find_off = [&](int row, int col) -> int {
    const int* b = &s.col_ind[s.row_ptr[row]];
    const int* e = &s.col_ind[s.row_ptr[row + 1]];
    const int* p = std::lower_bound(b, e, col);
    return (p != e && *p == col) ? (int)(p - b) : -1;
  };
  s.el_off.assign(36 * (size_t)m.nel, 0);
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
  for (int e = 0; e < m.nel; e++) {
    int k = nodes_of(m.type[e]);
    for (int w = 0; w < k; w++)
      for (int u = 0; u < k; u++) {
        int o = find_off(s.el_dof[(size_t)w * m.nel + e], s.el_dof[(size_t)u * m.nel + e]);
        if (o < 0) bad++;
        s.el_off[(size_t)(w * 6 + u) * m.nel + e] = (uint8_t)(o < 0 ? 0 : o);
      }
  }
  s.bf_off.assign(36 * (size_t)std::max(m.nbf, 1), 0);
  for (int f = 0; f < m.nbf; f++) {
    int e = m.bf_elem[f];
    int k = nodes_of(m.type[e]);
    for (int w = 0; w < k; w++)
      for (int u = 0; u < k; u++)
        s.bf_off[(size_t)(w * 6 + u) * m.nbf + f] =
            s.el_off[(size_t)(w * 6 + u) * m.nel + e];
  }
  if (bad) {
    set_error("assembly table lookup failed");
    return -1;
  }
  // keep the element incidence (ascending element order by construction)
  s.inc_ptr = inc_ptr;
  s.inc_el = inc;
  s.inc_loc.assign(inc.size(), 0);
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n; r++)
    for (int p = inc_ptr[r]; p < inc_ptr[r + 1]; p++) {
      int e = inc[p];
      for (int i = 0; i < nodes_of(m.type[e]); i++)
        if (s.el_dof[(size_t)i * m.nel + e] == r) s.inc_loc[p] = (uint8_t)i;
    }
  // boundary-face incidence through the owning element's DOFs
  s.finc_ptr.assign(n + 1, 0);
  for (int f = 0; f < m.nbf; f++) {
    int e = m.bf_elem[f];
    for (int i = 0; i < nodes_of(m.type[e]); i++) s.finc_ptr[s.el_dof[(size_t)i * m.nel + e] + 1]++;
  }
  for (int r = 0; r < n; r++) s.finc_ptr[r + 1] += s.finc_ptr[r];
  s.finc_f.assign(std::max(s.finc_ptr[n], 1), 0);
  s.finc_loc.assign(std::max(s.finc_ptr[n], 1), 0);
  {
    std::vector<int> fill(s.finc_ptr.begin(), s.finc_ptr.end() - 1);
    for (int f = 0; f < m.nbf; f++) {
      int e = m.bf_elem[f];
      for (int i = 0; i < nodes_of(m.type[e]); i++) {
        int r = s.el_dof[(size_t)i * m.nel + e];
        s.finc_f[fill[r]] = f;
        s.finc_loc[fill[r]] = (uint8_t)i;
        fill[r]++;
      }
    }
  }
  if (build_device_order(&m, s)) return -1;
  return build_sell(s);
}

namespace {
// 3 x 21-bit Morton interleave
uint64_t spread21(uint64_t v) {
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}
}  // namespace

int build_device_order(const Mesh* m, Setup& s) {
  const int n = s.n;
  s.perm.resize(n);
  s.iperm.resize(n);
  if (!m) {
    for (int i = 0; i < n; i++) s.perm[i] = s.iperm[i] = i;
    s.eperm.clear();
    s.ieperm.clear();
    return 0;
  }
  const int nel = m->nel;
  // element centroids -> Morton keys over the bounding box
  std::vector<double> cen(3 * (size_t)nel);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int e = 0; e < nel; e++) {
    const int k = nodes_of(m->type[e]);
    for (int d = 0; d < 3; d++) {
      double c = 0;
      for (int i = 0; i < k; i++) c += m->coor[3 * (size_t)m->nodes[6 * (size_t)e + i] + d];
      c /= k;
      cen[3 * (size_t)e + d] = c;
      lo[d] = std::min(lo[d], c);
      hi[d] = std::max(hi[d], c);
    }
  }
  std::vector<std::pair<uint64_t, int>> key(nel);
#pragma omp parallel for schedule(static)
  for (int e = 0; e < nel; e++) {
    uint64_t code = 0;
    for (int d = 0; d < 3; d++) {
      const double span = hi[d] - lo[d];
      double t = span > 0 ? (cen[3 * (size_t)e + d] - lo[d]) / span : 0.0;
      uint64_t q = (uint64_t)(t * 2097151.0);
      code |= spread21(q) << d;
    }
    key[e] = {code, e};
  }
  std::sort(key.begin(), key.end());
  s.eperm.resize(nel);
  s.ieperm.resize(nel);
  for (int i = 0; i < nel; i++) {
    s.eperm[i] = key[i].second;
    s.ieperm[key[i].second] = i;
  }
  // DOFs by first appearance over elements in device order
  std::fill(s.iperm.begin(), s.iperm.end(), -1);
  int next = 0;
  for (int i = 0; i < nel; i++) {
    const int e = s.eperm[i];
    for (int k = 0; k < nodes_of(m->type[e]); k++) {
      const int d = s.el_dof[(size_t)k * nel + e];
      if (s.iperm[d] < 0) {
        s.iperm[d] = next;
        s.perm[next++] = d;
      }
    }
  }
  if (next != n) {
    set_error("device order: DOF without an element");
    return -1;
  }
  return 0;
}

int build_sell(Setup& s) {
  const int n = s.n;
  const int ns = (n + 63) / 64;
  if ((int)s.perm.size() != n) build_device_order(nullptr, s);
  const std::vector<int>& P = s.perm;
  const std::vector<int>& IP = s.iperm;
  s.sl_ptr.assign(ns + 1, 0);
  for (int sl = 0; sl < ns; sl++) {
    int len = 0;
    for (int i = sl * 64; i < std::min(n, sl * 64 + 64); i++)
      len = std::max(len, s.row_ptr[P[i] + 1] - s.row_ptr[P[i]]);
    s.sl_ptr[sl + 1] = s.sl_ptr[sl] + 64LL * len;
  }
  s.sl_col.assign(s.sl_ptr[ns], 0);
  s.diag_k.assign(n, 0);
#pragma omp parallel for schedule(static)
  for (int sl = 0; sl < ns; sl++) {
    const long long base = s.sl_ptr[sl];
    const int len = (int)((s.sl_ptr[sl + 1] - base) / 64);
    for (int lane = 0; lane < 64; lane++) {
      const int i = sl * 64 + lane;  // device row
      const int r = i < n ? P[i] : -1;
      for (int k = 0; k < len; k++) {
        int c = i < n ? i : n - 1;  // padding points at the own row
        if (r >= 0 && k < s.row_ptr[r + 1] - s.row_ptr[r]) c = IP[s.col_ind[s.row_ptr[r] + k]];
        s.sl_col[base + 64LL * k + lane] = c;
      }
      if (r >= 0)
        for (int k = s.row_ptr[r]; k < s.row_ptr[r + 1]; k++)
          if (s.col_ind[k] == r) s.diag_k[i] = (uint8_t)(k - s.row_ptr[r]);
    }
  }
  return 0;
}

int fill_reference_table(const Mesh& m, const Setup& s, int ie, int* table) {
  if (ie < 0 || ie >= m.nel + m.nbf) {
    set_error("int-ent index out of range");
    return -1;
  }
  int e = ie < m.nel ? ie : m.bf_elem[ie - m.nel];
  int k = nodes_of(m.type[e]);
  int L[6];
  for (int i = 0; i < k; i++) L[i] = s.el_dof[(size_t)i * m.nel + e];
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++)
      table[j + i * k] = s.row_ptr[L[j]] + s.el_off[(size_t)(j * 6 + i) * m.nel + e];
  for (int j = 0; j < k; j++) table[k * k + j] = L[j];
  return (k + 1) * k;
}

}  // namespace mf
