// C ABI of libmodfem_hip.so (declarations and reference citations in
// include/modfem_hip.h).  Handles are small integers into process-wide
// tables; device memory is owned by the library.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>

#include "../../include/modfem_hip.h"
#include "mf_device.h"

using namespace mf;

namespace {

std::map<int, std::unique_ptr<Mesh>> g_meshes;
std::map<int, std::unique_ptr<Problem>> g_problems;
std::map<int, std::unique_ptr<Setup>> g_setups;
std::map<int, std::unique_ptr<Batch>> g_batches;
int g_next_mesh = 1, g_next_problem = 1, g_next_batch = 1;
bool g_inited = false;

Mesh* get_mesh(int id) {
  auto it = g_meshes.find(id);
  if (it == g_meshes.end()) {
    set_error("unknown mesh handle " + std::to_string(id));
    return nullptr;
  }
  return it->second.get();
}

Problem* get_problem(int id) {
  auto it = g_problems.find(id);
  if (it == g_problems.end()) {
    set_error("unknown problem handle " + std::to_string(id));
    return nullptr;
  }
  return it->second.get();
}

}  // namespace

Mesh* mf::mesh_by_id(int id) { return get_mesh(id); }

namespace {

int need_init() {
  if (!g_inited) {
    set_error("mf_init has not been called");
    return -1;
  }
  return 0;
}

int add_mesh(std::unique_ptr<Mesh> m) {
  int id = g_next_mesh++;
  g_meshes[id] = std::move(m);
  return id;
}

void free_problem_device(Problem& P) {
  void* ptrs[] = {P.d_coor, P.d_el_dof, P.d_el_off, P.d_bf_elem, P.d_bf_side, P.d_bf_off,
                  P.d_bf_g, P.d_row_ptr, P.d_col, P.d_val, P.d_rhs, P.d_val32,
                  P.d_coef, P.d_dinv, P.d_diag_ptr, P.d_dinv32, P.d_tile_row, P.d_V, P.d_H,
                  P.d_partials, P.d_counter, P.d_stat, P.d_x, P.d_work, P.d_work2, P.d_sl_ptr,
                  P.d_scol, P.d_diag_k, P.d_inc_ptr, P.d_inc_el, P.d_inc_loc,
                  P.d_finc_ptr, P.d_finc_f, P.d_finc_loc, P.d_el_offT, P.d_scol16, P.d_sbase,
                  P.d_erow, P.d_bf_type};
  for (void* p : ptrs)
    if (p) hipFree(p);
  if (P.h_stat) hipHostFree(P.h_stat);
  ilu_free(P);
  gs_free(P);
}

template <typename T>
int upload(T** dst, const T* src, size_t count) {
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  MF_HIP(hipMalloc(dst, bytes));
  if (src && count) MF_HIP(hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// SELL-64 arrays of a problem whose host setup has row_ptr/col_ind
int upload_sell(Problem& P) {
  Setup& S = P.setup;
  if (S.sl_ptr.empty()) build_sell(S);
  P.sell = true;
  P.nnz_sell = S.sl_ptr.back();
  if (upload(&P.d_sl_ptr, S.sl_ptr.data(), S.sl_ptr.size())) return -2;
  if (upload(&P.d_scol, S.sl_col.data(), S.sl_col.size())) return -2;
  if (upload(&P.d_diag_k, S.diag_k.data(), S.diag_k.size())) return -2;
  if (upload<double>(&P.d_val, nullptr, (size_t)P.nnz_sell)) return -2;
  return 0;
}

// CRS-ordered values <-> SELL-64 positions (host)
// (reference row r is device row iperm[r]; entry order inside a row is kept)
void crs_to_sell(const Setup& S, const double* crs, double* sell) {
  std::fill(sell, sell + S.sl_ptr.back(), 0.0);
  for (int r = 0; r < S.n; r++) {
    const int i = S.iperm[r];
    const long long base = S.sl_ptr[i >> 6] + (i & 63);
    for (int k = S.row_ptr[r]; k < S.row_ptr[r + 1]; k++)
      sell[base + 64LL * (k - S.row_ptr[r])] = crs[k];
  }
}

void sell_to_crs(const Setup& S, const double* sell, double* crs) {
  for (int r = 0; r < S.n; r++) {
    const int i = S.iperm[r];
    const long long base = S.sl_ptr[i >> 6] + (i & 63);
    for (int k = S.row_ptr[r]; k < S.row_ptr[r + 1]; k++)
      crs[k] = sell[base + 64LL * (k - S.row_ptr[r])];
  }
}

// reference-order vector -> device order and back
std::vector<double> to_device_order(const Setup& S, const double* x) {
  std::vector<double> d(S.n);
  for (int i = 0; i < S.n; i++) d[i] = x[S.perm[i]];
  return d;
}

void to_reference_order(const Setup& S, const double* d, double* x) {
  for (int i = 0; i < S.n; i++) x[S.perm[i]] = d[i];
}

// the assembly mode a problem runs: the row-wise and two-pass modes keep a
// row's sums in LDS for rows of up to 32 entries; a mesh with longer rows
// (vertices of high valence) is assembled by element atomics instead
static int effective_mode(const Problem& P) {
  return (P.assembly_mode != 1 && P.setup.max_row > 32) ? 1 : P.assembly_mode;
}

// a new system invalidates everything derived from the old one
void assembled(Problem& P) {
  P.dinv_valid = false;
  P.val32_valid = false;
  if (P.ilu) P.ilu->valid = false;
  if (P.gs) P.gs->valid = false;
}

// integrate + assemble in the problem's assembly mode
int assemble_problem(Problem& P, int precision) {
  const int mode = effective_mode(P);
  if (mode == 0) {
    if (launch_assemble_rows(P, precision)) return -2;
  } else if (mode == 2) {
    if (launch_assemble_two_pass(P, precision)) return -2;
  } else {
    if (launch_zero(P.d_val, P.nnz_sell) || launch_zero(P.d_rhs, P.n)) return -2;
    if (launch_integrate(P, precision, 0, nullptr, 0, P.nel)) return -2;
    if (launch_faces(P, 0, nullptr, 0, P.nbf)) return -2;
  }
  assembled(P);
  return 0;
}

int rhs_to_host(const Problem& P, double* rhs) {
  std::vector<double> d(P.n);
  MF_HIP(hipMemcpy(d.data(), P.d_rhs, (size_t)P.n * sizeof(double), hipMemcpyDeviceToHost));
  to_reference_order(P.setup, d.data(), rhs);
  return 0;
}

}  // namespace

extern "C" {

int mf_init(int device) {
  int rc = device_init(device);
  if (rc == 0) rc = noalloc_scratch_init(device);
  if (rc == 0) g_inited = true;
  return rc;
}

const char* mf_last_error(void) { return last_error(); }
int mf_version(void) { return 1; }

int mf_mesh_create(int nno, const double* coor, int nel, const int* el_type,
                   const int* el_nodes, int nbf, const int* bf_elem, const int* bf_side,
                   const int* bf_bc) {
  if (nno <= 0 || nel <= 0 || !coor || !el_type || !el_nodes || nbf < 0 ||
      (nbf > 0 && (!bf_elem || !bf_side || !bf_bc))) {
    set_error("mf_mesh_create: invalid arguments");
    return -1;
  }
  auto m = std::make_unique<Mesh>();
  m->nno = nno;
  m->coor.assign(coor, coor + 3 * (size_t)nno);
  m->nel = nel;
  m->type.assign(el_type, el_type + nel);
  m->nodes.assign(el_nodes, el_nodes + 6 * (size_t)nel);
  for (int e = 0; e < nel; e++) {
    if (m->type[e] != kPrism && m->type[e] != kTetra) {
      set_error("mf_mesh_create: element type must be MF_PRISM or MF_TETRA");
      return -1;
    }
    for (int i = 0; i < nodes_of(m->type[e]); i++) {
      int v = m->nodes[6 * (size_t)e + i];
      if (v < 0 || v >= nno) {
        set_error("mf_mesh_create: element node index out of range");
        return -1;
      }
    }
  }
  m->nbf = nbf;
  if (nbf) {
    m->bf_elem.assign(bf_elem, bf_elem + nbf);
    m->bf_side.assign(bf_side, bf_side + nbf);
    m->bf_bc.assign(bf_bc, bf_bc + nbf);
    for (int f = 0; f < nbf; f++) {
      int e = m->bf_elem[f];
      if (e < 0 || e >= nel || m->bf_side[f] < 0 ||
          m->bf_side[f] >= (m->type[e] == kPrism ? 5 : 4)) {
        set_error("mf_mesh_create: boundary face owner/side out of range");
        return -1;
      }
    }
  }
  return add_mesh(std::move(m));
}

int mf_mesh_read_jk(const char* filename) {
  auto m = std::make_unique<Mesh>();
  if (mesh_read_jk(filename, *m) < 0) return -1;
  return add_mesh(std::move(m));
}

int mf_mesh_read_nas(const char* filename) {
  auto m = std::make_unique<Mesh>();
  if (mesh_read_nas(filename, *m) < 0) return -1;
  return add_mesh(std::move(m));
}

int mf_mesh_read_dat(const char* filename) {
  auto m = std::make_unique<Mesh>();
  if (mesh_read_dat(filename, *m) < 0) return -1;
  return add_mesh(std::move(m));
}

int mf_mesh_write_jk(int mesh, const char* filename) {
  Mesh* m = get_mesh(mesh);
  if (!m) return -3;
  return mesh_write_jk(*m, filename);
}

int mf_mesh_box_prism(int nx, int ny, int nz, double lx, double ly, double lz,
                      unsigned long long seed, double jitter, int bc) {
  auto m = std::make_unique<Mesh>();
  if (mesh_box_prism(nx, ny, nz, lx, ly, lz, seed, jitter, bc, *m) < 0) return -1;
  return add_mesh(std::move(m));
}

int mf_mesh_box_tetra(int nx, int ny, int nz, double lx, double ly, double lz,
                      unsigned long long seed, double jitter, int bc) {
  auto m = std::make_unique<Mesh>();
  if (mesh_box_tetra(nx, ny, nz, lx, ly, lz, seed, jitter, bc, *m) < 0) return -1;
  return add_mesh(std::move(m));
}

int mf_mesh_box_hybrid(int nx, int ny, int nz, double lx, double ly, double lz,
                       unsigned long long seed, double jitter, int bc, double tet_frac) {
  auto m = std::make_unique<Mesh>();
  if (mesh_box_hybrid(nx, ny, nz, lx, ly, lz, seed, jitter, bc, tet_frac, *m) < 0) return -1;
  return add_mesh(std::move(m));
}

int mf_mesh_sizes(int mesh, int* nno, int* nel, int* nbf) {
  Mesh* m = get_mesh(mesh);
  if (!m) return -3;
  if (nno) *nno = m->nno;
  if (nel) *nel = m->nel;
  if (nbf) *nbf = m->nbf;
  return 0;
}

int mf_mesh_get(int mesh, double* coor, int* el_type, int* el_nodes, int* bf_elem,
                int* bf_side, int* bf_bc) {
  Mesh* m = get_mesh(mesh);
  if (!m) return -3;
  if (coor) std::memcpy(coor, m->coor.data(), m->coor.size() * sizeof(double));
  if (el_type) std::memcpy(el_type, m->type.data(), m->type.size() * sizeof(int));
  if (el_nodes) std::memcpy(el_nodes, m->nodes.data(), m->nodes.size() * sizeof(int));
  if (m->nbf) {
    if (bf_elem) std::memcpy(bf_elem, m->bf_elem.data(), m->nbf * sizeof(int));
    if (bf_side) std::memcpy(bf_side, m->bf_side.data(), m->nbf * sizeof(int));
    if (bf_bc) std::memcpy(bf_bc, m->bf_bc.data(), m->nbf * sizeof(int));
  }
  return 0;
}

int mf_mesh_free(int mesh) {
  if (!get_mesh(mesh)) return -3;
  g_setups.erase(mesh);
  g_meshes.erase(mesh);
  return 0;
}

// host-only set-up of a mesh (no device work): lets the integer maps be
// checked without a GPU
int mf_mesh_setup(int mesh, int* n, long long* nnz) {
  Mesh* m = get_mesh(mesh);
  if (!m) return -3;
  auto s = std::make_unique<Setup>();
  if (build_setup(*m, *s) < 0) return -1;
  if (n) *n = s->n;
  if (nnz) *nnz = s->nnz;
  g_setups[mesh] = std::move(s);
  return 0;
}

int mf_mesh_setup_get(int mesh, int* dof_node, int* row_ptr, int* col_ind) {
  auto it = g_setups.find(mesh);
  if (it == g_setups.end()) {
    set_error("mf_mesh_setup has not been called for this mesh");
    return -3;
  }
  const Setup& s = *it->second;
  if (dof_node) std::memcpy(dof_node, s.dof_node.data(), s.dof_node.size() * sizeof(int));
  if (row_ptr) std::memcpy(row_ptr, s.row_ptr.data(), s.row_ptr.size() * sizeof(int));
  if (col_ind) std::memcpy(col_ind, s.col_ind.data(), s.col_ind.size() * sizeof(int));
  return 0;
}

// the DOF-to-node map of a mesh (set up on first use)
static const Setup* field_setup(int mesh) {
  auto it = g_setups.find(mesh);
  if (it == g_setups.end()) {
    if (mf_mesh_setup(mesh, nullptr, nullptr) < 0) return nullptr;
    it = g_setups.find(mesh);
  }
  return it->second.get();
}

// field dump in R's standard format (apr_write_field, apd_std_lin/
// aps_std_lin_intf.c:532-650, one equation, one solution vector): "1 1",
// the number of DOF entities, then per node holding a DOF, in ascending node
// id (1-based), "id value " with %.12lg and values below accuracy as 0
int mf_field_write(int mesh, const double* x, double accuracy, const char* filename) {
  const Setup* S = field_setup(mesh);
  if (!S) return -3;
  if (!x || !filename) {
    set_error("mf_field_write: x and filename required");
    return -1;
  }
  Mesh* m = get_mesh(mesh);
  std::vector<int> dof_of(m->nno, -1);
  for (size_t d = 0; d < S->dof_node.size(); d++) dof_of[S->dof_node[d]] = (int)d;
  FILE* fp = std::fopen(filename, "w");
  if (!fp) {
    set_error(std::string("mf_field_write: cannot open ") + filename);
    return -1;
  }
  std::fprintf(fp, "%d %d\n", 1, 1);
  std::fprintf(fp, "%d\n", (int)S->dof_node.size());
  for (int v = 0; v < m->nno; v++) {
    if (dof_of[v] < 0) continue;
    const double val = x[dof_of[v]];
    std::fprintf(fp, "%d ", v + 1);
    std::fprintf(fp, "%.12lg ", std::fabs(val) < accuracy ? 0.0 : val);
    std::fprintf(fp, "\n");
  }
  const bool ok = std::fclose(fp) == 0;
  if (!ok) {
    set_error("mf_field_write: write failed");
    return -1;
  }
  return 1;
}

// read a field dump (apr_create_field with APC_READ, aps_std_lin_intf.c:
// 300-420): "nreq nr_sol", the number of DOF entities, then per entity its
// node id and nr_sol vectors of nreq values; the first vector is returned
// in DOF order.  One equation; every DOF must be present.
int mf_field_read(int mesh, const char* filename, double* x) {
  const Setup* S = field_setup(mesh);
  if (!S) return -3;
  Mesh* m = get_mesh(mesh);
  FILE* fp = filename ? std::fopen(filename, "r") : nullptr;
  if (!fp || !x) {
    if (fp) std::fclose(fp);
    set_error(std::string("mf_field_read: cannot open ") + (filename ? filename : "(null)"));
    return -1;
  }
  std::vector<int> dof_of(m->nno, -1);
  for (size_t d = 0; d < S->dof_node.size(); d++) dof_of[S->dof_node[d]] = (int)d;
  std::vector<char> seen(S->dof_node.size(), 0);
  int nreq = 0, nsol = 0, nent = 0, rc = 0;
  if (std::fscanf(fp, "%d %d", &nreq, &nsol) != 2 || std::fscanf(fp, "%d", &nent) != 1 ||
      nreq != 1 || nsol < 1 || nsol > 3 || nent < 0 || nent > m->nno) {
    set_error("mf_field_read: bad header (one equation, 1..3 solution vectors expected)");
    rc = -1;
  }
  for (int k = 0; k < nent && rc == 0; k++) {
    int id = 0;
    double v[3] = {0, 0, 0};
    bool ok = std::fscanf(fp, "%d", &id) == 1;
    for (int q = 0; q < nsol && ok; q++) ok = std::fscanf(fp, "%lg", &v[q]) == 1;
    if (!ok || id < 1 || id > m->nno || dof_of[id - 1] < 0) {
      set_error("mf_field_read: bad entry " + std::to_string(k) + " (node " + std::to_string(id) + ")");
      rc = -1;
      break;
    }
    x[dof_of[id - 1]] = v[0];
    seen[dof_of[id - 1]] = 1;
  }
  std::fclose(fp);
  if (rc == 0)
    for (size_t d = 0; d < seen.size(); d++)
      if (!seen[d]) {
        set_error("mf_field_read: no value for the DOF of node " + std::to_string(S->dof_node[d] + 1));
        return -1;
      }
  return rc == 0 ? nent : rc;
}

int mf_mesh_setup_order(int mesh, int* row_perm, int* el_perm) {
  auto it = g_setups.find(mesh);
  if (it == g_setups.end()) {
    set_error("mf_mesh_setup has not been called for this mesh");
    return -3;
  }
  const Setup& s = *it->second;
  if (row_perm) std::memcpy(row_perm, s.perm.data(), s.perm.size() * sizeof(int));
  if (el_perm) std::memcpy(el_perm, s.eperm.data(), s.eperm.size() * sizeof(int));
  return 0;
}

int mf_mesh_setup_table(int mesh, int int_ent, int* table) {
  auto it = g_setups.find(mesh);
  Mesh* m = get_mesh(mesh);
  if (!m) return -3;
  if (it == g_setups.end()) {
    set_error("mf_mesh_setup has not been called for this mesh");
    return -3;
  }
  return fill_reference_table(*m, *it->second, int_ent, table);
}

int mf_problem_create(int mesh) {
  if (need_init()) return -1;
  Mesh* m = get_mesh(mesh);
  if (!m) return -3;
  auto P = std::make_unique<Problem>();
  P->mesh = *m;
  if (build_setup(P->mesh, P->setup) < 0) return -1;
  const Mesh& M = P->mesh;
  const Setup& S = P->setup;
  P->n = S.n;
  P->nel = M.nel;
  P->nbf = M.nbf;
  P->nnz = S.nnz;
  P->el_type = M.type[0];
  for (int e = 0; e < M.nel; e++)
    if (M.type[e] != P->el_type) {
      set_error("mf_problem_create: mixed prism/tetra std meshes are not supported");
      return -1;
    }
  // device arrays in the internal locality numbering (rows perm, elements
  // eperm); every list keeps the reference's order within a row, so the
  // summation order of each matrix entry is unchanged
  const int n = S.n, nel = M.nel;
  const std::vector<int>&PR = S.perm, &IPR = S.iperm, &EP = S.eperm, &IEP = S.ieperm;
  {
    std::vector<double> cd(3 * (size_t)n);
    for (int i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) cd[3 * (size_t)i + k] = M.coor[3 * (size_t)S.dof_node[PR[i]] + k];
    if (upload(&P->d_coor, cd.data(), cd.size())) return -2;
  }
  {
    std::vector<int> ed(6 * (size_t)nel);
    std::vector<uint8_t> eo(36 * (size_t)nel), offT((size_t)nel * 36);
#pragma omp parallel for schedule(static)
    for (int e = 0; e < nel; e++) {
      const int r = EP[e];
      for (int k = 0; k < 6; k++) {
        const int d = S.el_dof[(size_t)k * nel + r];
        ed[(size_t)k * nel + e] = d < 0 ? -1 : IPR[d];
      }
      for (int q = 0; q < 36; q++) {
        const uint8_t o = S.el_off[(size_t)q * nel + r];
        eo[(size_t)q * nel + e] = o;
        offT[(size_t)e * 36 + q] = o;
      }
    }
    if (upload(&P->d_el_dof, ed.data(), ed.size())) return -2;
    if (upload(&P->d_el_off, eo.data(), eo.size())) return -2;
    if (upload(&P->d_el_offT, offT.data(), offT.size())) return -2;
  }
  {
    std::vector<int> bfe(M.nbf);
    for (int f = 0; f < M.nbf; f++) bfe[f] = IEP[M.bf_elem[f]];
    if (upload(&P->d_bf_elem, bfe.data(), bfe.size())) return -2;
  }
  if (upload(&P->d_bf_side, M.bf_side.data(), M.bf_side.size())) return -2;
  if (upload_sell(*P)) return -2;
  if (upload<double>(&P->d_rhs, nullptr, (size_t)n)) return -2;
  if (launch_zero(P->d_val, P->nnz_sell) || launch_zero(P->d_rhs, n)) return -2;
  {
    // incidence lists per device row, entries still in ascending reference
    // element / face order (the reference's summation order)
    std::vector<int> ip(n + 1, 0), fp(n + 1, 0);
    for (int i = 0; i < n; i++) {
      ip[i + 1] = ip[i] + S.inc_ptr[PR[i] + 1] - S.inc_ptr[PR[i]];
      fp[i + 1] = fp[i] + S.finc_ptr[PR[i] + 1] - S.finc_ptr[PR[i]];
    }
    std::vector<int> iel(S.inc_el.size()), ff(S.finc_f.size());
    std::vector<uint8_t> il(S.inc_loc.size()), fl(S.finc_loc.size());
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      const int r = PR[i];
      for (int p = S.inc_ptr[r], q = ip[i]; p < S.inc_ptr[r + 1]; p++, q++) {
        iel[q] = IEP[S.inc_el[p]];
        il[q] = S.inc_loc[p];
      }
      for (int p = S.finc_ptr[r], q = fp[i]; p < S.finc_ptr[r + 1]; p++, q++) {
        ff[q] = S.finc_f[p];
        fl[q] = S.finc_loc[p];
      }
    }
    if (upload(&P->d_inc_ptr, ip.data(), ip.size())) return -2;
    if (upload(&P->d_inc_el, iel.data(), iel.size())) return -2;
    if (upload(&P->d_inc_loc, il.data(), il.size())) return -2;
    if (upload(&P->d_finc_ptr, fp.data(), fp.size())) return -2;
    if (upload(&P->d_finc_f, ff.data(), ff.size())) return -2;
    if (upload(&P->d_finc_loc, fl.data(), fl.size())) return -2;
  }
  // default coefficients: -Laplace u = 1 (DIFF_IN_CUBE-style selection)
  P->mask = TermMask{0, 1, 0, 0, 0, 0, 0, 1};
  P->coef_const[C_A + 0] = P->coef_const[C_A + 4] = P->coef_const[C_A + 8] = 1.0;
  P->coef_const[C_S] = 1.0;
  MF_HIP(hipStreamSynchronize(stream()));
  int id = g_next_problem++;
  g_problems[id] = std::move(P);
  return id;
}

int mf_problem_create_device(int mesh) {
  if (need_init()) return -1;
  Mesh* m = get_mesh(mesh);
  if (!m) return -3;
  auto P = std::make_unique<Problem>();
  P->mesh = *m;
  const Mesh& M = P->mesh;
  for (int e = 0; e < M.nel; e++)
    if (M.type[e] != M.type[0]) {
      set_error("mf_problem_create_device: mixed prism/tetra std meshes are not supported");
      return -1;
    }
  const long long K = nodes_of(M.type[0]);
  if ((long long)M.nel * K * K > 2147483647LL || 6LL * M.nel > 4294967295LL) {
    set_error("mf_problem_create_device: more element-matrix entries than 2^31");
    return -1;
  }
  P->nel = M.nel;
  P->nbf = M.nbf;
  P->el_type = M.type[0];
  const int rc = problem_setup_device(M, *P);
  if (rc) return rc;
  P->mask = TermMask{0, 1, 0, 0, 0, 0, 0, 1};
  P->coef_const[C_A + 0] = P->coef_const[C_A + 4] = P->coef_const[C_A + 8] = 1.0;
  P->coef_const[C_S] = 1.0;
  int id = g_next_problem++;
  g_problems[id] = std::move(P);
  return id;
}

int mf_problem_create_crs(int n, const int* row_ptr, const int* col_ind, const double* val,
                          const double* rhs) {
  if (need_init()) return -1;
  if (n <= 0 || !row_ptr || !col_ind || !val) {
    set_error("mf_problem_create_crs: invalid arguments");
    return -1;
  }
  auto P = std::make_unique<Problem>();
  P->n = n;
  P->setup.n = n;
  P->setup.row_ptr.assign(row_ptr, row_ptr + n + 1);
  long long nnz = row_ptr[n];
  P->setup.col_ind.assign(col_ind, col_ind + nnz);
  P->setup.nnz = nnz;
  P->nnz = nnz;
  for (int r = 0; r < n; r++) {
    bool diag = false;
    for (int k = row_ptr[r]; k < row_ptr[r + 1]; k++) {
      if (col_ind[k] < 0 || col_ind[k] >= n || (k > row_ptr[r] && col_ind[k] <= col_ind[k - 1])) {
        set_error("mf_problem_create_crs: columns must be in range and strictly ascending");
        return -1;
      }
      if (col_ind[k] == r) diag = true;
    }
    if (!diag) {
      set_error("mf_problem_create_crs: every row needs its diagonal entry");
      return -1;
    }
  }
  {
    int mx = 0;
    for (int r = 0; r < n; r++) mx = std::max(mx, row_ptr[r + 1] - row_ptr[r]);
    if (mx > 255) {
      set_error("mf_problem_create_crs: rows longer than 255 entries are not supported");
      return -1;
    }
    P->setup.max_row = mx;
  }
  if (upload_sell(*P)) return -2;
  {
    std::vector<double> sv(P->nnz_sell);
    crs_to_sell(P->setup, val, sv.data());
    MF_HIP(hipMemcpy(P->d_val, sv.data(), sv.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  {
    std::vector<double> d;
    if (rhs) d = to_device_order(P->setup, rhs);
    if (upload<double>(&P->d_rhs, rhs ? d.data() : nullptr, (size_t)n)) return -2;
  }
  if (!rhs && launch_zero(P->d_rhs, n)) return -2;
  MF_HIP(hipStreamSynchronize(stream()));
  int id = g_next_problem++;
  g_problems[id] = std::move(P);
  return id;
}

int mf_problem_sizes(int pid, int* n, long long* nnz, int* nel, int* nbf, int* max_row) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (n) *n = P->n;
  if (nnz) *nnz = P->nnz;
  if (nel) *nel = P->nel;
  if (nbf) *nbf = P->nbf;
  if (max_row) *max_row = P->setup.max_row;
  return 0;
}

int mf_problem_get_dofs(int pid, int* dof_node) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  std::memcpy(dof_node, P->setup.dof_node.data(), P->setup.dof_node.size() * sizeof(int));
  return 0;
}

int mf_problem_get_pattern(int pid, int* row_ptr, int* col_ind) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (row_ptr) std::memcpy(row_ptr, P->setup.row_ptr.data(), P->setup.row_ptr.size() * sizeof(int));
  if (col_ind) std::memcpy(col_ind, P->setup.col_ind.data(), P->setup.col_ind.size() * sizeof(int));
  return 0;
}

int mf_problem_get_table(int pid, int int_ent, int* table) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (P->nel == 0) {
    set_error("problem has no mesh");
    return -1;
  }
  return fill_reference_table(P->mesh, P->setup, int_ent, table);
}

int mf_problem_set_coefficients(int pid, const int* ind, int coef_model,
                                const double* coef_const, const double* coef_el,
                                int src_model, const double* bf_g) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (!ind || coef_model < CM_UNIFORM || coef_model > CM_KAPPA_S || src_model < SRC_CONST ||
      src_model > SRC_DIFF_IN_BOX) {
    set_error("mf_problem_set_coefficients: invalid arguments");
    return -1;
  }
  if ((ind[11] || ind[12]) && !ind[13]) {
    set_error("2D convection coefficient sets (bx,by without bz) are not supported");
    return -1;
  }
  if (!ind[10] && (ind[2] || ind[6])) {
    set_error("2D diffusion coefficient sets (no azz) are not supported");
    return -1;
  }
  if ((ind[14] || ind[15]) && !ind[16]) {
    set_error("2D t-coefficient sets (tx,ty without tz) are not supported");
    return -1;
  }
  if ((ind[19] || ind[20]) && !ind[21]) {
    set_error("2D q-coefficient sets (qx,qy without qz) are not supported");
    return -1;
  }
  if (ind[18]) {
    set_error("the obsolete l (time rhs) coefficient is not supported");
    return -1;
  }
  TermMask M;
  M.m = ind[1] != 0;
  M.diff_full = ind[10] && ind[2];
  M.diff_diag = ind[10] && !ind[2];
  M.conv = ind[13] != 0;
  M.tconv = ind[16] != 0;
  M.reac = ind[17] != 0;
  M.q = ind[21] != 0;
  M.s = ind[22] != 0;
  P->mask = M;
  P->term_set = (M.m || M.tconv || M.reac || M.q) ? TS_ALL : (M.conv ? TS_DIFF_CONV : TS_DIFF);
  for (int k = 0; k < NCOEF; k++) P->coef_const[k] = coef_const ? coef_const[k] : 0.0;
  P->coef_model = coef_model;
  P->src_model = src_model;
  if (P->d_coef) {
    hipFree(P->d_coef);
    P->d_coef = nullptr;
  }
  size_t per = coef_model == CM_KAPPA ? 1 : coef_model == CM_KAPPA_B ? 4
             : coef_model == CM_FULL ? NCOEF : coef_model == CM_KAPPA_S ? 2 : 0;
  if (per) {
    if (!coef_el) {
      set_error("mf_problem_set_coefficients: coef_el required for this model");
      return -1;
    }
    // SoA [slot][nel] in device element order
    const size_t N = P->nel;
    std::vector<double> cd(per * N);
    for (size_t k = 0; k < per; k++)
      for (size_t e = 0; e < N; e++) cd[k * N + e] = coef_el[k * N + P->setup.eperm[e]];
    if (upload(&P->d_coef, cd.data(), cd.size())) return -2;
  }
  if (P->d_bf_g) {
    hipFree(P->d_bf_g);
    P->d_bf_g = nullptr;
  }
  if (bf_g && P->nbf)
    if (upload(&P->d_bf_g, bf_g, (size_t)P->nbf)) return -2;
  return 0;
}

int mf_problem_set_face_bc(int pid, const int* bc_type, const double* g) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (P->nbf == 0) return 0;
  if (!bc_type || !g) {
    set_error("mf_problem_set_face_bc: bc_type and g are required");
    return -1;
  }
  for (int f = 0; f < P->nbf; f++)
    if (bc_type[f] != BC_DIRICHLET && bc_type[f] != BC_NEUMANN) {
      set_error("mf_problem_set_face_bc: types are 1 (Dirichlet penalty) or 2 (Neumann)");
      return -1;
    }
  if (P->d_bf_type) hipFree(P->d_bf_type);
  if (P->d_bf_g) hipFree(P->d_bf_g);
  P->d_bf_type = nullptr;
  P->d_bf_g = nullptr;
  if (upload(&P->d_bf_type, bc_type, (size_t)P->nbf)) return -2;
  if (upload(&P->d_bf_g, g, (size_t)P->nbf)) return -2;
  return 0;
}

int mf_integrate_assemble(int pid, int precision) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (precision != 64 && precision != 32) {
    set_error("precision must be 64 or 32");
    return -1;
  }
  if (P->nel == 0) {
    set_error("problem has no mesh");
    return -1;
  }
  if (assemble_problem(*P, precision)) return -2;
  P->dinv_valid = false;  // refilled before the next preconditioned solve
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_integrate_assemble_async(int pid, int precision) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (precision != 64 && precision != 32) {
    set_error("precision must be 64 or 32");
    return -1;
  }
  if (P->nel == 0) {
    set_error("problem has no mesh");
    return -1;
  }
  return assemble_problem(*P, precision) ? -2 : 0;
}

int mf_element_matrices(int pid, int first, int count, int precision, double* out) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (first < 0 || count < 0 || first + count > P->nel || !out) {
    set_error("mf_element_matrices: range out of bounds");
    return -1;
  }
  if (count == 0) return 0;
  double* d = nullptr;
  MF_HIP(hipMalloc(&d, (size_t)count * 42 * sizeof(double)));
  MF_HIP(hipMemsetAsync(d, 0, (size_t)count * 42 * sizeof(double), stream()));
  // reference elements first..first+count-1 live at device slots ieperm[.]
  int* dl = nullptr;
  {
    std::vector<int> el(count);
    for (int i = 0; i < count; i++) el[i] = P->setup.ieperm[first + i];
    if (upload(&dl, el.data(), el.size())) {
      hipFree(d);
      return -2;
    }
  }
  int rc = launch_integrate(*P, precision, 1, d, 0, count, dl);
  if (rc == 0) {
    hipError_t e = hipMemcpyAsync(out, d, (size_t)count * 42 * sizeof(double),
                                  hipMemcpyDeviceToHost, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    if (e != hipSuccess) {
      set_error(hipGetErrorString(e));
      rc = -2;
    }
  }
  hipFree(d);
  hipFree(dl);
  return rc;
}

int mf_face_matrices(int pid, int first, int count, double* out) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (first < 0 || count < 0 || first + count > P->nbf || !out) {
    set_error("mf_face_matrices: range out of bounds");
    return -1;
  }
  if (count == 0) return 0;
  double* d = nullptr;
  MF_HIP(hipMalloc(&d, (size_t)count * 42 * sizeof(double)));
  MF_HIP(hipMemsetAsync(d, 0, (size_t)count * 42 * sizeof(double), stream()));
  int rc = launch_faces(*P, 1, d, first, count);
  if (rc == 0) {
    hipError_t e = hipMemcpyAsync(out, d, (size_t)count * 42 * sizeof(double),
                                  hipMemcpyDeviceToHost, stream());
    if (e == hipSuccess) e = hipStreamSynchronize(stream());
    if (e != hipSuccess) {
      set_error(hipGetErrorString(e));
      rc = -2;
    }
  }
  hipFree(d);
  return rc;
}

int mf_get_system(int pid, double* val, double* rhs) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  MF_HIP(hipStreamSynchronize(stream()));
  if (val) {
    std::vector<double> sv(P->nnz_sell);
    MF_HIP(hipMemcpy(sv.data(), P->d_val, sv.size() * sizeof(double), hipMemcpyDeviceToHost));
    sell_to_crs(P->setup, sv.data(), val);
  }
  if (rhs && rhs_to_host(*P, rhs)) return -2;
  return 0;
}

int mf_set_system(int pid, const double* val, const double* rhs) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  MF_HIP(hipStreamSynchronize(stream()));
  if (val) {
    std::vector<double> sv(P->nnz_sell);
    crs_to_sell(P->setup, val, sv.data());
    MF_HIP(hipMemcpy(P->d_val, sv.data(), sv.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  if (rhs) {
    std::vector<double> d = to_device_order(P->setup, rhs);
    MF_HIP(hipMemcpy(P->d_rhs, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  P->dinv_valid = false;  // refilled before the next preconditioned solve
  P->val32_valid = false;
  if (P->ilu) P->ilu->valid = false;
  if (P->gs) P->gs->valid = false;
  return 0;
}

int mf_spmv(int pid, const double* x, double* y) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  double *dx = nullptr, *dy = nullptr;
  MF_HIP(hipMalloc(&dx, (size_t)P->n * sizeof(double)));
  MF_HIP(hipMalloc(&dy, (size_t)P->n * sizeof(double)));
  int rc = 0;
  std::vector<double> xd = to_device_order(P->setup, x), yd(P->n);
  hipError_t e = hipMemcpyAsync(dx, xd.data(), (size_t)P->n * sizeof(double), hipMemcpyHostToDevice, stream());
  if (e == hipSuccess) rc = launch_spmv(*P, dx, dy);
  if (e == hipSuccess && rc == 0)
    e = hipMemcpyAsync(yd.data(), dy, (size_t)P->n * sizeof(double), hipMemcpyDeviceToHost, stream());
  if (e == hipSuccess) e = hipStreamSynchronize(stream());
  if (e == hipSuccess && rc == 0) to_reference_order(P->setup, yd.data(), y);
  if (e != hipSuccess) {
    set_error(hipGetErrorString(e));
    rc = -2;
  }
  hipFree(dx);
  hipFree(dy);
  return rc;
}

int mf_fill_precon(int pid, int precon) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (precon == 0) return 0;
  if (precon < 1 || precon > 3) {
    set_error("mf_fill_precon: 0 (none), 1 (Jacobi), 2 (ILU(0)) and 3 (Gauss-Seidel) are available");
    return -1;
  }
  if (precon == 3) {
    if (launch_fill_jacobi(*P)) return -2;
    if (!P->gs && gs_build(*P)) return -2;
    if (gs_refresh(*P)) return -2;
  } else if (precon == 2) {
    if (!P->ilu && ilu_build(*P, P->ilu_order)) return -2;
    if (ilu_factor(*P)) return -2;
  } else if (launch_fill_jacobi(*P)) {
    return -2;
  }
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_gmres(int pid, int precon, int krylbas, int max_iter, double tol, int ini_zero, double* x,
             int* iters, double* res_rel, double* resid) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (precon < 0 || precon > 3) {
    set_error("mf_gmres: precon must be 0 (none), 1 (Jacobi), 2 (ILU(0)) or 3 (Gauss-Seidel)");
    return -1;
  }
  if (max_iter < 1 || tol < 0) {
    set_error("mf_gmres: max_iter >= 1 and tol >= 0 required");
    return -1;
  }
  std::vector<double> xd;
  if (x) xd = ini_zero ? std::vector<double>(P->n, 0.0) : to_device_order(P->setup, x);
  int rc = gmres_solve(*P, precon, krylbas, max_iter, tol, 0.0, ini_zero, x ? xd.data() : nullptr,
                       iters, res_rel, resid);
  // rc: 1 converged, 0 not converged, <0 error
  if (rc >= 0 && x) to_reference_order(P->setup, xd.data(), x);
  return rc;
}

int mf_gmres_noalloc(int pid, int precon, int krylbas, int max_iter, double tol, int ini_zero,
                     double* x, int* iters, double* res_rel, double* resid) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  return gmres_noalloc(*P, precon, krylbas, max_iter, tol, ini_zero, x, iters, res_rel, resid);
}

int mf_mkb_read(const char* filename, int max_iter, int error_type, double error_tol,
                int monitoring, int* params, double* conv_meas) {
  if (!params || !conv_meas) {
    set_error("mf_mkb_read: params and conv_meas are required");
    return -1;
  }
  MkbParams p;
  if (mkb_read(filename, max_iter, error_type, error_tol, monitoring, p)) return -1;
  const int v[13] = {p.solver, p.nr_level, p.krylbas, p.precon, p.nr_prec, p.block_type,
                     p.storage_type, p.ilu_k, p.max_iter, p.conv_type, p.monitor,
                     p.nr_pre_smooth, p.nr_post_smooth};
  std::memcpy(params, v, sizeof(v));
  *conv_meas = p.conv_meas;
  return 0;
}

int mf_solve_mkb(int pid, const int* params, double conv_meas, int ini_zero, double* x,
                 int* iters, double* res_rel, double* resid) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (!params) {
    set_error("mf_solve_mkb: params is null");
    return -1;
  }
  const int solver = params[0], krylbas = params[2], nr_prec = params[4], ilu_k = params[7];
  const int storage = params[6], max_iter = params[8], conv_type = params[9];
  if (solver != 1) {
    set_error("mf_solve_mkb: only single-level GMRES (solver 1) is available");
    return -1;
  }
  // reference codes (lah_intf.h:45-50) -> this library's: NO_PRECON -> 0,
  // BLOCK_JACOBI with one sweep -> 1, BLOCK_ILU with k = 0 -> 2
  int precon = -1;
  if (params[3] == 0) precon = 0;
  else if (params[3] == 1 && nr_prec == 1) precon = 1;
  else if (params[3] == 5 && ilu_k == 0) precon = 2;
  else if (params[3] == 2 && nr_prec >= 1) precon = 3;  // BLOCK_GS, nr_prec sweeps
  if (precon < 0) {
    set_error("mf_solve_mkb: available preconditioners are none (0), one Jacobi sweep (1 1), "
              "block Gauss-Seidel (2 k) and ILU(0) as preconditioner (5 with k = 0)");
    return -1;
  }
  if (precon == 3) P->gs_sweeps = nr_prec;
  if (storage != -1 && storage != 1 && storage != 2) {
    set_error("mf_solve_mkb: storage must be undefined, block (size 1) or CRS");
    return -1;
  }
  if (conv_type < 0 || conv_type > 2) {
    set_error("mf_solve_mkb: conv_type must be 0, 1 or 2");
    return -1;
  }
  const double tol_rel = conv_type == 0 ? conv_meas : 0.0;
  const double tol_abs = conv_type == 1 ? conv_meas : 0.0;
  if (max_iter < 1) {
    set_error("mf_solve_mkb: max_iter >= 1 required");
    return -1;
  }
  std::vector<double> xd;
  if (x) xd = ini_zero ? std::vector<double>(P->n, 0.0) : to_device_order(P->setup, x);
  int rc = gmres_solve(*P, precon, krylbas, max_iter, tol_rel, tol_abs, ini_zero,
                       x ? xd.data() : nullptr, iters, res_rel, resid);
  if (rc >= 0 && x) to_reference_order(P->setup, xd.data(), x);
  return rc;
}

int mf_gmres_mixed(int pid, int precon, int krylbas, int max_iter, double tol, double* x,
                   int* iters, double* res_rel, double* resid, int* refinements) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (precon < 0 || precon > 2) {
    set_error("mf_gmres_mixed: precon must be 0 (none), 1 (Jacobi) or 2 (ILU(0))");
    return -1;
  }
  if (tol <= 0) {
    set_error("mf_gmres_mixed: tol > 0 required");
    return -1;
  }
  std::vector<double> xd(x ? P->n : 0);
  int rc = gmres_mixed(*P, precon, krylbas, max_iter, tol, x ? xd.data() : nullptr, iters,
                       res_rel, resid, refinements);
  if (rc >= 0 && x) to_reference_order(P->setup, xd.data(), x);
  return rc;
}

int mf_set_gs(int pid, int nsweeps) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (nsweeps < 1) {
    set_error("mf_set_gs: nsweeps >= 1 required");
    return -1;
  }
  P->gs_sweeps = nsweeps;
  return 0;
}

int mf_gs_get(int pid, int* elim_pos, int* ncolors) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (!P->gs && gs_build(*P)) return -2;
  if (elim_pos)
    for (int r = 0; r < P->n; r++) elim_pos[r] = P->gs->elim_pos[P->setup.iperm[r]];
  if (ncolors) *ncolors = P->gs->ncolors;
  return 0;
}

int mf_set_ilu_order(int pid, int order) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (order != 0 && order != 1) {
    set_error("mf_set_ilu_order: 0 (reference DOF order) or 1 (multicolour)");
    return -1;
  }
  if (P->batch_refs > 0) {
    set_error("mf_set_ilu_order: the problem is a member of a batch");
    return -1;
  }
  P->ilu_order = order;
  ilu_free(*P);
  return 0;
}

int mf_ilu_get(int pid, double* lu, int* elim_pos, int* levels) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (!P->ilu && ilu_build(*P, P->ilu_order)) return -2;
  if (!P->ilu->valid && ilu_factor(*P)) return -2;
  IluData& L = *P->ilu;
  MF_HIP(hipStreamSynchronize(stream()));
  if (lu) {
    std::vector<double> h(L.nnz);
    MF_HIP(hipMemcpy(h.data(), L.d_lu, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (long long q = 0; q < L.nnz; q++) lu[L.ref_pos[q]] = h[q];
  }
  if (elim_pos)
    for (int r = 0; r < P->n; r++) elim_pos[r] = L.pos_of_dev[P->setup.iperm[r]];
  if (levels) {
    levels[0] = L.nlev_l;
    levels[1] = L.nlev_u;
    levels[2] = L.ncolors;
  }
  return 0;
}

int mf_get_solution(int pid, double* x) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (!P->d_x || !x) {
    set_error("mf_get_solution: no solution (run mf_gmres first) or x is null");
    return -1;
  }
  std::vector<double> d(P->n);
  MF_HIP(hipStreamSynchronize(stream()));
  MF_HIP(hipMemcpy(d.data(), P->d_x, d.size() * sizeof(double), hipMemcpyDeviceToHost));
  to_reference_order(P->setup, d.data(), x);
  return 0;
}

int mf_true_residual(int pid, const double* x, double* rel) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  std::vector<double> y(P->n), b(P->n);
  int rc = mf_spmv(pid, x, y.data());
  if (rc) return rc;
  if (rhs_to_host(*P, b.data())) return -2;
  double nr = 0, nb = 0;
  for (int i = 0; i < P->n; i++) {
    double r = b[i] - y[i];
    nr += r * r;
    nb += b[i] * b[i];
  }
  *rel = nb > 0 ? std::sqrt(nr / nb) : std::sqrt(nr);
  return 0;
}

int mf_bench_integrate(int pid, int precision, int reps, double* ms_elems, double* ms_faces) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (reps < 1) reps = 1;
  hipEvent_t a, b, c;
  MF_HIP(hipEventCreate(&a));
  MF_HIP(hipEventCreate(&b));
  MF_HIP(hipEventCreate(&c));
  float te = 0, tf = 0;
  for (int r = 0; r < reps; r++) {
    const int mode = effective_mode(*P);
    if (mode == 0 || mode == 2) {
      MF_HIP(hipEventRecord(a, stream()));
      if (mode == 0 ? launch_assemble_rows(*P, precision)
                    : launch_assemble_two_pass(*P, precision))
        return -2;
      MF_HIP(hipEventRecord(b, stream()));
      MF_HIP(hipEventRecord(c, stream()));
    } else {
      if (launch_zero(P->d_val, P->nnz_sell) || launch_zero(P->d_rhs, P->n)) return -2;
      MF_HIP(hipEventRecord(a, stream()));
      if (launch_integrate(*P, precision, 0, nullptr, 0, P->nel)) return -2;
      MF_HIP(hipEventRecord(b, stream()));
      if (launch_faces(*P, 0, nullptr, 0, P->nbf)) return -2;
      MF_HIP(hipEventRecord(c, stream()));
    }
    MF_HIP(hipEventSynchronize(c));
    float x = 0, y = 0;
    hipEventElapsedTime(&x, a, b);
    hipEventElapsedTime(&y, b, c);
    te += x;
    tf += y;
  }
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipEventDestroy(c);
  // the system was re-assembled: derived data is remade before the next use
  P->dinv_valid = false;
  P->val32_valid = false;
  if (P->ilu) P->ilu->valid = false;
  if (P->gs) P->gs->valid = false;
  if (ms_elems) *ms_elems = te / reps;
  if (ms_faces) *ms_faces = tf / reps;
  return 0;
}

int mf_bench_spmv(int pid, int reps, double* ms) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  double *dx = nullptr, *dy = nullptr;
  MF_HIP(hipMalloc(&dx, (size_t)P->n * sizeof(double)));
  MF_HIP(hipMalloc(&dy, (size_t)P->n * sizeof(double)));
  if (launch_zero(dx, P->n)) return -2;
  hipEvent_t a, b;
  MF_HIP(hipEventCreate(&a));
  MF_HIP(hipEventCreate(&b));
  if (launch_spmv(*P, dx, dy)) return -2;
  MF_HIP(hipEventRecord(a, stream()));
  for (int r = 0; r < reps; r++)
    if (launch_spmv(*P, dx, dy)) return -2;
  MF_HIP(hipEventRecord(b, stream()));
  MF_HIP(hipEventSynchronize(b));
  float t = 0;
  hipEventElapsedTime(&t, a, b);
  *ms = t / std::max(reps, 1);
  hipEventDestroy(a);
  hipEventDestroy(b);
  hipFree(dx);
  hipFree(dy);
  return 0;
}

int mf_bench_spmv_ex(int pid, int prec, int ch, int reps, double* ms) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (!ms) {
    set_error("mf_bench_spmv_ex: ms is null");
    return -1;
  }
  return bench_spmv_variant(*P, prec, ch, reps, ms);
}

int mf_bench_copy(long long nbytes, int reps, double* ms) {
  if (need_init()) return -1;
  if (!ms) {
    set_error("mf_bench_copy: ms is null");
    return -1;
  }
  return bench_copy(nbytes, reps, ms);
}

int mf_last_solve_ms(int pid, double* ms) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  *ms = P->t_solve_ms;
  return 0;
}

int mf_set_assembly_mode(int pid, int mode) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (mode < 0 || mode > 2) {
    set_error("assembly mode: 0 row-wise, 1 element atomics, 2 two-pass (element rows + row gather)");
    return -1;
  }
  P->assembly_mode = mode;
  return 0;
}

int mf_gmres_kernel_stats(int enable, double* out) {
  if (out) kstats_get(out);
  kstats_enable(enable);
  return 0;
}

int mf_batch_create(int krylbas) {
  if (need_init()) return -1;
  if (krylbas < 1 || krylbas > 63) {
    set_error("mf_batch_create: krylbas must be in [1,63]");
    return -1;
  }
  auto B = std::make_unique<Batch>();
  B->m = krylbas;
  int id = g_next_batch++;
  g_batches[id] = std::move(B);
  return id;
}

static Batch* get_batch(int bid) {
  auto it = g_batches.find(bid);
  if (it == g_batches.end()) {
    set_error("unknown batch handle " + std::to_string(bid));
    return nullptr;
  }
  return it->second.get();
}

static bool batch_idle(const Batch* B) {
  if (B->pending) set_error("batch: a cycle is in flight (mf_batch_wait first)");
  return !B->pending;
}

int mf_batch_add(int bid, int pid, int precon, int max_iter, double tol) {
  Batch* B = get_batch(bid);
  Problem* P = get_problem(pid);
  if (!B || !P) return -3;
  if (!batch_idle(B)) return -1;
  return batch_add(*B, P, pid, precon, max_iter, tol);
}

int mf_batch_add_ex(int bid, int pid, int precon, int max_iter, double tol, int inner_precision) {
  Batch* B = get_batch(bid);
  Problem* P = get_problem(pid);
  if (!B || !P) return -3;
  if (inner_precision != 64 && inner_precision != 32) {
    set_error("mf_batch_add_ex: inner_precision must be 64 or 32");
    return -1;
  }
  if (!batch_idle(B)) return -1;
  return batch_add(*B, P, pid, precon, max_iter, tol, inner_precision == 32);
}

int mf_batch_add_dg(int bid, int did, int max_iter, double tol, int inner_precision) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  if (!batch_idle(B)) return -1;
  if (inner_precision != 64 && inner_precision != 32) {
    set_error("mf_batch_add_dg: inner_precision must be 64 or 32");
    return -1;
  }
  Problem* P = dg_batch_problem(did);
  if (!P) return -2;
  // members are told apart by handle: a DG problem joins as -did
  return batch_add(*B, P, -did, 0, max_iter, tol, inner_precision == 32);
}

int mf_batch_remove_dg(int bid, int did) { return mf_batch_remove(bid, -did); }

int mf_batch_remove(int bid, int pid) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  if (!batch_idle(B)) return -1;
  return batch_remove(*B, pid);
}

int mf_batch_integrate_assemble(int bid, int precision) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  if (!batch_idle(B)) return -1;
  if (precision != 64 && precision != 32) {
    set_error("precision must be 64 or 32");
    return -1;
  }
  // row-wise and two-pass members: one launch per kernel for all of them
  std::vector<Problem*> many;
  for (auto& M : B->mem) {
    if (M.P->nel <= 0) continue;
    const int mode = effective_mode(*M.P);
    if (mode == 0 || mode == 2) {
      many.push_back(M.P);
    } else if (assemble_problem(*M.P, precision)) {
      return -2;
    }
  }
  if (launch_assemble_many(many, precision)) return -2;
  for (Problem* P : many) assembled(*P);
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int mf_batch_run(int bid, int max_cycles, int use_graph, int* active) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  if (max_cycles < 0) {
    set_error("mf_batch_run: max_cycles >= 0 required");
    return -1;
  }
  return batch_run(*B, max_cycles, use_graph, active);
}

int mf_batch_set_stream(int bid) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  if (!batch_idle(B)) return -1;
  return batch_own_stream(*B);
}

int mf_batch_run_async(int bid, int use_graph) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  return batch_run_async(*B, use_graph);
}

int mf_batch_wait(int bid, int* active) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  return batch_wait(*B, active);
}

int mf_batch_set_tiles(int bid, int spb) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  if (!batch_idle(B)) return -1;
  return batch_set_tiles(*B, spb);
}

int mf_batch_tiles(int bid, int* spb, int* tuned, double* trial_ms) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  if (spb) *spb = B->spb_used;
  if (tuned) *tuned = B->spb_tuned ? 1 : 0;
  if (trial_ms)
    for (int k = 0; k < kBatchTileNum; k++) trial_ms[k] = B->spb_tuned ? B->tile_ms[k] : 0.0;
  return kBatchTileNum;
}

int mf_batch_tile_candidates(int* spb) {
  if (spb)
    for (int k = 0; k < kBatchTileNum; k++) spb[k] = kBatchTileCands[k];
  return kBatchTileNum;
}

int mf_batch_tile_cache_clear(void) {
  batch_tile_cache_clear();
  return 0;
}

int mf_batch_status_ex(int bid, int pid, int* done, int* converged, int* iters, double* res_rel,
                       int* refinements) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  if (!batch_idle(B)) return -1;
  for (size_t k = 0; k < B->mem.size(); k++)
    if (B->mem[k].pid == pid) {
      const GmresStat& st = B->h_stat[k];
      const MixedState& M = B->h_mix[k];
      if (M.mixed) {
        // refinement state: FP64 measure rho / rho_0, inner iterations so far
        if (done) *done = M.fin ? 1 : 0;
        if (converged) *converged = M.conv ? 1 : 0;
        if (iters) *iters = M.total + (M.phase == 0 && !M.fin ? st.gmiter : 0);
        if (res_rel) *res_rel = M.rho0 > 0 ? M.rho / M.rho0 : 0.0;
        if (refinements) *refinements = M.outer;
        return M.nan ? -4 : 0;
      }
      if (done) *done = st.done;
      if (converged) *converged = st.converged;
      if (iters) *iters = st.gmiter;
      if (res_rel) *res_rel = st.res_rel;
      if (refinements) *refinements = 0;
      return st.nan ? -4 : 0;
    }
  set_error("mf_batch_status: problem is not a member");
  return -1;
}

int mf_batch_status(int bid, int pid, int* done, int* converged, int* iters, double* res_rel) {
  return mf_batch_status_ex(bid, pid, done, converged, iters, res_rel, nullptr);
}

int mf_batch_status_dg(int bid, int did, int* done, int* converged, int* iters, double* res_rel,
                       int* refinements) {
  return mf_batch_status_ex(bid, -did, done, converged, iters, res_rel, refinements);
}

int mf_batch_free(int bid) {
  Batch* B = get_batch(bid);
  if (!B) return -3;
  batch_free(*B);
  g_batches.erase(bid);
  return 0;
}

int mf_problem_free(int pid) {
  Problem* P = get_problem(pid);
  if (!P) return -3;
  if (P->batch_refs > 0) {
    set_error("mf_problem_free: the problem is a member of a batch (mf_batch_remove first)");
    return -1;
  }
  hipStreamSynchronize(stream());
  free_problem_device(*P);
  g_problems.erase(pid);
  return 0;
}

}  // extern "C"
