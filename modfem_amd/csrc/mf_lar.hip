// Storage/operations seam with the reference's own names and signatures
// (lsd_mkb/lah_intf.h:90-289, dispatcher lsd_mkb/las_intf.c): block size 1
// (LAC_STORAGE_CRS) and 4x4 blocks (LAC_STORAGE_BCRS, chosen as in
// las_intf.c:90-97).  A ModFEM build can link these in place of las_intf.c +
// lad_crs + lad_bcrs: the matrix and vectors live in device memory and every
// operation runs on the GPU; vectors passed in may be host or device pointers.
//
// Per-element assembly calls (lar_assemble_SM_and_LV[_with_table]) come one
// element at a time from the host, so they accumulate into a host mirror of
// the values in call order -- the same order and rounding as
// lar_assemble_SM_and_LV_with_table_crs (lad_crs/las_crs_intf.c:440-513) and
// of the BCRS variants (lad_bcrs/las_bcrs_intf.c:407-623) -- and the mirror
// is uploaded before the next device operation.  Device-side batched assembly
// is mf_integrate_assemble (include/modfem_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "mf_device.h"

using namespace mf;

#define LAC_SCOPE_LV 1
#define LAC_SCOPE_SM 2
#define LAC_SCOPE_SM_AND_LV 0
#define NO_PRECON 0
#define BLOCK_JACOBI 1
#define BLOCK_GS 2
#define MULTI_ILU 4
#define BLOCK_ILU 5
#define LAC_STORAGE_UNDEFINED -1
#define LAC_STORAGE_BLOCK 1
#define LAC_STORAGE_CRS 2
#define LAC_STORAGE_BCRS 3
#define LAC_MAX_MATRICES 20
#define BLS 4

namespace {

struct LarMatrix {
  std::unique_ptr<Problem> P;   // CRS, block size 1
  std::unique_ptr<BcrsDev> B;   // BCRS, 4x4 blocks
  std::vector<int> rp, ci, diag;  // BCRS block pattern on the host
  std::vector<int> ilu_rp, ilu_ci;  // BCRS ILU(k > 0) pattern on the host
  std::vector<double> h_val, h_rhs;  // host mirror for per-element assembly
  bool dirty = false;
  bool filled = false;  // diagonal-block factors / diagonal inverses current
  int precon = -1;
  int gs_threads = 1;   // CRS BLOCK_GS: the reference's OpenMP thread count
  int ndof() const { return B ? BLS * B->nb : P->n; }
  // BCRS GMRES: workspace (Krylov basis, solution) and the operator
  std::unique_ptr<Problem> G;
  std::unique_ptr<ExtOp> op;
  int op_kind = -1, op_sweeps = -1;
};

std::vector<LarMatrix> g_mats;  // LIFO like itv_crs_matrices

int lar_bcrs_gmres(LarMatrix& L, int krylbas, int max_iter, double tol, double tol_abs,
                   int nr_prec, int ini_zero, double* x, int* iters, double* res_rel,
                   double* resid);
int lar_crs_gs_gmres(LarMatrix& L, int krylbas, int max_iter, double tol, double tol_abs,
                     int nr_prec, int ini_zero, double* x, int* iters, double* res_rel,
                     double* resid);

LarMatrix* lar_get(int id) {
  if (id < 0 || id >= (int)g_mats.size() || !(g_mats[id].P || g_mats[id].B)) {
    set_error("lar_: unknown matrix id " + std::to_string(id));
    return nullptr;
  }
  return &g_mats[id];
}

int flush(LarMatrix& M) {
  if (!M.dirty) return 0;
  double* d_val = M.B ? M.B->d_val : M.P->d_val;
  double* d_rhs = M.B ? M.B->d_rhs : M.P->d_rhs;
  MF_HIP(hipMemcpyAsync(d_val, M.h_val.data(), M.h_val.size() * sizeof(double),
                        hipMemcpyHostToDevice, stream()));
  MF_HIP(hipMemcpyAsync(d_rhs, M.h_rhs.data(), M.h_rhs.size() * sizeof(double),
                        hipMemcpyHostToDevice, stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  M.dirty = false;
  return 0;
}

bool is_device_ptr(const void* p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

// stage a vector of n doubles on the device (copy when it is host memory)
struct Staged {
  double* d = nullptr;
  bool owned = false;
  ~Staged() {
    if (owned && d) (void)hipFree(d);
  }
};

int stage_in(const double* p, int n, Staged& s, bool copy) {
  if (!p) return 0;
  if (is_device_ptr(p)) {
    s.d = const_cast<double*>(p);
    return 0;
  }
  MF_HIP(hipMalloc(&s.d, (size_t)n * sizeof(double)));
  s.owned = true;
  if (copy)
    MF_HIP(hipMemcpyAsync(s.d, p, (size_t)n * sizeof(double), hipMemcpyHostToDevice, stream()));
  return 0;
}

int stage_out(double* host, const Staged& s, int n) {
  if (s.owned)
    MF_HIP(hipMemcpyAsync(host, s.d, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, stream()));
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

// block pattern from neighbour lists (las_crs_intf.c:184-250,
// las_bcrs_intf.c:210-262): rows = blocks, columns = the sorted off-diagonal
// blocks with the diagonal inserted in place; ghost blocks (Nroffbl <= 0)
// have no stored row
int block_pattern(int n, const int* Nroffbl, int* const* L_offbl, std::vector<int>& rp,
                  std::vector<int>& ci) {
  rp.assign(n + 1, 0);
  ci.clear();
  for (int ib = 0; ib < n; ib++) {
    rp[ib] = (int)ci.size();
    if (Nroffbl[ib] <= 0) continue;
    bool inserted = false;
    for (int j = 0; j < Nroffbl[ib]; j++) {
      int c = L_offbl[ib][j];
      if (c < 0 || c >= n || c == ib || (j > 0 && L_offbl[ib][j - 1] >= c)) {
        set_error("lar_allocate_SM_and_LV: neighbour lists must be strictly ascending block "
                  "ids in range, without the block itself");
        return -1;
      }
      if (!inserted && c > ib) {
        ci.push_back(ib);
        inserted = true;
      }
      ci.push_back(c);
    }
    if (!inserted) ci.push_back(ib);
  }
  rp[n] = (int)ci.size();
  return 0;
}

int allocate_bcrs(int nb, const int* Nroffbl, int* const* L_offbl) {
  LarMatrix M;
  if (block_pattern(nb, Nroffbl, L_offbl, M.rp, M.ci)) return -1;
  M.B = std::make_unique<BcrsDev>();
  BcrsDev& B = *M.B;
  B.nb = nb;
  B.nblk = (long long)M.ci.size();
  for (int r = 0; r < nb; r++) B.max_row = std::max(B.max_row, M.rp[r + 1] - M.rp[r]);
  MF_HIP(hipMalloc(&B.d_rp, M.rp.size() * sizeof(int)));
  MF_HIP(hipMalloc(&B.d_ci, std::max<size_t>(M.ci.size(), 1) * sizeof(int)));
  MF_HIP(hipMalloc(&B.d_val, std::max<size_t>(M.ci.size(), 1) * BLS * BLS * sizeof(double)));
  MF_HIP(hipMalloc(&B.d_rhs, std::max(nb, 1) * BLS * sizeof(double)));
  MF_HIP(hipMemcpy(B.d_rp, M.rp.data(), M.rp.size() * sizeof(int), hipMemcpyHostToDevice));
  if (!M.ci.empty())
    MF_HIP(hipMemcpy(B.d_ci, M.ci.data(), M.ci.size() * sizeof(int), hipMemcpyHostToDevice));
  M.h_val.assign(M.ci.size() * BLS * BLS, 0.0);
  M.h_rhs.assign((size_t)nb * BLS, 0.0);
  M.dirty = true;
  g_mats.push_back(std::move(M));
  if (flush(g_mats.back())) return -2;
  return (int)g_mats.size() - 1;
}

void free_gmres(LarMatrix& M) {
  if (M.G) {
    Problem& P = *M.G;
    void* ptrs[] = {P.d_V, P.d_H, P.d_partials, P.d_stat, P.d_x};
    for (void* q : ptrs)
      if (q) (void)hipFree(q);
    if (P.h_stat) (void)hipHostFree(P.h_stat);
    M.G.reset();
  }
  M.op.reset();
  M.op_kind = M.op_sweeps = -1;
}

void free_bcrs_precon(BcrsDev& B) {
  void* ptrs[] = {B.d_diag, B.d_lu, B.d_piv, B.d_ilu, B.d_lev_rows[0], B.d_lev_rows[1],
                  B.d_ilu_rp, B.d_ilu_ci, B.d_ilu_diag, B.d_ilu_src, B.d_ilu_lev_rows[0],
                  B.d_ilu_lev_rows[1]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  B.d_diag = nullptr;
  B.d_lu = nullptr;
  B.d_piv = nullptr;
  B.d_ilu = nullptr;
  B.d_ilu_rp = B.d_ilu_ci = B.d_ilu_diag = B.d_ilu_src = nullptr;
  B.ilu_k = 0;
  B.ilu_nblk = 0;
  B.ilu_max_row = 0;
  for (int d = 0; d < 2; d++) {
    B.d_lev_rows[d] = nullptr;
    B.lev_ptr[d].clear();
    B.d_ilu_lev_rows[d] = nullptr;
    B.ilu_lev_ptr[d].clear();
  }
}

}  // namespace

extern "C" {

// lar_allocate_SM_and_LV (lah_intf.h:90; storage choice las_intf.c:90-97;
// CRS branch las_crs_intf.c:72-294, BCRS branch las_bcrs_intf.c:109-277).
// Returns the matrix id.
int lar_allocate_SM_and_LV(int Storage_type, int solver_type, int Nrdof_glob, int Max_SM_size,
                           int Nrblocks, int Block_size, int* Nrdofbl, int* Posglob,
                           int* Nroffbl, int** L_offbl) {
  (void)solver_type; (void)Max_SM_size; (void)Nrdofbl; (void)Posglob;
  int storage = Storage_type;
  if (storage == LAC_STORAGE_UNDEFINED)
    storage = Block_size == 1 ? LAC_STORAGE_CRS : Block_size > 1 ? LAC_STORAGE_BCRS : -1;
  // block storage (lad_block, the reference's choice for block-type > 0
  // control files) with a constant block size is held as CRS (1x1 blocks)
  // or BCRS (4x4 blocks): the same matrix and operations, whose sums run in
  // column order instead of lad_block's diagonal-block-first order
  if (storage == LAC_STORAGE_BLOCK && (Block_size == 1 || Block_size == BLS))
    storage = Block_size == 1 ? LAC_STORAGE_CRS : LAC_STORAGE_BCRS;
  if (storage != LAC_STORAGE_CRS && storage != LAC_STORAGE_BCRS) {
    set_error("lar_allocate_SM_and_LV: CRS (block size 1) and BCRS (block size 4) storage, and "
              "block storage of constant size 1 or 4, are available");
    return -1;
  }
  if (storage == LAC_STORAGE_CRS && (Block_size != 1 || Nrblocks != Nrdof_glob)) {
    set_error("lar_allocate_SM_and_LV: CRS storage needs block size 1");
    return -1;
  }
  if (storage == LAC_STORAGE_BCRS && (Block_size != BLS || Nrdof_glob != Nrblocks * BLS)) {
    set_error("lar_allocate_SM_and_LV: BCRS storage needs block size 4 and Nrdof_glob = 4 Nrblocks");
    return -1;
  }
  // a ModFEM build calls no mf_init: open device MF_DEVICE (default 0) here
  if (!stream()) {
    const char* dev = std::getenv("MF_DEVICE");
    if (device_init(dev ? std::atoi(dev) : 0)) return -2;
  }
  if ((int)g_mats.size() >= LAC_MAX_MATRICES) {
    set_error("lar_allocate_SM_and_LV: too many matrices");
    return -1;
  }
  if (storage == LAC_STORAGE_BCRS) return allocate_bcrs(Nrblocks, Nroffbl, L_offbl);
  const int n = Nrdof_glob;
  std::vector<int> rp, ci;
  if (block_pattern(n, Nroffbl, L_offbl, rp, ci)) return -1;
  LarMatrix M;
  M.P = std::make_unique<Problem>();
  Problem& P = *M.P;
  P.n = n;
  P.setup.n = n;
  P.setup.row_ptr = rp;
  P.setup.col_ind = ci;
  P.setup.nnz = P.nnz = (long long)ci.size();
  MF_HIP(hipMalloc(&P.d_row_ptr, rp.size() * sizeof(int)));
  MF_HIP(hipMalloc(&P.d_col, std::max<size_t>(ci.size(), 1) * sizeof(int)));
  MF_HIP(hipMalloc(&P.d_val, std::max<size_t>(ci.size(), 1) * sizeof(double)));
  MF_HIP(hipMalloc(&P.d_rhs, (size_t)n * sizeof(double)));
  MF_HIP(hipMemcpy(P.d_row_ptr, rp.data(), rp.size() * sizeof(int), hipMemcpyHostToDevice));
  MF_HIP(hipMemcpy(P.d_col, ci.data(), ci.size() * sizeof(int), hipMemcpyHostToDevice));
  if (build_tiles(P) < 0) return -1;
  M.h_val.assign(ci.size(), 0.0);
  M.h_rhs.assign(n, 0.0);
  M.dirty = true;
  g_mats.push_back(std::move(M));
  if (flush(g_mats.back())) return -2;
  return (int)g_mats.size() - 1;
}

// lar_initialize_SM_and_LV (las_crs_intf.c:256-288): rhs always, values for
// LAC_SCOPE_SM_AND_LV
int lar_initialize_SM_and_LV(int Matrix_id, int Scope) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  std::fill(M->h_rhs.begin(), M->h_rhs.end(), 0.0);
  if (Scope == LAC_SCOPE_SM_AND_LV) std::fill(M->h_val.begin(), M->h_val.end(), 0.0);
  M->dirty = true;
  return 0;
}

double lar_get_storage(int Matrix_id) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -1;
  if (M->B)  // values, column indices, row pointers, rhs (the reference returns -1)
    return ((double)M->B->nblk * (BLS * BLS * 8 + 4) + (double)M->B->nb * (BLS * 8 + 4)) /
           1024.0 / 1024.0;
  const Problem& P = *M->P;
  return ((double)P.nnz * 12 + (double)P.n * 20) / 1024.0 / 1024.0;
}

namespace {

// position of block (row, col) in the BCRS pattern, -1 when absent
int bcrs_find(const LarMatrix& M, int row, int col) {
  if (row < 0 || row >= M.B->nb) return -1;
  for (int s = M.rp[row]; s < M.rp[row + 1]; s++)
    if (M.ci[s] == col) return s;
  return -1;
}

bool blocks_in_range(const LarMatrix& M, int k, const int* L_bl_id) {
  for (int i = 0; i < k; i++)
    if (L_bl_id[i] < 0 || L_bl_id[i] >= M.B->nb) {
      set_error("lar_: block id out of range");
      return false;
    }
  return true;
}

}  // namespace

// lar_fill_assembly_table_int_ent (las_crs_intf.c:354-432)
int lar_fill_assembly_table_int_ent(int Matrix_id, int Nr_dof_bl, int* L_bl_id, int* L_bl_nrdof,
                                    int* Assembly_table_int_ent) {
  (void)L_bl_nrdof;
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (M->B) {  // las_bcrs_intf.c:325-399: entry jb + ib k = 16 * (position of
               // block (L_bl_id[ib], L_bl_id[jb])), then 4 * L_bl_id[jb]
    const int k = Nr_dof_bl;
    if (!blocks_in_range(*M, k, L_bl_id)) return -1;
    for (int ib = 0; ib < k; ib++)
      for (int jb = 0; jb < k; jb++) {
        const int s = bcrs_find(*M, L_bl_id[ib], L_bl_id[jb]);
        if (s < 0) {
          set_error("lar_fill_assembly_table_int_ent: block pair not in the pattern");
          return -1;
        }
        Assembly_table_int_ent[jb + ib * k] = s * BLS * BLS;
      }
    for (int jb = 0; jb < k; jb++) Assembly_table_int_ent[k * k + jb] = L_bl_id[jb] * BLS;
    return 1;
  }
  const Setup& S = M->P->setup;
  for (int i = 0; i < Nr_dof_bl; i++) {
    int col = L_bl_id[i];
    for (int j = 0; j < Nr_dof_bl; j++) {
      int row = L_bl_id[j];
      int found = -1;
      for (int c = S.row_ptr[row]; c < S.row_ptr[row + 1]; c++)
        if (S.col_ind[c] == col) {
          found = c;
          break;
        }
      if (found < 0) {
        set_error("lar_fill_assembly_table_int_ent: block pair not in the pattern");
        return -1;
      }
      Assembly_table_int_ent[j + i * Nr_dof_bl] = found;
    }
  }
  for (int j = 0; j < Nr_dof_bl; j++) Assembly_table_int_ent[Nr_dof_bl * Nr_dof_bl + j] = L_bl_id[j];
  return 1;
}

// lar_assemble_SM_and_LV_with_table (las_crs_intf.c:440-513)
int lar_assemble_SM_and_LV_with_table(int Matrix_id, int Scope, int Nr_dof_bl,
                                      int* Assembly_table_int_ent, double* Stiff_mat,
                                      double* Rhs_vect, char* Rewr_dofs) {
  (void)Scope; (void)Rewr_dofs;
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  const int k = Nr_dof_bl;
  if (M->B) {  // las_bcrs_intf.c:407-523: block (ib, jb) of the column-major
               // 4k x 4k element matrix added to the stored block
    const int ld = k * BLS;
    const long long nval = (long long)M->h_val.size(), nrhs = (long long)M->h_rhs.size();
    for (int ib = 0; ib < k; ib++)
      for (int jb = 0; jb < k; jb++) {
        const int pos = Assembly_table_int_ent[jb + ib * k];
        if (pos < 0 || pos + BLS * BLS > nval) {
          set_error("lar_assemble_SM_and_LV_with_table: table entry out of range");
          return -1;
        }
        const double* src = Stiff_mat + (jb * ld + ib) * BLS;
        for (int c = 0; c < BLS; c++)
          for (int r = 0; r < BLS; r++) M->h_val[pos + c * BLS + r] += src[c * ld + r];
      }
    for (int ib = 0; ib < k; ib++) {
      const int pos = Assembly_table_int_ent[k * k + ib];
      if (pos < 0 || pos + BLS > nrhs) {
        set_error("lar_assemble_SM_and_LV_with_table: table entry out of range");
        return -1;
      }
      for (int r = 0; r < BLS; r++) M->h_rhs[pos + r] += Rhs_vect[ib * BLS + r];
    }
    M->dirty = true;
    return 1;
  }
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) M->h_val[Assembly_table_int_ent[j + i * k]] += Stiff_mat[j + i * k];
  for (int i = 0; i < k; i++) M->h_rhs[Assembly_table_int_ent[k * k + i]] += Rhs_vect[i];
  M->dirty = true;
  return 1;
}

// lar_assemble_SM_and_LV (las_crs_intf.c:520-650): table-less variant with a
// column search per entry
int lar_assemble_SM_and_LV(int Matrix_id, int Scope, int Nr_dof_bl, int* L_bl_id,
                           int* L_bl_nrdof, double* Stiff_mat, double* Rhs_vect,
                           char* Rewr_dofs) {
  (void)Scope; (void)Rewr_dofs; (void)L_bl_nrdof;
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  const int k = Nr_dof_bl;
  if (M->B) {  // las_bcrs_intf.c:530-623: blocks missing from the pattern
               // (ghost rows) are skipped, the rhs is added for every block
    if (!blocks_in_range(*M, k, L_bl_id)) return -1;
    const int ld = k * BLS;
    for (int ib = 0; ib < k; ib++) {
      for (int jb = 0; jb < k; jb++) {
        const int s = bcrs_find(*M, L_bl_id[ib], L_bl_id[jb]);
        if (s < 0) continue;
        const double* src = Stiff_mat + (jb * ld + ib) * BLS;
        double* dst = M->h_val.data() + (size_t)s * BLS * BLS;
        for (int c = 0; c < BLS; c++)
          for (int r = 0; r < BLS; r++) dst[c * BLS + r] += src[c * ld + r];
      }
      for (int r = 0; r < BLS; r++) M->h_rhs[L_bl_id[ib] * BLS + r] += Rhs_vect[ib * BLS + r];
    }
    M->dirty = true;
    return 1;
  }
  const Setup& S = M->P->setup;
  for (int i = 0; i < k; i++) {
    int col = L_bl_id[i];
    for (int j = 0; j < k; j++) {
      int row = L_bl_id[j];
      for (int c = S.row_ptr[row]; c < S.row_ptr[row + 1]; c++)
        if (S.col_ind[c] == col) {
          M->h_val[c] += Stiff_mat[j + i * k];
          break;
        }
    }
  }
  for (int i = 0; i < k; i++) M->h_rhs[L_bl_id[i]] += Rhs_vect[i];
  M->dirty = true;
  return 1;
}

int lar_free_SM_and_LV(int Matrix_id) {
  if (Matrix_id != (int)g_mats.size() - 1) {
    set_error("lar_free_SM_and_LV: matrices must be freed in reverse order of creation");
    return -1;
  }
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (M->B) {
    BcrsDev& B = *M->B;
    free_gmres(*M);
    free_bcrs_precon(B);
    void* ptrs[] = {B.d_rp, B.d_ci, B.d_val, B.d_rhs, B.d_col_rows};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    g_mats.pop_back();
    return 0;
  }
  Problem& P = *M->P;
  void* ptrs[] = {P.d_row_ptr, P.d_col, P.d_val, P.d_rhs, P.d_dinv, P.d_diag_ptr, P.d_tile_row,
                  P.d_V, P.d_H, P.d_partials, P.d_stat, P.d_x};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (P.h_stat) (void)hipHostFree(P.h_stat);
  g_mats.pop_back();
  return 0;
}

// lar_allocate_preconditioner (las_crs_intf.c:658-700): Jacobi or none;
// BCRS (las_bcrs_intf.c:630-1016): block Jacobi or block Gauss-Seidel with
// the diagonal blocks' LU factors, or block ILU(k) (MULTI_ILU / BLOCK_ILU with
// Second_arg = k): k = 0 on the matrix pattern, k > 0 on the pattern grown
// by k rings of neighbours; the matrix pattern must be structurally
// symmetric.  The level schedules of the sweeps are built here
int lar_allocate_preconditioner(int Matrix_id, int Precon, int Second_arg) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (M->B) {
    const bool ilu = Precon == MULTI_ILU || Precon == BLOCK_ILU;
    if (Precon != NO_PRECON && Precon != BLOCK_JACOBI && Precon != BLOCK_GS && !ilu) {
      set_error("lar_allocate_preconditioner: BCRS storage offers NO_PRECON, BLOCK_JACOBI, "
                "BLOCK_GS, MULTI_ILU and BLOCK_ILU");
      return -1;
    }
    if (ilu && Second_arg < 0) {
      set_error("lar_allocate_preconditioner: ILU(k) needs k >= 0");
      return -1;
    }
    BcrsDev& B = *M->B;
    M->op.reset();
    M->op_kind = -1;
    free_bcrs_precon(B);
    M->filled = false;
    M->precon = Precon;
    if (Precon == NO_PRECON) return B.nb;
    M->diag.assign(B.nb, -1);
    for (int r = 0; r < B.nb; r++) {
      for (int s = M->rp[r]; s < M->rp[r + 1]; s++)
        if (M->ci[s] == r) M->diag[r] = s;
      if (M->diag[r] < 0) {
        set_error("lar_allocate_preconditioner: block row " + std::to_string(r) +
                  " has no stored diagonal block");
        return -1;
      }
    }
    if (ilu)  // every stored block (k, i) needs its mirror (i, k)
      for (int k = 0; k < B.nb; k++)
        for (int s = M->rp[k]; s < M->rp[k + 1]; s++) {
          const int i = M->ci[s];
          bool found = false;
          for (int t = M->rp[i]; t < M->rp[i + 1] && !found; t++) found = M->ci[t] == k;
          if (!found) {
            set_error("lar_allocate_preconditioner: block ILU needs a structurally symmetric "
                      "pattern");
            return -1;
          }
        }
    MF_HIP(hipMalloc(&B.d_diag, std::max(B.nb, 1) * sizeof(int)));
    MF_HIP(hipMemcpy(B.d_diag, M->diag.data(), B.nb * sizeof(int), hipMemcpyHostToDevice));
    M->ilu_rp.clear();
    M->ilu_ci.clear();
    if (ilu && Second_arg > 0) {
      // the reference's list holds (int)(0.5 (k+2)^3 max_nrneig) columns,
      // max_nrneig the most off-diagonal blocks of a row
      const int k = Second_arg;
      int max_off = 0;
      for (int r = 0; r < B.nb; r++) max_off = std::max(max_off, M->rp[r + 1] - M->rp[r] - 1);
      const int cap = (int)(0.5 * (k + 2) * (k + 2) * (k + 2) * max_off);
      int bad_row = -1;
      if (bcrs_iluk_pattern(B.nb, M->rp, M->ci, k, cap, M->ilu_rp, M->ilu_ci, &bad_row)) {
        set_error("lar_allocate_preconditioner: the ILU(" + std::to_string(k) + ") pattern of block row " +
                  std::to_string(bad_row) + " exceeds the reference's list of " +
                  std::to_string(cap) + " neighbours");
        free_bcrs_precon(B);
        M->precon = -1;
        return -1;
      }
      const std::vector<int>& prp = M->ilu_rp;
      const std::vector<int>& pci = M->ilu_ci;
      std::vector<int> pdiag(B.nb), src(pci.size());
      int max_row = 0;
      for (int r = 0; r < B.nb; r++) {
        max_row = std::max(max_row, prp[r + 1] - prp[r]);
        int s = M->rp[r];
        for (int s2 = prp[r]; s2 < prp[r + 1]; s2++) {  // both rows sorted
          while (s < M->rp[r + 1] && M->ci[s] < pci[s2]) s++;
          src[s2] = (s < M->rp[r + 1] && M->ci[s] == pci[s2]) ? s : -1;
          if (pci[s2] == r) pdiag[r] = s2;
        }
      }
      B.ilu_k = k;
      B.ilu_nblk = (long long)pci.size();
      B.ilu_max_row = max_row;
      MF_HIP(hipMalloc(&B.d_ilu_rp, prp.size() * sizeof(int)));
      MF_HIP(hipMalloc(&B.d_ilu_ci, std::max<size_t>(pci.size(), 1) * sizeof(int)));
      MF_HIP(hipMalloc(&B.d_ilu_diag, std::max(B.nb, 1) * sizeof(int)));
      MF_HIP(hipMalloc(&B.d_ilu_src, std::max<size_t>(src.size(), 1) * sizeof(int)));
      MF_HIP(hipMemcpy(B.d_ilu_rp, prp.data(), prp.size() * sizeof(int), hipMemcpyHostToDevice));
      MF_HIP(hipMemcpy(B.d_ilu_ci, pci.data(), pci.size() * sizeof(int), hipMemcpyHostToDevice));
      MF_HIP(hipMemcpy(B.d_ilu_diag, pdiag.data(), B.nb * sizeof(int), hipMemcpyHostToDevice));
      MF_HIP(hipMemcpy(B.d_ilu_src, src.data(), src.size() * sizeof(int), hipMemcpyHostToDevice));
      for (int dir = 0; dir < 2; dir++) {
        std::vector<int> rows;
        bcrs_levels(B.nb, prp, pci, pdiag, dir, B.ilu_lev_ptr[dir], rows);
        MF_HIP(hipMalloc(&B.d_ilu_lev_rows[dir], std::max(B.nb, 1) * sizeof(int)));
        MF_HIP(hipMemcpy(B.d_ilu_lev_rows[dir], rows.data(), B.nb * sizeof(int),
                         hipMemcpyHostToDevice));
      }
      MF_HIP(hipMalloc(&B.d_ilu, std::max<size_t>(B.ilu_nblk, 1) * BLS * BLS * sizeof(double)));
    } else if (ilu) {
      MF_HIP(hipMalloc(&B.d_ilu, std::max<size_t>(B.nblk, 1) * BLS * BLS * sizeof(double)));
    } else {
      MF_HIP(hipMalloc(&B.d_lu, std::max(B.nb, 1) * BLS * BLS * sizeof(double)));
      MF_HIP(hipMalloc(&B.d_piv, std::max(B.nb, 1) * BLS * sizeof(int)));
    }
    if (Precon == BLOCK_GS || ilu)
      for (int dir = 0; dir < 2; dir++) {
        std::vector<int> rows;
        bcrs_levels(B.nb, M->rp, M->ci, M->diag, dir, B.lev_ptr[dir], rows);
        MF_HIP(hipMalloc(&B.d_lev_rows[dir], std::max(B.nb, 1) * sizeof(int)));
        MF_HIP(hipMemcpy(B.d_lev_rows[dir], rows.data(), B.nb * sizeof(int),
                         hipMemcpyHostToDevice));
      }
    return B.nb;
  }
  if (Precon != NO_PRECON && Precon != BLOCK_JACOBI && Precon != BLOCK_GS) {
    set_error("lar_allocate_preconditioner: CRS storage offers NO_PRECON, BLOCK_JACOBI and "
              "BLOCK_GS");
    return -1;
  }
  M->precon = Precon;
  M->filled = false;
  M->op.reset();
  M->op_kind = -1;
  Problem& P = *M->P;
  if (!P.d_dinv) MF_HIP(hipMalloc(&P.d_dinv, (size_t)P.n * sizeof(double)));
  if (!P.d_diag_ptr) MF_HIP(hipMalloc(&P.d_diag_ptr, (size_t)P.n * sizeof(int)));
  return P.n;
}

// lar_fill_preconditioner (las_crs_intf.c:707-737): diag_precon = 1/a_ii;
// BCRS (las_bcrs_intf.c:1072-1148): LU of every diagonal block as LAPACK
// dgetrf computes it (an error, where dgetrf would report a zero pivot)
int lar_fill_preconditioner(int Matrix_id) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (flush(*M)) return -2;
  if (M->B) {
    if (M->precon == NO_PRECON) return 0;
    if (M->precon < 0 || !(M->B->d_lu || M->B->d_ilu)) {
      set_error("lar_fill_preconditioner: preconditioner not allocated");
      return -1;
    }
    if (M->B->d_ilu ? bcrs_ilu_factor(*M->B) : bcrs_fill_lu(*M->B)) return -1;
    M->filled = true;
    return 0;
  }
  Problem& P = *M->P;
  if (!P.d_dinv) {
    set_error("lar_fill_preconditioner: preconditioner not allocated");
    return -1;
  }
  if (launch_diag_ptr(P, P.d_diag_ptr) || launch_fill_jacobi(P)) return -2;
  MF_HIP(hipStreamSynchronize(stream()));
  M->filled = true;
  return 0;
}

int lar_free_preconditioner(int Matrix_id) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (M->B) {
    M->op.reset();
    M->op_kind = -1;
    free_bcrs_precon(*M->B);
    M->filled = false;
    M->precon = -1;
    return 0;
  }
  Problem& P = *M->P;
  M->op.reset();
  M->op_kind = -1;
  M->filled = false;
  M->precon = -1;
  if (P.d_dinv) (void)hipFree(P.d_dinv);
  if (P.d_diag_ptr) (void)hipFree(P.d_diag_ptr);
  P.d_dinv = nullptr;
  P.d_diag_ptr = nullptr;
  return 0;
}

// lar_compute_residual (las_crs_intf.c:788-849, las_bcrs_intf.c:2787-2930):
// V = (Use_rhs ? B|rhs : 0) - A X
void lar_compute_residual(int Matrix_id, int Use_rhs, int Ini_zero, int Ndof, double* X,
                          double* B, double* V) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M || Ndof != M->ndof()) {
    fprintf(stderr, "lar_compute_residual: bad matrix id or size\n");
    return;
  }
  if (flush(*M)) return;
  if (M->B) {
    Staged sx, sb, sv;
    if (!Ini_zero && stage_in(X, Ndof, sx, true)) return;
    if (Use_rhs == 1 && B && stage_in(B, Ndof, sb, true)) return;
    if (stage_in(V, Ndof, sv, false)) return;
    const double* db = (Use_rhs == 1) ? (B ? sb.d : M->B->d_rhs) : nullptr;
    if (bcrs_residual(*M->B, Ini_zero ? nullptr : sx.d, db, sv.d)) return;
    stage_out(V, sv, Ndof);
    return;
  }
  Problem& P = *M->P;
  Staged sx, sb, sv;
  if (!Ini_zero && stage_in(X, Ndof, sx, true)) return;
  if (Use_rhs == 1 && B && stage_in(B, Ndof, sb, true)) return;
  if (stage_in(V, Ndof, sv, false)) return;
  const double* db = (Use_rhs == 1) ? (B ? sb.d : P.d_rhs) : nullptr;
  if (launch_rowop(P, 0, Use_rhs == 1, Ini_zero ? nullptr : sx.d, db, nullptr, sv.d)) return;
  stage_out(V, sv, Ndof);
}

// lar_perform_BJ_or_GS_iterations (las_crs_intf.c:857-1013), BLOCK_JACOBI:
// V <- D^-1 (B - (A - D) V), Nr_prec sweeps; BLOCK_GS: Nr_prec forward
// sweeps in row chunks as the reference runs them on gs_threads OpenMP
// threads (las_crs_intf.c:905-998, mf_lar_set_gs_threads).  BCRS (las_bcrs_intf.c:
// 2938-3300): block Jacobi or block Gauss-Seidel in the reference's row order;
// even Nr_prec gives Nr_prec/2 forward + backward sweep pairs, odd gives
// Nr_prec/2 + 1 forward sweeps
void lar_perform_BJ_or_GS_iterations(int Matrix_id, int Use_rhs, int Ini_zero, int Nr_prec,
                                     int Ndof, double* V, double* B) {
  (void)Ini_zero;
  LarMatrix* M = lar_get(Matrix_id);
  if (M && M->B) {
    if (Ndof != M->ndof() || (M->precon != BLOCK_JACOBI && M->precon != BLOCK_GS) ||
        !M->filled) {
      fprintf(stderr, "lar_perform_BJ_or_GS_iterations: needs a filled BLOCK_JACOBI or "
                      "BLOCK_GS preconditioner\n");
      return;
    }
    if (flush(*M)) return;
    Staged sv, sb;
    if (stage_in(V, Ndof, sv, true)) return;
    if (Use_rhs == 1 && B && stage_in(B, Ndof, sb, true)) return;
    double* tmp = nullptr;
    if (hipMalloc(&tmp, (size_t)Ndof * sizeof(double)) != hipSuccess) return;
    const double* db = (Use_rhs == 1) ? (B ? sb.d : M->B->d_rhs) : nullptr;
    if (bcrs_sweeps(*M->B, M->precon == BLOCK_GS, Nr_prec, db, sv.d, tmp) == 0)
      stage_out(V, sv, Ndof);
    (void)hipFree(tmp);
    return;
  }
  if (!M || Ndof != M->P->n || (M->precon != BLOCK_JACOBI && M->precon != BLOCK_GS) ||
      !M->filled) {
    fprintf(stderr, "lar_perform_BJ_or_GS_iterations: needs a filled BLOCK_JACOBI or BLOCK_GS "
                    "preconditioner\n");
    return;
  }
  if (flush(*M)) return;
  Problem& P = *M->P;
  Staged sv, sb;
  if (stage_in(V, Ndof, sv, true)) return;
  if (Use_rhs == 1 && B && stage_in(B, Ndof, sb, true)) return;
  double* tmp = nullptr;
  if (hipMalloc(&tmp, (size_t)Ndof * sizeof(double)) != hipSuccess) return;
  const double* db = (Use_rhs == 1) ? (B ? sb.d : P.d_rhs) : nullptr;
  if (M->precon == BLOCK_GS) {
    if (crs_gs_chunked(P, M->gs_threads, Nr_prec, db, sv.d, tmp, nullptr) == 0)
      stage_out(V, sv, Ndof);
    else
      fprintf(stderr, "lar_perform_BJ_or_GS_iterations: %s\n", last_error());
    (void)hipFree(tmp);
    return;
  }
  for (int it = 0; it < Nr_prec; it++) {
    if (launch_rowop(P, 1, Use_rhs == 1, sv.d, db, P.d_diag_ptr, tmp)) break;
    if (hipMemcpyAsync(sv.d, tmp, (size_t)Ndof * sizeof(double), hipMemcpyDeviceToDevice,
                       stream()) != hipSuccess)
      break;
  }
  stage_out(V, sv, Ndof);
  (void)hipFree(tmp);
}

// lar_perform_rhsub: V = (LU)^-1 B (B null: the stored rhs) with the BCRS
// block ILU(k) factors (las_bcrs_intf.c:3319-3510, variant 0: forward
// reduction and back substitution in the reference's row order, bitwise its
// result).  The CRS body is empty in the reference (las_crs_intf.c:1019-1057):
// a no-op with a diagnostic there
void lar_perform_rhsub(int Matrix_id, int Ndof, double* V, double* B) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M || !M->B || !M->B->d_ilu || !M->filled || Ndof != M->ndof()) {
    set_error("lar_perform_rhsub: needs a BCRS matrix with filled ILU factors");
    fprintf(stderr, "lar_perform_rhsub: needs a BCRS matrix with filled ILU factors\n");
    return;
  }
  if (flush(*M)) return;
  Staged sb, sv;
  if (B && stage_in(B, Ndof, sb, true)) return;
  if (stage_in(V, Ndof, sv, false)) return;
  if (bcrs_ilu_solve(*M->B, B ? sb.d : M->B->d_rhs, sv.d)) return;
  stage_out(V, sv, Ndof);
}

// restarted GMRES on a BCRS matrix with its preconditioner (lsr_mkb_core_gmres,
// lss_mkb_core.c:2317, over the storage operations as compreres applies them,
// lss_mkb_core.c:1915-1990); b is the stored rhs, x host memory (the initial
// guess unless ini_zero; the solution on return)
int mf_lar_gmres(int Matrix_id, int krylbas, int max_iter, double tol, int nr_prec, int ini_zero,
                 double* x, int* iters, double* res_rel, double* resid) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (!M->B) {
    if (M->precon != BLOCK_GS) {
      set_error("mf_lar_gmres: needs a BCRS matrix, or a CRS matrix with BLOCK_GS (other CRS "
                "problems solve with mf_gmres)");
      return -1;
    }
    return lar_crs_gs_gmres(*M, krylbas, max_iter, tol, 0.0, nr_prec, ini_zero, x, iters,
                            res_rel, resid);
  }
  return lar_bcrs_gmres(*M, krylbas, max_iter, tol, 0.0, nr_prec, ini_zero, x, iters, res_rel,
                        resid);
}

}  // extern "C"

namespace {
// GMRES over a BCRS matrix's storage operations, tolerances relative to the
// first preconditioned residual and absolute (each floored at 1e-12)
int lar_bcrs_gmres(LarMatrix& L, int krylbas, int max_iter, double tol, double tol_abs,
                   int nr_prec, int ini_zero, double* x, int* iters, double* res_rel,
                   double* resid) {
  LarMatrix* M = &L;
  const int kind = M->precon < 0 ? NO_PRECON : M->precon;
  if (kind != NO_PRECON && !M->filled) {
    set_error("mf_lar_gmres: the preconditioner is not filled");
    return -1;
  }
  if ((kind == BLOCK_JACOBI || kind == BLOCK_GS) && nr_prec < 1) {
    set_error("mf_lar_gmres: nr_prec must be >= 1");
    return -1;
  }
  if (flush(*M)) return -2;
  BcrsDev& B = *M->B;
  if (!M->op || M->op_kind != kind || M->op_sweeps != nr_prec) {
    M->op.reset(bcrs_make_op(&B, kind, nr_prec));
    if (!M->op) return -2;
    M->op_kind = kind;
    M->op_sweeps = nr_prec;
  }
  if (!M->G) {
    M->G = std::make_unique<Problem>();
    M->G->n = BLS * B.nb;
    M->G->el_type = kPrism;
  }
  M->G->ext = M->op.get();
  M->G->d_rhs = B.d_rhs;
  return gmres_solve(*M->G, 0, krylbas, max_iter, tol, tol_abs, ini_zero, x, iters, res_rel,
                     resid);
}

// GMRES over a CRS matrix preconditioned by nr_prec chunked Gauss-Seidel
// sweeps per application (smoothing, as lsr_mkb_core_compreres applies
// BLOCK_GS), tolerances as lar_bcrs_gmres
int lar_crs_gs_gmres(LarMatrix& L, int krylbas, int max_iter, double tol, double tol_abs,
                     int nr_prec, int ini_zero, double* x, int* iters, double* res_rel,
                     double* resid) {
  LarMatrix* M = &L;
  if (M->precon != BLOCK_GS || !M->filled) {
    set_error("mf_lar_gmres: the BLOCK_GS preconditioner is not filled");
    return -1;
  }
  if (nr_prec < 1) {
    set_error("mf_lar_gmres: nr_prec must be >= 1");
    return -1;
  }
  if (flush(*M)) return -2;
  Problem& P = *M->P;
  if (!M->op || M->op_kind != (BLOCK_GS << 24 | M->gs_threads) || M->op_sweeps != nr_prec) {
    M->op.reset(crs_gs_make_op(&P, M->gs_threads, nr_prec));
    if (!M->op) return -2;
    M->op_kind = BLOCK_GS << 24 | M->gs_threads;
    M->op_sweeps = nr_prec;
  }
  if (!M->G) {
    M->G = std::make_unique<Problem>();
    M->G->n = P.n;
    M->G->el_type = kPrism;
  }
  M->G->ext = M->op.get();
  M->G->d_rhs = P.d_rhs;
  return gmres_solve(*M->G, 0, krylbas, max_iter, tol, tol_abs, ini_zero, x, iters, res_rel,
                     resid);
}
}  // namespace

extern "C" {

// HIP-event times on the library stream of one BCRS operation, averaged over
// reps, on device-resident vectors: what 0 residual v = b - A x, 1 one
// forward sweep of the allocated block Jacobi / Gauss-Seidel preconditioner,
// 2 the ILU(k) factorisation, 3 one ILU(k) solve
int mf_lar_bench(int Matrix_id, int what, int reps, double* ms) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (!M->B || reps < 1 || !ms || what < 0 || what > 3) {
    set_error("mf_lar_bench: needs a BCRS matrix, reps >= 1 and what in 0..3");
    return -1;
  }
  BcrsDev& B = *M->B;
  const bool ilu = B.d_ilu != nullptr;
  if ((what == 1 && (ilu || !B.d_lu || !M->filled)) || (what >= 2 && !ilu) ||
      (what == 3 && !M->filled)) {
    set_error("mf_lar_bench: the operation needs its (filled) preconditioner");
    return -1;
  }
  if (flush(*M)) return -2;
  const size_t n = (size_t)BLS * B.nb;
  double *x = nullptr, *v = nullptr, *t = nullptr;
  MF_HIP(hipMalloc(&x, std::max<size_t>(n, 1) * sizeof(double)));
  MF_HIP(hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(double)));
  MF_HIP(hipMalloc(&t, std::max<size_t>(n, 1) * sizeof(double)));
  hipStream_t s = stream();
  MF_HIP(hipMemcpyAsync(x, B.d_rhs, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  MF_HIP(hipMemcpyAsync(v, B.d_rhs, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  hipEvent_t e0, e1;
  MF_HIP(hipEventCreate(&e0));
  MF_HIP(hipEventCreate(&e1));
  int rc = 0;
  auto once = [&]() -> int {
    switch (what) {
      case 0: return bcrs_residual(B, x, B.d_rhs, v);
      case 1: return bcrs_sweeps(B, M->precon == BLOCK_GS, 1, B.d_rhs, v, t, nullptr, false);
      case 2: return bcrs_ilu_factor(B);
      default: return bcrs_ilu_solve(B, B.d_rhs, v);
    }
  };
  rc = once();  // warm-up
  if (rc == 0) {
    MF_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps && rc == 0; r++) rc = once();
    MF_HIP(hipEventRecord(e1, s));
    MF_HIP(hipEventSynchronize(e1));
    float t_ms = 0;
    MF_HIP(hipEventElapsedTime(&t_ms, e0, e1));
    *ms = t_ms / reps;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(x);
  (void)hipFree(v);
  (void)hipFree(t);
  return rc;
}

// order of the BCRS block Gauss-Seidel sweeps: 0 the reference's row order
// (default; bitwise the reference), 1 colour classes (rows of a class share
// no block; a sweep is one launch per class, classes ascending forward and
// descending backward).  Returns the number of colours (1 for order 0)
int mf_lar_set_gs_order(int Matrix_id, int order) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (!M->B || (order != 0 && order != 1)) {
    set_error("mf_lar_set_gs_order: needs a BCRS matrix and order 0 or 1");
    return -1;
  }
  BcrsDev& B = *M->B;
  if (order == 1 && !B.d_col_rows && B.nb > 0) {
    std::vector<int> rows;
    bcrs_colors(B.nb, M->rp, M->ci, B.col_ptr, rows);
    MF_HIP(hipMalloc(&B.d_col_rows, B.nb * sizeof(int)));
    MF_HIP(hipMemcpy(B.d_col_rows, rows.data(), B.nb * sizeof(int), hipMemcpyHostToDevice));
  }
  B.gs_order = order;
  return order == 1 ? (int)B.col_ptr.size() - 1 : 1;
}

// thread count whose row chunks the CRS BLOCK_GS sweeps reproduce
// (las_crs_intf.c:905-998 runs one chunk of n/T + 1 rows per OpenMP thread)
int mf_lar_set_gs_threads(int Matrix_id, int Nthreads) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (M->B || Nthreads < 1 || Nthreads > kMaxGsChunks) {
    set_error("mf_lar_set_gs_threads: needs a CRS matrix and 1 <= Nthreads <= " +
              std::to_string(kMaxGsChunks));
    return -1;
  }
  M->gs_threads = Nthreads;
  return 0;
}

// storage type (2 CRS, 3 BCRS), block size, rows of blocks and stored blocks
// of a matrix, to size the buffers of mf_lar_get_system
int mf_lar_get_sizes(int Matrix_id, int* storage, int* block_size, int* nrblocks,
                     long long* nr_stored) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (storage) *storage = M->B ? LAC_STORAGE_BCRS : LAC_STORAGE_CRS;
  if (block_size) *block_size = M->B ? BLS : 1;
  if (nrblocks) *nrblocks = M->B ? M->B->nb : M->P->n;
  if (nr_stored) *nr_stored = M->B ? M->B->nblk : M->P->nnz;
  return 0;
}

// host copies of the BCRS diagonal-block LU factors (16 per block row,
// column-major, unit lower part below the diagonal) and 0-based pivots (4 per
// block row: row j was interchanged with row piv[j]); for block ILU(k) the
// factors of every stored block (16 each: L blocks below the diagonal, U
// blocks above, inverted U diagonal blocks) and no pivots
int mf_lar_get_block_precon(int Matrix_id, double* lu, int* piv) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (!M->B || !M->filled) {
    set_error("mf_lar_get_block_precon: needs a BCRS matrix with a filled block preconditioner");
    return -1;
  }
  const BcrsDev& B = *M->B;
  if (B.d_ilu) {  // ILU(k): the factors of every block of its pattern
    const long long nf = B.ilu_k > 0 ? B.ilu_nblk : B.nblk;
    if (lu && nf)
      MF_HIP(hipMemcpy(lu, B.d_ilu, nf * BLS * BLS * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
  if (lu) MF_HIP(hipMemcpy(lu, B.d_lu, (size_t)B.nb * BLS * BLS * sizeof(double), hipMemcpyDeviceToHost));
  if (piv) MF_HIP(hipMemcpy(piv, B.d_piv, (size_t)B.nb * BLS * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

// the block pattern of the allocated BCRS ILU(k) factors: for k > 0 the
// grown pattern, for k = 0 the matrix's own; row_ptr [nrblocks+1], col_ind
// [*nr_stored] (either may be null); returns k
int mf_lar_get_ilu_pattern(int Matrix_id, int* row_ptr, int* col_ind, long long* nr_stored) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (!M->B || !M->B->d_ilu) {
    set_error("mf_lar_get_ilu_pattern: needs a BCRS matrix with block ILU allocated");
    return -1;
  }
  const bool grown = M->B->ilu_k > 0;
  const std::vector<int>& rp = grown ? M->ilu_rp : M->rp;
  const std::vector<int>& ci = grown ? M->ilu_ci : M->ci;
  if (row_ptr) memcpy(row_ptr, rp.data(), rp.size() * sizeof(int));
  if (col_ind) memcpy(col_ind, ci.data(), ci.size() * sizeof(int));
  if (nr_stored) *nr_stored = (long long)ci.size();
  return M->B->ilu_k;
}

// host copies of the stored system (lar_get_SM_and_LV_crs_from_crs,
// las_crs_intf.c:1063-1120 with offset 0); BCRS: block row pointers, block
// column indices, 16 values per block (column-major) and the rhs
int mf_lar_get_system(int Matrix_id, int* row_ptr, int* col_ind, double* val, double* rhs) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (flush(*M)) return -2;
  if (M->B) {
    const BcrsDev& B = *M->B;
    if (row_ptr) memcpy(row_ptr, M->rp.data(), M->rp.size() * sizeof(int));
    if (col_ind) memcpy(col_ind, M->ci.data(), M->ci.size() * sizeof(int));
    if (val && B.nblk)
      MF_HIP(hipMemcpy(val, B.d_val, B.nblk * BLS * BLS * sizeof(double), hipMemcpyDeviceToHost));
    if (rhs && B.nb)
      MF_HIP(hipMemcpy(rhs, B.d_rhs, (size_t)B.nb * BLS * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
  }
  Problem& P = *M->P;
  if (row_ptr) memcpy(row_ptr, P.setup.row_ptr.data(), (P.n + 1) * sizeof(int));
  if (col_ind) memcpy(col_ind, P.setup.col_ind.data(), P.nnz * sizeof(int));
  if (val) MF_HIP(hipMemcpy(val, P.d_val, P.nnz * sizeof(double), hipMemcpyDeviceToHost));
  if (rhs) MF_HIP(hipMemcpy(rhs, P.d_rhs, P.n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------
// Solver-module seam with the reference's own names and signatures
// (lsd_mkb/lsh_mkb_intf.h:165-338, implemented in lsd_mkb/lss_mkb_intf.c and
// lsd_mkb/lsd_mkb_core/lss_mkb_core.c): a solver instance reads its control
// file (lsr_mkb_core_init, single level), owns one matrix of the storage seam
// above, and solves with restarted GMRES on the device (lsr_mkb_core_solve /
// lsr_mkb_core_gmres).  Solver ids count from 0 in creation order and are
// destroyed in reverse order, as lsv_mkb_solver is (lss_mkb_intf.c:104-110).
// ---------------------------------------------------------------------
namespace {
#define LSC_SOLVE 0
#define LSC_RESOLVE 1
#define LSC_MAX_NUM_SOLV 10
#define REL_RES_INI 0
#define ABS_RES 1

struct LsrSolver {
  bool alive = false;
  int solver_type = 0;
  MkbParams prm;
  int matrix_id = -1;
  int nrdofgl = 0;
};
std::vector<LsrSolver> g_lsr;

LsrSolver* lsr_get(int id, int level = 0) {
  if (id < 0 || id >= (int)g_lsr.size() || !g_lsr[id].alive) {
    set_error("lsr_mkb_: unknown solver id " + std::to_string(id));
    return nullptr;
  }
  if (level != 0) {
    set_error("lsr_mkb_: single-level solvers only (level 0)");
    return nullptr;
  }
  return &g_lsr[id];
}
}  // namespace

extern "C" {

// lsr_mkb_init (lss_mkb_intf.c:88-175): control parameters from Filename
// (arguments >= 0 override them), single level; returns the solver id
int lsr_mkb_init(int Solver_type, int Parallel, int* Max_num_levels_p, int* Storage_type,
                 char* Filename, int Max_iter, int Error_type, double Error_tolerance,
                 int Monitoring_level) {
  if (Parallel != 0) {
    set_error("lsr_mkb_init: the sequential solver only (one subdomain per device)");
    return -1;
  }
  if (Solver_type <= 0 || Solver_type >= 100) {
    set_error("lsr_mkb_init: the MKB iterative solvers (types 1..99) only");
    return -1;
  }
  if ((int)g_lsr.size() >= LSC_MAX_NUM_SOLV) {
    set_error("lsr_mkb_init: too many solvers");
    return -1;
  }
  LsrSolver S;
  if (mkb_read(Filename, Max_iter, Error_type, Error_tolerance, Monitoring_level, S.prm) < 0)
    return -1;
  if (S.prm.solver != 1 && !(Max_num_levels_p && *Max_num_levels_p == 1)) {
    set_error("lsr_mkb_init: single-level GMRES only");
    return -1;
  }
  S.alive = true;
  S.solver_type = Solver_type;
  if (Max_num_levels_p) *Max_num_levels_p = 1;
  if (Storage_type) *Storage_type = S.prm.storage_type;
  g_lsr.push_back(S);
  return (int)g_lsr.size() - 1;
}

// lsr_mkb_destroy (lss_mkb_intf.c:355-410): the last created solver first
int lsr_mkb_destroy(int Solver_id) {
  if (!lsr_get(Solver_id)) return -3;
  if (Solver_id != (int)g_lsr.size() - 1) {
    set_error("lsr_mkb_destroy: solvers are destroyed in reverse order of creation");
    return -1;
  }
  if (g_lsr.back().matrix_id >= 0) {
    set_error("lsr_mkb_destroy: free the solver's matrix first (lsr_mkb_free_matrix)");
    return -1;
  }
  g_lsr.pop_back();
  return 1;
}

// lsr_mkb_create_matrix (lss_mkb_intf.c:417-520): the storage type enforced
// by the caller, else the control file's, through lar_allocate_SM_and_LV
int lsr_mkb_create_matrix(int Solver_id, int Level_id, int Storage_type, int Nrblocks,
                          int Nrdof_glob, int Block_size, int Max_sm_size, int* Nrdofbl,
                          int* Posglob, int* Nroffbl, int** L_offbl) {
  LsrSolver* S = lsr_get(Solver_id, Level_id);
  if (!S) return -3;
  if (S->matrix_id >= 0) {
    set_error("lsr_mkb_create_matrix: the solver already has a matrix");
    return -1;
  }
  const int storage = Storage_type != LAC_STORAGE_UNDEFINED ? Storage_type : S->prm.storage_type;
  const int id = lar_allocate_SM_and_LV(storage, S->solver_type, Nrdof_glob, Max_sm_size,
                                        Nrblocks, Block_size, Nrdofbl, Posglob, Nroffbl, L_offbl);
  if (id < 0) return id;
  S->matrix_id = id;
  S->nrdofgl = Nrdof_glob;
  return 0;
}

// lsr_mkb_clear_matrix (lss_mkb_intf.c:528-600): LSC_SOLVE clears values and
// rhs, LSC_RESOLVE the rhs
int lsr_mkb_clear_matrix(int Solver_id, int Level_id, int Comp_type) {
  LsrSolver* S = lsr_get(Solver_id, Level_id);
  if (!S) return -3;
  return lar_initialize_SM_and_LV(S->matrix_id,
                                  Comp_type == LSC_RESOLVE ? LAC_SCOPE_LV : LAC_SCOPE_SM_AND_LV);
}

// lsr_mkb_fill_assembly_table_int_ent (lss_mkb_intf.c:605-635)
int lsr_mkb_fill_assembly_table_int_ent(int Solver_id, int Level_id, int Nr_dof_bl, int* L_bl_id,
                                        int* L_bl_nrdof, int* Assembly_table_int_ent) {
  LsrSolver* S = lsr_get(Solver_id, Level_id);
  if (!S) return -3;
  return lar_fill_assembly_table_int_ent(S->matrix_id, Nr_dof_bl, L_bl_id, L_bl_nrdof,
                                         Assembly_table_int_ent);
}

// lsr_mkb_assemble_local_stiff_mat_with_table (lss_mkb_intf.c:641-700)
int lsr_mkb_assemble_local_stiff_mat_with_table(int Solver_id, int Level_id, int Comp_type,
                                                int Nr_dof_bl, int* Assembly_table_int_ent,
                                                double* Stiff_mat, double* Rhs_vect,
                                                char* Rewr_dofs) {
  LsrSolver* S = lsr_get(Solver_id, Level_id);
  if (!S) return -3;
  return lar_assemble_SM_and_LV_with_table(
      S->matrix_id, Comp_type == LSC_RESOLVE ? LAC_SCOPE_LV : LAC_SCOPE_SM_AND_LV, Nr_dof_bl,
      Assembly_table_int_ent, Stiff_mat, Rhs_vect, Rewr_dofs);
}

// lsr_mkb_assemble_local_sm (lss_mkb_intf.c:705-785)
int lsr_mkb_assemble_local_sm(int Solver_id, int Level_id, int Comp_type, int Nr_dof_bl,
                              int* L_bl_id, int* L_bl_nrdof, double* Stiff_mat, double* Rhs_vect,
                              char* Rewr_dofs) {
  LsrSolver* S = lsr_get(Solver_id, Level_id);
  if (!S) return -3;
  return lar_assemble_SM_and_LV(S->matrix_id,
                                Comp_type == LSC_RESOLVE ? LAC_SCOPE_LV : LAC_SCOPE_SM_AND_LV,
                                Nr_dof_bl, L_bl_id, L_bl_nrdof, Stiff_mat, Rhs_vect, Rewr_dofs);
}

// lsr_mkb_free_matrix (lss_mkb_intf.c:791-838): preconditioner and matrix
int lsr_mkb_free_matrix(int Solver_id) {
  LsrSolver* S = lsr_get(Solver_id);
  if (!S) return -3;
  if (S->matrix_id < 0) return 0;
  (void)lar_free_preconditioner(S->matrix_id);
  const int rc = lar_free_SM_and_LV(S->matrix_id);
  if (rc < 0) return rc;
  S->matrix_id = -1;
  return 1;
}

// lsr_mkb_create_precon (lss_mkb_intf.c:844-880 -> lsr_mkb_core_create_precon,
// lss_mkb_core.c:964-1020): the control file's preconditioner, with its
// block type as the second argument, or ILU's k for non-block storage;
// returns the number of diagonal blocks
int lsr_mkb_create_precon(int Solver_id, int Level_id) {
  LsrSolver* S = lsr_get(Solver_id, Level_id);
  if (!S) return -3;
  const int precon = S->prm.precon;
  if (precon == NO_PRECON) return 1;
  const int second = (precon == MULTI_ILU || precon == BLOCK_ILU) ? S->prm.ilu_k
                                                                   : S->prm.block_type;
  return lar_allocate_preconditioner(S->matrix_id, precon, second);
}

// lsr_mkb_fill_precon (lss_mkb_intf.c:888-940)
int lsr_mkb_fill_precon(int Solver_id, int Level_id) {
  LsrSolver* S = lsr_get(Solver_id, Level_id);
  if (!S) return -3;
  if (S->prm.precon == NO_PRECON) return 1;
  return lar_fill_preconditioner(S->matrix_id);
}

// lsr_mkb_solve (lss_mkb_intf.c:184-350 -> lsr_mkb_core_solve, lss_mkb_core.c:
// 712-925): *Nr_iter > 0 and *Toler > 0 override the control file; the
// tolerance is relative to the first preconditioned residual (REL_RES_INI),
// absolute (ABS_RES), or -- as the reference resets it -- none for
// REL_RES_RHS, each floored at 1e-12 (lss_mkb_core.c:2377-2384).  B != NULL
// replaces the stored rhs for this solve.  Returns 1 on convergence, 0
// otherwise; *Nr_iter, *Toler (the final measure of the chosen kind) and
// *Conv_rate (the mean reduction per iteration) come back
int lsr_mkb_solve(int Solver_id, int Comp_type, int Ini_zero, double* X, double* B,
                  int* Nr_iter, double* Toler, int Monitor, double* Conv_rate) {
  (void)Comp_type; (void)Monitor;
  LsrSolver* S = lsr_get(Solver_id);
  if (!S) return -3;
  LarMatrix* M = lar_get(S->matrix_id);
  if (!M) return -3;
  const MkbParams& p = S->prm;
  const int max_iter = (Nr_iter && *Nr_iter > 0) ? *Nr_iter : p.max_iter;
  const double meas = (Toler && *Toler > 0) ? *Toler : p.conv_meas;
  const double tol_rel = p.conv_type == REL_RES_INI ? meas : 0.0;
  const double tol_abs = p.conv_type == ABS_RES ? meas : 0.0;
  if (flush(*M)) return -2;
  const int n = M->ndof();
  // the rhs of this solve: the stored one, or B staged in its place
  double* d_rhs_saved = M->B ? M->B->d_rhs : M->P->d_rhs;
  Staged sb;
  if (B) {
    if (stage_in(B, n, sb, true)) return -2;
    if (M->B) M->B->d_rhs = sb.d;
    else M->P->d_rhs = sb.d;
  }
  int it = 0, rc = 0;
  double rr = 0.0, rs = 0.0;
  if (M->B) {
    rc = lar_bcrs_gmres(*M, p.krylbas, max_iter, tol_rel, tol_abs, p.nr_prec, Ini_zero, X, &it,
                        &rr, &rs);
  } else {
    int precon = 0;
    const bool gs = p.precon == BLOCK_GS && M->precon == BLOCK_GS;
    if (p.precon == BLOCK_JACOBI && M->precon == BLOCK_JACOBI && p.nr_prec == 1) precon = 1;
    else if (p.precon != NO_PRECON && !gs) {
      set_error("lsr_mkb_solve: CRS storage solves with no preconditioner, one block-Jacobi "
                "sweep or block Gauss-Seidel");
      rc = -1;
    }
    if (rc == 0 && gs)
      rc = lar_crs_gs_gmres(*M, p.krylbas, max_iter, tol_rel, tol_abs, p.nr_prec, Ini_zero, X,
                            &it, &rr, &rs);
    else if (rc == 0)
      rc = gmres_solve(*M->P, precon, p.krylbas, max_iter, tol_rel, tol_abs, Ini_zero, X, &it,
                       &rr, &rs);
  }
  if (B) {
    if (M->B) M->B->d_rhs = d_rhs_saved;
    else M->P->d_rhs = d_rhs_saved;
  }
  if (rc < 0) return rc;
  if (Nr_iter) *Nr_iter = it;
  if (Toler) *Toler = p.conv_type == ABS_RES ? rs : rr;
  if (Conv_rate) *Conv_rate = (it > 0 && rr > 0.0) ? std::exp(std::log(rr) / it) : 0.0;
  return rc;
}

}  // extern "C"
