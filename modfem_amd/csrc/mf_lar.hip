// Storage/operations seam with the reference's own names and signatures
// (lsd_mkb/lah_intf.h:90-289, dispatcher lsd_mkb/las_intf.c), block size 1
// (LAC_STORAGE_CRS, las_intf.c:95-97).  A ModFEM build can link these in
// place of las_intf.c + lad_crs: the matrix and vectors live in device memory
// and every operation runs on the GPU; vectors passed in may be host or
// device pointers.
//
// Per-element assembly calls (lar_assemble_SM_and_LV[_with_table]) come one
// element at a time from the host, so they accumulate into a host mirror of
// the values in call order -- the same order and rounding as
// lar_assemble_SM_and_LV_with_table_crs (lad_crs/las_crs_intf.c:440-513) --
// and the mirror is uploaded before the next device operation.  Device-side
// batched assembly is mf_integrate_assemble (include/modfem_hip.h).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "mf_device.h"

using namespace mf;

#define LAC_SCOPE_LV 1
#define LAC_SCOPE_SM 2
#define LAC_SCOPE_SM_AND_LV 0
#define NO_PRECON 0
#define BLOCK_JACOBI 1
#define LAC_STORAGE_UNDEFINED -1
#define LAC_STORAGE_CRS 2
#define LAC_MAX_MATRICES 20

namespace {

struct LarMatrix {
  std::unique_ptr<Problem> P;
  std::vector<double> h_val, h_rhs;  // host mirror for per-element assembly
  bool dirty = false;
  int precon = -1;
};

std::vector<LarMatrix> g_mats;  // LIFO like itv_crs_matrices

LarMatrix* lar_get(int id) {
  if (id < 0 || id >= (int)g_mats.size() || !g_mats[id].P) {
    set_error("lar_: unknown matrix id " + std::to_string(id));
    return nullptr;
  }
  return &g_mats[id];
}

int flush(LarMatrix& M) {
  if (!M.dirty) return 0;
  Problem& P = *M.P;
  MF_HIP(hipMemcpyAsync(P.d_val, M.h_val.data(), (size_t)P.nnz * sizeof(double),
                        hipMemcpyHostToDevice, stream()));
  MF_HIP(hipMemcpyAsync(P.d_rhs, M.h_rhs.data(), (size_t)P.n * sizeof(double),
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

}  // namespace

extern "C" {

// lar_allocate_SM_and_LV (lah_intf.h:90; CRS branch las_crs_intf.c:72-294):
// rows = blocks, columns = sorted off-diagonal blocks with the diagonal
// inserted in place.  Returns the matrix id.
int lar_allocate_SM_and_LV(int Storage_type, int solver_type, int Nrdof_glob, int Max_SM_size,
                           int Nrblocks, int Block_size, int* Nrdofbl, int* Posglob,
                           int* Nroffbl, int** L_offbl) {
  (void)solver_type; (void)Max_SM_size; (void)Nrdofbl; (void)Posglob;
  if (Storage_type != LAC_STORAGE_UNDEFINED && Storage_type != LAC_STORAGE_CRS) {
    set_error("lar_allocate_SM_and_LV: only CRS storage (block size 1) is available");
    return -1;
  }
  if (Block_size != 1 || Nrblocks != Nrdof_glob) {
    set_error("lar_allocate_SM_and_LV: block size must be 1");
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
  const int n = Nrdof_glob;
  std::vector<int> rp(n + 1, 0), ci;
  for (int ib = 0; ib < n; ib++) {
    rp[ib] = (int)ci.size();
    if (Nroffbl[ib] <= 0) continue;  // ghost block: no stored row
    bool inserted = false;
    for (int j = 0; j < Nroffbl[ib]; j++) {
      int c = L_offbl[ib][j];
      if (j > 0 && L_offbl[ib][j - 1] >= c) {
        set_error("lar_allocate_SM_and_LV: neighbour lists must be strictly ascending");
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
  const Problem& P = *M->P;
  return ((double)P.nnz * 12 + (double)P.n * 20) / 1024.0 / 1024.0;
}

// lar_fill_assembly_table_int_ent (las_crs_intf.c:354-432)
int lar_fill_assembly_table_int_ent(int Matrix_id, int Nr_dof_bl, int* L_bl_id, int* L_bl_nrdof,
                                    int* Assembly_table_int_ent) {
  (void)L_bl_nrdof;
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
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
  const Setup& S = M->P->setup;
  const int k = Nr_dof_bl;
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
  Problem& P = *M->P;
  void* ptrs[] = {P.d_row_ptr, P.d_col, P.d_val, P.d_rhs, P.d_dinv, P.d_diag_ptr, P.d_tile_row};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  g_mats.pop_back();
  return 0;
}

// lar_allocate_preconditioner (las_crs_intf.c:658-700): Jacobi or none
int lar_allocate_preconditioner(int Matrix_id, int Precon, int Second_arg) {
  (void)Second_arg;
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (Precon != NO_PRECON && Precon != BLOCK_JACOBI) {
    set_error("lar_allocate_preconditioner: CRS storage offers NO_PRECON and BLOCK_JACOBI");
    return -1;
  }
  M->precon = Precon;
  Problem& P = *M->P;
  if (!P.d_dinv) MF_HIP(hipMalloc(&P.d_dinv, (size_t)P.n * sizeof(double)));
  if (!P.d_diag_ptr) MF_HIP(hipMalloc(&P.d_diag_ptr, (size_t)P.n * sizeof(int)));
  return P.n;
}

// lar_fill_preconditioner (las_crs_intf.c:707-737): diag_precon = 1/a_ii
int lar_fill_preconditioner(int Matrix_id) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (flush(*M)) return -2;
  Problem& P = *M->P;
  if (!P.d_dinv) {
    set_error("lar_fill_preconditioner: preconditioner not allocated");
    return -1;
  }
  if (launch_diag_ptr(P, P.d_diag_ptr) || launch_fill_jacobi(P)) return -2;
  MF_HIP(hipStreamSynchronize(stream()));
  return 0;
}

int lar_free_preconditioner(int Matrix_id) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  Problem& P = *M->P;
  if (P.d_dinv) (void)hipFree(P.d_dinv);
  if (P.d_diag_ptr) (void)hipFree(P.d_diag_ptr);
  P.d_dinv = nullptr;
  P.d_diag_ptr = nullptr;
  return 0;
}

// lar_compute_residual (las_crs_intf.c:788-849): V = (Use_rhs ? B|rhs : 0) - A X
void lar_compute_residual(int Matrix_id, int Use_rhs, int Ini_zero, int Ndof, double* X,
                          double* B, double* V) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M || Ndof != M->P->n) {
    fprintf(stderr, "lar_compute_residual: bad matrix id or size\n");
    return;
  }
  if (flush(*M)) return;
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
// V <- D^-1 (B - (A - D) V), Nr_prec sweeps
void lar_perform_BJ_or_GS_iterations(int Matrix_id, int Use_rhs, int Ini_zero, int Nr_prec,
                                     int Ndof, double* V, double* B) {
  (void)Ini_zero;
  LarMatrix* M = lar_get(Matrix_id);
  if (!M || Ndof != M->P->n || M->precon != BLOCK_JACOBI || !M->P->d_dinv) {
    fprintf(stderr, "lar_perform_BJ_or_GS_iterations: needs a filled BLOCK_JACOBI preconditioner\n");
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
  for (int it = 0; it < Nr_prec; it++) {
    if (launch_rowop(P, 1, Use_rhs == 1, sv.d, db, P.d_diag_ptr, tmp)) break;
    if (hipMemcpyAsync(sv.d, tmp, (size_t)Ndof * sizeof(double), hipMemcpyDeviceToDevice,
                       stream()) != hipSuccess)
      break;
  }
  stage_out(V, sv, Ndof);
  (void)hipFree(tmp);
}

// lar_perform_rhsub: the CRS body is empty in the reference
// (las_crs_intf.c:1019-1057); kept as a no-op with a diagnostic
void lar_perform_rhsub(int Matrix_id, int Ndof, double* V, double* B) {
  (void)Matrix_id; (void)Ndof; (void)V; (void)B;
  set_error("lar_perform_rhsub: no ILU for CRS storage (the reference body is empty)");
}

// host copies of the stored system (lar_get_SM_and_LV_crs_from_crs,
// las_crs_intf.c:1063-1120 with offset 0)
int mf_lar_get_system(int Matrix_id, int* row_ptr, int* col_ind, double* val, double* rhs) {
  LarMatrix* M = lar_get(Matrix_id);
  if (!M) return -3;
  if (flush(*M)) return -2;
  Problem& P = *M->P;
  if (row_ptr) memcpy(row_ptr, P.setup.row_ptr.data(), (P.n + 1) * sizeof(int));
  if (col_ind) memcpy(col_ind, P.setup.col_ind.data(), P.nnz * sizeof(int));
  if (val) MF_HIP(hipMemcpy(val, P.d_val, P.nnz * sizeof(double), hipMemcpyDeviceToHost));
  if (rhs) MF_HIP(hipMemcpy(rhs, P.d_rhs, P.n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
