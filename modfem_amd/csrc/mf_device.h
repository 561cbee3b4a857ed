// Device-side data structures shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mf_internal.h"

namespace mf {

#define MF_HIP(call)                                                       \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess) {                                                \
      ::mf::set_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                      " at " + __FILE__ + ":" + std::to_string(__LINE__)); \
      return -2;                                                           \
    }                                                                      \
  } while (0)

// coefficient storage models (per-element arrays are SoA [slot][nel])
enum CoefModel {
  CM_UNIFORM = 0,  // all 21 values from coef_const
  CM_KAPPA = 1,    // a = kappa_e * I (diffusion branch), rest from coef_const
  CM_KAPPA_B = 2,  // kappa_e and b_e[3]; rest from coef_const
  CM_FULL = 3,     // 21 values per element
  CM_KAPPA_S = 4   // kappa_e and source s_e; rest from coef_const
};
// coefficient slots (same as the oracle): a[9] b[3] t[3] c q[3] s m
constexpr int C_A = 0, C_B = 9, C_T = 12, C_C = 15, C_Q = 16, C_S = 19, C_M = 20,
              NCOEF = 21;
// term sets compiled into the integration kernel
enum TermSet { TS_DIFF = 0, TS_DIFF_CONV = 1, TS_ALL = 2 };
// source models
enum SrcModel { SRC_CONST = 0, SRC_DIFF_IN_CUBE = 1, SRC_DIFF_IN_BOX = 2 };
// boundary-face types (pdr_get_bc_type; PDC_BC_DIRI / PDC_BC_NEUM)
enum BcType { BC_DIRICHLET = 1, BC_NEUMANN = 2 };

// exact solution u = exp(-|x|^2 / a1) of the DIFF_IN_CUBE (a1 = 1) and
// DIFF_IN_BOX (a1 = 100) problems and its Laplacian; the source is -lapl
// (pdr_exact_sol, pdd_conv_diff/main/pds_conv_diff_util.c)
__host__ __device__ inline void exact_sol(int model, const double* x, double* u,
                                          double* lapl) {
  double a1 = (model == SRC_DIFF_IN_BOX) ? 100.0 : 1.0;
  double X = x[0], Y = x[1], Z = x[2];
  double e = exp((-X * X - Y * Y - Z * Z) / a1);
  *u = e;
  *lapl = -(2.0 / a1) *
          ((1 - 2 * X * X / a1) + (1 - 2 * Y * Y / a1) + (1 - 2 * Z * Z / a1)) * e;
}

// exact normal flux grad u . ds / |ds| of the DIFF_IN_CUBE / DIFF_IN_BOX
// solution, the Neumann data of those problems (pdr_bc_neum_coeff,
// pdd_conv_diff/weak_formulation/pds_conv_diff_weakform.c; gradient as
// pdr_exact_sol, pds_conv_diff_util.c:1195-1260); ds is the face's area
// vector of apr_elem_calc_3D
__host__ __device__ inline double exact_flux(int model, const double* x, const double* ds) {
  double a1 = (model == SRC_DIFF_IN_BOX) ? 100.0 : 1.0;
  double X = x[0], Y = x[1], Z = x[2];
  double fx = -2.0 * X * exp((-X * X - Y * Y - Z * Z) / a1) / a1;
  double fy = -2.0 * Y * exp((-X * X - Y * Y - Z * Z) / a1) / a1;
  double fz = -2.0 * Z * exp((-X * X - Y * Y - Z * Z) / a1) / a1;
  double len = sqrt(ds[0] * ds[0] + ds[1] * ds[1] + ds[2] * ds[2]);
  return fx * (ds[0] / len) + fy * (ds[1] / len) + fz * (ds[2] / len);
}

// selected-term bits (reference coefficient indicator slots 1..22)
struct TermMask {
  int m, diff_full, diff_diag, conv, tconv, reac, q, s;
};

struct GmresStat {
  double beta0;     // norm of the first preconditioned residual (restart 0)
  double rnrm;      // norm at the start of the current restart
  double resid;     // current |g_{j+1}|
  double res_rel;
  double tol_rel, tol_res;
  int j_done;       // Arnoldi steps completed in this restart
  int gmiter;       // total iterations
  int max_iter;
  int stop;         // 1: stop this restart
  int converged;
  // restart bookkeeping kept on the device by the batched solver
  int restit, restrt, m;
  int done;         // 1: this problem's solve has finished
  int nan;          // 1: residual was not a number
  int ik_pending;   // coefficients computed this cycle, not yet applied to x
  double beta1;
};

// ILU(0) preconditioner on the device (mf_ilu.hip): the pattern in an
// elimination order (0: the reference's DOF order, 1: multicolour), rows of
// each forward / backward level, and the factors (strict L, strict U, and
// 1/u_ii on the diagonal as the reference keeps it).
struct IluData {
  int order = 0, n = 0;
  long long nnz = 0;
  int nlev_l = 0, nlev_u = 0, ncolors = 0;
  std::vector<int> lev_ptr_l, lev_ptr_u;   // host: level offsets
  std::vector<long long> ref_pos;          // host: reference CRS position of each entry
  std::vector<int> pos_of_dev;             // host: elimination position of a device row
  int *d_rp = nullptr, *d_ci = nullptr, *d_dg = nullptr;
  int *d_lev_l = nullptr, *d_lev_u = nullptr;
  long long* d_src = nullptr;              // SELL position of every entry
  double* d_lu = nullptr;
  bool valid = false;                      // factors match the current values
  // substitution copies of the factors, SELL-64 packed per level: forward
  // (strict L) and backward (strict U, 1/u_ii); slice s of a sweep holds 64
  // rows of one level, entry k of its row r at ptr[s] + 64 k + r % 64
  struct Sweep {
    std::vector<int> lev_sl;               // host: first slice of every level
    long long* d_ptr = nullptr;            // slices + 1
    int* d_col = nullptr;                  // device row of the column
    int* d_src = nullptr;                  // factor entry, -1 padding
    double* d_val = nullptr;
    int* d_row = nullptr;                  // device row, -1 padding
    int* d_len = nullptr;
    double* d_diag = nullptr;              // backward: 1/u_ii per packed row
    int* d_dpos = nullptr;                 // backward: factor entry of u_ii
    float* d_val32 = nullptr;              // FP32 copies for inner mixed-precision solves
    float* d_diag32 = nullptr;
    long long nent = 0;
    int nrows = 0;
  } fwd, bwd;
  bool valid32 = false;                    // FP32 copies match the factors
};

// Multicolour Gauss-Seidel (mf_gs.hip): off-diagonal rows of every colour
// packed in SELL-64 slices (rows and columns as device rows, values copied
// from the matrix), swept colour by colour in place
struct GsData {
  int ncolors = 0, n = 0;
  std::vector<int> col_sl;                 // host: first slice of every colour
  std::vector<int> elim_pos;               // host: position of a device row in the order
  long long* d_ptr = nullptr;
  int* d_col = nullptr;
  long long* d_src = nullptr;              // SELL position, -1 padding
  double* d_val = nullptr;
  int* d_row = nullptr;
  int* d_len = nullptr;
  long long nent = 0;
  int nrows = 0;
  bool valid = false;
};

struct Problem {
  // host side
  Mesh mesh;
  Setup setup;
  int n = 0, nel = 0, nbf = 0;
  long long nnz = 0;
  int el_type = kPrism;  // uniform element type (mixed handled per launch)
  // device arrays
  double* d_coor = nullptr;    // 3*n, device row order
  int* d_el_dof = nullptr;     // 6*nel SoA
  uint8_t* d_el_off = nullptr; // 36*nel SoA
  int* d_bf_elem = nullptr;
  int* d_bf_side = nullptr;
  uint8_t* d_bf_off = nullptr;
  double* d_bf_g = nullptr;
  int* d_bf_type = nullptr;    // per face BcType (null: all Dirichlet)
  int* d_row_ptr = nullptr;
  int* d_col = nullptr;
  double* d_val = nullptr;     // SELL-64 values when sell, else CRS
  // device-native SELL-64 storage (mf_ problems) and its incidence lists
  bool sell = false;
  long long nnz_sell = 0;
  long long* d_sl_ptr = nullptr;
  int* d_scol = nullptr;
  uint8_t* d_diag_k = nullptr;
  int* d_inc_ptr = nullptr;
  int* d_inc_el = nullptr;
  uint8_t* d_inc_loc = nullptr;
  int* d_finc_ptr = nullptr;
  int* d_finc_f = nullptr;
  uint8_t* d_finc_loc = nullptr;
  uint8_t* d_el_offT = nullptr; // [nel][36] offsets (row-major per element)
  void* d_erow = nullptr;       // element-row buffers of the two-pass assembly
  size_t erow_bytes = 0;
  int assembly_mode = 0;       // 0: row-wise (deterministic), 1: element atomics
  double* d_rhs = nullptr;
  float* d_val32 = nullptr;
  // coefficients
  int coef_model = CM_UNIFORM, src_model = SRC_CONST, term_set = TS_DIFF;
  TermMask mask{};
  double coef_const[NCOEF] = {0};
  double* d_coef = nullptr;    // SoA per model
  // preconditioner
  double* d_dinv = nullptr;
  bool dinv_valid = false;     // d_dinv matches the current values
  IluData* ilu = nullptr;      // precon 2
  int ilu_order = 0;           // elimination order for ILU(0)
  GsData* gs = nullptr;        // precon 3
  int gs_sweeps = 1;
  int batch_refs = 0;          // batches this problem is a member of
  int* d_diag_ptr = nullptr;
  float* d_dinv32 = nullptr;
  uint16_t* d_scol16 = nullptr;  // FP32 operator: 16-bit columns of narrow slices
  int* d_sbase = nullptr;        // per slice: their base column, or -1
  bool val32_valid = false;    // d_val32 (and d_dinv32) are the current values rounded
  // solver tiles
  std::vector<int> tile_row;   // ntiles+1
  int* d_tile_row = nullptr;
  int ntiles = 0;
  // GMRES workspace
  int kryl_alloc = 0;
  double* d_V = nullptr;       // (m+1)*n
  double* d_H = nullptr;       // (m+1)*(m+1) + 2*(m+1) givens + scale
  double* d_partials = nullptr;
  unsigned* d_counter = nullptr;
  GmresStat* d_stat = nullptr;
  GmresStat* h_stat = nullptr; // pinned
  double* d_x = nullptr;
  double* d_work = nullptr;    // n
  double* d_work2 = nullptr;   // n + 1: mixed-precision correction and a reduction slot
  // last measured times (ms)
  double t_integrate_ms = 0, t_faces_ms = 0, t_solve_ms = 0;
  // operator of another discretisation (not owned); replaces the CRS/SELL
  // operator and preconditioner inside gmres_core
  struct ExtOp* ext = nullptr;
};

// Batched restarted GMRES over resident problems (mf_solver.hip).  Members
// keep their own matrices, vectors and Krylov bases; one launch of each
// kernel serves all members through a block table (block -> member, rank).
// An operator supplied by another discretisation (the DG block system of
// mf_dg.hip): GMRES calls residual() for V_0 = M^-1 (b - A x) and apply()
// for w = s_j M^-1 A V_j (s_j = scale[j] read on the device; nothing to do
// once *stop is set), then forms the CGS dots, updates and restarts itself.
struct ExtOp {
  virtual ~ExtOp() {}
  virtual int residual(const double* b, const double* x, double* r) = 0;
  virtual int apply(const double* v, const double* scale, int j, double* w, const int* stop) = 0;
};

struct BatchMember {
  Problem* P = nullptr;
  int pid = -1;
  int precon = 1, max_iter = 1000;
  double tol = 1e-8;
  // mixed-precision member: FP64 refinement steps around inner solves with
  // FP32 values, basis and preconditioner data; the inner solve runs on
  // (r, d) = (d_work, d_work2)
  bool mixed = false;
};
// refinement state of a batch member, kept on the device (host copy after
// every cycle)
struct MixedState {
  int mixed;
  int phase;            // 1: the first refinement step is due
  int fin, conv, nan;   // ended; met the test; not a number
  int total, outer;     // inner iterations and refinement steps so far
  int max_iter;
  double tol, rho0, rho;
};
struct Batch {
  int m = 30;
  std::vector<BatchMember> mem;
  std::vector<GmresStat> h_stat;   // host copy, current after every cycle
  bool dirty = true;               // tables must be rebuilt
  bool retile = true;              // membership changed: choose the tile again
  // device tables (valid when !dirty)
  void* d_args = nullptr;          // SolveArgs per member
  int* d_blk_prob = nullptr;
  int* d_blk_rank = nullptr;
  int* d_nblk = nullptr;
  GmresStat* d_stat = nullptr;
  GmresStat* h_pinned = nullptr;
  int nblk = 0, cap = 0, blk_cap = 0;
  size_t lds = 0;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  int cycles = 0;
  // batch tiles: slices of 64 rows per block of every member (0 = tuned)
  int spb = 0, spb_used = 8;
  bool spb_tuned = false;          // spb_used came from timed trial cycles
  float tile_ms[8] = {};           // trial cycle times of the tile candidates
  // ILU(0) members: per-member factor pointers and, per substitution level,
  // the blocks (member, slice range) of every member that has the level
  bool has_ilu = false;
  void* d_ilu = nullptr;           // IluMember per member
  int* d_ilu_mem = nullptr;
  int* d_ilu_s0 = nullptr;
  int* d_ilu_s1 = nullptr;
  std::vector<int> ilu_off_f, ilu_off_b;  // level l's blocks: [off[l], off[l+1])
  // refinement steps of mixed-precision members
  bool has_mixed = false;
  std::vector<MixedState> h_mix;   // host copy, current after every cycle
  MixedState* d_mix = nullptr;
  MixedState* h_mix_pinned = nullptr;
  void* d_ref = nullptr;           // RefineArgs per member
  // own stream (optional): cycles run there after the library stream's work
  hipStream_t own = nullptr;
  hipEvent_t ev_ready = nullptr, ev_done = nullptr;
  bool pending = false;            // an asynchronous cycle has not been waited for
};
// tile candidates the tuner times (slices per block)
constexpr int kBatchTileCands[] = {2, 4, 8, 16, 32};
constexpr int kBatchTileNum = 5;
int batch_set_tiles(Batch& B, int spb);
void batch_tile_cache_clear();
int batch_add(Batch& B, Problem* P, int pid, int precon, int max_iter, double tol,
              bool mixed = false);
int batch_remove(Batch& B, int pid);
int batch_run(Batch& B, int max_cycles, int use_graph, int* active);
int batch_run_async(Batch& B, int use_graph);
// the batch form of a DG problem (block-Jacobi-preconditioned SELL system),
// made current; null and an error if the problem cannot join a batch
Problem* dg_batch_problem(int did);
int batch_wait(Batch& B, int* active);
int batch_own_stream(Batch& B);
void batch_free(Batch& B);

int device_init(int device);
hipStream_t stream();
// mesh object behind a C-ABI mesh handle (null and an error if unknown)
Mesh* mesh_by_id(int id);

// kernels (mf_device.hip)
int launch_zero(double* p, long long n);
// elements e0..e0+ecount-1 (device slots), or d_elist[0..ecount) when given
int launch_integrate(Problem& P, int precision, int write_elements,
                     double* d_emat, int e0, int ecount, const int* d_elist = nullptr);
int launch_assemble_rows(Problem& P, int precision);
int launch_assemble_two_pass(Problem& P, int precision);
// row-wise (mode 0) and two-pass (mode 2) problems with rows of at most 32
// entries, one launch per kernel and problem group; on the library stream
int launch_assemble_many(const std::vector<Problem*>& ps, int precision);
// the set-up maps of a std-linear problem built on the device, bitwise the
// host set-up's; fills the problem's device arrays and the host maps its
// later calls read (mf_setup_dev.hip)
int problem_setup_device(const Mesh& M, Problem& P);
int launch_faces(Problem& P, int write_faces, double* d_fmat, int f0, int fcount);
int launch_fill_jacobi(Problem& P);
// tol: relative to the initial preconditioned residual; tol_abs: absolute
int gmres_solve(Problem& P, int precon, int krylbas, int max_iter, double tol,
                double tol_abs, int ini_zero, double* h_x, int* iters, double* res_rel,
                double* resid);
// GMRES(m) with precon 0/1 whose work arrays live in a fixed device array
// (no device allocation); noalloc_scratch_init looks the array up for the
// current device (called by mf_init)
int noalloc_scratch_init(int device);
int gmres_noalloc(Problem& P, int precon, int krylbas, int max_iter, double tol, int ini_zero,
                  double* h_x, int* iters, double* res_rel, double* resid);
// FP64 refinement around GMRES with an FP32-valued operator
int gmres_mixed(Problem& P, int precon, int krylbas, int max_iter, double tol, double* h_x,
                int* iters, double* res_rel, double* resid, int* refinements);
int launch_spmv(Problem& P, const double* d_x, double* d_y);
// ms per operator application of a precision variant (mf_bench_spmv_ex)
int bench_spmv_variant(Problem& P, int prec, int ch, int reps, double* ms);
// ILU(0) (mf_ilu.hip): structure for an elimination order, factorization
// from the current values, and out = (LU)^-1 in (device-order vectors; out
// may alias in); stop: skip when *stop is set (GMRES stop flag) or null
int ilu_build(Problem& P, int order);
int ilu_factor(Problem& P);
int ilu_apply(Problem& P, const double* in, double* out, const int* stop);
// FP32 copies of the packed factors (mixed-precision inner solves), and
// the substitution on FP32 vectors with them
int ilu_round_f32(Problem& P);
int ilu_apply32(Problem& P, const float* in, float* out, const int* stop);
void ilu_free(Problem& P);
int greedy_colors(const Setup& S, std::vector<int>& color);
// multicolour Gauss-Seidel (mf_gs.hip): structure, value refresh, and
// nsweeps sweeps on v (device order) with rhs b (null: zero rhs)
int gs_build(Problem& P);
int gs_refresh(Problem& P);
int gs_sweeps(Problem& P, int nsweeps, const double* b, double* v, const int* stop);
void gs_free(Problem& P);
// CRS Gauss-Seidel as the reference computes it on nthreads OpenMP threads
// (row chunks of n/nthreads + 1, mf_gs.hip) on a CRS problem (row_ptr, col,
// val, diag_ptr, dinv): nr_prec sweeps on v with rhs b (null: none), tmp n
// doubles; and the GMRES operator preconditioned by it
constexpr int kMaxGsChunks = 1 << 20;
int crs_gs_chunked(Problem& P, int nthreads, int nr_prec, const double* b, double* v, double* tmp,
                   const int* stop);
struct ExtOp* crs_gs_make_op(Problem* P, int nthreads, int nr_prec);
// BCRS storage with 4x4 column-major blocks (mf_bcrs.hip): block rows in
// CRS order with the diagonal in place, rhs, and for the diagonal-block
// preconditioners the LU factors and pivots of every diagonal block and the
// Gauss-Seidel level schedules of the forward (0) and backward (1) sweeps
struct BcrsDev {
  int nb = 0;                     // block rows
  long long nblk = 0;             // stored blocks
  int max_row = 0;                // most blocks in one row
  int *d_rp = nullptr, *d_ci = nullptr, *d_diag = nullptr;
  double *d_val = nullptr, *d_rhs = nullptr, *d_lu = nullptr;
  double* d_ilu = nullptr;        // block ILU(0) factors (16 per stored block)
  int* d_piv = nullptr;
  std::vector<int> lev_ptr[2];
  int* d_lev_rows[2] = {nullptr, nullptr};
  // Gauss-Seidel order: 0 the reference's row order (level schedules), 1
  // colour classes of a greedy colouring (rows of a class share no block)
  int gs_order = 0;
  std::vector<int> col_ptr;
  int* d_col_rows = nullptr;
  // block ILU(k), k > 0: the grown pattern (ilu_k = 0 factors on the
  // matrix's own pattern), the matrix block behind every ILU block (-1 for
  // fill) and the level schedules of that pattern
  int ilu_k = 0;
  long long ilu_nblk = 0;
  int ilu_max_row = 0;
  int *d_ilu_rp = nullptr, *d_ilu_ci = nullptr, *d_ilu_diag = nullptr, *d_ilu_src = nullptr;
  std::vector<int> ilu_lev_ptr[2];
  int* d_ilu_lev_rows[2] = {nullptr, nullptr};
};
// ILU(k) pattern growth (host, lar_allocate_preconditioner_bcrs); returns 0
// or the row (>= 0) whose list outgrew cap
int bcrs_iluk_pattern(int nb, const std::vector<int>& rp, const std::vector<int>& ci, int k,
                      int cap, std::vector<int>& prp, std::vector<int>& pci, int* bad_row);
// greedy colouring of the block graph in row order; rows grouped by colour
int bcrs_colors(int nb, const std::vector<int>& rp, const std::vector<int>& ci,
                std::vector<int>& col_ptr, std::vector<int>& rows);
// level schedule of a sweep in direction dir; returns the number of levels
int bcrs_levels(int nb, const std::vector<int>& rp, const std::vector<int>& ci,
                const std::vector<int>& diag, int dir, std::vector<int>& lev_ptr,
                std::vector<int>& rows);
// v = (b ? b : 0) - A x (x null: zero); kernels skip their work once *stop
// is set (stop may be null)
int bcrs_residual(const BcrsDev& M, const double* x, const double* b, double* v,
                  const int* stop = nullptr);
// LU factors of the diagonal blocks (d_lu, d_piv allocated)
int bcrs_fill_lu(BcrsDev& M);
// block Jacobi (gs = 0) or Gauss-Seidel sweeps with the sweep-count rule of
// the reference; tmp: 4 nb doubles
int bcrs_sweeps(BcrsDev& M, int gs, int nr_prec, const double* b, double* v, double* tmp,
                const int* stop = nullptr, bool sync = true);
// block ILU(0) factors of the current values into d_ilu (levels of both
// directions built), and v = (LU)^-1 b (v may alias b)
int bcrs_ilu_factor(BcrsDev& M);
int bcrs_ilu_solve(const BcrsDev& M, const double* b, double* v, const int* stop = nullptr);
// GMRES operator on a BCRS matrix with preconditioner kind 0 (none), 1
// (block Jacobi), 2 (block Gauss-Seidel), 4 (MULTI_ILU), 5 (BLOCK_ILU);
// owned by the caller
struct ExtOp* bcrs_make_op(BcrsDev* B, int kind, int nr_prec);
int bench_copy(long long nbytes, int reps, double* ms);
int build_tiles(Problem& P);
int launch_rowop(Problem& P, int mode, int use_rhs, const double* d_x, const double* d_b,
                 const int* d_diag_ptr, double* d_v);
int launch_diag_ptr(Problem& P, int* d_dp);
void kstats_enable(int on);
void kstats_get(double* out);

}  // namespace mf
