// Internal declarations of the MI355X-native ModFEM hot-path library.
// Host code is C++17; device code lives in the .hip translation units.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace mf {

constexpr int kPrism = 5;  // matches MMC_PRISM (mmh_intf.h)
constexpr int kTetra = 7;  // matches MMC_TETRA

inline int nodes_of(int t) { return t == kTetra ? 4 : 6; }

// error reporting: every C entry point returns <0 and records a message
void set_error(const std::string& msg);
const char* last_error();

// ----------------------------------------------------------------------
// Mesh: what the reference's approximation module obtains through
// mmr_get_next_elem_all / mmr_el_node_coor / mmr_fa_neig / mmr_fa_bc
// (mmd_prism/mms_prism_intf.c:385, mms_prism_datstr.c:365,993).
// Indices are 0-based; reference IDs are index+1 in the same order.
// ----------------------------------------------------------------------
struct Mesh {
  int nno = 0;
  std::vector<double> coor;   // 3*nno
  int nel = 0;
  std::vector<int> type;      // nel
  std::vector<int> nodes;     // 6*nel, element-local order, -1 padded
  int nbf = 0;                // boundary faces with bc>0, face-ID order
  std::vector<int> bf_elem, bf_side, bf_bc;
  // jk-structure (kept when the mesh came from / maps onto a jk file)
  bool has_jk = false;
  int jk_nno2d = 0, jk_ntri = 0, jk_nlay = 0, jk_bottom_bc = 0, jk_top_bc = 0;
  double jk_zb = 0, jk_zt = 0;
  std::vector<double> jk_xy;   // 2*nno2d
  std::vector<int> jk_tri;     // 7 per triangle: mat n0 n1 n2 nb0 nb1 nb2 (1-based)
};

int mesh_from_jk_struct(Mesh& m);  // builds 3D mesh from the jk_* fields
int mesh_read_jk(const char* fname, Mesh& m);
int mesh_read_nas(const char* fname, Mesh& m);
int mesh_read_dat(const char* fname, Mesh& m);
int mesh_write_jk(const Mesh& m, const char* fname);
int mesh_box_prism(int nx, int ny, int nz, double lx, double ly, double lz,
                   uint64_t seed, double jitter, int bc, Mesh& m);
int mesh_box_tetra(int nx, int ny, int nz, double lx, double ly, double lz,
                   uint64_t seed, double jitter, int bc, Mesh& m);
int mesh_box_hybrid(int nx, int ny, int nz, double lx, double ly, double lz,
                    uint64_t seed, double jitter, int bc, double tet_frac, Mesh& m);

// ----------------------------------------------------------------------
// Setup (integer, bit-exact with the reference):
//   int-ent list   utr_get_list_int_ent   (utd_util/uts_ls_intf.c:102-172)
//   DOF list       utr_get_list_dof_ent   (uts_ls_intf.c:229-398)
//   neighbours     sir_create             (sid_mkb/sis_mkb_intf.c:762-1142)
//   CRS pattern    lar_allocate_SM_and_LV_crs (lsd_mkb/lad_crs/las_crs_intf.c:72-294)
//   tables         lar_fill_assembly_table_int_ent_crs (las_crs_intf.c:354-432)
// ----------------------------------------------------------------------
struct Setup {
  int n = 0;                       // DOFs (block size 1)
  std::vector<int> dof_node;       // n: node of DOF
  std::vector<int> node_dof;       // nno: DOF of node or -1
  std::vector<int> row_ptr;        // n+1
  std::vector<int> col_ind;        // nnz
  long long nnz = 0;
  int max_row = 0;
  // element connectivity in DOF numbering, SoA [k][nel] (k<6; tets pad -1)
  std::vector<int> el_dof;         // 6*nel
  // compact assembly table: for element e and pair (w,u) the offset of
  // column dof(u) inside row dof(w): off[(w*6+u)*nel + e]
  std::vector<uint8_t> el_off;     // 36*nel
  std::vector<uint8_t> bf_off;     // 36*nbf (faces reuse owner's DOFs)
  // device-native SELL-64 layout of the same matrix in device row order
  // (perm below): slice s holds device rows 64s..64s+63 column-major, padded
  // to its longest row (padding: column = own row, value 0).  Entry k of
  // device row i lives at sl_ptr[i/64] + 64*k + i%64; columns are device rows.
  std::vector<long long> sl_ptr;   // ns+1
  std::vector<int> sl_col;         // sl_ptr[ns]
  std::vector<uint8_t> diag_k;     // n: position of the diagonal in its row
  // DOF -> integration-entity incidence in ascending entity order (the
  // reference's summation order): elements, and boundary faces through their
  // owning element; loc = local index of the DOF in that element
  std::vector<int> inc_ptr, inc_el;
  std::vector<uint8_t> inc_loc;
  std::vector<int> finc_ptr, finc_f;
  std::vector<uint8_t> finc_loc;
  // device-internal numbering (no counterpart in the reference, whose DOF
  // and element numbering is kept at every host boundary): device row i holds
  // reference DOF perm[i] and device element slot e holds reference element
  // eperm[e]; iperm / ieperm are the inverses.  Rows keep the reference's
  // entry order, so per-entry and per-row sums are unchanged by it.
  std::vector<int> perm, iperm, eperm, ieperm;
};

// SELL-64 arrays (sl_ptr, sl_col, diag_k) in device row order
int build_sell(Setup& s);
// locality order: elements along a Morton curve of their centroids, DOFs by
// first appearance over elements in that order; identity when !mesh
int build_device_order(const Mesh* m, Setup& s);

int build_setup(const Mesh& m, Setup& s);

// single-level solver control (lsd_mkb/lsd_mkb_core/lsh_mkb_core.h:114-153)
struct MkbParams {
  int solver = 0, nr_level = 0, krylbas = 0, precon = 0, nr_prec = 0, block_type = 0;
  int storage_type = -1, ilu_k = 0, max_iter = 0, conv_type = 0, monitor = 0;
  int nr_pre_smooth = 0, nr_post_smooth = 0;
  double conv_meas = 0.0;
};
// mkb.dat reader (lss_mkb_core.c:148-705, single-level branch); arguments
// > 0 override the file, as in the reference
int mkb_read(const char* filename, int max_iter, int error_type, double error_tol,
             int monitoring, MkbParams& p);
// full reference-format table for int-ent ie (elements then faces)
int fill_reference_table(const Mesh& m, const Setup& s, int ie, int* table);

}  // namespace mf
