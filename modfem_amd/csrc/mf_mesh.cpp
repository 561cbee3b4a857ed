#include <sstream>
#include <fstream>
#include <array>
// Mesh construction: the jk ("gradmesh") half-structured prism format and
// seeded box generators.  The 3D mesh built from a jk structure follows the
// reference importer mmr_import_mesh_grad (mmd_prism/
// mms_prism_input_gradmesh.c:29-600) and element node order of
// mmr_el_node_coor (mmd_prism/mms_prism_datstr.c:365-540):
//   node ID     = istr + layer*N2D
//   element ID  = itri + layer*Ntri, nodes (n0,n1,n2) bottom then top
//   face IDs    = horizontal faces itri + layer*Ntri (layer 0..L), then
//                 vertical faces in creation order (triangle, edge), L each
//   boundary    = faces with bc>0 in face-ID order; owner = face neighbour 0
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <random>
#include <utility>

#include "mf_internal.h"

namespace mf {

static std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

int mesh_from_jk_struct(Mesh& m) {
  const int N2 = m.jk_nno2d, Nt = m.jk_ntri, L = m.jk_nlay;
  if (N2 <= 0 || Nt <= 0 || L <= 0) {
    set_error("jk mesh: empty structure");
    return -1;
  }
  const double dz = (m.jk_zt - m.jk_zb) / L;
  m.nno = N2 * (L + 1);
  m.coor.assign(3 * (size_t)m.nno, 0.0);
  for (int istr = 0; istr < N2; istr++)
    for (int lay = 0; lay <= L; lay++) {
      size_t id = istr + (size_t)lay * N2;
      m.coor[3 * id] = m.jk_xy[2 * istr];
      m.coor[3 * id + 1] = m.jk_xy[2 * istr + 1];
      m.coor[3 * id + 2] = m.jk_zb + lay * dz;
    }
  m.nel = Nt * L;
  m.type.assign(m.nel, kPrism);
  m.nodes.assign(6 * (size_t)m.nel, -1);
  for (int t = 0; t < Nt; t++) {
    const int* tr = &m.jk_tri[7 * (size_t)t];
    if (tr[0] <= 0) {
      set_error("jk mesh: only triangles (prisms) are supported");
      return -1;
    }
    for (int lay = 0; lay < L; lay++) {
      size_t e = t + (size_t)lay * Nt;
      for (int i = 0; i < 3; i++) {
        m.nodes[6 * e + i] = (tr[1 + i] - 1) + lay * N2;
        m.nodes[6 * e + 3 + i] = (tr[1 + i] - 1) + (lay + 1) * N2;
      }
    }
  }
  // boundary faces in face-ID order
  m.bf_elem.clear(); m.bf_side.clear(); m.bf_bc.clear();
  if (m.jk_bottom_bc > 0)
    for (int t = 0; t < Nt; t++) {
      m.bf_elem.push_back(t); m.bf_side.push_back(0); m.bf_bc.push_back(m.jk_bottom_bc);
    }
  if (m.jk_top_bc > 0)
    for (int t = 0; t < Nt; t++) {
      m.bf_elem.push_back(t + (L - 1) * Nt); m.bf_side.push_back(1);
      m.bf_bc.push_back(m.jk_top_bc);
    }
  // vertical faces: created by the first triangle (in jk order) that meets
  // the edge; boundary edges (neig<0) always create one.
  std::map<std::pair<int, int>, int> created;  // (min,max tri) -> 1
  for (int t = 0; t < Nt; t++) {
    const int* tr = &m.jk_tri[7 * (size_t)t];
    for (int ino = 0; ino < 3; ino++) {
      int neig = tr[4 + ino];
      if (neig > 0) {
        auto key = std::make_pair(std::min(t + 1, neig), std::max(t + 1, neig));
        if (!created.count(key)) created[key] = 1;
      } else if (neig < 0) {
        for (int lay = 0; lay < L; lay++) {
          m.bf_elem.push_back(t + lay * Nt);
          m.bf_side.push_back(2 + ino);
          m.bf_bc.push_back(-neig);
        }
      }
    }
  }
  m.nbf = (int)m.bf_elem.size();
  m.has_jk = true;
  return 0;
}

int mesh_read_jk(const char* fname, Mesh& m) {
  FILE* fp = fopen(fname, "r");
  if (!fp) {
    set_error(std::string("cannot open jk mesh file ") + fname);
    return -1;
  }
  int mx[4];
  int ok = fscanf(fp, "%d %d %d %d", &mx[0], &mx[1], &mx[2], &mx[3]) == 4;
  ok = ok && fscanf(fp, "%d %lg %lg %d %d %d", &m.jk_nno2d, &m.jk_zb, &m.jk_zt,
                    &m.jk_nlay, &m.jk_bottom_bc, &m.jk_top_bc) == 6;
  if (ok && m.jk_nlay < 1) m.jk_nlay = 1;
  if (ok) {
    m.jk_xy.resize(2 * (size_t)m.jk_nno2d);
    for (int i = 0; ok && i < m.jk_nno2d; i++)
      ok = fscanf(fp, "%lg %lg", &m.jk_xy[2 * i], &m.jk_xy[2 * i + 1]) == 2;
  }
  ok = ok && fscanf(fp, "%d", &m.jk_ntri) == 1;
  if (ok) {
    m.jk_tri.resize(7 * (size_t)m.jk_ntri);
    for (int t = 0; ok && t < m.jk_ntri; t++) {
      int* tr = &m.jk_tri[7 * (size_t)t];
      ok = fscanf(fp, "%d %d %d %d %d %d %d", &tr[0], &tr[1], &tr[2], &tr[3],
                  &tr[4], &tr[5], &tr[6]) == 7;
    }
  }
  fclose(fp);
  if (!ok) {
    set_error(std::string("malformed jk mesh file ") + fname);
    return -1;
  }
  return mesh_from_jk_struct(m);
}

int mesh_write_jk(const Mesh& m, const char* fname) {
  if (!m.has_jk) {
    set_error("mesh has no jk structure");
    return -1;
  }
  FILE* fp = fopen(fname, "w");
  if (!fp) {
    set_error(std::string("cannot write ") + fname);
    return -1;
  }
  fprintf(fp, "%d %d %d %d\n", m.nno, 4 * m.nno, 5 * m.nel, m.nel);
  fprintf(fp, "%d %.17g %.17g %d %d %d\n", m.jk_nno2d, m.jk_zb, m.jk_zt,
          m.jk_nlay, m.jk_bottom_bc, m.jk_top_bc);
  for (int i = 0; i < m.jk_nno2d; i++)
    fprintf(fp, "%.17g %.17g\n", m.jk_xy[2 * i], m.jk_xy[2 * i + 1]);
  fprintf(fp, "%d\n", m.jk_ntri);
  for (int t = 0; t < m.jk_ntri; t++) {
    const int* tr = &m.jk_tri[7 * (size_t)t];
    fprintf(fp, "%d %d %d %d %d %d %d\n", tr[0], tr[1], tr[2], tr[3], tr[4],
            tr[5], tr[6]);
  }
  fclose(fp);
  return 0;
}

// Seeded permutation (Fisher-Yates over mt19937_64, uniform via modulo of
// 64-bit draws; deterministic across platforms).
static std::vector<int> seeded_perm(int n, std::mt19937_64& rng) {
  std::vector<int> p(n);
  for (int i = 0; i < n; i++) p[i] = i;
  for (int i = n - 1; i > 0; i--) {
    int j = (int)(rng() % (uint64_t)(i + 1));
    std::swap(p[i], p[j]);
  }
  return p;
}

static double unit_draw(std::mt19937_64& rng) {
  return (double)(rng() >> 11) * (1.0 / 9007199254740992.0);  // [0,1)
}

static void jitter_interior(Mesh& m, double lx, double ly, double lz, int nx,
                            int ny, int nz, double jitter, uint64_t seed) {
  if (jitter <= 0) return;
  std::mt19937_64 rng(seed ^ 0x9e3779b97f4a7c15ULL);
  const double h[3] = {lx / nx, ly / ny, lz / nz}, L[3] = {lx, ly, lz};
  const double eps = 1e-9;
  for (int v = 0; v < m.nno; v++) {
    double* x = &m.coor[3 * (size_t)v];
    bool interior = true;
    for (int d = 0; d < 3; d++)
      if (x[d] < eps * L[d] || x[d] > L[d] * (1 - eps)) interior = false;
    double r[3] = {unit_draw(rng), unit_draw(rng), unit_draw(rng)};
    if (!interior) continue;
    for (int d = 0; d < 3; d++) x[d] += (2.0 * r[d] - 1.0) * jitter * h[d];
  }
}

int mesh_box_prism(int nx, int ny, int nz, double lx, double ly, double lz,
                   uint64_t seed, double jitter, int bc, Mesh& m) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error("box mesh: nx, ny, nz must be >= 1");
    return -1;
  }
  std::mt19937_64 rng(seed);
  const int N2 = (nx + 1) * (ny + 1), Nt = 2 * nx * ny;
  std::vector<int> pn = seeded_perm(N2, rng);  // grid node -> jk id-1
  std::vector<int> pt = seeded_perm(Nt, rng);  // grid tri  -> jk id-1
  std::vector<int> inv_t(Nt);
  for (int i = 0; i < Nt; i++) inv_t[pt[i]] = i;
  // grid triangles, counter-clockwise, split by the (0,0)-(1,1) diagonal
  std::vector<int> tri(3 * (size_t)Nt);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int a = i + j * (nx + 1), b = a + 1, c = a + (nx + 1) + 1, d = a + (nx + 1);
      int t = 2 * (i + j * nx);
      tri[3 * t] = a; tri[3 * t + 1] = b; tri[3 * t + 2] = c;
      tri[3 * (t + 1)] = a; tri[3 * (t + 1) + 1] = c; tri[3 * (t + 1) + 2] = d;
    }
  // edge -> triangles
  std::map<std::pair<int, int>, std::vector<int>> em;
  for (int t = 0; t < Nt; t++)
    for (int e = 0; e < 3; e++) {
      int u = tri[3 * t + e], v = tri[3 * t + (e + 1) % 3];
      em[std::make_pair(std::min(u, v), std::max(u, v))].push_back(t);
    }
  m = Mesh();
  m.jk_nno2d = N2; m.jk_ntri = Nt; m.jk_nlay = nz;
  m.jk_zb = 0.0; m.jk_zt = lz; m.jk_bottom_bc = bc; m.jk_top_bc = bc;
  m.jk_xy.assign(2 * (size_t)N2, 0.0);
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      int g = i + j * (nx + 1);
      m.jk_xy[2 * (size_t)pn[g]] = lx * i / nx;
      m.jk_xy[2 * (size_t)pn[g] + 1] = ly * j / ny;
    }
  m.jk_tri.assign(7 * (size_t)Nt, 0);
  for (int jt = 0; jt < Nt; jt++) {
    int t = inv_t[jt];
    int* out = &m.jk_tri[7 * (size_t)jt];
    out[0] = 1;
    for (int e = 0; e < 3; e++) {
      int u = tri[3 * t + e], v = tri[3 * t + (e + 1) % 3];
      out[1 + e] = pn[u] + 1;
      const auto& owners = em[std::make_pair(std::min(u, v), std::max(u, v))];
      int nb = -bc;
      for (int o : owners)
        if (o != t) nb = pt[o] + 1;
      out[4 + e] = nb;
    }
  }
  int rc = mesh_from_jk_struct(m);
  if (rc < 0) return rc;
  jitter_interior(m, lx, ly, lz, nx, ny, nz, jitter, seed);
  return 0;
}

// Tetrahedral box: every hex cell split into 6 tetrahedra around its main
// diagonal (Kuhn split, conforming).  Node and element IDs are permuted by the
// seed; boundary faces are listed element by element, local face order
// (mmd_prism/mms_prism_intf.c:97 face-node table {0,1,2},{0,1,3},{0,2,3},{1,2,3}).
// NASTRAN bulk data with tetrahedra (utr_io_read_nas, utd_util/
// uts_io_intf.cpp:812-1020), free field format: "GRID id x y z [cd]" nodes
// in id order 1..n, "CTRIAX|CTRIA3 id bc n1 n2 n3" boundary triangles (bc > 0
// kept as boundary faces in file order, bc <= 0 interior), "CTETRA id group
// n1 n2 n3 n4" elements in file order.  As the reference, element vertices
// are sorted ascending and then oriented positively by swapping local
// vertices 1 and 2 for a negative volume (ElemT4::init, mmd_t4_prism/
// MeshModule/ElemT4.cpp:24-26), the element-local order the face table of
// the hybrid mesh module expects.  Prisms (CPENTA, CPRIZM), quads and the
// fixed-field formats are not read.
int mesh_read_nas(const char* fname, Mesh& m) {
  std::ifstream in(fname);
  if (!in) {
    set_error(std::string("cannot open ") + fname);
    return -1;
  }
  m = Mesh();
  std::vector<std::array<int, 4>> tets;
  std::vector<std::array<int, 4>> tris;  // n1 n2 n3 bc
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ls(line);
    std::string key;
    if (!(ls >> key)) continue;
    if (key == "GRID") {
      long id;
      double x[3];
      if (!(ls >> id >> x[0] >> x[1] >> x[2]) || id != m.nno + 1) {
        set_error("NASTRAN: free-format GRID lines with ids 1..n in order expected");
        return -1;
      }
      m.coor.insert(m.coor.end(), x, x + 3);
      m.nno++;
    } else if (key == "CTRIAX" || key == "CTRIA3") {
      long id;
      int bc, v[3];
      if (!(ls >> id >> bc >> v[0] >> v[1] >> v[2])) {
        set_error("NASTRAN: bad triangle line");
        return -1;
      }
      if (bc > 0) tris.push_back({v[0], v[1], v[2], bc});
    } else if (key == "CTETRA") {
      long id;
      int grp;
      std::array<int, 4> v;
      if (!(ls >> id >> grp >> v[0] >> v[1] >> v[2] >> v[3])) {
        set_error("NASTRAN: bad CTETRA line");
        return -1;
      }
      tets.push_back(v);
    } else if (key == "CPENTA" || key == "CPRIZM" || key == "CQUADX" || key == "CQUAD4") {
      set_error("NASTRAN: only tetrahedra and triangles are read");
      return -1;
    }
  }
  if (m.nno == 0 || tets.empty()) {
    set_error("NASTRAN: no nodes or no tetrahedra");
    return -1;
  }
  m.nel = (int)tets.size();
  m.type.assign(m.nel, kTetra);
  m.nodes.assign(6 * (size_t)m.nel, -1);
  for (int e = 0; e < m.nel; e++) {
    std::array<int, 4> v = tets[e];
    std::sort(v.begin(), v.end());
    for (int a = 0; a < 4; a++)
      if (v[a] < 1 || v[a] > m.nno) {
        set_error("NASTRAN: element vertex out of range");
        return -1;
      }
    const double* p[4];
    for (int a = 0; a < 4; a++) p[a] = &m.coor[3 * (size_t)(v[a] - 1)];
    double a1[3], a2[3], a3[3];
    for (int d = 0; d < 3; d++) {
      a1[d] = p[1][d] - p[0][d]; a2[d] = p[2][d] - p[0][d]; a3[d] = p[3][d] - p[0][d];
    }
    const double det = a1[0] * (a2[1] * a3[2] - a2[2] * a3[1]) -
                       a1[1] * (a2[0] * a3[2] - a2[2] * a3[0]) +
                       a1[2] * (a2[0] * a3[1] - a2[1] * a3[0]);
    if (det < 0) std::swap(v[1], v[2]);
    for (int a = 0; a < 4; a++) m.nodes[6 * (size_t)e + a] = v[a] - 1;
  }
  // boundary triangles -> (element, side) by vertex set
  static const int fn[4][3] = {{0, 2, 1}, {0, 1, 3}, {0, 3, 2}, {1, 2, 3}};
  std::map<std::array<int, 3>, std::pair<int, int>> face_of;
  for (int e = 0; e < m.nel; e++)
    for (int f = 0; f < 4; f++) {
      std::array<int, 3> k = {m.nodes[6 * (size_t)e + fn[f][0]], m.nodes[6 * (size_t)e + fn[f][1]],
                              m.nodes[6 * (size_t)e + fn[f][2]]};
      std::sort(k.begin(), k.end());
      face_of.emplace(k, std::make_pair(e, f));
    }
  for (const auto& t : tris) {
    std::array<int, 3> k = {t[0] - 1, t[1] - 1, t[2] - 1};
    std::sort(k.begin(), k.end());
    auto it = face_of.find(k);
    if (it == face_of.end()) {
      set_error("NASTRAN: a boundary triangle is not a face of any tetrahedron");
      return -1;
    }
    m.bf_elem.push_back(it->second.first);
    m.bf_side.push_back(it->second.second);
    m.bf_bc.push_back(t[3]);
  }
  m.nbf = (int)m.bf_elem.size();
  return 0;
}

// ModFEM prism mesh dump (".dat" / ".dmp", the "prismatic mesh in standard
// format" of the problem files): mmr_read_mesh (mmd_prism/
// mms_prism_io_dump.c:26-270).  Sections in order: array sizes; nodes
// (nmno, first free, then x y z per node); edges (type n0 n1); faces (type
// bc neig0 neig1, then 3 or 4 signed edge IDs); elements (type mate fath
// refi, then 5 signed face IDs for a prism).  Only meshes of active
// entities are read (no refinement history, no free slots).  Element-local
// node order follows mmr_el_node_coor (mms_prism_datstr.c:365-510): the
// bottom face's nodes from its edges (mmr_fa_node_coor, :1488-1530, an edge
// entered with a negative sign contributes its second node), taken as
// (0, 2, 1) for a face of the element's orientation, else rotated by the
// shift stored as -bc of an interior face; the top face likewise as
// (0, 1, 2) or rotated and reflected.  Boundary faces are the faces with
// bc > 0 in face-ID order, owned by neighbour 0 on the side where the
// element lists the face (mmr_fa_eq_neig, :1140-1212).
int mesh_read_dat(const char* fname, Mesh& m) {
  FILE* fp = fopen(fname, "r");
  if (!fp) {
    set_error(std::string("cannot open mesh dump ") + fname);
    return -1;
  }
  auto fail = [&](const char* what) {
    fclose(fp);
    set_error(std::string("mesh dump ") + fname + ": " + what);
    return -1;
  };
  int mx[4], nmno, pf;
  if (fscanf(fp, "%d %d %d %d", &mx[0], &mx[1], &mx[2], &mx[3]) != 4) return fail("header");
  if (fscanf(fp, "%d %d", &nmno, &pf) != 2 || nmno < 1) return fail("node count");
  m = Mesh();
  m.nno = nmno;
  m.coor.assign(3 * (size_t)nmno, 0.0);
  for (int i = 0; i < nmno; i++) {
    double* c = &m.coor[3 * (size_t)i];
    if (fscanf(fp, "%lg", &c[0]) != 1) return fail("node record");
    if (c[0] < -1e10) return fail("free node slots are not read");
    if (fscanf(fp, "%lg %lg", &c[1], &c[2]) != 2) return fail("node record");
  }
  int nmed;
  if (fscanf(fp, "%d %d", &nmed, &pf) != 2 || nmed < 1) return fail("edge count");
  std::vector<int> ed(2 * (size_t)(nmed + 1), 0);
  for (int i = 1; i <= nmed; i++) {
    int t;
    if (fscanf(fp, "%d %d %d", &t, &ed[2 * i], &ed[2 * i + 1]) != 3) return fail("edge record");
    if (t <= 0) return fail("inactive or free edges are not read");
    if (ed[2 * i] < 1 || ed[2 * i] > nmno || ed[2 * i + 1] < 1 || ed[2 * i + 1] > nmno)
      return fail("edge node out of range");
  }
  int nmfa;
  if (fscanf(fp, "%d %d", &nmfa, &pf) != 2 || nmfa < 1) return fail("face count");
  std::vector<int> ftype(nmfa + 1), fbc(nmfa + 1), fneig(2 * (size_t)(nmfa + 1));
  std::vector<int> fedge(4 * (size_t)(nmfa + 1), 0);
  for (int i = 1; i <= nmfa; i++) {
    if (fscanf(fp, "%d %d %d %d", &ftype[i], &fbc[i], &fneig[2 * i], &fneig[2 * i + 1]) != 4)
      return fail("face record");
    const int ne = ftype[i] == 3 ? 3 : ftype[i] == 4 ? 4 : 0;
    if (ne == 0) return fail("inactive, free or unknown faces are not read");
    for (int k = 0; k < ne; k++) {
      if (fscanf(fp, "%d", &fedge[4 * i + k]) != 1) return fail("face edges");
      const int a = std::abs(fedge[4 * i + k]);
      if (a < 1 || a > nmed) return fail("face edge out of range");
    }
  }
  int nmel;
  if (fscanf(fp, "%d %d", &nmel, &pf) != 2 || nmel < 1) return fail("element count");
  std::vector<int> efa(5 * (size_t)nmel);
  for (int e = 0; e < nmel; e++) {
    int t, mate, fath, refi;
    if (fscanf(fp, "%d %d %d %d", &t, &mate, &fath, &refi) != 4) return fail("element record");
    if (t != kPrism) return fail("only active prisms are read");
    for (int k = 0; k < 5; k++) {
      if (fscanf(fp, "%d", &efa[5 * (size_t)e + k]) != 1) return fail("element faces");
      const int a = std::abs(efa[5 * (size_t)e + k]);
      if (a < 1 || a > nmfa) return fail("element face out of range");
    }
  }
  fclose(fp);
  // face node k: first node of edge k, or its second node when the edge is
  // entered with a negative sign
  auto fa_node = [&](int f, int k) {
    const int ied = fedge[4 * f + k];
    return ied > 0 ? ed[2 * ied] : ed[2 * (-ied) + 1];
  };
  m.nel = nmel;
  m.type.assign(nmel, kPrism);
  m.nodes.assign(6 * (size_t)nmel, -1);
  for (int e = 0; e < nmel; e++) {
    int* nd = &m.nodes[6 * (size_t)e];
    for (int b = 0; b < 2; b++) {
      const int sf = efa[5 * (size_t)e + b], f = std::abs(sf);
      if (ftype[f] != 3) {
        set_error("mesh dump: a prism base is not a triangle");
        return -1;
      }
      int v[3];
      if (sf > 0) {
        v[0] = fa_node(f, 0);
        v[1] = b == 0 ? fa_node(f, 2) : fa_node(f, 1);
        v[2] = b == 0 ? fa_node(f, 1) : fa_node(f, 2);
      } else {
        const int sh = -fbc[f];
        if (sh < 0 || sh > 2) {
          set_error("mesh dump: a base entered against its orientation has no node shift");
          return -1;
        }
        v[0] = fa_node(f, sh);
        v[b == 0 ? 1 : 2] = fa_node(f, (sh + 1) % 3);
        v[b == 0 ? 2 : 1] = fa_node(f, (sh + 2) % 3);
      }
      for (int k = 0; k < 3; k++) nd[3 * b + k] = v[k] - 1;
    }
  }
  for (int f = 1; f <= nmfa; f++) {
    if (fbc[f] <= 0) continue;
    const int owner = fneig[2 * f];
    if (owner < 1 || owner > nmel) {
      set_error("mesh dump: boundary face without an owning element");
      return -1;
    }
    int side = -1;
    for (int k = 0; k < 5 && side < 0; k++)
      if (std::abs(efa[5 * (size_t)(owner - 1) + k]) == f) side = k;
    if (side < 0) {
      set_error("mesh dump: boundary face not listed by its owner");
      return -1;
    }
    m.bf_elem.push_back(owner - 1);
    m.bf_side.push_back(side);
    m.bf_bc.push_back(fbc[f]);
  }
  m.nbf = (int)m.bf_elem.size();
  return 0;
}

int mesh_box_tetra(int nx, int ny, int nz, double lx, double ly, double lz,
                   uint64_t seed, double jitter, int bc, Mesh& m) {
  if (nx < 1 || ny < 1 || nz < 1) {
    set_error("box mesh: nx, ny, nz must be >= 1");
    return -1;
  }
  std::mt19937_64 rng(seed);
  const int NN = (nx + 1) * (ny + 1) * (nz + 1), NE = 6 * nx * ny * nz;
  std::vector<int> pn = seeded_perm(NN, rng);
  std::vector<int> pe = seeded_perm(NE, rng);
  m = Mesh();
  m.nno = NN;
  m.coor.assign(3 * (size_t)NN, 0.0);
  auto gid = [&](int i, int j, int k) { return i + (nx + 1) * (j + (ny + 1) * k); };
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        size_t v = pn[gid(i, j, k)];
        m.coor[3 * v] = lx * i / nx;
        m.coor[3 * v + 1] = ly * j / ny;
        m.coor[3 * v + 2] = lz * k / nz;
      }
  m.nel = NE;
  m.type.assign(NE, kTetra);
  m.nodes.assign(6 * (size_t)NE, -1);
  // Kuhn paths from corner 0 to corner 7 through the 6 axis permutations
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                  {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        for (int p = 0; p < 6; p++) {
          int c[3] = {i, j, k};
          int vv[4];
          vv[0] = gid(c[0], c[1], c[2]);
          for (int s = 0; s < 3; s++) {
            c[perms[p][s]]++;
            vv[s + 1] = gid(c[0], c[1], c[2]);
          }
          // orient positively
          double x[4][3];
          for (int a = 0; a < 4; a++) {
            int cc[3] = {vv[a] % (nx + 1), (vv[a] / (nx + 1)) % (ny + 1),
                         vv[a] / ((nx + 1) * (ny + 1))};
            x[a][0] = cc[0]; x[a][1] = cc[1]; x[a][2] = cc[2];
          }
          double a1[3], a2[3], a3[3];
          for (int d = 0; d < 3; d++) {
            a1[d] = x[1][d] - x[0][d]; a2[d] = x[2][d] - x[0][d]; a3[d] = x[3][d] - x[0][d];
          }
          double det = a1[0] * (a2[1] * a3[2] - a2[2] * a3[1]) -
                       a1[1] * (a2[0] * a3[2] - a2[2] * a3[0]) +
                       a1[2] * (a2[0] * a3[1] - a2[1] * a3[0]);
          if (det < 0) std::swap(vv[2], vv[3]);
          size_t e = pe[((k * ny + j) * nx + i) * 6 + p];
          for (int a = 0; a < 4; a++) m.nodes[6 * e + a] = pn[vv[a]];
        }
  // boundary faces
  static const int fn[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  const double L3[3] = {lx, ly, lz};
  for (int e = 0; e < NE; e++)
    for (int f = 0; f < 4; f++) {
      for (int d = 0; d < 3; d++) {
        bool lo = true, hi = true;
        for (int a = 0; a < 3; a++) {
          double xv = m.coor[3 * (size_t)m.nodes[6 * (size_t)e + fn[f][a]] + d];
          if (std::fabs(xv) > 1e-12 * L3[d]) lo = false;
          if (std::fabs(xv - L3[d]) > 1e-12 * L3[d]) hi = false;
        }
        if (lo || hi) {
          m.bf_elem.push_back(e); m.bf_side.push_back(f); m.bf_bc.push_back(bc);
          break;
        }
      }
    }
  m.nbf = (int)m.bf_elem.size();
  jitter_interior(m, lx, ly, lz, nx, ny, nz, jitter, seed);
  return 0;
}


// Hybrid box: every z-layer of hex cells is split either into 2 prisms per
// cell (along the (0,0)-(1,1) xy-diagonal, as the prism box) or into 6
// tetrahedra per cell (Kuhn split, as the tetrahedral box).  Both splits cut
// the horizontal faces along that same diagonal, so prism and tet layers
// meet on matching triangles and the mesh is conforming; tets touch prisms
// only through those triangles.  A layer is tetrahedral with probability
// tet_frac (seeded draws, layer by layer).  Node and element IDs are
// permuted by the seed; boundary faces are listed element by element in
// local face order (prisms: bottom, top, quads {0,1,4,3}, {1,2,5,4},
// {2,0,3,5}; tets: {0,1,2}, {0,1,3}, {0,2,3}, {1,2,3}).
int mesh_box_hybrid(int nx, int ny, int nz, double lx, double ly, double lz, uint64_t seed,
                    double jitter, int bc, double tet_frac, Mesh& m) {
  if (nx < 1 || ny < 1 || nz < 1 || tet_frac < 0 || tet_frac > 1) {
    set_error("hybrid box mesh: nx, ny, nz >= 1 and 0 <= tet_frac <= 1 required");
    return -1;
  }
  std::mt19937_64 rng(seed);
  std::vector<int> tet_layer(nz);
  long long ne = 0;
  for (int k = 0; k < nz; k++) {
    tet_layer[k] = unit_draw(rng) < tet_frac ? 1 : 0;
    ne += (long long)nx * ny * (tet_layer[k] ? 6 : 2);
  }
  if (ne > 2147483647LL / 6) {
    set_error("hybrid box mesh: too many elements");
    return -1;
  }
  const int NN = (nx + 1) * (ny + 1) * (nz + 1), NE = (int)ne;
  std::vector<int> pn = seeded_perm(NN, rng);
  std::vector<int> pe = seeded_perm(NE, rng);
  m = Mesh();
  m.nno = NN;
  m.coor.assign(3 * (size_t)NN, 0.0);
  auto gid = [&](int i, int j, int k) { return i + (nx + 1) * (j + (ny + 1) * k); };
  for (int k = 0; k <= nz; k++)
    for (int j = 0; j <= ny; j++)
      for (int i = 0; i <= nx; i++) {
        const size_t v = pn[gid(i, j, k)];
        m.coor[3 * v] = lx * i / nx;
        m.coor[3 * v + 1] = ly * j / ny;
        m.coor[3 * v + 2] = lz * k / nz;
      }
  m.nel = NE;
  m.type.assign(NE, kPrism);
  m.nodes.assign(6 * (size_t)NE, -1);
  std::vector<int> grid(6 * (size_t)NE, -1);  // grid node ids, element order
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                  {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  int t = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        if (!tet_layer[k]) {
          const int a = gid(i, j, k), b = gid(i + 1, j, k), c = gid(i + 1, j + 1, k),
                    d = gid(i, j + 1, k), up = (nx + 1) * (ny + 1);
          const int pr[2][3] = {{a, b, c}, {a, c, d}};
          for (int q = 0; q < 2; q++, t++) {
            const size_t e = pe[t];
            for (int r = 0; r < 3; r++) {
              grid[6 * e + r] = pr[q][r];
              grid[6 * e + 3 + r] = pr[q][r] + up;
            }
          }
        } else {
          for (int q = 0; q < 6; q++, t++) {
            int c[3] = {i, j, k}, vv[4];
            vv[0] = gid(c[0], c[1], c[2]);
            for (int s = 0; s < 3; s++) {
              c[perms[q][s]]++;
              vv[s + 1] = gid(c[0], c[1], c[2]);
            }
            // the Kuhn path i -> j -> k order is positive for even permutations
            if (q == 1 || q == 2 || q == 5) std::swap(vv[2], vv[3]);
            const size_t e = pe[t];
            m.type[e] = kTetra;
            for (int r = 0; r < 4; r++) grid[6 * e + r] = vv[r];
          }
        }
      }
  for (size_t q = 0; q < grid.size(); q++)
    if (grid[q] >= 0) m.nodes[q] = pn[grid[q]];
  // boundary faces: every vertex of the face on one side of the box
  static const int pf[5][4] = {{0, 1, 2, -1}, {3, 4, 5, -1}, {0, 1, 4, 3}, {1, 2, 5, 4},
                               {2, 0, 3, 5}};
  static const int tf[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  const int hi[3] = {nx, ny, nz};
  for (int e = 0; e < NE; e++) {
    const bool tet = m.type[e] == kTetra;
    for (int s = 0; s < (tet ? 4 : 5); s++) {
      const int nv = tet || s < 2 ? 3 : 4;
      bool on = false;
      for (int d = 0; d < 3 && !on; d++) {
        bool lo = true, up = true;
        for (int r = 0; r < nv; r++) {
          const int g = grid[6 * (size_t)e + (tet ? tf[s][r] : pf[s][r])];
          const int ci[3] = {g % (nx + 1), (g / (nx + 1)) % (ny + 1), g / ((nx + 1) * (ny + 1))};
          if (ci[d] != 0) lo = false;
          if (ci[d] != hi[d]) up = false;
        }
        on = lo || up;
      }
      if (on) {
        m.bf_elem.push_back(e); m.bf_side.push_back(s); m.bf_bc.push_back(bc);
      }
    }
  }
  m.nbf = (int)m.bf_elem.size();
  jitter_interior(m, lx, ly, lz, nx, ny, nz, jitter, seed);
  return 0;
}

}  // namespace mf
