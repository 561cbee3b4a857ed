/*
 * A ModFEM-style driver in plain C that reaches the device only through the
 * reference's own seams (include/modfem_lsr.h, modfem_lar.h, modfem_tmr.h).
 * The problem-module role (element matrices, coefficients, mesh) is played by
 * libmodfem_hip.so's mf_* calls and by the formulas below.
 *
 *   seam_driver sir <jk mesh> <mkb.dat> <out>
 *     the call sequence of sir_create + sir_solve (sid_mkb/sis_mkb_intf.c:
 *     396-1692, 1698-2178) with the solver assembling (SIC_SOLVER_ASSEMBLE):
 *     lsr_mkb_init, lsr_mkb_create_matrix (neighbour-block lists of the DOF
 *     structure), lsr_mkb_create_precon, lsr_mkb_fill_assembly_table_int_ent
 *     per integration entity; lsr_mkb_clear_matrix, per entity the local
 *     matrices (pdr_comp_stiff_mat's role) and
 *     lsr_mkb_assemble_local_stiff_mat_with_table, lsr_mkb_fill_precon,
 *     lsr_mkb_solve; lsr_mkb_free_matrix, lsr_mkb_destroy.  Problem:
 *     DIFF_IN_BOX, penalty faces.  out: n, iterations, convergence measure,
 *     then the solution in DOF order.
 *
 *   seam_driver accel <nx> <ny> <nz> <packing 1|2> <out>
 *     the GPU_ASSEMBLING create-assemble sequence (utd_util/uts_ls_intf.c:
 *     3480-3560, uts_accel_intf.c:127-470): the CRS arrays registered
 *     (tmr_get_crs_structure), tmr_prepare_create_assemble_accel,
 *     tmr_perform_creation_crs, the elements streamed in three colours
 *     packed as the reference packs them (gauss, shape functions, vertex
 *     coordinates; PDE data 1: one source value per integration point,
 *     DIFF_IN_BOX; 2: 20 coefficients per element), tmr_prepare_final_crs,
 *     tmr_cleanup_crs.  Box of prisms without boundary faces.  out: n, nnz,
 *     values, rhs.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "modfem_hip.h"
#include "modfem_lar.h"
#include "modfem_lsr.h"
#include "modfem_tmr.h"

#define CHECK(x)                                                              \
  do {                                                                        \
    int rc_ = (x);                                                            \
    if (rc_ < 0) {                                                            \
      fprintf(stderr, "%s:%d %s -> %d: %s\n", __FILE__, __LINE__, #x, rc_, \
              mf_last_error());                                               \
      exit(2);                                                                \
    }                                                                         \
  } while (0)

/* coefficient slots of mf_problem_set_coefficients (23-slot indicator of
   pdr_select_el_coeff_vect) */
static void diffusion_indicator(int* ind) {
  memset(ind, 0, 23 * sizeof(int));
  ind[0] = 1;
  ind[2] = ind[6] = ind[10] = 1; /* axx ayy azz */
  ind[22] = 1;                   /* s */
}

static int run_sir(const char* mesh_file, const char* mkb_file, const char* out_file) {
  CHECK(mf_init(0));
  int mesh = mf_mesh_read_jk(mesh_file);
  CHECK(mesh);
  int pid = mf_problem_create(mesh);
  CHECK(pid);
  int ind[23];
  double cc[21] = {0};
  diffusion_indicator(ind);
  cc[0] = cc[4] = cc[8] = 1.0;
  CHECK(mf_problem_set_coefficients(pid, ind, 0, cc, NULL, 2 /* DIFF_IN_BOX */, NULL));
  int n = 0, nel = 0, nbf = 0, maxrow = 0;
  long long nnz = 0;
  CHECK(mf_problem_sizes(pid, &n, &nnz, &nel, &nbf, &maxrow));
  int* rp = malloc((n + 1) * sizeof(int));
  int* ci = malloc(nnz * sizeof(int));
  CHECK(mf_problem_get_pattern(pid, rp, ci));
  /* DOF structure: neighbour blocks without the block itself, ascending */
  int* nro = malloc(n * sizeof(int));
  int** lof = malloc(n * sizeof(int*));
  int* ones = malloc(n * sizeof(int));
  int* pos = malloc(n * sizeof(int));
  for (int r = 0; r < n; r++) {
    ones[r] = 1;
    pos[r] = r;
    lof[r] = malloc((rp[r + 1] - rp[r]) * sizeof(int));
    nro[r] = 0;
    for (int p = rp[r]; p < rp[r + 1]; p++)
      if (ci[p] != r) lof[r][nro[r]++] = ci[p];
  }
  int levels = 1, storage = -1;
  int solver = lsr_mkb_init(1, 0, &levels, &storage, (char*)mkb_file, -1, -1, -1.0, 0);
  CHECK(solver);
  CHECK(lsr_mkb_create_matrix(solver, 0, -1, n, n, 1, 36, ones, pos, nro, lof));
  CHECK(lsr_mkb_create_precon(solver, 0));
  /* integration entities: elements, then boundary faces; tables as sir_create
     fills them (sis_mkb_intf.c:1562-1688) */
  const int nint = nel + nbf;
  int* el_type = malloc(nel * sizeof(int));
  int* el_nodes = malloc(6 * nel * sizeof(int));
  int* bf_elem = malloc((nbf + 1) * sizeof(int));
  int nno = 0, mel = 0, mbf = 0;
  CHECK(mf_mesh_sizes(mesh, &nno, &mel, &mbf));
  double* coor = malloc(3 * nno * sizeof(double));
  int* bf_side = malloc((mbf + 1) * sizeof(int));
  int* bf_bc = malloc((mbf + 1) * sizeof(int));
  CHECK(mf_mesh_get(mesh, coor, el_type, el_nodes, bf_elem, bf_side, bf_bc));
  int* dof_node = malloc(n * sizeof(int));
  int* node_dof = malloc(nno * sizeof(int));
  CHECK(mf_problem_get_dofs(pid, dof_node));
  for (int i = 0; i < n; i++) node_dof[dof_node[i]] = i;
  /* assembly tables only for non-block storage and at least 1000 DOFs
     (sis_mkb_intf.c:1560-1570); otherwise the solver assembles from the
     entities' DOF lists */
  const int use_tables = storage != 1 && n >= 1000;
  int* tables = malloc((size_t)nint * 42 * sizeof(int));
  for (int ie = 0; use_tables && ie < nint; ie++) {
    const int e = ie < nel ? ie : bf_elem[ie - nel];
    const int k = el_type[e] == MF_TETRA ? 4 : 6;
    int lbl[6], lnr[6];
    for (int i = 0; i < k; i++) {
      lbl[i] = node_dof[el_nodes[6 * e + i]];
      lnr[i] = 1;
    }
    CHECK(lsr_mkb_fill_assembly_table_int_ent(solver, 0, k, lbl, lnr, tables + (size_t)ie * 42));
  }
  /* sir_solve: clear, local matrices of every entity, assemble with tables */
  double* em = malloc((size_t)nel * 42 * sizeof(double));
  double* fm = malloc((size_t)(nbf + 1) * 42 * sizeof(double));
  CHECK(mf_element_matrices(pid, 0, nel, 64, em));
  if (nbf) CHECK(mf_face_matrices(pid, 0, nbf, fm));
  CHECK(lsr_mkb_clear_matrix(solver, 0, 0));
  char rewr = 'F';
  for (int ie = 0; ie < nint; ie++) {
    const int e = ie < nel ? ie : bf_elem[ie - nel];
    const int k = el_type[e] == MF_TETRA ? 4 : 6;
    double* m = ie < nel ? em + (size_t)ie * 42 : fm + (size_t)(ie - nel) * 42;
    if (use_tables) {
      CHECK(lsr_mkb_assemble_local_stiff_mat_with_table(solver, 0, 0, k, tables + (size_t)ie * 42,
                                                        m, m + k * k, &rewr));
    } else {
      int lbl[6], lnr[6];
      for (int i = 0; i < k; i++) {
        lbl[i] = node_dof[el_nodes[6 * e + i]];
        lnr[i] = 1;
      }
      CHECK(lsr_mkb_assemble_local_sm(solver, 0, 0, k, lbl, lnr, m, m + k * k, &rewr));
    }
  }
  CHECK(lsr_mkb_fill_precon(solver, 0));
  double* x = calloc(n, sizeof(double));
  int nr_iter = 2000;
  double toler = 1e-12, rate = 0.0;
  int conv = lsr_mkb_solve(solver, 0, 1, x, NULL, &nr_iter, &toler, 0, &rate);
  CHECK(conv);
  printf("sir sequence: n %d, %d int-ents, storage %d, tables %d, converged %d in %d "
         "iterations, measure %.3e, rate %.4f\n",
         n, nint, storage, use_tables, conv, nr_iter, toler, rate);
  FILE* f = fopen(out_file, "wb");
  double hdr[3] = {(double)n, (double)nr_iter, toler};
  fwrite(hdr, sizeof(double), 3, f);
  fwrite(x, sizeof(double), n, f);
  fclose(f);
  CHECK(lsr_mkb_free_matrix(solver));
  CHECK(lsr_mkb_destroy(solver));
  CHECK(mf_problem_free(pid));
  CHECK(mf_mesh_free(mesh));
  return conv == 1 ? 0 : 3;
}

/* the reference prism's 6-point rule and P1 shape functions
   (aps_gauss_util_std.c, aps_std_util.c:463-511) */
static void prism_shape(const double* xi, double* s) {
  const double b[3] = {1.0 - xi[0] - xi[1], xi[0], xi[1]};
  const double db[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
  for (int lev = 0; lev < 2; lev++) {
    const double h = lev == 0 ? 0.5 * (1.0 - xi[2]) : 0.5 * (1.0 + xi[2]);
    const double dh = lev == 0 ? -0.5 : 0.5;
    for (int i = 0; i < 3; i++) {
      double* o = s + 4 * (3 * lev + i);
      o[0] = b[i] * h;
      o[1] = db[i][0] * h;
      o[2] = db[i][1] * h;
      o[3] = b[i] * dh;
    }
  }
}

static double lapl_diff_in_box(const double* x) {
  const double a1 = 100.0, X = x[0], Y = x[1], Z = x[2];
  const double e = exp((-X * X - Y * Y - Z * Z) / a1);
  return -(2.0 / a1) * ((1 - 2 * X * X / a1) + (1 - 2 * Y * Y / a1) + (1 - 2 * Z * Z / a1)) * e;
}

static int run_accel(int nx, int ny, int nz, int packing, const char* out_file) {
  CHECK(mf_init(0));
  int mesh = mf_mesh_box_prism(nx, ny, nz, 1.0, 1.3, 0.7, 11ULL, 0.2, 0);
  CHECK(mesh);
  int pid = mf_problem_create(mesh);
  CHECK(pid);
  int n = 0, nel = 0, nbf = 0, maxrow = 0;
  long long nnz = 0;
  CHECK(mf_problem_sizes(pid, &n, &nnz, &nel, &nbf, &maxrow));
  if (nbf != 0) {
    fprintf(stderr, "accel: the box must have no boundary faces\n");
    return 2;
  }
  /* the CRS module's arrays (lar_allocate_SM_and_LV_crs registers them) */
  int* rp = malloc((n + 1) * sizeof(int));
  int* ci = malloc(nnz * sizeof(int));
  double* val = calloc(nnz, sizeof(double));
  double* rhs = calloc(n, sizeof(double));
  CHECK(mf_problem_get_pattern(pid, rp, ci));
  CHECK(tmr_get_crs_structure(n, (int)nnz, ci, rp, val, rhs));
  int nno = 0, mel = 0, mbf = 0;
  CHECK(mf_mesh_sizes(mesh, &nno, &mel, &mbf));
  double* coor = malloc(3 * nno * sizeof(double));
  int* el_type = malloc(nel * sizeof(int));
  int* el_nodes = malloc(6 * nel * sizeof(int));
  CHECK(mf_mesh_get(mesh, coor, el_type, el_nodes, NULL, NULL, NULL));
  int* pos = malloc(nel * sizeof(int));
  int* table = malloc((size_t)nel * 42 * sizeof(int));
  for (int e = 0; e < nel; e++) {
    pos[e] = 42 * e;
    CHECK(mf_problem_get_table(pid, e, table + 42 * e));
  }
  CHECK(tmr_prepare_create_assemble_accel(1, 0));
  CHECK(tmr_perform_creation_crs(1, nel, pos, table, 6));
  /* quadrature and shape functions of the reference element */
  const double G2 = 0.577350269189626, T1 = 0.166666666666667, T2 = 0.666666666666667;
  const double TW = 0.333333333333333;
  double gauss[6 * 4], shape[6 * 6 * 4];
  for (int g = 0; g < 6; g++) {
    const double xi[3] = {g % 3 == 2 ? T2 : T1, g % 3 == 1 ? T2 : T1, g < 3 ? -G2 : G2};
    gauss[4 * g] = xi[0];
    gauss[4 * g + 1] = xi[1];
    gauss[4 * g + 2] = xi[2];
    gauss[4 * g + 3] = 0.5 * TW * 1.0;
    prism_shape(xi, shape + 24 * g);
  }
  double* geo = malloc((size_t)nel * 18 * sizeof(double));
  for (int e = 0; e < nel; e++)
    for (int i = 0; i < 6; i++)
      for (int d = 0; d < 3; d++) geo[18 * e + 3 * i + d] = coor[3 * el_nodes[6 * e + i] + d];
  const int per_el = packing == 2 ? 20 : 0, per_pt = packing == 1 ? 1 : 0;
  double* pde = malloc((size_t)nel * (packing == 2 ? 20 : 6) * sizeof(double));
  for (int e = 0; e < nel; e++) {
    if (packing == 1) {
      for (int g = 0; g < 6; g++) {
        double xg[3] = {0, 0, 0};
        for (int i = 0; i < 6; i++)
          for (int d = 0; d < 3; d++) xg[d] += geo[18 * e + 3 * i + d] * shape[24 * g + 4 * i];
        pde[6 * e + g] = -lapl_diff_in_box(xg);
      }
    } else {
      double* p = pde + 20 * e;
      const double kap = 1.0 + 0.5 * sin(0.37 * e);
      memset(p, 0, 20 * sizeof(double));
      p[0] = 2.0 * kap; p[1] = 0.1; p[2] = -0.05; p[3] = 0.2; p[4] = 1.5 * kap; p[5] = 0.03;
      p[6] = -0.1; p[7] = 0.04; p[8] = kap;                        /* a */
      p[9] = 0.3 * cos(e); p[10] = -0.7; p[11] = 0.2;               /* b */
      p[12] = 0.1; p[13] = 0.05 * sin(e); p[14] = -0.2;            /* t */
      p[15] = 0.8;                                                  /* c */
      p[16] = -0.3; p[17] = 0.6; p[18] = 0.1 * cos(2.0 * e);       /* q */
      p[19] = 1.0 + 0.1 * e / nel;                                  /* s */
    }
  }
  /* three colours of consecutive elements */
  const int bounds[4] = {0, nel / 3, (2 * nel) / 3, nel};
  for (int c = 0; c < 3; c++) {
    const int e0 = bounds[c], ne = bounds[c + 1] - bounds[c];
    int* ids = malloc((ne + 1) * sizeof(int));
    for (int i = 0; i < ne; i++) ids[i] = e0 + i + 1;
    CHECK(tmr_perform_create_assemble_accel(
        1, 0, 3, ne, e0, ids, pos, table, 6, gauss, 6, shape, 6, geo + 18 * (size_t)e0, 6, 0,
        per_el, per_pt, pde + (size_t)e0 * (packing == 2 ? 20 : 6), 6));
    free(ids);
  }
  CHECK(tmr_prepare_final_crs());
  CHECK(tmr_cleanup_crs());
  printf("accel sequence: n %d, nnz %lld, %d elements in 3 colours, packing %d\n", n, nnz, nel,
         packing);
  FILE* f = fopen(out_file, "wb");
  double hdr[2] = {(double)n, (double)nnz};
  fwrite(hdr, sizeof(double), 2, f);
  fwrite(val, sizeof(double), nnz, f);
  fwrite(rhs, sizeof(double), n, f);
  fwrite(pde, sizeof(double), (size_t)nel * (packing == 2 ? 20 : 6), f);
  fclose(f);
  CHECK(mf_problem_free(pid));
  CHECK(mf_mesh_free(mesh));
  return 0;
}

int main(int argc, char** argv) {
  if (argc == 5 && !strcmp(argv[1], "sir")) return run_sir(argv[2], argv[3], argv[4]);
  if (argc == 7 && !strcmp(argv[1], "accel"))
    return run_accel(atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), argv[6]);
  fprintf(stderr, "usage: seam_driver sir <jk mesh> <mkb.dat> <out> | "
                  "accel <nx> <ny> <nz> <1|2> <out>\n");
  return 1;
}
