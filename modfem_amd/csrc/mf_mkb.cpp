// Solver control file (mkb.dat) reader, following the single-level branch of
// lsr_mkb_core_init (lsd_mkb/lsd_mkb_core/lss_mkb_core.c:148-705):
//
//   FINE_LEVEL                        keyword, rest of the line ignored
//   solver nr_level krylbas           (:424)
//   precon nr_prec block_type         (:469-481)  block_type < 0 selects the
//                                      storage |block_type|; precon 4x/5x is
//                                      ILU(k) with k = precon % 10
//   max_iter conv_type conv_meas      (:474-476)
//   monitor nr_pre_smooth nr_post     (:477-480)
//
// Each line is read with the same stdio conversions as the reference and the
// rest of the line is skipped (lsr_mkb_core_util_skip_rest_of_line, :3005),
// so comment text after the numbers is ignored the same way.  Arguments
// > 0 override the file (:483-487); an empty file name gives the defaults of
// :237-266.  Multigrid levels, AMG and the NS-SUPG branch are not part of this
// build and are reported as errors instead of exiting.
#include <cstdio>
#include <cstring>
#include <string>

#include "mf_internal.h"

namespace mf {

namespace {
// constants of lsh_mkb_intf.h:116-122, lah_intf.h:45-50,68-73, lsh_mkb_core.h:111
constexpr int GMRES = 1, MULTI_GMRES = 2, STANDARD_IT = 10, MULTI_GRID = 20;
constexpr int BLOCK_GS = 2, MULTI_ILU = 4, BLOCK_ILU = 5;
constexpr int STORAGE_UNDEFINED = -1, STORAGE_BLOCK = 1;
constexpr int REL_RES_INI = 0;

void skip_line(FILE* fp) {
  int c;
  while ((c = fgetc(fp)) != '\n' && c != EOF) {
  }
}
}  // namespace

int mkb_read(const char* filename, int max_iter, int error_type, double error_tol,
             int monitoring, MkbParams& p) {
  p = MkbParams();
  if (!filename || !filename[0]) {
    p.solver = GMRES;
    p.nr_level = 1;
    p.storage_type = STORAGE_UNDEFINED;
    p.max_iter = max_iter < 0 ? 100 : max_iter;
    p.conv_type = error_type < 0 ? REL_RES_INI : error_type;
    p.conv_meas = error_tol < 0.0 ? 1e-6 : error_tol;
    p.monitor = monitoring < 0 ? 0 : monitoring;
    p.krylbas = 50;
    p.precon = BLOCK_GS;
    p.nr_prec = 1;
    p.ilu_k = 0;
    p.block_type = 1;
    p.nr_pre_smooth = 1;
    p.nr_post_smooth = 0;
    return 0;
  }
  FILE* fp = std::fopen(filename, "r");
  if (!fp) {
    set_error(std::string("mkb: cannot open solver control file ") + filename);
    return -1;
  }
  char keyword[64] = {0};
  int ok = std::fscanf(fp, "%63s", keyword) == 1;
  skip_line(fp);
  if (!ok || std::strcmp(keyword, "FINE_LEVEL") != 0) {
    std::fclose(fp);
    set_error(std::string("mkb: only FINE_LEVEL single-level files are supported, found '") +
              keyword + "'");
    return -1;
  }
  int solver = 0, nr_level = 0, krylbas = 0;
  ok = std::fscanf(fp, "%d %d %d", &solver, &nr_level, &krylbas) == 3;
  skip_line(fp);
  if (nr_level > 1) nr_level = 1;  // one level on the device (:427-440 clamp)
  if (nr_level == 1) {
    if (solver == MULTI_GMRES) solver = GMRES;
    if (solver == MULTI_GRID) solver = STANDARD_IT;
  }
  if (!ok || (solver != GMRES && solver != STANDARD_IT)) {
    std::fclose(fp);
    set_error("mkb: solver line must be 'solver nr_level krylbas' with a GMRES or standard "
              "iterations solver");
    return -1;
  }
  p.solver = solver;
  p.nr_level = nr_level;
  p.krylbas = krylbas;
  int precon = 0, nr_prec = 0, block_type = 0;
  ok = std::fscanf(fp, "%d %d %d", &precon, &nr_prec, &block_type) == 3;
  if (block_type < 0) {
    p.storage_type = -block_type;
    block_type = -1;
  } else {
    p.storage_type = STORAGE_BLOCK;
  }
  if ((precon >= 10 * MULTI_ILU && precon < 10 * MULTI_ILU + 9) ||
      (precon >= 10 * BLOCK_ILU && precon < 10 * BLOCK_ILU + 9)) {
    p.ilu_k = precon % 10;
    precon /= 10;
  } else {
    p.ilu_k = 0;
  }
  p.precon = precon;
  p.nr_prec = nr_prec;
  p.block_type = block_type;
  skip_line(fp);
  int mi = 0, ct = 0;
  double cm = 0;
  ok = ok && std::fscanf(fp, "%d %d %lg", &mi, &ct, &cm) == 3;
  skip_line(fp);
  p.max_iter = mi;
  p.conv_type = ct;
  p.conv_meas = cm;
  int mon = 0, pre = 0, post = 0;
  int got = std::fscanf(fp, "%d %d %d", &mon, &pre, &post);
  std::fclose(fp);
  if (!ok || got < 1) {
    set_error("mkb: malformed preconditioner, tolerance or output line");
    return -1;
  }
  p.monitor = mon;
  p.nr_pre_smooth = got >= 2 ? pre : 0;
  p.nr_post_smooth = got >= 3 ? post : 0;
  if (max_iter > 0) p.max_iter = max_iter;
  if (error_type > 0) p.conv_type = error_type;
  if (error_tol > 0) p.conv_meas = error_tol;
  if (monitoring > 0) p.monitor = monitoring;
  return 0;
}

}  // namespace mf
