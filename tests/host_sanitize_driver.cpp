// Host-only exercise of the parameter and scan readers, the box mesher, the
// mesh reader (also on truncated files), assembly and BFS, built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_host.py (no
// device code is linked).  argv[1]: a scratch directory; argv[2], optional:
// a directory of parameter files to read.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dirent.h>
#include <string>
#include <vector>
#include "srs_abi.h"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string tmp = argv[1];
  int32_t ints[SRS_N_INT];
  double reals[SRS_N_REAL];
  char title[256];
  int n_ok = 0, n_bad = 0;
  // every parameter file of the given directory and its walker2014/
  std::vector<std::string> dirs;
  if (argc > 2) dirs = {argv[2], std::string(argv[2]) + "/walker2014"};
  for (const std::string& d : dirs) {
    DIR* dp = opendir(d.c_str());
    if (!dp) continue;
    while (dirent* e = readdir(dp)) {
      std::string f = d + "/" + e->d_name;
      if (f.size() < 4 || f.substr(f.size() - 4) != ".xml") continue;
      if (srs_read_parameters(f.c_str(), ints, reals, title) == SRS_OK) ++n_ok; else ++n_bad;
    }
    closedir(dp);
  }
  printf("parameter files: %d ok, %d errors\n", n_ok, n_bad);
  // malformed XML snippets
  const char* bad[] = {"", "<", "<simulation>", "<simulation><parameter><symbol>t_final</symbol></parameter>",
                       "<simulation><parameter><symbol>t_final</symbol><value>1e</value></parameter></simulation>",
                       "<simulation><parameter><symbol>", "<?xml version='1.0'?><simulation></simulation>",
                       "<simulation><parameter><symbol>t_final</symbol><value>5</value></parameter><parameter>"};
  for (const char* b : bad) {
    const std::string f = tmp + "/bad.xml";
    FILE* fp = fopen(f.c_str(), "w");
    fputs(b, fp);
    fclose(fp);
    int rc = srs_read_parameters(f.c_str(), ints, reals, title);
    printf("bad xml rc %d\n", rc);
  }
  // scan parameters
  srs_scan_desc sd;
  printf("scan defaults rc %d\n", srs_scan_read_parameters(nullptr, &sd));
  // box mesh -> files -> reader -> assembly -> BFS
  srs_box_spec spec;
  memset(&spec, 0, sizeof spec);
  std::vector<double> xs = {0, 0.04, 0.08, 0.12, 0.16}, ys = xs, zs = {0, 0.015, 0.055, 0.095};
  spec.nx = (int)xs.size() - 1; spec.xs = xs.data();
  spec.ny = (int)ys.size() - 1; spec.ys = ys.data();
  spec.nz = (int)zs.size() - 1; spec.zs = zs.data();
  spec.jsr_lo[0] = 0.04; spec.jsr_lo[1] = 0.04; spec.jsr_lo[2] = 0.015;
  spec.jsr_hi[0] = 0.12; spec.jsr_hi[1] = 0.12; spec.jsr_hi[2] = 0.055;
  spec.sl_face_z0 = 1;
  spec.trpn_radius = -1.0;
  spec.ryr_rows = 2; spec.ryr_cols = 2; spec.ryr_spacing = 0.031;
  spec.ryr_center[0] = 0.08; spec.ryr_center[1] = 0.08; spec.ryr_z_cyt = 0.0075; spec.ryr_z_jsr = 0.035;
  double lcc[3] = {0.09, 0.05, 0.0075};
  spec.n_lcc = 1; spec.lcc_xyz = lcc;
  const std::string base = tmp + "/mesh";
  int rc = srs_box_mesh_write(&spec, base.c_str());
  printf("mesh write rc %d\n", rc);
  srs_grid* g = nullptr;
  int32_t warn = 0;
  rc = srs_grid_read(base.c_str(), 1, &g, &warn);
  printf("grid read rc %d\n", rc);
  if (rc == SRS_OK) {
    std::vector<double> aij(5 * (size_t)g->n_ele);
    std::vector<int32_t> nij(4 * (size_t)g->n_ele);
    rc = srs_assemble_operator(g->n_ele, g->n_face, g->ele1, g->ele2, g->face_type, g->face_c, g->V0,
                               aij.data(), nij.data());
    printf("assemble rc %d\n", rc);
    srs_grid_free(g);
  }
  // truncated files
  for (const char* suf : {".node", ".ele", "_tet_properties.txt", "_face_properties.txt", "_ryr.txt", "_lcc.txt"}) {
    std::string f = base + suf;
    FILE* fp = fopen(f.c_str(), "rb");
    if (!fp) continue;
    fseek(fp, 0, SEEK_END);
    long n = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    std::vector<char> buf(n);
    size_t got = fread(buf.data(), 1, n, fp);
    fclose(fp);
    for (long cut : {0L, 1L, n / 3, n / 2, n - 2}) {
      if (cut < 0 || cut > (long)got) continue;
      fp = fopen(f.c_str(), "wb");
      fwrite(buf.data(), 1, cut, fp);
      fclose(fp);
      srs_grid* g2 = nullptr;
      int rc2 = srs_grid_read(base.c_str(), 1, &g2, &warn);
      if (rc2 == SRS_OK) srs_grid_free(g2);
    }
    fp = fopen(f.c_str(), "wb");
    fwrite(buf.data(), 1, got, fp);
    fclose(fp);
  }
  printf("truncation sweep done\n");
  return (rc == SRS_OK) ? 0 : 1;
}
