/*
 * bridge_check.cpp -- runs the per-step bridge of the integration notes in
 * the reference's own step loop and compares it with the reference's own
 * ExplicitStep.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/build.py compiles it together with the
 * bridge code extracted from INTEGRATION.md (oracle/_ref/srs_bridge.inc), the
 * reference's own header cpu/parameters.h and its unmodified
 * cpu/RunSimulation.cpp (read where they lie), linked against
 * libsrs_amd.so, into oracle/_ref/bridge_check.
 *
 *   bridge_check PARAM.xml MESH_BASE N_STEPS SEED
 *
 * Two SimData of the same mesh and parameters (flag_nogating set): one is
 * stepped by SrsExplicitStep (the bridge: srs_explicit_step on the GPU),
 * the other by the reference's ExplicitStep (cpu/RunSimulation.cpp:591),
 * both with the RyR states of every step overwritten from the same random
 * schedule before the call, as the reference's loop lets a caller do
 * between steps (cpu/RunSimulation.cpp:540-542, 569).  After N_STEPS the
 * bridge's fields (SrsPullStates) must equal the reference's bit for bit.
 * Exit status 0: equal; 2: different; 3: set-up failure.
 */
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "srs_bridge.inc"

static_assert(N_PARAMETERS_INT == SRS_N_INT && N_PARAMETERS_REAL == SRS_N_REAL, "parameter vector layout");

namespace {

struct Owned {
  std::vector<double> X, buf, zeros, R, aij;
  std::vector<int> nij, rs, ls, lvs;
  std::vector<double*> rows, brows;
};

void bind(SimData& sd, Owned& o, const srs_grid* g, const std::vector<double>& X0, const std::vector<double>& aij,
          const std::vector<int32_t>& nij, const std::vector<double>& R0) {
  const int n = g->n_ele;
  memset(&sd, 0, sizeof sd);
  o.X = X0;
  o.buf.assign(6 * (size_t)n, 0.0);
  o.zeros.assign((size_t)n, 0.0);
  o.aij = aij;
  o.nij.assign(nij.begin(), nij.end());
  o.rows.resize(6);
  o.brows.resize(6);
  for (int s = 0; s < 6; ++s) {
    o.rows[s] = o.X.data() + (size_t)s * n;
    o.brows[s] = o.zeros.data();
  }
  GridStruct& G = sd.Grid;
  G.N_States = 6;
  G.N_Ele = n;
  G.States = o.rows.data();
  G.States_Buffer = o.buf.data();
  G.Boundary_Rates = o.brows.data();
  G.aij = o.aij.data();
  G.nij = o.nij.data();
  G.Domain = g->domain;
  G.TropC = g->tropc;
  G.V0 = g->V0;
  G.TTSurfaceArea = g->tt_area;
  G.SRSurfaceArea = g->sr_area;
  sd.N_RyR = g->n_ryr;
  sd.N_LCC = g->n_lcc;
  sd.N_Channels = g->n_ryr + g->n_lcc;
  o.rs.assign((size_t)g->n_ryr, 0);
  o.ls.assign((size_t)g->n_lcc, 1);
  o.lvs.assign((size_t)g->n_lcc, 2);
  o.R = R0;
  sd.RyR_States = o.rs.data();
  sd.RyR_Ele = g->ryr_ele;
  sd.RyR_JSR_Ele = g->ryr_jsr;
  sd.RyR_Neighb = g->ryr_nb;
  sd.LCC_States = o.ls.data();
  sd.LCCV_States = o.lvs.data();
  sd.LCC_Ele = g->lcc_ele;
  sd.R = o.R.data();
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 5) {
    fprintf(stderr, "usage: bridge_check PARAM.xml MESH_BASE N_STEPS SEED\n");
    return 3;
  }
  const int n_steps = atoi(argv[3]);
  uint64_t lcg = strtoull(argv[4], nullptr, 10) * 2862933555777941757ull + 3037000493ull;
  ParamStruct ps;
  memset(&ps, 0, sizeof ps);
  char title[SRS_TITLE_LEN];
  srs_grid* g = nullptr;
  int32_t warn = 0;
  if (srs_read_parameters(argv[1], ps.Integers, ps.Reals, title) || srs_grid_read(argv[2], 1, &g, &warn)) {
    fprintf(stderr, "set-up: %s\n", srs_last_error());
    return 3;
  }
  ps.Integers[I_FLAG_NO_GATING] = 1;
  const int n = g->n_ele;
  std::vector<double> aij(5 * (size_t)n), X0(6 * (size_t)n), bnd(6);
  std::vector<int32_t> nij(4 * (size_t)n);
  if (srs_assemble_operator(n, g->n_face, g->ele1, g->ele2, g->face_type, g->face_c, g->V0, aij.data(), nij.data()) ||
      srs_equilibrium_state(n, g->domain, g->tropc, ps.Reals, X0.data(), bnd.data())) {
    fprintf(stderr, "set-up: %s\n", srs_last_error());
    return 3;
  }
  std::vector<double> R0((size_t)(g->n_ryr + g->n_lcc), -1.0);
  SimData a, b;
  Owned oa, ob;
  bind(a, oa, g, X0, aij, nij, R0);
  bind(b, ob, g, X0, aij, nij, R0);
  a.Global_Vars[GLOBAL_VM] = b.Global_Vars[GLOBAL_VM] = 0.0;
  a.Global_Vars[GLOBAL_CA_NSR] = b.Global_Vars[GLOBAL_CA_NSR] = ps.Reals[I_DEFAULT_CA_NSR];
  const double dt = ps.Reals[I_T_STEP];
  SrsInit(&a, &ps);
  int opened = 0;
  for (int k = 0; k < n_steps; ++k) {
    for (int i = 0; i < g->n_ryr; ++i) {
      lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
      const int st = (lcg >> 33) % 4 == 0;
      a.RyR_States[i] = b.RyR_States[i] = st;
      opened += st;
    }
    SrsExplicitStep(dt, &a, &ps);
    ExplicitStep(dt, &b, &ps);
  }
  SrsPullStates(&a);
  SrsFree();
  size_t diff = 0;
  for (int s = 0; s < 6; ++s)
    for (int e = 0; e < n; ++e)
      if (memcmp(&a.Grid.States[s][e], &b.Grid.States[s][e], sizeof(double)) != 0) ++diff;
  size_t moved = 0;
  for (size_t q = 0; q < X0.size(); ++q) moved += oa.X[q] != X0[q];
  printf("bridge_check: %d elements, %d RyRs, %d steps, %d RyR-steps open, %zu values moved, %zu differ\n", n,
         g->n_ryr, n_steps, opened, moved, diff);
  srs_grid_free(g);
  return diff == 0 && moved > 0 && opened > 0 ? 0 : 2;
}
