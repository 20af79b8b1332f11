/*
 * ref_harness.cpp -- thin C entry points around the reference's own CPU step.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/build_ref.sh together with the
 * unmodified reference source cpu/RunSimulation.cpp (read where it lies; it
 * is never copied into this repository).  The reference header
 * cpu/parameters.h supplies the SimData/GridStruct/ParamStruct layouts.
 *
 * Only ExplicitStep (cpu/RunSimulation.cpp:591) and UpdateChannelGating
 * (:656) are reached.  RunSimulation (:402) also needs the MPI-based output
 * functions, which are not built; RunSimulation.o is compiled with hidden
 * visibility and the link drops unreferenced sections, so it never enters
 * the library.
 */
#include <stdlib.h>
#include <string.h>
#include "parameters.h"

namespace {

struct Bound {
  SimData sd;
  ParamStruct ps;
  double* rows[N_STATES];
  double* brows[N_STATES];
  double* zeros;
};

void bind(Bound& b, int n_ele, const double* aij, const int* nij, const double* br, const int* domain,
          const int* tropc, const double* V0, const double* tt, int n_ryr, int n_lcc, int* ryr_state,
          const int* ryr_ele, const int* ryr_jsr, const int* ryr_nb, int* lcc_state, int* lccv_state,
          const int* lcc_ele, double* R, const int* ints, const double* reals, double* gv, double* X) {
  memset(&b.sd, 0, sizeof(b.sd));
  memset(&b.ps, 0, sizeof(b.ps));
  memcpy(b.ps.Integers, ints, sizeof(int) * N_PARAMETERS_INT);
  memcpy(b.ps.Reals, reals, sizeof(double) * N_PARAMETERS_REAL);
  GridStruct& g = b.sd.Grid;
  g.N_States = N_STATES;
  g.N_Ele = n_ele;
  for (int s = 0; s < N_STATES; ++s) b.rows[s] = X + (size_t)s * n_ele;
  g.States = b.rows;
  g.States_Buffer = (double*)calloc((size_t)N_STATES * n_ele, sizeof(double));
  b.zeros = nullptr;
  if (br) {
    for (int s = 0; s < N_STATES; ++s) b.brows[s] = const_cast<double*>(br) + (size_t)s * n_ele;
  } else {
    b.zeros = (double*)calloc((size_t)n_ele, sizeof(double));
    for (int s = 0; s < N_STATES; ++s) b.brows[s] = b.zeros;
  }
  g.Boundary_Rates = b.brows;
  g.aij = const_cast<double*>(aij);
  g.nij = const_cast<int*>(nij);
  g.Domain = const_cast<int*>(domain);
  g.TropC = const_cast<int*>(tropc);
  g.V0 = const_cast<double*>(V0);
  g.TTSurfaceArea = const_cast<double*>(tt);
  g.SRSurfaceArea = nullptr;
  b.sd.N_RyR = n_ryr;
  b.sd.N_LCC = n_lcc;
  b.sd.N_Channels = n_ryr + n_lcc;
  b.sd.RyR_States = ryr_state;
  b.sd.RyR_Ele = const_cast<int*>(ryr_ele);
  b.sd.RyR_JSR_Ele = const_cast<int*>(ryr_jsr);
  b.sd.RyR_Neighb = const_cast<int*>(ryr_nb);
  b.sd.LCC_States = lcc_state;
  b.sd.LCCV_States = lccv_state;
  b.sd.LCC_Ele = const_cast<int*>(lcc_ele);
  b.sd.R = R;
  b.sd.Global_Vars[0] = gv[0];
  b.sd.Global_Vars[1] = gv[1];
}

void unbind(Bound& b, double* gv) {
  gv[0] = b.sd.Global_Vars[0];
  gv[1] = b.sd.Global_Vars[1];
  free(b.sd.Grid.States_Buffer);
  free(b.zeros);
}

}  // namespace

extern "C" {

__attribute__((visibility("default"))) void ref_srand(unsigned seed) { srand(seed); }

/* Runs n_steps reference ExplicitStep calls.  If schedule != NULL, RyR states
 * are overwritten from schedule[k*n_ryr + i] before step k (the replay mode:
 * flag_nogating must then be set in ints).  Returns 0. */
__attribute__((visibility("default"))) int ref_explicit_steps(
    int n_steps, double dt, int n_ele, const double* aij, const int* nij, const double* br,
    const int* domain, const int* tropc, const double* V0, const double* tt, int n_ryr, int n_lcc,
    int* ryr_state, const int* ryr_ele, const int* ryr_jsr, const int* ryr_nb, int* lcc_state,
    int* lccv_state, const int* lcc_ele, double* R, const int* ints, const double* reals, double* gv,
    double* X, const int* schedule) {
  Bound b;
  bind(b, n_ele, aij, nij, br, domain, tropc, V0, tt, n_ryr, n_lcc, ryr_state, ryr_ele, ryr_jsr, ryr_nb,
       lcc_state, lccv_state, lcc_ele, R, ints, reals, gv, X);
  for (int k = 0; k < n_steps; ++k) {
    if (schedule)
      for (int i = 0; i < n_ryr; ++i) ryr_state[i] = schedule[(size_t)k * n_ryr + i];
    ExplicitStep(dt, &b.sd, &b.ps);
  }
  unbind(b, gv);
  return 0;
}

/* One reference UpdateChannelGating call on the given fields. */
__attribute__((visibility("default"))) int ref_channel_gating(
    double dt, int n_ele, const double* V0, int n_ryr, int n_lcc, int* ryr_state, const int* ryr_ele,
    const int* ryr_jsr, const int* ryr_nb, int* lcc_state, int* lccv_state, const int* lcc_ele, double* R,
    const int* ints, const double* reals, double* gv, double* X) {
  Bound b;
  static const double dummy_a[5] = {0, 0, 0, 0, 0};
  bind(b, n_ele, dummy_a, nullptr, nullptr, nullptr, nullptr, V0, nullptr, n_ryr, n_lcc, ryr_state, ryr_ele,
       ryr_jsr, ryr_nb, lcc_state, lccv_state, lcc_ele, R, ints, reals, gv, X);
  UpdateChannelGating(dt, &b.sd, &b.ps);
  unbind(b, gv);
  return 0;
}

}  // extern "C"
