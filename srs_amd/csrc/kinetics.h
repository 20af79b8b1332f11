// kinetics.h -- channel kinetics and the states record, shared by the host
// C++ code and the HIP kernels of libsrs_amd.  Expressions keep the
// reference's operation order (the library is built with FP contraction
// off), so that host and device agree with the reference's CPU build.
#pragma once
#include <math.h>
#include <stdint.h>

#include "philox.h"
#include "srs_abi.h"

// LCC constants, cpu/RunSimulation.cpp:149-151.
#define SRS_P_CAL ((9.13e-13) * (1e12) * (1e-3))
#define SRS_F_CONST 96.5
#define SRS_RT_CONST (310 * 8.314)

// GHK-type LCC flux numerator/denominator, cpu/RunSimulation.cpp:153-161.
// Returns the flux before the division by the element volume.
SRS_HD double srs_lcc_flux_raw(double V, double c, double ca_o) {
  const double F_over_RT = SRS_F_CONST / SRS_RT_CONST;
  if (fabs(V) < 1e-6) V = 1e-6;
  return -SRS_P_CAL * V * (2 * F_over_RT) * (c * exp(2 * V * F_over_RT) - 0.34 * ca_o) /
         (exp(2 * V * F_over_RT) - 1);
}

SRS_HD bool srs_lcc_open(int32_t s, int32_t vs) { return (s == 6 || s == 12) && vs == 2; }

// RyR transition rate, cpu/RunSimulation.cpp:172-202.
SRS_HD double srs_ryr_rate(int32_t state, int n_open, int n_closed, double ca_ss, double ca_jsr, const double* P) {
  double X, rate;
  if (state) {
    X = exp(0.5 * P[SRS_R_RYR_A_STAR] * (n_open * P[SRS_R_RYR_EPS_OO] - n_closed * P[SRS_R_RYR_EPS_CC]));
    rate = X * P[SRS_R_RYR_K_MINUS];
  } else {
    double phi = (1.0 - pow(1000.0 / P[SRS_R_RYR_PHI_B], P[SRS_R_RYR_PHI_M])) +
                 pow(ca_jsr / P[SRS_R_RYR_PHI_B], P[SRS_R_RYR_PHI_M]);
    X = exp(0.5 * P[SRS_R_RYR_A_STAR] * (n_closed * P[SRS_R_RYR_EPS_CC] - n_open * P[SRS_R_RYR_EPS_OO]));
    rate = X * phi * P[SRS_R_RYR_K_PLUS] * pow(ca_ss, P[SRS_R_RYR_ETA]);
  }
  return rate;
}

// LCC 12-state mode-switching model plus V-dependent gate,
// cpu/RunSimulation.cpp:235-377.  Fills the three transition rates and
// targets of state s and returns the gate rate.
SRS_HD double srs_lcc_rates(int32_t s, int32_t vs, double V, double ca_ss, double tr[3], int32_t tg[3]) {
  const double fL = 0.85, gL = 2.0, fLprime = 0.005, gLprime = 7.0, bL = 1.9356;
  const double bL2 = bL * bL, bL3 = bL * bL * bL, bL4 = bL * bL * bL * bL;
  const double aL = 2.0, aL2 = aL * aL, aL3 = aL * aL * aL, aL4 = aL * aL * aL * aL;
  const double omega = 0.83 * 2.0 * 1.3 * 0.01;
  const double alphacf = 4.0 * 1.2 * 0.416, betacf = 4.0 * 0.45 * 0.049;
  const double gammacf = 0.83 * 1.9 * 1.3 * 0.31 * 7.5 * 0.09233 / 1000.0;
  const double yCa_frac = 0.4;
  const double alpha = alphacf * exp(0.012 * (V - 35.0));
  const double beta = betacf * exp(-0.05 * (V - 35.0));
  const double ap = aL * alpha, bp = beta / bL;
  const double g = gammacf * ca_ss;
  tr[0] = tr[1] = tr[2] = 0;
  tg[0] = tg[1] = tg[2] = 0;
  switch (s) {
    case 1: tr[0] = 4.0 * alpha; tr[1] = g; tg[0] = 2; tg[1] = 7; break;
    case 2: tr[0] = beta; tr[1] = 3.0 * alpha; tr[2] = aL * g; tg[0] = 1; tg[1] = 3; tg[2] = 8; break;
    case 3: tr[0] = 2.0 * beta; tr[1] = 2.0 * alpha; tr[2] = aL2 * g; tg[0] = 2; tg[1] = 4; tg[2] = 9; break;
    case 4: tr[0] = 3.0 * beta; tr[1] = alpha; tr[2] = aL3 * g; tg[0] = 3; tg[1] = 5; tg[2] = 10; break;
    case 5: tr[0] = 4.0 * beta; tr[1] = fL; tr[2] = aL4 * g; tg[0] = 4; tg[1] = 6; tg[2] = 11; break;
    case 6: tr[0] = gL; tg[0] = 5; break;
    case 7: tr[0] = omega; tr[1] = 4.0 * ap; tg[0] = 1; tg[1] = 8; break;
    case 8: tr[0] = bp; tr[1] = omega / bL; tr[2] = 3.0 * ap; tg[0] = 7; tg[1] = 2; tg[2] = 9; break;
    case 9: tr[0] = 2.0 * bp; tr[1] = omega / bL2; tr[2] = 2.0 * ap; tg[0] = 8; tg[1] = 3; tg[2] = 10; break;
    case 10: tr[0] = 3.0 * bp; tr[1] = omega / bL3; tr[2] = ap; tg[0] = 9; tg[1] = 4; tg[2] = 11; break;
    case 11: tr[0] = 4.0 * bp; tr[1] = omega / bL4; tr[2] = fLprime; tg[0] = 10; tg[1] = 5; tg[2] = 12; break;
    case 12: tr[0] = gLprime; tg[0] = 11; break;
    default: break;
  }
  const double yCa_inf = yCa_frac / (1.0 + exp((V + 12.5) / 5.0)) + (1.0 - yCa_frac);
  const double tau_yCa = 60.0 + 340.0 / (1.0 + exp((V + 30.0) / 12.0));
  return (vs == 1) ? yCa_inf / tau_yCa : (1.0 - yCa_inf) / tau_yCa;
}

// States record, cpu/OutputFuncs.cpp:244-350 (gpu_layout: plus Ca_cyt at each
// RyR, gpu/OutputFuncs.cpp:277-284).  Returns the number of floats.
SRS_HD int srs_record_fill(double t, int32_t n_ryr, const int32_t* ryr_state, const double* ca_cyt,
                           const double* ca_jsr, int32_t n_lcc, const int32_t* lcc_state,
                           const int32_t* lccv_state, const double* ca_lcc, double vm, const double* P,
                           int ryr_out, int gpu_layout, float* rec) {
  int k = 0;
  rec[k++] = (float)t;
  double i_ryr = 0;
  for (int i = 0; i < n_ryr; ++i)
    i_ryr += ryr_state[i] * P[SRS_R_V_RYR] * (ca_jsr[i] - ca_cyt[i]) * (P[SRS_R_V_CELL]) * 96485 * 2 * (1e-6);
  rec[k++] = (float)i_ryr;
  int n_open = 0;
  for (int i = 0; i < n_ryr; ++i) n_open += ryr_state[i];
  rec[k++] = (float)n_open;
  int n_open_lcc = 0;
  for (int i = 0; i < n_lcc; ++i)
    if (lccv_state[i] == 2 && (lcc_state[i] == 6 || lcc_state[i] == 12)) n_open_lcc++;
  rec[k++] = (float)n_open_lcc;
  double i_lcc = 0;
  for (int i = 0; i < n_lcc; ++i)
    if (srs_lcc_open(lcc_state[i], lccv_state[i]))
      i_lcc += srs_lcc_flux_raw(vm, ca_lcc[i], P[SRS_R_CA_0]) * 96485 * 2 * (1e-6);
  rec[k++] = (float)i_lcc;
  double m = 0;
  for (int i = 0; i < n_ryr; ++i) m += ca_jsr[i];
  m /= n_ryr;
  rec[k++] = (float)m;
  m = 0;
  for (int i = 0; i < n_ryr; ++i) m += ca_cyt[i];
  m /= n_ryr;
  rec[k++] = (float)m;
  if (gpu_layout)
    for (int i = 0; i < n_ryr; ++i) rec[k++] = (float)ca_cyt[i];
  if (ryr_out)
    for (int i = 0; i < n_ryr; ++i) rec[k++] = (float)ryr_state[i];
  return k;
}
