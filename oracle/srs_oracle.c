/*
 * srs_oracle.c -- CPU restatement of the reference spark time-step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY: only tests/, bench.py's cpu_baseline leg and
 * __graft_entry__.smoke() may load this.  Compiled with -O2 -ffp-contract=off
 * so that the operation order below is the order executed.
 *
 * Functions that have no check beyond the ones listed in srs_oracle.h are
 * marked "parity unpinned" there.
 */
#include "srs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* Philox4x64-10 (Salmon et al., SC'11).  Round function and constants as     */
/* published with Random123; pinned in tests against numpy.random.Philox.    */
/* ------------------------------------------------------------------------ */
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
  unsigned __int128 p = (unsigned __int128)a * (unsigned __int128)b;
  *hi = (uint64_t)(p >> 64);
  *lo = (uint64_t)p;
}

void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
  uint64_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint64_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += W0; k1 += W1; }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(M0, c0, &hi0, &lo0);
    mulhilo64(M1, c2, &hi1, &lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 52 random bits mapped to the open interval (0,1): the extremes are 2^-53 and 1 - 2^-53. */
double orc_u64_to_open01(uint64_t x) {
  return ((double)(x >> 12) + 0.5) * (1.0 / 4503599627370496.0);
}

/* ------------------------------------------------------------------------ */
/* Step: cpu/RunSimulation.cpp:591-648                                        */
/* ------------------------------------------------------------------------ */

/* cpu/RunSimulation.cpp:25-125 (UpdateReactionTerms) */
static void reaction_terms(int j, const orc_grid* g, const double* P, const double* X, double* buf,
                           const double* gv) {
  const int N = g->n_ele;
  const double* Ca = X + ORC_CA * N;
  const double c_ca = Ca[j];
  double r_net, c_b;
  buf[ORC_TRPN * N + j] = 0;
  buf[ORC_CSQN * N + j] = 0;

  const double a0 = g->aij[5 * j], a1 = g->aij[5 * j + 1], a2 = g->aij[5 * j + 2],
               a3 = g->aij[5 * j + 3], a4 = g->aij[5 * j + 4];
  const int n1 = g->nij[4 * j], n2 = g->nij[4 * j + 1], n3 = g->nij[4 * j + 2], n4 = g->nij[4 * j + 3];

  /* :61-71, order Ca, Dye, ATP, CaM; Ca uses D_CA in both domains */
  static const int mob[4] = {ORC_CA, ORC_DYE, ORC_ATP, ORC_CMDN};
  static const int dix[4] = {ORC_D_CA, ORC_D_DYE, ORC_D_ATP, ORC_D_CMDN};
  for (int m = 0; m < 4; ++m) {
    const double* S = X + mob[m] * N;
    double br = g->br ? g->br[mob[m] * N + j] : 0.0;
    buf[mob[m] * N + j] = P[dix[m]] * (S[j] * a0 + S[n1] * a1 + S[n2] * a2 + S[n3] * a3 + S[n4] * a4) + br;
  }

  double* bca = buf + ORC_CA * N + j;
  if (g->domain[j] == 0) {
    if (g->tropc[j]) { /* :75-91 */
      c_b = X[ORC_TRPN * N + j];
      r_net = P[ORC_K_ON_TRPN] * (P[ORC_B_TOT_TRPN] - c_b) * c_ca - P[ORC_K_OFF_TRPN] * c_b;
      *bca += -r_net;
      buf[ORC_TRPN * N + j] = r_net;
      double K_i = c_ca / P[ORC_K_D_I];
      K_i = K_i * K_i;
      double K_sr = gv[1] / P[ORC_K_D_SR];
      K_sr = K_sr * K_sr;
      double D_cycle = 0.104217 + 17.923 * K_sr + K_i * (1.75583e6 + K_sr * 7.61673e6) +
                       K_i * K_i * (6.08463e11 + K_sr * 4.50544e11);
      double v_cycle = (K_i * K_i * 3.24873e12 + K_i * (9.17846e6 - 11478.2 * K_sr) - 0.329904 * K_sr) / D_cycle;
      double d_ca = (2e-3) * v_cycle * P[ORC_A_P];
      *bca -= d_ca;
    }
    /* :94-109 dye, ATP, calmodulin */
    c_b = X[ORC_DYE * N + j];
    r_net = P[ORC_K_ON_DYE] * (P[ORC_B_TOT_DYE] - c_b) * c_ca - P[ORC_K_OFF_DYE] * c_b;
    *bca += -r_net;
    buf[ORC_DYE * N + j] += r_net;
    c_b = X[ORC_ATP * N + j];
    r_net = P[ORC_K_ON_ATP] * (P[ORC_B_TOT_ATP] - c_b) * c_ca - P[ORC_K_OFF_ATP] * c_b;
    *bca += -r_net;
    buf[ORC_ATP * N + j] += r_net;
    c_b = X[ORC_CMDN * N + j];
    r_net = P[ORC_K_ON_CMDN] * (P[ORC_B_TOT_CMDN] - c_b) * c_ca - P[ORC_K_OFF_CMDN] * c_b;
    *bca += -r_net;
    buf[ORC_CMDN * N + j] += r_net;
  } else { /* :113-121 calsequestrin, refill */
    c_b = X[ORC_CSQN * N + j];
    r_net = P[ORC_K_ON_CSQN] * (P[ORC_B_T_JSR] - c_b) * c_ca - P[ORC_K_OFF_CSQN] * c_b;
    *bca += -r_net;
    buf[ORC_CSQN * N + j] = r_net;
    *bca += (gv[1] - c_ca) * P[ORC_V_REFILL];
  }
}

/* cpu/RunSimulation.cpp:143-164 constants */
#define ORC_P_CAL ((9.13e-13) * (1e12) * (1e-3))
#define ORC_F 96.5
#define ORC_RT (310 * 8.314)

void orc_explicit_step(double dt, const orc_grid* g, const double* P, double* X, double* buf,
                       double* gv, const orc_chan* ch) {
  const int N = g->n_ele;
  for (int j = 0; j < N; ++j) reaction_terms(j, g, P, X, buf, gv);

  const double* Ca = X + ORC_CA * N;
  double* bca = buf + ORC_CA * N;
  /* :127-141 RyR flux, in channel order */
  for (int i = 0; i < ch->n_ryr; ++i) {
    if (ch->ryr_state[i]) {
      int ce = ch->ryr_ele[i], je = ch->ryr_jsr[i];
      double J = P[ORC_V_RYR] * (Ca[je] - Ca[ce]) * P[ORC_V_CELL];
      bca[ce] += J / g->V0[ce];
      bca[je] -= J / g->V0[je];
    }
  }
  /* :143-164 LCC flux */
  for (int i = 0; i < ch->n_lcc; ++i) {
    int s = ch->lcc_state[i];
    if ((s == 6 || s == 12) && ch->lccv_state[i] == 2) {
      const double F_over_RT = ORC_F / ORC_RT;
      double V = gv[0];
      int e = ch->lcc_ele[i];
      double c = Ca[e];
      if (fabs(V) < 1e-6) V = 1e-6;
      double J = -ORC_P_CAL * V * (2 * F_over_RT) * (c * exp(2 * V * F_over_RT) - 0.34 * P[ORC_CA_0]) /
                 (exp(2 * V * F_over_RT) - 1) / g->V0[e];
      bca[e] += J;
    }
  }
  /* :15-22 rapid sarcolemmal buffer, applied last */
  for (int j = 0; j < N; ++j) {
    double btothi = g->tt_area[j] * P[ORC_B_TOT_SL_HI] / (g->V0[j] * (1e-15));
    double btotlo = g->tt_area[j] * P[ORC_B_TOT_SL_LO] / (g->V0[j] * (1e-15));
    double c = Ca[j];
    double B = 1 + (btothi * P[ORC_K_D_SL_HI] / ((P[ORC_K_D_SL_HI] + c) * (P[ORC_K_D_SL_HI] + c))) +
               (btotlo * P[ORC_K_D_SL_LO] / ((P[ORC_K_D_SL_LO] + c) * (P[ORC_K_D_SL_LO] + c)));
    bca[j] /= B;
  }
  /* :640-647 forward Euler, species-major; global sources are always zero */
  for (int s = 0; s < ORC_NS; ++s)
    for (int j = 0; j < N; ++j) X[s * N + j] += buf[s * N + j] * dt;
  gv[0] += 0.0 * dt;
  gv[1] += 0.0 * dt;
}

/* ------------------------------------------------------------------------ */
/* Gating: cpu/RunSimulation.cpp:166-400, 656-674                             */
/* ------------------------------------------------------------------------ */
static double draw(orc_rng* rng, uint64_t channel, int which) {
  if (rng->kind == ORC_RNG_GLIBC) return ((double)rand()) / ((double)RAND_MAX);
  uint64_t ctr[4] = {rng->step, channel, rng->realization, 0};
  uint64_t key[2] = {rng->key0, rng->key1};
  uint64_t out[4];
  orc_philox4x64(ctr, key, out);
  return orc_u64_to_open01(out[which]);
}

void orc_gating(double dt, const orc_grid* g, const double* P, const double* X, const double* gv,
                orc_chan* ch, orc_rng* rng) {
  const double* Ca = X + ORC_CA * g->n_ele;
  /* RyRs in index order; later RyRs see earlier flips (:661-666) */
  for (int idx = 0; idx < ch->n_ryr; ++idx) {
    int N_Open = 0, N_Closed = 0;
    for (int i = 0; i < 4; ++i) {
      int nb = ch->ryr_nb[idx * 4 + i];
      if (nb >= 0) {
        if (ch->ryr_state[nb]) N_Open++; else N_Closed++;
      }
    }
    double X_, rate;
    double ca_ss = Ca[ch->ryr_ele[idx]];
    if (ch->ryr_state[idx]) {
      X_ = exp(0.5 * P[ORC_RYR_A_STAR] * (N_Open * P[ORC_RYR_EPS_OO] - N_Closed * P[ORC_RYR_EPS_CC]));
      rate = X_ * P[ORC_RYR_K_MINUS];
    } else {
      double phi = (1.0 - pow(1000.0 / P[ORC_RYR_PHI_B], P[ORC_RYR_PHI_M])) +
                   pow(Ca[ch->ryr_jsr[idx]] / P[ORC_RYR_PHI_B], P[ORC_RYR_PHI_M]);
      X_ = exp(0.5 * P[ORC_RYR_A_STAR] * (N_Closed * P[ORC_RYR_EPS_CC] - N_Open * P[ORC_RYR_EPS_OO]));
      rate = X_ * phi * P[ORC_RYR_K_PLUS] * pow(ca_ss, P[ORC_RYR_ETA]);
    }
    ch->R[idx] += rate * dt;
    if (ch->R[idx] > 0) {
      ch->R[idx] = log(draw(rng, (uint64_t)idx, 0));
      ch->ryr_state[idx] = ch->ryr_state[idx] == 1 ? 0 : 1;
    }
  }

  /* LCCs (:229-400) */
  const double fL = 0.85, gL = 2.0, fLprime = 0.005, gLprime = 7.0, bL = 1.9356;
  const double bL2 = bL * bL, bL3 = bL * bL * bL, bL4 = bL * bL * bL * bL;
  const double aL = 2.0, aL2 = aL * aL, aL3 = aL * aL * aL, aL4 = aL * aL * aL * aL;
  const double omega = 0.83 * 2.0 * 1.3 * 0.01;
  const double alphacf = 4.0 * 1.2 * 0.416, betacf = 4.0 * 0.45 * 0.049;
  const double gammacf = 0.83 * 1.9 * 1.3 * 0.31 * 7.5 * 0.09233 / 1000.0;
  const double CCa0_to_C0 = omega, CCa1_to_C1 = omega / bL, CCa2_to_C2 = omega / bL2,
               CCa3_to_C3 = omega / bL3, CCa4_to_C4 = omega / bL4;
  const double yCa_frac = 0.4;
  for (int idx = 0; idx < ch->n_lcc; ++idx) {
    double V = gv[0];
    double ca_ss = Ca[ch->lcc_ele[idx]];
    double alpha = alphacf * exp(0.012 * (V - 35.0));
    double beta = betacf * exp(-0.05 * (V - 35.0));
    double alpha_prime = aL * alpha, beta_prime = beta / bL;
    double gamma_rate = gammacf * ca_ss;
    int i_lcc = idx + ch->n_ryr;
    int s = ch->lcc_state[idx];
    double tr[3] = {0, 0, 0};
    int tg[3] = {0, 0, 0};
    switch (s) {
      case 1: tr[0] = 4.0 * alpha; tr[1] = gamma_rate; tg[0] = 2; tg[1] = 7; break;
      case 2: tr[0] = beta; tr[1] = 3.0 * alpha; tr[2] = aL * gamma_rate; tg[0] = 1; tg[1] = 3; tg[2] = 8; break;
      case 3: tr[0] = 2.0 * beta; tr[1] = 2.0 * alpha; tr[2] = aL2 * gamma_rate; tg[0] = 2; tg[1] = 4; tg[2] = 9; break;
      case 4: tr[0] = 3.0 * beta; tr[1] = alpha; tr[2] = aL3 * gamma_rate; tg[0] = 3; tg[1] = 5; tg[2] = 10; break;
      case 5: tr[0] = 4.0 * beta; tr[1] = fL; tr[2] = aL4 * gamma_rate; tg[0] = 4; tg[1] = 6; tg[2] = 11; break;
      case 6: tr[0] = gL; tg[0] = 5; break;
      case 7: tr[0] = CCa0_to_C0; tr[1] = 4.0 * alpha_prime; tg[0] = 1; tg[1] = 8; break;
      case 8: tr[0] = beta_prime; tr[1] = CCa1_to_C1; tr[2] = 3.0 * alpha_prime; tg[0] = 7; tg[1] = 2; tg[2] = 9; break;
      case 9: tr[0] = 2.0 * beta_prime; tr[1] = CCa2_to_C2; tr[2] = 2.0 * alpha_prime; tg[0] = 8; tg[1] = 3; tg[2] = 10; break;
      case 10: tr[0] = 3.0 * beta_prime; tr[1] = CCa3_to_C3; tr[2] = alpha_prime; tg[0] = 9; tg[1] = 4; tg[2] = 11; break;
      case 11: tr[0] = 4.0 * beta_prime; tr[1] = CCa4_to_C4; tr[2] = fLprime; tg[0] = 10; tg[1] = 5; tg[2] = 12; break;
      case 12: tr[0] = gLprime; tg[0] = 11; break;
      default: break;
    }
    int vs = ch->lccv_state[idx];
    double yCa_inf = yCa_frac / (1.0 + exp((V + 12.5) / 5.0)) + (1.0 - yCa_frac);
    double tau_yCa = 60.0 + 340.0 / (1.0 + exp((V + 30.0) / 12.0));
    double lccv_rate = (vs == 1) ? yCa_inf / tau_yCa : (1.0 - yCa_inf) / tau_yCa;
    ch->R[i_lcc] += (tr[0] + tr[1] + tr[2] + lccv_rate) * dt;
    if (ch->R[i_lcc] > 0) {
      ch->R[i_lcc] = log(draw(rng, (uint64_t)i_lcc, 0));
      double runif = draw(rng, (uint64_t)i_lcc, 1) * (tr[0] + tr[1] + tr[2] + lccv_rate);
      if (runif < tr[0]) ch->lcc_state[idx] = tg[0];
      else if (runif < tr[0] + tr[1]) ch->lcc_state[idx] = tg[1];
      else if (runif < tr[0] + tr[1] + tr[2]) ch->lcc_state[idx] = tg[2];
      else ch->lccv_state[idx] = (vs == 1) ? 2 : 1;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Operator: cpu/InitFuncs.cpp:1058-1112 (primary gradient), 1273-1322 (ELL)  */
/* ------------------------------------------------------------------------ */
typedef struct { int col; double val; } orc_entry;

int orc_assemble_operator(int n_ele, int n_face, const int* ele1, const int* ele2, const int* ftype,
                          const double* face_c, const double* V0, double* aij, int* nij) {
  /* Per-row entry lists; at most 5 distinct columns per row for a tet mesh,
   * but a malformed mesh may have more, so rows grow on demand. */
  int* cnt = (int*)calloc((size_t)n_ele, sizeof(int));
  int* cap = (int*)malloc((size_t)n_ele * sizeof(int));
  orc_entry** rows = (orc_entry**)malloc((size_t)n_ele * sizeof(orc_entry*));
  for (int i = 0; i < n_ele; ++i) { cap[i] = 6; rows[i] = (orc_entry*)malloc(6 * sizeof(orc_entry)); }
#define ORC_ADD(r, c, v)                                                              \
  do {                                                                                \
    int _r = (r), _c = (c), _k;                                                       \
    for (_k = 0; _k < cnt[_r]; ++_k) if (rows[_r][_k].col == _c) break;               \
    if (_k == cnt[_r]) {                                                              \
      if (cnt[_r] == cap[_r]) { cap[_r] *= 2; rows[_r] = (orc_entry*)realloc(rows[_r], cap[_r] * sizeof(orc_entry)); } \
      rows[_r][_k].col = _c; rows[_r][_k].val = 0.0; cnt[_r]++;                       \
    }                                                                                 \
    rows[_r][_k].val += (v);                                                          \
  } while (0)
  for (int f = 0; f < n_face; ++f) {
    int t1 = ele1[f], t2 = ele2[f];
    double diff = 1;
    double k1 = diff * (face_c[f]) / V0[t1];
    if (ftype[f] == 0) {
      ORC_ADD(t1, t2, k1);
      ORC_ADD(t1, t1, -k1);
      double k2 = diff * (face_c[f]) / V0[t2];
      ORC_ADD(t2, t1, k2);
      ORC_ADD(t2, t2, -k2);
    } else if (ftype[f] == 1) {
      ORC_ADD(t1, t1, -k1);
    }
  }
#undef ORC_ADD
  int rc = 0;
  for (int i = 0; i < 5 * n_ele; ++i) aij[i] = 0;
  for (int i = 0; i < n_ele; ++i) nij[4 * i] = nij[4 * i + 1] = nij[4 * i + 2] = nij[4 * i + 3] = i;
  for (int r = 0; r < n_ele; ++r) {
    /* map iteration order: columns ascending (insertion sort, rows are tiny) */
    for (int a = 1; a < cnt[r]; ++a) {
      orc_entry e = rows[r][a];
      int b = a - 1;
      while (b >= 0 && rows[r][b].col > e.col) { rows[r][b + 1] = rows[r][b]; --b; }
      rows[r][b + 1] = e;
    }
    int nf = 0;
    for (int k = 0; k < cnt[r]; ++k) {
      if (rows[r][k].val == 0) continue;
      if (rows[r][k].col == r) {
        aij[5 * r] = rows[r][k].val;
      } else {
        if (nf >= 4) { rc = -1; continue; }
        aij[5 * r + 1 + nf] = rows[r][k].val;
        nij[4 * r + nf] = rows[r][k].col;
        nf++;
      }
    }
  }
  for (int i = 0; i < n_ele; ++i) free(rows[i]);
  free(rows); free(cap); free(cnt);
  return rc;
}

void orc_boundary_rates(int n_ele, int n_face, const int* ele1, const int* ftype, const double* face_c,
                        const double* V0, const double* diff, const double* bnd, double* br) {
  for (int k = 0; k < ORC_NS; ++k) {
    for (int i = 0; i < n_ele; ++i) br[k * n_ele + i] = 0;
    for (int f = 0; f < n_face; ++f) {
      int t1 = ele1[f];
      if (ftype[f] == 1) br[k * n_ele + t1] += diff[k] * (face_c[f]) * bnd[k] / V0[t1];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Renumbering: cpu/InitFuncs.cpp:679-731 (BFS from element 0 over faces     */
/* with Ele_2 >= 0, neighbours discovered in face-file order).               */
/* Written as the reference writes it (linear scans), for small meshes.      */
/* ------------------------------------------------------------------------ */
int orc_bfs_order(int n_ele, int n_face, const int* ele1, const int* ele2, int* visited) {
  int* e1 = (int*)malloc((size_t)n_face * sizeof(int));
  int* e2 = (int*)malloc((size_t)n_face * sizeof(int));
  int ne = 0;
  for (int i = 0; i < n_face; ++i) if (ele2[i] >= 0) { e1[ne] = ele1[i]; e2[ne] = ele2[i]; ++ne; }
  int* queue = (int*)malloc((size_t)(n_ele + 1) * sizeof(int));
  int qh = 0, qt = 0, nv = 0;
  queue[qt++] = 0;
  while (qh < qt) {
    int v = queue[qh++];
    visited[nv++] = v;
    for (int i = 0; i < ne; ++i) {
      if (e1[i] == v || e2[i] == v) {
        int nb = (e1[i] == v) ? e2[i] : e1[i];
        int skip = 0;
        for (int j = qh; j < qt; ++j) if (queue[j] == nb) { skip = 1; break; }
        if (!skip) for (int j = 0; j < nv; ++j) if (visited[j] == nb) { skip = 1; break; }
        if (!skip) queue[qt++] = nb;
      }
    }
  }
  free(queue); free(e1); free(e2);
  return nv;
}

/* cpu/InitFuncs.cpp:890-954 */
void orc_equilibrium_state(int n_ele, const int* domain, const int* tropc, const double* P, double* X,
                           double* boundaries) {
  double ca0 = P[ORC_C_0_CA], ca0_jsr = P[ORC_C_0_CA_JSR], Kd;
  Kd = P[ORC_K_OFF_TRPN] / P[ORC_K_ON_TRPN];
  double c_trpn = ca0 * P[ORC_B_TOT_TRPN] / (Kd + ca0);
  Kd = P[ORC_K_OFF_CSQN] / P[ORC_K_ON_CSQN];
  double c_csqn = ca0_jsr * P[ORC_B_T_JSR] / (Kd + ca0_jsr);
  Kd = P[ORC_K_OFF_ATP] / P[ORC_K_ON_ATP];
  double c_atp = ca0 * P[ORC_B_TOT_ATP] / (Kd + ca0);
  Kd = P[ORC_K_OFF_CMDN] / P[ORC_K_ON_CMDN];
  double c_cmdn = ca0 * P[ORC_B_TOT_CMDN] / (Kd + ca0);
  Kd = P[ORC_K_OFF_DYE] / P[ORC_K_ON_DYE];
  double c_dye = ca0 * P[ORC_B_TOT_DYE] / (Kd + ca0);
  if (boundaries) {
    boundaries[ORC_CA] = ca0; boundaries[ORC_TRPN] = c_trpn; boundaries[ORC_CSQN] = 0;
    boundaries[ORC_ATP] = c_atp; boundaries[ORC_CMDN] = c_cmdn; boundaries[ORC_DYE] = c_dye;
  }
  const int N = n_ele;
  for (int j = 0; j < N; ++j) {
    if (domain[j] == 0) {
      X[ORC_CA * N + j] = ca0;
      X[ORC_TRPN * N + j] = c_trpn * tropc[j];
      X[ORC_ATP * N + j] = c_atp;
      X[ORC_CMDN * N + j] = c_cmdn;
      X[ORC_DYE * N + j] = c_dye;
      /* CSQN is not assigned for cytosol by the reference; its buffer is
       * allocated with new[] and in practice starts at 0 in the configs. */
      X[ORC_CSQN * N + j] = 0;
    } else if (domain[j] == 1) {
      X[ORC_CA * N + j] = ca0_jsr;
      X[ORC_TRPN * N + j] = 0;
      X[ORC_CSQN * N + j] = c_csqn;
      X[ORC_ATP * N + j] = 0;
      X[ORC_CMDN * N + j] = 0;
      X[ORC_DYE * N + j] = 0;
    }
  }
}

/* cpu/OutputFuncs.cpp:223-350; gpu/OutputFuncs.cpp adds Ca_cyt per RyR before the RyR states */
int orc_states_record(double t, const double* X, int n_ele, const double* gv, const orc_chan* ch,
                      const double* P, int ryr_out, int gpu_layout, float* rec) {
  const double* Ca = X + ORC_CA * n_ele;
  int k = 0;
  rec[k++] = (float)t;
  double i_ryr = 0;
  for (int i = 0; i < ch->n_ryr; ++i) {
    double CaJSR = Ca[ch->ryr_jsr[i]], CaSS = Ca[ch->ryr_ele[i]];
    i_ryr += ch->ryr_state[i] * P[ORC_V_RYR] * (CaJSR - CaSS) * (P[ORC_V_CELL]) * 96485 * 2 * (1e-6);
  }
  rec[k++] = (float)i_ryr;
  int n_open = 0;
  for (int i = 0; i < ch->n_ryr; ++i) n_open += ch->ryr_state[i];
  rec[k++] = (float)n_open;
  int n_open_lcc = 0;
  for (int i = 0; i < ch->n_lcc; ++i)
    if (ch->lccv_state[i] == 2 && (ch->lcc_state[i] == 6 || ch->lcc_state[i] == 12)) n_open_lcc++;
  rec[k++] = (float)n_open_lcc;
  const double F_over_RT = ORC_F / ORC_RT;
  double V = gv[0];
  if (fabs(V) < 1e-6) V = 1e-6;
  double i_lcc = 0;
  for (int i = 0; i < ch->n_lcc; ++i) {
    double c = Ca[ch->lcc_ele[i]];
    if ((ch->lcc_state[i] == 6 || ch->lcc_state[i] == 12) && ch->lccv_state[i] == 2)
      i_lcc += -ORC_P_CAL * V * (2 * F_over_RT) * (c * exp(2 * V * F_over_RT) - 0.34 * P[ORC_CA_0]) /
               (exp(2 * V * F_over_RT) - 1) * 96485 * 2 * (1e-6);
  }
  rec[k++] = (float)i_lcc;
  double m = 0;
  for (int i = 0; i < ch->n_ryr; ++i) m += Ca[ch->ryr_jsr[i]];
  m /= ch->n_ryr;
  rec[k++] = (float)m;
  m = 0;
  for (int i = 0; i < ch->n_ryr; ++i) m += Ca[ch->ryr_ele[i]];
  m /= ch->n_ryr;
  rec[k++] = (float)m;
  if (gpu_layout)
    for (int i = 0; i < ch->n_ryr; ++i) rec[k++] = (float)Ca[ch->ryr_ele[i]];
  if (ryr_out)
    for (int i = 0; i < ch->n_ryr; ++i) rec[k++] = (float)ch->ryr_state[i];
  return k;
}

/* cpu/InitFuncs.cpp:1420-1475; the reference draws r from rand() after srand(seed). */
int orc_open_random_ryr(const double* X, int n_ele, const double* P, orc_chan* ch, double u01) {
  const double* Ca = X + ORC_CA * n_ele;
  double* rates = (double*)malloc((size_t)(ch->n_ryr > 0 ? ch->n_ryr : 1) * sizeof(double));
  double total = 0;
  for (int idx = 0; idx < ch->n_ryr; ++idx) {
    int N_Open = 0, N_Closed = 0;
    for (int i = 0; i < 4; ++i) {
      int nb = ch->ryr_nb[idx * 4 + i];
      if (nb >= 0) { if (ch->ryr_state[nb]) N_Open++; else N_Closed++; }
    }
    double X_, rate, ca_ss = Ca[ch->ryr_ele[idx]];
    if (ch->ryr_state[idx]) {
      X_ = exp(0.5 * P[ORC_RYR_A_STAR] * (N_Open * P[ORC_RYR_EPS_OO] - N_Closed * P[ORC_RYR_EPS_CC]));
      rate = X_ * P[ORC_RYR_K_MINUS];
    } else {
      double phi = P[ORC_RYR_PHI_M] * Ca[ch->ryr_jsr[idx]] + P[ORC_RYR_PHI_B];
      X_ = exp(0.5 * P[ORC_RYR_A_STAR] * (N_Closed * P[ORC_RYR_EPS_CC] - N_Open * P[ORC_RYR_EPS_OO]));
      rate = X_ * phi * P[ORC_RYR_K_PLUS] * pow(ca_ss, P[ORC_RYR_ETA]);
    }
    rates[idx] = rate;
    total += rate;
  }
  double r = total * u01;
  int open = 0;
  double sum = 0;
  for (int i = 0; i < ch->n_ryr; ++i) {
    sum += rates[i];
    if (sum >= r) { open = i; break; }
  }
  if (ch->n_ryr > 0) ch->ryr_state[open] = 1;
  free(rates);
  return open;
}

/* ------------------------------------------------------------- linescan */
/* linescan/parameters.h:45-62: the reference's own value of pi. */
#define ORC_SCAN_PI 3.14159

typedef struct { double* c; double* v; double* ca; double* caf; int n, cap; } orc_cells;

static void orc_cells_push(orc_cells* m, double x, double y, double z, double v, double ca, double caf) {
  if (m->n == m->cap) {
    m->cap = m->cap ? 2 * m->cap : 1024;
    m->c = (double*)realloc(m->c, 3 * (size_t)m->cap * sizeof(double));
    m->v = (double*)realloc(m->v, (size_t)m->cap * sizeof(double));
    m->ca = (double*)realloc(m->ca, (size_t)m->cap * sizeof(double));
    m->caf = (double*)realloc(m->caf, (size_t)m->cap * sizeof(double));
  }
  m->c[3 * m->n] = x; m->c[3 * m->n + 1] = y; m->c[3 * m->n + 2] = z;
  m->v[m->n] = v; m->ca[m->n] = ca; m->caf[m->n] = caf;
  m->n++;
}

/* linescan/main.cpp:268-306 */
static void orc_add_point_grid(orc_cells* m, const orc_scan* s, double xmin, double xmax, double ymin, double ymax,
                               double zmin, double zmax) {
  double dx = s->pad_res, dy = s->pad_res, dz = s->pad_res;
  if (xmin > xmax) dx = -dx;
  if (ymin > ymax) dy = -dy;
  if (zmin > zmax) dz = -dz;
  int nx = (int)((xmax - xmin) / dx), ny = (int)((ymax - ymin) / dy), nz = (int)((zmax - zmin) / dz);
  for (int i = 0; i < nx; i++)
    for (int j = 0; j < ny; j++)
      for (int k = 0; k < nz; k++)
        orc_cells_push(m, xmin + i * dx + 0.5 * dx, ymin + j * dy + 0.5 * dy, zmin + k * dz + 0.5 * dz,
                       fabs(dx * dy * dz), s->c0, s->f0);
}

void orc_linescan_size(const orc_scan* s, int* n_f, int* n_ca) {
  *n_f = (int)((s->x_max - s->x_min) / s->res_f);
  *n_ca = (int)((s->x_max - s->x_min) / s->res_ca);
}

void orc_linescan(const orc_scan* s, int n, const double* centers, const double* V0, const double* ca,
                  const double* caf, double* out_ca, double* out_caf, double* out_cafb) {
  orc_cells m = {0};
  for (int j = 0; j < n; ++j)
    orc_cells_push(&m, centers[3 * j], centers[3 * j + 1], centers[3 * j + 2], V0[j], ca[j], caf[j]);
  /* padding, linescan/main.cpp:119-141 */
  const double pmin = s->pad_min, pmax = s->pad_max;
  const double xlo = -s->dom[0] / 2, xhi = s->dom[0] / 2, ylo = -s->dom[1] / 2, yhi = s->dom[1] / 2;
  const double zlo = -s->dom[2] / 2, zhi = s->dom[2] / 2;
  orc_add_point_grid(&m, s, pmin, pmax, pmin, pmax, zhi, pmax);
  orc_add_point_grid(&m, s, pmin, pmax, pmin, pmax, zlo, pmin);
  orc_add_point_grid(&m, s, xlo, pmin, pmin, pmax, zlo, zhi);
  orc_add_point_grid(&m, s, xhi, pmax, pmin, pmax, zlo, zhi);
  orc_add_point_grid(&m, s, xlo, xhi, ylo, pmin, zlo, zhi);
  orc_add_point_grid(&m, s, xlo, xhi, yhi, pmax, zlo, zhi);
  int n_f, n_ca;
  orc_linescan_size(s, &n_f, &n_ca);
  /* axis mapping, linescan/main.cpp:169-194 */
  double c0[3] = {0, 0, 0}, sx, sy, sz;
  if (s->axis == 0) {
    c0[1] = s->off1; c0[2] = s->off2;
    sx = sqrt(s->var[0]); sy = sqrt(s->var[1]); sz = sqrt(s->var[2]);
  } else if (s->axis == 1) {
    c0[0] = s->off1; c0[2] = s->off2;
    sx = sqrt(s->var[0]); sy = sqrt(s->var[1]); sz = sqrt(s->var[2]);
  } else {
    c0[0] = s->off1; c0[1] = s->off2;
    sx = sqrt(s->var[2]); sy = sqrt(s->var[0]); sz = sqrt(s->var[1]);
  }
  const double alpha = 1.0 / pow(2 * ORC_SCAN_PI, 1.5);
  const double r_thresh = 1;
  for (int i = 0; i < n_f; ++i) { /* :197-225 */
    c0[s->axis] = s->x_min + i * s->res_f;
    double sum = 0;
    double rx = m.c[0] - c0[0], ry = m.c[1] - c0[1], rz = m.c[2] - c0[2];
    double r_min = rx * rx + ry * ry + rz * rz;
    int cell_min = 0;
    for (int j = 0; j < m.n; ++j) {
      rx = m.c[3 * j] - c0[0]; ry = m.c[3 * j + 1] - c0[1]; rz = m.c[3 * j + 2] - c0[2];
      double dist = rx * rx + ry * ry + rz * rz;
      if (dist < r_thresh) {
        double w = alpha * exp(-0.5 * ((rx * rx / (sx * sx)) + (ry * ry / (sy * sy)) + (rz * rz / (sz * sz)))) /
                   (sx * sy * sz);
        sum += m.v[j] * m.caf[j] * w;
      }
      if (dist < r_min) { r_min = dist; cell_min = j; }
    }
    out_cafb[i] = sum;
    out_caf[i] = m.caf[cell_min];
  }
  for (int i = 0; i < n_ca; ++i) { /* :227-249 */
    c0[s->axis] = s->x_min + i * s->res_ca;
    double rx = m.c[0] - c0[0], ry = m.c[1] - c0[1], rz = m.c[2] - c0[2];
    double r_min = rx * rx + ry * ry + rz * rz;
    int cell_min = 0;
    for (int j = 0; j < m.n; ++j) {
      rx = m.c[3 * j] - c0[0]; ry = m.c[3 * j + 1] - c0[1]; rz = m.c[3 * j + 2] - c0[2];
      double dist = rx * rx + ry * ry + rz * rz;
      if (dist < r_min) { r_min = dist; cell_min = j; }
    }
    out_ca[i] = m.ca[cell_min];
  }
  free(m.c); free(m.v); free(m.ca); free(m.caf);
}
