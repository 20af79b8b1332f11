// adjoint.hip -- schedule recording, replayed segments and the adjoint of
// the frozen-channel step for line-scan fitting (single realizations and
// batches).  No reference counterpart.
#include "dev_common.h"

#include <functional>

namespace srs_dev {

// Reverse mode of the explicit step with the channel trajectory frozen
// (RyR states from a recorded schedule, LCC states held): for
// X_{n+1} = X_n + dt r(X_n; theta), lambda_n = lambda_{n+1} + dt J^T lambda_{n+1}
// and dL/dtheta += dt (dr/dtheta)^T lambda_{n+1}.  The diffusion operator
// satisfies V_i A_ij = V_j A_ji (both equal the face coefficient), so
// (A^T y)_j = V_j (A (y / V))_j: the transpose is another gather over the same
// rows.  Single-realization layout (quads, then the TnC and CSQ planes); a
// launch covers the realizations of a table, one workgroup row (blockIdx.y)
// each, with row y's adjoint, state and weight arrays at y times their size.
constexpr int ADJ_NT = 13;   // D_CA, D_DYE, kon/koff of CaM, TnC, ATP, Fluo, V_RYR, A_P, K_D_SL_HI
constexpr int ADJ_BS = 256;
// Reverse-step launches (tiles x members) below which w is written by each
// step for the next instead of formed in the kernel (SRS_TUNE_ADJ_W 0).
constexpr size_t kAdjWrittenWBlocks = 4096;
__constant__ const int kAdjTheta[ADJ_NT] = {SRS_R_D_CA,     SRS_R_D_DYE,     SRS_R_K_ON_CMDN, SRS_R_K_OFF_CMDN,
                                            SRS_R_K_ON_TRPN, SRS_R_K_OFF_TRPN, SRS_R_K_ON_ATP, SRS_R_K_OFF_ATP,
                                            SRS_R_K_ON_DYE, SRS_R_K_OFF_DYE, SRS_R_V_RYR,     SRS_R_A_P,
                                            SRS_R_K_D_SL_HI};

// 1 / beta of the rapid SL buffer at element j and its derivatives in c and
// in K_D_SL_HI (cpu/RunSimulation.cpp:15-22).
struct SlBuf { double beta, dbeta_dc, dbeta_dkh; };
__device__ __forceinline__ SlBuf sl_buffer(const RealDesc& d, int j, double c, uint8_t f) {
  SlBuf s{1.0, 0.0, 0.0};
  if (!(f & F_SL)) return s;
  const double* P = d.P;
  const double ttj = d.tt[j], v0j = d.V0[j];
  const double bhi = ttj * P[SRS_R_B_TOT_SL_HI] / (v0j * (1e-15));
  const double blo = ttj * P[SRS_R_B_TOT_SL_LO] / (v0j * (1e-15));
  const double kh = P[SRS_R_K_D_SL_HI], kl = P[SRS_R_K_D_SL_LO];
  const double eh = kh + c, el = kl + c;
  s.beta = 1 + (bhi * kh / (eh * eh)) + (blo * kl / (el * el));
  s.dbeta_dc = -2 * bhi * kh / (eh * eh * eh) - 2 * blo * kl / (el * el * el);
  s.dbeta_dkh = bhi * (c - kh) / (eh * eh * eh);
  return s;
}

// The loss source of one scan sample at element j: sum_i gcoef_i w(c_j - p_i)
// over points within 1 um (the blur's PSF, linescan/parameters.h:60-62), with
// the blur's own points and weights (the cell's run of points within the
// cut-off).  Every lane of the wavefront calls it (valid false: no element,
// sum zero).  The loop runs over the points any lane can reach, the same for
// the whole wavefront, so gcoef_i is one uniform load per point; each lane's
// weights walk from anchors at multiples of the anchor spacing, so a lane's
// weight at a point of its own run is the one its own walk would give (a
// weight outside the run is never used, so it may be anything).
__device__ __forceinline__ double scan_adj_sum(const double* __restrict__ cen, const ScanGeo& g, int n_f,
                                               const double* __restrict__ gcoef, size_t np, int j, bool valid) {
  int i_lo = n_f, i_hi = -1;
  ScanCell c{0.0, 0.0, 0.0, 0.0};
  if (valid) {
    const double cx = cen[j], cy = cen[np + j], cz = cen[2 * np + j];
    scan_point_range(g, n_f, cx, cy, cz, &i_lo, &i_hi);
    c = scan_cell(g, cx, cy, cz);
    scan_exact_range(g, c, &i_lo, &i_hi);
  }
  int lo = i_lo <= i_hi ? i_lo : n_f, hi = i_lo <= i_hi ? i_hi : -1;
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  hi = __builtin_amdgcn_readfirstlane(hi);
  double s = 0.0;
  if (lo > hi) return s;
  const int i0 = __builtin_amdgcn_readfirstlane(scan_walk_start(g, lo));
  double w = 0.0, r = 0.0;
  for (int i = i0; i <= hi; ++i) {
    const double pa = scan_pa(g, i);
    if (scan_is_anchor(g, i, i0)) scan_anchor(g, c, pa, &w, &r);
    const double gi = gcoef[i];
    if (i >= i_lo && i <= i_hi) s += gi * w;
    w *= r;
    r *= g.q;
  }
  return s;
}

// w = (lambda_Ca / beta, lambda_ATP, lambda_CaM, lambda_Fluo) / V0 of element j
// (c, the element's Ca, is read only where the SL buffer makes beta depend on
// it; callers pass 0 elsewhere and skip the load).
__device__ __forceinline__ d4 adj_weight(const RealDesc& d, int j, double c, uint8_t f, double la_c, double la_a,
                                         double la_m, double la_d) {
  const SlBuf sb = sl_buffer(d, j, c, f);
  const double iv = 1.0 / d.V0[j];
  return d4{la_c / sb.beta * iv, la_a * iv, la_m * iv, la_d * iv};
}

// Written w (small launches): w of the first step of a recomputed segment;
// the step kernel writes each later one for the step before.
__global__ __launch_bounds__(ADJ_BS) void k_adj_w(const RealDesc* __restrict__ D, const int32_t* __restrict__ rid,
                                                  const double* __restrict__ X, const double* __restrict__ lam,
                                                  d4* __restrict__ w) {
  const RealDesc& d = D[rid[blockIdx.y]];
  const int j = blockIdx.x * ADJ_BS + threadIdx.x;
  if (j >= d.n) return;
  X += blockIdx.y * 6 * (size_t)d.np;
  lam += blockIdx.y * 6 * (size_t)d.np;
  w += blockIdx.y * (size_t)d.np;
  const d4 m = ((const d4*)lam)[j];
  const uint8_t f = d.flags[j];
  w[j] = adj_weight(d, j, (f & F_SL) ? X[4 * (size_t)j] : 0.0, f, m.x, m.y, m.z, m.w);
}

// What the step kernel does after lambda_n: add the loss source of a scan
// sample taken after n steps (SCAN), and, with written w, form w for step
// n - 1 from lambda_n and that step's state (Xprev, when it is in the
// recomputed segment).
struct AdjNext {
  const double* Xprev;  // state of step n - 1, or null
  d4* wout;             // its w
  const double* cen;    // scan cell centres [3][np]
  const double* gcoef;  // the sample's loss coefficients per point
  ScanGeo g;
  int n_f, gstride;
};

// lambda_n and the theta-gradient partials of one step.  The transposed
// diffusion gathers w = (lambda_Ca / beta, lambda_ATP, lambda_CaM,
// lambda_Fluo) / V0 of lambda_{n+1} and this step's state.  WMAT: w was
// written by the step before (or k_adj_w) and is gathered as it is, the
// shortest dependency chain for a launch too small to fill the GPU.  Else
// each element forms its own w and stages it in LDS for the tile, and a
// neighbour outside the tile forms it from its lambda: no w buffer is read
// or written.  The operations are the same, so are the bits.
template <bool COMPACT, bool SCAN, bool WMAT>
__device__ __forceinline__ void adj_step(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                         const int32_t* __restrict__ rid, const double* __restrict__ X,
                                         const double* __restrict__ lam, const d4* __restrict__ w, int64_t step,
                                         double* __restrict__ lam_out, double* __restrict__ gpart, AdjNext nx) {
  __shared__ double red[ADJ_BS / 16][ADJ_NT];  // gradient partials per row of 16 lanes
  __shared__ d4 tw[ADJ_BS];  // the tile's own w: neighbours inside the tile read it from here
  const int r = rid[blockIdx.y];
  const RealDesc& d = D[r];
  const RealLive& l = L[r];
  // Tiles in grid order: an XCD-aware order (consecutive tiles on one XCD)
  // cuts this kernel's L2 misses by 40% but measured 6% slower.
  const int tile = blockIdx.x;
  const int j = tile * ADJ_BS + threadIdx.x;
  const size_t np = d.np;
  X += blockIdx.y * 6 * np;
  lam += blockIdx.y * 6 * np;
  lam_out += blockIdx.y * 6 * np;
  if (WMAT) w += blockIdx.y * np;
  gpart += (size_t)blockIdx.y * gridDim.x * ADJ_NT;
  // the tile's running partials, loaded now so the final update waits on no load
  const double gp_old = threadIdx.x < ADJ_NT ? gpart[(size_t)tile * ADJ_NT + threadIdx.x] : 0.0;
  // RyR states during this step: row `step` of the frozen schedule
  const int32_t* ryr = d.n_ryr ? d.sched + step * d.n_ryr : nullptr;
  double gr[ADJ_NT];
#pragma unroll
  for (int k = 0; k < ADJ_NT; ++k) gr[k] = 0.0;
  // the element's operator row, state, lambda_{n+1} and w; with w formed
  // here, the tile's w staged in LDS
  const bool in = j < d.n;
  const d4* Q = (const d4*)X;
  const d4* Lm = (const d4*)lam;
  Geo g{};
  d4 m0 = {0.0, 0.0, 0.0, 0.0}, mu = {0.0, 0.0, 0.0, 0.0}, w0 = {0.0, 0.0, 0.0, 0.0};
  double V0 = 1.0;
  SlBuf sb{1.0, 0.0, 0.0};
  if (in) {
    g = load_geo<false, COMPACT, -1>(d, j);
    m0 = Q[j];
    mu = Lm[j];
    V0 = d.V0[j];
    sb = sl_buffer(d, j, m0.x, g.f);
    if constexpr (WMAT) {
      w0 = w[j];
    } else {
      const double iv = 1.0 / V0;
      w0 = d4{mu.x / sb.beta * iv, mu.y * iv, mu.z * iv, mu.w * iv};
    }
  }
  if constexpr (!WMAT) {
    tw[threadIdx.x] = w0;
    __syncthreads();
  }
  // the scan sample's loss source (all lanes of the wavefront take part)
  const double src =
      SCAN ? scan_adj_sum(nx.cen, nx.g, nx.n_f, nx.gcoef + (size_t)blockIdx.y * nx.gstride, np, j, in) : 0.0;
  if (in) {
    const double* P = d.P;
    const double dt = P[SRS_R_T_STEP];
    const uint8_t f = g.f;
    const int base = tile * ADJ_BS;
    auto nw = [&](int nb) -> d4 {
      if constexpr (WMAT) {
        return w[nb];
      } else {
        const unsigned k = (unsigned)(nb - base);
        if (k < (unsigned)ADJ_BS) return tw[k];
        const uint8_t fn = d.flags[nb];
        const d4 ln = Lm[nb];
        return adj_weight(d, nb, (fn & F_SL) ? Q[nb].x : 0.0, fn, ln.x, ln.y, ln.z, ln.w);
      }
    };
    const d4 w1 = nw(g.n1), w2 = nw(g.n2), w3 = nw(g.n3), w4 = nw(g.n4);
    const bool live_tr = (f & (F_JSR | F_TROPC)) == F_TROPC;
    const bool live_cs = (f & F_JSR) != 0;
    const double bT = live_tr ? X[4 * np + j] : 0.0, bC = live_cs ? X[5 * np + j] : 0.0;
    // the TnC and CSQ adjoints change only where the buffer lives (troponin
    // cytosol, jSR); both lambda buffers start at zero, so elsewhere they stay
    // zero and those planes are neither read nor written
    const double muT = live_tr ? lam[4 * np + j] : 0.0, muC = live_cs ? lam[5 * np + j] : 0.0;
    const double c = m0.x, bA = m0.y, bM = m0.z, bD = m0.w;
    const double u = mu.x / sb.beta;

    // transposed diffusion: dt D_s V0 (A w_s)_j
    const double aw_c = g.a0 * w0.x + g.a1 * w1.x + g.a2 * w2.x + g.a3 * w3.x + g.a4 * w4.x;
    const double aw_a = g.a0 * w0.y + g.a1 * w1.y + g.a2 * w2.y + g.a3 * w3.y + g.a4 * w4.y;
    const double aw_m = g.a0 * w0.z + g.a1 * w1.z + g.a2 * w2.z + g.a3 * w3.z + g.a4 * w4.z;
    const double aw_d = g.a0 * w0.w + g.a1 * w1.w + g.a2 * w2.w + g.a3 * w3.w + g.a4 * w4.w;
    double lc = mu.x + dt * P[SRS_R_D_CA] * V0 * aw_c;
    double la = mu.y + dt * P[SRS_R_D_ATP] * V0 * aw_a;
    double lm = mu.z + dt * P[SRS_R_D_CMDN] * V0 * aw_m;
    double ld = mu.w + dt * P[SRS_R_D_DYE] * V0 * aw_d;
    double lT = muT, lC = muC;

    // The D gradients sum u_j (A c)_j and mu_Fluo,j (A b_Fluo)_j over the mesh;
    // V_i A_ij = V_j A_ji turns those sums into sum_j c_j V0_j (A w_Ca)_j and
    // sum_j b_Fluo,j V0_j (A w_Fluo)_j, terms of the element alone.  So the
    // neighbours' states are gathered only for an SL element, whose SL terms
    // need its full Ca rate before the division (rca; unused elsewhere).
    double ac = 0.0;
    if (f & F_SL)
      ac = c * g.a0 + Q[g.n1].x * g.a1 + Q[g.n2].x * g.a2 + Q[g.n3].x * g.a3 + Q[g.n4].x * g.a4;
    // boundary rates are D times a D-free plane: they enter the D gradients
    const double bu_c = (f & F_BR) ? br_unit(d.br_mesh[j], d.brd[0]) : 0.0;
    const double bu_d = (f & F_BR) ? br_unit(d.br_mesh[np + j], d.brd[1]) : 0.0;
    double rca = P[SRS_R_D_CA] * ac + ((f & F_BR) ? d.br[j] : 0.0);
    gr[0] += dt * (c * V0 * aw_c + u * bu_c);       // D_CA
    gr[1] += dt * (bD * V0 * aw_d + mu.w * bu_d);   // D_DYE
    if (!(f & F_JSR)) {
      if (f & F_TROPC) {
        const double kon = P[SRS_R_K_ON_TRPN], koff = P[SRS_R_K_OFF_TRPN], B = P[SRS_R_B_TOT_TRPN];
        const double r = kon * (B - bT) * c - koff * bT;
        rca -= r;
        lc += dt * (muT - u) * kon * (B - bT);
        lT += dt * (muT - u) * (-kon * c - koff);
        gr[4] += dt * (muT - u) * (B - bT) * c;
        gr[5] += dt * (muT - u) * (-bT);
        // SERCA, cpu/RunSimulation.cpp:82-90
        const double kdi = P[SRS_R_K_D_I];
        double K = c / kdi;
        K = K * K;
        const double dK = 2.0 * (c / kdi) / kdi;
        const double Nn = K * K * 3.24873e12 + K * l.serca[3] - l.serca[4];
        const double Dn = l.serca[0] + K * l.serca[1] + K * K * l.serca[2];
        const double v = Nn / Dn;
        const double dv = ((2 * K * 3.24873e12 + l.serca[3]) * Dn - Nn * (l.serca[1] + 2 * K * l.serca[2])) / (Dn * Dn);
        rca -= (2e-3) * v * P[SRS_R_A_P];
        lc += dt * (-u) * (2e-3) * P[SRS_R_A_P] * dv * dK;
        gr[11] += dt * u * (-(2e-3) * v);
      }
      const double konD = P[SRS_R_K_ON_DYE], koffD = P[SRS_R_K_OFF_DYE], BD = P[SRS_R_B_TOT_DYE];
      const double rD = konD * (BD - bD) * c - koffD * bD;
      rca -= rD;
      lc += dt * (mu.w - u) * konD * (BD - bD);
      ld += dt * (mu.w - u) * (-konD * c - koffD);
      gr[8] += dt * (mu.w - u) * (BD - bD) * c;
      gr[9] += dt * (mu.w - u) * (-bD);
      const double konA = P[SRS_R_K_ON_ATP], koffA = P[SRS_R_K_OFF_ATP], BA = P[SRS_R_B_TOT_ATP];
      const double rA = konA * (BA - bA) * c - koffA * bA;
      rca -= rA;
      lc += dt * (mu.y - u) * konA * (BA - bA);
      la += dt * (mu.y - u) * (-konA * c - koffA);
      gr[6] += dt * (mu.y - u) * (BA - bA) * c;
      gr[7] += dt * (mu.y - u) * (-bA);
      const double konM = P[SRS_R_K_ON_CMDN], koffM = P[SRS_R_K_OFF_CMDN], BM = P[SRS_R_B_TOT_CMDN];
      const double rM = konM * (BM - bM) * c - koffM * bM;
      rca -= rM;
      lc += dt * (mu.z - u) * konM * (BM - bM);
      lm += dt * (mu.z - u) * (-konM * c - koffM);
      gr[2] += dt * (mu.z - u) * (BM - bM) * c;
      gr[3] += dt * (mu.z - u) * (-bM);
    } else {
      const double kon = P[SRS_R_K_ON_CSQN], koff = P[SRS_R_K_OFF_CSQN], B = P[SRS_R_B_T_JSR];
      const double r = kon * (B - bC) * c - koff * bC;
      rca -= r;
      rca += (l.ca_nsr - c) * P[SRS_R_V_REFILL];
      lc += dt * ((muC - u) * kon * (B - bC) - u * P[SRS_R_V_REFILL]);
      lC += dt * (muC - u) * (-kon * c - koff);
    }
    if (f & F_CHAN) {
      int lo = 0, hi = d.n_celem;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (d.celem[mid] < j) lo = mid + 1; else hi = mid;
      }
      if (lo < d.n_celem && d.celem[lo] == j) {
        const double a = P[SRS_R_V_RYR] * P[SRS_R_V_CELL];
        for (int e = d.cstart[lo]; e < d.cstart[lo + 1]; ++e) {
          const int ch = d.centry[e];
          if (ch >= 0) {
            const int i = ch >> 1;
            if (!ryr[i]) continue;
            const int ce = d.ryr_ele[i], je = d.ryr_jsr[i];
            const int p = (ch & 1) ? ce : je;  // the partner element
            const double cp = X[4 * (size_t)p];
            const double up = lam[4 * (size_t)p] / sl_buffer(d, p, cp, d.flags[p]).beta;
            const double dc = X[4 * (size_t)je] - X[4 * (size_t)ce];
            const double J = a * dc;
            if (ch & 1) {  // jSR side: r -= J / V0_jsr
              rca -= J / V0;
              lc += dt * (u * (-a / V0) + up * (a / d.V0[p]));
            } else {       // cleft side: r += J / V0_cyt
              rca += J / V0;
              lc += dt * (u * (-a / V0) + up * (a / d.V0[p]));
              gr[10] += dt * (u * P[SRS_R_V_CELL] * dc / V0 - up * P[SRS_R_V_CELL] * dc / d.V0[p]);
            }
          } else {
            const int i = -ch - 1;
            if (srs_lcc_open(d.lcc_state[i], d.lccv_state[i])) {
              rca += srs_lcc_flux_raw(l.vm, c, P[SRS_R_CA_0]) / V0;
              lc += dt * u * srs_lcc_flux_raw(l.vm, 1.0, 0.0) / V0;
            }
          }
        }
      }
    }
    if (f & F_SL) {
      const double b2 = sb.beta * sb.beta;
      lc += dt * mu.x * (-rca * sb.dbeta_dc / b2);
      gr[12] += dt * mu.x * (-rca * sb.dbeta_dkh / b2);
    }
    if (SCAN) ld += V0 * src;
    ((d4*)lam_out)[j] = d4{lc, la, lm, ld};
    if (live_tr) lam_out[4 * np + j] = lT;
    if (live_cs) lam_out[5 * np + j] = lC;
    if (WMAT && nx.Xprev) {
      const double* Xp = nx.Xprev + blockIdx.y * 6 * np;
      nx.wout[blockIdx.y * np + j] = adj_weight(d, j, (f & F_SL) ? Xp[4 * (size_t)j] : 0.0, f, lc, la, lm, ld);
    }
  }
  // block reduction of the gradient partials, accumulated per block: DPP
  // sums within each row of 16 lanes (a component that is zero in every lane
  // of the wavefront, as the channel and SL terms are in most, is not
  // summed), then the block's rows in LDS in a fixed order
#pragma unroll
  for (int k = 0; k < ADJ_NT; ++k) {
    const double v = __builtin_amdgcn_ballot_w64(gr[k] != 0.0) ? row_sum_dpp(gr[k]) : 0.0;
    if ((threadIdx.x & 15) == 15) red[threadIdx.x >> 4][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < ADJ_NT) {
    double s = 0.0;
    for (int q = 0; q < ADJ_BS / 16; ++q) s += red[q][threadIdx.x];
    gpart[(size_t)tile * ADJ_NT + threadIdx.x] = gp_old + s;
  }
}

// The compiler's register budget (126 VGPRs: four wavefronts per SIMD), or
// at most 128 (four) asked for explicitly; the two compile alike.
template <bool COMPACT, bool SCAN, bool WMAT>
__global__ __launch_bounds__(ADJ_BS) void k_adj_step(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                                     const int32_t* __restrict__ rid, const double* __restrict__ X,
                                                     const double* __restrict__ lam, const d4* __restrict__ w,
                                                     int64_t step, double* __restrict__ lam_out,
                                                     double* __restrict__ gpart, AdjNext nx) {
  adj_step<COMPACT, SCAN, WMAT>(D, L, rid, X, lam, w, step, lam_out, gpart, nx);
}
template <bool COMPACT, bool SCAN, bool WMAT>
__global__ __launch_bounds__(ADJ_BS) __attribute__((amdgpu_waves_per_eu(4))) void k_adj_step_w4(
    const RealDesc* __restrict__ D, const RealLive* __restrict__ L, const int32_t* __restrict__ rid,
    const double* __restrict__ X, const double* __restrict__ lam, const d4* __restrict__ w, int64_t step,
    double* __restrict__ lam_out, double* __restrict__ gpart, AdjNext nx) {
  adj_step<COMPACT, SCAN, WMAT>(D, L, rid, X, lam, w, step, lam_out, gpart, nx);
}

using AdjStepFn = void (*)(const RealDesc*, const RealLive*, const int32_t*, const double*, const double*,
                           const d4*, int64_t, double*, double*, AdjNext);
template <bool COMPACT, bool SCAN, bool WMAT>
static AdjStepFn adj_step_fn(bool waves4) {
  return waves4 ? k_adj_step_w4<COMPACT, SCAN, WMAT> : k_adj_step<COMPACT, SCAN, WMAT>;
}
static AdjStepFn adj_step_kernel(bool compact, bool scan, bool wmat, bool waves4) {
  if (compact) {
    if (scan) return wmat ? adj_step_fn<true, true, true>(waves4) : adj_step_fn<true, true, false>(waves4);
    return wmat ? adj_step_fn<true, false, true>(waves4) : adj_step_fn<true, false, false>(waves4);
  }
  if (scan) return wmat ? adj_step_fn<false, true, true>(waves4) : adj_step_fn<false, true, false>(waves4);
  return wmat ? adj_step_fn<false, false, true>(waves4) : adj_step_fn<false, false, false>(waves4);
}

// The loss gradient of one scan sample in the Fluo adjoint:
// lambda_Fluo,j += V0_j sum_i gcoef_i w(c_j - p_i) over points within 1 um.
__global__ __launch_bounds__(ADJ_BS) void k_scan_adj(const RealDesc* __restrict__ D, const int32_t* __restrict__ rid,
                                                     const double* __restrict__ cen, ScanGeo g, int n_f,
                                                     const double* __restrict__ gcoef, int gstride,
                                                     double* __restrict__ lam) {
  const RealDesc& d = D[rid[blockIdx.y]];
  const int j = blockIdx.x * ADJ_BS + threadIdx.x;
  const size_t np = d.np;
  const double s = scan_adj_sum(cen, g, n_f, gcoef + (size_t)blockIdx.y * gstride, np, j, j < d.n);
  if (j >= d.n) return;
  lam += blockIdx.y * 6 * np;
  lam[4 * (size_t)j + 3] += d.V0[j] * s;
}

// Copies the current state (X[cur]) of each realization of a table to slot y
// of a contiguous buffer (to_buf) or back.
__global__ void k_batch_state(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                              const int32_t* __restrict__ rid, double* __restrict__ buf, int to_buf) {
  const int r = rid[blockIdx.y];
  const RealDesc& d = D[r];
  const size_t n6 = 6 * (size_t)d.np;
  double* X = d.X[L[r].cur];
  double* B = buf + blockIdx.y * n6;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n6; i += (size_t)gridDim.x * blockDim.x) {
    if (to_buf) B[i] = X[i];
    else X[i] = B[i];
  }
}

// Points the ping-pong buffers of each realization y of a table at slots of a
// recomputed segment (slot i of y at seg + (i * nb + y) * 6 np): the current
// state at slot step - s0, the next one at the slot after, so the sweep writes
// the segment in place.
__global__ void k_seg_ptrs(RealDesc* __restrict__ D, const RealLive* __restrict__ L, const int32_t* __restrict__ rid,
                           double* seg, int64_t s0, int nb) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= nb) return;
  const int r = rid[y];
  const int cur = L[r].cur;
  const size_t n6 = 6 * (size_t)D[r].np;
  const size_t i = (size_t)(L[r].step - s0);
  D[r].X[cur] = seg + (i * nb + y) * n6;
  D[r].X[cur ^ 1] = seg + ((i + 1) * nb + y) * n6;
}

// Starts a replayed segment at `step` for each realization of a table: the
// step counter, schedule rows counted from zero, RyR states of row `step`.
__global__ void k_seg_start(const RealDesc* __restrict__ D, RealLive* __restrict__ L, const int32_t* __restrict__ rid,
                            int64_t step) {
  const int r = rid[blockIdx.x];
  const RealDesc& d = D[r];
  for (int i = threadIdx.x; i < d.n_ryr; i += blockDim.x) d.ryr_state[i] = d.sched[step * d.n_ryr + i];
  if (threadIdx.x == 0) {
    L[r].step = step;
    L[r].sched_base = 0;
  }
}

// Writes each step's RyR states (the ones the next step's fluxes use) into a
// schedule, for srs_record_schedule.
__global__ void k_sched_out(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                            const int64_t* __restrict__ base, const size_t* __restrict__ off, int32_t* out,
                            int64_t n_rows) {
  const int r = blockIdx.x;
  const RealDesc& d = D[r];
  const int64_t row = L[r].step - base[r];  // k_gate already advanced the step counter
  if (row < 1 || row >= n_rows) return;
  out += off[r];
  for (int i = threadIdx.x; i < d.n_ryr; i += blockDim.x) out[row * d.n_ryr + i] = d.ryr_state[i];
}


}  // namespace srs_dev

extern "C" {

int srs_adjoint_theta(int32_t* idx) {
  srs::clear_error();
  if (!idx) return srs::fail(SRS_E_ARG, "null output");
  const int t[ADJ_NT] = {SRS_R_D_CA,     SRS_R_D_DYE,     SRS_R_K_ON_CMDN, SRS_R_K_OFF_CMDN, SRS_R_K_ON_TRPN,
                         SRS_R_K_OFF_TRPN, SRS_R_K_ON_ATP, SRS_R_K_OFF_ATP, SRS_R_K_ON_DYE,  SRS_R_K_OFF_DYE,
                         SRS_R_V_RYR,     SRS_R_A_P,       SRS_R_K_D_SL_HI};
  for (int k = 0; k < ADJ_NT; ++k) idx[k] = t[k];
  return SRS_N_THETA == ADJ_NT ? SRS_OK : srs::fail(SRS_E_STATE, "theta count mismatch");
}

int srs_params_set(srs_ctx* c, int32_t r, const double* reals) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (!reals) return srs::fail(SRS_E_ARG, "null parameters");
  HostReal& hr = c->reals[r];
  memcpy(hr.reals, reals, sizeof(hr.reals));
  memcpy(hr.desc.P, reals, sizeof(hr.desc.P));
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  const HostMesh& hm = *c->meshes[hr.mesh];
  if (!hm.br.empty()) {
    // boundary rates are D Face_C b / V0: at a D other than the one the
    // mesh's rates were built with, the realization steps with rates of its own
    const double D[4] = {reals[SRS_R_D_CA], reals[SRS_R_D_DYE], reals[SRS_R_D_ATP], reals[SRS_R_D_CMDN]};
    if (std::equal(D, D + 4, hm.br_d)) {
      hr.desc.br = hm.d_br;
    } else {
      if (!hr.d_br_own && (rc = dev_alloc(c, &hr.d_br_own, hm.br.size()))) return rc;
      std::vector<double> own(hm.br.size());
      const size_t np = hm.np;
      for (int q = 0; q < 4; ++q)
        for (size_t k = 0; k < np; ++k) own[q * np + k] = br_scaled(hm.br[q * np + k], D[q], hm.br_d[q]);
      HIPCHK(hipMemcpy(hr.d_br_own, own.data(), own.size() * sizeof(double), hipMemcpyHostToDevice));
      hr.desc.br = hr.d_br_own;
    }
  }
  HIPCHK(hipMemcpy(c->d_desc + r, &hr.desc, sizeof(hr.desc), hipMemcpyHostToDevice));
  RealLive lv;
  if ((rc = live_rw(c, r, &lv, false))) return rc;
  serca_terms(lv.ca_nsr, hr.reals, lv.serca);
  return live_rw(c, r, &lv, true);
}

static int single_real(srs_ctx* c, int32_t r) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (c->reals.size() != 1 || c->group != 1)
    return srs::fail(SRS_E_ARG, "this call needs a context with one realization and group size 1");
  return SRS_OK;
}

// n_steps gated steps of the whole context, recording every realization's RyR
// states per step: realization r's n_steps x n_ryr rows at out + off[r].
static int record_run(srs_ctx* c, int64_t n_steps, int32_t* out_host) {
  const int n_real = (int)c->reals.size();
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<RealLive> lv(n_real);
  HIPCHK(hipMemcpy(lv.data(), c->d_live, n_real * sizeof(RealLive), hipMemcpyDeviceToHost));
  std::vector<int64_t> base(n_real);
  std::vector<size_t> off(n_real + 1, 0);
  for (int r = 0; r < n_real; ++r) {
    if (lv[r].done) return srs::fail(SRS_E_ARG, "record schedule: realization %d has finished its protocol", r);
    base[r] = lv[r].step;
    off[r + 1] = off[r] + (size_t)n_steps * c->reals[r].ryr_ele.size();
  }
  DevMem M;
  int32_t* d_out = nullptr;
  int64_t* d_base = nullptr;
  size_t* d_off = nullptr;
  HIPCHK(M.alloc(&d_out, off[n_real]));
  HIPCHK(M.alloc(&d_base, (size_t)n_real));
  HIPCHK(M.alloc(&d_off, (size_t)n_real));
  HIPCHK(hipMemcpy(d_base, base.data(), n_real * sizeof(int64_t), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d_off, off.data(), n_real * sizeof(size_t), hipMemcpyHostToDevice));
  for (int r = 0; r < n_real; ++r)
    if (off[r + 1] > off[r])
      HIPCHK(hipMemcpy(d_out + off[r], c->reals[r].desc.ryr_state, c->reals[r].ryr_ele.size() * sizeof(int32_t),
                       hipMemcpyDeviceToDevice));
  hipError_t e = hipSuccess;
  for (int64_t k = 0; k < n_steps && e == hipSuccess; ++k) {
    launch_step(c, G_AFTER_SWEEP);
    hipLaunchKernelGGL(k_sched_out, dim3(n_real), dim3(256), 0, c->stream, c->d_desc, c->d_live, d_base, d_off,
                       d_out, n_steps);
    e = hipGetLastError();
    if (e == hipSuccess && (k & 1023) == 1023) e = hipStreamSynchronize(c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(out_host, d_out, off[n_real] * sizeof(int32_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return srs::fail(SRS_E_HIP, "record schedule: %s", hipGetErrorString(e));
  return SRS_OK;
}

int srs_record_schedule(srs_ctx* c, int32_t r, int64_t n_steps, int32_t* schedule) {
  int rc = single_real(c, r);
  if (rc) return rc;
  if (n_steps <= 0) return SRS_OK;
  if (c->reals[r].ryr_ele.empty()) return srs::fail(SRS_E_ARG, "realization has no RyRs");
  if (!schedule) return srs::fail(SRS_E_ARG, "null schedule");
  return record_run(c, n_steps, schedule);
}

int srs_record_schedules(srs_ctx* c, int64_t n_steps, int32_t* schedules) {
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (!c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (n_steps <= 0) return SRS_OK;
  if (!schedules) return srs::fail(SRS_E_ARG, "null schedules");
  return record_run(c, n_steps, schedules);
}

// Checkpoint plan of the adjoint: L levels of stored states with strides
// K[0] > K[1] > ... > K[L-1], the last being the length of a recomputed
// segment whose every state is kept.  The forward run keeps the state every
// K[0] steps; going backward, each level-l segment is recomputed from its
// checkpoint keeping the states every K[l+1] steps, down to the segments of
// the last level.  A member holds slots[0] + ... + slots[L-1] + K[L-1]
// states, about (L + 1) n^(1/(L+1)) for equal ratios, and every level below
// the first costs one more forward pass.  Each recomputation replays the same
// kernels from the same state, so the plan changes memory and time, never a
// bit of the losses or gradients.
constexpr int ADJ_MAX_LEVELS = 6;
struct AdjPlan {
  int L = 1;
  int64_t K[ADJ_MAX_LEVELS] = {};
  int64_t slots[ADJ_MAX_LEVELS] = {};
  int64_t kept() const {
    int64_t s = 0;
    for (int l = 0; l < L; ++l) s += slots[l];
    return s;
  }
};
// leaf: the recomputed segment's length, 0 = (n_steps)^(1/(L+1)) rounded up.
static AdjPlan adj_plan(int64_t n_steps, int64_t leaf, int L) {
  AdjPlan p;
  p.L = L;
  const double n = (double)std::max<int64_t>(n_steps, 1);
  if (leaf <= 0) leaf = (int64_t)std::ceil(std::pow(n, 1.0 / (L + 1)) - 1e-9);
  leaf = std::max<int64_t>(1, std::min<int64_t>(leaf, std::max<int64_t>(n_steps, 1)));
  p.K[L - 1] = leaf;
  const int64_t r = L > 1 ? std::max<int64_t>(2, (int64_t)std::ceil(std::pow(n / leaf, 1.0 / L) - 1e-9)) : 0;
  for (int l = L - 2; l >= 0; --l) p.K[l] = p.K[l + 1] * r;
  p.slots[0] = (std::max<int64_t>(n_steps, 1) + p.K[0] - 1) / p.K[0];
  for (int l = 1; l < L; ++l) p.slots[l] = (p.K[l - 1] + p.K[l] - 1) / p.K[l];
  return p;
}

// Line-scan loss and adjoint gradient of the realizations `ids` (all on the
// scan's mesh, group size 1; arguments already checked).  Members run in
// sub-batches of at most `max_batch` that step together; every other
// realization is held (done) meanwhile, and all are left as they were.
// Schedules are each member's n_steps x n_ryr rows, concatenated in member
// order; targets are [member][sample][point].
static int loss_grad_run(srs_ctx* c, const std::vector<int32_t>& ids, int32_t scan_id, int64_t n_steps,
                         const int32_t* schedules, int32_t n_samples, const int64_t* sample_steps,
                         const double* targets, int32_t ckpt, int32_t max_batch, double* losses, double* grads,
                         double* scans) {
  const HostScan& S = c->scans[scan_id];
  const HostMesh& hm = *c->meshes[S.mesh];
  const size_t np = hm.np, n6 = 6 * np;
  const int nf = S.n_f, nca = S.n_ca, nout = nca + 2 * nf;
  const int n_mem = (int)ids.size(), n_real = (int)c->reals.size();
  const bool want_grad = grads != nullptr;
  const int nblk = (hm.n + ADJ_BS - 1) / ADJ_BS;
  std::vector<size_t> soff(n_mem + 1, 0);
  size_t max_sched = 0;
  for (int m = 0; m < n_mem; ++m) {
    const size_t k = (size_t)n_steps * c->reals[ids[m]].ryr_ele.size();
    soff[m + 1] = soff[m] + k;
    max_sched = std::max(max_sched, k);
  }
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  size_t mem_free = 0, mem_total = 0;
  HIPCHK(hipMemGetInfo(&mem_free, &mem_total));
  const size_t budget = mem_free / 10 * 9;
  // written w (two buffers) on launches of fewer than kAdjWrittenWBlocks
  // workgroups, unless SRS_TUNE_ADJ_W forces one form
  const bool wmat_possible = want_grad && (c->adj_w == 1 || (c->adj_w == 0 && (size_t)nblk < kAdjWrittenWBlocks));
  // device bytes per member of a sub-batch under a plan (a loss alone keeps
  // only the start state)
  auto per_member = [&](const AdjPlan& p) {
    return 8 * n6 * (size_t)(want_grad ? p.kept() + p.K[p.L - 1] + 2 : 1) + (wmat_possible ? 64 * np : 0) +
           8 * (size_t)n_samples * nout + 8 * (size_t)S.n_blocks * nf + 8 * (size_t)nblk * ADJ_NT +
           8 * (size_t)n_samples * nf + 4 * max_sched + 64;
  };
  // the plan: the tuned number of levels, or the fewest levels (least
  // recomputation) with which the wanted sub-batch fits
  const int want = max_batch > 0 ? std::min(max_batch, n_mem) : 1;
  AdjPlan plan = adj_plan(n_steps, ckpt, std::max(1, c->adj_levels));
  if (want_grad && c->adj_levels == 0) {
    for (int L = 1; L <= ADJ_MAX_LEVELS; ++L) {
      const AdjPlan p = adj_plan(n_steps, L == 1 ? ckpt : 0, L);
      if (L == 1 || budget / per_member(p) > budget / per_member(plan)) plan = p;
      if (budget / per_member(p) >= (size_t)want) {
        plan = p;
        break;
      }
    }
  }
  const size_t per = per_member(plan);
  int bs = max_batch > 0 ? std::min(max_batch, n_mem) : n_mem;
  bs = (int)std::min<size_t>((size_t)bs, budget / per);
  if (bs < 1)
    return srs::fail(SRS_E_ARG, "checkpoint plan (%d level(s), segments of %lld steps) needs %.3g GB per "
                     "realization, %.3g GB free", plan.L, (long long)plan.K[plan.L - 1], per / 1e9, mem_free / 1e9);
  const int64_t K0 = plan.K[0];
  const int64_t n_ck_kept = want_grad ? plan.slots[0] : 1;
  const int64_t leaf = plan.K[plan.L - 1];
  c->adj_last_levels = plan.L;
  for (int l = 0; l < ADJ_MAX_LEVELS; ++l) c->adj_last_strides[l] = l < plan.L ? plan.K[l] : 0;
  c->adj_last_batch = bs;
  const bool wmat = wmat_possible && (c->adj_w == 1 || (size_t)nblk * bs < kAdjWrittenWBlocks);

  // everything restored at the end
  std::vector<RealLive> live0(n_real);
  HIPCHK(hipMemcpy(live0.data(), c->d_live, n_real * sizeof(RealLive), hipMemcpyDeviceToHost));
  std::vector<std::vector<int32_t>> ryr0(n_mem);
  for (int m = 0; m < n_mem; ++m) {
    const HostReal& hr = c->reals[ids[m]];
    ryr0[m].resize(hr.ryr_ele.size());
    if (!ryr0[m].empty())
      HIPCHK(hipMemcpy(ryr0[m].data(), hr.desc.ryr_state, ryr0[m].size() * sizeof(int32_t), hipMemcpyDeviceToHost));
  }

  DevMem M;
  double *ck = nullptr, *samp = nullptr, *part = nullptr, *dg = nullptr, *gpart = nullptr;
  double *lam0 = nullptr, *lam1 = nullptr, *seg = nullptr;
  std::vector<double*> ckl(plan.L, nullptr);  // checkpoint levels; ckl[0] = ck, the forward's
  d4 *w0 = nullptr, *w1 = nullptr;            // written w of steps n and n - 1
  int32_t *d_sched = nullptr, *d_rid = nullptr;
  int2* d_sub_tiles = nullptr;
  const int nt_mesh = (hm.n + c->tile - 1) / c->tile;
  HIPCHK(M.alloc(&ck, (size_t)n_ck_kept * bs * n6));
  ckl[0] = ck;
  HIPCHK(M.alloc(&d_sub_tiles, (size_t)bs * nt_mesh));
  HIPCHK(M.alloc(&samp, (size_t)n_samples * bs * nout));
  HIPCHK(M.alloc(&part, (size_t)bs * S.n_blocks * nf));
  HIPCHK(M.alloc(&d_sched, std::max<size_t>(1, (size_t)bs * max_sched)));
  HIPCHK(M.alloc(&d_rid, (size_t)bs));
  if (want_grad) {
    HIPCHK(M.alloc(&lam0, (size_t)bs * n6));
    HIPCHK(M.alloc(&lam1, (size_t)bs * n6));
    if (wmat) {
      HIPCHK(M.alloc(&w0, (size_t)bs * np));
      HIPCHK(M.alloc(&w1, (size_t)bs * np));
    }
    HIPCHK(M.alloc(&seg, (size_t)leaf * bs * n6));
    for (int l = 1; l < plan.L; ++l) HIPCHK(M.alloc(&ckl[l], (size_t)plan.slots[l] * bs * n6));
    HIPCHK(M.alloc(&gpart, (size_t)bs * nblk * ADJ_NT));
    HIPCHK(M.alloc(&dg, (size_t)bs * n_samples * nf));
  }
  const uint32_t gmode = G_AFTER_SWEEP | G_NO_GATING | G_REPLAY;
  const unsigned cp_blocks = (unsigned)std::min<size_t>((n6 + 255) / 256, 2048);
  std::vector<double> sh, gco, gp;
  std::vector<RealLive> lv(n_real);
  std::vector<int2> sub_tiles;
  // the sweep visits only the members' tiles (a held realization's tiles would
  // still cost a dispatch each); the context's own list is put back on every
  // return, before the buffers above are freed
  struct TileRestore {
    srs_ctx* c;
    int2* tiles;
    int n_tiles, grid;
    ~TileRestore() { c->d_tiles = tiles; c->n_tiles = n_tiles; c->grid_sweep = grid; }
  } tile_restore{c, c->d_tiles, c->n_tiles, c->grid_sweep};
  hipError_t e = hipSuccess;
  const int rc = [&]() -> int {
  for (int m0 = 0; m0 < n_mem; m0 += bs) {
    const int nb = std::min(bs, n_mem - m0);
    const dim3 g_adj(nblk, nb), g_cp(cp_blocks, nb);
    if (n_real > 1) {
      sub_tiles.clear();
      for (int y = 0; y < nb; ++y)
        for (int t = 0; t < nt_mesh; ++t) sub_tiles.push_back(make_int2(c->reals[ids[m0 + y]].group, t));
      HIPCHK(hipMemcpy(d_sub_tiles, sub_tiles.data(), sub_tiles.size() * sizeof(int2), hipMemcpyHostToDevice));
      c->d_tiles = d_sub_tiles;
      c->n_tiles = (int)sub_tiles.size();
      c->grid_sweep = (c->n_tiles + NXCD - 1) / NXCD * NXCD;
    }
    // members replay their schedules from row 0; the rest are held
    for (int q = 0; q < n_real; ++q) {
      lv[q] = live0[q];
      lv[q].done = 1;
    }
    std::vector<RealDesc> dsc_sub(nb);  // the members' descriptors with their schedules, own state buffers
    for (int y = 0; y < nb; ++y) {
      const int r = ids[m0 + y];
      lv[r].done = 0;
      lv[r].ffwd = 0;
      lv[r].step = 0;
      lv[r].sched_base = 0;
      HostReal& hr = c->reals[r];
      RealDesc& dsc = dsc_sub[y];
      dsc = hr.desc;
      dsc.sched = d_sched + (size_t)y * max_sched;
      dsc.n_sched = n_steps;
      HIPCHK(hipMemcpy(d_sched + (size_t)y * max_sched, schedules + soff[m0 + y],
                       (soff[m0 + y + 1] - soff[m0 + y]) * sizeof(int32_t), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(c->d_desc + r, &dsc, sizeof(dsc), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->d_live, lv.data(), n_real * sizeof(RealLive), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_rid, ids.data() + m0, nb * sizeof(int32_t), hipMemcpyHostToDevice));
    auto scan_into = [&](double* out) {
      hipLaunchKernelGGL(k_scan_blur, dim3(S.n_blocks, nb), dim3(SCAN_BS), 0, c->stream, c->d_desc, c->d_live, 0,
                         d_rid, S.d_cen, S.geo, nf, part);
      hipLaunchKernelGGL(k_scan_final, dim3(1, nb), dim3(256), 0, c->stream, c->d_desc, c->d_live, 0, d_rid,
                         S.n_blocks, nf, nca, part, S.d_pad, S.d_near_f, S.d_near_ca, S.desc.f0, S.desc.c0, out, nout);
    };
    auto ck_at = [&](int l, int64_t k) { return ckl[l] + (size_t)k * nb * n6; };

    // forward, with checkpoints and scan samples; a single realization runs
    // the steps between two of those through the gated sweep (one launch per
    // step, the replay of each step in the next one's workgroup 0)
    hipLaunchKernelGGL(k_seg_start, dim3(nb), dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid, (int64_t)0);
    const bool gated = c->gated && n_real == 1 && c->group == 1 && c->lds && c->occ <= 6;
    int cur = lv[ids[m0]].cur;
    int ks = 0;
    for (int64_t n = 0; n < n_steps && e == hipSuccess;) {
      if (n % K0 == 0 && n / K0 < n_ck_kept)
        hipLaunchKernelGGL(k_batch_state, g_cp, dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid,
                           ck_at(0, n / K0), 1);
      int64_t m = 1;
      if (gated) {
        int64_t next = std::min<int64_t>(n_steps, (n / K0 + 1) * K0);
        if (ks < n_samples && sample_steps[ks] > n) next = std::min<int64_t>(next, sample_steps[ks]);
        m = std::max<int64_t>(1, next - n);
        launch_gated_block(c, m, cur, gmode);
        cur ^= (int)(m & 1);
      } else {
        launch_step(c, gmode);
      }
      n += m;
      if (ks < n_samples && sample_steps[ks] == n) scan_into(samp + (size_t)(ks++) * nb * nout);
      e = hipGetLastError();
      if (e == hipSuccess && n / 4096 != (n - m) / 4096) e = hipStreamSynchronize(c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return srs::fail(SRS_E_HIP, "adjoint forward: %s", hipGetErrorString(e));
    sh.resize((size_t)n_samples * nb * nout);
    if ((e = hipMemcpy(sh.data(), samp, sh.size() * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
      return srs::fail(SRS_E_HIP, "adjoint forward: %s", hipGetErrorString(e));
    // losses and their gradients in the blurred fluorescence
    gco.assign((size_t)nb * n_samples * nf, 0.0);
    const double f0 = S.desc.f0;
    for (int y = 0; y < nb; ++y) {
      const double* tg = targets + (size_t)(m0 + y) * n_samples * nf;
      double Ls = 0.0;
      for (int k = 0; k < n_samples; ++k)
        for (int i = 0; i < nf; ++i) {
          const double dff = sh[((size_t)k * nb + y) * nout + nca + nf + i] / f0 - 1.0;
          const double res = dff - tg[(size_t)k * nf + i];
          Ls += res * res;
          gco[((size_t)y * n_samples + k) * nf + i] = 2.0 * res / f0;
          if (scans) scans[((size_t)(m0 + y) * n_samples + k) * nf + i] = dff;
        }
      losses[m0 + y] = Ls;
    }

    if (want_grad) {
      e = hipMemcpy(dg, gco.data(), gco.size() * sizeof(double), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemsetAsync(lam0, 0, (size_t)nb * n6 * sizeof(double), c->stream);
      if (e == hipSuccess) e = hipMemsetAsync(lam1, 0, (size_t)nb * n6 * sizeof(double), c->stream);
      if (e == hipSuccess) e = hipMemsetAsync(gpart, 0, (size_t)nb * nblk * ADJ_NT * sizeof(double), c->stream);
      double* lam = lam0;   // lambda_{n+1}
      double* lout = lam1;  // lambda_n
      d4* wcur = w0;        // written w of step n
      d4* wnext = w1;       // written w of step n - 1
      ks = n_samples - 1;
      // the loss source of a sample after the last step; the step kernel
      // adds the others to the lambda it produces
      if (sample_steps[ks] == n_steps) {
        hipLaunchKernelGGL(k_scan_adj, g_adj, dim3(ADJ_BS), 0, c->stream, c->d_desc, d_rid, S.d_cen, S.geo, nf,
                           dg + (size_t)ks * nf, n_samples * nf, lam);
        --ks;
      }
      // a segment of the last level: recompute it from its checkpoint, the
      // sweep writing each state into its slot of the segment, then step
      // backward through it
      auto leaf_back = [&](int64_t s0, int64_t s1, const double* ckq) -> hipError_t {
        hipLaunchKernelGGL(k_seg_start, dim3(nb), dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid, s0);
        hipError_t e2 = hipMemcpyAsync(seg, ckq, (size_t)nb * n6 * sizeof(double), hipMemcpyDeviceToDevice,
                                       c->stream);
        if (gated) {
          // a single realization: one gated-sweep launch per step, from slot
          // n - s0 into slot n - s0 + 1, whose workgroup 0 replays the step
          // before; k_gate replays the last one (the same live state as
          // below, and the same fields)
          for (int64_t n = s0; n < s1 - 1 && e2 == hipSuccess; ++n) {
            launch_gated(c, 0, n > s0 ? gmode : 0u, false, seg + (size_t)(n - s0) * n6,
                         seg + (size_t)(n - s0 + 1) * n6);
            e2 = hipGetLastError();
          }
          if (s1 - 1 > s0 && e2 == hipSuccess) {
            launch_gate_single(c, gmode);
            e2 = hipGetLastError();
          }
        }
        for (int64_t n = s0; !gated && n < s1 - 1 && e2 == hipSuccess; ++n) {
          hipLaunchKernelGGL(k_seg_ptrs, dim3((nb + 63) / 64), dim3(64), 0, c->stream, c->d_desc, c->d_live, d_rid,
                             seg, s0, nb);
          launch_step(c, gmode);
          e2 = hipGetLastError();
        }
        for (int64_t n = s1 - 1; n >= s0 && e2 == hipSuccess; --n) {
          const double* Xn = seg + (size_t)(n - s0) * nb * n6;
          if (wmat && n == s1 - 1)
            hipLaunchKernelGGL(k_adj_w, g_adj, dim3(ADJ_BS), 0, c->stream, c->d_desc, d_rid, Xn, lam, wcur);
          const bool scan = ks >= 0 && sample_steps[ks] == n;
          AdjNext nx{wmat && n > s0 ? Xn - nb * n6 : nullptr, wnext, S.d_cen, scan ? dg + (size_t)ks * nf : nullptr,
                     S.geo, nf, n_samples * nf};
          if (scan) --ks;
          const AdjStepFn kern = adj_step_kernel(c->compact_all, scan, wmat, c->adj_waves == 4);
          hipLaunchKernelGGL(kern, g_adj, dim3(ADJ_BS), 0, c->stream, c->d_desc, c->d_live, d_rid, Xn, lam, wcur, n,
                             lout, gpart, nx);
          std::swap(lam, lout);
          std::swap(wcur, wnext);
          e2 = hipGetLastError();
        }
        if (e2 == hipSuccess) e2 = hipStreamSynchronize(c->stream);
        return e2;
      };
      // the level-l segments of [a, b), last first; the states at a + q K[l]
      // are in slot q of level l
      std::function<hipError_t(int, int64_t, int64_t)> back = [&](int l, int64_t a, int64_t b) -> hipError_t {
        const int64_t K = plan.K[l];
        hipError_t e2 = hipSuccess;
        for (int64_t q = (b - a + K - 1) / K - 1; q >= 0 && e2 == hipSuccess; --q) {
          const int64_t s0 = a + q * K, s1 = std::min<int64_t>(b, s0 + K);
          if (l == plan.L - 1) {
            e2 = leaf_back(s0, s1, ck_at(l, q));
            continue;
          }
          // recompute [s0, s1) in the members' own state buffers (the last
          // level's recomputation pointed them into the segment), keeping
          // the next level's checkpoints
          e2 = hipStreamSynchronize(c->stream);
          for (int y = 0; y < nb && e2 == hipSuccess; ++y)
            e2 = hipMemcpy(c->d_desc + ids[m0 + y], &dsc_sub[y], sizeof(RealDesc), hipMemcpyHostToDevice);
          if (e2 != hipSuccess) break;
          hipLaunchKernelGGL(k_seg_start, dim3(nb), dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid, s0);
          hipLaunchKernelGGL(k_batch_state, g_cp, dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid,
                             (double*)ck_at(l, q), 0);
          const int64_t K1 = plan.K[l + 1];
          for (int64_t n = s0; n < s1 && e2 == hipSuccess; ++n) {
            if ((n - s0) % K1 == 0)
              hipLaunchKernelGGL(k_batch_state, g_cp, dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid,
                                 ck_at(l + 1, (n - s0) / K1), 1);
            launch_step(c, gmode);
            e2 = hipGetLastError();
            if (e2 == hipSuccess && (n - s0) % 4096 == 4095) e2 = hipStreamSynchronize(c->stream);
          }
          if (e2 == hipSuccess) e2 = back(l + 1, s0, s1);
        }
        return e2;
      };
      if (e == hipSuccess) e = back(0, 0, n_steps);
      if (e == hipSuccess) {
        gp.resize((size_t)nb * nblk * ADJ_NT);
        e = hipMemcpy(gp.data(), gpart, gp.size() * sizeof(double), hipMemcpyDeviceToHost);
      }
      if (e != hipSuccess) return srs::fail(SRS_E_HIP, "adjoint backward: %s", hipGetErrorString(e));
      for (int y = 0; y < nb; ++y)
        for (int k = 0; k < ADJ_NT; ++k) {
          double sum = 0.0;
          for (int b = 0; b < nblk; ++b) sum += gp[((size_t)y * nblk + b) * ADJ_NT + k];
          grads[(size_t)(m0 + y) * ADJ_NT + k] = sum;
        }
    }
    // the members' states as they were: descriptors (the recomputation
    // pointed their buffers into the segment) and live state (it names the
    // current buffer) first, then the start state into that buffer
    for (int y = 0; y < nb; ++y) {
      const HostReal& hr = c->reals[ids[m0 + y]];
      HIPCHK(hipMemcpy(c->d_desc + ids[m0 + y], &hr.desc, sizeof(hr.desc), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(c->d_live, live0.data(), n_real * sizeof(RealLive), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_batch_state, g_cp, dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid, ck_at(0, 0), 0);
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return SRS_OK;
  }();
  // restore descriptors, live state and RyR states (also after a failure)
  for (int m = 0; m < n_mem; ++m) {
    const HostReal& hr = c->reals[ids[m]];
    (void)hipMemcpy(c->d_desc + ids[m], &hr.desc, sizeof(hr.desc), hipMemcpyHostToDevice);
    if (!ryr0[m].empty())
      (void)hipMemcpy(hr.desc.ryr_state, ryr0[m].data(), ryr0[m].size() * sizeof(int32_t), hipMemcpyHostToDevice);
  }
  (void)hipMemcpy(c->d_live, live0.data(), n_real * sizeof(RealLive), hipMemcpyHostToDevice);
  e = hipDeviceSynchronize();
  if (rc == SRS_OK && e != hipSuccess) return srs::fail(SRS_E_HIP, "adjoint restore: %s", hipGetErrorString(e));
  return rc;
}

static int loss_grad_check(srs_ctx* c, int32_t scan_id, int64_t n_steps, int32_t n_samples,
                           const int64_t* sample_steps, const double* target, int32_t ckpt, const void* loss,
                           bool auto_ckpt = false) {
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (!c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (c->group != 1) return srs::fail(SRS_E_ARG, "the adjoint needs realization groups of size 1");
  if (scan_id < 0 || scan_id >= (int)c->scans.size()) return srs::fail(SRS_E_ARG, "bad scan id %d", scan_id);
  if (c->scans[scan_id].n_f <= 0) return srs::fail(SRS_E_ARG, "the scan has no fluorescence points");
  if (n_steps <= 0 || n_samples <= 0 || !sample_steps || !target || !loss) return srs::fail(SRS_E_ARG, "bad arguments");
  if ((ckpt <= 0 && !(auto_ckpt && ckpt == 0)) || ckpt > 65536)
    return srs::fail(SRS_E_ARG, "checkpoint interval must be 1..65536%s", auto_ckpt ? " or 0" : "");
  for (int k = 0; k < n_samples; ++k)
    if (sample_steps[k] < 1 || sample_steps[k] > n_steps || (k && sample_steps[k] <= sample_steps[k - 1]))
      return srs::fail(SRS_E_ARG, "sample steps must increase within 1..n_steps");
  return SRS_OK;
}

int srs_linescan_loss_grad(srs_ctx* c, int32_t r, int32_t scan_id, int64_t n_steps, const int32_t* schedule,
                           int32_t n_samples, const int64_t* sample_steps, const double* target, int32_t ckpt,
                           double* loss, double* grad) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if ((rc = loss_grad_check(c, scan_id, n_steps, n_samples, sample_steps, target, ckpt, loss))) return rc;
  if (c->reals[r].mesh != c->scans[scan_id].mesh) return srs::fail(SRS_E_ARG, "realization is not on the scan's mesh");
  if (c->reals[r].ext.jsr_diffusion || c->reals[r].ext.b_sr_sites != 0.0)
    return srs::fail(SRS_E_ARG, "realization %d has a model extension on; the adjoint covers the reference model", r);
  if (!c->reals[r].ryr_ele.empty() && !schedule) return srs::fail(SRS_E_ARG, "null schedule");
  return loss_grad_run(c, {r}, scan_id, n_steps, schedule, n_samples, sample_steps, target, ckpt, 1, loss, grad,
                       nullptr);
}

int srs_linescan_loss_grad_batch(srs_ctx* c, int32_t n_real, const int32_t* real_ids, int32_t scan_id,
                                 int64_t n_steps, const int32_t* schedules, int32_t n_samples,
                                 const int64_t* sample_steps, const double* targets, int32_t ckpt, int32_t max_batch,
                                 double* losses, double* grads, double* scans) {
  int rc = c ? SRS_OK : srs::fail(SRS_E_ARG, "null context");
  if (rc) return rc;
  if (n_real <= 0 || !real_ids) return srs::fail(SRS_E_ARG, "no realizations");
  if ((rc = loss_grad_check(c, scan_id, n_steps, n_samples, sample_steps, targets, ckpt, losses, true))) return rc;
  std::vector<int32_t> ids(real_ids, real_ids + n_real);
  std::vector<char> seen(c->reals.size(), 0);
  bool any_ryr = false;
  for (int32_t r : ids) {
    if ((rc = check_real(c, r))) return rc;
    if (seen[r]++) return srs::fail(SRS_E_ARG, "realization %d listed twice", r);
    if (c->reals[r].mesh != c->scans[scan_id].mesh)
      return srs::fail(SRS_E_ARG, "realization %d is not on the scan's mesh", r);
    if (c->reals[r].ext.jsr_diffusion || c->reals[r].ext.b_sr_sites != 0.0)
      return srs::fail(SRS_E_ARG, "realization %d has a model extension on; the adjoint covers the reference model", r);
    any_ryr |= !c->reals[r].ryr_ele.empty();
  }
  if (any_ryr && !schedules) return srs::fail(SRS_E_ARG, "null schedules");
  return loss_grad_run(c, ids, scan_id, n_steps, schedules, n_samples, sample_steps, targets, ckpt, max_batch, losses,
                       grads, scans);
}

int srs_adjoint_last_plan(srs_ctx* c, int32_t* levels, int64_t* strides, int32_t* sub_batch) {
  srs::clear_error();
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (levels) *levels = c->adj_last_levels;
  if (strides)
    for (int l = 0; l < ADJ_MAX_LEVELS; ++l) strides[l] = c->adj_last_strides[l];
  if (sub_batch) *sub_batch = c->adj_last_batch;
  return SRS_OK;
}


}  // extern "C"
