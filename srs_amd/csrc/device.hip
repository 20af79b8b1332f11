// device.hip -- gfx950 kernels and the device half of the C ABI.
//
// One spark time-step (cpu/RunSimulation.cpp:591-654) is two kernels:
//   k_sweep : every element of every live realization, fused: 5-point ELL
//             diffusion of Ca/Fluo/ATP/CaM, binding reactions, SERCA, jSR
//             refill, RyR/LCC flux injection, rapid SL buffer and the forward
//             Euler update, old state -> new state (ping-pong buffers).
//   k_gate  : one workgroup per realization: RyR and LCC Markov gating on the
//             new Ca (integrated-hazard residuals, Philox draws), then,
//             optionally, the RunSimulation loop body for the next time
//             (while-condition, LCC trigger, voltage clamp, RyR hold,
//             states-record output).
// A single realization stepping without the loop body runs each step as one
// launch instead, k_sweep_gated: its workgroup 0 runs k_gate's body for the
// previous step before updating the channel elements.
// The reference runs the same step as nine kernels with a state-sized rate
// buffer (gpu/RunSimulation.cu:781-852); here the rates never leave registers.
//
// Arithmetic keeps the reference's operation order and this file is compiled
// without FP contraction, so the field update is bit-identical to the
// reference CPU path for the same row order of each element's neighbours.
// The line-scan operator is in linescan.hip, recording and the adjoint in
// adjoint.hip; the shared types are in dev_common.h.
#include "dev_common.h"

#ifndef SRS_UNIT
#define SRS_UNIT 0
#endif

#include <functional>
#include <mutex>

namespace srs_dev {

// The state of one element of one realization: its own and its neighbours'
// diffusing quads, TnC and CSQ where they are live.
struct Loc {
  d4 m0, m1, m2, m3, m4;
  double c_tr, c_cs;
};

// A realization's state buffers for one step: read X[cur], write X[cur ^ 1].
// Both pointers are loaded and one is selected, so that the live index and
// the pointers arrive in the same round of scalar loads.
struct Bufs {
  const SRS_GLOBAL double* xo;
  SRS_GLOBAL double* xn;
};
__device__ __forceinline__ Bufs bufs(double* x0, double* x1, int cur) {
  return cur ? Bufs{gp(x1), gpw(x0)} : Bufs{gp(x0), gpw(x1)};
}

// Pins uniform values to this point of the program: the scalar loads that
// produce them are issued as one batch ahead of the branches that follow,
// instead of one memory round trip after another where each is first used.
// The values are folded into one operand of an empty asm statement.
template <typename T>
__device__ __forceinline__ uint64_t fold_bits(const T v) {
  if constexpr (std::is_pointer<T>::value) return (uint64_t)(uintptr_t)v;
  else return (uint64_t)(int64_t)v;
}
template <typename... T>
__device__ __forceinline__ uint64_t fold(const T... v) {
  return (fold_bits(v) ^ ...);
}
__device__ __forceinline__ void scalars_here(const uint64_t folded) { __asm__ volatile("" ::"s"(folded)); }

template <int FK>
__device__ __forceinline__ Loc load_loc(const RealDesc& d, const SRS_GLOBAL double* __restrict__ Xo, const int j,
                                        const Geo& g) {
  // Realization-group layout: element e of the realization in slot s of a
  // group of K is quad gidx(e, K, s); the TnC and CSQ planes follow the quads.
  const int K = d.kst, s = d.slot;
  const size_t np = d.np;
  const SRS_GLOBAL d4* Mo = (const SRS_GLOBAL d4*)Xo;
  Loc c;
  c.m0 = Mo[gidx(j, K, s)];
  // TnC has a nonzero rate only in cytosol elements with troponin, CSQ only
  // in the jSR (cpu/RunSimulation.cpp:45-46 reset both rates to zero).  A
  // zero rate leaves the value unchanged, and both state buffers hold the same
  // value, so those planes are neither read nor written elsewhere.
  const uint8_t f = FK >= 0 ? (uint8_t)FK : g.f;
  const bool live_tr = (f & (F_JSR | F_TROPC)) == F_TROPC;
  const bool live_cs = (f & F_JSR) != 0;
  c.c_tr = live_tr ? Xo[4 * K * np + gidx(j, K, s)] : 0.0;
  c.c_cs = live_cs ? Xo[5 * K * np + gidx(j, K, s)] : 0.0;
  c.m1 = Mo[gidx(g.n1, K, s)];
  c.m2 = Mo[gidx(g.n2, K, s)];
  c.m3 = Mo[gidx(g.n3, K, s)];
  c.m4 = Mo[gidx(g.n4, K, s)];
  return c;
}

// As load_loc for a single-realization layout, with the neighbour quads that
// lie in the workgroup's own tile read from the copy staged in LDS.
template <int FK, int BS>
__device__ __forceinline__ Loc load_loc_lds(const RealDesc& d, const SRS_GLOBAL double* __restrict__ Xo, const int j,
                                            const Geo& g, const d4* tq, const int base, const d4& m0,
                                            const double tr) {
  const size_t np = d.np;
  const SRS_GLOBAL d4* Mo = (const SRS_GLOBAL d4*)Xo;
  Loc c;
  c.m0 = m0;
  const uint8_t f = FK >= 0 ? (uint8_t)FK : g.f;
  const bool live_tr = (f & (F_JSR | F_TROPC)) == F_TROPC;
  const bool live_cs = (f & F_JSR) != 0;
  c.c_tr = live_tr ? tr : 0.0;
  c.c_cs = live_cs ? Xo[5 * np + j] : 0.0;
  auto nq = [&](int nb) -> d4 {
    const unsigned k = (unsigned)(nb - base);
    return k < (unsigned)BS ? tq[k] : Mo[nb];
  };
  c.m1 = nq(g.n1);
  c.m2 = nq(g.n2);
  c.m3 = nq(g.n3);
  c.m4 = nq(g.n4);
  return c;
}

// The fused step of one element of one realization, from its loaded row and
// state, in the reference's operation order (cpu/RunSimulation.cpp:591-648).
template <bool NTS, int FK>
__device__ __forceinline__ void step_elem(const RealDesc& d, const RealLive& l, const Bufs bf, const int j,
                                          const Geo& g, const Loc& c) {
  const int K = d.kst, s = d.slot;
  const size_t np = d.np;
  const SRS_GLOBAL double* __restrict__ Xo = bf.xo;
  SRS_GLOBAL double* __restrict__ Xn = bf.xn;
  const double* P = d.P;
  const double dt = P[SRS_R_T_STEP];
  const uint8_t f = FK >= 0 ? (uint8_t)FK : g.f;
  const double a0 = g.a0, a1 = g.a1, a2 = g.a2, a3 = g.a3, a4 = g.a4;
  const double c_ca = c.m0.x, c_at = c.m0.y, c_cm = c.m0.z, c_dy = c.m0.w;
  const double c_tr = c.c_tr, c_cs = c.c_cs;
  const d4 m1 = c.m1, m2 = c.m2, m3 = c.m3, m4 = c.m4;
  const bool live_tr = (f & (F_JSR | F_TROPC)) == F_TROPC;
  const bool live_cs = (f & F_JSR) != 0;

  double br_ca = 0.0, br_dy = 0.0, br_at = 0.0, br_cm = 0.0;
  if (f & F_BR) {
    const SRS_GLOBAL double* br = gp(d.br);
    br_ca = br[j];
    br_dy = br[np + j];
    br_at = br[2 * np + j];
    br_cm = br[3 * np + j];
  }
  // cpu/RunSimulation.cpp:61-71 (Ca uses D_CA in both domains; with the jSR
  // diffusion extension, D_CAJSR in jSR elements)
  const double d_ca = (live_cs && d.jsr_diff) ? P[SRS_R_D_CA_JSR] : P[SRS_R_D_CA];
  double rca = d_ca * (c_ca * a0 + m1.x * a1 + m2.x * a2 + m3.x * a3 + m4.x * a4) + br_ca;
  double rdy = P[SRS_R_D_DYE] * (c_dy * a0 + m1.w * a1 + m2.w * a2 + m3.w * a3 + m4.w * a4) + br_dy;
  double rat = P[SRS_R_D_ATP] * (c_at * a0 + m1.y * a1 + m2.y * a2 + m3.y * a3 + m4.y * a4) + br_at;
  double rcm = P[SRS_R_D_CMDN] * (c_cm * a0 + m1.z * a1 + m2.z * a2 + m3.z * a3 + m4.z * a4) + br_cm;
  double rtr = 0, rcs = 0;
  const double nsr = l.ca_nsr;

  if (!(f & F_JSR)) {
    if (f & F_TROPC) {  // :75-91
      const double r = P[SRS_R_K_ON_TRPN] * (P[SRS_R_B_TOT_TRPN] - c_tr) * c_ca - P[SRS_R_K_OFF_TRPN] * c_tr;
      rca += -r;
      rtr = r;
      double K_i = c_ca / P[SRS_R_K_D_I];
      K_i = K_i * K_i;
      const double D_cycle = l.serca[0] + K_i * l.serca[1] + K_i * K_i * l.serca[2];
      const double v_cycle = (K_i * K_i * 3.24873e12 + K_i * l.serca[3] - l.serca[4]) / D_cycle;
      const double d_ca = (2e-3) * v_cycle * P[SRS_R_A_P];
      rca -= d_ca;
    }
    double r = P[SRS_R_K_ON_DYE] * (P[SRS_R_B_TOT_DYE] - c_dy) * c_ca - P[SRS_R_K_OFF_DYE] * c_dy;
    rca += -r;
    rdy += r;
    r = P[SRS_R_K_ON_ATP] * (P[SRS_R_B_TOT_ATP] - c_at) * c_ca - P[SRS_R_K_OFF_ATP] * c_at;
    rca += -r;
    rat += r;
    r = P[SRS_R_K_ON_CMDN] * (P[SRS_R_B_TOT_CMDN] - c_cm) * c_ca - P[SRS_R_K_OFF_CMDN] * c_cm;
    rca += -r;
    rcm += r;
  } else {  // :113-121
    const double r = P[SRS_R_K_ON_CSQN] * (P[SRS_R_B_T_JSR] - c_cs) * c_ca - P[SRS_R_K_OFF_CSQN] * c_cs;
    rca += -r;
    rcs = r;
    rca += (nsr - c_ca) * P[SRS_R_V_REFILL];
  }

  if (f & F_CHAN) {  // :619-632, contributions in the reference's channel order
    int lo = 0, hi = d.n_celem;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (d.celem[mid] < j) lo = mid + 1; else hi = mid;
    }
    if (lo < d.n_celem && d.celem[lo] == j) {
      const double v0 = d.V0[j];
      for (int e = d.cstart[lo]; e < d.cstart[lo + 1]; ++e) {
        const int ch = d.centry[e];
        if (ch >= 0) {
          const int i = ch >> 1;
          if (d.ryr_state[i]) {
            const double J = P[SRS_R_V_RYR] * (Xo[4 * gidx(d.ryr_jsr[i], K, s)] -
                                               Xo[4 * gidx(d.ryr_ele[i], K, s)]) * P[SRS_R_V_CELL];
            if (ch & 1) rca -= J / v0; else rca += J / v0;
          }
        } else {
          const int i = -ch - 1;
          if (srs_lcc_open(d.lcc_state[i], d.lccv_state[i])) rca += srs_lcc_flux_raw(l.vm, c_ca, P[SRS_R_CA_0]) / v0;
        }
      }
    }
  }

  // SR-site rapid buffer of the cytosol (extension; b_sr == 0 in the reference model)
  const bool sr = !live_cs && d.b_sr != 0.0;
  if ((f & F_SL) || sr) {
    double B = 1.0;
    if (f & F_SL) {  // :15-22; B == 1 exactly where the TT area is zero
      const double ttj = gp(d.tt)[j], v0j = gp(d.V0)[j];
      const double btothi = ttj * P[SRS_R_B_TOT_SL_HI] / (v0j * (1e-15));
      const double btotlo = ttj * P[SRS_R_B_TOT_SL_LO] / (v0j * (1e-15));
      const double kh = P[SRS_R_K_D_SL_HI], kl = P[SRS_R_K_D_SL_LO];
      B = 1 + (btothi * kh / ((kh + c_ca) * (kh + c_ca))) + (btotlo * kl / ((kl + c_ca) * (kl + c_ca)));
    }
    if (sr) B = B + d.b_sr * d.k_sr / ((d.k_sr + c_ca) * (d.k_sr + c_ca));
    rca /= B;
  }

  // :640-644 forward Euler
  const d4 mn = {c_ca + rca * dt, c_at + rat * dt, c_cm + rcm * dt, c_dy + rdy * dt};
  const size_t q = gidx(j, K, s);
  stg<NTS>((SRS_GLOBAL d4*)Xn + q, mn);
  if (live_tr) stg<NTS>(Xn + 4 * K * np + q, c_tr + rtr * dt);
  if (live_cs) stg<NTS>(Xn + 5 * K * np + q, c_cs + rcs * dt);
}

// Flags shared by every lane of a wavefront-sized element group: the low byte
// is the OR of the lanes' flags, the high byte the AND.  A wavefront of plain
// cytosol elements with troponin takes the path with the flags known.
constexpr uint8_t F_SPECIAL = F_JSR | F_SL | F_BR | F_CHAN;

// One tile of one realization group: the operator rows are loaded once and
// applied to the KG realizations of the group, one after another; -1 marks an
// empty slot or a member that does not step.
template <bool NTS, int KG, int FK>
__device__ __forceinline__ void sweep_group(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                            const int* mem, const int j, const Geo& g) {
  for (int s = 0; s < KG; ++s)
    if (KG == 1 || mem[s] >= 0) {
      const RealDesc& d = D[mem[s]];
      const Bufs bf = bufs(d.X[0], d.X[1], L[mem[s]].cur);
      const Loc c = load_loc<FK>(d, bf.xo, j, g);
      step_elem<NTS, FK>(d, L[mem[s]], bf, j, g, c);
    }
}

// The sizes and operator-row pointers of a mesh, folded for scalars_here.
template <bool COMPACT>
__device__ __forceinline__ uint64_t geo_fold(const RealDesc& d) {
  if constexpr (COMPACT) return fold(d.n, d.np, d.flags, d.off, d.cls, d.dq, d.tab);
  else return fold(d.n, d.np, d.flags, d.coef, d.nbr);
}

// Single-realization groups with LDS staging: the tile's own quads are
// loaded coalesced into LDS, and neighbour quads inside the tile come from
// there instead of through the L1 cache.
template <int BS, bool NTG, bool NTS, bool COMPACT, bool SKIP_CHAN = false>
__device__ __forceinline__ void sweep_tile_lds(const RealDesc& d, const RealLive& l, const Bufs bf, int t) {
  __shared__ d4 tq[BS];
  const int base = t * BS;
  const int j = base + threadIdx.x;
  const bool in = j < d.n;
  Geo g{};
  d4 m0 = {0.0, 0.0, 0.0, 0.0};
  double tr = 0.0;
  if (in) {
    // the element's TnC is loaded with its row and quad, before its flags are
    // known (it is live in all but jSR and troponin-free cytosol elements), so
    // that the loads after the barrier are only the neighbour gathers
    g = load_geo<NTG, COMPACT, -1>(d, j);
    m0 = ((const SRS_GLOBAL d4*)bf.xo)[j];
    tr = bf.xo[4 * (size_t)d.np + j];
  }
  tq[threadIdx.x] = m0;
  __syncthreads();
  if (!in) return;
  if (SKIP_CHAN && (g.f & F_CHAN)) return;  // updated by the gating workgroup (sweep_gated)
  const bool plain = (g.f & F_SPECIAL) == 0 && (g.f & F_TROPC);
  if (__builtin_amdgcn_ballot_w64(!plain) == 0) {
    const Loc c = load_loc_lds<F_TROPC, BS>(d, bf.xo, j, g, tq, base, m0, tr);
    step_elem<NTS, F_TROPC>(d, l, bf, j, g, c);
  } else {
    const Loc c = load_loc_lds<-1, BS>(d, bf.xo, j, g, tq, base, m0, tr);
    step_elem<NTS, -1>(d, l, bf, j, g, c);
  }
}

// The element's flags are loaded with its operator row; a wavefront whose
// active lanes are all plain cytosol with troponin (the bulk of a mesh) then
// runs the specialisation with the flags known at compile time.
template <int BS, bool NTG, bool NTS, bool COMPACT, int KG>
__device__ __forceinline__ void sweep_tile(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                           const int2* __restrict__ tiles, int n_tiles, const int* __restrict__ groups,
                                           int single) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  if (b >= n_tiles) return;
  const int2 tl = single ? make_int2(0, b) : tiles[b];  // {group, tile}
  // members that step: a finished member (done) or one in the gating-only
  // ffwd phase keeps its fields, so its part of the tile is skipped
  int mem[KG];
  int first = -1;
#pragma unroll
  for (int s = 0; s < KG; ++s) {
    const int r = single ? (s == 0 ? 0 : -1) : groups[KG * tl.x + s];
    mem[s] = (r >= 0 && !(L[r].done | L[r].ffwd)) ? r : -1;
    if (first < 0) first = mem[s];
  }
  if (first < 0) return;
  const RealDesc& d = D[first];
  const int j = tl.y * BS + threadIdx.x;
  if (j >= d.n) return;
  const Geo g = load_geo<NTG, COMPACT, -1>(d, j);
  const bool plain = (g.f & F_SPECIAL) == 0 && (g.f & F_TROPC);
  if (__builtin_amdgcn_ballot_w64(!plain) == 0) sweep_group<NTS, KG, F_TROPC>(D, L, mem, j, g);
  else sweep_group<NTS, KG, -1>(D, L, mem, j, g);
}

template <int BS, bool NTG, bool NTS, bool COMPACT>
__device__ __forceinline__ void sweep_lds(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                          const int2* __restrict__ tiles, int n_tiles, const int* __restrict__ groups,
                                          int single) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  if (b >= n_tiles) return;
  const int2 tl = single ? make_int2(0, b) : tiles[b];
  const int r = single ? 0 : groups[tl.x];
  const RealDesc& d = D[r];
  const RealLive& l = L[r];
  const int skip = l.done | l.ffwd;
  const Bufs bf = bufs(d.X[0], d.X[1], l.cur);
  scalars_here(fold(skip, bf.xo, bf.xn) ^ geo_fold<COMPACT>(d));
  if (skip) return;
  sweep_tile_lds<BS, NTG, NTS, COMPACT>(d, l, bf, tl.y);
}

// Register budgets: the compiler's choice, or the largest that lets W
// wavefronts share each SIMD (64 VGPRs for eight, 80 for six).
template <int BS, bool NTG, bool NTS, bool COMPACT, int KG>
__global__ __launch_bounds__(BS) void k_sweep(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                              const int2* __restrict__ tiles, int n_tiles,
                                              const int* __restrict__ groups, int single) {
  sweep_tile<BS, NTG, NTS, COMPACT, KG>(D, L, tiles, n_tiles, groups, single);
}
template <int BS, bool NTG, bool NTS, bool COMPACT, int KG, int W>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(W))) void k_sweep_w(
    const RealDesc* __restrict__ D, const RealLive* __restrict__ L, const int2* __restrict__ tiles, int n_tiles,
    const int* __restrict__ groups, int single) {
  sweep_tile<BS, NTG, NTS, COMPACT, KG>(D, L, tiles, n_tiles, groups, single);
}
template <int BS, bool NTG, bool NTS, bool COMPACT>
__global__ __launch_bounds__(BS) void k_sweep_lds(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                                  const int2* __restrict__ tiles, int n_tiles,
                                                  const int* __restrict__ groups, int single) {
  sweep_lds<BS, NTG, NTS, COMPACT>(D, L, tiles, n_tiles, groups, single);
}
template <int BS, bool NTG, bool NTS, bool COMPACT, int W>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(W))) void k_sweep_lds_w(
    const RealDesc* __restrict__ D, const RealLive* __restrict__ L, const int2* __restrict__ tiles, int n_tiles,
    const int* __restrict__ groups, int single) {
  sweep_lds<BS, NTG, NTS, COMPACT>(D, L, tiles, n_tiles, groups, single);
}

__device__ void write_record(const RealDesc& d, RealLive& l, const double* Ca, double t) {
  // Called by one thread: sums run sequentially, in the reference's order.
  if (l.rec_count >= d.rec_cap) { l.rec_overflow = 1; return; }
  float* rec = d.rec + (size_t)l.rec_count * d.rec_len;
  int k = 0;
  rec[k++] = (float)t;
  double i_ryr = 0;
  for (int i = 0; i < d.n_ryr; ++i)
    i_ryr += d.ryr_state[i] * d.P[SRS_R_V_RYR] * (Ca[cq(d, d.ryr_jsr[i])] - Ca[cq(d, d.ryr_ele[i])]) * (d.P[SRS_R_V_CELL]) * 96485 *
             2 * (1e-6);
  rec[k++] = (float)i_ryr;
  int n_open = 0;
  for (int i = 0; i < d.n_ryr; ++i) n_open += d.ryr_state[i];
  rec[k++] = (float)n_open;
  int n_open_lcc = 0;
  for (int i = 0; i < d.n_lcc; ++i)
    if (d.lccv_state[i] == 2 && (d.lcc_state[i] == 6 || d.lcc_state[i] == 12)) n_open_lcc++;
  rec[k++] = (float)n_open_lcc;
  double i_lcc = 0;
  for (int i = 0; i < d.n_lcc; ++i)
    if (srs_lcc_open(d.lcc_state[i], d.lccv_state[i]))
      i_lcc += srs_lcc_flux_raw(l.vm, Ca[cq(d, d.lcc_ele[i])], d.P[SRS_R_CA_0]) * 96485 * 2 * (1e-6);
  rec[k++] = (float)i_lcc;
  double m = 0;
  for (int i = 0; i < d.n_ryr; ++i) m += Ca[cq(d, d.ryr_jsr[i])];
  m /= d.n_ryr;
  rec[k++] = (float)m;
  m = 0;
  for (int i = 0; i < d.n_ryr; ++i) m += Ca[cq(d, d.ryr_ele[i])];
  m /= d.n_ryr;
  rec[k++] = (float)m;
  if (d.gpu_layout)
    for (int i = 0; i < d.n_ryr; ++i) rec[k++] = (float)Ca[cq(d, d.ryr_ele[i])];
  if (d.ryr_out)
    for (int i = 0; i < d.n_ryr; ++i) rec[k++] = (float)d.ryr_state[i];
  l.rec_count++;
}

// The gating and loop body of realization r by one workgroup of NT threads.
// LEAN loads each LCC's inputs when it is updated rather than with the RyRs'
// before the first barrier: one more memory round trip, fewer live registers.
// PROTO = false leaves out the loop body (callers that never pass G_PROTOCOL).
template <int NT, bool LEAN = false, bool PROTO = true>
__device__ __forceinline__ void gate_workgroup(const RealDesc* __restrict__ D, RealLive* __restrict__ L, const int r,
                                               const uint32_t mode) {
  static_assert(MAX_RYR % NT == 0 && MAX_LCC % NT == 0, "channel slots per thread");
  const RealDesc& d = D[r];
  RealLive& l = L[r];
  if (l.done) return;
  __shared__ int s_old[MAX_RYR];
  __shared__ int s_a[MAX_RYR];
  __shared__ int s_b[MAX_RYR];
  const int tid = threadIdx.x;
  const int cur = l.cur;
  // In the flag_ffwd phase the sweep skipped this realization: the fields are
  // unchanged and only the channels gate (cpu/RunSimulation.cpp:474-476).
  const bool ff = l.ffwd != 0;
  const bool swept = (mode & G_AFTER_SWEEP) && !ff;
  const double* Xn = swept ? d.X[cur ^ 1] : d.X[cur];
  const double* Ca = Xn;  // Ca of element e at Ca[cq(d, e)] (interleaved diffusing species)
  const double* P = d.P;
  const double dt = P[SRS_R_T_STEP];
  const int64_t step = l.step;
  // read by every thread before the first barrier: thread 0 rewrites the
  // live state at the end, and the time decides below whether the whole
  // workgroup takes part in the loop-body counts
  const double t = l.t + dt;  // cpu/RunSimulation.cpp:572-573 (ffwd: :477-478)
  // ExplicitStep honours flag_nogating; the ffwd loop calls
  // UpdateChannelGating directly (cpu/RunSimulation.cpp:476, 650-652).
  const bool gating = !(mode & G_NO_GATING) && (ff || !d.I[SRS_I_FLAG_NO_GATING]);
  const int nr = d.n_ryr;

  if (gating) {
    // Everything a channel's update reads is loaded before the first barrier
    // (RyR i by thread i mod 256, LCC i by thread 255 - i mod 256, so RyRs
    // and LCCs are updated side by side), leaving one memory round trip in
    // front of the arithmetic.
    constexpr int RPT = MAX_RYR / NT, LPT = MAX_LCC / NT;
    double ca_c[RPT], ca_j[RPT], Rr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int i = tid + k * NT;
      if (i < nr) {
        const int so = d.ryr_state[i];
        s_old[i] = so;
        s_a[i] = so;
        ca_c[k] = Ca[cq(d, d.ryr_ele[i])];
        ca_j[k] = Ca[cq(d, d.ryr_jsr[i])];
        Rr[k] = d.R[i];
      }
    }
    int ls[LPT], lvs[LPT];
    double ca_l[LPT], Rl[LPT];
    auto load_lcc = [&](const int k) {
      const int i = (NT - 1 - tid) + k * NT;
      if (i < d.n_lcc) {
        ls[k] = d.lcc_state[i];
        lvs[k] = d.lccv_state[i];
        ca_l[k] = Ca[cq(d, d.lcc_ele[i])];
        Rl[k] = d.R[nr + i];
      }
    };
    if constexpr (!LEAN) {
#pragma unroll
      for (int k = 0; k < LPT; ++k) load_lcc(k);
    }
    __syncthreads();
    // RyRs: the reference updates them one after another (later RyRs see
    // earlier flips through the neighbour counts, cpu/RunSimulation.cpp:661-666).
    // Each sweep below recomputes every RyR with lower-index neighbours taken
    // from the previous sweep's result and higher-index ones from the old
    // states; the first sweep that changes nothing is the sequential result.
    const bool coupled = P[SRS_R_RYR_A_STAR] != 0.0;
    int* prev = s_a;
    int* next = s_b;
    if (coupled) {
      for (int it = 0; it <= nr; ++it) {
        int changed = 0;
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
          const int i = tid + k * NT;
          if (i >= nr) continue;
          int no = 0, ncl = 0;
          for (int q = 0; q < 4; ++q) {
            const int nb = d.ryr_nb[4 * i + q];
            if (nb >= 0) {
              const int s = (nb < i) ? prev[nb] : s_old[nb];
              if (s) no++; else ncl++;
            }
          }
          const double rate = srs_ryr_rate(s_old[i], no, ncl, ca_c[k], ca_j[k], P);
          const double Rn = Rr[k] + rate * dt;
          const int ns = (Rn > 0) ? (s_old[i] == 1 ? 0 : 1) : s_old[i];
          next[i] = ns;
          changed |= (ns != prev[i]);
        }
        changed = __syncthreads_or(changed);
        int* tmp = prev; prev = next; next = tmp;
        if (!changed) break;
      }
    }
    // Commit with neighbour states from the fixed point (for uncoupled RyRs
    // the neighbour counts do not enter the rate, A* = 0 gives X = 1).
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int i = tid + k * NT;
      if (i >= nr) continue;
      int no = 0, ncl = 0;
      if (coupled) {
        for (int q = 0; q < 4; ++q) {
          const int nb = d.ryr_nb[4 * i + q];
          if (nb >= 0) {
            const int s = (nb < i) ? prev[nb] : s_old[nb];
            if (s) no++; else ncl++;
          }
        }
      }
      const double rate = srs_ryr_rate(s_old[i], no, ncl, ca_c[k], ca_j[k], P);
      double Rn = Rr[k] + rate * dt;
      int ns = s_old[i];
      if (Rn > 0) {
        uint64_t o[4];
        srs_philox4x64_10((uint64_t)step, (uint64_t)i, d.stream, SRS_PURPOSE_GATING, d.key, SRS_PHILOX_KEY1, o);
        Rn = log(srs_open01(o[0]));
        ns = (s_old[i] == 1) ? 0 : 1;
      }
      d.R[i] = Rn;
      d.ryr_state[i] = ns;
    }
    // LCCs, cpu/RunSimulation.cpp:229-400
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
      const int i = (NT - 1 - tid) + k * NT;
      if (i >= d.n_lcc) continue;
      if constexpr (LEAN) load_lcc(k);
      const int s = ls[k], vs = lvs[k];
      double tr[3];
      int32_t tg[3];
      const double lr = srs_lcc_rates(s, vs, l.vm, ca_l[k], tr, tg);
      const int ic = nr + i;
      double Rn = Rl[k] + (tr[0] + tr[1] + tr[2] + lr) * dt;
      if (Rn > 0) {
        uint64_t o[4];
        srs_philox4x64_10((uint64_t)step, (uint64_t)ic, d.stream, SRS_PURPOSE_GATING, d.key, SRS_PHILOX_KEY1, o);
        Rn = log(srs_open01(o[0]));
        const double u = srs_open01(o[1]) * (tr[0] + tr[1] + tr[2] + lr);
        if (u < tr[0]) d.lcc_state[i] = tg[0];
        else if (u < tr[0] + tr[1]) d.lcc_state[i] = tg[1];
        else if (u < tr[0] + tr[1] + tr[2]) d.lcc_state[i] = tg[2];
        else d.lccv_state[i] = (vs == 1) ? 2 : 1;
      }
      d.R[ic] = Rn;
    }
  }
  __syncthreads();

  if (mode & G_REPLAY) {
    const int64_t k = step + 1 - l.sched_base;  // schedule row of the next step
    if (k < d.n_sched)
      for (int i = tid; i < nr; i += NT) d.ryr_state[i] = d.sched[k * nr + i];
  }

  const double iv = P[SRS_R_OUTPUT_STATES_INTERVAL];
  const bool out_due = floor(t / iv) > floor((t - dt) / iv);
  // The loop body's counts (open RyRs, open LCCs, the largest cleft Ca over
  // the RyRs) are needed in the ffwd phase and at states outputs.  The
  // workgroup reduces them with wavefront ballots and a cross-lane max
  // (exact: integer counts and a maximum, so any order gives the sequential
  // result).  Each thread reads the channel entries it wrote itself above.
  int n_open_all = 0, n_lcc_all = 0;
  double cmax_all = 0.0;
  if (PROTO && (mode & G_PROTOCOL) && (ff || out_due)) {
    __shared__ int s_cnt[2][NT / 64];
    __shared__ double s_max[NT / 64];
    int cr = 0, cl = 0;
    double m = 0.0;
    for (int b = 0; b < nr; b += NT) {
      const int i = b + tid;
      const bool open = i < nr && d.ryr_state[i] != 0;
      cr += __builtin_popcountll(__builtin_amdgcn_ballot_w64(open));
      if (i < nr) {
        const double c = Ca[cq(d, d.ryr_ele[i])];
        if (c > m) m = c;
      }
    }
    for (int b = 0; b < d.n_lcc; b += NT) {
      const int i = b + (NT - 1 - tid);
      const bool open = i < d.n_lcc && d.lccv_state[i] == 2 && (d.lcc_state[i] == 6 || d.lcc_state[i] == 12);
      cl += __builtin_popcountll(__builtin_amdgcn_ballot_w64(open));
    }
    for (int o = 32; o > 0; o >>= 1) {
      const double v = __shfl_xor(m, o);
      if (v > m) m = v;
    }
    const int wave = tid >> 6;
    if ((tid & 63) == 0) {
      s_cnt[0][wave] = cr;
      s_cnt[1][wave] = cl;
      s_max[wave] = m;
    }
    __syncthreads();
    if (tid == 0)
      for (int w = 0; w < NT / 64; ++w) {
        n_open_all += s_cnt[0][w];
        n_lcc_all += s_cnt[1][w];
        if (s_max[w] > cmax_all) cmax_all = s_max[w];
      }
  }

  if (tid != 0) return;
  if (swept) l.cur = cur ^ 1;
  l.step = step + 1;
  l.t = t;
  l.dt_last = dt;
  if (!PROTO || !(mode & G_PROTOCOL)) return;

  if (ff) {  // the rest of the ffwd loop body, cpu/RunSimulation.cpp:480-518
    const int n_open = n_open_all, n_open_lcc = n_lcc_all;
    const bool stop = n_open > 0 || n_open_lcc > 0 || t > P[SRS_R_T_FINAL];
    if (!stop && out_due) write_record(d, l, Ca, t);
    if (d.I[SRS_I_FLAG_V_CLAMP]) l.vm = (t <= P[SRS_R_T_CLAMP]) ? P[SRS_R_V_CLAMP] : P[SRS_R_DEFAULT_VM];
    if (!stop) return;
    // The main loop starts at this t with N_RyR_Open from the last ffwd
    // count; ca_ss_max keeps its initial value.
    l.ffwd = 0;
    l.n_open_last = n_open;
  }

  // While-condition at the top of the next iteration (:522).
  const bool fid = d.I[SRS_I_FLAG_FIDELITY_SIM] != 0;
  const bool cont = t <= P[SRS_R_T_FINAL] &&
                    (!fid || ((l.lcc_open || l.n_open_last > 0 || l.ca_max_last > P[SRS_R_C_MIN]) &&
                              l.n_open_last < d.I[SRS_I_N_RYR_FIDELITY]));
  if (!cont) {
    l.done = 1;
    write_record(d, l, Ca, t);  // final states output, :584
    return;
  }
  if (d.I[SRS_I_FLAG_LCC] && l.lcc_open && t >= l.t_lcc_close) {  // :525-528
    l.lcc_open = 0;
    d.lcc_state[d.I[SRS_I_LCC_INDEX]] = 1;
  }
  if (d.I[SRS_I_FLAG_V_CLAMP]) l.vm = (t <= P[SRS_R_T_CLAMP]) ? P[SRS_R_V_CLAMP] : P[SRS_R_DEFAULT_VM];  // :531-537
  const int ir = d.I[SRS_I_INIT_RYR];
  if (ir != -1 && t < P[SRS_R_INIT_RYR_HOLD] && !d.ryr_state[ir]) {  // :540-542
    d.ryr_state[ir] = 1;
    ++n_open_all;  // counted closed above
  }
  if (out_due) {  // :552-566
    l.n_open_last = n_open_all;
    l.ca_max_last = cmax_all;
    write_record(d, l, Ca, t);
  }
}

#if SRS_UNIT == 0
__global__ __launch_bounds__(GATE_THREADS) void k_gate(const RealDesc* __restrict__ D, RealLive* __restrict__ L,
                                                       uint32_t mode) {
  gate_workgroup<GATE_THREADS>(D, L, blockIdx.x, mode);
}
// The gating of the realizations of a list (one subset of the context).
__global__ __launch_bounds__(GATE_THREADS) void k_gate_list(const RealDesc* __restrict__ D, RealLive* __restrict__ L,
                                                            const int32_t* __restrict__ rids, uint32_t mode) {
  gate_workgroup<GATE_THREADS>(D, L, rids[blockIdx.x], mode);
}
#endif

// One step of a single-realization context with the previous step's gating
// folded in (cpu/RunSimulation.cpp:650-652 then :591-648).  Workgroup 0 gates
// on the state this step reads, which is the previous step's result, exactly
// as k_gate would after that step, then updates the elements that carry
// channel flux: only their rates read the channel states.  The other
// workgroups sweep the tiles without those elements.  cur, the state buffer
// this step reads, is an argument because workgroup 0 flips the live index
// while the others run.  With the loop body (PROTO, G_PROTOCOL) workgroup 0
// also changes Vm and channel states, which only the channel elements read,
// and may finish the realization: workgroups that see `done` skip their
// tile, the others write the buffer the finished realization never reads
// again.  A block starts only outside the ffwd phase, which the loop body
// can leave but not enter, so every step of it flips the live index.
// gmode 0 skips the gating (the first step of a block).
// curx: the live state index in bit 0; bit 1 set sweeps the tiles in the
// reverse order when the index is 1 (alternate steps), so that a step first
// reads the tiles the step before wrote last, while they may still be in the
// last-level cache.  The order changes no result.
template <int BS, bool NTG, bool NTS, bool COMPACT, bool LEAN, bool PROTO>
__device__ __forceinline__ void sweep_gated(const RealDesc* __restrict__ D, RealLive* __restrict__ L, int n_tiles,
                                            int curx, uint32_t gmode, const double* xo_at, double* xn_at) {
  const int cur = curx & 1;
  int b = xcd_tile(blockIdx.x, gridDim.x);
  if (b > n_tiles) return;  // b = 0: the gating workgroup; b = 1..n_tiles: tile b - 1
  if (b > 0 && curx == 3) b = n_tiles + 1 - b;
  const RealDesc& d = D[0];
  const RealLive& l = L[0];
  const int skip = l.done | l.ffwd;
  // the state buffers: X[cur] -> X[cur ^ 1], or the caller's (xo_at, xn_at)
  // (a replayed segment whose steps each have a slot of their own)
  const Bufs bf = xo_at ? Bufs{gp(xo_at), gpw(xn_at)} : bufs(d.X[0], d.X[1], cur);
  scalars_here(fold(skip, bf.xo, bf.xn) ^ geo_fold<COMPACT>(d));
  if (b == 0) {
    if (gmode) gate_workgroup<BS, LEAN, PROTO>(D, L, 0, gmode);
    __syncthreads();  // the new channel states, written by this workgroup, before the flux reads them
    if (skip) return;
    for (int k = threadIdx.x; k < d.n_celem; k += BS) {
      const int j = gp(d.celem)[k];
      const Geo g = load_geo<NTG, COMPACT, -1>(d, j);
      const Loc c = load_loc<-1>(d, bf.xo, j, g);
      step_elem<NTS, -1>(d, l, bf, j, g, c);
    }
    return;
  }
  if (skip) return;
  sweep_tile_lds<BS, NTG, NTS, COMPACT, true>(d, l, bf, b - 1);
}
// Register budgets as for the sweep.  The gating workgroup's latency is on
// the critical path of a small mesh's step, so the compiler's budget keeps
// all of its loads in one round trip; with a budget the LCC loads wait for
// the RyR update (LEAN), which is what lets the kernel fit 80 VGPRs.
template <int BS, bool NTG, bool NTS, bool COMPACT>
__global__ __launch_bounds__(BS) void k_sweep_gated(const RealDesc* __restrict__ D, RealLive* __restrict__ L,
                                                    int n_tiles, int cur, uint32_t gmode, const double* xo_at,
                                                    double* xn_at) {
  sweep_gated<BS, NTG, NTS, COMPACT, false, false>(D, L, n_tiles, cur, gmode, xo_at, xn_at);
}
template <int BS, bool NTG, bool NTS, bool COMPACT, int W>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(W))) void k_sweep_gated_w(
    const RealDesc* __restrict__ D, RealLive* __restrict__ L, int n_tiles, int cur, uint32_t gmode,
    const double* xo_at, double* xn_at) {
  sweep_gated<BS, NTG, NTS, COMPACT, true, false>(D, L, n_tiles, cur, gmode, xo_at, xn_at);
}
// with the loop body (SRS_STEP_PROTOCOL), compiler's register budget only
template <int BS, bool NTG, bool NTS, bool COMPACT>
__global__ __launch_bounds__(BS) void k_sweep_gated_p(const RealDesc* __restrict__ D, RealLive* __restrict__ L,
                                                      int n_tiles, int cur, uint32_t gmode, const double* xo_at,
                                                      double* xn_at) {
  sweep_gated<BS, NTG, NTS, COMPACT, false, true>(D, L, n_tiles, cur, gmode, xo_at, xn_at);
}

// ------------------------------------------------------------ kernel choice
// The sweep kernels are instantiated in five compilation units of this file
// (SRS_UNIT 0: host code and the gating kernel; 1: the sweeps; 2, 3, 4: the
// gated sweeps of 128-, 256- and 512-element tiles), which the build
// compiles in parallel.

using SweepFn = void (*)(const RealDesc*, const RealLive*, const int2*, int, const int*, int);
using GatedFn = void (*)(const RealDesc*, RealLive*, int, int, uint32_t, const double*, double*);
SweepFn pick_sweep(const srs_ctx* c);
template <int BS>
GatedFn pick_gated_bs(const srs_ctx* c, bool proto);

#if SRS_UNIT == 1
// The sweep kernel for a tile size, cache hints, register budget (0, 6 or 8
// wavefronts per SIMD) and LDS staging (single-realization groups only).
template <int BS, bool CMP, int KG, int W>
SweepFn pick_w(int ntg, int nts, int lds) {
  if constexpr (KG == 1) {
    if (lds) {
      if constexpr (W == 0) {
        if (ntg) return nts ? k_sweep_lds<BS, true, true, CMP> : k_sweep_lds<BS, true, false, CMP>;
        return nts ? k_sweep_lds<BS, false, true, CMP> : k_sweep_lds<BS, false, false, CMP>;
      } else {
        if (ntg) return nts ? k_sweep_lds_w<BS, true, true, CMP, W> : k_sweep_lds_w<BS, true, false, CMP, W>;
        return nts ? k_sweep_lds_w<BS, false, true, CMP, W> : k_sweep_lds_w<BS, false, false, CMP, W>;
      }
    }
  }
  if constexpr (W == 0) {
    if (ntg) return nts ? k_sweep<BS, true, true, CMP, KG> : k_sweep<BS, true, false, CMP, KG>;
    return nts ? k_sweep<BS, false, true, CMP, KG> : k_sweep<BS, false, false, CMP, KG>;
  } else {
    if (ntg) return nts ? k_sweep_w<BS, true, true, CMP, KG, W> : k_sweep_w<BS, true, false, CMP, KG, W>;
    return nts ? k_sweep_w<BS, false, true, CMP, KG, W> : k_sweep_w<BS, false, false, CMP, KG, W>;
  }
}

template <int BS, bool CMP, int KG>
SweepFn pick3(int ntg, int nts, int occ, int lds) {
  if (occ == 8) return pick_w<BS, CMP, KG, 8>(ntg, nts, lds);
  if (occ == 6) return pick_w<BS, CMP, KG, 6>(ntg, nts, lds);
  return pick_w<BS, CMP, KG, 0>(ntg, nts, lds);
}

template <bool CMP, int KG>
SweepFn pick_tile(const srs_ctx* c) {
  switch (c->tile) {
    case 128: return pick3<128, CMP, KG>(c->nt_geometry, c->nt_state, c->occ, c->lds);
    case 512: return pick3<512, CMP, KG>(c->nt_geometry, c->nt_state, c->occ, c->lds);
    default: return pick3<256, CMP, KG>(c->nt_geometry, c->nt_state, c->occ, c->lds);
  }
}

template <int KG>
SweepFn pick_kg(const srs_ctx* c) { return c->compact_all ? pick_tile<true, KG>(c) : pick_tile<false, KG>(c); }

SweepFn pick_sweep(const srs_ctx* c) {
  return c->group == 4 ? pick_kg<4>(c) : c->group == 2 ? pick_kg<2>(c) : pick_kg<1>(c);
}
#endif  // SRS_UNIT == 1

#if SRS_UNIT >= 2
template <int BS, bool CMP, int W>
GatedFn pick_gated_w(int ntg, int nts) {
  if constexpr (W == 0) {
    if (ntg) return nts ? k_sweep_gated<BS, true, true, CMP> : k_sweep_gated<BS, true, false, CMP>;
    return nts ? k_sweep_gated<BS, false, true, CMP> : k_sweep_gated<BS, false, false, CMP>;
  } else {
    if (ntg) return nts ? k_sweep_gated_w<BS, true, true, CMP, W> : k_sweep_gated_w<BS, true, false, CMP, W>;
    return nts ? k_sweep_gated_w<BS, false, true, CMP, W> : k_sweep_gated_w<BS, false, false, CMP, W>;
  }
}

template <int BS, bool CMP>
GatedFn pick_gated_p(int ntg, int nts) {
  if (ntg) return nts ? k_sweep_gated_p<BS, true, true, CMP> : k_sweep_gated_p<BS, true, false, CMP>;
  return nts ? k_sweep_gated_p<BS, false, true, CMP> : k_sweep_gated_p<BS, false, false, CMP>;
}

template <int BS, bool CMP>
GatedFn pick_gated_cmp(const srs_ctx* c, bool proto) {
  if (proto) return pick_gated_p<BS, CMP>(c->nt_geometry, c->nt_state);
  return c->occ == 6 ? pick_gated_w<BS, CMP, 6>(c->nt_geometry, c->nt_state)
                     : pick_gated_w<BS, CMP, 0>(c->nt_geometry, c->nt_state);
}

template <int BS>
GatedFn pick_gated_bs(const srs_ctx* c, bool proto) {
  return c->compact_all ? pick_gated_cmp<BS, true>(c, proto) : pick_gated_cmp<BS, false>(c, proto);
}
// one tile size per unit: 128 in unit 2, 256 in unit 3, 512 in unit 4
template GatedFn pick_gated_bs<SRS_UNIT == 2 ? 128 : SRS_UNIT == 3 ? 256 : 512>(const srs_ctx*, bool);
#endif  // SRS_UNIT >= 2

#if SRS_UNIT == 0
// Gathers Ca at the channel elements of one realization.
__global__ void k_channel_ca(const RealDesc* __restrict__ D, const RealLive* __restrict__ L, int r, double* out) {
  const RealDesc& d = D[r];
  const double* Ca = d.X[L[r].cur];  // Ca of element e at Ca[cq(d, e)]
  for (int i = threadIdx.x; i < d.n_ryr; i += blockDim.x) {
    out[i] = Ca[cq(d, d.ryr_ele[i])];
    out[d.n_ryr + i] = Ca[cq(d, d.ryr_jsr[i])];
  }
  for (int i = threadIdx.x; i < d.n_lcc; i += blockDim.x) out[2 * d.n_ryr + i] = Ca[cq(d, d.lcc_ele[i])];
}


// Moves one realization's state between the single-realization layout
// (np quads, TnC[np], CSQ[np]) and its slot of a group buffer.
__global__ void k_slot_copy(const double* __restrict__ src, double* __restrict__ dst, size_t np, int K, int slot,
                            int to_group) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 6 * np) return;
  size_t g;
  if (i < 4 * np) g = 4 * gidx(i / 4, K, slot) + i % 4;
  else if (i < 5 * np) g = 4 * (size_t)K * np + gidx(i - 4 * np, K, slot);
  else g = 5 * (size_t)K * np + gidx(i - 5 * np, K, slot);
  if (to_group) dst[g] = src[i];
  else dst[i] = src[g];
}

// ----------------------------------------------------------------- host side

// Rounds the significand to `bits` bits (class key of a coefficient).
double quantize(double a, int bits) {
  if (a == 0.0 || !std::isfinite(a)) return a;
  int e;
  const double m = std::frexp(a, &e);
  return std::ldexp(std::nearbyint(std::ldexp(m, bits)), e - bits);
}

// Builds the compact operator layout: neighbours as int16 offsets, and each
// row as a class row plus int16 multiples of the class row's ulps.  Leaves
// hm.compact false unless every row of the mesh reconstructs exactly.
void build_compact(HostMesh& hm) {
  const int n = hm.n, np = hm.np;
  for (int q = 0; q < 4; ++q)
    for (int k = 0; k < n; ++k) {
      const long o = (long)hm.nbr[(size_t)q * np + k] - k;
      if (o < -32767 || o > 32767) return;
    }
  std::map<std::array<double, 5>, int> classes;
  std::vector<int> row_class(n);
  for (int k = 0; k < n; ++k) {
    std::array<double, 5> key;
    for (int q = 0; q < 5; ++q) key[q] = quantize(hm.coef[(size_t)q * np + k], 40);
    auto it = classes.find(key);
    if (it == classes.end()) {
      if (classes.size() >= 65535) return;
      it = classes.emplace(key, (int)classes.size()).first;
    }
    row_class[k] = it->second;
  }
  const int nc = (int)classes.size();
  std::vector<double> tab(10 * (size_t)nc);
  for (auto& kv : classes)
    for (int q = 0; q < 5; ++q) {
      const double b = kv.first[q];
      tab[10 * (size_t)kv.second + q] = b;
      int e;
      std::frexp(b, &e);
      tab[10 * (size_t)kv.second + 5 + q] = (b == 0.0) ? 0.0 : std::ldexp(1.0, e - 53);  // ulp of b
    }
  std::vector<int16_t> off(4 * (size_t)np, 0), dq(5 * (size_t)np, 0);
  std::vector<uint16_t> cls(np, 0);
  for (int k = 0; k < n; ++k) {
    for (int q = 0; q < 4; ++q) off[(size_t)q * np + k] = (int16_t)(hm.nbr[(size_t)q * np + k] - k);
    const double* t = &tab[10 * (size_t)row_class[k]];
    int16_t d[5];
    for (int q = 0; q < 5; ++q) {
      const double a = hm.coef[(size_t)q * np + k];
      if (t[5 + q] == 0.0) {
        if (!(a == t[q] && !std::signbit(a))) return;
        d[q] = 0;
        continue;
      }
      const double m = (a - t[q]) / t[5 + q];
      if (!(std::fabs(m) <= 32767.0) || m != std::nearbyint(m)) return;
      d[q] = (int16_t)m;
      if ((t[q] + (double)d[q] * t[5 + q]) != a) return;  // exact, as the kernel computes it
    }
    cls[k] = (uint16_t)row_class[k];
    for (int q = 0; q < 5; ++q) dq[(size_t)q * np + k] = d[q];
  }
  hm.off.swap(off);
  hm.dq.swap(dq);
  hm.cls.swap(cls);
  hm.tab.swap(tab);
  hm.n_class = nc;
  hm.compact = true;
}

// cpu/RunSimulation.cpp:85-88, the parts that depend on CaNSR only.
void serca_terms(double nsr, const double* P, double* out) {
  double K_sr = nsr / P[SRS_R_K_D_SR];
  K_sr = K_sr * K_sr;
  out[0] = 0.104217 + 17.923 * K_sr;
  out[1] = 1.75583e6 + K_sr * 7.61673e6;
  out[2] = 6.08463e11 + K_sr * 4.50544e11;
  out[3] = 9.17846e6 - 11478.2 * K_sr;
  out[4] = 0.329904 * K_sr;
}

int check_real(srs_ctx* c, int32_t r) {
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (!c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (r < 0 || r >= (int32_t)c->reals.size()) return srs::fail(SRS_E_ARG, "realization %d out of range", r);
  return SRS_OK;
}

void launch_sweep(srs_ctx* c) {
  if (c->grid_sweep == 0) return;  // every realization has finished
  // one realization: tile b is tile b of realization 0, no table lookups
  const int single = c->reals.size() == 1 ? 1 : 0;
  hipLaunchKernelGGL(pick_sweep(c), dim3(c->grid_sweep), dim3(c->tile), 0, c->stream, c->d_desc, c->d_live,
                     c->d_tiles, c->n_tiles, (const int*)c->d_groups, single);
}

// m steps of every subset, one subset after another: the sweep over the
// subset's run of tiles and the gating of its realizations, m times.
// Realizations are independent and each one's Philox counters follow its own
// step count, so the order changes no result.
void launch_subset_block(srs_ctx* c, int64_t m, uint32_t gmode) {
  for (const srs_ctx::Subset& sb : c->subsets)
    for (int64_t k = 0; k < m; ++k) {
      if (sb.nt > 0)
        hipLaunchKernelGGL(pick_sweep(c), dim3((sb.nt + NXCD - 1) / NXCD * NXCD), dim3(c->tile), 0, c->stream,
                           c->d_desc, c->d_live, c->d_tiles + sb.t0, sb.nt, (const int*)c->d_groups, 0);
      hipLaunchKernelGGL(k_gate_list, dim3((unsigned)sb.nr), dim3(GATE_THREADS), 0, c->stream, c->d_desc, c->d_live,
                         (const int32_t*)c->d_sub_rids + sb.r0, gmode);
    }
}

// Splits the current tile list into subsets: runs of whole groups (in list
// order) whose realizations' state buffers (two of 48 B per element each)
// take at most subset_mb MB; realizations whose groups have no tile left
// (all finished) are gated with the last subset, as k_gate would.
int plan_subsets(srs_ctx* c) {
  c->subsets.clear();
  c->subsets_valid = true;
  if (c->subset_mb <= 0 || c->reals.size() < 2) return SRS_OK;
  const size_t budget = (size_t)c->subset_mb << 20;
  const int KG = c->group;
  std::vector<int32_t> rids;
  std::vector<char> listed(c->reals.size(), 0);
  srs_ctx::Subset cur{0, 0, 0, 0};
  size_t bytes = 0;
  const int n = (int)c->tiles_cur.size();
  for (int t = 0; t < n;) {
    const int g = c->tiles_cur[t].x;
    int t1 = t;
    while (t1 < n && c->tiles_cur[t1].x == g) ++t1;
    size_t gb = 0;
    for (int s = 0; s < KG; ++s) {
      const int r = c->groups[(size_t)KG * g + s];
      if (r >= 0) gb += 2 * 48 * (size_t)c->meshes[c->reals[r].mesh]->np;
    }
    if (cur.nt > 0 && bytes + gb > budget) {
      c->subsets.push_back(cur);
      cur = srs_ctx::Subset{t, 0, (int)rids.size(), 0};
      bytes = 0;
    }
    for (int s = 0; s < KG; ++s) {
      const int r = c->groups[(size_t)KG * g + s];
      if (r >= 0 && !listed[r]) {
        listed[r] = 1;
        rids.push_back(r);
        cur.nr++;
      }
    }
    cur.nt += t1 - t;
    bytes += gb;
    t = t1;
  }
  for (size_t r = 0; r < c->reals.size(); ++r)
    if (!listed[r]) {
      rids.push_back((int32_t)r);
      cur.nr++;
    }
  if (cur.nt > 0 || cur.nr > 0) c->subsets.push_back(cur);
  if (c->subsets.size() < 2) {  // one subset is the whole context: the plain step is the same launches
    c->subsets.clear();
    return SRS_OK;
  }
  HIPCHK(hipMemcpy(c->d_sub_rids, rids.data(), rids.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  return SRS_OK;
}

int launch_step(srs_ctx* c, uint32_t gmode) {
  launch_sweep(c);
  hipLaunchKernelGGL(k_gate, dim3((unsigned)c->reals.size()), dim3(GATE_THREADS), 0, c->stream, c->d_desc,
                     c->d_live, gmode);
  return SRS_OK;
}

// One step through k_sweep_gated: tile b - 1 in workgroup b, workgroup 0
// gates (gmode) and updates the channel elements.  proto picks the kernel
// with the loop body.
void launch_gated(srs_ctx* c, int cur, uint32_t gmode, bool proto, const double* xo_at, double* xn_at) {
  const GatedFn f = c->tile == 128 ? pick_gated_bs<128>(c, proto)
                    : c->tile == 512 ? pick_gated_bs<512>(c, proto) : pick_gated_bs<256>(c, proto);
  const int grid = (c->n_tiles + 1 + NXCD - 1) / NXCD * NXCD;
  hipLaunchKernelGGL(f, dim3(grid), dim3(c->tile), 0, c->stream, c->d_desc, c->d_live, c->n_tiles,
                     cur | (c->alt_order ? 2 : 0), gmode, xo_at, xn_at);
}

// m steps through k_sweep_gated: the first without gating (nothing is
// pending when a block starts), each later one gating the step before it,
// then k_gate for the last step, so that a block leaves no gating pending.
// cur is the live state index when the block starts; gmode is the step's
// gating mode (G_AFTER_SWEEP, with G_PROTOCOL for the loop body).
// The gating of a single realization's last gated step (k_gate alone).
void launch_gate_single(srs_ctx* c, uint32_t gmode) {
  hipLaunchKernelGGL(k_gate, dim3(1), dim3(GATE_THREADS), 0, c->stream, c->d_desc, c->d_live, gmode);
}

void launch_gated_block(srs_ctx* c, int64_t m, int cur, uint32_t gmode) {
  const bool proto = (gmode & G_PROTOCOL) != 0;
  for (int64_t k = 0; k < m; ++k) launch_gated(c, cur ^ (int)(k & 1), k ? gmode : 0u, proto);
  hipLaunchKernelGGL(k_gate, dim3(1), dim3(GATE_THREADS), 0, c->stream, c->d_desc, c->d_live, gmode);
}

// Captured step blocks hold the sweep's tile list and grid; they are dropped
// when either changes.
int drop_graphs(srs_ctx* c) {
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
  return SRS_OK;
}

void set_tiles(srs_ctx* c, int2* d, int n) {
  c->d_tiles = d;
  c->n_tiles = n;
  c->grid_sweep = (n + NXCD - 1) / NXCD * NXCD;
  c->subsets_valid = false;
}

// The full tile list again, before the host changes live state (a finished
// realization may be restarted).
int tiles_full(srs_ctx* c) {
  if (!c->tiles_compacted) return SRS_OK;
  int rc = drop_graphs(c);
  if (rc) return rc;
  set_tiles(c, c->d_tiles_all, (int)c->tiles_all.size());
  c->tiles_cur = c->tiles_all;
  c->tiles_compacted = false;
  return SRS_OK;
}

// A tile whose realizations have all finished still costs the sweep about
// what a live one does (its workgroup is dispatched, reads the tile table and
// the done flags, and exits), so once such tiles make up an eighth of the
// list the sweep gets a list without them.  A finished realization stays
// finished until the host changes its live state, which puts the full list
// back (tiles_full).
int tiles_compact(srs_ctx* c) {
  const int n_real = (int)c->reals.size();
  if (n_real < 2 || !c->drop_finished) return SRS_OK;
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<RealLive> lv(n_real);
  HIPCHK(hipMemcpy(lv.data(), c->d_live, n_real * sizeof(RealLive), hipMemcpyDeviceToHost));
  const int KG = c->group;
  const int n_groups = (int)c->groups.size() / KG;
  std::vector<char> live(n_groups, 0);
  for (int g = 0; g < n_groups; ++g)
    for (int s = 0; s < KG; ++s) {
      const int r = c->groups[(size_t)KG * g + s];
      if (r >= 0 && !lv[r].done) live[g] = 1;
    }
  std::vector<int2> t;
  for (const int2& x : c->tiles_all)
    if (live[x.x]) t.push_back(x);
  if ((int64_t)t.size() == c->n_tiles || (int64_t)t.size() * 8 > (int64_t)c->n_tiles * 7) return SRS_OK;
  int rc = drop_graphs(c);
  if (rc) return rc;
  if (!t.empty()) HIPCHK(hipMemcpy(c->d_tiles_live, t.data(), t.size() * sizeof(int2), hipMemcpyHostToDevice));
  set_tiles(c, c->d_tiles_live, (int)t.size());
  c->tiles_cur = t;
  c->tiles_compacted = true;
  return SRS_OK;
}


// Copies one realization's state between a host buffer in the single
// realization layout (np quads, then the TnC and CSQ planes) and its slot of
// the group buffer: contiguous through a device staging buffer, then a
// scatter (or gather) kernel for the slot.
int slot_copy(srs_ctx* c, const RealDesc& d, double* dev, double* host, bool to_dev) {
  const size_t np = d.np, n6 = 6 * np;
  const hipMemcpyKind kind = to_dev ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  if (d.kst == 1) {
    if (to_dev) HIPCHK(hipMemcpy(dev, host, n6 * sizeof(double), kind));
    else HIPCHK(hipMemcpy(host, dev, n6 * sizeof(double), kind));
    return SRS_OK;
  }
  if (c->stage_len < n6) {
    if (c->stage) HIPCHK(hipFree(c->stage));
    c->stage = nullptr;
    c->stage_len = 0;
    HIPCHK(hipMalloc((void**)&c->stage, n6 * sizeof(double)));
    c->stage_len = n6;
  }
  const unsigned blocks = (unsigned)((n6 + 255) / 256);
  if (to_dev) {
    HIPCHK(hipMemcpy(c->stage, host, n6 * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_slot_copy, dim3(blocks), dim3(256), 0, c->stream, (const double*)c->stage, dev, np, d.kst,
                       d.slot, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
  } else {
    hipLaunchKernelGGL(k_slot_copy, dim3(blocks), dim3(256), 0, c->stream, (const double*)dev, c->stage, np, d.kst,
                       d.slot, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(host, c->stage, n6 * sizeof(double), hipMemcpyDeviceToHost));
  }
  return SRS_OK;
}

int live_rw(srs_ctx* c, int32_t r, RealLive* lv, bool write) {
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (write) {
    int rc = tiles_full(c);
    if (rc) return rc;
  }
  if (write) HIPCHK(hipMemcpy(c->d_live + r, lv, sizeof(*lv), hipMemcpyHostToDevice));
  else HIPCHK(hipMemcpy(lv, c->d_live + r, sizeof(*lv), hipMemcpyDeviceToHost));
  return SRS_OK;
}

#endif  // SRS_UNIT == 0

}  // namespace srs_dev

#if SRS_UNIT == 0
// Replay schedules of srs_explicit_step_replay_static: file-scope device
// memory that every GPU the library runs on holds until the process ends
// (8 MiB: 2,097,152 RyR states, steps x RyRs).  In host code `scratch` names
// a host-side stand-in; the device address of each GPU's copy comes from
// hipGetSymbolAddress with that GPU current and is the only form used.
constexpr size_t kStaticSchedInts = (size_t)1 << 21;
static __device__ __align__(256) int32_t scratch[kStaticSchedInts];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
int32_t* g_scratch_ptr[kMaxDevices] = {};

// Once per device, with that device current (srs_ctx_create sets it).
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return srs::fail(SRS_E_ARG, "device %d beyond %d", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return SRS_OK;
  void* p = nullptr;
  HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = (int32_t*)p;
  return SRS_OK;
}

int32_t* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}
}  // namespace

extern "C" {

int srs_ctx_create(int device, srs_ctx** out) {
  srs::clear_error();
  if (!out) return srs::fail(SRS_E_ARG, "null out");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return srs::fail(SRS_E_NODEV, "no HIP device");
  if (device < 0 || device >= n) return srs::fail(SRS_E_ARG, "device %d out of range (%d)", device, n);
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return srs::fail(SRS_E_NODEV, "device %d is %s; this library is built for gfx950", device, prop.gcnArchName);
  HIPCHK(hipSetDevice(device));
  int rc = scratch_setup(device);
  if (rc) return rc;
  auto c = new srs_ctx();
  c->device = device;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return srs::fail(SRS_E_HIP, "stream/event creation failed");
  }
  *out = c;
  return SRS_OK;
}

int srs_ctx_tune(srs_ctx* c, int32_t key, int32_t value) {
  srs::clear_error();
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  switch (key) {
    case SRS_TUNE_TILE:
      if (value != 128 && value != 256 && value != 512) return srs::fail(SRS_E_ARG, "tile must be 128, 256 or 512");
      if (c->finalized) return srs::fail(SRS_E_STATE, "tile must be set before srs_ctx_finalize");
      c->tile = value;
      return SRS_OK;
    case SRS_TUNE_PLANE_PAD:
      if (value < 0 || value % 64 || value > 1 << 20) return srs::fail(SRS_E_ARG, "plane pad must be a multiple of 64");
      if (c->layout_fixed) return srs::fail(SRS_E_STATE, "plane pad must be set before the first mesh");
      c->plane_pad = value;
      return SRS_OK;
    case SRS_TUNE_COMPACT:
      if (c->layout_fixed) return srs::fail(SRS_E_STATE, "compact layout must be chosen before the first mesh");
      c->compact = value != 0;
      return SRS_OK;
    // kernel keys: captured step graphs hold the old kernels, so they are dropped
    case SRS_TUNE_NT_GEOMETRY:
      c->nt_geometry = value != 0;
      return c->finalized ? drop_graphs(c) : SRS_OK;
    case SRS_TUNE_OCCUPANCY:
      if (value != 0 && value != 6 && value != 8) return srs::fail(SRS_E_ARG, "occupancy must be 0, 6 or 8");
      c->occ = value;
      return c->finalized ? drop_graphs(c) : SRS_OK;
    case SRS_TUNE_NT_STATE:
      c->nt_state = value != 0;
      return c->finalized ? drop_graphs(c) : SRS_OK;
    case SRS_TUNE_LDS:
      c->lds = value != 0;
      return c->finalized ? drop_graphs(c) : SRS_OK;
    case SRS_TUNE_GATED_SWEEP:
      c->gated = value != 0;
      return c->finalized ? drop_graphs(c) : SRS_OK;
    case SRS_TUNE_DROP_FINISHED:
      c->drop_finished = value != 0;
      return c->finalized ? tiles_full(c) : SRS_OK;
    case SRS_TUNE_ADJ_WAVES:
      if (value != 0 && value != 4) return srs::fail(SRS_E_ARG, "adjoint register budget must be 0 or 4");
      c->adj_waves = value;
      return SRS_OK;
    case SRS_TUNE_ADJ_W:
      if (value < 0 || value > 2) return srs::fail(SRS_E_ARG, "adjoint w form must be 0, 1 or 2");
      c->adj_w = value;
      return SRS_OK;
    case SRS_TUNE_ALT_ORDER:
      c->alt_order = value != 0;
      return c->finalized ? drop_graphs(c) : SRS_OK;
    case SRS_TUNE_SUBSET_MB:
      if (value < 0 || value > (1 << 20)) return srs::fail(SRS_E_ARG, "subset size must be 0..2^20 MB");
      c->subset_mb = value;
      c->subsets_valid = false;
      return c->finalized ? drop_graphs(c) : SRS_OK;
    case SRS_TUNE_ADJ_LEVELS:
      if (value < 0 || value > 6) return srs::fail(SRS_E_ARG, "adjoint checkpoint levels must be 0..6");
      c->adj_levels = value;
      return SRS_OK;
    case SRS_TUNE_GROUP:
      if (value != 1 && value != 2 && value != 4) return srs::fail(SRS_E_ARG, "group must be 1, 2 or 4");
      if (c->finalized) return srs::fail(SRS_E_STATE, "group must be set before srs_ctx_finalize");
      c->group = value;
      return SRS_OK;
    default: return srs::fail(SRS_E_ARG, "unknown tuning key %d", key);
  }
}

int srs_ctx_record_layout(srs_ctx* c, int32_t gpu_layout) {
  srs::clear_error();
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (c->finalized) return srs::fail(SRS_E_STATE, "the record layout is fixed at srs_ctx_finalize");
  if (gpu_layout != 0 && gpu_layout != 1) return srs::fail(SRS_E_ARG, "gpu_layout must be 0 or 1");
  c->gpu_layout = gpu_layout;
  return SRS_OK;
}

int srs_ctx_extensions(srs_ctx* c, int32_t r, const srs_extensions* e) {
  srs::clear_error();
  if (!c || !e) return srs::fail(SRS_E_ARG, "null argument");
  if (r < 0 || r >= (int32_t)c->reals.size()) return srs::fail(SRS_E_ARG, "realization %d out of range", r);
  if (e->jsr_diffusion != 0 && e->jsr_diffusion != 1) return srs::fail(SRS_E_ARG, "jsr_diffusion must be 0 or 1");
  if (!(e->b_sr_sites >= 0.0) || (e->b_sr_sites > 0.0 && !(e->k_d_sr_sites > 0.0)))
    return srs::fail(SRS_E_ARG, "SR sites need b >= 0 and, when b > 0, a dissociation constant > 0");
  HostReal& hr = c->reals[r];
  hr.ext = *e;
  if (!c->finalized) return SRS_OK;
  hr.desc.jsr_diff = e->jsr_diffusion;
  hr.desc.b_sr = e->b_sr_sites;
  hr.desc.k_sr = e->k_d_sr_sites;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->d_desc + r, &hr.desc, sizeof(hr.desc), hipMemcpyHostToDevice));
  return SRS_OK;
}

int srs_ctx_destroy(srs_ctx* c) {
  if (!c) return SRS_OK;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  for (void* p : c->allocs) (void)hipFree(p);
  if (c->stage) (void)hipFree(c->stage);
  (void)hipEventDestroy(c->ev0);
  (void)hipEventDestroy(c->ev1);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return SRS_OK;
}

int srs_ctx_add_mesh(srs_ctx* c, const srs_mesh_desc* m, int32_t* mesh_id) {
  srs::clear_error();
  if (!c || !m || !mesh_id) return srs::fail(SRS_E_ARG, "null argument");
  if (c->finalized) return srs::fail(SRS_E_STATE, "context already finalized");
  const int32_t n = m->n_ele;
  if (n <= 0 || !m->aij || !m->nij || !m->domain || !m->tropc || !m->V0 || !m->tt_area)
    return srs::fail(SRS_E_ARG, "mesh descriptor incomplete");
  for (int64_t k = 0; k < 4LL * n; ++k)
    if (m->nij[k] < 0 || m->nij[k] >= n) return srs::fail(SRS_E_ARG, "nij[%lld] = %d out of range", (long long)k, m->nij[k]);
  auto hm = std::make_unique<HostMesh>();
  hm->n = n;
  c->layout_fixed = true;
  hm->np = (n + MAX_TILE - 1) / MAX_TILE * MAX_TILE + c->plane_pad;
  const int np = hm->np;
  // element order inside the context
  if (m->order == SRS_ORDER_RCM) {
    srs::rcm_order(n, m->nij, hm->perm);
  } else if (m->order == SRS_ORDER_CUSTOM) {
    if (!m->perm) return srs::fail(SRS_E_ARG, "SRS_ORDER_CUSTOM needs a permutation");
    hm->perm.assign(m->perm, m->perm + n);
    std::vector<uint8_t> seen(n, 0);
    for (int i = 0; i < n; ++i) {
      if (hm->perm[i] < 0 || hm->perm[i] >= n || seen[hm->perm[i]]) return srs::fail(SRS_E_ARG, "perm is not a permutation");
      seen[hm->perm[i]] = 1;
    }
  } else {
    hm->perm.resize(n);
    std::iota(hm->perm.begin(), hm->perm.end(), 0);
  }
  hm->inv.assign(n, 0);
  for (int i = 0; i < n; ++i) hm->inv[hm->perm[i]] = i;
  hm->nbr.assign(4 * (size_t)np, 0);
  hm->coef.assign(5 * (size_t)np, 0.0);
  hm->flags.assign(np, 0);
  hm->V0.assign(np, 1.0);
  hm->tt.assign(np, 0.0);
  bool any_br = false;
  if (m->br)
    for (int s : {srs::CA, srs::DYE, srs::ATP, srs::CMDN})
      for (int i = 0; i < n; ++i) any_br |= (m->br[(size_t)s * n + i] != 0.0);
  if (any_br) hm->br.assign(4 * (size_t)np, 0.0);
  for (int k = 0; k < np; ++k) {
    if (k >= n) {  // padding rows point at themselves
      for (int q = 0; q < 4; ++q) hm->nbr[(size_t)q * np + k] = k < n ? k : 0;
      continue;
    }
    const int i = hm->perm[k];
    for (int q = 0; q < 5; ++q) hm->coef[(size_t)q * np + k] = m->aij[5 * (size_t)i + q];
    for (int q = 0; q < 4; ++q) hm->nbr[(size_t)q * np + k] = hm->inv[m->nij[4 * (size_t)i + q]];
    uint8_t f = 0;
    if (m->domain[i] != 0) f |= F_JSR;
    if (m->tropc[i]) f |= F_TROPC;
    if (m->tt_area[i] != 0.0) f |= F_SL;
    hm->V0[k] = m->V0[i];
    hm->tt[k] = m->tt_area[i];
    if (any_br) {
      const int sp[4] = {srs::CA, srs::DYE, srs::ATP, srs::CMDN};
      bool nz = false;
      for (int q = 0; q < 4; ++q) {
        const double v = m->br[(size_t)sp[q] * n + i];
        hm->br[(size_t)q * np + k] = v;
        nz |= (v != 0.0);
      }
      if (nz) f |= F_BR;
    }
    hm->flags[k] = f;
  }
  if (c->compact) build_compact(*hm);
  *mesh_id = (int32_t)c->meshes.size();
  c->meshes.push_back(std::move(hm));
  return SRS_OK;
}

int srs_ctx_add_realization(srs_ctx* c, int32_t mesh_id, const srs_chan_desc* ch, const int32_t ints[SRS_N_INT],
                            const double reals[SRS_N_REAL], uint64_t seed_key, uint64_t stream, int32_t* real_id) {
  srs::clear_error();
  if (!c || !ch || !ints || !reals || !real_id) return srs::fail(SRS_E_ARG, "null argument");
  if (c->finalized) return srs::fail(SRS_E_STATE, "context already finalized");
  if (mesh_id < 0 || mesh_id >= (int32_t)c->meshes.size()) return srs::fail(SRS_E_ARG, "bad mesh id %d", mesh_id);
  HostMesh& hm = *c->meshes[mesh_id];
  if (ch->n_ryr < 0 || ch->n_ryr > MAX_RYR || ch->n_lcc < 0 || ch->n_lcc > MAX_LCC)
    return srs::fail(SRS_E_ARG, "channel counts %d/%d exceed %d/%d", ch->n_ryr, ch->n_lcc, MAX_RYR, MAX_LCC);
  if (!(reals[SRS_R_T_STEP] > 0)) return srs::fail(SRS_E_ARG, "t_step must be > 0");
  HostReal hr;
  hr.mesh = mesh_id;
  memcpy(hr.ints, ints, sizeof(hr.ints));
  memcpy(hr.reals, reals, sizeof(hr.reals));
  hr.key = seed_key;
  hr.stream = stream;
  std::vector<std::pair<int32_t, int32_t>> entries;  // (dev element, code) in reference order
  for (int i = 0; i < ch->n_ryr; ++i) {
    const int ce = ch->ryr_ele[i], je = ch->ryr_jsr[i];
    if (ce < 0 || ce >= hm.n || je < 0 || je >= hm.n) return srs::fail(SRS_E_ARG, "RyR %d element out of range", i);
    hr.ryr_ele.push_back(hm.inv[ce]);
    hr.ryr_jsr.push_back(hm.inv[je]);
    for (int q = 0; q < 4; ++q) {
      const int nb = ch->ryr_nb ? ch->ryr_nb[4 * i + q] : -1;
      if (nb >= ch->n_ryr || nb < -1) return srs::fail(SRS_E_ARG, "RyR %d neighbour %d out of range", i, nb);
      hr.ryr_nb.push_back(nb);
    }
    entries.push_back({hm.inv[ce], 2 * i});
    entries.push_back({hm.inv[je], 2 * i + 1});
  }
  for (int i = 0; i < ch->n_lcc; ++i) {
    const int e = ch->lcc_ele[i];
    if (e < 0 || e >= hm.n) return srs::fail(SRS_E_ARG, "LCC %d element out of range", i);
    hr.lcc_ele.push_back(hm.inv[e]);
    entries.push_back({hm.inv[e], -(i + 1)});
  }
  if (ints[SRS_I_INIT_RYR] != -1 && (ints[SRS_I_INIT_RYR] < 0 || ints[SRS_I_INIT_RYR] >= ch->n_ryr))
    return srs::fail(SRS_E_ARG, "ryr_open_init %d must be -1 or in [0, %d)", ints[SRS_I_INIT_RYR], ch->n_ryr);
  if (ints[SRS_I_FLAG_LCC] && (ints[SRS_I_LCC_INDEX] < 0 || ints[SRS_I_LCC_INDEX] >= ch->n_lcc))
    return srs::fail(SRS_E_ARG, "LCC index %d out of range", ints[SRS_I_LCC_INDEX]);
  std::stable_sort(entries.begin(), entries.end(),
                   [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) { return a.first < b.first; });
  for (size_t k = 0; k < entries.size(); ++k) {
    if (k == 0 || entries[k].first != entries[k - 1].first) {
      hr.celem.push_back(entries[k].first);
      hr.cstart.push_back((int32_t)k);
      hm.flags[entries[k].first] |= F_CHAN;
    }
    hr.centry.push_back(entries[k].second);
  }
  hr.cstart.push_back((int32_t)entries.size());
  if (!hm.br.empty()) {
    // The mesh's boundary rates were built with one set of D; every
    // realization on the mesh must start from it (srs_params_set may move it
    // later, and the rates follow).
    const double D[4] = {reals[SRS_R_D_CA], reals[SRS_R_D_DYE], reals[SRS_R_D_ATP], reals[SRS_R_D_CMDN]};
    if (!hm.br_d_set) {
      std::copy(D, D + 4, hm.br_d);
      hm.br_d_set = true;
    } else if (!std::equal(D, D + 4, hm.br_d)) {
      return srs::fail(SRS_E_ARG, "realizations on a mesh with boundary rates must share the D the rates were "
                                  "built with (D_Ca %g, D_Fluo %g, D_ATP %g, D_CaM %g)", hm.br_d[0], hm.br_d[1],
                       hm.br_d[2], hm.br_d[3]);
    }
  }
  *real_id = (int32_t)c->reals.size();
  c->reals.push_back(std::move(hr));
  return SRS_OK;
}

int srs_ctx_finalize(srs_ctx* c, int32_t record_capacity) {
  srs::clear_error();
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (c->finalized) return srs::fail(SRS_E_STATE, "already finalized");
  if (c->reals.empty()) return srs::fail(SRS_E_STATE, "no realizations");
  HIPCHK(hipSetDevice(c->device));
  int rc;
  if (c->tile == 0) {
    // 512-element tiles pay on meshes of a million elements and more (the
    // 3M-element dyad); smaller meshes keep more workgroups in flight with 256
    int n_max = 0;
    for (const HostReal& hr : c->reals) n_max = std::max(n_max, c->meshes[hr.mesh]->n);
    c->tile = n_max >= (1 << 20) ? 512 : 256;
  }
  c->compact_all = !c->meshes.empty();
  for (auto& hmp : c->meshes) c->compact_all &= hmp->compact;
  for (auto& hmp : c->meshes) {
    HostMesh& hm = *hmp;
    if (!c->compact_all) {
      if ((rc = dev_upload(c, &hm.d_nbr, hm.nbr))) return rc;
      if ((rc = dev_upload(c, &hm.d_coef, hm.coef))) return rc;
    } else {
      if ((rc = dev_upload(c, &hm.d_off, hm.off))) return rc;
      if ((rc = dev_upload(c, &hm.d_cls, hm.cls))) return rc;
      if ((rc = dev_upload(c, &hm.d_dq, hm.dq))) return rc;
      if ((rc = dev_upload(c, &hm.d_tab, hm.tab))) return rc;
    }
    if ((rc = dev_upload(c, &hm.d_flags, hm.flags))) return rc;
    if ((rc = dev_upload(c, &hm.d_V0, hm.V0))) return rc;
    if ((rc = dev_upload(c, &hm.d_tt, hm.tt))) return rc;
    if (!hm.br.empty() && (rc = dev_upload(c, &hm.d_br, hm.br))) return rc;
  }
  std::vector<RealDesc> descs;
  std::vector<RealLive> lives;
  std::vector<int2> tiles;
  // Realization groups: up to c->group realizations of one mesh, in the
  // order they were added, share interleaved state buffers and each sweep
  // tile of the group loads the operator rows once for all of them.
  const int KG = c->group;
  std::vector<int32_t> groups;     // [n_groups][KG], -1 pads
  std::vector<int32_t> real_group(c->reals.size()), real_slot(c->reals.size());
  std::vector<int32_t> group_mesh;
  for (size_t m = 0; m < c->meshes.size(); ++m) {
    int cur_g = -1, fill = KG;
    for (size_t r = 0; r < c->reals.size(); ++r) {
      if (c->reals[r].mesh != (int)m) continue;
      if (fill == KG) {
        cur_g = (int)group_mesh.size();
        group_mesh.push_back((int)m);
        groups.insert(groups.end(), KG, -1);
        fill = 0;
      }
      groups[(size_t)KG * cur_g + fill] = (int32_t)r;
      real_group[r] = cur_g;
      real_slot[r] = fill++;
    }
  }
  const int n_groups = (int)group_mesh.size();
  std::vector<double*> gx0(n_groups), gx1(n_groups);
  for (int gi = 0; gi < n_groups; ++gi) {
    const HostMesh& hm = *c->meshes[group_mesh[gi]];
    const size_t cnt = 6 * (size_t)KG * hm.np;
    if ((rc = dev_alloc(c, &gx0[gi], cnt)) || (rc = dev_alloc(c, &gx1[gi], cnt))) return rc;
    HIPCHK(hipMemset(gx0[gi], 0, cnt * sizeof(double)));
    HIPCHK(hipMemset(gx1[gi], 0, cnt * sizeof(double)));
    const int nt = (hm.n + c->tile - 1) / c->tile;
    for (int t = 0; t < nt; ++t) tiles.push_back(make_int2(gi, t));
  }
  for (size_t r = 0; r < c->reals.size(); ++r) {
    HostReal& hr = c->reals[r];
    HostMesh& hm = *c->meshes[hr.mesh];
    RealDesc d{};
    d.nbr = hm.d_nbr; d.coef = hm.d_coef; d.flags = hm.d_flags; d.V0 = hm.d_V0; d.tt = hm.d_tt; d.br = hm.d_br;
    d.br_mesh = hm.d_br;
    std::copy(hm.br_d, hm.br_d + 4, d.brd);
    d.off = hm.d_off; d.cls = hm.d_cls; d.dq = hm.d_dq; d.tab = hm.d_tab;
    d.n_class = hm.n_class;
    d.n = hm.n; d.np = hm.np;
    d.X[0] = gx0[real_group[r]]; d.X[1] = gx1[real_group[r]];
    hr.group = real_group[r];
    d.kst = KG;
    d.slot = real_slot[r];
    d.n_ryr = (int32_t)hr.ryr_ele.size();
    d.n_lcc = (int32_t)hr.lcc_ele.size();
    int32_t *p;
    if ((rc = dev_upload(c, &p, hr.ryr_ele))) return rc; d.ryr_ele = p;
    if ((rc = dev_upload(c, &p, hr.ryr_jsr))) return rc; d.ryr_jsr = p;
    if ((rc = dev_upload(c, &p, hr.ryr_nb))) return rc; d.ryr_nb = p;
    if ((rc = dev_upload(c, &p, hr.lcc_ele))) return rc; d.lcc_ele = p;
    if ((rc = dev_upload(c, &p, hr.celem))) return rc; d.celem = p;
    if ((rc = dev_upload(c, &p, hr.cstart))) return rc; d.cstart = p;
    if ((rc = dev_upload(c, &p, hr.centry))) return rc; d.centry = p;
    d.n_celem = (int32_t)hr.celem.size();
    std::vector<int32_t> zeros_r(d.n_ryr, 0), ones_l(d.n_lcc, 1), twos_l(d.n_lcc, 2);
    if ((rc = dev_upload(c, &p, zeros_r))) return rc; d.ryr_state = p;
    if ((rc = dev_upload(c, &p, ones_l))) return rc; d.lcc_state = p;
    if ((rc = dev_upload(c, &p, twos_l))) return rc; d.lccv_state = p;
    std::vector<double> zr(d.n_ryr + d.n_lcc, -1.0);
    double* pr;
    if ((rc = dev_upload(c, &pr, zr))) return rc; d.R = pr;
    d.ryr_out = hr.ints[SRS_I_FLAG_RYR_OUT] ? 1 : 0;
    d.gpu_layout = c->gpu_layout;
    d.rec_len = 7 + (d.gpu_layout ? d.n_ryr : 0) + (d.ryr_out ? d.n_ryr : 0);
    if (record_capacity > 0 && d.rec_len > 255)  // cpu/OutputFuncs.cpp:441-444
      return srs::fail(SRS_E_ARG, "states record of %d floats exceeds the 255-float record limit", d.rec_len);
    d.rec_cap = record_capacity;
    if (record_capacity > 0) {
      float* fr;
      if ((rc = dev_alloc(c, &fr, (size_t)record_capacity * d.rec_len))) return rc;
      d.rec = fr;
    }
    d.key = hr.key;
    d.stream = hr.stream;
    d.jsr_diff = hr.ext.jsr_diffusion;
    d.b_sr = hr.ext.b_sr_sites;
    d.k_sr = hr.ext.k_d_sr_sites;
    memcpy(d.P, hr.reals, sizeof(d.P));
    memcpy(d.I, hr.ints, sizeof(d.I));
    hr.desc = d;
    descs.push_back(d);
    RealLive lv{};
    lv.vm = hr.reals[SRS_R_DEFAULT_VM];
    lv.ca_nsr = hr.reals[SRS_R_DEFAULT_CA_NSR];
    serca_terms(lv.ca_nsr, hr.reals, lv.serca);
    lives.push_back(lv);
  }
  c->n_tiles = (int)tiles.size();
  c->grid_sweep = (c->n_tiles + NXCD - 1) / NXCD * NXCD;
  if ((rc = dev_upload(c, &c->d_desc, descs))) return rc;
  if ((rc = dev_upload(c, &c->d_live, lives))) return rc;
  if ((rc = dev_upload(c, &c->d_tiles, tiles))) return rc;
  if ((rc = dev_upload(c, &c->d_groups, groups))) return rc;
  c->tiles_all = tiles;
  c->tiles_cur = tiles;
  c->d_tiles_all = c->d_tiles;
  if (c->reals.size() > 1 && (rc = dev_alloc(c, &c->d_sub_rids, c->reals.size()))) return rc;
  c->groups = groups;
  if (c->reals.size() > 1 && (rc = dev_alloc(c, &c->d_tiles_live, tiles.size()))) return rc;
  // meshes' channel flags were updated by add_realization; refresh them
  for (auto& hmp : c->meshes)
    HIPCHK(hipMemcpy(hmp->d_flags, hmp->flags.data(), hmp->flags.size(), hipMemcpyHostToDevice));
  HIPCHK(hipDeviceSynchronize());
  c->finalized = true;
  return SRS_OK;
}


int srs_state_set(srs_ctx* c, int32_t r, const double* X) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (!X) return srs::fail(SRS_E_ARG, "null state");
  const HostReal& hr = c->reals[r];
  const HostMesh& hm = *c->meshes[hr.mesh];
  std::vector<double> buf(6 * (size_t)hm.np, 0.0);
  for (int s = 0; s < 6; ++s)
    for (int i = 0; i < hm.n; ++i) buf[xoff(s, hm.inv[i], hm.np)] = X[(size_t)s * hm.n + i];
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if ((rc = slot_copy(c, hr.desc, hr.desc.X[0], buf.data(), true))) return rc;
  return slot_copy(c, hr.desc, hr.desc.X[1], buf.data(), true);
}

int srs_state_get(srs_ctx* c, int32_t r, double* X) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (!X) return srs::fail(SRS_E_ARG, "null state");
  const HostReal& hr = c->reals[r];
  const HostMesh& hm = *c->meshes[hr.mesh];
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  RealLive lv;
  HIPCHK(hipMemcpy(&lv, c->d_live + r, sizeof(lv), hipMemcpyDeviceToHost));
  std::vector<double> buf(6 * (size_t)hm.np);
  if ((rc = slot_copy(c, hr.desc, hr.desc.X[lv.cur], buf.data(), false))) return rc;
  for (int s = 0; s < 6; ++s)
    for (int i = 0; i < hm.n; ++i) X[(size_t)s * hm.n + i] = buf[xoff(s, hm.inv[i], hm.np)];
  return SRS_OK;
}

int srs_channels_set(srs_ctx* c, int32_t r, const int32_t* ryr_state, const int32_t* lcc_state,
                     const int32_t* lccv_state, const double* R) {
  int rc = check_real(c, r);
  if (rc) return rc;
  const RealDesc& d = c->reals[r].desc;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (ryr_state && d.n_ryr) HIPCHK(hipMemcpy(d.ryr_state, ryr_state, d.n_ryr * sizeof(int32_t), hipMemcpyHostToDevice));
  if (lcc_state && d.n_lcc) HIPCHK(hipMemcpy(d.lcc_state, lcc_state, d.n_lcc * sizeof(int32_t), hipMemcpyHostToDevice));
  if (lccv_state && d.n_lcc) HIPCHK(hipMemcpy(d.lccv_state, lccv_state, d.n_lcc * sizeof(int32_t), hipMemcpyHostToDevice));
  if (R && d.n_ryr + d.n_lcc) HIPCHK(hipMemcpy(d.R, R, (d.n_ryr + d.n_lcc) * sizeof(double), hipMemcpyHostToDevice));
  return SRS_OK;
}

int srs_channels_get(srs_ctx* c, int32_t r, int32_t* ryr_state, int32_t* lcc_state, int32_t* lccv_state, double* R) {
  int rc = check_real(c, r);
  if (rc) return rc;
  const RealDesc& d = c->reals[r].desc;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (ryr_state && d.n_ryr) HIPCHK(hipMemcpy(ryr_state, d.ryr_state, d.n_ryr * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (lcc_state && d.n_lcc) HIPCHK(hipMemcpy(lcc_state, d.lcc_state, d.n_lcc * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (lccv_state && d.n_lcc) HIPCHK(hipMemcpy(lccv_state, d.lccv_state, d.n_lcc * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (R && d.n_ryr + d.n_lcc) HIPCHK(hipMemcpy(R, d.R, (d.n_ryr + d.n_lcc) * sizeof(double), hipMemcpyDeviceToHost));
  return SRS_OK;
}


int srs_globals_set(srs_ctx* c, int32_t r, double vm, double ca_nsr) {
  int rc = check_real(c, r);
  if (rc) return rc;
  RealLive lv;
  if ((rc = live_rw(c, r, &lv, false))) return rc;
  lv.vm = vm;
  lv.ca_nsr = ca_nsr;
  serca_terms(ca_nsr, c->reals[r].reals, lv.serca);
  return live_rw(c, r, &lv, true);
}

int srs_globals_get(srs_ctx* c, int32_t r, double* vm, double* ca_nsr) {
  int rc = check_real(c, r);
  if (rc) return rc;
  RealLive lv;
  if ((rc = live_rw(c, r, &lv, false))) return rc;
  if (vm) *vm = lv.vm;
  if (ca_nsr) *ca_nsr = lv.ca_nsr;
  return SRS_OK;
}

int srs_loop_set(srs_ctx* c, int32_t r, const srs_loop_state* ls) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (!ls) return srs::fail(SRS_E_ARG, "null loop state");
  RealLive lv;
  if ((rc = live_rw(c, r, &lv, false))) return rc;
  lv.t = ls->t; lv.dt_last = ls->dt_last; lv.step = ls->step; lv.done = ls->done; lv.lcc_open = ls->lcc_open;
  lv.t_lcc_close = ls->t_lcc_close; lv.n_open_last = ls->n_ryr_open_last; lv.ca_max_last = ls->ca_ss_max_last;
  lv.ffwd = ls->ffwd ? 1 : 0;
  return live_rw(c, r, &lv, true);
}

int srs_loop_get(srs_ctx* c, int32_t r, srs_loop_state* ls) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (!ls) return srs::fail(SRS_E_ARG, "null loop state");
  RealLive lv;
  if ((rc = live_rw(c, r, &lv, false))) return rc;
  ls->t = lv.t; ls->dt_last = lv.dt_last; ls->step = lv.step; ls->done = lv.done; ls->lcc_open = lv.lcc_open;
  ls->t_lcc_close = lv.t_lcc_close; ls->n_ryr_open_last = lv.n_open_last; ls->ca_ss_max_last = lv.ca_max_last;
  ls->ffwd = lv.ffwd;
  return SRS_OK;
}

// ---------------------------------------------------------- checkpoints
// Layout of a realization's checkpoint: header, fields in the reference
// numbering ([6][n_ele] FP64, as srs_state_get), RyR states, LCC states, LCC
// voltage-gate states (int32, padded to 8 bytes), residuals, then the loop
// state.
namespace {
struct CkHeader {
  char magic[8];
  int32_t version, n_ele, n_ryr, n_lcc;
  uint64_t key, stream;
};
struct CkLive {
  double t, dt_last, vm, ca_nsr, t_lcc_close, ca_max_last;
  int64_t step;
  int32_t done, lcc_open, n_open_last, ffwd;
};
constexpr char kCkMagic[8] = {'S', 'R', 'S', 'C', 'K', 'P', 'T', '1'};

size_t ck_int_bytes(size_t nr, size_t nl) { return (sizeof(int32_t) * (nr + 2 * nl) + 7) / 8 * 8; }

size_t ck_bytes(const HostReal& hr, int32_t n_ele) {
  const size_t nr = hr.ryr_ele.size(), nl = hr.lcc_ele.size();
  return sizeof(CkHeader) + 6 * sizeof(double) * (size_t)n_ele + ck_int_bytes(nr, nl) + sizeof(double) * (nr + nl) +
         sizeof(CkLive);
}
}  // namespace

int srs_checkpoint_size(srs_ctx* c, int32_t r, size_t* bytes) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (!bytes) return srs::fail(SRS_E_ARG, "null size");
  *bytes = ck_bytes(c->reals[r], c->meshes[c->reals[r].mesh]->n);
  return SRS_OK;
}

int srs_checkpoint_save(srs_ctx* c, int32_t r, void* buf, size_t bytes) {
  int rc = check_real(c, r);
  if (rc) return rc;
  const HostReal& hr = c->reals[r];
  const int32_t n = c->meshes[hr.mesh]->n;
  if (!buf || bytes < ck_bytes(hr, n)) return srs::fail(SRS_E_ARG, "checkpoint buffer smaller than srs_checkpoint_size");
  const size_t nr = hr.ryr_ele.size(), nl = hr.lcc_ele.size();
  char* p = (char*)buf;
  CkHeader h{};
  memcpy(h.magic, kCkMagic, 8);
  h.version = 1;
  h.n_ele = n;
  h.n_ryr = (int32_t)nr;
  h.n_lcc = (int32_t)nl;
  h.key = hr.key;
  h.stream = hr.stream;
  memcpy(p, &h, sizeof h);
  p += sizeof h;
  if ((rc = srs_state_get(c, r, (double*)p))) return rc;
  p += 6 * sizeof(double) * (size_t)n;
  int32_t* rs = (int32_t*)p;
  int32_t* ls = rs + nr;
  int32_t* lv = ls + nl;
  memset(lv + nl, 0, ck_int_bytes(nr, nl) - sizeof(int32_t) * (nr + 2 * nl));
  double* R = (double*)(p + ck_int_bytes(nr, nl));
  if ((rc = srs_channels_get(c, r, rs, ls, lv, R))) return rc;
  p = (char*)(R + nr + nl);
  RealLive L;
  if ((rc = live_rw(c, r, &L, false))) return rc;
  const CkLive k{L.t, L.dt_last, L.vm, L.ca_nsr, L.t_lcc_close, L.ca_max_last, L.step,
                 L.done, L.lcc_open, L.n_open_last, L.ffwd};
  memcpy(p, &k, sizeof k);
  return SRS_OK;
}

int srs_checkpoint_load(srs_ctx* c, int32_t r, const void* buf, size_t bytes) {
  int rc = check_real(c, r);
  if (rc) return rc;
  const HostReal& hr = c->reals[r];
  const int32_t n = c->meshes[hr.mesh]->n;
  const size_t need = ck_bytes(hr, n);
  if (!buf || bytes < sizeof(CkHeader)) return srs::fail(SRS_E_ARG, "checkpoint buffer too small");
  const char* p = (const char*)buf;
  CkHeader h;
  memcpy(&h, p, sizeof h);
  if (memcmp(h.magic, kCkMagic, 8) || h.version != 1) return srs::fail(SRS_E_IO, "not an srs checkpoint (version 1)");
  if (h.n_ele != n || h.n_ryr != (int32_t)hr.ryr_ele.size() || h.n_lcc != (int32_t)hr.lcc_ele.size())
    return srs::fail(SRS_E_ARG, "checkpoint of %d elements, %d RyR, %d LCC does not fit this realization", h.n_ele,
                     h.n_ryr, h.n_lcc);
  if (h.key != hr.key || h.stream != hr.stream)
    return srs::fail(SRS_E_ARG, "checkpoint of another seed or stream: its gating would not continue the same run");
  if (bytes < need) return srs::fail(SRS_E_ARG, "checkpoint buffer truncated");
  p += sizeof h;
  const size_t nr = h.n_ryr, nl = h.n_lcc;
  const double* X = (const double*)p;
  const int32_t* rs = (const int32_t*)(X + 6 * (size_t)n);
  const int32_t* ls = rs + nr;
  const int32_t* lv = ls + nl;
  const double* R = (const double*)((const char*)rs + ck_int_bytes(nr, nl));
  CkLive k;
  memcpy(&k, R + nr + nl, sizeof k);
  if ((rc = srs_state_set(c, r, X))) return rc;
  if ((rc = srs_channels_set(c, r, rs, ls, lv, R))) return rc;
  RealLive L;
  if ((rc = live_rw(c, r, &L, false))) return rc;
  L.t = k.t; L.dt_last = k.dt_last; L.vm = k.vm; L.ca_nsr = k.ca_nsr; L.t_lcc_close = k.t_lcc_close;
  L.ca_max_last = k.ca_max_last; L.step = k.step; L.sched_base = k.step;
  L.done = k.done; L.lcc_open = k.lcc_open; L.n_open_last = k.n_open_last; L.ffwd = k.ffwd;
  L.rec_count = 0;
  L.rec_overflow = 0;
  serca_terms(L.ca_nsr, hr.reals, L.serca);
  return live_rw(c, r, &L, true);
}

// The captured block of `block` steps in gate mode `gmode` (made on first use
// and cached; tiles_compact and tiles_full drop the cache).
// Graph-cache key bits of the gated blocks (host only, never passed to a kernel).
constexpr uint32_t G_KEY_GATED = 1u << 8, G_KEY_CUR1 = 1u << 9, G_KEY_SUBSETS = 1u << 10;

static int graph_cached(srs_ctx* c, int64_t block, uint32_t key_mode, const std::function<void()>& capture,
                        hipGraphExec_t* out) {
  auto key = std::make_pair(block, key_mode);
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    capture();
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t ge;
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    it = c->graphs.emplace(key, ge).first;
  }
  *out = it->second;
  return SRS_OK;
}

static int graph_for(srs_ctx* c, int64_t block, uint32_t gmode, hipGraphExec_t* out) {
  if (!c->subsets.empty())
    return graph_cached(c, block, gmode | G_KEY_SUBSETS, [&] { launch_subset_block(c, block, gmode); }, out);
  return graph_cached(c, block, gmode, [&] { for (int64_t k = 0; k < block; ++k) launch_step(c, gmode); }, out);
}

// A block of steps through k_sweep_gated starting from live state index cur.
static int graph_gated(srs_ctx* c, int64_t block, int cur, uint32_t gmode, hipGraphExec_t* out) {
  return graph_cached(c, block, gmode | G_KEY_GATED | (cur ? G_KEY_CUR1 : 0u),
                      [&] { launch_gated_block(c, block, cur, gmode); }, out);
}

// Whether steps run through k_sweep_gated, and the live state index they
// start from: single-realization contexts with the LDS sweep of groups of
// one and a register budget of 0 or 6 (0 with the loop body), while the
// realization is neither finished nor in the ffwd phase.  Reading the live
// state waits for the stream.
static int gated_start(srs_ctx* c, uint32_t flags, bool* ok, int* cur) {
  *ok = false;
  const int occ_max = (flags & SRS_STEP_PROTOCOL) ? 0 : 6;
  if (!c->gated || c->reals.size() != 1 || c->group != 1 || !c->lds || c->occ > occ_max) return SRS_OK;
  RealLive lv;
  int rc = live_rw(c, 0, &lv, false);
  if (rc) return rc;
  if (lv.done || lv.ffwd || c->n_tiles == 0) return SRS_OK;
  *ok = true;
  *cur = lv.cur;
  return SRS_OK;
}

constexpr int64_t kGraphBlock = 128, kGraphMin = 16;
static_assert(kGraphBlock % 2 == 0, "a full gated block returns the live state index to where it started");

int srs_explicit_step_prepare(srs_ctx* c, int64_t n_steps, uint32_t flags) {
  srs::clear_error();
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (!c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (n_steps < 0) return srs::fail(SRS_E_ARG, "negative step count");
  HIPCHK(hipSetDevice(c->device));
  const uint32_t gmode = G_AFTER_SWEEP | ((flags & SRS_STEP_PROTOCOL) ? G_PROTOCOL : 0u);
  int rc = tiles_compact(c);
  if (rc) return rc;
  if (!c->subsets_valid && (rc = plan_subsets(c))) return rc;
  hipGraphExec_t ge;
  bool gated;
  int cur;
  if ((rc = gated_start(c, flags, &gated, &cur))) return rc;
  if (gated) {
    // both starting live indices: the one a later call starts from is not known here
    for (int p = 0; p < 2; ++p) {
      if (n_steps >= kGraphBlock && (rc = graph_gated(c, kGraphBlock, p, gmode, &ge))) return rc;
      if (n_steps % kGraphBlock >= kGraphMin && (rc = graph_gated(c, n_steps % kGraphBlock, p, gmode, &ge)))
        return rc;
    }
    return SRS_OK;
  }
  if (n_steps >= kGraphBlock && (rc = graph_for(c, kGraphBlock, gmode, &ge))) return rc;
  if (n_steps % kGraphBlock >= kGraphMin && (rc = graph_for(c, n_steps % kGraphBlock, gmode, &ge))) return rc;
  return SRS_OK;
}

int srs_explicit_step(srs_ctx* c, int64_t n_steps, uint32_t flags) {
  srs::clear_error();
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (!c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (n_steps < 0) return srs::fail(SRS_E_ARG, "negative step count");
  HIPCHK(hipSetDevice(c->device));
  const uint32_t gmode = G_AFTER_SWEEP | ((flags & SRS_STEP_PROTOCOL) ? G_PROTOCOL : 0u);
  int rc = tiles_compact(c);
  if (rc) return rc;
  if (!c->subsets_valid && (rc = plan_subsets(c))) return rc;
  bool gated = false;
  int cur = 0;
  if (n_steps > 0 && (rc = gated_start(c, flags, &gated, &cur))) return rc;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int64_t done = 0;
  if (flags & SRS_STEP_GRAPH) {
    // 128-step blocks, then the rest as one block of its own when it is long enough
    hipGraphExec_t ge;
    if (n_steps >= kGraphBlock) {
      if ((rc = gated ? graph_gated(c, kGraphBlock, cur, gmode, &ge) : graph_for(c, kGraphBlock, gmode, &ge)))
        return rc;
      for (; n_steps - done >= kGraphBlock; done += kGraphBlock) HIPCHK(hipGraphLaunch(ge, c->stream));
    }
    if (n_steps - done >= kGraphMin) {
      const int64_t m = n_steps - done;
      if ((rc = gated ? graph_gated(c, m, cur, gmode, &ge) : graph_for(c, m, gmode, &ge))) return rc;
      HIPCHK(hipGraphLaunch(ge, c->stream));
      cur ^= (int)(m & 1);
      done = n_steps;
    }
  }
  if (gated) {
    if (done < n_steps) launch_gated_block(c, n_steps - done, cur, gmode);
  } else if (!c->subsets.empty()) {
    if (done < n_steps) launch_subset_block(c, n_steps - done, gmode);
  } else {
    for (; done < n_steps; ++done) launch_step(c, gmode);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_total_ms = ms;
  return SRS_OK;
}

int srs_channel_gating(srs_ctx* c) {
  srs::clear_error();
  if (!c || !c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  HIPCHK(hipSetDevice(c->device));
  hipLaunchKernelGGL(k_gate, dim3((unsigned)c->reals.size()), dim3(GATE_THREADS), 0, c->stream, c->d_desc, c->d_live,
                     0u);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return SRS_OK;
}

// n_steps replayed steps with the schedule already in device memory at ds
// (null when the realization has no RyRs).
static int replay_from(srs_ctx* c, int32_t r, int64_t n_steps, const int32_t* schedule, int32_t* ds) {
  HostReal& hr = c->reals[r];
  const int nr = (int)hr.ryr_ele.size();
  if (nr) HIPCHK(hipMemcpy(hr.desc.ryr_state, schedule, nr * sizeof(int32_t), hipMemcpyHostToDevice));
  RealDesc d = hr.desc;
  d.sched = ds;
  d.n_sched = n_steps;
  HIPCHK(hipMemcpy(c->d_desc + r, &d, sizeof(d), hipMemcpyHostToDevice));
  RealLive lv;
  HIPCHK(hipMemcpy(&lv, c->d_live + r, sizeof(lv), hipMemcpyDeviceToHost));
  lv.sched_base = lv.step;
  HIPCHK(hipMemcpy(c->d_live + r, &lv, sizeof(lv), hipMemcpyHostToDevice));
  const uint32_t gmode = G_AFTER_SWEEP | G_NO_GATING | G_REPLAY;
  for (int64_t k = 0; k < n_steps; ++k) launch_step(c, gmode);
  hipError_t e = hipGetLastError();
  (void)hipStreamSynchronize(c->stream);
  HIPCHK(hipMemcpy(c->d_desc + r, &hr.desc, sizeof(hr.desc), hipMemcpyHostToDevice));
  if (e != hipSuccess) return srs::fail(SRS_E_HIP, "replay launch: %s", hipGetErrorString(e));
  return SRS_OK;
}

static int replay_check(srs_ctx* c, int32_t r, int64_t n_steps, const int32_t* schedule) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (c->reals.size() != 1) return srs::fail(SRS_E_ARG, "replay needs a single-realization context");
  if (n_steps < 0) return srs::fail(SRS_E_ARG, "negative step count");
  if (!schedule && n_steps > 0 && !c->reals[r].ryr_ele.empty()) return srs::fail(SRS_E_ARG, "null schedule");
  return SRS_OK;
}

int srs_explicit_step_replay(srs_ctx* c, int32_t r, int64_t n_steps, const int32_t* schedule) {
  int rc = replay_check(c, r, n_steps, schedule);
  if (rc || n_steps == 0) return rc;
  const int nr = (int)c->reals[r].ryr_ele.size();
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  int32_t* ds = nullptr;
  if (nr) {
    HIPCHK(hipMalloc(&ds, (size_t)n_steps * nr * sizeof(int32_t)));
    hipError_t e = hipMemcpy(ds, schedule, (size_t)n_steps * nr * sizeof(int32_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(ds);
      return srs::fail(SRS_E_HIP, "replay schedule copy: %s", hipGetErrorString(e));
    }
  }
  rc = replay_from(c, r, n_steps, schedule, ds);
  if (ds) (void)hipFree(ds);
  return rc;
}

int srs_explicit_step_replay_static(srs_ctx* c, int32_t r, int64_t n_steps, const int32_t* schedule) {
  int rc = replay_check(c, r, n_steps, schedule);
  if (rc || n_steps == 0) return rc;
  const size_t nr = c->reals[r].ryr_ele.size();
  if (nr && (size_t)n_steps > kStaticSchedInts / nr)
    return srs::fail(SRS_E_ARG, "replay of %lld steps x %zu RyRs exceeds the static schedule buffer of %zu RyR states",
                     (long long)n_steps, nr, kStaticSchedInts);
  int32_t* ds = scratch_ptr(c->device);
  if (!ds) return srs::fail(SRS_E_STATE, "no static schedule buffer on device %d", c->device);
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (nr) HIPCHK(hipMemcpy(ds, schedule, (size_t)n_steps * nr * sizeof(int32_t), hipMemcpyHostToDevice));
  return replay_from(c, r, n_steps, schedule, nr ? ds : nullptr);
}

int srs_channel_ca(srs_ctx* c, int32_t r, double* ca_ryr_cyt, double* ca_ryr_jsr, double* ca_lcc) {
  int rc = check_real(c, r);
  if (rc) return rc;
  const RealDesc& d = c->reals[r].desc;
  const int n = 2 * d.n_ryr + d.n_lcc;
  if (n == 0) return SRS_OK;
  HIPCHK(hipSetDevice(c->device));
  double* tmp;
  HIPCHK(hipMalloc(&tmp, n * sizeof(double)));
  hipLaunchKernelGGL(k_channel_ca, dim3(1), dim3(256), 0, c->stream, c->d_desc, c->d_live, r, tmp);
  std::vector<double> h(n);
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(h.data(), tmp, n * sizeof(double), hipMemcpyDeviceToHost);
  (void)hipFree(tmp);
  if (e != hipSuccess) return srs::fail(SRS_E_HIP, "channel Ca: %s", hipGetErrorString(e));
  if (ca_ryr_cyt) memcpy(ca_ryr_cyt, h.data(), d.n_ryr * sizeof(double));
  if (ca_ryr_jsr) memcpy(ca_ryr_jsr, h.data() + d.n_ryr, d.n_ryr * sizeof(double));
  if (ca_lcc) memcpy(ca_lcc, h.data() + 2 * d.n_ryr, d.n_lcc * sizeof(double));
  return SRS_OK;
}

int srs_records_read(srs_ctx* c, int32_t r, float* out, int32_t max_records, int32_t* n_records, int32_t* rec_len) {
  int rc = check_real(c, r);
  if (rc) return rc;
  const RealDesc& d = c->reals[r].desc;
  RealLive lv;
  if ((rc = live_rw(c, r, &lv, false))) return rc;
  if (rec_len) *rec_len = d.rec_len;
  if (lv.rec_overflow) return srs::fail(SRS_E_STATE, "device record buffer overflowed (capacity %d)", d.rec_cap);
  const int n = std::min(lv.rec_count, max_records);
  if (n > 0 && out) HIPCHK(hipMemcpy(out, d.rec, (size_t)n * d.rec_len * sizeof(float), hipMemcpyDeviceToHost));
  if (n_records) *n_records = n;
  // keep any records that did not fit
  const int rest = lv.rec_count - n;
  if (n == 0) return SRS_OK;
  if (rest > 0) {
    std::vector<float> tail((size_t)rest * d.rec_len);
    HIPCHK(hipMemcpy(tail.data(), d.rec + (size_t)n * d.rec_len, tail.size() * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(d.rec, tail.data(), tail.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  // Only the record count changes: a finished realization whose records
  // are drained stays finished, so the sweep's tile list and the captured
  // graphs remain valid (a full live-state write would restore the tile
  // list and drop them).
  const int32_t cnt = rest;
  HIPCHK(hipMemcpy((char*)(c->d_live + r) + offsetof(RealLive, rec_count), &cnt, sizeof cnt, hipMemcpyHostToDevice));
  return SRS_OK;
}

int srs_ctx_sync(srs_ctx* c) {
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  return SRS_OK;
}

int srs_real_info(srs_ctx* c, int32_t r, int32_t* n_ele, int32_t* n_ryr, int32_t* n_lcc) {
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (r < 0 || r >= (int32_t)c->reals.size()) return srs::fail(SRS_E_ARG, "realization %d out of range", r);
  const HostReal& hr = c->reals[r];
  if (n_ele) *n_ele = c->meshes[hr.mesh]->n;
  if (n_ryr) *n_ryr = (int32_t)hr.ryr_ele.size();
  if (n_lcc) *n_lcc = (int32_t)hr.lcc_ele.size();
  return SRS_OK;
}

int srs_last_timing(srs_ctx* c, double* total_ms, double* sweep_ms) {
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (total_ms) *total_ms = c->last_total_ms;
  if (sweep_ms) *sweep_ms = c->last_sweep_ms;
  return SRS_OK;
}

int srs_time_sweep(srs_ctx* c, int32_t reps, double* sweep_ms, double* gating_ms) {
  // Real steps (sweep + gating, state advances), with HIP events on the
  // context's stream bracketing each kernel separately.
  srs::clear_error();
  if (!c || !c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (reps <= 0) return srs::fail(SRS_E_ARG, "reps must be > 0");
  HIPCHK(hipSetDevice(c->device));
  bool gated;
  int cur;
  int rc = gated_start(c, 0, &gated, &cur);
  if (rc) return rc;
  std::vector<hipEvent_t> ev(3 * reps);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  const uint32_t gmode = G_AFTER_SWEEP;
  for (int k = 0; k < reps; ++k) {
    HIPCHK(hipEventRecord(ev[3 * k], c->stream));
    // the kernels srs_explicit_step runs: with the gated sweep, each launch
    // also gates the step before it, and the gating after the last one is
    // the whole gating time of the reps
    if (gated) launch_gated(c, cur ^ (k & 1), k ? gmode : 0u, false);
    else launch_sweep(c);
    HIPCHK(hipEventRecord(ev[3 * k + 1], c->stream));
    if (!gated || k == reps - 1) {
      hipLaunchKernelGGL(k_gate, dim3((unsigned)c->reals.size()), dim3(GATE_THREADS), 0, c->stream, c->d_desc,
                         c->d_live, gmode);
      HIPCHK(hipEventRecord(ev[3 * k + 2], c->stream));
    }
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  double s = 0, g = 0;
  for (int k = 0; k < reps; ++k) {
    float a = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]));
    if (!gated || k == reps - 1) HIPCHK(hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]));
    s += a;
    g += b;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  c->last_sweep_ms = s / reps;
  if (sweep_ms) *sweep_ms = s / reps;
  if (gating_ms) *gating_ms = g / reps;
  return SRS_OK;
}

}  // extern "C"
#endif  // SRS_UNIT == 0
