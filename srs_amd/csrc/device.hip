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
// The reference runs the same step as nine kernels with a state-sized rate
// buffer (gpu/RunSimulation.cu:781-852); here the rates never leave registers.
//
// Arithmetic keeps the reference's operation order and this file is compiled
// without FP contraction, so the field update is bit-identical to the
// reference CPU path for the same row order of each element's neighbours.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

#include "kinetics.h"
#include "philox.h"
#include "srs_internal.h"

#pragma clang fp contract(off)

namespace {

constexpr int MAX_TILE = 512;  // largest k_sweep workgroup (elements per workgroup)
constexpr int GATE_THREADS = 256;
constexpr int MAX_RYR = 1024;  // the reference GPU path launches <<<1, N_RyR>>>
constexpr int MAX_LCC = 1024;
constexpr int NXCD = 8;        // MI355X: 8 XCDs, workgroups dispatched round-robin

enum : uint8_t { F_JSR = 1, F_TROPC = 2, F_SL = 4, F_BR = 8, F_CHAN = 16 };

// State layout of one realization (6*np doubles): the four diffusing species
// interleaved per element, (Ca, ATP, CaM, Fluo) in 32 bytes, so that one
// neighbour gather fetches all four with one contiguous access; then the TnC
// and CSQ planes, which are never gathered.
constexpr int MOB_SLOT[6] = {0, 1, 2, -1, -1, 3};  // species -> slot in the interleaved quad
// A realization group holds K realizations of one mesh in blocks of 64
// elements: block b holds the 64 quads of slot 0, then slot 1, ..., so a
// wavefront's loads of one realization stay contiguous.  Element e of slot
// `slot` is entry ((e/64)*K + slot)*64 + e%64 of the quad array, and of the
// TnC and CSQ planes (K*np entries each) that follow the 4*K*np quad doubles.
__host__ __device__ inline size_t gidx(size_t e, int K, int slot) {
  return ((e >> 6) * (size_t)K + (size_t)slot) * 64 + (e & 63);
}
__host__ __device__ inline size_t xoff(int s, size_t e, size_t np, int K = 1, int slot = 0) {
  const size_t q = gidx(e, K, slot);
  return s == srs::TRPN ? 4 * (size_t)K * np + q : s == srs::CSQN ? 5 * (size_t)K * np + q : 4 * q + MOB_SLOT[s];
}
typedef double d4 __attribute__((ext_vector_type(4)));

// Immutable per-realization descriptor (read with scalar loads).
struct RealDesc {
  const int32_t* nbr;     // [4][np]
  const double* coef;     // [5][np] (plain layout)
  const int16_t* off;     // compact layout: [4][np] neighbour offsets j' - j
  const uint16_t* cls;    // compact layout: [np] operator-row class
  const int16_t* dq;      // compact layout: [5][np] coefficient deltas in ulps of the class row
  const double* tab;      // compact layout: [n_class][10] class row (5 values) and its ulps (5 values)
  int32_t n_class;
  const uint8_t* flags;   // [np]
  const double* V0;       // [np]
  const double* tt;       // [np]
  const double* br;       // [4][np] Ca, Fluo, ATP, CaM boundary rates, or null
  double* X[2];           // [6][np] ping-pong state
  int32_t n, np;
  int32_t kst, slot;      // realization group: K realizations interleaved, this one in slot `slot`
  int32_t n_ryr, n_lcc;
  const int32_t* ryr_ele;
  const int32_t* ryr_jsr;
  const int32_t* ryr_nb;
  int32_t* ryr_state;
  const int32_t* lcc_ele;
  int32_t* lcc_state;
  int32_t* lccv_state;
  double* R;
  int32_t n_celem;
  const int32_t* celem;   // sorted element indices that carry channel flux
  const int32_t* cstart;  // [n_celem+1]
  const int32_t* centry;  // >=0: 2*ryr + side(0 cleft, 1 jSR); <0: -(lcc+1)
  const int32_t* sched;   // replay schedule [n_sched][n_ryr] or null
  int64_t n_sched;
  float* rec;
  int32_t rec_cap, rec_len, ryr_out, gpu_layout;
  uint64_t key, stream;
  double P[SRS_N_REAL];
  int32_t I[SRS_N_INT];
};

// Offset of element e's Ca in a realization's state buffer (its quad's first
// double); Fluo is at +3.
__device__ __forceinline__ size_t cq(const RealDesc& d, int e) { return 4 * gidx((size_t)e, d.kst, d.slot); }

// Mutable per-realization loop state.
struct RealLive {
  double t, dt_last, vm, ca_nsr, t_lcc_close, ca_max_last;
  // CaNSR-dependent SERCA subexpressions of cpu/RunSimulation.cpp:85-88, the
  // same for every element: 0.104217 + 17.923 K_sr, 1.75583e6 + K_sr 7.61673e6,
  // 6.08463e11 + K_sr 4.50544e11, 9.17846e6 - 11478.2 K_sr, 0.329904 K_sr
  double serca[5];
  int64_t step;
  int64_t sched_base;
  int32_t done, lcc_open, n_open_last, cur, rec_count, rec_overflow;
  int32_t ffwd;  // gating-only phase of flag_ffwd: the sweep skips this realization
};

enum : uint32_t {
  G_AFTER_SWEEP = 1,  // new state is X[1-cur]; flip cur afterwards
  G_PROTOCOL = 2,     // run the loop body for the next time
  G_NO_GATING = 4,    // skip gating regardless of the parameter flag
  G_REPLAY = 8,       // load RyR states for the next step from the schedule
};

__device__ __forceinline__ int xcd_tile(int bid, int grid) {
  // Consecutive tiles of a realization land on one XCD, so neighbour gathers
  // that cross a tile boundary hit that XCD's L2.
  return (bid % NXCD) * (grid / NXCD) + bid / NXCD;
}

// Loads of data read once per step (operator rows, flags) may bypass cache
// allocation so that the state planes stay resident between steps.  Pointers
// read from the descriptor are generic; casting them to the global address
// space lets the compiler use global (not flat) memory instructions.
#define SRS_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const SRS_GLOBAL T* gp(const T* p) { return (const SRS_GLOBAL T*)p; }
template <class T>
__device__ __forceinline__ SRS_GLOBAL T* gpw(T* p) { return (SRS_GLOBAL T*)p; }

template <bool NT, class T>
__device__ __forceinline__ T ldg(const SRS_GLOBAL T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool NT, class T>
__device__ __forceinline__ void stg(SRS_GLOBAL T* p, T v) {
  if constexpr (NT) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// Operator row of one element: coefficients (diagonal first), neighbour
// indices and flags.  Mesh-level data, shared by every realization of a
// realization group.
struct Geo {
  double a0, a1, a2, a3, a4;
  int n1, n2, n3, n4;
  uint8_t f;
};

// FK >= 0: the caller knows the element's flags (a wavefront whose lanes are
// all alike), so none are loaded.
template <bool NTG, bool COMPACT, int FK>
__device__ __forceinline__ Geo load_geo(const RealDesc& d, const int j) {
  Geo g;
  const int np = d.np;
  g.f = FK >= 0 ? (uint8_t)FK : ldg<NTG>(gp(d.flags) + j);
  if constexpr (COMPACT) {
    // Operator row = class row + small integer multiples of its ulps, which
    // reproduces the assembled FP64 values exactly (every row was checked
    // when the mesh was added).
    const SRS_GLOBAL int16_t* off = gp(d.off);
    g.n1 = j + ldg<NTG>(off + j);
    g.n2 = j + ldg<NTG>(off + np + j);
    g.n3 = j + ldg<NTG>(off + 2 * np + j);
    g.n4 = j + ldg<NTG>(off + 3 * np + j);
    const SRS_GLOBAL double* t = gp(d.tab) + 10 * (int)ldg<NTG>(gp(d.cls) + j);
    const SRS_GLOBAL int16_t* dq = gp(d.dq);
    g.a0 = t[0] + (double)ldg<NTG>(dq + j) * t[5];
    g.a1 = t[1] + (double)ldg<NTG>(dq + np + j) * t[6];
    g.a2 = t[2] + (double)ldg<NTG>(dq + 2 * np + j) * t[7];
    g.a3 = t[3] + (double)ldg<NTG>(dq + 3 * np + j) * t[8];
    g.a4 = t[4] + (double)ldg<NTG>(dq + 4 * np + j) * t[9];
  } else {
    const SRS_GLOBAL double* cf = gp(d.coef);
    const SRS_GLOBAL int32_t* nb = gp(d.nbr);
    g.a0 = ldg<NTG>(cf + j); g.a1 = ldg<NTG>(cf + np + j); g.a2 = ldg<NTG>(cf + 2 * np + j);
    g.a3 = ldg<NTG>(cf + 3 * np + j); g.a4 = ldg<NTG>(cf + 4 * np + j);
    g.n1 = ldg<NTG>(nb + j); g.n2 = ldg<NTG>(nb + np + j); g.n3 = ldg<NTG>(nb + 2 * np + j);
    g.n4 = ldg<NTG>(nb + 3 * np + j);
  }
  return g;
}

// The state of one element of one realization: its own and its neighbours'
// diffusing quads, TnC and CSQ where they are live.
struct Loc {
  d4 m0, m1, m2, m3, m4;
  double c_tr, c_cs;
};

template <int FK>
__device__ __forceinline__ Loc load_loc(const RealDesc& d, const RealLive& l, const int j, const Geo& g) {
  // Realization-group layout: element e of the realization in slot s of a
  // group of K is quad gidx(e, K, s); the TnC and CSQ planes follow the quads.
  const int K = d.kst, s = d.slot;
  const size_t np = d.np;
  const SRS_GLOBAL double* __restrict__ Xo = gp(d.X[l.cur]);
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
__device__ __forceinline__ Loc load_loc_lds(const RealDesc& d, const RealLive& l, const int j, const Geo& g,
                                            const d4* tq, const int base, const d4& m0) {
  const size_t np = d.np;
  const SRS_GLOBAL double* __restrict__ Xo = gp(d.X[l.cur]);
  const SRS_GLOBAL d4* Mo = (const SRS_GLOBAL d4*)Xo;
  Loc c;
  c.m0 = m0;
  const uint8_t f = FK >= 0 ? (uint8_t)FK : g.f;
  const bool live_tr = (f & (F_JSR | F_TROPC)) == F_TROPC;
  const bool live_cs = (f & F_JSR) != 0;
  c.c_tr = live_tr ? Xo[4 * np + j] : 0.0;
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
__device__ __forceinline__ void step_elem(const RealDesc& d, const RealLive& l, const int j, const Geo& g,
                                          const Loc& c) {
  const int K = d.kst, s = d.slot;
  const size_t np = d.np;
  const SRS_GLOBAL double* __restrict__ Xo = gp(d.X[l.cur]);
  SRS_GLOBAL double* __restrict__ Xn = gpw(d.X[l.cur ^ 1]);
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
  // cpu/RunSimulation.cpp:61-71 (Ca uses D_CA in both domains)
  double rca = P[SRS_R_D_CA] * (c_ca * a0 + m1.x * a1 + m2.x * a2 + m3.x * a3 + m4.x * a4) + br_ca;
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

  if (f & F_SL) {  // :15-22; B == 1 exactly where the TT area is zero
    const double ttj = gp(d.tt)[j], v0j = gp(d.V0)[j];
    const double btothi = ttj * P[SRS_R_B_TOT_SL_HI] / (v0j * (1e-15));
    const double btotlo = ttj * P[SRS_R_B_TOT_SL_LO] / (v0j * (1e-15));
    const double kh = P[SRS_R_K_D_SL_HI], kl = P[SRS_R_K_D_SL_LO];
    const double B = 1 + (btothi * kh / ((kh + c_ca) * (kh + c_ca))) + (btotlo * kl / ((kl + c_ca) * (kl + c_ca)));
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
// applied to the KG realizations of the group, one after another.  Realizations that are done or in the ffwd phase
// are computed too, into the buffer they do not read next: their state
// pointer does not flip, so the result is never used.
template <bool NTS, int KG, int FK>
__device__ __forceinline__ void sweep_group(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                            const int* mem, const int j, const Geo& g) {
  if constexpr (KG == 1) {
    const Loc c = load_loc<FK>(D[mem[0]], L[mem[0]], j, g);
    step_elem<NTS, FK>(D[mem[0]], L[mem[0]], j, g, c);
  } else {
    // members one after another: the register budget of a single realization;
    // -1 marks an empty slot or a member that does not step
    for (int s = 0; s < KG; ++s)
      if (mem[s] >= 0) {
        const Loc c = load_loc<FK>(D[mem[s]], L[mem[s]], j, g);
        step_elem<NTS, FK>(D[mem[s]], L[mem[s]], j, g, c);
      }
  }
}

// Single-realization groups with LDS staging: the tile's own quads are
// loaded coalesced into LDS, and neighbour quads inside the tile come from
// there instead of through the L1 cache.
template <int BS, bool NTG, bool NTS, bool COMPACT>
__device__ __forceinline__ void sweep_tile_lds(const RealDesc* __restrict__ D, const RealLive* __restrict__ L, int r,
                                               int t) {
  __shared__ d4 tq[BS];
  const RealDesc& d = D[r];
  const RealLive& l = L[r];
  const int base = t * BS;
  const int j = base + threadIdx.x;
  const bool in = j < d.n;
  Geo g{};
  d4 m0 = {0.0, 0.0, 0.0, 0.0};
  if (in) {
    g = load_geo<NTG, COMPACT, -1>(d, j);
    m0 = ((const SRS_GLOBAL d4*)gp(d.X[l.cur]))[j];
  }
  tq[threadIdx.x] = m0;
  __syncthreads();
  if (!in) return;
  const bool plain = (g.f & F_SPECIAL) == 0 && (g.f & F_TROPC);
  if (__builtin_amdgcn_ballot_w64(!plain) == 0) {
    const Loc c = load_loc_lds<F_TROPC, BS>(d, l, j, g, tq, base, m0);
    step_elem<NTS, F_TROPC>(d, l, j, g, c);
  } else {
    const Loc c = load_loc_lds<-1, BS>(d, l, j, g, tq, base, m0);
    step_elem<NTS, -1>(d, l, j, g, c);
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

// Two register budgets: the compiler's choice, or at most 64 VGPRs so that
// eight wavefronts share each SIMD.
template <int BS, bool NTG, bool NTS, bool COMPACT, int KG>
__global__ __launch_bounds__(BS) void k_sweep(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                              const int2* __restrict__ tiles, int n_tiles,
                                              const int* __restrict__ groups, int single) {
  sweep_tile<BS, NTG, NTS, COMPACT, KG>(D, L, tiles, n_tiles, groups, single);
}
template <int BS, bool NTG, bool NTS, bool COMPACT>
__global__ __launch_bounds__(BS) void k_sweep_lds(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                                  const int2* __restrict__ tiles, int n_tiles,
                                                  const int* __restrict__ groups, int single) {
  const int b = xcd_tile(blockIdx.x, gridDim.x);
  if (b >= n_tiles) return;
  const int2 tl = single ? make_int2(0, b) : tiles[b];
  const int r = single ? 0 : groups[tl.x];
  if (L[r].done | L[r].ffwd) return;
  sweep_tile_lds<BS, NTG, NTS, COMPACT>(D, L, r, tl.y);
}
template <int BS, bool NTG, bool NTS, bool COMPACT, int KG>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(8))) void k_sweep_o8(
    const RealDesc* __restrict__ D, const RealLive* __restrict__ L, const int2* __restrict__ tiles, int n_tiles,
    const int* __restrict__ groups, int single) {
  sweep_tile<BS, NTG, NTS, COMPACT, KG>(D, L, tiles, n_tiles, groups, single);
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

__global__ __launch_bounds__(GATE_THREADS) void k_gate(const RealDesc* __restrict__ D, RealLive* __restrict__ L,
                                                       uint32_t mode) {
  const RealDesc& d = D[blockIdx.x];
  RealLive& l = L[blockIdx.x];
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
    constexpr int RPT = MAX_RYR / GATE_THREADS, LPT = MAX_LCC / GATE_THREADS;
    double ca_c[RPT], ca_j[RPT], Rr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
      const int i = tid + k * GATE_THREADS;
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
#pragma unroll
    for (int k = 0; k < LPT; ++k) {
      const int i = (GATE_THREADS - 1 - tid) + k * GATE_THREADS;
      if (i < d.n_lcc) {
        ls[k] = d.lcc_state[i];
        lvs[k] = d.lccv_state[i];
        ca_l[k] = Ca[cq(d, d.lcc_ele[i])];
        Rl[k] = d.R[nr + i];
      }
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
          const int i = tid + k * GATE_THREADS;
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
      const int i = tid + k * GATE_THREADS;
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
      const int i = (GATE_THREADS - 1 - tid) + k * GATE_THREADS;
      if (i >= d.n_lcc) continue;
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
      for (int i = tid; i < nr; i += GATE_THREADS) d.ryr_state[i] = d.sched[k * nr + i];
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
  if ((mode & G_PROTOCOL) && (ff || out_due)) {
    __shared__ int s_cnt[2][GATE_THREADS / 64];
    __shared__ double s_max[GATE_THREADS / 64];
    int cr = 0, cl = 0;
    double m = 0.0;
    for (int b = 0; b < nr; b += GATE_THREADS) {
      const int i = b + tid;
      const bool open = i < nr && d.ryr_state[i] != 0;
      cr += __builtin_popcountll(__builtin_amdgcn_ballot_w64(open));
      if (i < nr) {
        const double c = Ca[cq(d, d.ryr_ele[i])];
        if (c > m) m = c;
      }
    }
    for (int b = 0; b < d.n_lcc; b += GATE_THREADS) {
      const int i = b + (GATE_THREADS - 1 - tid);
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
      for (int w = 0; w < GATE_THREADS / 64; ++w) {
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
  if (!(mode & G_PROTOCOL)) return;

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

// ------------------------------------------------------------------ linescan
// linescan/main.cpp:195-249.  The blurred fluorescence of scan point i is
// sum_j V0_j CaF_j w(c_j - p_i) over cells with |c_j - p_i|^2 < 1, with the
// Gaussian PSF w of linescan/parameters.h:60-62.  Mesh cells are summed on the
// device, one partial sum per workgroup and point, then the partials in
// workgroup order; the padding cells' part is a per-point constant.
struct ScanGeo {
  double base[3];   // scan line point with the axis coordinate zero
  int axis;
  double x_min, res;
  double sx, sy, sz, alpha;
};

constexpr int SCAN_BS = 256;
constexpr int SCAN_MAX_POINTS = 1024;
constexpr int SCAN_CELLS_PER_THREAD = 4;

// Realization of workgroup row blockIdx.y in a batched launch: rid[blockIdx.y],
// or r when there is no table (a launch for one realization).
__device__ __forceinline__ int batch_real(const int32_t* __restrict__ rid, int r) { return rid ? rid[blockIdx.y] : r; }

__global__ __launch_bounds__(SCAN_BS) void k_scan_blur(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                                      int r, const int32_t* __restrict__ rid,
                                                      const double* __restrict__ cen, ScanGeo g, int n_f,
                                                      double* __restrict__ part) {
  __shared__ double red[SCAN_BS / 64][SCAN_MAX_POINTS];
  r = batch_real(rid, r);
  part += (size_t)blockIdx.y * gridDim.x * n_f;
  const RealDesc& d = D[r];
  const double* X = d.X[L[r].cur];
  const int n = d.n, np = d.np;
  double cx[SCAN_CELLS_PER_THREAD], cy[SCAN_CELLS_PER_THREAD], cz[SCAN_CELLS_PER_THREAD], vf[SCAN_CELLS_PER_THREAD];
  for (int q = 0; q < SCAN_CELLS_PER_THREAD; ++q) {
    const int j = (blockIdx.x * SCAN_CELLS_PER_THREAD + q) * SCAN_BS + threadIdx.x;
    if (j < n) {
      cx[q] = cen[j]; cy[q] = cen[np + j]; cz[q] = cen[2 * np + j];
      vf[q] = d.V0[j] * X[cq(d, j) + 3];  // V0 * Ca-bound Fluo
    } else {
      cx[q] = cy[q] = cz[q] = 1e30;  // beyond the 1 um cut-off of every point
      vf[q] = 0.0;
    }
  }
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = 0; i < n_f; ++i) {
    double p[3] = {g.base[0], g.base[1], g.base[2]};
    p[g.axis] = g.x_min + i * g.res;
    double acc = 0.0;
    for (int q = 0; q < SCAN_CELLS_PER_THREAD; ++q) {
      const double rx = cx[q] - p[0], ry = cy[q] - p[1], rz = cz[q] - p[2];
      const double dist = rx * rx + ry * ry + rz * rz;
      if (dist < 1.0) {
        const double w = g.alpha * exp(-0.5 * ((rx * rx / (g.sx * g.sx)) + (ry * ry / (g.sy * g.sy)) +
                                               (rz * rz / (g.sz * g.sz)))) / (g.sx * g.sy * g.sz);
        acc += vf[q] * w;
      }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if (lane == 0) red[wave][i] = acc;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n_f; i += SCAN_BS) {
    double s = 0.0;
    for (int w = 0; w < SCAN_BS / 64; ++w) s += red[w][i];
    part[(size_t)blockIdx.x * n_f + i] = s;
  }
}

__global__ void k_scan_final(const RealDesc* __restrict__ D, const RealLive* __restrict__ L, int r,
                             const int32_t* __restrict__ rid, int n_blocks, int n_f, int n_ca,
                             const double* __restrict__ part, const double* __restrict__ pad,
                             const int32_t* __restrict__ near_f, const int32_t* __restrict__ near_ca, double f0,
                             double c0, double* __restrict__ out, int out_stride) {
  r = batch_real(rid, r);
  part += (size_t)blockIdx.y * n_blocks * n_f;
  out += (size_t)blockIdx.y * out_stride;
  const double* X = D[r].X[L[r].cur];
  for (int i = threadIdx.x; i < n_f; i += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < n_blocks; ++b) s += part[(size_t)b * n_f + i];
    out[n_ca + n_f + i] = s + pad[i];
    const int e = near_f[i];
    out[n_ca + i] = e >= 0 ? X[cq(D[r], e) + 3] : f0;
  }
  for (int i = threadIdx.x; i < n_ca; i += blockDim.x) {
    const int e = near_ca[i];
    out[i] = e >= 0 ? X[cq(D[r], e)] : c0;
  }
}

// ------------------------------------------------------------------- adjoint
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
// over points within 1 um (the blur's PSF, linescan/parameters.h:60-62).
__device__ __forceinline__ double scan_adj_sum(const double* __restrict__ cen, const ScanGeo& g, int n_f,
                                               const double* __restrict__ gcoef, size_t np, int j) {
  const double cx = cen[j], cy = cen[np + j], cz = cen[2 * np + j];
  double s = 0.0;
  for (int i = 0; i < n_f; ++i) {
    double p[3] = {g.base[0], g.base[1], g.base[2]};
    p[g.axis] = g.x_min + i * g.res;
    const double rx = cx - p[0], ry = cy - p[1], rz = cz - p[2];
    if (rx * rx + ry * ry + rz * rz < 1.0)
      s += gcoef[i] * (g.alpha * exp(-0.5 * ((rx * rx / (g.sx * g.sx)) + (ry * ry / (g.sy * g.sy)) +
                                             (rz * rz / (g.sz * g.sz)))) / (g.sx * g.sy * g.sz));
  }
  return s;
}

// w = (lambda_Ca / beta, lambda_ATP, lambda_CaM, lambda_Fluo) / V0 of element j.
__device__ __forceinline__ d4 adj_weight(const RealDesc& d, int j, double c, uint8_t f, double la_c, double la_a,
                                         double la_m, double la_d) {
  const SlBuf sb = sl_buffer(d, j, c, f);
  const double iv = 1.0 / d.V0[j];
  return d4{la_c / sb.beta * iv, la_a * iv, la_m * iv, la_d * iv};
}

// Pass 1 (the first step of a recomputed segment; later steps get w from the
// step kernel): w = (lambda_Ca / beta, lambda_ATP, lambda_CaM, lambda_Fluo) / V0.
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
  w[j] = adj_weight(d, j, X[4 * (size_t)j], d.flags[j], m.x, m.y, m.z, m.w);
}

// What the step kernel does after lambda_n: add the loss source of a scan
// sample taken after n steps (SCAN), and form w for step n - 1 from lambda_n
// and that step's state (Xprev, when it is in the recomputed segment).
struct AdjNext {
  const double* Xprev;  // state of step n - 1, or null
  d4* wout;             // its w
  const double* cen;    // scan cell centres [3][np]
  const double* gcoef;  // the sample's loss coefficients per point
  ScanGeo g;
  int n_f, gstride;
};

// Pass 2: lambda_n and the theta-gradient partials of one step.
template <bool COMPACT, bool SCAN>
__global__ __launch_bounds__(ADJ_BS) void k_adj_step(const RealDesc* __restrict__ D, const RealLive* __restrict__ L,
                                                     const int32_t* __restrict__ rid, const double* __restrict__ X,
                                                     const double* __restrict__ lam, const d4* __restrict__ w,
                                                     int64_t step, double* __restrict__ lam_out,
                                                     double* __restrict__ gpart, AdjNext nx) {
  __shared__ double red[ADJ_BS / 64][ADJ_NT];
  const int r = rid[blockIdx.y];
  const RealDesc& d = D[r];
  const RealLive& l = L[r];
  const int j = blockIdx.x * ADJ_BS + threadIdx.x;
  const size_t np = d.np;
  X += blockIdx.y * 6 * np;
  lam += blockIdx.y * 6 * np;
  lam_out += blockIdx.y * 6 * np;
  w += blockIdx.y * np;
  gpart += (size_t)blockIdx.y * gridDim.x * ADJ_NT;
  // the block's running partials, loaded now so the final update waits on no load
  const double gp_old = threadIdx.x < ADJ_NT ? gpart[(size_t)blockIdx.x * ADJ_NT + threadIdx.x] : 0.0;
  // RyR states during this step: row `step` of the frozen schedule
  const int32_t* ryr = d.n_ryr ? d.sched + step * d.n_ryr : nullptr;
  double gr[ADJ_NT];
#pragma unroll
  for (int k = 0; k < ADJ_NT; ++k) gr[k] = 0.0;
  if (j < d.n) {
    const double* P = d.P;
    const double dt = P[SRS_R_T_STEP];
    const Geo g = load_geo<false, COMPACT, -1>(d, j);
    const uint8_t f = g.f;
    const d4* Q = (const d4*)X;
    const d4 m0 = Q[j], m1 = Q[g.n1], m2 = Q[g.n2], m3 = Q[g.n3], m4 = Q[g.n4];
    const d4 w0 = w[j], w1 = w[g.n1], w2 = w[g.n2], w3 = w[g.n3], w4 = w[g.n4];
    const d4 mu = ((const d4*)lam)[j];
    const bool live_tr = (f & (F_JSR | F_TROPC)) == F_TROPC;
    const bool live_cs = (f & F_JSR) != 0;
    const double bT = live_tr ? X[4 * np + j] : 0.0, bC = live_cs ? X[5 * np + j] : 0.0;
    const double muT = lam[4 * np + j], muC = lam[5 * np + j];
    const double c = m0.x, bA = m0.y, bM = m0.z, bD = m0.w;
    const double V0 = d.V0[j];
    const SlBuf sb = sl_buffer(d, j, c, f);
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

    // forward Ca rate before the SL division (for the SL terms) and the
    // diffusion sums (for the D gradients)
    const double ac = c * g.a0 + m1.x * g.a1 + m2.x * g.a2 + m3.x * g.a3 + m4.x * g.a4;
    const double ad = bD * g.a0 + m1.w * g.a1 + m2.w * g.a2 + m3.w * g.a3 + m4.w * g.a4;
    double rca = P[SRS_R_D_CA] * ac + ((f & F_BR) ? d.br[j] : 0.0);
    gr[0] += dt * u * ac;          // D_CA
    gr[1] += dt * mu.w * ad;       // D_DYE
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
    if (SCAN) ld += V0 * scan_adj_sum(nx.cen, nx.g, nx.n_f, nx.gcoef + (size_t)blockIdx.y * nx.gstride, np, j);
    ((d4*)lam_out)[j] = d4{lc, la, lm, ld};
    lam_out[4 * np + j] = lT;
    lam_out[5 * np + j] = lC;
    if (nx.Xprev) {
      const double* Xp = nx.Xprev + blockIdx.y * 6 * np;
      nx.wout[blockIdx.y * np + j] = adj_weight(d, j, Xp[4 * (size_t)j], f, lc, la, lm, ld);
    }
  }
  // block reduction of the gradient partials, accumulated per block
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < ADJ_NT; ++k) {
    double v = gr[k];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if (lane == 0) red[wave][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < ADJ_NT) {
    double s = 0.0;
    for (int q = 0; q < ADJ_BS / 64; ++q) s += red[q][threadIdx.x];
    gpart[(size_t)blockIdx.x * ADJ_NT + threadIdx.x] = gp_old + s;
  }
}

// The loss gradient of one scan sample in the Fluo adjoint:
// lambda_Fluo,j += V0_j sum_i gcoef_i w(c_j - p_i) over points within 1 um.
__global__ __launch_bounds__(ADJ_BS) void k_scan_adj(const RealDesc* __restrict__ D, const int32_t* __restrict__ rid,
                                                     const double* __restrict__ cen, ScanGeo g, int n_f,
                                                     const double* __restrict__ gcoef, int gstride,
                                                     double* __restrict__ lam) {
  const RealDesc& d = D[rid[blockIdx.y]];
  const int j = blockIdx.x * ADJ_BS + threadIdx.x;
  if (j >= d.n) return;
  const size_t np = d.np;
  lam += blockIdx.y * 6 * np;
  lam[4 * (size_t)j + 3] += d.V0[j] * scan_adj_sum(cen, g, n_f, gcoef + (size_t)blockIdx.y * gstride, np, j);
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

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) return srs::fail(SRS_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, __LINE__); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

struct HostMesh {
  int32_t n = 0, np = 0;
  std::vector<int32_t> perm;  // dev -> ref
  std::vector<int32_t> inv;   // ref -> dev
  std::vector<int32_t> nbr;   // [4][np] dev numbering
  std::vector<double> coef;   // [5][np]
  std::vector<uint8_t> flags;
  std::vector<double> V0, tt, br;  // br [4][np] or empty
  // compact operator layout
  bool compact = false;
  std::vector<int16_t> off, dq;
  std::vector<uint16_t> cls;
  std::vector<double> tab;
  int32_t n_class = 0;
  int32_t *d_nbr = nullptr;
  double *d_coef = nullptr, *d_V0 = nullptr, *d_tt = nullptr, *d_br = nullptr, *d_tab = nullptr;
  uint8_t* d_flags = nullptr;
  int16_t *d_off = nullptr, *d_dq = nullptr;
  uint16_t* d_cls = nullptr;
};

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

struct HostReal {
  int32_t mesh = 0;
  std::vector<int32_t> ryr_ele, ryr_jsr, ryr_nb, lcc_ele;  // dev numbering
  std::vector<int32_t> celem, cstart, centry;
  int32_t ints[SRS_N_INT];
  double reals[SRS_N_REAL];
  uint64_t key = 0, stream = 0;
  int32_t group = 0;  // realization group (index into the sweep's group table)
  RealDesc desc{};
};

struct HostScan {
  int32_t mesh = 0, n_f = 0, n_ca = 0, n_blocks = 0;
  srs_scan_desc desc{};
  ScanGeo geo{};
  double *d_cen = nullptr, *d_pad = nullptr, *d_part = nullptr, *d_out = nullptr;
  int32_t *d_near_f = nullptr, *d_near_ca = nullptr;
};

}  // namespace

struct srs_ctx {
  int device = 0;
  int tile = 256;        // elements per k_sweep workgroup
  int nt_geometry = 0;   // non-temporal loads of operator rows and flags
  int nt_state = 0;      // non-temporal stores of the new state
  int plane_pad = 0;     // extra elements per SoA plane
  int compact = 1;       // compress operator rows when a mesh allows it
  int occ8 = 0;          // sweep compiled for eight wavefronts per SIMD
  int group = 1;         // realizations per realization group (1, 2 or 4)
  int lds = 0;           // stage each tile's quads in LDS for the neighbour gathers (groups of 1)
  int drop_finished = 1; // sweep without the tiles of groups that have all finished
  bool compact_all = false;
  bool layout_fixed = false;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool finalized = false;
  std::vector<std::unique_ptr<HostMesh>> meshes;
  std::vector<HostReal> reals;
  std::vector<void*> allocs;
  RealDesc* d_desc = nullptr;
  RealLive* d_live = nullptr;
  int2* d_tiles = nullptr;   // {group, tile} of the sweep
  int n_tiles = 0, grid_sweep = 0;
  // Every group's tiles; when the groups whose members have all finished hold
  // enough of them, the sweep runs over a list without them (d_tiles_live).
  std::vector<int2> tiles_all;
  int2* d_tiles_all = nullptr;
  int2* d_tiles_live = nullptr;
  bool tiles_compacted = false;
  int32_t* d_groups = nullptr;  // [n_groups][group] realization ids, -1 pads
  std::vector<int32_t> groups;
  double* stage = nullptr;       // staging buffer for state transfers of grouped realizations
  size_t stage_len = 0;
  double last_total_ms = 0, last_sweep_ms = 0;
  // captured step blocks, keyed by (steps, gate mode)
  std::map<std::pair<int64_t, uint32_t>, hipGraphExec_t> graphs;
  int32_t* d_sched = nullptr;
  std::vector<HostScan> scans;
};

namespace {

template <class T>
int dev_alloc(srs_ctx* c, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return SRS_OK;
  HIPCHK(hipMalloc((void**)p, count * sizeof(T)));
  c->allocs.push_back((void*)*p);
  return SRS_OK;
}

template <class T>
int dev_upload(srs_ctx* c, T** p, const std::vector<T>& v) {
  int rc = dev_alloc(c, p, v.size());
  if (rc) return rc;
  if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return SRS_OK;
}

int check_real(srs_ctx* c, int32_t r) {
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (!c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (r < 0 || r >= (int32_t)c->reals.size()) return srs::fail(SRS_E_ARG, "realization %d out of range", r);
  return SRS_OK;
}

using SweepFn = void (*)(const RealDesc*, const RealLive*, const int2*, int, const int*, int);

template <int BS, bool CMP, int KG>
SweepFn pick3(int ntg, int nts, int occ8, int lds) {
  if (KG == 1 && lds) {
    if (ntg) return nts ? k_sweep_lds<BS, true, true, CMP> : k_sweep_lds<BS, true, false, CMP>;
    return nts ? k_sweep_lds<BS, false, true, CMP> : k_sweep_lds<BS, false, false, CMP>;
  }
  if (occ8) {
    if (ntg) return nts ? k_sweep_o8<BS, true, true, CMP, KG> : k_sweep_o8<BS, true, false, CMP, KG>;
    return nts ? k_sweep_o8<BS, false, true, CMP, KG> : k_sweep_o8<BS, false, false, CMP, KG>;
  }
  if (ntg) return nts ? k_sweep<BS, true, true, CMP, KG> : k_sweep<BS, true, false, CMP, KG>;
  return nts ? k_sweep<BS, false, true, CMP, KG> : k_sweep<BS, false, false, CMP, KG>;
}

template <bool CMP, int KG>
SweepFn pick_tile(const srs_ctx* c) {
  switch (c->tile) {
    case 128: return pick3<128, CMP, KG>(c->nt_geometry, c->nt_state, c->occ8, c->lds);
    case 512: return pick3<512, CMP, KG>(c->nt_geometry, c->nt_state, c->occ8, c->lds);
    default: return pick3<256, CMP, KG>(c->nt_geometry, c->nt_state, c->occ8, c->lds);
  }
}

template <int KG>
SweepFn pick_kg(const srs_ctx* c) { return c->compact_all ? pick_tile<true, KG>(c) : pick_tile<false, KG>(c); }

SweepFn pick_sweep(const srs_ctx* c) {
  return c->group == 4 ? pick_kg<4>(c) : c->group == 2 ? pick_kg<2>(c) : pick_kg<1>(c);
}

void launch_sweep(srs_ctx* c) {
  if (c->grid_sweep == 0) return;  // every realization has finished
  // one realization: tile b is tile b of realization 0, no table lookups
  const int single = c->reals.size() == 1 ? 1 : 0;
  hipLaunchKernelGGL(pick_sweep(c), dim3(c->grid_sweep), dim3(c->tile), 0, c->stream, c->d_desc, c->d_live,
                     c->d_tiles, c->n_tiles, (const int*)c->d_groups, single);
}

int launch_step(srs_ctx* c, uint32_t gmode) {
  launch_sweep(c);
  hipLaunchKernelGGL(k_gate, dim3((unsigned)c->reals.size()), dim3(GATE_THREADS), 0, c->stream, c->d_desc,
                     c->d_live, gmode);
  return SRS_OK;
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
}

// The full tile list again, before the host changes live state (a finished
// realization may be restarted).
int tiles_full(srs_ctx* c) {
  if (!c->tiles_compacted) return SRS_OK;
  int rc = drop_graphs(c);
  if (rc) return rc;
  set_tiles(c, c->d_tiles_all, (int)c->tiles_all.size());
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
  c->tiles_compacted = true;
  return SRS_OK;
}

}  // namespace

namespace {
// Temporary device buffers freed when the call returns.
struct DevMem {
  std::vector<void*> ptrs;
  ~DevMem() { for (void* p : ptrs) (void)hipFree(p); }
  template <class T>
  hipError_t alloc(T** p, size_t count) {
    hipError_t e = hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back((void*)*p);
    return e;
  }
};
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
    case SRS_TUNE_NT_GEOMETRY: c->nt_geometry = value != 0; return SRS_OK;
    case SRS_TUNE_OCCUPANCY: c->occ8 = value == 8; return SRS_OK;
    case SRS_TUNE_NT_STATE: c->nt_state = value != 0; return SRS_OK;
    case SRS_TUNE_LDS: c->lds = value != 0; return SRS_OK;
    case SRS_TUNE_DROP_FINISHED:
      c->drop_finished = value != 0;
      return c->finalized ? tiles_full(c) : SRS_OK;
    case SRS_TUNE_GROUP:
      if (value != 1 && value != 2 && value != 4) return srs::fail(SRS_E_ARG, "group must be 1, 2 or 4");
      if (c->finalized) return srs::fail(SRS_E_STATE, "group must be set before srs_ctx_finalize");
      c->group = value;
      return SRS_OK;
    default: return srs::fail(SRS_E_ARG, "unknown tuning key %d", key);
  }
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
    d.gpu_layout = 0;
    d.rec_len = 7 + (d.ryr_out ? d.n_ryr : 0);
    d.rec_cap = record_capacity;
    if (record_capacity > 0) {
      float* fr;
      if ((rc = dev_alloc(c, &fr, (size_t)record_capacity * d.rec_len))) return rc;
      d.rec = fr;
    }
    d.key = hr.key;
    d.stream = hr.stream;
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
  c->d_tiles_all = c->d_tiles;
  c->groups = groups;
  if (c->reals.size() > 1 && (rc = dev_alloc(c, &c->d_tiles_live, tiles.size()))) return rc;
  // meshes' channel flags were updated by add_realization; refresh them
  for (auto& hmp : c->meshes)
    HIPCHK(hipMemcpy(hmp->d_flags, hmp->flags.data(), hmp->flags.size(), hipMemcpyHostToDevice));
  HIPCHK(hipDeviceSynchronize());
  c->finalized = true;
  return SRS_OK;
}

// Copies one realization's state between a host buffer in the single
// realization layout (np quads, then the TnC and CSQ planes) and its slot of
// the group buffer: contiguous through a device staging buffer, then a
// scatter (or gather) kernel for the slot.
static int slot_copy(srs_ctx* c, const RealDesc& d, double* dev, double* host, bool to_dev) {
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

static int live_rw(srs_ctx* c, int32_t r, RealLive* lv, bool write) {
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
static int graph_for(srs_ctx* c, int64_t block, uint32_t gmode, hipGraphExec_t* out) {
  auto key = std::make_pair(block, gmode);
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    for (int64_t k = 0; k < block; ++k) launch_step(c, gmode);
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t ge;
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    HIPCHK(hipGraphDestroy(g));
    it = c->graphs.emplace(key, ge).first;
  }
  *out = it->second;
  return SRS_OK;
}

constexpr int64_t kGraphBlock = 128, kGraphMin = 16;

int srs_explicit_step_prepare(srs_ctx* c, int64_t n_steps, uint32_t flags) {
  srs::clear_error();
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (!c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (n_steps < 0) return srs::fail(SRS_E_ARG, "negative step count");
  HIPCHK(hipSetDevice(c->device));
  const uint32_t gmode = G_AFTER_SWEEP | ((flags & SRS_STEP_PROTOCOL) ? G_PROTOCOL : 0u);
  int rc = tiles_compact(c);
  if (rc) return rc;
  hipGraphExec_t ge;
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
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int64_t done = 0;
  if (flags & SRS_STEP_GRAPH) {
    // 128-step blocks, then the rest as one block of its own when it is long enough
    hipGraphExec_t ge;
    if (n_steps >= kGraphBlock) {
      if ((rc = graph_for(c, kGraphBlock, gmode, &ge))) return rc;
      for (; n_steps - done >= kGraphBlock; done += kGraphBlock) HIPCHK(hipGraphLaunch(ge, c->stream));
    }
    if (n_steps - done >= kGraphMin) {
      if ((rc = graph_for(c, n_steps - done, gmode, &ge))) return rc;
      HIPCHK(hipGraphLaunch(ge, c->stream));
      done = n_steps;
    }
  }
  for (; done < n_steps; ++done) launch_step(c, gmode);
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

int srs_explicit_step_replay(srs_ctx* c, int32_t r, int64_t n_steps, const int32_t* schedule) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (c->reals.size() != 1) return srs::fail(SRS_E_ARG, "replay needs a single-realization context");
  HostReal& hr = c->reals[r];
  const int nr = (int)hr.ryr_ele.size();
  if (n_steps <= 0) return SRS_OK;
  if (!schedule && nr) return srs::fail(SRS_E_ARG, "null schedule");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  int32_t* ds = nullptr;
  if (nr) {
    HIPCHK(hipMalloc(&ds, (size_t)n_steps * nr * sizeof(int32_t)));
    HIPCHK(hipMemcpy(ds, schedule, (size_t)n_steps * nr * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(hr.desc.ryr_state, schedule, nr * sizeof(int32_t), hipMemcpyHostToDevice));
  }
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
  if (ds) (void)hipFree(ds);
  if (e != hipSuccess) return srs::fail(SRS_E_HIP, "replay launch: %s", hipGetErrorString(e));
  return SRS_OK;
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
  lv.rec_count = rest;
  return live_rw(c, r, &lv, true);
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
  std::vector<hipEvent_t> ev(3 * reps);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  const uint32_t gmode = G_AFTER_SWEEP;
  for (int k = 0; k < reps; ++k) {
    HIPCHK(hipEventRecord(ev[3 * k], c->stream));
    launch_sweep(c);
    HIPCHK(hipEventRecord(ev[3 * k + 1], c->stream));
    hipLaunchKernelGGL(k_gate, dim3((unsigned)c->reals.size()), dim3(GATE_THREADS), 0, c->stream, c->d_desc,
                       c->d_live, gmode);
    HIPCHK(hipEventRecord(ev[3 * k + 2], c->stream));
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  double s = 0, g = 0;
  for (int k = 0; k < reps; ++k) {
    float a = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]));
    HIPCHK(hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]));
    s += a;
    g += b;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  c->last_sweep_ms = s / reps;
  if (sweep_ms) *sweep_ms = s / reps;
  if (gating_ms) *gating_ms = g / reps;
  return SRS_OK;
}

// ----------------------------------------------------------------- linescan

int srs_ctx_add_scan(srs_ctx* c, int32_t mesh_id, const double* centers, const srs_scan_desc* sd, int32_t* scan_id,
                     int32_t* n_f_out, int32_t* n_ca_out) {
  srs::clear_error();
  if (!c || !centers || !sd || !scan_id) return srs::fail(SRS_E_ARG, "null argument");
  if (!c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (mesh_id < 0 || mesh_id >= (int)c->meshes.size()) return srs::fail(SRS_E_ARG, "bad mesh id %d", mesh_id);
  if (sd->axis < 0 || sd->axis > 2) return srs::fail(SRS_E_ARG, "scan axis must be 0, 1 or 2");
  if (!(sd->res_x_f > 0) || !(sd->res_x_ca > 0) || !(sd->pad_res > 0) || !(sd->var[0] > 0) || !(sd->var[1] > 0) ||
      !(sd->var[2] > 0))
    return srs::fail(SRS_E_ARG, "scan resolutions and PSF variances must be positive");
  const HostMesh& hm = *c->meshes[mesh_id];
  HostScan S;
  S.mesh = mesh_id;
  S.desc = *sd;
  // linescan/main.cpp:143-152
  S.n_f = (int)((sd->x_max - sd->x_min) / sd->res_x_f);
  S.n_ca = (int)((sd->x_max - sd->x_min) / sd->res_x_ca);
  if (S.n_f < 0 || S.n_ca < 0 || S.n_f > SCAN_MAX_POINTS || S.n_ca > (1 << 20))
    return srs::fail(SRS_E_ARG, "scan has %d fluorescence points (at most %d) and %d Ca points", S.n_f,
                     SCAN_MAX_POINTS, S.n_ca);
  // axis mapping and PSF widths, linescan/main.cpp:169-194
  ScanGeo& g = S.geo;
  g.base[0] = g.base[1] = g.base[2] = 0.0;
  g.axis = sd->axis;
  g.x_min = sd->x_min;
  g.res = sd->res_x_f;
  if (sd->axis == 0) { g.base[1] = sd->offset1; g.base[2] = sd->offset2; }
  else if (sd->axis == 1) { g.base[0] = sd->offset1; g.base[2] = sd->offset2; }
  else { g.base[0] = sd->offset1; g.base[1] = sd->offset2; }
  if (sd->axis == 2) { g.sx = std::sqrt(sd->var[2]); g.sy = std::sqrt(sd->var[0]); g.sz = std::sqrt(sd->var[1]); }
  else { g.sx = std::sqrt(sd->var[0]); g.sy = std::sqrt(sd->var[1]); g.sz = std::sqrt(sd->var[2]); }
  g.alpha = 1.0 / std::pow(2 * 3.14159, 1.5);  // the reference's pi, linescan/parameters.h:45-46

  // padding cells, linescan/main.cpp:119-141 and AddPointGrid :268-306
  std::vector<std::array<double, 3>> pc;
  std::vector<double> pv;
  auto add_grid = [&](double xmin, double xmax, double ymin, double ymax, double zmin, double zmax) {
    double dx = sd->pad_res, dy = sd->pad_res, dz = sd->pad_res;
    if (xmin > xmax) dx = -dx;
    if (ymin > ymax) dy = -dy;
    if (zmin > zmax) dz = -dz;
    const int nx = (int)((xmax - xmin) / dx), ny = (int)((ymax - ymin) / dy), nz = (int)((zmax - zmin) / dz);
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          pc.push_back({xmin + i * dx + 0.5 * dx, ymin + j * dy + 0.5 * dy, zmin + k * dz + 0.5 * dz});
          pv.push_back(std::fabs(dx * dy * dz));
        }
  };
  const double pmin = sd->pad_min, pmax = sd->pad_max;
  const double xlo = -sd->domain[0] / 2, xhi = sd->domain[0] / 2, ylo = -sd->domain[1] / 2, yhi = sd->domain[1] / 2;
  const double zlo = -sd->domain[2] / 2, zhi = sd->domain[2] / 2;
  add_grid(pmin, pmax, pmin, pmax, zhi, pmax);
  add_grid(pmin, pmax, pmin, pmax, zlo, pmin);
  add_grid(xlo, pmin, pmin, pmax, zlo, zhi);
  add_grid(xhi, pmax, pmin, pmax, zlo, zhi);
  add_grid(xlo, xhi, ylo, pmin, zlo, zhi);
  add_grid(xlo, xhi, yhi, pmax, zlo, zhi);

  // Nearest cell of each point (first minimum over mesh cells in reference
  // order, then padding cells), and the padding part of the blurred sum.
  const int n = hm.n, np = hm.np;
  auto nearest = [&](const double p[3]) -> int32_t {
    double rx = centers[0] - p[0], ry = centers[1] - p[1], rz = centers[2] - p[2];
    double r_min = rx * rx + ry * ry + rz * rz;
    int best = 0;
    for (int j = 1; j < n; ++j) {
      rx = centers[3 * (size_t)j] - p[0]; ry = centers[3 * (size_t)j + 1] - p[1]; rz = centers[3 * (size_t)j + 2] - p[2];
      const double dist = rx * rx + ry * ry + rz * rz;
      if (dist < r_min) { r_min = dist; best = j; }
    }
    int pbest = -1;
    for (size_t k = 0; k < pc.size(); ++k) {
      rx = pc[k][0] - p[0]; ry = pc[k][1] - p[1]; rz = pc[k][2] - p[2];
      const double dist = rx * rx + ry * ry + rz * rz;
      if (dist < r_min) { r_min = dist; pbest = (int)k; }
    }
    return pbest >= 0 ? -1 : hm.inv[best];
  };
  std::vector<int32_t> near_f(S.n_f), near_ca(S.n_ca);
  std::vector<double> pad(S.n_f);
  for (int i = 0; i < S.n_f; ++i) {
    double p[3] = {g.base[0], g.base[1], g.base[2]};
    p[g.axis] = sd->x_min + i * sd->res_x_f;
    near_f[i] = nearest(p);
    double sum = 0.0;
    for (size_t k = 0; k < pc.size(); ++k) {
      const double rx = pc[k][0] - p[0], ry = pc[k][1] - p[1], rz = pc[k][2] - p[2];
      if (rx * rx + ry * ry + rz * rz < 1.0)
        sum += pv[k] * sd->f0 *
               (g.alpha * std::exp(-0.5 * ((rx * rx / (g.sx * g.sx)) + (ry * ry / (g.sy * g.sy)) + (rz * rz / (g.sz * g.sz)))) /
                (g.sx * g.sy * g.sz));
    }
    pad[i] = sum;
  }
  for (int i = 0; i < S.n_ca; ++i) {
    double p[3] = {g.base[0], g.base[1], g.base[2]};
    p[g.axis] = sd->x_min + i * sd->res_x_ca;
    near_ca[i] = nearest(p);
  }
  std::vector<double> cen(3 * (size_t)np, 1e30);
  for (int k = 0; k < n; ++k)
    for (int q = 0; q < 3; ++q) cen[(size_t)q * np + k] = centers[3 * (size_t)hm.perm[k] + q];
  S.n_blocks = (n + SCAN_BS * SCAN_CELLS_PER_THREAD - 1) / (SCAN_BS * SCAN_CELLS_PER_THREAD);
  HIPCHK(hipSetDevice(c->device));
  int rc;
  if ((rc = dev_upload(c, &S.d_cen, cen))) return rc;
  if ((rc = dev_upload(c, &S.d_pad, pad))) return rc;
  if ((rc = dev_upload(c, &S.d_near_f, near_f))) return rc;
  if ((rc = dev_upload(c, &S.d_near_ca, near_ca))) return rc;
  if ((rc = dev_alloc(c, &S.d_part, (size_t)S.n_blocks * std::max(S.n_f, 1)))) return rc;
  if ((rc = dev_alloc(c, &S.d_out, (size_t)S.n_ca + 2 * (size_t)S.n_f + 1))) return rc;
  *scan_id = (int32_t)c->scans.size();
  c->scans.push_back(S);
  if (n_f_out) *n_f_out = S.n_f;
  if (n_ca_out) *n_ca_out = S.n_ca;
  return SRS_OK;
}

int srs_scan_sample(srs_ctx* c, int32_t r, int32_t scan_id, double* ca, double* caf, double* cafb) {
  int rc = check_real(c, r);
  if (rc) return rc;
  if (scan_id < 0 || scan_id >= (int)c->scans.size()) return srs::fail(SRS_E_ARG, "bad scan id %d", scan_id);
  const HostScan& S = c->scans[scan_id];
  if (c->reals[r].mesh != S.mesh) return srs::fail(SRS_E_ARG, "realization %d is not on the scan's mesh", r);
  HIPCHK(hipSetDevice(c->device));
  if (S.n_f > 0)
    hipLaunchKernelGGL(k_scan_blur, dim3(S.n_blocks), dim3(SCAN_BS), 0, c->stream, c->d_desc, c->d_live, r, nullptr,
                       S.d_cen, S.geo, S.n_f, S.d_part);
  hipLaunchKernelGGL(k_scan_final, dim3(1), dim3(256), 0, c->stream, c->d_desc, c->d_live, r, nullptr, S.n_blocks,
                     S.n_f, S.n_ca, S.d_part, S.d_pad, S.d_near_f, S.d_near_ca, S.desc.f0, S.desc.c0, S.d_out, 0);
  HIPCHK(hipGetLastError());
  std::vector<double> out((size_t)S.n_ca + 2 * (size_t)S.n_f);
  HIPCHK(hipMemcpyAsync(out.data(), S.d_out, out.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (ca) std::copy(out.begin(), out.begin() + S.n_ca, ca);
  if (caf) std::copy(out.begin() + S.n_ca, out.begin() + S.n_ca + S.n_f, caf);
  if (cafb) std::copy(out.begin() + S.n_ca + S.n_f, out.end(), cafb);
  return SRS_OK;
}

// ----------------------------------------------------------------- adjoint

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
  const int64_t n_ck = (n_steps + ckpt - 1) / ckpt;
  const int64_t n_ck_kept = want_grad ? n_ck : 1;  // a loss alone keeps only the start state
  const int nblk = (hm.n + ADJ_BS - 1) / ADJ_BS;
  std::vector<size_t> soff(n_mem + 1, 0);
  size_t max_sched = 0;
  for (int m = 0; m < n_mem; ++m) {
    const size_t k = (size_t)n_steps * c->reals[ids[m]].ryr_ele.size();
    soff[m + 1] = soff[m] + k;
    max_sched = std::max(max_sched, k);
  }
  // device bytes per member of a sub-batch
  const size_t per = 8 * n6 * (size_t)(n_ck_kept + (want_grad ? ckpt + 2 : 0)) + (want_grad ? 64 * np : 0) +
                     8 * (size_t)n_samples * nout + 8 * (size_t)S.n_blocks * nf + 8 * (size_t)nblk * ADJ_NT +
                     8 * (size_t)n_samples * nf + 4 * max_sched + 64;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  size_t mem_free = 0, mem_total = 0;
  HIPCHK(hipMemGetInfo(&mem_free, &mem_total));
  const size_t budget = mem_free / 10 * 9;
  int bs = max_batch > 0 ? std::min(max_batch, n_mem) : n_mem;
  bs = (int)std::min<size_t>((size_t)bs, budget / per);
  if (bs < 1)
    return srs::fail(SRS_E_ARG, "checkpoint plan needs %.3g GB per realization, %.3g GB free; raise or lower the "
                     "checkpoint interval (about sqrt(n_steps) needs least)", per / 1e9, mem_free / 1e9);

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
  d4 *w0 = nullptr, *w1 = nullptr;
  int32_t *d_sched = nullptr, *d_rid = nullptr;
  int2* d_sub_tiles = nullptr;
  const int nt_mesh = (hm.n + c->tile - 1) / c->tile;
  HIPCHK(M.alloc(&ck, (size_t)n_ck_kept * bs * n6));
  HIPCHK(M.alloc(&d_sub_tiles, (size_t)bs * nt_mesh));
  HIPCHK(M.alloc(&samp, (size_t)n_samples * bs * nout));
  HIPCHK(M.alloc(&part, (size_t)bs * S.n_blocks * nf));
  HIPCHK(M.alloc(&d_sched, std::max<size_t>(1, (size_t)bs * max_sched)));
  HIPCHK(M.alloc(&d_rid, (size_t)bs));
  if (want_grad) {
    HIPCHK(M.alloc(&lam0, (size_t)bs * n6));
    HIPCHK(M.alloc(&lam1, (size_t)bs * n6));
    HIPCHK(M.alloc(&w0, (size_t)bs * np));
    HIPCHK(M.alloc(&w1, (size_t)bs * np));
    HIPCHK(M.alloc(&seg, (size_t)ckpt * bs * n6));
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
    for (int y = 0; y < nb; ++y) {
      const int r = ids[m0 + y];
      lv[r].done = 0;
      lv[r].ffwd = 0;
      lv[r].step = 0;
      lv[r].sched_base = 0;
      HostReal& hr = c->reals[r];
      RealDesc dsc = hr.desc;
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
    auto ck_at = [&](int64_t k) { return ck + (size_t)k * nb * n6; };

    // forward, with checkpoints and scan samples
    hipLaunchKernelGGL(k_seg_start, dim3(nb), dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid, (int64_t)0);
    int ks = 0;
    for (int64_t n = 0; n < n_steps && e == hipSuccess; ++n) {
      if (n % ckpt == 0 && n / ckpt < n_ck_kept)
        hipLaunchKernelGGL(k_batch_state, g_cp, dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid,
                           ck_at(n / ckpt), 1);
      launch_step(c, gmode);
      if (ks < n_samples && sample_steps[ks] == n + 1) scan_into(samp + (size_t)(ks++) * nb * nout);
      e = hipGetLastError();
      if (e == hipSuccess && (n & 4095) == 4095) e = hipStreamSynchronize(c->stream);
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
      if (e == hipSuccess) e = hipMemsetAsync(gpart, 0, (size_t)nb * nblk * ADJ_NT * sizeof(double), c->stream);
      double* lam = lam0;   // lambda_{n+1}
      double* lout = lam1;  // lambda_n
      d4* wcur = w0;        // w of step n
      d4* wnext = w1;       // w of step n - 1
      ks = n_samples - 1;
      // the loss source of a sample after the last step; the step kernel
      // adds the others to the lambda it produces
      if (sample_steps[ks] == n_steps) {
        hipLaunchKernelGGL(k_scan_adj, g_adj, dim3(ADJ_BS), 0, c->stream, c->d_desc, d_rid, S.d_cen, S.geo, nf,
                           dg + (size_t)ks * nf, n_samples * nf, lam);
        --ks;
      }
      for (int64_t s0 = (n_ck - 1) * ckpt; s0 >= 0 && e == hipSuccess; s0 -= ckpt) {
        const int64_t s1 = std::min<int64_t>(n_steps, s0 + ckpt);
        // recompute the segment from its checkpoint, the sweep writing each
        // state into its slot of the segment
        hipLaunchKernelGGL(k_seg_start, dim3(nb), dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid, s0);
        e = hipMemcpyAsync(seg, ck_at(s0 / ckpt), (size_t)nb * n6 * sizeof(double), hipMemcpyDeviceToDevice,
                           c->stream);
        for (int64_t n = s0; n < s1 - 1 && e == hipSuccess; ++n) {
          hipLaunchKernelGGL(k_seg_ptrs, dim3((nb + 63) / 64), dim3(64), 0, c->stream, c->d_desc, c->d_live, d_rid,
                             seg, s0, nb);
          launch_step(c, gmode);
          e = hipGetLastError();
        }
        // backward through the segment
        for (int64_t n = s1 - 1; n >= s0 && e == hipSuccess; --n) {
          const double* Xn = seg + (size_t)(n - s0) * nb * n6;
          if (n == s1 - 1)
            hipLaunchKernelGGL(k_adj_w, g_adj, dim3(ADJ_BS), 0, c->stream, c->d_desc, d_rid, Xn, lam, wcur);
          const bool scan = ks >= 0 && sample_steps[ks] == n;
          AdjNext nx{n > s0 ? Xn - nb * n6 : nullptr, wnext, S.d_cen, scan ? dg + (size_t)ks * nf : nullptr,
                     S.geo, nf, n_samples * nf};
          if (scan) --ks;
          auto kern = c->compact_all ? (scan ? k_adj_step<true, true> : k_adj_step<true, false>)
                                     : (scan ? k_adj_step<false, true> : k_adj_step<false, false>);
          hipLaunchKernelGGL(kern, g_adj, dim3(ADJ_BS), 0, c->stream, c->d_desc, c->d_live, d_rid, Xn, lam, wcur, n,
                             lout, gpart, nx);
          std::swap(lam, lout);
          std::swap(wcur, wnext);
          e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      }
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
    hipLaunchKernelGGL(k_batch_state, g_cp, dim3(256), 0, c->stream, c->d_desc, c->d_live, d_rid, ck_at(0), 0);
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
                           const int64_t* sample_steps, const double* target, int32_t ckpt, const void* loss) {
  if (!c) return srs::fail(SRS_E_ARG, "null context");
  if (!c->finalized) return srs::fail(SRS_E_STATE, "context not finalized");
  if (c->group != 1) return srs::fail(SRS_E_ARG, "the adjoint needs realization groups of size 1");
  if (scan_id < 0 || scan_id >= (int)c->scans.size()) return srs::fail(SRS_E_ARG, "bad scan id %d", scan_id);
  if (c->scans[scan_id].n_f <= 0) return srs::fail(SRS_E_ARG, "the scan has no fluorescence points");
  if (n_steps <= 0 || n_samples <= 0 || !sample_steps || !target || !loss) return srs::fail(SRS_E_ARG, "bad arguments");
  if (ckpt <= 0 || ckpt > 65536) return srs::fail(SRS_E_ARG, "checkpoint interval must be 1..65536");
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
  if (ckpt == 0) ckpt = (int32_t)std::max<double>(1.0, std::ceil(std::sqrt((double)std::max<int64_t>(n_steps, 1))));
  if ((rc = loss_grad_check(c, scan_id, n_steps, n_samples, sample_steps, targets, ckpt, losses))) return rc;
  std::vector<int32_t> ids(real_ids, real_ids + n_real);
  std::vector<char> seen(c->reals.size(), 0);
  bool any_ryr = false;
  for (int32_t r : ids) {
    if ((rc = check_real(c, r))) return rc;
    if (seen[r]++) return srs::fail(SRS_E_ARG, "realization %d listed twice", r);
    if (c->reals[r].mesh != c->scans[scan_id].mesh)
      return srs::fail(SRS_E_ARG, "realization %d is not on the scan's mesh", r);
    any_ryr |= !c->reals[r].ryr_ele.empty();
  }
  if (any_ryr && !schedules) return srs::fail(SRS_E_ARG, "null schedules");
  return loss_grad_run(c, ids, scan_id, n_steps, schedules, n_samples, sample_steps, targets, ckpt, max_batch, losses,
                       grads, scans);
}

}  // extern "C"
