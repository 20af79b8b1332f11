// dev_common.h -- types, layout helpers and host-side declarations shared by
// the HIP translation units of libsrs_amd (device.hip: the step kernels and the
// context; linescan.hip: the line-scan operator; adjoint.hip: recording,
// replayed segments and the adjoint).  Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstddef>
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

namespace srs_dev {

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
  const double* br;       // [4][np] Ca, Fluo, ATP, CaM boundary rates at this realization's D, or null
  const double* br_mesh;  // [4][np] the mesh's boundary rates, built with the D in brd
  double brd[4];          // the D_Ca, D_Fluo, D_ATP, D_CaM the mesh's boundary rates were built with
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
  // model extensions (srs_extensions): Ca diffusion with D_CAJSR in jSR
  // elements, SR-site rapid buffer (b_sr uM, dissociation constant k_sr uM)
  int32_t jsr_diff;
  double b_sr, k_sr;
  double P[SRS_N_REAL];
  int32_t I[SRS_N_INT];
};

// A boundary rate at a realization's current D.  The mesh's planes hold
// D_b Face_C b / V0 summed over an element's bound faces
// (cpu/InitFuncs.cpp:1019-1024), with D_b the coefficient they were built
// with (brd); the rate is linear in D, so a realization whose D has since been
// changed (srs_params_set) gets planes of its own scaled by D / D_b, and one
// whose D is unchanged uses the mesh's bits as they are.
__host__ __device__ inline double br_scaled(double v, double D, double D_b) {
  return (D == D_b || D_b == 0.0) ? v : v / D_b * D;
}
// The D-free part Face_C b / V0 of a boundary rate (its derivative in D).
__host__ __device__ inline double br_unit(double v, double D_b) { return D_b == 0.0 ? 0.0 : v / D_b; }

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

// Sum of v over the 64 lanes of a wavefront, left in lane 63, by DPP lane
// moves (no LDS traffic): sums within quads, then within each row of 16
// lanes (row_shr 4 and 8 leave the row's sum in its lane 15), then row 0
// into row 1 and rows 0-1 into rows 2-3 by row broadcasts.  Every lane must
// be active.  A fixed order, so the result is deterministic.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWS, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWS, 0xf, true);
  return __hiloint2double(hi, lo);
}
// The first four steps alone: each row's sum in its lane 15.
__device__ __forceinline__ double row_sum_dpp(double v) {
  v += dpp_f64<0xb1, 0xf>(v);   // quad_perm [1,0,3,2]
  v += dpp_f64<0x4e, 0xf>(v);   // quad_perm [2,3,0,1]
  v += dpp_f64<0x114, 0xf>(v);  // row_shr:4
  v += dpp_f64<0x118, 0xf>(v);  // row_shr:8
  return v;
}
__device__ __forceinline__ double wave_sum_dpp(double v) {
  v = row_sum_dpp(v);
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  return v;
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

// --------------------------------------------- line-scan geometry and sizes
// linescan/main.cpp:195-249.  The blurred fluorescence of scan point i is
// sum_j V0_j CaF_j w(c_j - p_i) over cells with |c_j - p_i|^2 < 1, with the
// Gaussian PSF w of linescan/parameters.h:60-62.  Mesh cells are summed on the
// device, one partial sum per workgroup and point, then the partials in
// workgroup order; the padding cells' part is a per-point constant.
//
// Along the line only the axis term of the PSF's exponent changes, by a
// quadratic in the point index, so a cell's weights at consecutive points
// follow w(i+1) = w(i) r(i), r(i+1) = r(i) q with q = exp(-res^2 / s_a^2): two
// products per point instead of an exponential.  The walk restarts from the
// closed form at every SCAN_ANCHOR-th point, which keeps the rounding of a
// weight within a few tens of ulps; a weight is a function of the cell and
// the point alone, so the blur and the adjoint's loss source use the same
// numbers.  PSFs so narrow that an anchor could underflow while its segment
// holds a weight above exp(-260), or a ratio overflow, within a cell's reach
// take the closed form at every point (walk = 0).
struct ScanGeo {
  double base[3];   // scan line point with the axis coordinate zero
  int axis;
  double x_min, res;
  double sx, sy, sz, alpha;
  double inv[3];    // 1 / s^2 per coordinate (s = sx, sy, sz)
  double ia;        // inv[axis] (device code indexes no array by axis: the
                    // struct then stays in scalar registers)
  double kw;        // alpha / (sx sy sz)
  double q;         // exp(-res^2 inv[axis])
  int walk;         // 1: the recurrence above; 0: the closed form everywhere
};

constexpr int SCAN_ANCHOR = 8;

// One cell's terms of the walk: its axis coordinate, the perpendicular
// squares in the reference's summation order (distance^2 = (da^2 + u) + v,
// da the axis difference) and the perpendicular part of the PSF exponent.
struct ScanCell {
  double ca, u, v, eperp;
};
__device__ __forceinline__ ScanCell scan_cell(const ScanGeo& g, double cx, double cy, double cz) {
  const double r0 = cx - g.base[0], r1 = cy - g.base[1], r2 = cz - g.base[2];
  const double t0 = r0 * r0, t1 = r1 * r1, t2 = r2 * r2;
  ScanCell c;
  if (g.axis == 0) { c.ca = cx; c.u = t1; c.v = t2; c.eperp = t1 * g.inv[1] + t2 * g.inv[2]; }
  else if (g.axis == 1) { c.ca = cy; c.u = t0; c.v = t2; c.eperp = t0 * g.inv[0] + t2 * g.inv[2]; }
  else { c.ca = cz; c.u = t0 + t1; c.v = 0.0; c.eperp = t0 * g.inv[0] + t1 * g.inv[1]; }
  return c;
}
// Axis coordinate of scan point i (linescan/main.cpp:199-200).
__device__ __forceinline__ double scan_pa(const ScanGeo& g, int i) { return g.x_min + i * g.res; }
// The reference's distance^2 between the cell and point i, bit for bit.
__device__ __forceinline__ double scan_dist2(const ScanCell& c, double pa) {
  const double da = c.ca - pa;
  return (da * da + c.u) + c.v;
}
// Weight at point i from the closed form and the ratio to point i + 1.
__device__ __forceinline__ void scan_anchor(const ScanGeo& g, const ScanCell& c, double pa, double* w, double* r) {
  const double da = c.ca - pa, ia = g.ia;
  *w = g.kw * exp(-0.5 * (da * da * ia + c.eperp));
  *r = exp(g.res * (da - 0.5 * g.res) * ia);
}
// The points within the reference's exact cut-off of a cell, [lo, hi] (none
// if lo > hi), inside a superset [lo, hi] (scan_point_range).  The point
// coordinate x_min + i res rises with i and every rounding step is monotonic,
// so the distance^2 falls and then rises along the line: the points that pass
// are one run, found by testing its ends, and the callers' per-point test
// becomes a range check (the same points, the same bits).
__device__ __forceinline__ void scan_exact_range(const ScanGeo& g, const ScanCell& c, int* lo, int* hi) {
  while (*lo <= *hi && !(scan_dist2(c, scan_pa(g, *lo)) < 1.0)) ++*lo;
  while (*hi >= *lo && !(scan_dist2(c, scan_pa(g, *hi)) < 1.0)) --*hi;
}
// The first point of a walk that covers lo: its anchor.
__device__ __forceinline__ int scan_walk_start(const ScanGeo& g, int lo) {
  return g.walk ? lo - lo % SCAN_ANCHOR : lo;
}
__device__ __forceinline__ bool scan_is_anchor(const ScanGeo& g, int i, int i0) {
  return !g.walk || i == i0 || i % SCAN_ANCHOR == 0;
}

// The scan points a cell at (cx, cy, cz) can reach (distance < 1 um): none
// (lo > hi) when its distance to the line is 1 um or more, else those whose
// axis coordinate lies within sqrt(1 - t^2) of the cell's, t the distance to
// the line, widened by one point each way and by 1e-9 on t^2 against the
// rounding of the exact test, which the callers still apply.
__device__ __forceinline__ void scan_point_range(const ScanGeo& g, int n_f, double cx, double cy, double cz, int* lo,
                                                 int* hi) {
  const double ca = g.axis == 0 ? cx : g.axis == 1 ? cy : cz;
  double t2 = 0.0;
  if (g.axis != 0) t2 += (cx - g.base[0]) * (cx - g.base[0]);
  if (g.axis != 1) t2 += (cy - g.base[1]) * (cy - g.base[1]);
  if (g.axis != 2) t2 += (cz - g.base[2]) * (cz - g.base[2]);
  if (!(t2 < 1.0 + 1e-9)) {
    *lo = n_f;
    *hi = -1;
    return;
  }
  const double half = sqrt(1.0 + 1e-9 - t2);
  *lo = max(0, (int)floor((ca - half - g.x_min) / g.res) - 1);
  *hi = min(n_f - 1, (int)ceil((ca + half - g.x_min) / g.res) + 1);
}

constexpr int SCAN_BS = 256;
constexpr int SCAN_MAX_POINTS = 1024;
constexpr int SCAN_CELLS_PER_THREAD = 4;

// Realization of workgroup row blockIdx.y in a batched launch: rid[blockIdx.y],
// or r when there is no table (a launch for one realization).
__device__ __forceinline__ int batch_real(const int32_t* __restrict__ rid, int r) { return rid ? rid[blockIdx.y] : r; }

// ----------------------------------------------------------- host-side state

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
  // The D of the four diffusing species that br was built with
  // (D Face_C b / V0, cpu/InitFuncs.cpp:1019-1024): those of the first
  // realization added on the mesh, which every later one must share.
  double br_d[4] = {0, 0, 0, 0};
  bool br_d_set = false;
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

struct HostReal {
  int32_t mesh = 0;
  double* d_br_own = nullptr;  // boundary rates rescaled to this realization's D (srs_params_set)
  srs_extensions ext{0, 0.0, 0.6};
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

// cpu/RunSimulation.cpp:85-88, the parts that depend on CaNSR only.
void serca_terms(double nsr, const double* P, double* out);

}  // namespace srs_dev

using namespace srs_dev;

struct srs_ctx {
  int device = 0;
  int tile = 0;          // elements per k_sweep workgroup; 0: chosen by srs_ctx_finalize
  int nt_geometry = 1;   // non-temporal loads of operator rows and flags
  int nt_state = 0;      // non-temporal stores of the new state
  int plane_pad = 0;     // extra elements per SoA plane
  int compact = 1;       // compress operator rows when a mesh allows it
  int occ = 0;           // sweep register budget: 0 compiler's choice, 6 or 8 wavefronts per SIMD
  int group = 1;         // realizations per realization group (1, 2 or 4)
  int lds = 1;           // stage each tile's quads in LDS for the neighbour gathers (groups of 1)
  int drop_finished = 1; // sweep without the tiles of groups that have all finished
  int gated = 1;         // fold each step's gating into the next step's sweep (k_sweep_gated) where it applies
  int alt_order = 0;      // single-realization gated sweep: tiles in reverse order on alternate steps
  int gpu_layout = 0;     // states records in the GPU program's layout (Ca_cyt at each RyR appended)
  int adj_levels = 0;     // adjoint checkpoint levels, 0: the fewest that fit the sub-batch
  int adj_waves = 0;      // adjoint reverse-step register budget: 0 compiler's choice, 4 wavefronts per SIMD
  int adj_w = 0;          // adjoint w: 0 by launch size, 1 written by each step, 2 formed in the kernel
  int adj_last_levels = 0, adj_last_batch = 0;  // the plan of the last loss/gradient call
  int64_t adj_last_strides[6] = {};
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
  std::vector<int2> tiles_cur;    // host copy of the tile list the sweep runs over (d_tiles)
  // Realization subsets (SRS_TUNE_SUBSET_MB): consecutive groups of the tile
  // list whose state takes at most subset_mb MB, each stepped through a
  // whole block of steps before the next, so that its state stays in the
  // last-level cache between steps.
  struct Subset { int t0, nt, r0, nr; };
  int subset_mb = 0;
  std::vector<Subset> subsets;
  bool subsets_valid = false;
  int32_t* d_sub_rids = nullptr;  // realization ids of the subsets, concatenated
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

namespace srs_dev {

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

// Context helpers (device.hip).
int check_real(srs_ctx* c, int32_t r);
void launch_sweep(srs_ctx* c);
int launch_step(srs_ctx* c, uint32_t gmode);
// m steps of a single-realization context through the gated sweep, starting
// from live state index cur, ending with the gating of the last step
void launch_gated_block(srs_ctx* c, int64_t m, int cur, uint32_t gmode);
// One step through the single-realization gated sweep; xo_at/xn_at (both or
// neither) replace X[cur] -> X[cur ^ 1] as the step's state buffers.
void launch_gated(srs_ctx* c, int cur, uint32_t gmode, bool proto, const double* xo_at = nullptr,
                  double* xn_at = nullptr);
void launch_gate_single(srs_ctx* c, uint32_t gmode);
int drop_graphs(srs_ctx* c);
int tiles_full(srs_ctx* c);
int tiles_compact(srs_ctx* c);
int plan_subsets(srs_ctx* c);
void launch_subset_block(srs_ctx* c, int64_t m, uint32_t gmode);
int live_rw(srs_ctx* c, int32_t r, RealLive* lv, bool write);
int slot_copy(srs_ctx* c, const RealDesc& d, double* dev, double* host, bool to_dev);

// The line-scan kernels (linescan.hip), also launched by the adjoint.
__global__ void k_scan_blur(const RealDesc* __restrict__ D, const RealLive* __restrict__ L, int r,
                            const int32_t* __restrict__ rid, const double* __restrict__ cen, ScanGeo g, int n_f,
                            double* __restrict__ part);
__global__ void k_scan_final(const RealDesc* __restrict__ D, const RealLive* __restrict__ L, int r,
                             const int32_t* __restrict__ rid, int n_blocks, int n_f, int n_ca,
                             const double* __restrict__ part, const double* __restrict__ pad,
                             const int32_t* __restrict__ near_f, const int32_t* __restrict__ near_ca, double f0,
                             double c0, double* __restrict__ out, int out_stride);

}  // namespace srs_dev
