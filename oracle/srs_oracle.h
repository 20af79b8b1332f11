/*
 * srs_oracle.h -- CPU restatement of the reference spark time-step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product path (srs_amd/) may
 * include, link or call this code; it exists so that tests/, bench.py's
 * cpu_baseline leg and __graft_entry__.smoke() can check the HIP path.
 *
 * Every function follows a named piece of the reference (mwalker174/srs),
 * cited as file:line.  Arithmetic is written in the reference's operation
 * order and the file is compiled with -ffp-contract=off so that, on x86-64,
 * results are bit-identical to the reference's own cpu/ build.
 *
 * Pinning (what checks each function, see tests/):
 *   orc_explicit_step / orc_gating   -> bit-exact against oracle/_ref (the
 *                                       reference cpu/RunSimulation.cpp itself,
 *                                       compiled by oracle/build.py), and
 *                                       against closed forms (equilibrium fixed
 *                                       point, mass conservation, Po law).
 *   orc_philox4x64                   -> numpy.random.Philox (independent impl.)
 *   orc_assemble_operator            -> dense brute-force assembly, row-sum and
 *                                       V-weighted symmetry identities.
 *   orc_bfs_order                    -> BFS characterisation (levels, FIFO parent
 *                                       order, face-order tie break).
 *   orc_equilibrium_state            -> binding-reaction fixed point.
 *   orc_boundary_rates               -> bit-exact against oracle/_ref, and the
 *                                       closed form D Face_C b / V0 per face.
 *   orc_states_record                -> a hand-computed record (every field
 *                                       from its definition).
 *   orc_linescan                     -> closed forms: one cell, the PSF
 *                                       integral over a uniform padding grid,
 *                                       brute-force nearest cell.
 *   orc_open_random_ryr              -> selection law: RyR i is picked with
 *                                       frequency rate_i / sum(rates), checked
 *                                       in two limits where the rate is known
 *                                       in closed form (linear phi alone, then
 *                                       cytosolic Ca alone).
 */
#ifndef SRS_ORACLE_H_
#define SRS_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameter vector layout: same indices as the reference ParamStruct
 * (cpu/parameters.h:18-96). */
enum {
  ORC_I_OUTPUT_GRID_FLAG = 0, ORC_I_FLAG_FIDELITY_SIM = 1, ORC_I_FLAG_CLOCK_SEED = 2,
  ORC_I_RAND_SEED_0 = 3, ORC_I_RAND_ADD_A = 4, ORC_I_RAND_ADD_B = 5,
  ORC_I_INITIALIZE_FROM_FILE = 6, ORC_I_CUDA_DEVICE = 7, ORC_I_FLAG_SECONDARY_GRADIENT = 8,
  ORC_I_FLAG_NO_FLUX_BOUNDARY = 9, ORC_I_N_RYR_FIDELITY = 10, ORC_I_FLAG_LCC = 11,
  ORC_I_LCC_INDEX = 12, ORC_I_FLAG_V_CLAMP = 13, ORC_I_FLAG_NO_GATING = 14,
  ORC_I_GATING_FFWD = 15, ORC_I_FLAG_RYR_OUT = 16, ORC_I_N_SIMS = 17, ORC_I_INIT_RYR = 18,
  ORC_N_INT = 19
};
enum {
  ORC_T_FINAL = 0, ORC_T_STEP = 1, ORC_OUTPUT_STATES_INTERVAL = 2, ORC_OUTPUT_GRID_INTERVAL = 3,
  ORC_C_MIN = 4, ORC_D_CA = 5, ORC_D_CA_JSR = 6, ORC_D_ATP = 7, ORC_D_CMDN = 8, ORC_D_TRPN = 9,
  ORC_D_CSQN = 10, ORC_D_DYE = 11, ORC_D_EGTA = 12, ORC_C_0_CA = 13, ORC_C_0_CA_JSR = 14,
  ORC_DEFAULT_VM = 15, ORC_DEFAULT_CA_NSR = 16, ORC_CA_0 = 17, ORC_B_TOT_ATP = 18,
  ORC_K_OFF_ATP = 19, ORC_K_ON_ATP = 20, ORC_B_TOT_CMDN = 21, ORC_K_OFF_CMDN = 22,
  ORC_K_ON_CMDN = 23, ORC_B_TOT_TRPN = 24, ORC_K_OFF_TRPN = 25, ORC_K_ON_TRPN = 26,
  ORC_B_TOT_SL_HI = 27, ORC_K_D_SL_HI = 28, ORC_B_TOT_SL_LO = 29, ORC_K_D_SL_LO = 30,
  ORC_B_TOT_DYE = 31, ORC_K_OFF_DYE = 32, ORC_K_ON_DYE = 33, ORC_B_T_JSR = 34,
  ORC_K_M_JSR = 35, ORC_V_REFILL = 36, ORC_V_RYR = 37, ORC_V_CELL = 38, ORC_K_ON_CSQN = 39,
  ORC_K_OFF_CSQN = 40, ORC_RYR_ETA = 41, ORC_RYR_A_STAR = 42, ORC_RYR_EPS_CC = 43,
  ORC_RYR_EPS_OO = 44, ORC_RYR_K_PLUS = 45, ORC_RYR_K_MINUS = 46, ORC_RYR_PHI_M = 47,
  ORC_RYR_PHI_B = 48, ORC_K_D_I = 49, ORC_K_D_SR = 50, ORC_A_P = 51, ORC_LCC_DURATION = 52,
  ORC_T_CLAMP = 53, ORC_V_CLAMP = 54, ORC_INIT_RYR_HOLD = 55, ORC_N_REAL = 56
};

/* Species order (cpu/parameters.h:136-141). */
enum { ORC_CA = 0, ORC_ATP = 1, ORC_CMDN = 2, ORC_TRPN = 3, ORC_CSQN = 4, ORC_DYE = 5, ORC_NS = 6 };

typedef struct {
  int n_ele;
  const double* aij;     /* [5*n_ele] diag first, then up to 4 off-diagonals      */
  const int* nij;        /* [4*n_ele] neighbour columns, self where absent        */
  const double* br;      /* [6*n_ele] species-major boundary rates (may be NULL)  */
  const int* domain;     /* [n_ele] 0 cytosol, 1 jSR                              */
  const int* tropc;      /* [n_ele] 0/1                                           */
  const double* V0;      /* [n_ele] element volumes                               */
  const double* tt_area; /* [n_ele] TT surface area                               */
} orc_grid;

typedef struct {
  int n_ryr, n_lcc;
  int* ryr_state; const int* ryr_ele; const int* ryr_jsr; const int* ryr_nb; /* nb: [4*n_ryr] */
  int* lcc_state; int* lccv_state; const int* lcc_ele;
  double* R;             /* [n_ryr + n_lcc] integrated-hazard residuals */
} orc_chan;

/* Uniform source for gating: glibc rand() in the reference's call order, or
 * counter-based Philox4x64-10. */
enum { ORC_RNG_GLIBC = 0, ORC_RNG_PHILOX = 1 };
typedef struct {
  int kind;
  uint64_t key0, key1;   /* Philox key */
  uint64_t realization;  /* Philox counter word 2 */
  uint64_t step;         /* Philox counter word 0 (set by caller per step) */
} orc_rng;

void orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
double orc_u64_to_open01(uint64_t x);

/* cpu/RunSimulation.cpp:591-648 (ExplicitStep without the gating call).
 * X, buf: species-major [6][n_ele].  gv: {Vm, CaNSR}. */
void orc_explicit_step(double dt, const orc_grid* g, const double* reals, double* X, double* buf,
                       double* gv, const orc_chan* ch);

/* cpu/RunSimulation.cpp:656-674, 166-400 (UpdateChannelGating). */
void orc_gating(double dt, const orc_grid* g, const double* reals, const double* X,
                const double* gv, orc_chan* ch, orc_rng* rng);

/* cpu/InitFuncs.cpp:1058-1112, 1273-1322.  face_type uses the simulation codes
 * (0 default, 1 bound, 2 noflux).  Returns 0, or -1 if a row has > 4 neighbours. */
int orc_assemble_operator(int n_ele, int n_face, const int* ele1, const int* ele2,
                          const int* face_type, const double* face_c, const double* V0,
                          double* aij, int* nij);

/* cpu/InitFuncs.cpp:1012-1026. diff[6], boundaries[6]; br out [6][n_ele]. */
void orc_boundary_rates(int n_ele, int n_face, const int* ele1, const int* face_type,
                        const double* face_c, const double* V0, const double* diff,
                        const double* boundaries, double* br);

/* cpu/InitFuncs.cpp:679-731.  visited[i] = original index of new element i.
 * Returns the number of visited elements (== n_ele for a connected mesh). */
int orc_bfs_order(int n_ele, int n_face, const int* ele1, const int* ele2, int* visited);

/* cpu/InitFuncs.cpp:890-954.  Writes X[6][n_ele]; boundaries[6] out (may be NULL). */
void orc_equilibrium_state(int n_ele, const int* domain, const int* tropc, const double* reals,
                           double* X, double* boundaries);

/* cpu/OutputFuncs.cpp:223-350 (and gpu/OutputFuncs.cpp per-RyR Ca variant).
 * Returns the number of floats written to rec. */
int orc_states_record(double t, const double* X, int n_ele, const double* gv, const orc_chan* ch,
                      const double* reals, int ryr_out, int gpu_layout, float* rec);

/* cpu/InitFuncs.cpp:1420-1475 with the uniform supplied by the caller. */
int orc_open_random_ryr(const double* X, int n_ele, const double* reals, orc_chan* ch, double u01);

/* Linescan of one state (linescan/main.cpp:96-249, AddPointGrid :268-306,
 * normpdf3d linescan/parameters.h:60-62).  axis 0/1/2 = X/Y/Z. */
typedef struct orc_scan {
  int axis;
  double x_min, x_max, res_f, res_ca, off1, off2;
  double dom[3], pad_min, pad_max, pad_res, var[3], f0, c0;
} orc_scan;
/* Number of fluorescence and Ca points. */
void orc_linescan_size(const orc_scan* s, int* n_f, int* n_ca);
/* centers [n][3]; outputs ca[n_ca], caf[n_f], cafb[n_f]. */
void orc_linescan(const orc_scan* s, int n, const double* centers, const double* V0, const double* ca,
                  const double* caf, double* out_ca, double* out_caf, double* out_cafb);

#ifdef __cplusplus
}
#endif
#endif
