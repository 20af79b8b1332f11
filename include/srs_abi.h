/*
 * srs_abi.h -- C ABI of the MI355X spark time-step (srs_amd).
 *
 * The reference (mwalker174/srs) has no plugin API: its seam is a set of
 * plain C-style functions over C structs, declared in cpu/parameters.h:252-275
 * (CPU path) and gpu/DeviceStructs.h:60-71 (CUDA path).  Each entry point
 * below names the reference function(s) it replaces.  Everything is plain C:
 * pointers, sizes and status codes; no C++ or torch types cross this line.
 *
 * Conventions
 *  - Every call returns SRS_OK (0) or a negative SRS_E_* code; the message of
 *    the last failure on the calling thread is returned by srs_last_error().
 *  - Element indices, face lists and state arrays crossing the ABI are in the
 *    reference's numbering (after the reference's BFS renumbering,
 *    cpu/InitFuncs.cpp:679-793).  A context may store elements in another
 *    order internally; that permutation never leaks out.
 *  - State arrays are species-major double[6][n_ele] in the reference species
 *    order Ca, ATP, CaM, TnC, CSQ, Fluo (cpu/parameters.h:136-141).
 *  - Parameter vectors are int32[19] and double[56] with the reference's
 *    index layout (cpu/parameters.h:18-92); SRS_I_* / SRS_R_* below are those
 *    indices.  t_step is stored in ms, as ReadParameters leaves it.
 *  - A context owns all device memory; caller buffers are copied, never kept.
 *    One context per host thread; each context has its own HIP stream.
 */
#ifndef SRS_ABI_H_
#define SRS_ABI_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- status */
#define SRS_OK 0
#define SRS_E_ARG -1      /* invalid argument / size */
#define SRS_E_IO -2       /* file missing, truncated or malformed */
#define SRS_E_HIP -3      /* HIP runtime error (message has the HIP string) */
#define SRS_E_STATE -4    /* call out of order (e.g. before finalize) */
#define SRS_E_DIVERGED -5 /* non-finite or negative field detected */
#define SRS_E_NODEV -6    /* no usable gfx950 device */

const char* srs_last_error(void);
const char* srs_version(void);

/* ------------------------------------------------- parameter vector layout */
/* cpu/parameters.h:18-36 */
#define SRS_N_INT 19
#define SRS_I_OUTPUT_GRID_FLAG 0
#define SRS_I_FLAG_FIDELITY_SIM 1
#define SRS_I_FLAG_CLOCK_SEED 2
#define SRS_I_RAND_SEED_0 3
#define SRS_I_RAND_ADD_A 4
#define SRS_I_RAND_ADD_B 5
#define SRS_I_INITIALIZE_FROM_FILE 6
#define SRS_I_CUDA_DEVICE 7
#define SRS_I_FLAG_SECONDARY_GRADIENT 8
#define SRS_I_FLAG_NO_FLUX_BOUNDARY 9
#define SRS_I_N_RYR_FIDELITY 10
#define SRS_I_FLAG_LCC 11
#define SRS_I_LCC_INDEX 12
#define SRS_I_FLAG_V_CLAMP 13
#define SRS_I_FLAG_NO_GATING 14
#define SRS_I_GATING_FFWD 15
#define SRS_I_FLAG_RYR_OUT 16
#define SRS_I_N_SIMS 17
#define SRS_I_INIT_RYR 18
/* cpu/parameters.h:37-92 */
#define SRS_N_REAL 56
#define SRS_R_T_FINAL 0
#define SRS_R_T_STEP 1
#define SRS_R_OUTPUT_STATES_INTERVAL 2
#define SRS_R_OUTPUT_GRID_INTERVAL 3
#define SRS_R_C_MIN 4
#define SRS_R_D_CA 5
#define SRS_R_D_CA_JSR 6
#define SRS_R_D_ATP 7
#define SRS_R_D_CMDN 8
#define SRS_R_D_TRPN 9
#define SRS_R_D_CSQN 10
#define SRS_R_D_DYE 11
#define SRS_R_D_EGTA 12
#define SRS_R_C_0_CA 13
#define SRS_R_C_0_CA_JSR 14
#define SRS_R_DEFAULT_VM 15
#define SRS_R_DEFAULT_CA_NSR 16
#define SRS_R_CA_0 17
#define SRS_R_B_TOT_ATP 18
#define SRS_R_K_OFF_ATP 19
#define SRS_R_K_ON_ATP 20
#define SRS_R_B_TOT_CMDN 21
#define SRS_R_K_OFF_CMDN 22
#define SRS_R_K_ON_CMDN 23
#define SRS_R_B_TOT_TRPN 24
#define SRS_R_K_OFF_TRPN 25
#define SRS_R_K_ON_TRPN 26
#define SRS_R_B_TOT_SL_HI 27
#define SRS_R_K_D_SL_HI 28
#define SRS_R_B_TOT_SL_LO 29
#define SRS_R_K_D_SL_LO 30
#define SRS_R_B_TOT_DYE 31
#define SRS_R_K_OFF_DYE 32
#define SRS_R_K_ON_DYE 33
#define SRS_R_B_T_JSR 34
#define SRS_R_K_M_JSR 35
#define SRS_R_V_REFILL 36
#define SRS_R_V_RYR 37
#define SRS_R_V_CELL 38
#define SRS_R_K_ON_CSQN 39
#define SRS_R_K_OFF_CSQN 40
#define SRS_R_RYR_ETA 41
#define SRS_R_RYR_A_STAR 42
#define SRS_R_RYR_EPS_CC 43
#define SRS_R_RYR_EPS_OO 44
#define SRS_R_RYR_K_PLUS 45
#define SRS_R_RYR_K_MINUS 46
#define SRS_R_RYR_PHI_M 47
#define SRS_R_RYR_PHI_B 48
#define SRS_R_K_D_I 49
#define SRS_R_K_D_SR 50
#define SRS_R_A_P 51
#define SRS_R_LCC_DURATION 52
#define SRS_R_T_CLAMP 53
#define SRS_R_V_CLAMP 54
#define SRS_R_INIT_RYR_HOLD 55

#define SRS_N_SPECIES 6
#define SRS_TITLE_LEN 256

/* ======================================================= host-side layer */

/* Replaces ReadParameters (cpu/InitFuncs.cpp:25-381): defaults (:28-108),
 * then every <parameter><symbol/><value/> element of the file, mapped as the
 * reference maps them (:191-349; atoi/atof semantics, t_step ns -> ms).
 * Unknown symbols are ignored.  A file with no <parameter> is SRS_E_IO, as
 * the reference throws at :164-166.  path == NULL returns the defaults. */
int srs_read_parameters(const char* path, int32_t ints[SRS_N_INT], double reals[SRS_N_REAL],
                        char title[SRS_TITLE_LEN]);

/* Mesh as loaded and renumbered by ReadGrid (cpu/InitFuncs.cpp:384-813,
 * GridStruct cpu/parameters.h:161-206, channel arrays of SimData :216-226).
 * Face types are already mapped to simulation codes 0 default, 1 bound,
 * 2 no-flux (cpu/InitFuncs.cpp:494-514). */
typedef struct srs_grid {
  int32_t n_ele, n_face, n_nodes;
  double *V0, *tt_area, *sr_area;       /* [n_ele] */
  int32_t *domain, *tropc;              /* [n_ele] */
  int32_t *ele_nodes;                   /* [4*n_ele], 0-based */
  double *node_xyz;                     /* [3*n_nodes] */
  int32_t *ele1, *ele2, *face_type;     /* [n_face] */
  double *face_c;                       /* [n_face] */
  int32_t n_ryr, n_lcc;
  int32_t *ryr_ele, *ryr_jsr, *ryr_nb;  /* [n_ryr], [n_ryr], [4*n_ryr] */
  int32_t *lcc_ele;                     /* [n_lcc] */
  int32_t *perm;                        /* [n_ele] new -> file index (the BFS order) */
} srs_grid;

/* Replaces ReadGrid.  Reads base_{tet,face}_properties.txt, base.node,
 * base.ele, base_{ryr,lcc}.txt, maps face types with the no-flux flag, and
 * renumbers by the reference's BFS from element 0 (in O(N) rather than the
 * reference's O(N_e*N_f)).  Fails with SRS_E_IO on a missing or truncated
 * file (the reference prints and continues with uninitialised arrays) and
 * with SRS_E_ARG on a disconnected face graph (where the reference reads out
 * of bounds).  Duplicate channel elements are reported through *warnings. */
int srs_grid_read(const char* base, int no_flux_boundary, srs_grid** out, int32_t* warnings);
void srs_grid_free(srs_grid* g);

/* The reference BFS order alone (cpu/InitFuncs.cpp:679-731): visited[i] is the
 * original index of new element i.  Returns SRS_E_ARG if not connected. */
int srs_bfs_order(int32_t n_ele, int32_t n_face, const int32_t* ele1, const int32_t* ele2,
                  int32_t* visited);

/* Replaces InitializeExplicitSparseMatrixFast (cpu/InitFuncs.cpp:1058-1112,
 * 1273-1322): the explicit FV operator packed as the reference packs it,
 * aij[5*i] diagonal then up to four off-diagonals with ascending columns,
 * nij[4*i] their columns, self and 0.0 where absent. */
int srs_assemble_operator(int32_t n_ele, int32_t n_face, const int32_t* ele1, const int32_t* ele2,
                          const int32_t* face_type, const double* face_c, const double* V0,
                          double* aij, int32_t* nij);

/* InitializeDomainStruct boundary rates (cpu/InitFuncs.cpp:1012-1026), br out
 * double[6][n_ele]; and ResetDefaultGridStates (cpu/InitFuncs.cpp:890-954),
 * X out double[6][n_ele], boundaries out double[6] (may be NULL). */
int srs_boundary_rates(int32_t n_ele, int32_t n_face, const int32_t* ele1, const int32_t* face_type,
                       const double* face_c, const double* V0, const double reals[SRS_N_REAL],
                       const double boundaries[SRS_N_SPECIES], double* br);
int srs_equilibrium_state(int32_t n_ele, const int32_t* domain, const int32_t* tropc,
                          const double reals[SRS_N_REAL], double* X, double boundaries[SRS_N_SPECIES]);

/* Synthetic geometry.  Writes base.{node,ele,face}, base_{tet,face}_properties.txt
 * and base_{ryr,lcc}.txt in the formats of mesh/main.cpp:499-567, 1305-1364 for
 * a rectilinear box split into 6 Kuhn tets per cell, with optional jSR slab,
 * sarcolemmal face and T-tubule cylinder.  Face coefficients follow
 * mesh/main.cpp:334-482.  No reference counterpart (the reference meshes with
 * tetgen, which is not rebuilt). */
typedef struct srs_box_spec {
  int32_t nx, ny, nz;
  const double *xs, *ys, *zs;   /* plane coordinates, nx+1 / ny+1 / nz+1 values, um */
  double jsr_lo[3], jsr_hi[3];  /* jSR slab box (elements whose centroid lies inside), hi<lo: none */
  int32_t sl_face_z0;           /* 1: the z = zs[0] face is sarcolemma (face type 2) */
  double tt_axis_x, tt_axis_z, tt_radius; /* T-tubule cylinder along y; radius <= 0: none */
  double trpn_radius;           /* TnC in cytosol elements farther than this from the TT axis */
  int32_t ryr_rows, ryr_cols;   /* RyR lattice (rows along y, cols along x) */
  double ryr_spacing, ryr_center[2], ryr_z_cyt, ryr_z_jsr;
  const uint8_t* ryr_mask;      /* rows*cols occupancy, row-major; NULL = all */
  int32_t n_lcc;
  const double* lcc_xyz;        /* [3*n_lcc] LCC positions (nearest cytosol element) */
} srs_box_spec;
int srs_box_mesh_write(const srs_box_spec* spec, const char* base);
/* The same mesh built in memory and returned as srs_grid_read would return it
 * after reading the files (face types mapped, BFS-renumbered), for meshes too
 * large to round-trip through text files. */
int srs_box_grid(const srs_box_spec* spec, int no_flux_boundary, srs_grid** out);

/* States record of WriteToStateFiles (cpu/OutputFuncs.cpp:223-350): t, I_RyR,
 * N_RyR open, N_LCC open, I_LCC, mean Ca_jSR at RyRs, mean Ca_cyt at RyRs,
 * [gpu_layout: Ca_cyt at each RyR, gpu/OutputFuncs.cpp:277-284],
 * [ryr_out: RyR states].  ca_ryr_cyt/ca_ryr_jsr/ca_lcc are Ca at the channel
 * elements.  Returns the number of floats written (> 255 is SRS_E_ARG, the
 * reference's buffer limit, cpu/OutputFuncs.cpp:441-444). */
int srs_states_record(double t, int32_t n_ryr, const int32_t* ryr_state, const double* ca_ryr_cyt,
                      const double* ca_ryr_jsr, int32_t n_lcc, const int32_t* lcc_state,
                      const int32_t* lccv_state, const double* ca_lcc, double vm,
                      const double reals[SRS_N_REAL], int32_t ryr_out, int32_t gpu_layout, float* rec);

/* Replaces main + RunSimulation (cpu/main.cpp:18-274, cpu/RunSimulation.cpp:402-589):
 * the cru3d driver.  Reads the parameter file and mesh, runs sims_per_proc
 * realizations from start_index on the GPU, writes states_<rank>_<run>.out
 * and the VTK grids into out_dir (created if needed).  rank plays the role of
 * the MPI rank in seeding and file names.  The states file keeps the CPU
 * record layout unless gpu_layout is set. */
int srs_run_cru3d(const char* param_path, const char* mesh_base, const char* out_dir, int32_t rank,
                  int32_t start_index, int32_t gpu_layout, double* wall_seconds);

/* Element order the cru3d driver gives its context: elements sorted along a
 * Morton curve through their centroids (1024 cells per axis over the bounding
 * box), perm[k] = element stored at position k.  No reference counterpart
 * (the reference keeps its BFS order on the device, cpu/InitFuncs.cpp:679-731);
 * the order moves data, never changes results. */
int srs_morton_order(int32_t n_ele, const int32_t* ele_nodes, const double* node_xyz, int32_t* perm);

/* Model extensions the reference cannot express (no counterpart there).
 * All are off by default, and a run with them off is the reference's model
 * bit for bit.
 *   jsr_diffusion  1: Ca diffuses with D_CAJSR (reals[SRS_R_D_CA_JSR]) in jSR
 *                  elements.  The reference parses D_CAJSR but assembles the
 *                  operator with D = 1 (cpu/InitFuncs.cpp:1081) and scales Ca
 *                  by D_CA in both domains (cpu/RunSimulation.cpp:62); jSR and
 *                  cytosol are joined by no-flux faces only, so a per-domain
 *                  coefficient is a per-row scale.
 *   b_sr_sites     SR Ca-binding sites in the cytosol (uM), a rapid buffer of
 *   k_d_sr_sites   dissociation constant k_d_sr_sites (uM) in the form of the
 *                  sarcolemmal rapid buffer (cpu/RunSimulation.cpp:15-22):
 *                  beta += b K / (K + c)^2 in every cytosol element.  0: none.
 * srs_read_extensions reads the XML symbols flag_jsr_diffusion ("true"),
 * B_TOT_SRSITES and K_D_SRSITES from a parameter file (the reference ignores
 * unknown symbols; defaults 0, 0 and 0.6 uM). */
typedef struct srs_extensions {
  int32_t jsr_diffusion;
  double b_sr_sites, k_d_sr_sites;
} srs_extensions;
int srs_read_extensions(const char* path, srs_extensions* out);

/* ======================================================= device layer */

typedef struct srs_ctx srs_ctx;

/* Element order inside a context (never visible through the ABI). */
#define SRS_ORDER_AS_GIVEN 0
#define SRS_ORDER_RCM 1
#define SRS_ORDER_CUSTOM 2   /* srs_mesh_desc.perm gives the order */

/* Replaces InitializeDeviceConstants / cudaSetDevice (gpu/RunSimulation.cu:503-506, 869-872). */
int srs_ctx_create(int device, srs_ctx** out);
/* Kernel and layout tunables (no reference counterpart; the reference fixes
 * 256-thread blocks, gpu/RunSimulation.cu:784).  They change speed only,
 * never results.  Layout keys must be set before the first mesh is added,
 * the tile before srs_ctx_finalize; the rest at any time.  The defaults are
 * the fastest setting measured on the 3M-element dyad mesh in either element
 * order, short of the register budget (which helps in the block order only). */
#define SRS_TUNE_TILE 1          /* elements per sweep workgroup: 128, 256, 512; by default srs_ctx_finalize
                                    takes 512 when a realization's mesh has 2^20 elements or more, else 256 */
#define SRS_TUNE_NT_GEOMETRY 2   /* 1 (default): non-temporal loads of operator rows and flags */
#define SRS_TUNE_NT_STATE 3      /* 1: non-temporal stores of the new state */
#define SRS_TUNE_PLANE_PAD 4     /* extra elements (multiple of 64) between SoA planes */
#define SRS_TUNE_COMPACT 5       /* 1 (default): store operator rows as class + int16 ulp deltas and
                                    neighbours as int16 offsets where that reproduces them exactly */
#define SRS_TUNE_OCCUPANCY 6     /* sweep register budget: 0 compiler's choice; 6 or 8: at most 80 or 64
                                    VGPRs, so that six or eight wavefronts share each SIMD */
#define SRS_TUNE_GROUP 7         /* 1 (default), 2 or 4: realizations of one mesh whose states are interleaved
                                    per element so that each sweep tile loads the operator rows once for all
                                    of them (set before srs_ctx_finalize) */
#define SRS_TUNE_LDS 8           /* 1 (default): stage each sweep tile's diffusing species in LDS and read
                                    neighbours inside the tile from there (groups of 1) */
#define SRS_TUNE_DROP_FINISHED 9 /* 1 (default): srs_explicit_step leaves the tiles of groups whose
                                    realizations have all finished out of the sweep; 0: sweep them all */
#define SRS_TUNE_GATED_SWEEP 10  /* 1 (default): a single-realization context with the LDS sweep runs each
                                    step's gating (and, with SRS_STEP_PROTOCOL, the loop body) inside the
                                    next step's sweep launch (register budget 0 or 6; 0 with the loop
                                    body); 0: a gating launch after every sweep */
#define SRS_TUNE_SUBSET_MB 12    /* 0 (default): every step sweeps all realizations; > 0: srs_explicit_step
                                    steps runs of realization groups whose state takes at most this many MB
                                    through a whole block of steps (128 with SRS_STEP_GRAPH) each, one after
                                    another, so that a run's state stays in the last-level cache */
#define SRS_TUNE_ALT_ORDER 13    /* 1: the single-realization step (SRS_TUNE_GATED_SWEEP) sweeps its tiles
                                    in reverse order on alternate steps, reading first what the step before
                                    wrote last; 0 (default): one order */
#define SRS_TUNE_ADJ_WAVES 14    /* adjoint reverse step's register budget: 0 (default) the compiler's, 4: at
                                    most 128 VGPRs so that four wavefronts share each SIMD */
#define SRS_TUNE_ADJ_W 15        /* adjoint reverse step's w = (lambda_Ca / beta, lambda_ATP, lambda_CaM,
                                    lambda_Fluo) / V0: 0 (default) written by each step for the next on
                                    launches of fewer than 4096 workgroups (tiles x members), else formed
                                    in the kernel from lambda; 1: always written (two more state-sized
                                    buffers per member); 2: always formed.  Same results */
#define SRS_TUNE_ADJ_LEVELS 11   /* adjoint checkpoint levels (srs_linescan_loss_grad*): 0 (default) the
                                    fewest whose memory fits the sub-batch, 1..6 forced; each level below the
                                    first is one more forward pass and needs about (L+1) n^(1/(L+1)) states */
int srs_ctx_tune(srs_ctx* ctx, int32_t key, int32_t value);

/* The model extensions of one realization (srs_extensions; no reference
 * counterpart), at any time after srs_ctx_add_realization.  The adjoint
 * (srs_linescan_loss_grad*) covers the reference model only and refuses a
 * realization with an extension on. */
int srs_ctx_extensions(srs_ctx* ctx, int32_t real, const srs_extensions* ext);

/* Replaces FreeDomainStruct / FreeGlobalStruct (gpu/RunSimulation.cu:874-901). */
int srs_ctx_destroy(srs_ctx* ctx);
/* States records in the GPU program's layout: WriteToStateFiles of gpu/
 * (gpu/OutputFuncs.cpp:277-284) appends Ca_cyt at each RyR after the seven
 * fields, before the optional RyR states.  0 (the default) keeps the CPU
 * layout (cpu/OutputFuncs.cpp:244-350).  Before srs_ctx_finalize. */
int srs_ctx_record_layout(srs_ctx* ctx, int32_t gpu_layout);

/* Geometry of one mesh, shared by all realizations on it.  Replaces the
 * DeviceDomainStruct upload of InitializeDomainStruct (gpu/InitFuncs.cu:1001-1093,
 * gpu/DeviceStructs.h:19-38).  aij/nij as from srs_assemble_operator; br may
 * be NULL (all zero, as with flag_noflux).  order: SRS_ORDER_*. */
typedef struct srs_mesh_desc {
  int32_t n_ele;
  const double* aij;       /* [5*n_ele] */
  const int32_t* nij;      /* [4*n_ele] */
  const double* br;        /* [6*n_ele] or NULL */
  const int32_t* domain;   /* [n_ele] */
  const int32_t* tropc;    /* [n_ele] */
  const double* V0;        /* [n_ele] */
  const double* tt_area;   /* [n_ele] */
  int32_t order;
  const int32_t* perm;     /* SRS_ORDER_CUSTOM: perm[k] = element stored at position k */
} srs_mesh_desc;
int srs_ctx_add_mesh(srs_ctx* ctx, const srs_mesh_desc* mesh, int32_t* mesh_id);

/* One realization: a channel set on a mesh plus its own parameter vector.
 * Replaces InitializeDeviceGlobalStruct (gpu/InitFuncs.cu:1736-1782,
 * gpu/DeviceStructs.h:40-58).  Channel element indices are mesh indices. */
typedef struct srs_chan_desc {
  int32_t n_ryr;
  const int32_t* ryr_ele;  /* [n_ryr] cleft element */
  const int32_t* ryr_jsr;  /* [n_ryr] jSR element */
  const int32_t* ryr_nb;   /* [4*n_ryr] lattice neighbours, -1 none */
  int32_t n_lcc;
  const int32_t* lcc_ele;  /* [n_lcc] */
} srs_chan_desc;
/* seed_key: the Philox key for this realization's gating draws; stream: the
 * realization's counter word (distinct streams for distinct realizations). */
int srs_ctx_add_realization(srs_ctx* ctx, int32_t mesh_id, const srs_chan_desc* ch,
                            const int32_t ints[SRS_N_INT], const double reals[SRS_N_REAL],
                            uint64_t seed_key, uint64_t stream, int32_t* real_id);

/* Allocates and uploads everything; after this the context is runnable.
 * record_capacity: states records kept on the device per realization
 * between srs_records_read calls (0: no device-side records). */
int srs_ctx_finalize(srs_ctx* ctx, int32_t record_capacity);

/* Field and channel state transfer (the cudaMemcpy calls of
 * gpu/RunSimulation.cu:712-723, 755-763).  X is double[6][n_ele] of the
 * realization's mesh, reference numbering. */
int srs_state_set(srs_ctx* ctx, int32_t real, const double* X);
int srs_state_get(srs_ctx* ctx, int32_t real, double* X);
int srs_channels_set(srs_ctx* ctx, int32_t real, const int32_t* ryr_state, const int32_t* lcc_state,
                     const int32_t* lccv_state, const double* R);
int srs_channels_get(srs_ctx* ctx, int32_t real, int32_t* ryr_state, int32_t* lcc_state,
                     int32_t* lccv_state, double* R);
int srs_globals_set(srs_ctx* ctx, int32_t real, double vm, double ca_nsr);
int srs_globals_get(srs_ctx* ctx, int32_t real, double* vm, double* ca_nsr);

/* Loop state of a realization as RunSimulation keeps it (cpu/RunSimulation.cpp:410-460):
 * simulated time, output bookkeeping and protocol flags. */
typedef struct srs_loop_state {
  double t, dt_last;
  int64_t step;              /* Philox counter word: number of steps taken */
  int32_t done;              /* 1 once the reference's while-condition failed */
  int32_t lcc_open;          /* bLCC_Open */
  double t_lcc_close;        /* T_LCC_Close */
  int32_t n_ryr_open_last;   /* N_RyR_Open as of the last states output */
  double ca_ss_max_last;     /* ca_ss_max as of the last states output */
  int32_t ffwd;              /* 1 while the flag_ffwd gating-only phase runs
                                (cpu/RunSimulation.cpp:469-519): steps of this
                                realization gate channels on the unchanged fields */
} srs_loop_state;
int srs_loop_set(srs_ctx* ctx, int32_t real, const srs_loop_state* ls);
int srs_loop_get(srs_ctx* ctx, int32_t real, srs_loop_state* ls);

/* Checkpoint and resume of one realization (the reference restarts from a
 * states file only through the unreachable INITIALIZE_FROM_FILE path,
 * cpu/InitFuncs.cpp:815-888; these calls replace it): fields, channel states,
 * residuals and loop state in one versioned buffer.  Gating draws are counted
 * by step, so a realization loaded from a checkpoint continues bit for bit as
 * the saved run would have; loading needs the same mesh size, channel counts,
 * seed and stream.  Records not yet read with srs_records_read are not saved
 * (the record buffer restarts empty). */
int srs_checkpoint_size(srs_ctx* ctx, int32_t real, size_t* bytes);
int srs_checkpoint_save(srs_ctx* ctx, int32_t real, void* buf, size_t bytes);
int srs_checkpoint_load(srs_ctx* ctx, int32_t real, const void* buf, size_t bytes);

/* Step flags */
#define SRS_STEP_PLAIN 0u      /* exactly ExplicitStep: sweep, then gating unless flag_nogating */
#define SRS_STEP_PROTOCOL 1u   /* also run the RunSimulation loop body on the device: time
                                  accumulation, while-condition, LCC trigger, voltage clamp,
                                  RyR hold and states-record output (cpu/RunSimulation.cpp:522-573) */
#define SRS_STEP_GRAPH 2u      /* replay the step block as a captured hipGraph: 128-step blocks, the
                                  rest as one block when it is 16 steps or more (graphs are cached per
                                  size and mode) */

/* Replaces ExplicitStep (cpu/RunSimulation.cpp:591-654; gpu/RunSimulation.cu:781-852),
 * called n_steps times for every realization that is not done, with each
 * realization's own dt = reals[SRS_R_T_STEP]. */
int srs_explicit_step(srs_ctx* ctx, int64_t n_steps, uint32_t flags);

/* Captures the graphs srs_explicit_step(ctx, n_steps, flags | SRS_STEP_GRAPH)
 * will replay, without stepping, so that a timed or latency-sensitive run
 * does not pay for the capture.  No reference counterpart. */
int srs_explicit_step_prepare(srs_ctx* ctx, int64_t n_steps, uint32_t flags);

/* Replaces UpdateChannelGating (cpu/RunSimulation.cpp:656-674; gpu/RunSimulation.cu:854-866)
 * alone, for every realization: the gating-only loop of flag_ffwd. */
int srs_channel_gating(srs_ctx* ctx);

/* Deterministic replay: n_steps ExplicitStep calls with gating off and the RyR
 * states of realization `real` overwritten before step k by
 * schedule[k*n_ryr + i] (a recorded open/closed schedule). */
int srs_explicit_step_replay(srs_ctx* ctx, int32_t real, int64_t n_steps, const int32_t* schedule);
/* srs_explicit_step_replay without any device allocation, for callers that
 * cannot allocate (code under graph capture, serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The schedule is staged
 * in a fixed device buffer of 2,097,152 RyR states (8 MiB) that each GPU holds
 * from its first srs_ctx_create until the process ends; n_steps x n_ryr above
 * that is SRS_E_ARG.  Every call on one GPU shares that buffer (it keeps the
 * last call's schedule): calls that overlap on one GPU, from several threads
 * or contexts, race.  Results equal srs_explicit_step_replay's bit for bit. */
int srs_explicit_step_replay_static(srs_ctx* ctx, int32_t real, int64_t n_steps, const int32_t* schedule);

/* Ca at the channel elements (for the states record on the host). */
int srs_channel_ca(srs_ctx* ctx, int32_t real, double* ca_ryr_cyt, double* ca_ryr_jsr, double* ca_lcc);

/* Device-side states records written by SRS_STEP_PROTOCOL steps since the last
 * call; copies up to max_records records of rec_len floats. */
int srs_records_read(srs_ctx* ctx, int32_t real, float* out, int32_t max_records, int32_t* n_records,
                     int32_t* rec_len);

/* Waits for all work of the context. */
int srs_ctx_sync(srs_ctx* ctx);

/* Sizes of a realization. */
int srs_real_info(srs_ctx* ctx, int32_t real, int32_t* n_ele, int32_t* n_ryr, int32_t* n_lcc);

/* Timing of the last srs_explicit_step call, measured with HIP events on the
 * context's stream: total milliseconds and the mean duration of one sweep
 * kernel (ms, from a separately timed single-step launch when requested). */
int srs_last_timing(srs_ctx* ctx, double* total_ms, double* sweep_ms);
int srs_time_sweep(srs_ctx* ctx, int32_t reps, double* sweep_ms, double* gating_ms);

/* ================================================== linescan (dF/F0 operator) */

/* Parameters of the reference's linescan program (linescan/parameters.h:17-43;
 * defaults and XML symbols linescan/main.cpp:353-508). */
typedef struct srs_scan_desc {
  int32_t n_scans;        /* scan_n */
  int32_t n_scans_skip;   /* scan_n_per_line (flag_x_res_ca also lands here, as in the reference) */
  int32_t axis;           /* scan_axis "X" 0, "Y" 1, "Z" 2 */
  double res_x_f;         /* scan_x_res_f: fluorescence pixel [um] */
  double res_x_ca;        /* Ca pixel [um] (no XML symbol reaches it; default 0.01) */
  double res_t;
  double x_min, x_max;    /* scan_x_min, scan_x_max */
  double offset1, offset2;/* scan_offset_1, scan_offset_2 */
  double domain[3];       /* scan_domain_x/y/z */
  double pad_min, pad_max, pad_res;  /* scan_padding_* */
  double var[3];          /* scan_var_x/y/z: PSF variances */
  double f0, c0;          /* scan_f0, scan_c0: values of the padding cells */
} srs_scan_desc;

/* Replaces the linescan ReadParameters (linescan/main.cpp:353-508); path NULL
 * gives the defaults.  The root element must be <analysis>. */
int srs_scan_read_parameters(const char* path, srs_scan_desc* out);

/* Prepares a scan line over mesh `mesh_id` of the context (after
 * srs_ctx_finalize): centers[3*i + k] is coordinate k of element i's centroid
 * (reference numbering).  Replaces the geometry half of the linescan main
 * (linescan/main.cpp:85-141: cell centres, padding grid, scan points).
 * n_f / n_ca: number of fluorescence and Ca points. */
int srs_ctx_add_scan(srs_ctx* ctx, int32_t mesh_id, const double* centers, const srs_scan_desc* scan,
                     int32_t* scan_id, int32_t* n_f, int32_t* n_ca);

/* One scan line of realization `real`'s current state, computed on the device
 * (linescan/main.cpp:195-249): ca[n_ca] = Ca of the nearest cell,
 * caf[n_f] = Ca-bound Fluo of the nearest cell, cafb[n_f] = the PSF-blurred
 * Ca-bound Fluo (sum of V0 * CaF * Gaussian PSF over cells within 1 um,
 * padding cells included).  dF/F0 = cafb / f0 - 1. */
int srs_scan_sample(srs_ctx* ctx, int32_t real, int32_t scan_id, double* ca, double* caf, double* cafb);

/* Replaces the linescan program's WriteOutputFile (linescan/main.cpp:316-336),
 * which writes scan_output_Ca.txt, scan_output_CaF.txt and scan_output.txt
 * (:482-485, :252-254): rows[j * n_cols + i] as printf "%g " and a newline
 * after each of the n_rows rows; one row per scan (the ca, caf or cafb of an
 * srs_scan_sample call), one column per point.  SRS_E_IO if the file cannot
 * be written (the reference prints a warning). */
int srs_scan_write(const char* path, const double* rows, int32_t n_rows, int32_t n_cols);

/* ====================================================== adjoint (fitting) */
/* These calls have no counterpart in the reference: they serve parameter
 * fitting against line scans with the channel trajectory held fixed. */

/* Number of parameters with gradients, and their SRS_R_* indices in order:
 * D_CA, D_DYE, K_ON/K_OFF of CaM, TnC, ATP and Fluo, V_RYR, A_P (SERCA pump
 * density, the Vmax scale), K_D_SL_HI. */
#define SRS_N_THETA 13
int srs_adjoint_theta(int32_t idx[SRS_N_THETA]);

/* Replaces a realization's parameter vector after srs_ctx_finalize. */
int srs_params_set(srs_ctx* ctx, int32_t real, const double reals[SRS_N_REAL]);

/* n_steps plain steps with gating of a one-realization context, recording
 * schedule[k*n_ryr + i] = state of RyR i during step k (row 0 = the states
 * at the call). */
int srs_record_schedule(srs_ctx* ctx, int32_t real, int64_t n_steps, int32_t* schedule);

/* The same for every realization of a context, stepped together: each
 * realization's n_steps x n_ryr rows, concatenated in realization order. */
int srs_record_schedules(srs_ctx* ctx, int64_t n_steps, int32_t* schedules);

/* From the realization's current state, n_steps steps with the RyR states of
 * each step taken from `schedule` and gating off (LCC states held), sampling
 * scan line `scan_id` after sample_steps[k] steps (increasing, 1..n_steps).
 * loss = sum over samples and fluorescence points of (cafb/f0 - 1 - target)^2.
 * If grad is not NULL, the adjoint computes d loss / d theta (srs_adjoint_theta
 * order) with checkpoints every `ckpt` steps (1..65536) and recomputation.
 * Needs realization groups of size 1; other realizations of the context are
 * held meanwhile.  The realization's state, channel states and loop state are
 * restored. */
int srs_linescan_loss_grad(srs_ctx* ctx, int32_t real, int32_t scan_id, int64_t n_steps, const int32_t* schedule,
                           int32_t n_samples, const int64_t* sample_steps, const double* target, int32_t ckpt,
                           double* loss, double* grad);

/* srs_linescan_loss_grad for the realizations real_ids[0..n_real), each with
 * its own schedule, target and parameters, stepped together in sub-batches:
 *   schedules  the members' n_steps x n_ryr(member) schedules, concatenated in
 *              real_ids order;
 *   targets    [n_real][n_samples][n_f];
 *   ckpt       length of a recomputed segment (every state of which is
 *              kept), 0 = chosen with the checkpoint levels: n_steps^(1/(L+1))
 *              for L levels (ceil(sqrt(n_steps)) for one level);
 *   max_batch  most members stepped at once, 0 = as many as device memory
 *              holds.  A member takes about 48 B x elements x (states kept + 2):
 *              with one level n_steps/ckpt + ckpt, with L levels about
 *              (L+1) n_steps^(1/(L+1)).  Unless SRS_TUNE_ADJ_LEVELS forces it,
 *              the call takes the fewest levels with which max_batch members
 *              (one, if max_batch is 0) fit; the plan changes memory and time,
 *              never a bit of the result;
 *   losses     [n_real];  grads [n_real][SRS_N_THETA] or NULL;
 *   scans      NULL or [n_real][n_samples][n_f]: the simulated dF/F0.
 * Results for a member equal srs_linescan_loss_grad on that member alone, bit
 * for bit.  All members must be on the scan's mesh. */
int srs_linescan_loss_grad_batch(srs_ctx* ctx, int32_t n_real, const int32_t* real_ids, int32_t scan_id,
                                 int64_t n_steps, const int32_t* schedules, int32_t n_samples,
                                 const int64_t* sample_steps, const double* targets, int32_t ckpt, int32_t max_batch,
                                 double* losses, double* grads, double* scans);

/* The checkpoint plan of the context's last srs_linescan_loss_grad* call:
 * levels, the strides of its levels (strides[levels-1] is the recomputed
 * segment's length; 6 entries, unused ones 0) and the members stepped per
 * sub-batch.  No reference counterpart. */
int srs_adjoint_last_plan(srs_ctx* ctx, int32_t* levels, int64_t* strides, int32_t* sub_batch);

#ifdef __cplusplus
}
#endif
#endif /* SRS_ABI_H_ */
