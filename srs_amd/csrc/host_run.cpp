// host_run.cpp -- the RunSimulation driver on top of the device layer, and
// the states/VTK writers.
//
//   srs_states_record  cpu/OutputFuncs.cpp:223-483 (WriteToStateFiles record)
//   srs_run_cru3d      cpu/main.cpp:18-274 + cpu/RunSimulation.cpp:402-589
//
// The ensemble loop of main (cpu/main.cpp:209-258), which the reference runs
// one realization after another per MPI rank, becomes one batched context:
// every realization of the rank steps in the same kernel launches.  The
// RunSimulation loop body (while-condition, protocols, states output) runs on
// the device after each step; the host only stops the device at grid-output
// times and to drain the states records.
#include <sys/stat.h>
#include <sys/types.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>
#include <utility>
#include <vector>

#include "kinetics.h"
#include "philox.h"
#include "srs_internal.h"

extern "C" int srs_states_record(double t, int32_t n_ryr, const int32_t* ryr_state, const double* ca_ryr_cyt,
                                 const double* ca_ryr_jsr, int32_t n_lcc, const int32_t* lcc_state,
                                 const int32_t* lccv_state, const double* ca_lcc, double vm,
                                 const double reals[SRS_N_REAL], int32_t ryr_out, int32_t gpu_layout, float* rec) {
  srs::clear_error();
  const int len = 7 + (gpu_layout ? n_ryr : 0) + (ryr_out ? n_ryr : 0);
  if (len > 255) return srs::fail(SRS_E_ARG, "states record of %d floats exceeds the 255-float record limit", len);
  if (!rec || !reals) return srs::fail(SRS_E_ARG, "null argument");
  return srs_record_fill(t, n_ryr, ryr_state, ca_ryr_cyt, ca_ryr_jsr, n_lcc, lcc_state, lccv_state, ca_lcc, vm, reals,
                         ryr_out, gpu_layout, rec);
}

namespace {

struct Guard {
  srs_ctx* ctx = nullptr;
  srs_grid* grid = nullptr;
  ~Guard() {
    if (ctx) srs_ctx_destroy(ctx);
    if (grid) srs_grid_free(grid);
  }
};

// cpu/OutputFuncs.cpp:42-220, streamed instead of a fixed 10 MB buffer.
int write_vtk(const std::string& path, int run, int rank, int n, double t, const srs_grid* g, const double* X,
              const int32_t* ryr_state) {
  FILE* f = fopen(path.c_str(), "w");
  if (!f) return srs::fail(SRS_E_IO, "cannot open grid file %s", path.c_str());
  const int N = g->n_ele;
  fprintf(f, "# vtk DataFile Version 3.0\n");
  fprintf(f, "Grid data for run %d-%d-%d at t = %f\n", run, rank, n, t);
  fprintf(f, "ASCII\n\n");
  fprintf(f, "DATASET UNSTRUCTURED_GRID\n");
  fprintf(f, "POINTS %d float\n", g->n_nodes);
  for (int j = 0; j < g->n_nodes; ++j) fprintf(f, "%f %f %f\n", g->node_xyz[3 * j], g->node_xyz[3 * j + 1], g->node_xyz[3 * j + 2]);
  fprintf(f, "\nCELLS %d %d\n", N, 5 * N);
  for (int j = 0; j < N; ++j)
    fprintf(f, "4 %d %d %d %d\n", g->ele_nodes[4 * j], g->ele_nodes[4 * j + 1], g->ele_nodes[4 * j + 2], g->ele_nodes[4 * j + 3]);
  fprintf(f, "\nCELL_TYPES %d\n", N);
  for (int j = 0; j < N; ++j) fprintf(f, "10\n");
  fprintf(f, "\nCELL_DATA %d\n", N);
  fprintf(f, "SCALARS C_CA float\nLOOKUP_TABLE default\n");
  for (int j = 0; j < N; ++j) fprintf(f, "%f\n", X[srs::CA * (size_t)N + j]);
  fprintf(f, "SCALARS C_DYE float\nLOOKUP_TABLE default\n");
  for (int j = 0; j < N; ++j) fprintf(f, "%f\n", X[srs::DYE * (size_t)N + j]);
  std::vector<int> b(N, 0);
  for (int i = 0; i < g->n_ryr; ++i) b[g->ryr_ele[i]] = ryr_state[i] ? 2 : 1;
  fprintf(f, "SCALARS B_RYR int\nLOOKUP_TABLE default\n");
  for (int j = 0; j < N; ++j) fprintf(f, "%d\n", b[j]);
  fprintf(f, "SCALARS B_JSR int\nLOOKUP_TABLE default\n");
  for (int j = 0; j < N; ++j) fprintf(f, "%d\n", g->domain[j] == 1 ? 1 : 0);
  std::fill(b.begin(), b.end(), 0);
  for (int i = 0; i < g->n_lcc; ++i) b[g->lcc_ele[i]] = 1;
  fprintf(f, "SCALARS B_LCC int\nLOOKUP_TABLE default\n");
  for (int j = 0; j < N; ++j) fprintf(f, "%d\n", b[j]);
  const bool ok = ferror(f) == 0;
  fclose(f);
  return ok ? SRS_OK : srs::fail(SRS_E_IO, "write error on %s", path.c_str());
}

int append_records(const std::string& path, const float* rec, size_t n_floats) {
  FILE* f = fopen(path.c_str(), "ab");
  if (!f) return srs::fail(SRS_E_IO, "cannot open states file %s", path.c_str());
  const size_t w = fwrite(rec, sizeof(float), n_floats, f);
  fclose(f);
  return w == n_floats ? SRS_OK : srs::fail(SRS_E_IO, "short write on %s", path.c_str());
}

}  // namespace

extern "C" int srs_morton_order(int32_t n_ele, const int32_t* ele_nodes, const double* node_xyz, int32_t* perm) {
  // Device element order for the drop-in: elements sorted along a Morton
  // (Z-order) curve through their centroids on a 1024^3 grid over the mesh's
  // bounding box, so that a sweep tile covers a compact patch of the mesh and
  // its neighbour gathers stay in cache.  The order changes where elements are
  // stored, never the arithmetic.  perm[k] = element stored at k.
  srs::clear_error();
  if (n_ele < 0 || (n_ele > 0 && (!ele_nodes || !node_xyz || !perm))) return srs::fail(SRS_E_ARG, "null argument");
  const int n = n_ele;
  std::vector<double> cen(3 * (size_t)n);
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int e = 0; e < n; ++e)
    for (int a = 0; a < 3; ++a) {
      double c = 0.0;
      for (int q = 0; q < 4; ++q) c += node_xyz[3 * (size_t)ele_nodes[4 * (size_t)e + q] + a];
      c *= 0.25;
      cen[3 * (size_t)e + a] = c;
      lo[a] = std::min(lo[a], c);
      hi[a] = std::max(hi[a], c);
    }
  auto spread = [](uint64_t v) {  // bit b of v to bit 3b
    uint64_t r = 0;
    for (int b = 0; b < 10; ++b) r |= ((v >> b) & 1u) << (3 * b);
    return r;
  };
  std::vector<std::pair<uint64_t, int32_t>> key(n);
  for (int e = 0; e < n; ++e) {
    uint64_t k = 0;
    for (int a = 0; a < 3; ++a) {
      const double w = hi[a] > lo[a] ? (cen[3 * (size_t)e + a] - lo[a]) / (hi[a] - lo[a]) : 0.0;
      k |= spread((uint64_t)std::lround(w * 1023.0)) << a;
    }
    key[e] = {k, e};
  }
  std::sort(key.begin(), key.end());
  for (int k = 0; k < n; ++k) perm[k] = key[k].second;
  return SRS_OK;
}

extern "C" int srs_run_cru3d(const char* param_path, const char* mesh_base, const char* out_dir, int32_t rank,
                             int32_t start_index, int32_t gpu_layout, double* wall_seconds) {
  srs::clear_error();
  if (!param_path || !mesh_base) return srs::fail(SRS_E_ARG, "parameter file and mesh base are required");
  int32_t I[SRS_N_INT];
  double P[SRS_N_REAL];
  char title[SRS_TITLE_LEN];
  int rc = srs_read_parameters(param_path, I, P, title);
  if (rc) return rc;
  srs_extensions ext;
  if ((rc = srs_read_extensions(param_path, &ext))) return rc;
  if (I[SRS_I_INITIALIZE_FROM_FILE]) return srs::fail(SRS_E_ARG, "initial states from file are not supported");

  // Output directory, cpu/main.cpp:30-133
  std::string dir;
  if (out_dir) {
    dir = out_dir;
  } else {
    time_t timer = time(nullptr);
    struct tm y2k = {};
    y2k.tm_year = 100; y2k.tm_mon = 0; y2k.tm_mday = 1;
    const int seconds = (int)difftime(timer, mktime(&y2k));
    srand(seconds);
    const int id = rand() % 1000;
    char buf[600];
    snprintf(buf, sizeof(buf), "output/%s_%d_%d", title, seconds, id);
    dir = buf;
    mkdir("output", 0700);
  }
  mkdir(dir.c_str(), 0700);
  if (!out_dir) {  // copy of the parameter file, cpu/main.cpp:144-169
    const char* slash = strrchr(param_path, '/');
    std::string dst = dir + "/" + (slash ? slash + 1 : param_path);
    FILE* a = fopen(param_path, "rb");
    FILE* b = fopen(dst.c_str(), "wb");
    if (a && b) {
      char buf[8192];
      size_t n;
      while ((n = fread(buf, 1, sizeof(buf), a)) > 0) fwrite(buf, 1, n, b);
    }
    if (a) fclose(a);
    if (b) fclose(b);
  }

  Guard G;
  int32_t warn = 0;
  if ((rc = srs_grid_read(mesh_base, I[SRS_I_FLAG_NO_FLUX_BOUNDARY], &G.grid, &warn))) return rc;
  if (warn) fprintf(stderr, "**Warning: %d channel(s) share an element with another channel.\n", warn);
  const srs_grid* g = G.grid;
  const int N = g->n_ele;
  std::vector<double> aij(5 * (size_t)N), X0(6 * (size_t)N), br(6 * (size_t)N), bnd(6);
  std::vector<int32_t> nij(4 * (size_t)N);
  if ((rc = srs_assemble_operator(N, g->n_face, g->ele1, g->ele2, g->face_type, g->face_c, g->V0, aij.data(), nij.data())))
    return rc;
  if ((rc = srs_equilibrium_state(N, g->domain, g->tropc, P, X0.data(), bnd.data()))) return rc;
  if ((rc = srs_boundary_rates(N, g->n_face, g->ele1, g->face_type, g->face_c, g->V0, P, bnd.data(), br.data()))) return rc;

  const int n_sims = I[SRS_I_N_SIMS] - start_index;
  if (n_sims <= 0) return SRS_OK;
  if ((rc = srs_ctx_create(I[SRS_I_CUDA_DEVICE], &G.ctx))) return rc;
  std::vector<int32_t> perm(N);
  if ((rc = srs_morton_order(N, g->ele_nodes, g->node_xyz, perm.data()))) return rc;
  srs_mesh_desc md{N, aij.data(), nij.data(), br.data(), g->domain, g->tropc, g->V0, g->tt_area, SRS_ORDER_CUSTOM,
                   perm.data()};
  int32_t mesh_id;
  if ((rc = srs_ctx_add_mesh(G.ctx, &md, &mesh_id))) return rc;
  srs_chan_desc cd{g->n_ryr, g->ryr_ele, g->ryr_jsr, g->ryr_nb, g->n_lcc, g->lcc_ele};
  std::vector<uint64_t> seeds(n_sims);
  for (int k = 0; k < n_sims; ++k) {
    const int run = start_index + k;
    // cpu/RunSimulation.cpp:446-449
    int seed = 93843 + I[SRS_I_RAND_SEED_0] + rank * I[SRS_I_RAND_ADD_A] + run * I[SRS_I_RAND_ADD_B];
    if (I[SRS_I_FLAG_CLOCK_SEED]) seed += (int)time(nullptr);
    seeds[k] = (uint64_t)(uint32_t)seed;
    int32_t rid;
    if ((rc = srs_ctx_add_realization(G.ctx, mesh_id, &cd, I, P, seeds[k], (uint64_t)run, &rid))) return rc;
    if ((rc = srs_ctx_extensions(G.ctx, rid, &ext))) return rc;
  }
  const double dt = P[SRS_R_T_STEP];
  const double iv = P[SRS_R_OUTPUT_STATES_INTERVAL];
  const int64_t cap = (int64_t)std::ceil(P[SRS_R_T_FINAL] / iv) + 16;
  if (cap > (1 << 26)) return srs::fail(SRS_E_ARG, "too many states outputs requested (%lld)", (long long)cap);
  const int nr = g->n_ryr, nl = g->n_lcc;
  const int ryr_out = I[SRS_I_FLAG_RYR_OUT] ? 1 : 0;
  gpu_layout = gpu_layout ? 1 : 0;
  const int rec_len = 7 + (gpu_layout ? nr : 0) + (ryr_out ? nr : 0);
  if (rec_len > 255) return srs::fail(SRS_E_ARG, "states record of %d floats exceeds the 255-float limit", rec_len);
  if ((rc = srs_ctx_record_layout(G.ctx, gpu_layout))) return rc;
  if ((rc = srs_ctx_finalize(G.ctx, (int32_t)cap))) return rc;
  std::vector<std::string> state_files(n_sims);
  std::vector<int> grid_counter(n_sims, 0);
  std::vector<float> rec(rec_len);
  std::vector<double> ca_c(nr), ca_j(nr), ca_l(nl);
  const std::string grid_base = dir + "/grid_" + std::to_string(rank);
  auto grid_name = [&](int run, int n) { return grid_base + "_" + std::to_string(run) + "_" + std::to_string(n) + ".vtk"; };

  for (int k = 0; k < n_sims; ++k) {
    const int run = start_index + k;
    std::vector<int32_t> rs(nr, 0), ls(nl, 1), lvs(nl, 2);
    std::vector<double> R(nr + nl);
    // initial channel protocol, cpu/main.cpp:217-250
    if (I[SRS_I_FLAG_LCC]) {
      ls[I[SRS_I_LCC_INDEX]] = 6;
    } else if (I[SRS_I_FLAG_V_CLAMP]) {
      // no channel is opened for a voltage-clamp run
    } else if (I[SRS_I_INIT_RYR] != -1) {
      if (I[SRS_I_INIT_RYR] < 0 || I[SRS_I_INIT_RYR] >= nr)
        return srs::fail(SRS_E_ARG, "Invalid initial RyR index (%d), must be -1 to ignore, or between 0 and %d",
                         I[SRS_I_INIT_RYR], nr);
      rs[I[SRS_I_INIT_RYR]] = 1;
    } else if (nr > 0) {
      // OpenRandomRyR, cpu/InitFuncs.cpp:1420-1475 (linear phi, rate-weighted)
      std::vector<double> rates(nr);
      double total = 0;
      for (int i = 0; i < nr; ++i) {
        int no = 0, nc = 0;
        for (int q = 0; q < 4; ++q) {
          const int nb = g->ryr_nb[4 * i + q];
          if (nb >= 0) { if (rs[nb]) no++; else nc++; }
        }
        const double ca_ss = X0[g->ryr_ele[i]];
        double rate;
        if (rs[i]) {
          rate = exp(0.5 * P[SRS_R_RYR_A_STAR] * (no * P[SRS_R_RYR_EPS_OO] - nc * P[SRS_R_RYR_EPS_CC])) * P[SRS_R_RYR_K_MINUS];
        } else {
          const double phi = P[SRS_R_RYR_PHI_M] * X0[g->ryr_jsr[i]] + P[SRS_R_RYR_PHI_B];
          const double Xc = exp(0.5 * P[SRS_R_RYR_A_STAR] * (nc * P[SRS_R_RYR_EPS_CC] - no * P[SRS_R_RYR_EPS_OO]));
          rate = Xc * phi * P[SRS_R_RYR_K_PLUS] * pow(ca_ss, P[SRS_R_RYR_ETA]);
        }
        rates[i] = rate;
        total += rate;
      }
      uint64_t o[4];
      srs_philox4x64_10(0, 0, (uint64_t)run, SRS_PURPOSE_RANDOM_RYR, seeds[k], SRS_PHILOX_KEY1, o);
      const double r = total * srs_open01(o[0]);
      int open = 0;
      double sum = 0;
      for (int i = 0; i < nr; ++i) {
        sum += rates[i];
        if (sum >= r) { open = i; break; }
      }
      rs[open] = 1;
    }
    // grid #0 and the t = 0 states record, before the loop (cpu/RunSimulation.cpp:421-424)
    if (I[SRS_I_OUTPUT_GRID_FLAG] || (rank == 0 && run == 0)) {
      if ((rc = write_vtk(grid_name(run, 0), run, rank, 0, 0.0, g, X0.data(), rs.data()))) return rc;
    }
    grid_counter[k] = 1;
    for (int i = 0; i < nr; ++i) { ca_c[i] = X0[g->ryr_ele[i]]; ca_j[i] = X0[g->ryr_jsr[i]]; }
    for (int i = 0; i < nl; ++i) ca_l[i] = X0[g->lcc_ele[i]];
    srs_record_fill(0.0, nr, rs.data(), ca_c.data(), ca_j.data(), nl, ls.data(), lvs.data(), ca_l.data(),
                    P[SRS_R_DEFAULT_VM], P, ryr_out, gpu_layout, rec.data());
    state_files[k] = dir + "/states_" + std::to_string(rank) + "_" + std::to_string(run) + ".out";
    {
      FILE* f = fopen(state_files[k].c_str(), "wb");
      if (!f) return srs::fail(SRS_E_IO, "cannot create %s", state_files[k].c_str());
      fclose(f);
    }
    if ((rc = append_records(state_files[k], rec.data(), rec_len))) return rc;
    // residuals and LCC-trigger close time, cpu/RunSimulation.cpp:452-460
    for (int i = 0; i < nr + nl; ++i) {
      uint64_t o[4];
      srs_philox4x64_10(0, (uint64_t)i, (uint64_t)run, SRS_PURPOSE_INIT_R, seeds[k], SRS_PHILOX_KEY1, o);
      R[i] = log(srs_open01(o[0]));
    }
    uint64_t o[4];
    srs_philox4x64_10(0, 0, (uint64_t)run, SRS_PURPOSE_LCC_CLOSE, seeds[k], SRS_PHILOX_KEY1, o);
    const double t_lcc_close = -P[SRS_R_LCC_DURATION] * log(srs_open01(o[0]));
    if (I[SRS_I_FLAG_LCC]) printf("LCC will be closed at %g ms.\n", t_lcc_close);
    // first iteration of the loop (t = 0) on the host: condition, then protocols (:522-542)
    int n_open = 0;
    for (int i = 0; i < nr; ++i) n_open += rs[i];
    double vm = P[SRS_R_DEFAULT_VM];
    int lcc_open = I[SRS_I_FLAG_LCC];
    const bool fid = I[SRS_I_FLAG_FIDELITY_SIM] != 0;
    const bool ffwd = I[SRS_I_GATING_FFWD] != 0;
    if (ffwd) printf("Using accelerated steady-state gating...\n");
    const double t0 = 0.0;
    // With flag_ffwd the gating-only loop runs first (:469-519, on the device);
    // the main loop's first iteration follows at the time it stops.
    const bool cont = ffwd || (t0 <= P[SRS_R_T_FINAL] &&
                               (!fid || ((lcc_open || n_open > 0 || 0.0 > P[SRS_R_C_MIN]) &&
                                         n_open < I[SRS_I_N_RYR_FIDELITY])));
    if (ffwd) {
      // Vm stays at its default until the first ffwd iteration ends (:510-517)
    } else if (cont) {
      if (I[SRS_I_FLAG_LCC] && lcc_open && t0 >= t_lcc_close) { lcc_open = 0; ls[I[SRS_I_LCC_INDEX]] = 1; }
      if (I[SRS_I_FLAG_V_CLAMP]) vm = (t0 <= P[SRS_R_T_CLAMP]) ? P[SRS_R_V_CLAMP] : P[SRS_R_DEFAULT_VM];
      if (I[SRS_I_INIT_RYR] != -1 && t0 < P[SRS_R_INIT_RYR_HOLD] && !rs[I[SRS_I_INIT_RYR]]) rs[I[SRS_I_INIT_RYR]] = 1;
    } else {
      // the loop never runs: the final record is at t = 0
      if ((rc = append_records(state_files[k], rec.data(), rec_len))) return rc;
    }
    if ((rc = srs_state_set(G.ctx, k, X0.data()))) return rc;
    if ((rc = srs_channels_set(G.ctx, k, rs.data(), ls.data(), lvs.data(), R.data()))) return rc;
    if ((rc = srs_globals_set(G.ctx, k, vm, P[SRS_R_DEFAULT_CA_NSR]))) return rc;
    srs_loop_state lsd{};
    lsd.t = 0; lsd.dt_last = 0; lsd.step = 0; lsd.done = cont ? 0 : 1; lsd.lcc_open = lcc_open;
    lsd.t_lcc_close = t_lcc_close; lsd.n_ryr_open_last = n_open; lsd.ca_ss_max_last = 0;
    lsd.ffwd = ffwd ? 1 : 0;
    if ((rc = srs_loop_set(G.ctx, k, &lsd))) return rc;
  }

  // main loop, cpu/RunSimulation.cpp:522-575, in device-resident chunks
  const auto clock0 = std::chrono::steady_clock::now();
  const bool out_grid = I[SRS_I_OUTPUT_GRID_FLAG] != 0;
  const double giv = P[SRS_R_OUTPUT_GRID_INTERVAL];
  const int64_t chunk_max = 20000;
  double t = 0.0;
  std::vector<float> recbuf((size_t)cap * rec_len);
  std::vector<double> Xh(6 * (size_t)N);
  std::vector<int32_t> rsh(nr);
  auto drain = [&](int k) {
    int32_t n = 0, len = 0;
    int r2 = srs_records_read(G.ctx, k, recbuf.data(), (int32_t)cap, &n, &len);
    if (r2) return r2;
    return n ? append_records(state_files[k], recbuf.data(), (size_t)n * len) : SRS_OK;
  };
  for (;;) {
    bool all_done = true;
    for (int k = 0; k < n_sims; ++k) {
      srs_loop_state ls;
      if ((rc = srs_loop_get(G.ctx, k, &ls))) return rc;
      all_done &= ls.done != 0;
    }
    if (all_done) break;
    int64_t n = 0;
    double tn = t;
    bool grid_due = false;
    while (n < chunk_max) {
      tn += dt;
      ++n;
      if (out_grid && srs::output_due(tn, dt, giv)) { grid_due = true; break; }
      if (tn > P[SRS_R_T_FINAL]) break;
    }
    if ((rc = srs_explicit_step(G.ctx, n, SRS_STEP_PROTOCOL | SRS_STEP_GRAPH))) return rc;
    t = tn;
    for (int k = 0; k < n_sims; ++k)
      if ((rc = drain(k))) return rc;
    if (grid_due) {
      for (int k = 0; k < n_sims; ++k) {
        srs_loop_state ls;
        if ((rc = srs_loop_get(G.ctx, k, &ls))) return rc;
        if (ls.done) continue;
        if ((rc = srs_state_get(G.ctx, k, Xh.data()))) return rc;
        if ((rc = srs_channels_get(G.ctx, k, rsh.data(), nullptr, nullptr, nullptr))) return rc;
        if ((rc = write_vtk(grid_name(start_index + k, grid_counter[k]), start_index + k, rank, grid_counter[k], t, g,
                            Xh.data(), rsh.data())))
          return rc;
        grid_counter[k]++;
      }
    }
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - clock0).count();
  // final grid output (cpu/RunSimulation.cpp:582; written only with the grid flag since n > 0)
  if (out_grid) {
    for (int k = 0; k < n_sims; ++k) {
      srs_loop_state ls;
      if ((rc = srs_loop_get(G.ctx, k, &ls))) return rc;
      if ((rc = srs_state_get(G.ctx, k, Xh.data()))) return rc;
      if ((rc = srs_channels_get(G.ctx, k, rsh.data(), nullptr, nullptr, nullptr))) return rc;
      if ((rc = write_vtk(grid_name(start_index + k, grid_counter[k]), start_index + k, rank, grid_counter[k], ls.t, g,
                          Xh.data(), rsh.data())))
        return rc;
    }
  }
  printf("Simulation completed in %f secs (rank %d).\n", secs, rank);
  if (wall_seconds) *wall_seconds = secs;
  return SRS_OK;
}
