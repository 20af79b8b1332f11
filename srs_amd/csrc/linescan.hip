// linescan.hip -- the line-scan operator of linescan/main.cpp on the live
// device state: scan-line set-up (nearest cells, padding grid) on the host,
// the PSF-blurred fluorescence per sample on the device.
#include "dev_common.h"

namespace srs_dev {

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
  // The points any cell of this wavefront can reach: outside them every
  // lane's sum is zero, so the wavefront's partial is written as zero
  // without the loop (the same bits as the full loop).
  int lo = n_f, hi = -1;
  ScanCell sc[SCAN_CELLS_PER_THREAD];
  int clo[SCAN_CELLS_PER_THREAD], chi[SCAN_CELLS_PER_THREAD];  // each cell's points within the cut-off
  for (int q = 0; q < SCAN_CELLS_PER_THREAD; ++q) {
    sc[q] = scan_cell(g, cx[q], cy[q], cz[q]);
    const int j = (blockIdx.x * SCAN_CELLS_PER_THREAD + q) * SCAN_BS + threadIdx.x;
    clo[q] = n_f;
    chi[q] = -1;
    if (j < n) {
      scan_point_range(g, n_f, cx[q], cy[q], cz[q], &clo[q], &chi[q]);
      scan_exact_range(g, sc[q], &clo[q], &chi[q]);
      if (clo[q] <= chi[q]) {
        lo = min(lo, clo[q]);
        hi = max(hi, chi[q]);
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  if (lane == 0) {
    for (int i = 0; i < min(lo, n_f); ++i) red[wave][i] = 0.0;
    for (int i = max(hi + 1, 0); i < n_f; ++i) red[wave][i] = 0.0;
  }
  // Each cell's weights by the walk from the anchor at or below lo (a weight
  // outside the cell's run of points within the cut-off is never used, so it
  // may be anything).
  const int i0 = scan_walk_start(g, lo);
  double ww[SCAN_CELLS_PER_THREAD], wr[SCAN_CELLS_PER_THREAD];
  for (int i = i0; i <= hi; ++i) {
    const double pa = scan_pa(g, i);
    const bool anchor = scan_is_anchor(g, i, i0);  // uniform across the wavefront
    double acc = 0.0;
    for (int q = 0; q < SCAN_CELLS_PER_THREAD; ++q) {
      if (anchor) scan_anchor(g, sc[q], pa, &ww[q], &wr[q]);
      if (i >= clo[q] && i <= chi[q]) acc += vf[q] * ww[q];
      ww[q] *= wr[q];
      wr[q] *= g.q;
    }
    if (i < lo) continue;
    acc = wave_sum_dpp(acc);
    if (lane == 63) red[wave][i] = acc;
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
  // a wavefront per point: each lane sums every 64th workgroup's partial,
  // then the lanes' sums by DPP (a fixed order: deterministic)
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, n_waves = blockDim.x >> 6;
  for (int i = wave; i < n_f; i += n_waves) {
    double s = 0.0;
    for (int b = lane; b < n_blocks; b += 64) s += part[(size_t)b * n_f + i];
    s = wave_sum_dpp(s);
    if (lane == 63) out[n_ca + n_f + i] = s + pad[i];
  }
  for (int i = threadIdx.x; i < n_f; i += blockDim.x) {
    const int e = near_f[i];
    out[n_ca + i] = e >= 0 ? X[cq(D[r], e) + 3] : f0;
  }
  for (int i = threadIdx.x; i < n_ca; i += blockDim.x) {
    const int e = near_ca[i];
    out[i] = e >= 0 ? X[cq(D[r], e)] : c0;
  }
}


}  // namespace srs_dev

extern "C" {

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
  g.inv[0] = 1.0 / (g.sx * g.sx);
  g.inv[1] = 1.0 / (g.sy * g.sy);
  g.inv[2] = 1.0 / (g.sz * g.sz);
  g.ia = g.inv[g.axis];
  g.kw = g.alpha / (g.sx * g.sy * g.sz);
  {
    // The walk is safe when an anchor SCAN_ANCHOR - 1 points before a point
    // of weight ~1 has an exponent below 100 (then a weight exp(-x) of the
    // segment can be lost to the anchor's underflow only for x > 260) and
    // the ratio stays finite over a cell's reach (|da| < 1 + (SCAN_ANCHOR +
    // 1) res).
    const double ia = g.inv[g.axis], h = g.res;
    const double far = (SCAN_ANCHOR - 1) * h;
    g.walk = 0.5 * far * far * ia < 100.0 && h * (1.0 + (SCAN_ANCHOR + 1) * h) * ia < 600.0;
    g.q = std::exp(-h * h * ia);
  }

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


}  // extern "C"
