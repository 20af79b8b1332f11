"""srs_amd -- MI355X-native spark time-step of the superresolution Ca-spark model.

Python face of libsrs_amd.so (C ABI in include/srs_abi.h).  Function names
mirror the reference's C entry points (cpu/parameters.h:252-275):

=====================  ==========================================================
reference              here
=====================  ==========================================================
ReadParameters         read_parameters(path) -> Params
ReadGrid               grid_read(base, no_flux) -> Grid
InitializeExplicit...  assemble_operator(grid) -> (aij, nij)
InitializeDomainStr..  boundary_rates(...), Context(...) upload
ResetDefaultGridSt..   equilibrium_state(grid, reals)
ExplicitStep           Context.explicit_step(n)
UpdateChannelGating    Context.channel_gating()
WriteToStateFiles      states_record(...)
main / RunSimulation   run_cru3d(...)
=====================  ==========================================================

There is no CPU fallback: every compute call goes through the HIP library and
raises if it is missing or no gfx950 device is present.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _abi
from ._abi import SrsError, check, lib

N_INT = _abi.N_INT
N_REAL = _abi.N_REAL
NS = 6
CA, ATP, CMDN, TRPN, CSQN, DYE = range(6)

# parameter indices (cpu/parameters.h:18-92)
I = dict(OUTPUT_GRID_FLAG=0, FLAG_FIDELITY_SIM=1, FLAG_CLOCK_SEED=2, RAND_SEED_0=3, RAND_ADD_A=4, RAND_ADD_B=5,
         INITIALIZE_FROM_FILE=6, CUDA_DEVICE=7, FLAG_SECONDARY_GRADIENT=8, FLAG_NO_FLUX_BOUNDARY=9,
         N_RYR_FIDELITY=10, FLAG_LCC=11, LCC_INDEX=12, FLAG_V_CLAMP=13, FLAG_NO_GATING=14, GATING_FFWD=15,
         FLAG_RYR_OUT=16, N_SIMS=17, INIT_RYR=18)
R = dict(T_FINAL=0, T_STEP=1, OUTPUT_STATES_INTERVAL=2, OUTPUT_GRID_INTERVAL=3, C_MIN=4, D_CA=5, D_CA_JSR=6,
         D_ATP=7, D_CMDN=8, D_TRPN=9, D_CSQN=10, D_DYE=11, D_EGTA=12, C_0_CA=13, C_0_CA_JSR=14, DEFAULT_VM=15,
         DEFAULT_CA_NSR=16, CA_0=17, B_TOT_ATP=18, K_OFF_ATP=19, K_ON_ATP=20, B_TOT_CMDN=21, K_OFF_CMDN=22,
         K_ON_CMDN=23, B_TOT_TRPN=24, K_OFF_TRPN=25, K_ON_TRPN=26, B_TOT_SL_HI=27, K_D_SL_HI=28, B_TOT_SL_LO=29,
         K_D_SL_LO=30, B_TOT_DYE=31, K_OFF_DYE=32, K_ON_DYE=33, B_T_JSR=34, K_M_JSR=35, V_REFILL=36, V_RYR=37,
         V_CELL=38, K_ON_CSQN=39, K_OFF_CSQN=40, RYR_ETA=41, RYR_A_STAR=42, RYR_EPS_CC=43, RYR_EPS_OO=44,
         RYR_K_PLUS=45, RYR_K_MINUS=46, RYR_PHI_M=47, RYR_PHI_B=48, K_D_I=49, K_D_SR=50, A_P=51,
         LCC_DURATION=52, T_CLAMP=53, V_CLAMP=54, INIT_RYR_HOLD=55)

STEP_PLAIN = 0
STEP_PROTOCOL = 1
STEP_GRAPH = 2
ORDER_AS_GIVEN = 0
ORDER_RCM = 1
ORDER_CUSTOM = 2
TUNE_TILE = 1
TUNE_NT_GEOMETRY = 2
TUNE_NT_STATE = 3
TUNE_PLANE_PAD = 4
TUNE_COMPACT = 5
TUNE_OCCUPANCY = 6
TUNE_GROUP = 7
TUNE_LDS = 8
TUNE_DROP_FINISHED = 9
TUNE_GATED_SWEEP = 10
TUNE_ADJ_LEVELS = 11
TUNE_SUBSET_MB = 12
TUNE_ALT_ORDER = 13
TUNE_ADJ_WAVES = 14
TUNE_ADJ_W = 15


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _i32(a):
    return _p(a, C.c_int32)


def _f64(a):
    return _p(a, C.c_double)


def _f32(a):
    return _p(a, C.c_float)


@dataclass
class Params:
    ints: np.ndarray
    reals: np.ndarray
    title: str = "Untitled"

    def copy(self) -> "Params":
        return Params(self.ints.copy(), self.reals.copy(), self.title)

    def seti(self, **kw):
        for k, v in kw.items():
            self.ints[I[k]] = v
        return self

    def setr(self, **kw):
        for k, v in kw.items():
            self.reals[R[k]] = v
        return self


def read_parameters(path: str | None = None) -> Params:
    """ReadParameters (cpu/InitFuncs.cpp:25-381); path None gives the defaults."""
    ints = np.zeros(N_INT, np.int32)
    reals = np.zeros(N_REAL, np.float64)
    title = C.create_string_buffer(_abi.TITLE_LEN)
    check(lib().srs_read_parameters(path.encode() if path else None, _i32(ints), _f64(reals), title))
    return Params(ints, reals, title.value.decode())


@dataclass
class Grid:
    n_ele: int
    n_face: int
    n_nodes: int
    V0: np.ndarray
    tt_area: np.ndarray
    sr_area: np.ndarray
    domain: np.ndarray
    tropc: np.ndarray
    ele_nodes: np.ndarray
    node_xyz: np.ndarray
    ele1: np.ndarray
    ele2: np.ndarray
    face_type: np.ndarray
    face_c: np.ndarray
    ryr_ele: np.ndarray
    ryr_jsr: np.ndarray
    ryr_nb: np.ndarray
    lcc_ele: np.ndarray
    perm: np.ndarray

    def centroids(self) -> np.ndarray:
        """Element centres [n_ele][3], the mean of the four nodes in the order
        the linescan program adds them (linescan/main.cpp:96-103)."""
        X = self.node_xyz.reshape(-1, 3)
        t = self.ele_nodes.reshape(-1, 4)
        return (((X[t[:, 0]] + X[t[:, 1]]) + X[t[:, 2]]) + X[t[:, 3]]) / 4.0

    @property
    def n_ryr(self):
        return len(self.ryr_ele)

    @property
    def n_lcc(self):
        return len(self.lcc_ele)


def _grid_from(gp) -> Grid:
    g = gp.contents
    n, nf, nn = g.n_ele, g.n_face, g.n_nodes

    def arr(ptr, cnt, dt):
        if cnt == 0:
            return np.zeros(0, dt)
        return np.ctypeslib.as_array(ptr, shape=(cnt,)).astype(dt, copy=True)

    out = Grid(n, nf, nn, arr(g.V0, n, np.float64), arr(g.tt_area, n, np.float64), arr(g.sr_area, n, np.float64),
               arr(g.domain, n, np.int32), arr(g.tropc, n, np.int32), arr(g.ele_nodes, 4 * n, np.int32),
               arr(g.node_xyz, 3 * nn, np.float64), arr(g.ele1, nf, np.int32), arr(g.ele2, nf, np.int32),
               arr(g.face_type, nf, np.int32), arr(g.face_c, nf, np.float64), arr(g.ryr_ele, g.n_ryr, np.int32),
               arr(g.ryr_jsr, g.n_ryr, np.int32), arr(g.ryr_nb, 4 * g.n_ryr, np.int32),
               arr(g.lcc_ele, g.n_lcc, np.int32), arr(g.perm, n, np.int32))
    lib().srs_grid_free(gp)
    return out


def grid_read(base: str, no_flux: bool = True) -> tuple[Grid, int]:
    """ReadGrid (cpu/InitFuncs.cpp:384-813): returns the grid and the number of
    duplicate channel elements (the reference's race-condition warnings)."""
    gp = C.POINTER(_abi.SrsGrid)()
    warn = C.c_int32(0)
    check(lib().srs_grid_read(base.encode(), int(no_flux), C.byref(gp), C.byref(warn)))
    return _grid_from(gp), warn.value


@dataclass
class BoxSpec:
    """Rectilinear box split into Kuhn tets (see srs_box_spec)."""
    xs: np.ndarray
    ys: np.ndarray
    zs: np.ndarray
    jsr_lo: tuple = (1.0, 1.0, 1.0)
    jsr_hi: tuple = (0.0, 0.0, 0.0)
    sl_face_z0: bool = False
    tt_axis_x: float = 0.0
    tt_axis_z: float = 0.0
    tt_radius: float = 0.0
    trpn_radius: float = -1.0
    ryr_rows: int = 0
    ryr_cols: int = 0
    ryr_spacing: float = 0.031
    ryr_center: tuple = (0.0, 0.0)
    ryr_z_cyt: float = 0.0
    ryr_z_jsr: float = 0.0
    ryr_mask: np.ndarray | None = None
    lcc_xyz: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))

    def _c(self):
        self._keep = [np.ascontiguousarray(self.xs, np.float64), np.ascontiguousarray(self.ys, np.float64),
                      np.ascontiguousarray(self.zs, np.float64),
                      np.ascontiguousarray(np.asarray(self.lcc_xyz, np.float64).reshape(-1))]
        mask = None
        if self.ryr_mask is not None:
            mask = np.ascontiguousarray(self.ryr_mask, np.uint8).reshape(-1)
            self._keep.append(mask)
        s = _abi.SrsBoxSpec()
        s.nx, s.ny, s.nz = len(self.xs) - 1, len(self.ys) - 1, len(self.zs) - 1
        s.xs, s.ys, s.zs = _f64(self._keep[0]), _f64(self._keep[1]), _f64(self._keep[2])
        s.jsr_lo = (C.c_double * 3)(*self.jsr_lo)
        s.jsr_hi = (C.c_double * 3)(*self.jsr_hi)
        s.sl_face_z0 = int(self.sl_face_z0)
        s.tt_axis_x, s.tt_axis_z, s.tt_radius = self.tt_axis_x, self.tt_axis_z, self.tt_radius
        s.trpn_radius = self.trpn_radius
        s.ryr_rows, s.ryr_cols, s.ryr_spacing = self.ryr_rows, self.ryr_cols, self.ryr_spacing
        s.ryr_center = (C.c_double * 2)(*self.ryr_center)
        s.ryr_z_cyt, s.ryr_z_jsr = self.ryr_z_cyt, self.ryr_z_jsr
        s.ryr_mask = mask.ctypes.data_as(_abi.u8p) if mask is not None else None
        s.n_lcc = len(self._keep[3]) // 3
        s.lcc_xyz = _f64(self._keep[3])
        return s


def box_mesh_write(spec: BoxSpec, base: str) -> None:
    s = spec._c()
    check(lib().srs_box_mesh_write(C.byref(s), base.encode()))


def box_grid(spec: BoxSpec, no_flux: bool = True) -> Grid:
    s = spec._c()
    gp = C.POINTER(_abi.SrsGrid)()
    check(lib().srs_box_grid(C.byref(s), int(no_flux), C.byref(gp)))
    return _grid_from(gp)


def bfs_order(n_ele, ele1, ele2) -> np.ndarray:
    e1 = np.ascontiguousarray(ele1, np.int32)
    e2 = np.ascontiguousarray(ele2, np.int32)
    out = np.zeros(n_ele, np.int32)
    check(lib().srs_bfs_order(n_ele, len(e1), _i32(e1), _i32(e2), _i32(out)))
    return out


def morton_order(g: Grid) -> np.ndarray:
    """The device element order of the cru3d driver: perm[k] = element stored at k."""
    perm = np.zeros(g.n_ele, np.int32)
    check(lib().srs_morton_order(g.n_ele, _i32(np.ascontiguousarray(g.ele_nodes, np.int32)),
                                 _f64(np.ascontiguousarray(g.node_xyz, np.float64)), _i32(perm)))
    return perm


def assemble_operator(g: Grid) -> tuple[np.ndarray, np.ndarray]:
    """InitializeExplicitSparseMatrixFast: aij[5N], nij[4N] in the reference packing."""
    aij = np.zeros(5 * g.n_ele, np.float64)
    nij = np.zeros(4 * g.n_ele, np.int32)
    check(lib().srs_assemble_operator(g.n_ele, g.n_face, _i32(g.ele1), _i32(g.ele2), _i32(g.face_type),
                                      _f64(g.face_c), _f64(g.V0), _f64(aij), _i32(nij)))
    return aij, nij


def equilibrium_state(g: Grid, reals: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    X = np.zeros(6 * g.n_ele, np.float64)
    bnd = np.zeros(6, np.float64)
    check(lib().srs_equilibrium_state(g.n_ele, _i32(g.domain), _i32(g.tropc), _f64(reals), _f64(X), _f64(bnd)))
    return X, bnd


def boundary_rates(g: Grid, reals: np.ndarray, bnd: np.ndarray) -> np.ndarray:
    br = np.zeros(6 * g.n_ele, np.float64)
    check(lib().srs_boundary_rates(g.n_ele, g.n_face, _i32(g.ele1), _i32(g.face_type), _f64(g.face_c),
                                   _f64(g.V0), _f64(reals), _f64(np.ascontiguousarray(bnd, np.float64)), _f64(br)))
    return br


def states_record(t, ryr_state, ca_cyt, ca_jsr, lcc_state, lccv_state, ca_lcc, vm, reals, ryr_out=False,
                  gpu_layout=False) -> np.ndarray:
    rs = np.ascontiguousarray(ryr_state, np.int32)
    ls = np.ascontiguousarray(lcc_state, np.int32)
    lv = np.ascontiguousarray(lccv_state, np.int32)
    cc, cj, cl = (np.ascontiguousarray(a, np.float64) for a in (ca_cyt, ca_jsr, ca_lcc))
    rec = np.zeros(255, np.float32)
    n = lib().srs_states_record(float(t), len(rs), _i32(rs), _f64(cc), _f64(cj), len(ls), _i32(ls), _i32(lv),
                                _f64(cl), float(vm), _f64(np.ascontiguousarray(reals, np.float64)), int(ryr_out),
                                int(gpu_layout), _f32(rec))
    if n < 0:
        check(n)
    return rec[:n].copy()


def run_cru3d(param: str, mesh: str, out_dir: str | None, rank: int = 0, start: int = 0,
              gpu_layout: bool = False) -> float:
    secs = C.c_double(0)
    check(lib().srs_run_cru3d(param.encode(), mesh.encode(), out_dir.encode() if out_dir else None, rank, start,
                              int(gpu_layout), C.byref(secs)))
    return secs.value


def read_extensions(path: str | None = None) -> dict:
    """The model extensions of a parameter file (srs_read_extensions):
    jsr_diffusion, b_sr_sites (uM), k_d_sr_sites (uM)."""
    e = _abi.SrsExtensions()
    check(lib().srs_read_extensions(path.encode() if path else None, C.byref(e)))
    return dict(jsr_diffusion=e.jsr_diffusion, b_sr_sites=e.b_sr_sites, k_d_sr_sites=e.k_d_sr_sites)


def read_scan_parameters(path: str | None = None) -> _abi.SrsScanDesc:
    """The linescan program's ReadParameters (linescan/main.cpp:353-508)."""
    d = _abi.SrsScanDesc()
    check(lib().srs_scan_read_parameters(path.encode() if path else None, C.byref(d)))
    return d


N_THETA = 13


def adjoint_theta() -> np.ndarray:
    """SRS_R_* indices of the parameters srs_linescan_loss_grad differentiates."""
    idx = np.zeros(N_THETA, np.int32)
    check(lib().srs_adjoint_theta(_i32(idx)))
    return idx


def scan_write(path: str, rows) -> None:
    """One of the linescan program's output files (WriteOutputFile,
    linescan/main.cpp:316-336): one row per scan, one column per point."""
    a = np.ascontiguousarray(np.atleast_2d(np.asarray(rows, np.float64)))
    check(lib().srs_scan_write(str(path).encode(), _f64(a), a.shape[0], a.shape[1]))


def dff0(cafb: np.ndarray, f0: float) -> np.ndarray:
    """dF/F0 of a blurred fluorescence scan line against the resting signal."""
    return np.asarray(cafb) / f0 - 1.0


class Context:
    """A device context: meshes, realizations, and the step.

    Replaces the reference's DeviceDomainStruct / DeviceGlobalStruct pair and
    the process-global parameter constant (gpu/DeviceStructs.h:19-58,
    gpu/RunSimulation.cu:16)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        check(lib().srs_ctx_create(device, C.byref(self._h)))
        self._keep = []
        self.n_real = 0
        self._sizes = []
        self._scans = {}

    def tune(self, **kw):
        """Sweep tunables (srs_ctx_tune): tile=, plane_pad=, compact=, group=
        before the meshes / srs_ctx_finalize; nt_geometry=, nt_state=,
        occupancy=, lds=, drop_finished=, gated_sweep= at any time.  None changes
        results."""
        keys = dict(tile=TUNE_TILE, nt_geometry=TUNE_NT_GEOMETRY, nt_state=TUNE_NT_STATE, plane_pad=TUNE_PLANE_PAD,
                    compact=TUNE_COMPACT, occupancy=TUNE_OCCUPANCY, group=TUNE_GROUP, lds=TUNE_LDS,
                    drop_finished=TUNE_DROP_FINISHED, gated_sweep=TUNE_GATED_SWEEP, adj_levels=TUNE_ADJ_LEVELS,
                    subset_mb=TUNE_SUBSET_MB, alt_order=TUNE_ALT_ORDER, adj_waves=TUNE_ADJ_WAVES,
                    adj_w=TUNE_ADJ_W)
        for k, v in kw.items():
            check(lib().srs_ctx_tune(self._h, keys[k], int(v)))
        return self

    def close(self):
        if self._h:
            lib().srs_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_mesh(self, aij, nij, domain, tropc, V0, tt_area, br=None, order=ORDER_RCM, perm=None) -> int:
        """order: ORDER_AS_GIVEN, ORDER_RCM, or a permutation array (perm[k] =
        element stored at device position k)."""
        if perm is not None:
            order = ORDER_CUSTOM
        arrs = [np.ascontiguousarray(aij, np.float64), np.ascontiguousarray(nij, np.int32),
                np.ascontiguousarray(domain, np.int32), np.ascontiguousarray(tropc, np.int32),
                np.ascontiguousarray(V0, np.float64), np.ascontiguousarray(tt_area, np.float64)]
        brr = np.ascontiguousarray(br, np.float64) if br is not None else None
        d = _abi.SrsMeshDesc()
        d.n_ele = len(arrs[2])
        d.aij, d.nij = _f64(arrs[0]), _i32(arrs[1])
        d.br = _f64(brr) if brr is not None else None
        d.domain, d.tropc, d.V0, d.tt_area = _i32(arrs[2]), _i32(arrs[3]), _f64(arrs[4]), _f64(arrs[5])
        d.order = order
        if perm is not None:
            pa = np.ascontiguousarray(perm, np.int32)
            self._keep.append(pa)
            d.perm = _i32(pa)
        mid = C.c_int32()
        check(lib().srs_ctx_add_mesh(self._h, C.byref(d), C.byref(mid)))
        return mid.value

    def add_mesh_grid(self, g: Grid, aij=None, nij=None, br=None, order=ORDER_RCM, perm=None) -> int:
        if aij is None:
            aij, nij = assemble_operator(g)
        return self.add_mesh(aij, nij, g.domain, g.tropc, g.V0, g.tt_area, br=br, order=order, perm=perm)

    def add_realization(self, mesh_id, ryr_ele, ryr_jsr, ryr_nb, lcc_ele, params: Params, seed: int,
                        stream: int) -> int:
        arrs = [np.ascontiguousarray(a, np.int32) for a in (ryr_ele, ryr_jsr, ryr_nb, lcc_ele)]
        d = _abi.SrsChanDesc()
        d.n_ryr, d.ryr_ele, d.ryr_jsr, d.ryr_nb = len(arrs[0]), _i32(arrs[0]), _i32(arrs[1]), _i32(arrs[2])
        d.n_lcc, d.lcc_ele = len(arrs[3]), _i32(arrs[3])
        ints = np.ascontiguousarray(params.ints, np.int32)
        reals = np.ascontiguousarray(params.reals, np.float64)
        rid = C.c_int32()
        check(lib().srs_ctx_add_realization(self._h, mesh_id, C.byref(d), _i32(ints), _f64(reals),
                                            C.c_uint64(seed & (2**64 - 1)), C.c_uint64(stream), C.byref(rid)))
        self.n_real += 1
        return rid.value

    def record_layout(self, gpu_layout: bool):
        """States records with Ca_cyt at each RyR (the GPU program's layout); before finalize."""
        check(lib().srs_ctx_record_layout(self._h, int(gpu_layout)))

    def extensions(self, r: int, jsr_diffusion: int = 0, b_sr_sites: float = 0.0, k_d_sr_sites: float = 0.6):
        """Model extensions of realization r (srs_ctx_extensions): Ca diffusion
        with D_CAJSR in jSR elements, SR-site rapid buffer in the cytosol."""
        e = _abi.SrsExtensions(int(jsr_diffusion), float(b_sr_sites), float(k_d_sr_sites))
        check(lib().srs_ctx_extensions(self._h, r, C.byref(e)))

    def finalize(self, record_capacity: int = 0):
        check(lib().srs_ctx_finalize(self._h, record_capacity))

    def info(self, r):
        n, nr, nl = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().srs_real_info(self._h, r, C.byref(n), C.byref(nr), C.byref(nl)))
        return n.value, nr.value, nl.value

    def state_set(self, r, X):
        X = np.ascontiguousarray(X, np.float64)
        check(lib().srs_state_set(self._h, r, _f64(X)))

    def state_get(self, r) -> np.ndarray:
        n, _, _ = self.info(r)
        X = np.zeros(6 * n, np.float64)
        check(lib().srs_state_get(self._h, r, _f64(X)))
        return X

    def channels_set(self, r, ryr_state=None, lcc_state=None, lccv_state=None, R_=None):
        a = [None if v is None else np.ascontiguousarray(v, t)
             for v, t in ((ryr_state, np.int32), (lcc_state, np.int32), (lccv_state, np.int32), (R_, np.float64))]
        check(lib().srs_channels_set(self._h, r, _i32(a[0]) if a[0] is not None else None,
                                     _i32(a[1]) if a[1] is not None else None,
                                     _i32(a[2]) if a[2] is not None else None,
                                     _f64(a[3]) if a[3] is not None else None))

    def channels_get(self, r):
        _, nr, nl = self.info(r)
        rs, ls, lv, Rr = (np.zeros(nr, np.int32), np.zeros(nl, np.int32), np.zeros(nl, np.int32),
                          np.zeros(nr + nl, np.float64))
        check(lib().srs_channels_get(self._h, r, _i32(rs), _i32(ls), _i32(lv), _f64(Rr)))
        return rs, ls, lv, Rr

    def globals_set(self, r, vm, ca_nsr):
        check(lib().srs_globals_set(self._h, r, float(vm), float(ca_nsr)))

    def globals_get(self, r):
        a, b = C.c_double(), C.c_double()
        check(lib().srs_globals_get(self._h, r, C.byref(a), C.byref(b)))
        return a.value, b.value

    def loop_set(self, r, t=0.0, dt_last=0.0, step=0, done=0, lcc_open=0, t_lcc_close=0.0, n_open=0, ca_max=0.0,
                 ffwd=0):
        s = _abi.SrsLoopState(t, dt_last, step, done, lcc_open, t_lcc_close, n_open, ca_max, ffwd)
        check(lib().srs_loop_set(self._h, r, C.byref(s)))

    def loop_get(self, r) -> dict:
        s = _abi.SrsLoopState()
        check(lib().srs_loop_get(self._h, r, C.byref(s)))
        return {k: getattr(s, k) for k, _ in _abi.SrsLoopState._fields_}

    def checkpoint(self, r) -> bytes:
        """The realization's fields, channels and loop state as one buffer."""
        n = C.c_size_t()
        check(lib().srs_checkpoint_size(self._h, r, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        check(lib().srs_checkpoint_save(self._h, r, buf, n.value))
        return buf.raw

    def restore(self, r, data: bytes):
        """Loads a buffer from checkpoint() into realization r."""
        buf = C.create_string_buffer(bytes(data), len(data))
        check(lib().srs_checkpoint_load(self._h, r, buf, len(data)))

    def explicit_step(self, n_steps=1, flags=STEP_PLAIN):
        check(lib().srs_explicit_step(self._h, int(n_steps), flags))

    def prepare(self, n_steps, flags=STEP_GRAPH):
        """Captures the graphs explicit_step(n_steps, flags) will replay."""
        check(lib().srs_explicit_step_prepare(self._h, int(n_steps), flags))

    def channel_gating(self):
        check(lib().srs_channel_gating(self._h))

    def replay(self, r, schedule: np.ndarray):
        sch = np.ascontiguousarray(schedule, np.int32)
        check(lib().srs_explicit_step_replay(self._h, r, sch.shape[0], _i32(sch)))

    def channel_ca(self, r):
        _, nr, nl = self.info(r)
        a, b, c = np.zeros(nr), np.zeros(nr), np.zeros(nl)
        check(lib().srs_channel_ca(self._h, r, _f64(a), _f64(b), _f64(c)))
        return a, b, c

    def records(self, r, max_records=1 << 20) -> np.ndarray:
        n, ln = C.c_int32(), C.c_int32()
        check(lib().srs_records_read(self._h, r, None, 0, C.byref(n), C.byref(ln)))
        # first call only reports the record length; read for real now
        buf = np.zeros(max_records * max(ln.value, 1), np.float32)
        check(lib().srs_records_read(self._h, r, _f32(buf), max_records, C.byref(n), C.byref(ln)))
        return buf[: n.value * ln.value].reshape(n.value, ln.value)

    def sync(self):
        check(lib().srs_ctx_sync(self._h))

    def last_timing(self):
        a, b = C.c_double(), C.c_double()
        check(lib().srs_last_timing(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def add_scan(self, mesh_id: int, centers: np.ndarray, desc: _abi.SrsScanDesc) -> tuple[int, int, int]:
        """A scan line over a mesh (after finalize); returns (scan id, n_f, n_ca)."""
        c = np.ascontiguousarray(centers, np.float64).reshape(-1)
        sid, nf, nca = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().srs_ctx_add_scan(self._h, mesh_id, _f64(c), C.byref(desc), C.byref(sid), C.byref(nf),
                                     C.byref(nca)))
        self._scans[sid.value] = (nf.value, nca.value)
        return sid.value, nf.value, nca.value

    def scan_sample(self, r: int, scan_id: int):
        """(Ca nearest [n_ca], CaF nearest [n_f], CaF blurred [n_f]) of the current state."""
        nf, nca = self._scans[scan_id]
        ca, caf, cafb = np.zeros(nca), np.zeros(nf), np.zeros(nf)
        check(lib().srs_scan_sample(self._h, r, scan_id, _f64(ca), _f64(caf), _f64(cafb)))
        return ca, caf, cafb

    def params_set(self, r: int, reals: np.ndarray):
        """Replaces realization r's parameter vector (SRS_R_* layout)."""
        P = np.ascontiguousarray(reals, np.float64)
        check(lib().srs_params_set(self._h, r, _f64(P)))

    def record_schedule(self, r: int, n_steps: int) -> np.ndarray:
        """n_steps steps with gating; row k = RyR states during step k."""
        _, nr, _ = self.info(r)
        out = np.zeros((n_steps, nr), np.int32)
        check(lib().srs_record_schedule(self._h, r, int(n_steps), _i32(out)))
        return out

    def record_schedules(self, n_steps: int) -> list:
        """record_schedule for every realization of the context at once."""
        nrs = [self.info(r)[1] for r in range(self.n_real)]
        flat = np.zeros(n_steps * sum(nrs), np.int32)
        check(lib().srs_record_schedules(self._h, int(n_steps), _i32(flat)))
        out, o = [], 0
        for nr in nrs:
            out.append(flat[o:o + n_steps * nr].reshape(n_steps, nr))
            o += n_steps * nr
        return out

    def loss_grad(self, r: int, scan_id: int, schedule: np.ndarray, sample_steps, target: np.ndarray, ckpt: int = 256,
                  grad: bool = True):
        """Line-scan loss of a replayed run and (optionally) its adjoint
        gradient in adjoint_theta() order; the realization is left as it was."""
        sch = np.ascontiguousarray(schedule, np.int32)
        ss = np.ascontiguousarray(sample_steps, np.int64)
        tg = np.ascontiguousarray(target, np.float64)
        loss = C.c_double()
        g = np.zeros(N_THETA) if grad else None
        check(lib().srs_linescan_loss_grad(self._h, r, scan_id, int(sch.shape[0]), _i32(sch), len(ss),
                                           ss.ctypes.data_as(C.POINTER(C.c_int64)), _f64(tg), int(ckpt),
                                           C.byref(loss), _f64(g) if grad else None))
        return (loss.value, g) if grad else loss.value

    def loss_grad_batch(self, reals, scan_id: int, schedules, sample_steps, targets, ckpt: int = 0,
                        max_batch: int = 0, grad: bool = True, scans: bool = False):
        """loss_grad for several realizations stepped together: schedules[m]
        is realization reals[m]'s (n_steps, n_ryr) schedule, targets[m] its
        (n_samples, n_f) target.  Returns losses (n,), gradients (n, 13) if
        grad, and the simulated dF/F0 (n, n_samples, n_f) if scans."""
        ids = np.ascontiguousarray(reals, np.int32)
        n = len(ids)
        n_steps = int(np.shape(schedules[0])[0])
        for sc in schedules:
            if np.shape(sc)[0] != n_steps:
                raise ValueError("every schedule needs the same number of steps")
        sch = np.ascontiguousarray(np.concatenate([np.asarray(sc, np.int32).ravel() for sc in schedules]), np.int32)
        ss = np.ascontiguousarray(sample_steps, np.int64)
        tg = np.ascontiguousarray(targets, np.float64)
        if tg.shape[:2] != (n, len(ss)):
            raise ValueError("targets must be (n_real, n_samples, n_f)")
        losses = np.zeros(n)
        g = np.zeros((n, N_THETA)) if grad else None
        sc = np.zeros(tg.shape) if scans else None
        check(lib().srs_linescan_loss_grad_batch(self._h, n, _i32(ids), scan_id, n_steps, _i32(sch), len(ss),
                                                 ss.ctypes.data_as(C.POINTER(C.c_int64)), _f64(tg), int(ckpt),
                                                 int(max_batch), _f64(losses), _f64(g) if grad else None,
                                                 _f64(sc) if scans else None))
        out = (losses,) + ((g,) if grad else ()) + ((sc,) if scans else ())
        return out if len(out) > 1 else losses

    def adjoint_plan(self) -> dict:
        """Checkpoint plan of the last loss/gradient call: levels, their
        strides (the last is the recomputed segment's length), sub-batch."""
        lv, sb = C.c_int32(), C.c_int32()
        st = (C.c_int64 * 6)()
        check(lib().srs_adjoint_last_plan(self._h, C.byref(lv), st, C.byref(sb)))
        return dict(levels=lv.value, strides=[int(x) for x in st[:lv.value]], sub_batch=sb.value)

    def time_sweep(self, reps: int):
        a, b = C.c_double(), C.c_double()
        check(lib().srs_time_sweep(self._h, reps, C.byref(a), C.byref(b)))
        return a.value, b.value


__all__ = ["Params", "Grid", "BoxSpec", "Context", "SrsError", "read_parameters", "read_scan_parameters", "scan_write",
           "dff0", "grid_read", "box_grid",
           "box_mesh_write", "bfs_order", "assemble_operator", "equilibrium_state", "boundary_rates",
           "states_record", "run_cru3d", "I", "R"]
