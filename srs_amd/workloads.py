"""Synthetic spark workloads on Kuhn-tet box meshes.

* ``spark_box``: cytosol box 1.0 x 1.0 x 0.6 um, jSR slab 0.4 x 0.4 x 0.04 um
  15 nm above the sarcolemmal z = 0 face, 7x7 RyR lattice at 31 nm, 4 LCCs,
  cubes of h = 40 nm (56,250 tets).  The plain CPU-runnable spark.
* ``dyad_box``: cytosol box 2 x 2 x 1 um with a 200 nm T-tubule cylinder
  along y and a 1.0 x 1.0 x 0.04 um jSR slab 15 nm above it, h = 20 nm
  (about 3M tets), 10x10 RyR lattice at 70% occupancy, 8 LCCs.  The plain
  GPU workload.
* ``ensemble``: many spark_box realizations on meshes of h = 40/30/25 nm
  with irregular RyR clusters (9-100 RyRs, 50-90% occupancy) and SR loads
  of 0.5-1.25 mM.

Species and kinetics: Ca D = 0.22 um^2/ms, ATP 455 uM (D 0.14), CaM 24 uM
(D 0.025), Fluo-4 50 uM (D 0.042), TnC 70 uM, CSQ 30 mM in the jSR,
Ca 0.1 uM cytosol / 1 mM jSR at equilibrium, 2-state RyR
k_open = 1e-4 [Ca]^2.1 /ms (phi = 1, A* = 0), k_close = 0.5 /ms, 0.15 pA
at 1 mM gradient, LCCs clamped at 0 mV, explicit step 0.9 x the mesh
stability bound.

Two workload features have no counterpart in the reference's equations and
are left out of the parameter vectors so that the reference reproduces every
run: a jSR Ca diffusion coefficient different from the cytosolic one (the
reference assembles the operator with D = 1 and scales Ca by D_CA in both
domains, cpu/InitFuncs.cpp:1081, cpu/RunSimulation.cpp:62) and SR binding
sites (the reference loads SR surface areas but uses them in no rate).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import BoxSpec, Grid, Params, assemble_operator, box_grid, equilibrium_state, read_parameters

V_CELL = 25.84e3


def spark_params(t_final=20.0, states_interval=0.1, dt=None) -> Params:
    p = read_parameters(None)
    p.title = "spark_box"
    p.setr(D_CA=0.22, D_CA_JSR=0.22, D_ATP=0.14, D_CMDN=0.025, D_DYE=0.042,
           B_TOT_ATP=455.0, B_TOT_CMDN=24.0, B_TOT_DYE=50.0, B_TOT_TRPN=70.0, B_T_JSR=30000.0,
           C_0_CA=0.1, C_0_CA_JSR=1000.0, DEFAULT_CA_NSR=1000.0,
           RYR_K_PLUS=1e-4, RYR_ETA=2.1, RYR_K_MINUS=0.5, RYR_PHI_M=0.0, RYR_A_STAR=0.0,
           V_RYR=0.15 / (1000.0 * V_CELL * 96485 * 2 * 1e-6),
           V_CLAMP=0.0, T_CLAMP=1e9, T_FINAL=t_final, OUTPUT_STATES_INTERVAL=states_interval)
    p.seti(FLAG_V_CLAMP=1, FLAG_FIDELITY_SIM=0, FLAG_CLOCK_SEED=0, INIT_RYR=24)
    if dt is not None:
        p.setr(T_STEP=dt, INIT_RYR_HOLD=dt)
    return p


def stable_dt(aij: np.ndarray, D: float, safety: float = 0.9) -> float:
    """0.9 x the positivity-preserving explicit bound 1/(D max|a_ii|)."""
    return safety / (D * float(np.max(np.abs(aij[0::5]))))


def stable_dt_domains(aij: np.ndarray, domain: np.ndarray, d_cyt: float, d_jsr: float, safety: float = 0.9) -> float:
    """The explicit bound with a per-domain Ca coefficient: 0.9 / max over
    rows of D_row |a_ii| (jSR rows at d_jsr, cytosol rows at d_cyt; the
    rapid buffers only slow the Ca rate, so they leave the bound as it is)."""
    diag = np.abs(aij[0::5])
    jsr = np.asarray(domain) == 1
    m = max(d_cyt * float(diag[~jsr].max(initial=0.0)), d_jsr * float(diag[jsr].max(initial=0.0)))
    return safety / m


# The configurations' stated physics that the reference cannot express: Ca
# diffuses at 0.35 um^2/ms in the jSR, and the cytosol holds 47 uM of SR
# Ca-binding sites (dissociation constant 0.6 uM, 60/ms off over 100/mM/ms
# on), a rapid buffer.
CONFIG_D_JSR = 0.35
CONFIG_SR_SITES = dict(jsr_diffusion=1, b_sr_sites=47.0, k_d_sr_sites=0.6)


def with_config_physics(pr: "Prepared") -> "Prepared":
    """The prepared workload with D_CAJSR = 0.35 and dt from the per-domain
    bound; the SR sites are a realization's extension (CONFIG_SR_SITES)."""
    p = pr.params.copy()
    p.setr(D_CA_JSR=CONFIG_D_JSR)
    dt = stable_dt_domains(pr.aij, pr.grid.domain, p.reals[5], CONFIG_D_JSR)
    p.setr(T_STEP=dt, INIT_RYR_HOLD=dt)
    return Prepared(pr.grid, pr.aij, pr.nij, pr.X0, pr.bnd, p)


def _planes(lo, hi, n):
    return np.linspace(lo, hi, n + 1)


def spark_box_spec(h=0.04, n_xy=None, n_z_rest=None, jsr_layers=1, ryr_rows=7, ryr_cols=7, ryr_mask=None,
                   n_lcc=4) -> BoxSpec:
    """Cytosol 1.0 x 1.0 x 0.6 um: cleft 0 - 0.015 um, jSR 0.015 - 0.055 um."""
    n_xy = n_xy or int(round(1.0 / h))
    xs = _planes(0.0, 1.0, n_xy)
    zr = n_z_rest or int(round(0.545 / h)) - 1
    zs = np.concatenate([[0.0, 0.015], 0.015 + np.arange(1, jsr_layers + 1) * (0.04 / jsr_layers),
                         np.linspace(0.055, 0.6, zr + 1)[1:]])
    c = 0.52 if n_xy == 25 else 0.5
    half = 0.2
    sp = 0.031
    lcc = []
    for dx, dy in ((-1.5, -1.5), (1.5, -1.5), (-1.5, 1.5), (1.5, 1.5))[:n_lcc]:
        lcc.append((c + dx * sp, c + dy * sp, 0.0075))
    return BoxSpec(xs=xs, ys=xs.copy(), zs=zs, jsr_lo=(c - half, c - half, 0.015), jsr_hi=(c + half, c + half, 0.055),
                   sl_face_z0=True, trpn_radius=-1.0, ryr_rows=ryr_rows, ryr_cols=ryr_cols, ryr_spacing=sp,
                   ryr_center=(c, c), ryr_z_cyt=0.0075, ryr_z_jsr=0.015 + 0.02 / jsr_layers, ryr_mask=ryr_mask,
                   lcc_xyz=np.array(lcc))


def spark_box(h=0.04) -> BoxSpec:
    """h = 40 nm: 25 x 25 x 15 cubes, 56,250 tets."""
    return spark_box_spec(h=h)


def dyad_box_spec(h=0.02, seed=38376427, occupancy=0.7) -> BoxSpec:
    """2 x 2 x 1 um, T-tubule r = 100 nm along y at x = 1.0, z = 0.1; cleft
    0.200 - 0.215 um above its top; jSR 0.215 - 0.255 um over 1.0 x 1.0 um."""
    n_xy = int(round(2.0 / h))
    xs = _planes(0.0, 2.0, n_xy)
    n_low = int(round(0.2 / h))
    n_top = 50 - n_low - 3 if abs(h - 0.02) < 1e-12 else int(round(0.745 / h))
    zs = np.concatenate([np.linspace(0.0, 0.2, n_low + 1), [0.215, 0.235, 0.255], np.linspace(0.255, 1.0, n_top + 1)[1:]])
    rng = np.random.Generator(np.random.Philox(key=seed))
    n = 100
    mask = np.zeros(n, np.uint8)
    mask[rng.permutation(n)[: int(round(occupancy * n))]] = 1
    sp = 0.031
    lcc = [(1.0 + dx * sp, 1.0 + dy * sp, 0.2075) for dx in (-0.5, 0.5) for dy in (-1.5, -0.5, 0.5, 1.5)]
    return BoxSpec(xs=xs, ys=xs.copy(), zs=zs, jsr_lo=(0.5, 0.5, 0.215), jsr_hi=(1.5, 1.5, 0.255), sl_face_z0=False,
                   tt_axis_x=1.0, tt_axis_z=0.1, tt_radius=0.1, trpn_radius=-1.0, ryr_rows=10, ryr_cols=10,
                   ryr_spacing=sp, ryr_center=(1.0, 1.0), ryr_z_cyt=0.2075, ryr_z_jsr=0.225,
                   ryr_mask=mask.reshape(10, 10), lcc_xyz=np.array(lcc))


def tiny_box_spec(nx=4, ny=4, nz=3, tt=False, jsr=True, ryr=2, lcc=1, sl=True) -> BoxSpec:
    """A few hundred tets with every feature the step has: jSR, SL face, T-tubule
    hole, RyRs on the jSR face, LCCs."""
    xs = np.linspace(0.0, 0.04 * nx, nx + 1)
    ys = np.linspace(0.0, 0.04 * ny, ny + 1)
    zs = np.concatenate([[0.0, 0.015], 0.015 + 0.04 * np.arange(1, nz)])
    cx, cy = 0.02 * nx, 0.02 * ny
    spec = BoxSpec(xs=xs, ys=ys, zs=zs, sl_face_z0=sl, trpn_radius=-1.0, ryr_rows=ryr, ryr_cols=ryr,
                   ryr_spacing=0.031, ryr_center=(cx, cy), ryr_z_cyt=0.0075, ryr_z_jsr=0.035,
                   lcc_xyz=np.array([(cx + 0.01, cy - 0.03, 0.0075)] * lcc).reshape(-1, 3))
    if jsr:
        spec.jsr_lo = (cx - 0.06, cy - 0.06, 0.015)
        spec.jsr_hi = (cx + 0.06, cy + 0.06, 0.055)
    if tt:
        spec.tt_axis_x = 0.0
        spec.tt_axis_z = 0.12
        spec.tt_radius = 0.05
        spec.trpn_radius = 0.08
    return spec


@dataclass
class Prepared:
    grid: Grid
    aij: np.ndarray
    nij: np.ndarray
    X0: np.ndarray
    bnd: np.ndarray
    params: Params


def prepare(spec: BoxSpec, params: Params | None = None, dt=None, t_final=None) -> Prepared:
    g = box_grid(spec)
    aij, nij = assemble_operator(g)
    p = (params or spark_params()).copy()
    if dt is None and not p.reals[1] > 0:
        dt = stable_dt(aij, p.reals[5])
    if dt is not None:
        p.setr(T_STEP=dt, INIT_RYR_HOLD=dt)
    if t_final is not None:
        p.setr(T_FINAL=t_final)
    if p.ints[18] >= g.n_ryr:  # the initial RyR must exist on this mesh
        p.ints[18] = 0 if g.n_ryr > 0 else -1
    X0, bnd = equilibrium_state(g, p.reals)
    return Prepared(g, aij, nij, X0, bnd, p)


def prepare_spark(h=0.04, **kw) -> Prepared:
    spec = spark_box(h)
    g = box_grid(spec)
    aij, nij = assemble_operator(g)
    p = spark_params(**{k: v for k, v in kw.items() if k in ("t_final", "states_interval")})
    p.setr(T_STEP=stable_dt(aij, p.reals[5]))
    p.setr(INIT_RYR_HOLD=p.reals[1])
    X0, bnd = equilibrium_state(g, p.reals)
    return Prepared(g, aij, nij, X0, bnd, p)


def prepare_dyad(h=0.02, t_final=100.0) -> Prepared:
    g = box_grid(dyad_box_spec(h))
    aij, nij = assemble_operator(g)
    p = spark_params(t_final=t_final)
    p.title = "dyad_box"
    p.setr(T_STEP=stable_dt(aij, p.reals[5]))
    p.setr(INIT_RYR_HOLD=p.reals[1])
    p.seti(INIT_RYR=0)
    X0, bnd = equilibrium_state(g, p.reals)
    return Prepared(g, aij, nij, X0, bnd, p)


def ensemble_members(n_real: int, seed: int = 38376427):
    """Per-realization draws: mesh index (round robin over h = 40/30/25 nm),
    RyR count in 9..100, occupancy in 0.5-0.9, SR load in 500-1250 uM, and
    the lattice mask (rows x cols, up to 12 x 12 so it fits the jSR)."""
    rng = np.random.Generator(np.random.Philox(key=seed + 1))
    out = []
    for r in range(n_real):
        n = int(rng.integers(9, 101))
        occ = float(rng.uniform(0.5, 0.9))
        L = int(min(12, np.ceil(np.sqrt(n / occ))))
        mask = np.zeros(L * L, np.uint8)
        mask[rng.permutation(L * L)[:n]] = 1
        sr = float(rng.uniform(500.0, 1250.0))
        out.append(dict(mesh=r % 3, n_ryr=n, L=L, mask=mask.reshape(L, L), ca_sr=sr))
    return out


ENSEMBLE_H = (0.04, 0.0303030303030303, 0.025)
ENSEMBLE_SHAPES = ((25, 13, 1), (33, 18, 1), (40, 21, 2))  # (n_xy, z layers above the jSR, jSR layers)


def ensemble_mesh_spec(m: int) -> BoxSpec:
    """Mesh m of the ensemble with a full 12 x 12 RyR lattice; realizations
    pick subsets of its positions."""
    n_xy, zr, jl = ENSEMBLE_SHAPES[m]
    return spark_box_spec(h=1.0 / n_xy, n_xy=n_xy, n_z_rest=zr, jsr_layers=jl, ryr_rows=12, ryr_cols=12)


def lattice_subset(mask: np.ndarray, full_rows=12, full_cols=12):
    """Indices into the full 12x12 lattice (centred) of an L x L mask, and the
    up/down/left/right neighbour table of the occupied sites."""
    L = mask.shape[0]
    off = (full_rows - L) // 2
    lab = -np.ones((L, L), np.int64)
    idx = []
    k = 0
    for i in range(L):
        for j in range(L):
            if mask[i, j]:
                lab[i, j] = k
                idx.append((i + off) * full_cols + (j + off))
                k += 1
    nb = []
    for i in range(L):
        for j in range(L):
            if mask[i, j]:
                nb += [lab[i - 1, j] if i > 0 else -1, lab[i + 1, j] if i < L - 1 else -1,
                       lab[i, j - 1] if j > 0 else -1, lab[i, j + 1] if j < L - 1 else -1]
    return np.array(idx, np.int64), np.array(nb, np.int32)


def initial_channels(pr: Prepared, open_ryr=None, seed=0):
    """Channel states at the start of a run: the initial RyR open (the
    parameter vector's ryr_open_init unless given), LCCs closed-resting
    (state 1, gate open 2, cpu/InitFuncs.cpp:971-982) and residuals ln U."""
    g = pr.grid
    rs = np.zeros(g.n_ryr, np.int32)
    i0 = int(pr.params.ints[18]) if open_ryr is None else open_ryr
    if 0 <= i0 < g.n_ryr:
        rs[i0] = 1
    rng = np.random.default_rng(seed)
    R = np.log(rng.random(g.n_ryr + g.n_lcc) * (1 - 2**-53) + 2**-54)
    return rs, np.ones(g.n_lcc, np.int32), np.full(g.n_lcc, 2, np.int32), R


def ensemble_meshes():
    """The three ensemble meshes (h = 40/30/25 nm): (grid, aij, nij, dt) each,
    dt = 0.9 x the stability bound at D_Ca = 0.22."""
    out = []
    for m in range(3):
        g = box_grid(ensemble_mesh_spec(m))
        aij, nij = assemble_operator(g)
        out.append((g, aij, nij, stable_dt(aij, 0.22)))
    return out


def ensemble_member(mem: dict, mesh, r: int, seed: int = 38376427, t_final: float = 50.0, with_state: bool = True):
    """Everything one ensemble realization starts from: its parameters (dt of
    its mesh, SR load as CA_SR, the middle RyR of its cluster as the initial
    one), its channel arrays (a subset of the mesh's 12 x 12 lattice), its
    Philox key and stream, and (with_state) the equilibrium fields, initial
    channel states, ln U residuals and globals."""
    g, aij, nij, dt = mesh
    idx, nb = lattice_subset(mem["mask"])
    nr = len(idx)
    p = spark_params(t_final=t_final)
    p.setr(T_STEP=dt, INIT_RYR_HOLD=dt, C_0_CA_JSR=mem["ca_sr"], DEFAULT_CA_NSR=mem["ca_sr"])
    p.seti(INIT_RYR=int(nr // 2))
    out = dict(p=p, ryr_ele=g.ryr_ele[idx], ryr_jsr=g.ryr_jsr[idx], ryr_nb=nb, lcc_ele=g.lcc_ele, key=seed + r,
               stream=r, n_ryr=nr)
    if with_state:
        X0, _ = equilibrium_state(g, p.reals)
        rs = np.zeros(nr, np.int32)
        rs[nr // 2] = 1
        rng = np.random.default_rng(seed + r)
        R = np.log(rng.random(nr + g.n_lcc) * (1 - 2**-53) + 2**-54)
        out.update(X0=X0, ryr_state=rs, lcc_state=np.ones(g.n_lcc, np.int32),
                   lccv_state=np.full(g.n_lcc, 2, np.int32), R=R, gv=(0.0, float(p.reals[16])))
    return out


def ensemble_context(n_real: int, seed: int = 38376427, record_capacity: int = 0, t_final: float = 50.0,
                     order: str = "block", tune: dict | None = None, only=None, log=None):
    """The batched Monte Carlo workload in one context: returns (ctx, dt per
    realization, total elements over realizations, algorithmic sweep bytes
    per step).  Algorithmic bytes: 96 B per element of every realization (6
    FP64 species read and written) plus 65 B per element of each distinct
    mesh (the 4 int32 + 5 FP64 operator row and 9 B of per-element
    constants), counted once per step however many realizations share it.
    only: indices of the members to include (default all n_real), e.g. one
    member at a time for the plain sequential comparison."""
    from . import ORDER_AS_GIVEN, ORDER_RCM, Context

    members = ensemble_members(n_real, seed)
    picked = list(range(n_real)) if only is None else list(only)
    ctx = Context(0)
    tune = dict(tune or {})
    ctx.tune(**{k: tune[k] for k in ("tile", "plane_pad", "compact", "group") if k in tune})
    meshes = ensemble_meshes()
    mids = []
    for g, aij, nij, _ in meshes:
        if order == "block":
            mids.append(ctx.add_mesh(aij, nij, g.domain, g.tropc, g.V0, g.tt_area, perm=kuhn_block_order(g)))
        else:
            mids.append(ctx.add_mesh(aij, nij, g.domain, g.tropc, g.V0, g.tt_area,
                                     order=ORDER_RCM if order == "rcm" else ORDER_AS_GIVEN))
    dts, n_ele, nbytes = [], 0, 0
    for m in sorted({members[r]["mesh"] for r in picked}):
        nbytes += 65 * meshes[m][0].n_ele
    for k, r in enumerate(picked):
        m = members[r]["mesh"]
        e = ensemble_member(members[r], meshes[m], r, seed, t_final, with_state=False)
        ctx.add_realization(mids[m], e["ryr_ele"], e["ryr_jsr"], e["ryr_nb"], e["lcc_ele"], e["p"], seed=e["key"],
                            stream=e["stream"])
        dts.append(meshes[m][3])
        n_ele += meshes[m][0].n_ele
        nbytes += 96 * meshes[m][0].n_ele
        if log and (k + 1) % 256 == 0:
            log(f"  added {k + 1} realizations")
    ctx.finalize(record_capacity)
    for k, r in enumerate(picked):
        m = members[r]["mesh"]
        e = ensemble_member(members[r], meshes[m], r, seed, t_final)
        ctx.state_set(k, e["X0"])
        ctx.channels_set(k, e["ryr_state"], e["lcc_state"], e["lccv_state"], e["R"])
        ctx.globals_set(k, *e["gv"])
        if log and (k + 1) % 256 == 0:
            log(f"  initialised {k + 1} realizations")
    ctx.tune(**{k: tune[k] for k in ("nt_geometry", "nt_state", "occupancy", "lds") if k in tune})
    return ctx, np.array(dts), n_ele, nbytes


def kuhn_block_order(g: Grid, block: int = 64) -> np.ndarray:
    """Device element order for a Kuhn box mesh: cubes ordered with z fastest,
    then x, then y (the longest stride is nz*nx cubes), grouped in blocks of
    `block` cubes, and inside a block the six tets of one Kuhn type stored
    together.  A wavefront then holds one tet type over consecutive cubes, so
    its neighbour gathers fall at (nearly) uniform offsets.  Returns perm with
    perm[k] = reference-numbered element stored at position k."""
    n = g.n_ele
    ftype = g.perm % 6  # file order is cube-by-cube, 6 tets per cube, one per axis permutation
    nodes = g.ele_nodes.reshape(-1, 4)
    xyz = g.node_xyz.reshape(-1, 3)
    cen = xyz[nodes].mean(axis=1)
    planes = []
    for ax in range(3):
        planes.append(np.unique(xyz[:, ax]))
    idx = [np.clip(np.searchsorted(planes[ax], cen[:, ax]) - 1, 0, len(planes[ax]) - 2) for ax in range(3)]
    nx, nz = len(planes[0]) - 1, len(planes[2]) - 1
    key = (idx[1].astype(np.int64) * nx + idx[0]) * nz + idx[2]
    ucubes, cube_rank = np.unique(key, return_inverse=True)
    if len(ucubes) * 6 != n:
        raise ValueError("not a Kuhn box mesh (6 tets per cube expected)")
    nc = len(ucubes)
    blk = cube_rank // block
    start = blk * block
    width = np.minimum(block, nc - start)
    pos = start * 6 + ftype * width + (cube_rank - start)
    perm = np.empty(n, np.int32)
    perm[pos] = np.arange(n, dtype=np.int32)
    return perm


def spark_adjoint_problem(t_sim: float = 5.0, n_points: int = 64, n_samples: int = 50, ckpt: int = 256,
                          seed: int = 7, perturb: float = 0.1, tune=None):
    """The fitting workload on the h = 40 nm spark box: t_sim ms with the RyR
    trajectory frozen from one gated run (LCC states held), a line scan of
    n_points along x through the cluster centre sampled n_samples times, and a
    target made by the same frozen-channel run with every parameter of theta
    moved by +-perturb (signs drawn from `seed`); `tune`: sweep tunables of
    the context.  Returns the LinescanProblem (its context holds the
    realization and the scan) and a dict with the sizes and the loss at the
    generating parameters."""
    from . import R, adjoint_theta, dff0, read_scan_parameters
    from . import Context
    from .fit import LinescanProblem

    pr = prepare_spark()
    g = pr.grid
    p = pr.params.copy()
    dt = float(p.reals[R["T_STEP"]])
    n_steps = int(round(t_sim / dt))
    ctx = Context(0)
    ctx.tune(**(tune or {}))
    mid = ctx.add_mesh_grid(g, pr.aij, pr.nij, perm=kuhn_block_order(g))
    ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=38376427, stream=0)
    ctx.finalize()
    rs, ls, lvs, Rr = initial_channels(pr)
    ctx.state_set(0, pr.X0)
    ctx.channels_set(0, rs, ls, lvs, Rr)
    ctx.globals_set(0, 0.0, p.reals[R["DEFAULT_CA_NSR"]])
    schedule = ctx.record_schedule(0, n_steps)
    ctx.channels_set(0, schedule[0], ls, lvs, Rr)
    cen = g.centroids()
    c0 = np.array([0.5, 0.5, 0.3])  # the linescan program assumes a centred mesh
    d = read_scan_parameters(None)
    d.axis = 0
    rc = cen[g.ryr_ele].mean(axis=0) - c0
    d.offset1, d.offset2 = float(rc[1]), float(rc[2])
    d.x_min, d.x_max = -0.48, 0.48
    d.res_x_f = (d.x_max - d.x_min) / (n_points + 0.5)  # the point count truncates (x_max - x_min) / res
    d.res_x_ca = 0.01
    d.domain[0], d.domain[1], d.domain[2] = 1.0, 1.0, 0.6
    d.pad_min, d.pad_max, d.pad_res = -1.5, 1.5, 0.1
    sid, nf, _ = ctx.add_scan(mid, cen - c0, d)
    samples = np.linspace(n_steps // n_samples, n_steps, n_samples).astype(np.int64)
    idx = adjoint_theta()
    sgn = np.where(np.random.default_rng(seed).random(len(idx)) < 0.5, -1.0, 1.0)
    Pt = p.reals.copy()
    Pt[idx] *= 1 + perturb * sgn
    # target: dF/F0 of the same frozen-channel run at the moved parameters
    ctx.params_set(0, Pt)
    ctx.state_set(0, pr.X0)
    tgt, prev = [], 0
    for smp in samples:
        ctx.replay(0, schedule[prev:smp])
        prev = int(smp)
        tgt.append(dff0(ctx.scan_sample(0, sid)[2], d.f0))
    tgt = np.array(tgt)
    ctx.state_set(0, pr.X0)
    ctx.channels_set(0, schedule[0], ls, lvs, Rr)
    L0 = ctx.loss_grad(0, sid, schedule, samples, tgt, ckpt, grad=False)
    ctx.params_set(0, p.reals)
    prob = LinescanProblem(ctx, sid, schedule, samples, tgt, p, pr.X0, ckpt=ckpt)
    return prob, dict(n_steps=n_steps, n_ele=g.n_ele, n_f=nf, theta_true=Pt[idx], loss_at_target=L0,
                      n_openings=int(schedule.sum()))


def dyad_adjoint_batch(n_real: int = 256, n_steps: int = 256, n_points: int = 64, n_samples: int = 8,
                       max_batch: int = 0, h: float = 0.03, seed: int = 7, perturb: float = 0.1, log=None,
                       zero_target: bool = False):
    """The batched fitting workload: n_real realizations on the dyad-box
    geometry at h = 30 nm (927,816 elements), each with its own RyR trajectory
    (Philox stream r) recorded by one gated run of all of them and then
    frozen; a line scan of n_points along x through the cluster centre sampled
    n_samples times; each member's target made by its own frozen-channel run
    with theta moved by +-perturb (signs drawn from `seed`), or zero
    (zero_target: the loss is then the squared dF/F0, and no forward run is
    spent on targets).  Returns the
    context, the scan id, the schedules, the sample steps, the targets, the
    parameters and a dict of sizes."""
    from . import R, adjoint_theta, read_scan_parameters
    from . import Context

    pr = prepare_dyad(h=h)
    g = pr.grid
    p = pr.params.copy()
    ctx = Context(0)
    mid = ctx.add_mesh_grid(g, pr.aij, pr.nij, perm=kuhn_block_order(g))
    for r in range(n_real):
        ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=38376427, stream=r)
    ctx.finalize()
    for r in range(n_real):
        rs, ls, lvs, Rr = initial_channels(pr, seed=r)
        ctx.state_set(r, pr.X0)
        ctx.channels_set(r, rs, ls, lvs, Rr)
        ctx.globals_set(r, 0.0, p.reals[R["DEFAULT_CA_NSR"]])
    if log:
        log(f"  {n_real} realizations of {g.n_ele} elements set up; recording {n_steps} steps of schedules")
    scheds = ctx.record_schedules(n_steps)
    for r in range(n_real):
        rs, ls, lvs, Rr = initial_channels(pr, seed=r)
        ctx.channels_set(r, scheds[r][0], ls, lvs, Rr)
    cen = g.centroids()
    c0 = np.array([1.0, 1.0, 0.5])  # the linescan program assumes a centred mesh
    d = read_scan_parameters(None)
    d.axis = 0
    rc = cen[g.ryr_ele].mean(axis=0) - c0
    d.offset1, d.offset2 = float(rc[1]), float(rc[2])
    d.x_min, d.x_max = -0.96, 0.96
    d.res_x_f = (d.x_max - d.x_min) / (n_points + 0.5)  # the point count truncates (x_max - x_min) / res
    d.res_x_ca = 0.01
    d.domain[0], d.domain[1], d.domain[2] = 2.0, 2.0, 1.0
    d.pad_min, d.pad_max, d.pad_res = -2.5, 2.5, 0.1
    sid, nf, _ = ctx.add_scan(mid, cen - c0, d)
    samples = np.linspace(n_steps // n_samples, n_steps, n_samples).astype(np.int64)
    idx = adjoint_theta()
    sgn = np.where(np.random.default_rng(seed).random(len(idx)) < 0.5, -1.0, 1.0)
    Pt = p.reals.copy()
    Pt[idx] *= 1 + perturb * sgn
    zero = np.zeros((n_real, n_samples, nf))
    if zero_target:
        tgt = zero
    else:
        for r in range(n_real):
            ctx.params_set(r, Pt)
        if log:
            log("  targets")
        _, tgt = ctx.loss_grad_batch(range(n_real), sid, scheds, samples, zero, grad=False, scans=True,
                                     max_batch=max_batch)
        for r in range(n_real):
            ctx.params_set(r, p.reals)
    return ctx, sid, scheds, samples, tgt, p, dict(n_ele=g.n_ele, n_f=nf, dt=float(p.reals[R["T_STEP"]]),
                                                   theta_true=Pt[idx])
