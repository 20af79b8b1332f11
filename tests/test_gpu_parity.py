"""HIP path vs the oracle, through the C ABI (needs an MI355X).

Tolerance: rel <= 1e-10 for FP64 fields (the deterministic-mode bar).  The sweep is
built without FP contraction and in the reference's operation order, so most
runs are bit-identical; the bound leaves room for the one-ulp differences of
device exp/pow/log against glibc in the LCC flux and the gating rates.
Channel trajectories (integer states) must match exactly.
"""
import os

import numpy as np
import pytest

import srs_amd as S
from oracle.oracle import PHILOX_KEY1, RNG_GLIBC, RNG_PHILOX, Fields, Oracle, grid_dict
from srs_amd import workloads as W
from tests.helpers import fast_gating, gpu_context, initial_fields, oracle_loop, output_due, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _tiny(**kw):
    return W.prepare(W.tiny_box_spec(**kw), W.spark_params())


@pytest.mark.parametrize("tt", [False, True])
def test_replay_schedule_matches_oracle(tt):
    """Deterministic mode: RyR states replayed from a recorded schedule, gating off."""
    pr = _tiny(tt=tt)
    p = pr.params.copy().seti(FLAG_NO_GATING=1)
    F = initial_fields(pr)
    nsteps = 300
    sch = (np.random.default_rng(3).random((nsteps, pr.grid.n_ryr)) < 0.5).astype(np.int32)
    print(f"replay tt={tt} n_ele={pr.grid.n_ele}", flush=True)
    with gpu_context([pr], [F], params=[p]) as ctx:
        ctx.replay(0, sch)
        Xg = ctx.state_get(0)
    Fo = Oracle().steps(grid_dict(pr.grid, pr.aij, pr.nij), p.reals, F.copy(), nsteps, p.reals[1], gating=False,
                        schedule=sch)
    assert rel_err(Xg, Fo.X) <= TOL


@pytest.mark.parametrize("order", [S.ORDER_AS_GIVEN, S.ORDER_RCM, "block"])
def test_gating_trajectory_matches_oracle(order):
    """Stochastic mode with the same Philox stream: identical channel trajectories."""
    pr = _tiny(tt=True)
    p = fast_gating(pr.params)
    F = initial_fields(pr, seed=5)
    n = 600
    print(f"gating order={order}", flush=True)
    with gpu_context([pr], [F], params=[p], seeds=[77], order=order) as ctx:
        ctx.explicit_step(n)
        Xg = ctx.state_get(0)
        rs, ls, lvs, Rg = ctx.channels_get(0)
    Fo = Oracle().steps(grid_dict(pr.grid, pr.aij, pr.nij), p.reals, F.copy(), n, p.reals[1], gating=True,
                        rng_kind=RNG_PHILOX, key=77, stream=0)
    assert np.array_equal(rs, Fo.ryr_state)
    assert np.array_equal(ls, Fo.lcc_state) and np.array_equal(lvs, Fo.lccv_state)
    assert rel_err(Rg, Fo.R) <= 1e-9
    assert rel_err(Xg, Fo.X) <= TOL


def test_coupled_ryr_gating_is_sequential():
    """A* != 0: later RyRs see earlier flips (cpu/RunSimulation.cpp:661-666)."""
    pr = W.prepare(W.tiny_box_spec(nx=6, ny=6, ryr=3), W.spark_params())
    p = fast_gating(pr.params).setr(RYR_A_STAR=0.8, RYR_K_MINUS=60.0, RYR_K_PLUS=2.0)
    F = initial_fields(pr, ryr_open=(0, 4, 8), seed=9)
    n = 500
    with gpu_context([pr], [F], params=[p], seeds=[5]) as ctx:
        ctx.explicit_step(n)
        Xg = ctx.state_get(0)
        rs, _, _, Rg = ctx.channels_get(0)
    Fo = Oracle().steps(grid_dict(pr.grid, pr.aij, pr.nij), p.reals, F.copy(), n, p.reals[1], gating=True,
                        rng_kind=RNG_PHILOX, key=5, stream=0)
    assert np.array_equal(rs, Fo.ryr_state)
    assert rel_err(Xg, Fo.X) <= TOL


def test_graph_replay_equals_launches():
    """300 steps as 2 captured 128-step blocks and a 44-step block (also
    after capturing them ahead with prepare) equal 300 single launches; calls
    of 7 steps (no graph) and 20 steps (one block of 20) too."""
    pr = _tiny(tt=True)
    p = fast_gating(pr.params)
    F = initial_fields(pr, seed=2)
    out = []
    for flags, pre, split in ((S.STEP_PLAIN, False, (300,)), (S.STEP_GRAPH, False, (300,)),
                              (S.STEP_GRAPH, True, (300,)), (S.STEP_GRAPH, False, (7, 20, 273))):
        with gpu_context([pr], [F], params=[p], seeds=[3]) as ctx:
            if pre:
                ctx.prepare(300, S.STEP_GRAPH)
            for n in split:
                ctx.explicit_step(n, flags)
            out.append((ctx.state_get(0), ctx.channels_get(0)[0]))
    for X, rs in out[1:]:
        assert np.array_equal(out[0][0], X) and np.array_equal(out[0][1], rs)


@pytest.mark.parametrize("lds", [0, 1])
def test_batch_equals_single_runs(lds):
    """Realizations of different meshes and parameters in one launch reproduce
    their single-realization runs bit for bit."""
    preps = [_tiny(tt=False), _tiny(tt=True), W.prepare(W.tiny_box_spec(nx=5, ny=3), W.spark_params())]
    params = [fast_gating(pr.params) for pr in preps]
    params[2].setr(C_0_CA_JSR=700.0, DEFAULT_CA_NSR=700.0)
    fields = [initial_fields(pr, seed=k) for k, pr in enumerate(preps)]
    seeds = [11, 12, 13]
    with gpu_context(preps, fields, params=params, seeds=seeds) as ctx:
        ctx.tune(lds=lds)
        ctx.explicit_step(250, S.STEP_GRAPH)
        batch = [(ctx.state_get(k), ctx.channels_get(k)[0]) for k in range(3)]
    for k, pr in enumerate(preps):
        Fo = Oracle().steps(grid_dict(pr.grid, pr.aij, pr.nij), params[k].reals, fields[k].copy(), 250,
                            params[k].reals[1], gating=True, rng_kind=RNG_PHILOX, key=seeds[k], stream=k)
        assert np.array_equal(batch[k][1], Fo.ryr_state)
        assert rel_err(batch[k][0], Fo.X) <= TOL


@pytest.mark.parametrize("group", [2, 4])
def test_realization_groups_equal_single_runs(group):
    """Realizations of one mesh interleaved in groups (operator rows loaded
    once per group): each equals its oracle run; a realization marked done
    halfway keeps its state while its group partners go on."""
    pr = _tiny(tt=True)
    n_real = 5
    params = [fast_gating(pr.params) for _ in range(n_real)]
    for k, p in enumerate(params):
        p.setr(C_0_CA_JSR=600.0 + 100 * k, DEFAULT_CA_NSR=600.0 + 100 * k)
    fields = [initial_fields(pr, ryr_open=(k % pr.grid.n_ryr,), seed=20 + k) for k in range(n_real)]
    seeds = [40 + k for k in range(n_real)]
    with gpu_context([pr] * n_real, fields, params=params, seeds=seeds, share_mesh=True, group=group) as ctx:
        ctx.explicit_step(120, S.STEP_GRAPH)
        X1 = ctx.state_get(1)
        ctx.loop_set(1, step=120, done=1)
        ctx.explicit_step(130, S.STEP_GRAPH)
        got = [(ctx.state_get(k), ctx.channels_get(k)[0]) for k in range(n_real)]
    assert np.array_equal(got[1][0], X1)
    for k in range(n_real):
        n = 120 if k == 1 else 250
        Fo = Oracle().steps(grid_dict(pr.grid, pr.aij, pr.nij), params[k].reals, fields[k].copy(), n,
                            params[k].reals[1], gating=True, rng_kind=RNG_PHILOX, key=seeds[k], stream=k)
        assert np.array_equal(got[k][1], Fo.ryr_state), k
        assert rel_err(got[k][0], Fo.X) <= TOL, k


@pytest.mark.parametrize("gated", [0, 1])
def test_protocol_loop_records_match_oracle_loop(gated):
    """The device-side RunSimulation loop body (clamp, RyR hold, states output,
    while-condition, final record) against an oracle-driven loop, with the
    loop body in the gating launch after each sweep or in workgroup 0 of the
    next sweep launch (the realization finishes inside a block)."""
    pr = _tiny(tt=True)
    p = fast_gating(pr.params)
    dt = p.reals[1]
    p.setr(T_FINAL=900 * dt, OUTPUT_STATES_INTERVAL=37 * dt, INIT_RYR_HOLD=100 * dt, T_CLAMP=450 * dt,
           V_CLAMP=-10.0, DEFAULT_VM=-80.0)
    p.seti(INIT_RYR=1, FLAG_V_CLAMP=1, FLAG_RYR_OUT=1)
    F = initial_fields(pr, ryr_open=(), vm=-80.0, seed=4)
    O = Oracle()
    recs_o, t_end, nsteps = oracle_loop(O, pr, p, F.copy(), key=21, stream=0)
    # device: apply the t = 0 iteration on the host, as srs_run_cru3d does
    Fd = F.copy()
    Fd.gv[0] = p.reals[S.R["V_CLAMP"]]
    Fd.ryr_state[1] = 1
    with gpu_context([pr], [Fd], params=[p], seeds=[21], record_capacity=256) as ctx:
        ctx.tune(gated_sweep=gated)
        ctx.loop_set(0, n_open=int(F.ryr_state.sum()))
        ctx.explicit_step(nsteps + 5, S.STEP_PROTOCOL | S.STEP_GRAPH)
        recs_d = ctx.records(0)
        ls = ctx.loop_get(0)
    assert ls["done"] == 1 and ls["step"] == nsteps and ls["t"] == t_end
    assert recs_d.shape[0] == len(recs_o)
    np.testing.assert_array_equal(recs_d, np.array(recs_o))



@pytest.mark.parametrize("gated", [0, 1])
def test_lcc_trigger_protocol_matches_oracle_loop(gated):
    """The LCC trigger protocol: an LCC held open from t = 0 is closed by the
    loop body once t reaches T_LCC_Close (cpu/RunSimulation.cpp:459-463,
    525-528), with the loop body in separate gating launches or in the gated
    sweep."""
    pr = _tiny(tt=True)
    p = fast_gating(pr.params)
    dt = p.reals[1]
    p.setr(T_FINAL=400 * dt, OUTPUT_STATES_INTERVAL=13 * dt, V_CLAMP=0.0, DEFAULT_VM=0.0)
    p.seti(INIT_RYR=-1, FLAG_V_CLAMP=0, FLAG_RYR_OUT=1, FLAG_LCC=1, LCC_INDEX=0)
    F = initial_fields(pr, ryr_open=(), vm=0.0, lcc_state=1, seed=6)
    F.lcc_state[0] = 6
    tc = 150.5 * dt
    O = Oracle()
    recs_o, t_end, nsteps = oracle_loop(O, pr, p, F.copy(), key=31, stream=0, t_lcc_close=tc)
    with gpu_context([pr], [F.copy()], params=[p], seeds=[31], record_capacity=256) as ctx:
        ctx.tune(gated_sweep=gated)
        ctx.loop_set(0, lcc_open=1, t_lcc_close=tc)
        ctx.explicit_step(nsteps + 5, S.STEP_PROTOCOL | S.STEP_GRAPH)
        recs_d = ctx.records(0)
        ls = ctx.loop_get(0)
    assert ls["done"] == 1 and ls["lcc_open"] == 0 and ls["step"] == nsteps and ls["t"] == t_end
    assert recs_d.shape[0] == len(recs_o)
    np.testing.assert_array_equal(recs_d, np.array(recs_o))
    assert nsteps > 151


@pytest.mark.parametrize("gated", [0, 1])
@pytest.mark.parametrize("c_min", [0.0, 0.5])
def test_fidelity_stop_matches_oracle_loop(gated, c_min):
    """The fidelity protocol's while-condition (cpu/RunSimulation.cpp:522):
    the run stops at the first states output with N_RYR_FIDELITY RyRs open
    (c_min 0) or once no RyR is open and the largest cleft Ca has fallen to
    C_MIN (c_min 0.5, above the resting Ca), with N_open and the maximum
    refreshed only at outputs."""
    pr = _tiny(tt=True)
    p = pr.params.copy()
    dt = p.reals[1]
    c0 = float(pr.X0[pr.grid.ryr_ele[0]])
    p.setr(RYR_K_PLUS=1.0 / (40 * dt * c0 ** p.reals[S.R["RYR_ETA"]]), RYR_K_MINUS=1.0 / (90 * dt),
           T_FINAL=3000 * dt, OUTPUT_STATES_INTERVAL=11 * dt, C_MIN=c_min, V_CLAMP=0.0, DEFAULT_VM=0.0)
    p.seti(INIT_RYR=-1, FLAG_V_CLAMP=0, FLAG_RYR_OUT=1, FLAG_FIDELITY_SIM=1, N_RYR_FIDELITY=2)
    F = initial_fields(pr, ryr_open=(0,), vm=0.0, lcc_state=1, seed=12)
    O = Oracle()
    recs_o, t_end, nsteps = oracle_loop(O, pr, p, F.copy(), key=41, stream=0)
    with gpu_context([pr], [F.copy()], params=[p], seeds=[41], record_capacity=512) as ctx:
        ctx.tune(gated_sweep=gated)
        ctx.loop_set(0, n_open=int(F.ryr_state.sum()))
        ctx.explicit_step(nsteps + 5, S.STEP_PROTOCOL | S.STEP_GRAPH)
        recs_d = ctx.records(0)
        ls = ctx.loop_get(0)
    assert ls["done"] == 1 and ls["step"] == nsteps and ls["t"] == t_end
    assert recs_d.shape[0] == len(recs_o)
    np.testing.assert_array_equal(recs_d, np.array(recs_o))
    assert t_end < p.reals[S.R["T_FINAL"]]  # the fidelity condition, not T_FINAL, ended it

@pytest.mark.parametrize("lds,chunk", [(0, 0), (1, 0), (1, 37)])
@pytest.mark.parametrize("case", ["opens", "t_final"])
def test_ffwd_gating_phase_matches_oracle_loop(case, lds, chunk):
    """flag_ffwd: gating-only steps on the unchanged fields until a channel
    opens (or T_FINAL passes), then the main loop from that time
    (cpu/RunSimulation.cpp:469-519).  Both sweep kernels skip the
    realization while it is in the ffwd phase and after it is done.  In
    calls of 37 steps the calls after the ffwd phase run through the gated
    sweep, the first ones through separate gating launches."""
    pr = _tiny(tt=True)
    p = pr.params.copy()
    dt = p.reals[1]
    c0 = float(pr.X0[pr.grid.ryr_ele[0]])
    # about one RyR opening per 60 steps over the whole cluster
    kp = 1.0 / (60 * dt * pr.grid.n_ryr * c0 ** p.reals[S.R["RYR_ETA"]]) if case == "opens" else 0.0
    p.setr(RYR_K_PLUS=kp, RYR_K_MINUS=20.0, T_FINAL=(400 if case == "opens" else 45) * dt,
           OUTPUT_STATES_INTERVAL=7 * dt, T_CLAMP=1e9, V_CLAMP=0.0, DEFAULT_VM=-80.0)
    p.seti(INIT_RYR=-1, FLAG_V_CLAMP=1, GATING_FFWD=1, FLAG_RYR_OUT=1)
    F = initial_fields(pr, ryr_open=(), vm=-80.0, lcc_state=1, seed=9)
    O = Oracle()
    recs_o, t_end, nsteps = oracle_loop(O, pr, p, F.copy(), key=33, stream=0)
    with gpu_context([pr], [F.copy()], params=[p], seeds=[33], record_capacity=256) as ctx:
        if lds:
            ctx.tune(lds=1, occupancy=6)
        ctx.loop_set(0, n_open=0, ffwd=1)
        if chunk:
            for _ in range((nsteps + 5 + chunk - 1) // chunk):
                ctx.explicit_step(chunk, S.STEP_PROTOCOL | S.STEP_GRAPH)
        else:
            ctx.explicit_step(nsteps + 5, S.STEP_PROTOCOL | S.STEP_GRAPH)
        recs_d = ctx.records(0)
        ls = ctx.loop_get(0)
        X = ctx.state_get(0)
    assert ls["done"] == 1 and ls["ffwd"] == 0 and ls["step"] == nsteps and ls["t"] == t_end
    assert recs_d.shape[0] == len(recs_o)
    np.testing.assert_array_equal(recs_d, np.array(recs_o))
    if case == "t_final":
        # nothing opened: the fields never moved
        np.testing.assert_array_equal(X, F.X)
    else:
        assert nsteps > 20 and np.any(X != F.X)


def _spark_stats(recs, nr):
    """Per-realization statistics of a states-record sequence (one record per
    step): first closing step of RyR 0 (its open time), integrated open RyR
    count, peak mean cleft Ca at the RyRs (amplitude proxy), number of RyR
    openings."""
    st = recs[:, 7:7 + nr] > 0.5
    closed0 = np.nonzero(~st[:, 0])[0]
    first_close = int(closed0[0]) + 1 if closed0.size else recs.shape[0] + 1
    openings = int(np.sum(st[1:] & ~st[:-1]))
    return first_close, float(recs[:, 2].sum()), float(recs[:, 6].max()), openings


def test_stochastic_gating_matches_reference_in_distribution():
    """Philox gating on the device against the reference's own generator
    (glibc rand(), as cpu/RunSimulation.cpp:208-209, 382-397 draw it; the
    oracle's glibc mode is bit-exact with the reference step): two-sample KS
    tests on open time, integrated open count, spark amplitude proxy and
    number of openings, p > 0.01, plus the open-time law of a lone open RyR."""
    from scipy import stats

    import ctypes

    n_real, n_steps = 160, 500
    pr = _tiny()
    p = pr.params.copy()
    dt = p.reals[1]
    # mean open time 120 steps; a closed RyR opens about every 300 steps at 0.2 uM
    p.setr(RYR_K_MINUS=1 / (120 * dt), RYR_K_PLUS=1 / (300 * dt * 0.2 ** 2.1))
    p.setr(T_FINAL=(n_steps - 0.5) * dt, OUTPUT_STATES_INTERVAL=dt * (1 - 1e-9), V_CLAMP=0.0, DEFAULT_VM=0.0)
    p.seti(INIT_RYR=-1, FLAG_V_CLAMP=0, FLAG_RYR_OUT=1)
    nr = pr.grid.n_ryr
    fields = [initial_fields(pr, ryr_open=(0,), vm=0.0, lcc_state=1, seed=100 + k) for k in range(n_real)]
    with gpu_context([pr] * n_real, [F.copy() for F in fields], params=[p] * n_real, seeds=[77] * n_real,
                     record_capacity=n_steps + 8, share_mesh=True) as ctx:
        for k in range(n_real):
            ctx.loop_set(k, n_open=1)
        ctx.explicit_step(n_steps + 4, S.STEP_PROTOCOL | S.STEP_GRAPH)
        gpu = [_spark_stats(ctx.records(k), nr) for k in range(n_real)]
    O = Oracle()
    libc = ctypes.CDLL(None)
    ref = []
    for k in range(n_real):
        libc.srand(ctypes.c_uint(93843 + 38376427 + 83 * k))
        recs, _, _ = oracle_loop(O, pr, p, fields[k].copy(), key=0, stream=0, rng_kind=RNG_GLIBC)
        ref.append(_spark_stats(np.array(recs), nr))
    gpu, ref = np.array(gpu), np.array(ref)
    for q, name in enumerate(("open time", "open RyR-steps", "amplitude", "openings")):
        pv = stats.ks_2samp(gpu[:, q], ref[:, q]).pvalue
        assert pv > 0.01, (name, pv, gpu[:, q].mean(), ref[:, q].mean())
    # a lone open RyR closes after ceil(E / (k_minus dt)) steps, E ~ Exp(1):
    # geometric, here conditioned on closing within the run
    q = 1.0 - np.exp(-p.reals[S.R["RYR_K_MINUS"]] * dt)
    n = np.arange(1, n_steps + 1)
    pmf = (1 - q) ** (n - 1) * q
    pmf /= pmf.sum()
    mean_law = float(np.sum(n * pmf))
    sd_law = float(np.sqrt(np.sum((n - mean_law) ** 2 * pmf)))
    for arr in (gpu, ref):
        ft = arr[:, 0][arr[:, 0] <= n_steps]
        assert abs(ft.mean() - mean_law) < 4 * sd_law / np.sqrt(ft.size)
    assert ref[:, 3].mean() > 0.5  # the cluster does recruit openings


@pytest.mark.parametrize("axis,psf", [(0, "default"), (2, "default"), (0, "narrow"), (2, "fine"), (0, "edge"),
                                      (0, "far")])
def test_linescan_matches_oracle(axis, psf):
    """The device linescan (nearest-cell Ca and CaF, PSF-blurred CaF with the
    padding) on the 56,250-element spark box against the oracle's restatement
    of linescan/main.cpp.  Nearest values are copies (bit-exact); the blurred
    sum is summed in another order, and its weights walk along the line by
    products between closed-form anchors (rel <= 1e-12).  "narrow": a PSF of
    s = 10 nm, too narrow for the walk, so every weight is the closed form;
    "fine": 5 nm points (200 per um) under the default PSF, the longest
    walks; "edge": the line 0.8 um above the mesh, so only its top cells
    reach it and most wavefronts reach no point; "far": more than 1 um from
    every cell, so the blurred sum is the padding's alone."""
    pr = W.prepare_spark()
    g = pr.grid
    n = g.n_ele
    cen = g.centroids() - np.array([0.5, 0.5, 0.3])  # the linescan program assumes a centred mesh
    rng = np.random.default_rng(11)
    X = pr.X0.copy()
    X[:n] *= 1 + 50 * rng.random(n)
    X[5 * n:] *= 1 + 3 * rng.random(n)
    d = S.read_scan_parameters(None)
    d.axis = axis
    d.domain[0], d.domain[1], d.domain[2] = 1.0, 1.0, 0.6
    d.pad_min, d.pad_max, d.pad_res = -1.5, 1.5, 0.1
    if axis == 0:
        d.x_min, d.x_max, d.res_x_f, d.res_x_ca = -0.7, 0.7, 0.025, 0.01
        d.offset1, d.offset2 = 0.02, -0.2925
    else:
        d.x_min, d.x_max, d.res_x_f, d.res_x_ca = -0.45, 0.45, 0.02, 0.005
        d.offset1, d.offset2 = 0.02, 0.01
    if psf == "narrow":
        for k in range(3):
            d.var[k] = 1e-4
    elif psf == "fine":
        d.res_x_f = 0.005
    elif psf == "edge":
        d.offset2 = 1.1
    elif psf == "far":
        d.offset2 = 1.5
    with S.Context(0) as ctx:
        mid = ctx.add_mesh_grid(g, pr.aij, pr.nij, perm=W.kuhn_block_order(g))
        ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, pr.params, seed=1, stream=0)
        ctx.finalize()
        ctx.state_set(0, X)
        sid, nf, nca = ctx.add_scan(mid, cen, d)
        ca, caf, cafb = ctx.scan_sample(0, sid)
    oca, ocaf, ocafb = Oracle().linescan(d, cen, g.V0, X[:n], X[5 * n:])
    assert nf == len(ocafb) > 40 and nca == len(oca) > 100
    np.testing.assert_array_equal(ca, oca)
    np.testing.assert_array_equal(caf, ocaf)
    assert rel_err(cafb, ocafb) <= 1e-12
    dff = S.dff0(cafb, d.f0)
    assert np.all(np.isfinite(dff))
    if psf not in ("edge", "far"):  # off the mesh the nearest points are padding
        assert np.any(caf != d.f0) and np.any(ca != d.c0) and dff.max() > 0


def _write_params(path, entries):
    body = "".join(f"  <parameter><symbol>{k}</symbol><value>{v}</value></parameter>\n" for k, v in entries.items())
    with open(path, "w") as f:
        f.write(f'<?xml version="1.0" encoding="UTF-8"?>\n<simulation>\n{body}</simulation>\n')


@pytest.mark.parametrize("ffwd", [False, True])
def test_cru3d_run_matches_oracle_loop(tmp_path, ffwd):
    """The cru3d program path end to end (parameter XML, mesh files, initial
    channel protocol, Philox residuals, device loop, states file) against
    the oracle-driven RunSimulation replica started from the same state."""
    from types import SimpleNamespace

    base = str(tmp_path / "mesh")
    S.box_mesh_write(W.tiny_box_spec(tt=True), base)
    g, _ = S.grid_read(base)
    aij, nij = S.assemble_operator(g)
    dt = W.stable_dt(aij, 0.25)
    c0 = 0.1
    kp = 1.0 / (60 * dt * g.n_ryr * c0 ** 2.1) if ffwd else 0.5
    entries = dict(t_step=repr(dt * 1e6), t_final=repr(300 * dt), states_interval=repr(9 * dt),
                   RYR_K_PLUS=repr(kp), RYR_K_MINUS="20", flag_use_clock="false", seed_base="1234",
                   flag_output_ryr_states="true", sims_per_proc="1", paramset_title="tiny")
    if ffwd:
        entries.update(protocol="gain", flag_ffwd="true", v_clamp="0", t_clamp="1e9")
    else:
        entries.update(protocol="none", ryr_open_init="1", ryr_open_time=repr(50 * dt))
    xml = str(tmp_path / "p.xml")
    _write_params(xml, entries)
    out = str(tmp_path / "out")
    S.run_cru3d(xml, base, out)
    p = S.read_parameters(xml)
    assert p.reals[1] == float(repr(dt * 1e6)) * 1e-6
    recs_d = np.fromfile(os.path.join(out, "states_0_0.out"), np.float32).reshape(-1, 7 + g.n_ryr)

    O = Oracle()
    P = p.reals
    seed = (93843 + 1234) & 0xFFFFFFFF
    nr, nl = g.n_ryr, g.n_lcc
    rs = np.zeros(nr, np.int32)
    if not ffwd:
        rs[1] = 1
    R = np.array([np.log(O.open01(O.philox([0, i, 0, 1], [seed, PHILOX_KEY1])[0])) for i in range(nr + nl)])
    X0, _ = S.equilibrium_state(g, P)
    F = Fields(X0, rs, np.ones(nl, np.int32), np.full(nl, 2, np.int32), R, [P[S.R["DEFAULT_VM"]],
                                                                          P[S.R["DEFAULT_CA_NSR"]]])
    G = grid_dict(g, aij, nij)
    rec0 = O.record(0.0, G, F, P, ryr_out=True)
    pr = SimpleNamespace(grid=g, aij=aij, nij=nij)
    recs_o, t_end, nsteps = oracle_loop(O, pr, p, F, key=seed, stream=0)
    exp = np.array([rec0] + recs_o)
    assert recs_d.shape == exp.shape
    np.testing.assert_array_equal(recs_d, exp)
    assert nsteps > 100



def test_cru3d_random_initial_ryr_and_gpu_layout(tmp_path):
    """ryr_open_init = -1 without a clamp or LCC trigger: the driver opens a
    rate-weighted random RyR (OpenRandomRyR, cpu/InitFuncs.cpp:1420-1475,
    linear phi) with its Philox draw, and writes the states file in the GPU
    program's layout (Ca_cyt at each RyR after the seven fields,
    gpu/OutputFuncs.cpp:277-284; the -layout gpu option of cru3d_amd).  The
    oracle's OpenRandomRyR with the same uniform picks the same RyR, and the
    records equal the oracle loop's in that layout."""
    import ctypes as C
    import subprocess
    from types import SimpleNamespace

    base = str(tmp_path / "mesh")
    S.box_mesh_write(W.tiny_box_spec(nx=6, ny=6, tt=True, ryr=3), base)
    g, _ = S.grid_read(base)
    aij, nij = S.assemble_operator(g)
    dt = W.stable_dt(aij, 0.25)
    entries = dict(t_step=repr(dt * 1e6), t_final=repr(250 * dt), states_interval=repr(9 * dt),
                   RYR_K_PLUS="0.5", RYR_K_MINUS="20", flag_use_clock="false", seed_base="777",
                   flag_output_ryr_states="true", sims_per_proc="2", paramset_title="tiny", protocol="none",
                   ryr_open_init="-1")
    xml = str(tmp_path / "p.xml")
    _write_params(xml, entries)
    out = str(tmp_path / "out")
    S.run_cru3d(xml, base, out, gpu_layout=True)
    out2 = str(tmp_path / "out2")
    exe = os.path.join(os.path.dirname(S.__file__), "cru3d_amd")
    subprocess.run([exe, "-param", xml, "-mesh", base, "-out", out2, "-layout", "gpu"], check=True,
                   stdout=subprocess.DEVNULL)
    p = S.read_parameters(xml)
    P = p.reals
    O = Oracle()
    nr, nl = g.n_ryr, g.n_lcc
    X0, _ = S.equilibrium_state(g, P)
    G = grid_dict(g, aij, nij)
    pr = SimpleNamespace(grid=g, aij=aij, nij=nij)
    picked = []
    for run in range(2):
        recs_d = np.fromfile(os.path.join(out, f"states_0_{run}.out"), np.float32).reshape(-1, 7 + 2 * nr)
        recs_e = np.fromfile(os.path.join(out2, f"states_0_{run}.out"), np.float32)
        np.testing.assert_array_equal(recs_d.ravel(), recs_e)
        seed = (93843 + 777 + 83 * run) & 0xFFFFFFFF  # seed_B's default is 83
        rs = np.zeros(nr, np.int32)
        F = Fields(X0.copy(), rs, np.ones(nl, np.int32), np.full(nl, 2, np.int32), np.zeros(nr + nl),
                   [P[S.R["DEFAULT_VM"]], P[S.R["DEFAULT_CA_NSR"]]])
        u = O.open01(O.philox([0, 0, run, 3], [seed, PHILOX_KEY1])[0])
        ch = O._chan(G, F)
        k = O.L.orc_open_random_ryr(F.X.ctypes.data_as(C.POINTER(C.c_double)), g.n_ele,
                                    P.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ch), u)
        assert F.ryr_state.sum() == 1 and F.ryr_state[k] == 1
        picked.append(k)
        F.R[:] = [np.log(O.open01(O.philox([0, i, run, 1], [seed, PHILOX_KEY1])[0])) for i in range(nr + nl)]
        rec0 = O.record(0.0, G, F, P, ryr_out=True, gpu_layout=True)
        recs_o, _, nsteps = oracle_loop(O, pr, p, F, key=seed, stream=run, gpu_layout=True)
        exp = np.array([rec0] + recs_o)
        assert recs_d.shape == exp.shape and nsteps > 100
        np.testing.assert_array_equal(recs_d, exp)
        np.testing.assert_array_equal(recs_d[0, 7:7 + nr], X0[g.ryr_ele].astype(np.float32))
    assert recs_d[0, 7 + nr:].sum() == 1  # one RyR open at t = 0

def _vtk_text(g, X, ryr_state, run, rank, n, t):
    """WriteToGridFiles restated (cpu/OutputFuncs.cpp:42-220): legacy-ASCII
    VTK, points and cells, then cell data C_CA, C_DYE (printf %f), B_RYR (0
    none, 1 closed, 2 open), B_JSR and B_LCC."""
    N = g.n_ele
    xyz = g.node_xyz.reshape(-1, 3)
    en = g.ele_nodes.reshape(-1, 4)
    out = ["# vtk DataFile Version 3.0\n", "Grid data for run %d-%d-%d at t = %f\n" % (run, rank, n, t), "ASCII\n\n",
           "DATASET UNSTRUCTURED_GRID\n", "POINTS %d float\n" % len(xyz)]
    out += ["%f %f %f\n" % tuple(p) for p in xyz]
    out.append("\nCELLS %d %d\n" % (N, 5 * N))
    out += ["4 %d %d %d %d\n" % tuple(e) for e in en]
    out.append("\nCELL_TYPES %d\n" % N)
    out += ["10\n"] * N
    out.append("\nCELL_DATA %d\n" % N)
    out.append("SCALARS C_CA float\nLOOKUP_TABLE default\n")
    out += ["%f\n" % v for v in X[:N]]
    out.append("SCALARS C_DYE float\nLOOKUP_TABLE default\n")
    out += ["%f\n" % v for v in X[5 * N:6 * N]]
    b = np.zeros(N, int)
    for i, e in enumerate(g.ryr_ele):
        b[e] = 2 if ryr_state[i] else 1
    out.append("SCALARS B_RYR int\nLOOKUP_TABLE default\n")
    out += ["%d\n" % v for v in b]
    out.append("SCALARS B_JSR int\nLOOKUP_TABLE default\n")
    out += ["%d\n" % (1 if d == 1 else 0) for d in g.domain]
    b = np.zeros(N, int)
    b[g.lcc_ele] = 1
    out.append("SCALARS B_LCC int\nLOOKUP_TABLE default\n")
    out += ["%d\n" % v for v in b]
    return "".join(out)


def test_cru3d_grid_files_match_restated_writer(tmp_path):
    """flag_output_grid: the driver writes grid #0 before the loop, one at
    every floor crossing of grid_interval (at the top of the loop, after the
    protocols, cpu/RunSimulation.cpp:545-550) and the final one; each file
    equals the restated writer applied to the oracle loop's state at that
    time, byte for byte."""
    from types import SimpleNamespace

    base = str(tmp_path / "mesh")
    S.box_mesh_write(W.tiny_box_spec(tt=True), base)
    g, _ = S.grid_read(base)
    aij, nij = S.assemble_operator(g)
    dt = W.stable_dt(aij, 0.25)
    entries = dict(t_step=repr(dt * 1e6), t_final=repr(130 * dt), states_interval=repr(9 * dt),
                   grid_interval=repr(41 * dt), flag_output_grid="true", RYR_K_PLUS="0.5", RYR_K_MINUS="20",
                   flag_use_clock="false", seed_base="55", sims_per_proc="1", paramset_title="tiny",
                   protocol="none", ryr_open_init="0", ryr_open_time=repr(30 * dt))
    xml = str(tmp_path / "p.xml")
    _write_params(xml, entries)
    out = str(tmp_path / "out")
    S.run_cru3d(xml, base, out)
    p = S.read_parameters(xml)
    P = p.reals
    O = Oracle()
    nr, nl = g.n_ryr, g.n_lcc
    seed = (93843 + 55) & 0xFFFFFFFF
    rs = np.zeros(nr, np.int32)
    rs[0] = 1
    R = np.array([np.log(O.open01(O.philox([0, i, 0, 1], [seed, PHILOX_KEY1])[0])) for i in range(nr + nl)])
    X0, _ = S.equilibrium_state(g, P)
    F = Fields(X0, rs, np.ones(nl, np.int32), np.full(nl, 2, np.int32), R, [P[S.R["DEFAULT_VM"]],
                                                                        P[S.R["DEFAULT_CA_NSR"]]])
    want = [_vtk_text(g, F.X, F.ryr_state, 0, 0, 0, 0.0)]
    giv = P[S.R["OUTPUT_GRID_INTERVAL"]]

    def on_top(t, dt_last, Fc):
        if output_due(t, dt_last, giv):
            want.append(_vtk_text(g, Fc.X, Fc.ryr_state, 0, 0, len(want), t))

    pr = SimpleNamespace(grid=g, aij=aij, nij=nij)
    _, t_end, nsteps = oracle_loop(O, pr, p, F, key=seed, stream=0, on_top=on_top)
    want.append(_vtk_text(g, F.X, F.ryr_state, 0, 0, len(want), t_end))
    assert len(want) == 5 and nsteps > 120
    for n, text in enumerate(want):
        with open(os.path.join(out, f"grid_0_0_{n}.vtk")) as f:
            got = f.read()
        assert got == text, n
    assert not os.path.exists(os.path.join(out, f"grid_0_0_{len(want)}.vtk"))

def test_cru3d_ensemble_runs_match_oracle_loops(tmp_path):
    """sims_per_proc > 1: the driver runs the ensemble's realizations together
    in one context (cpu/main.cpp:209-258 runs them one after another); run k
    has seed 93843 + seed_base + k seed_B and its own Philox stream, and each
    states file equals the oracle loop of that run alone."""
    from types import SimpleNamespace

    base = str(tmp_path / "mesh")
    S.box_mesh_write(W.tiny_box_spec(tt=True), base)
    g, _ = S.grid_read(base)
    aij, nij = S.assemble_operator(g)
    dt = W.stable_dt(aij, 0.25)
    entries = dict(t_step=repr(dt * 1e6), t_final=repr(200 * dt), states_interval=repr(7 * dt),
                   RYR_K_PLUS="0.5", RYR_K_MINUS="20", flag_use_clock="false", seed_base="4321", seed_B="83",
                   flag_output_ryr_states="true", sims_per_proc="3", paramset_title="tiny", protocol="none",
                   ryr_open_init="1", ryr_open_time=repr(40 * dt))
    xml = str(tmp_path / "p.xml")
    _write_params(xml, entries)
    out = str(tmp_path / "out")
    S.run_cru3d(xml, base, out)
    p = S.read_parameters(xml)
    P = p.reals
    O = Oracle()
    nr, nl = g.n_ryr, g.n_lcc
    X0, _ = S.equilibrium_state(g, P)
    G = grid_dict(g, aij, nij)
    pr = SimpleNamespace(grid=g, aij=aij, nij=nij)
    finals = []
    for run in range(3):
        recs_d = np.fromfile(os.path.join(out, f"states_0_{run}.out"), np.float32).reshape(-1, 7 + nr)
        seed = (93843 + 4321 + 83 * run) & 0xFFFFFFFF
        rs = np.zeros(nr, np.int32)
        rs[1] = 1
        R = np.array([np.log(O.open01(O.philox([0, i, run, 1], [seed, PHILOX_KEY1])[0])) for i in range(nr + nl)])
        F = Fields(X0.copy(), rs, np.ones(nl, np.int32), np.full(nl, 2, np.int32), R,
                   [P[S.R["DEFAULT_VM"]], P[S.R["DEFAULT_CA_NSR"]]])
        rec0 = O.record(0.0, G, F, P, ryr_out=True)
        recs_o, _, nsteps = oracle_loop(O, pr, p, F, key=seed, stream=run)
        exp = np.array([rec0] + recs_o)
        assert recs_d.shape == exp.shape and nsteps > 100
        np.testing.assert_array_equal(recs_d, exp)
        finals.append(recs_d)
    # the runs' own draws: their RyR trajectories differ
    assert not np.array_equal(finals[0][:, 7:], finals[1][:, 7:])

def test_spark_box_replay_full_mesh():
    """The 56,250-tet spark box, replayed schedule, 200 steps."""
    pr = W.prepare_spark()
    p = pr.params.copy().seti(FLAG_NO_GATING=1)
    F = initial_fields(pr, ryr_open=(24,))
    sch = np.zeros((200, pr.grid.n_ryr), np.int32)
    sch[:, 24] = 1
    sch[100:, 17] = 1
    print(f"spark box n_ele={pr.grid.n_ele}", flush=True)
    for order in (S.ORDER_RCM, "block"):
        with gpu_context([pr], [F], params=[p], order=order) as ctx:
            ctx.replay(0, sch)
            Xg = ctx.state_get(0)
        Fo = Oracle().steps(grid_dict(pr.grid, pr.aij, pr.nij), p.reals, F.copy(), 200, p.reals[1], gating=False,
                            schedule=sch)
        assert rel_err(Xg, Fo.X) <= TOL


def test_no_channels_and_done_realizations():
    """Edge cases: a realization without channels; a realization marked done is not stepped."""
    pr = _tiny(ryr=0, lcc=0)
    F = initial_fields(pr)
    pr2 = _tiny()
    F2 = initial_fields(pr2)
    with gpu_context([pr, pr2], [F, F2]) as ctx:
        ctx.loop_set(1, done=1)
        ctx.explicit_step(50)
        X1 = ctx.state_get(1)
        X0 = ctx.state_get(0)
    assert np.array_equal(X1, F2.X)
    Fo = Oracle().steps(grid_dict(pr.grid, pr.aij, pr.nij), pr.params.reals, F.copy(), 50, pr.params.reals[1],
                        gating=True, rng_kind=RNG_PHILOX, key=1000, stream=0)
    assert rel_err(X0, Fo.X) <= TOL



def test_single_realization_without_channels_and_done():
    """The gated sweep's edge cases on a single realization: no channel
    elements at all (the gating workgroup has nothing to update) against the
    oracle, and a finished realization, which no call steps."""
    pr = _tiny(ryr=0, lcc=0)
    F = initial_fields(pr)
    with gpu_context([pr], [F]) as ctx:
        ctx.explicit_step(20)
        ctx.explicit_step(150, S.STEP_GRAPH)
        X0 = ctx.state_get(0)
        ls = ctx.loop_get(0)
        assert ls["step"] == 170
        ctx.loop_set(0, ls["t"], ls["dt_last"], ls["step"], 1)
        ctx.explicit_step(40, S.STEP_GRAPH)
        ctx.explicit_step(3)
        assert np.array_equal(ctx.state_get(0), X0) and ctx.loop_get(0)["step"] == 170
    Fo = Oracle().steps(grid_dict(pr.grid, pr.aij, pr.nij), pr.params.reals, F.copy(), 170, pr.params.reals[1],
                        gating=True, rng_kind=RNG_PHILOX, key=1000, stream=0)
    assert rel_err(X0, Fo.X) <= TOL

@pytest.mark.parametrize("group", [1, 2])
def test_finished_realizations_leave_the_sweep(group):
    """Protocol runs that end at different times, stepped in chunks: the
    sweep drops the tiles of finished groups, and every realization's records
    and final state equal its own one-realization run; a finished realization
    restarted by the host (longer T_FINAL, done cleared) is stepped again."""
    pr = _tiny(tt=True)
    dt = pr.params.reals[1]
    n_real, chunk = 6, 40
    params, fields = [], []
    for k in range(n_real):
        p = fast_gating(pr.params)
        p.setr(T_FINAL=(60 + 55 * k) * dt, OUTPUT_STATES_INTERVAL=9 * dt, INIT_RYR_HOLD=20 * dt)
        p.seti(INIT_RYR=k % pr.grid.n_ryr, FLAG_RYR_OUT=1)
        params.append(p)
        fields.append(initial_fields(pr, ryr_open=(k % pr.grid.n_ryr,), seed=50 + k))

    def build(ks, grp):
        ctx = S.Context(0)
        ctx.tune(group=grp)
        mid = ctx.add_mesh_grid(pr.grid, pr.aij, pr.nij)
        g = pr.grid
        for k in ks:
            ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, params[k], seed=70 + k, stream=k)
        ctx.finalize(256)
        for q, k in enumerate(ks):
            F = fields[k]
            ctx.state_set(q, F.X)
            ctx.channels_set(q, F.ryr_state, F.lcc_state, F.lccv_state, F.R)
            ctx.globals_set(q, F.gv[0], F.gv[1])
        return ctx

    def run(ctx, n):
        for _ in range(200):
            if all(ctx.loop_get(q)["done"] for q in range(n)):
                return
            ctx.explicit_step(chunk, S.STEP_PROTOCOL | S.STEP_GRAPH)
        raise AssertionError("runs did not finish")

    ctx = build(range(n_real), group)
    run(ctx, n_real)
    got = [(ctx.state_get(k), ctx.records(k), ctx.loop_get(k)) for k in range(n_real)]
    # restart realization 0 for 100 more steps of protocol
    P0 = params[0].reals.copy()
    P0[S.R["T_FINAL"]] += 100 * dt
    ctx.params_set(0, P0)
    ls = got[0][2]
    ctx.loop_set(0, t=ls["t"], dt_last=ls["dt_last"], step=ls["step"], done=0, n_open=ls["n_ryr_open_last"],
                 ca_max=ls["ca_ss_max_last"])
    ctx.explicit_step(chunk, S.STEP_PROTOCOL | S.STEP_GRAPH)
    restarted = ctx.state_get(0)
    ctx.close()
    assert len({g[2]["step"] for g in got}) > 1  # the runs ended at different steps
    for k in range(n_real):
        one = build([k], 1)
        run(one, 1)
        assert np.array_equal(one.state_get(0), got[k][0]), k
        np.testing.assert_array_equal(one.records(0), got[k][1])
        assert one.loop_get(0) == got[k][2], k
        if k == 0:
            one.params_set(0, P0)
            one.loop_set(0, t=ls["t"], dt_last=ls["dt_last"], step=ls["step"], done=0,
                         n_open=ls["n_ryr_open_last"], ca_max=ls["ca_ss_max_last"])
            one.explicit_step(chunk, S.STEP_PROTOCOL | S.STEP_GRAPH)
            assert np.array_equal(one.state_get(0), restarted)
            assert not np.array_equal(restarted, got[0][0])
        one.close()


@pytest.mark.parametrize("cfg", [dict(tile=128, occupancy=8), dict(tile=128, nt_geometry=1), dict(tile=512, nt_state=1),
                                 dict(tile=256, plane_pad=448, nt_geometry=1, nt_state=1), dict(compact=0),
                                 dict(group=2), dict(group=4, compact=0, occupancy=8), dict(lds=1),
                                 dict(lds=1, tile=128, compact=0, nt_state=1), dict(lds=1, tile=512, nt_geometry=1),
                                 dict(lds=1, tile=512, occupancy=6), dict(lds=1, tile=256, occupancy=8),
                                 dict(tile=256, occupancy=6), dict(group=2, occupancy=6),
                                 dict(tile=256, nt_geometry=0, lds=0), dict(lds=1, tile=512, occupancy=6, alt_order=1),
                                 dict(lds=1, tile=128, alt_order=1)])
def test_tunables_do_not_change_results(cfg):
    pr = _tiny(tt=True)
    p = fast_gating(pr.params)
    F = initial_fields(pr, seed=8)
    outs = []
    for tune in ({}, cfg):
        ctx = S.Context(0)
        ctx.tune(**{k: v for k, v in tune.items() if k in ("tile", "plane_pad", "compact", "group")})
        mid = ctx.add_mesh_grid(pr.grid, pr.aij, pr.nij, perm=W.kuhn_block_order(pr.grid))
        g = pr.grid
        ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=4, stream=0)
        ctx.finalize()
        ctx.tune(**{k: v for k, v in tune.items() if k in ("nt_geometry", "nt_state", "occupancy", "lds", "alt_order")})
        ctx.state_set(0, F.X)
        ctx.channels_set(0, F.ryr_state, F.lcc_state, F.lccv_state, F.R)
        ctx.globals_set(0, F.gv[0], F.gv[1])
        ctx.explicit_step(200, S.STEP_GRAPH)
        ctx.explicit_step(13, S.STEP_PROTOCOL)
        outs.append((ctx.state_get(0), ctx.channels_get(0)[0]))
        ctx.close()
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])



@pytest.mark.parametrize("cfg", [dict(tile=512), dict(tile=256, occupancy=6), dict(tile=128, nt_state=1, nt_geometry=0),
                                 dict(tile=128, occupancy=6, alt_order=1)])
def test_gated_sweep_equals_separate_gating(cfg):
    """Each step's gating run inside the next step's sweep launch
    (k_sweep_gated) against a gating launch after every sweep: identical
    fields, channel states, residuals and loop state after graph blocks of
    128 and of odd length, single launches and timed steps, with RyRs and
    LCCs flipping throughout."""
    pr = _tiny(tt=True)
    p = fast_gating(pr.params)
    F = initial_fields(pr, seed=8)
    outs = []
    for gated in (0, 1):
        ctx = S.Context(0)
        ctx.tune(tile=cfg["tile"])
        mid = ctx.add_mesh_grid(pr.grid, pr.aij, pr.nij, perm=W.kuhn_block_order(pr.grid))
        g = pr.grid
        ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=4, stream=0)
        ctx.finalize()
        ctx.tune(lds=1, gated_sweep=gated, **{k: v for k, v in cfg.items() if k != "tile"})
        ctx.state_set(0, F.X)
        ctx.channels_set(0, F.ryr_state, F.lcc_state, F.lccv_state, F.R)
        ctx.globals_set(0, F.gv[0], F.gv[1])
        seen = []
        for n, flags in ((300, S.STEP_GRAPH), (7, 0), (17, S.STEP_GRAPH), (1, 0), (129, S.STEP_GRAPH)):
            ctx.explicit_step(n, flags)
            seen.append(ctx.channels_get(0)[0].copy())
        ctx.time_sweep(5)
        ctx.explicit_step(3)
        outs.append((ctx.state_get(0), ctx.channels_get(0), ctx.loop_get(0), seen))
        ctx.close()
    (x0, c0, l0, s0), (x1, c1, l1, s1) = outs
    assert np.array_equal(x0, x1)
    for a, b in zip(c0, c1):
        assert np.array_equal(a, b)
    assert l0 == l1 and l1["step"] == 300 + 7 + 17 + 1 + 129 + 5 + 3
    for a, b in zip(s0, s1):
        assert np.array_equal(a, b)
    assert any(not np.array_equal(a, F.ryr_state) for a in s1)  # the RyRs did move

def test_ryr_open_probability_law():
    """At fixed Ca (no channel flux, SERCA and refill off, buffers at
    equilibrium) an uncoupled RyR's open and closed dwell times are geometric
    on the step grid (hazard integration with ln U residuals), so its long-run
    open fraction is m_o / (m_o + m_c) with m = 1 / (1 - exp(-rate dt)).
    576 independent channels (64 realizations x 9 RyRs, own Philox streams),
    20,000 recorded steps each; the mean of the per-channel open fractions is
    held to 4 standard errors."""
    pr = _tiny(nx=6, ny=6, nz=4, tt=False, ryr=3, lcc=0)
    g = pr.grid
    p = pr.params.copy()
    dt = p.reals[1]
    c = float(pr.X0[g.ryr_ele[0]])
    km = 1.0 / (40 * dt)
    kp = 0.6 * km / c ** p.reals[S.R["RYR_ETA"]]  # closed -> open rate 0.6 km at this Ca
    p.setr(RYR_K_MINUS=km, RYR_K_PLUS=kp, V_RYR=0.0, A_P=0.0, V_REFILL=0.0, V_CLAMP=-80.0, DEFAULT_VM=-80.0)
    n_real, n_steps = 64, 20000
    fields = [initial_fields(pr, ryr_open=(), vm=-80.0, seed=300 + k) for k in range(n_real)]
    assert np.all(pr.X0[g.ryr_ele] == c)
    with gpu_context([pr] * n_real, fields, params=[p] * n_real, seeds=[91] * n_real, share_mesh=True) as ctx:
        sched = ctx.record_schedules(n_steps)
        Xe = ctx.state_get(0)
    assert np.max(np.abs(Xe[g.ryr_ele] / c - 1)) < 1e-9  # Ca stayed put (round-off of the equilibrium only)
    frac = np.concatenate([s[1000:].mean(axis=0) for s in sched])  # drop the approach from all-closed
    m_o = 1.0 / (1.0 - np.exp(-km * dt))
    m_c = 1.0 / (1.0 - np.exp(-kp * c ** p.reals[S.R["RYR_ETA"]] * dt))
    po = m_o / (m_o + m_c)
    se = frac.std(ddof=1) / np.sqrt(frac.size)
    assert abs(frac.mean() - po) < 4 * se, (frac.mean(), po, se)
    assert 0.3 < po < 0.45 and se < 0.01


@pytest.mark.parametrize("group", [1, 2])
def test_checkpoint_resume_continues_bit_for_bit(group):
    """Protocol run with gating and states records: 300 steps straight, or
    150 steps, a checkpoint, a fresh context restored from it and 150 more
    steps -- the same fields, channel states, loop state and records."""
    pr = _tiny(tt=True)
    dt = pr.params.reals[1]
    p = fast_gating(pr.params)
    p.setr(T_FINAL=1e3 * dt, OUTPUT_STATES_INTERVAL=7 * dt)
    p.seti(FLAG_RYR_OUT=1)
    Fs = [initial_fields(pr, seed=60 + k) for k in range(2)]
    n_real = 2

    def build():
        return gpu_context([pr] * n_real, [F.copy() for F in Fs], params=[p] * n_real, seeds=[5, 6],
                           record_capacity=256, share_mesh=True, group=group)

    with build() as a:
        a.explicit_step(300, S.STEP_PROTOCOL | S.STEP_GRAPH)
        want = [(a.state_get(k), a.channels_get(k), a.loop_get(k), a.records(k)) for k in range(n_real)]
    with build() as b:
        b.explicit_step(150, S.STEP_PROTOCOL | S.STEP_GRAPH)
        cks = [b.checkpoint(k) for k in range(n_real)]
        recs_b = [b.records(k) for k in range(n_real)]
    with build() as c:
        for k in range(n_real):
            c.restore(k, cks[k])
        c.explicit_step(150, S.STEP_PROTOCOL | S.STEP_GRAPH)
        got = [(c.state_get(k), c.channels_get(k), c.loop_get(k), c.records(k)) for k in range(n_real)]
        # a checkpoint belongs to its realization's stream; a cut buffer is refused
        with pytest.raises(S.SrsError, match="seed or stream"):
            c.restore(1, cks[0])
        with pytest.raises(S.SrsError, match="truncated"):
            c.restore(0, cks[0][:-8])
        with pytest.raises(S.SrsError, match="not an srs checkpoint"):
            c.restore(0, b"XXXXXXXX" + cks[0][8:])
    for k in range(n_real):
        X, ch, lp, rec = want[k]
        assert np.array_equal(got[k][0], X)
        assert all(np.array_equal(u, v) for u, v in zip(got[k][1], ch))
        assert got[k][2] == lp
        np.testing.assert_array_equal(np.vstack([recs_b[k], got[k][3]]), rec)
    assert lp["step"] == 300 and len(rec) > 10


@pytest.mark.parametrize("compact", [1, 0])
def test_irregular_plane_spacing_gating_run(compact):
    """A box whose plane spacings are drawn at random (every element row
    differs, so the compact layout needs many classes and ulp deltas), with a
    T-tubule hole, the jSR slab and an SL face: a gated run matches the oracle
    in both row layouts, bit for bit in the channel states."""
    rng = np.random.default_rng(17)
    xs = np.concatenate([[0.0], np.cumsum(rng.uniform(0.025, 0.055, 9))])
    ys = np.concatenate([[0.0], np.cumsum(rng.uniform(0.025, 0.055, 8))])
    zs = np.concatenate([[0.0, 0.015], 0.015 + np.cumsum(rng.uniform(0.03, 0.05, 5))])
    cx, cy = xs[-1] / 2, ys[-1] / 2
    spec = S.BoxSpec(xs=xs, ys=ys, zs=zs, jsr_lo=(cx - 0.07, cy - 0.07, 0.015), jsr_hi=(cx + 0.07, cy + 0.07, 0.06),
                     sl_face_z0=True, tt_axis_x=0.0, tt_axis_z=0.15, tt_radius=0.06, trpn_radius=0.09, ryr_rows=3,
                     ryr_cols=3, ryr_spacing=0.031, ryr_center=(cx, cy), ryr_z_cyt=0.0075, ryr_z_jsr=0.03,
                     lcc_xyz=np.array([[cx + 0.02, cy - 0.03, 0.0075]]))
    pr = W.prepare(spec, W.spark_params())
    g = pr.grid
    p = fast_gating(pr.params)
    F = initial_fields(pr, seed=23)
    ctx = S.Context(0)
    ctx.tune(compact=compact)
    mid = ctx.add_mesh_grid(g, pr.aij, pr.nij)
    ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=29, stream=0)
    ctx.finalize()
    ctx.state_set(0, F.X)
    ctx.channels_set(0, F.ryr_state, F.lcc_state, F.lccv_state, F.R)
    ctx.globals_set(0, F.gv[0], F.gv[1])
    ctx.explicit_step(400, S.STEP_GRAPH)
    X, ch = ctx.state_get(0), ctx.channels_get(0)
    ctx.close()
    Fo = Oracle().steps(grid_dict(g, pr.aij, pr.nij), p.reals, F.copy(), 400, p.reals[1], gating=True,
                        rng_kind=RNG_PHILOX, key=29, stream=0)
    assert len(np.unique(pr.aij.reshape(-1, 5), axis=0)) > g.n_ele // 4  # the rows really vary
    assert np.array_equal(ch[0], Fo.ryr_state) and np.array_equal(ch[1], Fo.lcc_state)
    assert rel_err(X, Fo.X) <= TOL


def test_integration_bridge_runs_in_the_reference_step_loop(tmp_path):
    """The per-step bridge of the integration notes (SrsInit / SrsExplicitStep
    / SrsPullStates over the reference's own SimData and ParamStruct,
    cpu/parameters.h) stepped beside the reference's own ExplicitStep
    (cpu/RunSimulation.cpp:591, compiled from the reference source into
    oracle/_ref/bridge_check), with the RyR states of every step set from
    one random schedule before the call: the fields agree bit for bit after
    200 steps."""
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "bridge_check")
    if not os.path.exists(exe):
        pytest.skip("oracle/_ref/bridge_check is not built (needs the reference sources at build time)")
    base = str(tmp_path / "mesh")
    S.box_mesh_write(W.tiny_box_spec(nx=5, ny=5, tt=True, ryr=3), base)
    g, _ = S.grid_read(base)
    aij, _ = S.assemble_operator(g)
    dt = W.stable_dt(aij, 0.25)
    xml = str(tmp_path / "p.xml")
    _write_params(xml, dict(t_step=repr(dt * 1e6), V_RYR="5e-5", paramset_title="bridge"))
    out = subprocess.run([exe, xml, base, "200", "11"], capture_output=True, text=True, timeout=120)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, (out.returncode, out.stdout, out.stderr)
    assert " 0 differ" in out.stdout


@pytest.mark.parametrize("group", [1, 2])
def test_realization_subsets_equal_all_at_once(group):
    """SRS_TUNE_SUBSET_MB: runs of realization groups stepped through whole
    blocks of steps one after another (each run's state small enough to stay
    in the last-level cache) give every realization the fields, channel
    states, loop state and records of the all-at-once step, with protocol
    runs that finish at different times (the tile list shrinks, the subsets
    follow), graph blocks and single launches."""
    preps = [_tiny(tt=True), W.prepare(W.tiny_box_spec(nx=5, ny=4), W.spark_params())]
    dt = preps[0].params.reals[1]
    n_real = 40
    fields, params, prs = [], [], []
    for k in range(n_real):
        pr = preps[k % 2]
        p = fast_gating(pr.params)
        p.setr(T_FINAL=(150 + 9 * k) * dt, OUTPUT_STATES_INTERVAL=13 * dt)
        p.seti(FLAG_RYR_OUT=1)
        prs.append(pr)
        params.append(p)
        fields.append(initial_fields(pr, ryr_open=(k % pr.grid.n_ryr,), seed=400 + k))
    runs = []
    for mb in (0, 1):
        ctx = S.Context(0)
        ctx.tune(group=group)
        mids = [ctx.add_mesh_grid(pr.grid, pr.aij, pr.nij) for pr in preps]
        for k in range(n_real):
            g = prs[k].grid
            ctx.add_realization(mids[k % 2], g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, params[k], seed=90 + k,
                                stream=k)
        ctx.finalize(128)
        ctx.tune(subset_mb=mb)
        for k, F in enumerate(fields):
            ctx.state_set(k, F.X)
            ctx.channels_set(k, F.ryr_state, F.lcc_state, F.lccv_state, F.R)
            ctx.globals_set(k, F.gv[0], F.gv[1])
        for n in (200, 7, 150, 130):
            ctx.explicit_step(n, S.STEP_PROTOCOL | S.STEP_GRAPH)
        ctx.explicit_step(5, S.STEP_PROTOCOL)
        runs.append([(ctx.state_get(k), ctx.channels_get(k), ctx.loop_get(k), ctx.records(k)) for k in range(n_real)])
        ctx.close()
    steps = set()
    for a, b in zip(*runs):
        assert np.array_equal(a[0], b[0])
        assert all(np.array_equal(u, v) for u, v in zip(a[1], b[1]))
        assert a[2] == b[2]
        np.testing.assert_array_equal(a[3], b[3])
        steps.add(a[2]["step"])
    assert len(steps) > 10  # the runs ended at different steps


@pytest.mark.timeout(600)
def test_spark_box_gating_matches_reference_in_distribution():
    """Stochastic gating on the plain CPU workload's mesh: the 56,250-element
    spark box with its 7x7 cluster, the centre RyR open at t = 0, 1,500 steps,
    with RyR kinetics fast enough that the cluster recruits within that window
    (k_open 1.0 [Ca]^2.1 /ms, k_close 20 /ms).  1,024 device realizations
    (Philox, own streams, one batched context) against 128 runs of the
    oracle step with the reference's own glibc rand() gating (bit-exact with
    the reference's step): two-sample KS and Welch t tests, p > 0.01, on the
    centre RyR's first closing time, the number of openings, the number of
    RyRs recruited and the peak number open at once, and one-sample KS tests
    of both sets' first closing times against their exact geometric law.
    Seeds are fixed, so the outcome is reproducible.  tests/spark_box_power.py
    plants a 1.5x k_open or k_close in the oracle's Philox restatement of
    the device's gating and shows that this test then fails."""
    import multiprocessing as mp

    from tests.helpers import (SPARK_DIST, oracle_glibc_spark, oracle_spark_init, spark_dist_check,
                               spark_dist_fields, spark_stats)

    c = SPARK_DIST
    pr = W.prepare_spark()
    g = pr.grid
    p = pr.params.copy().setr(RYR_K_PLUS=c["k_plus"], RYR_K_MINUS=c["k_minus"])
    n_dev, n_ref, n_steps, first = c["n_dev"], c["n_ref"], c["n_steps"], c["first"]
    fields = [spark_dist_fields(pr, k) for k in range(n_dev)]
    print(f"spark box in distribution: {n_dev} device realizations", flush=True)
    with gpu_context([pr] * n_dev, fields, params=[p] * n_dev, seeds=[c["key"]] * n_dev, share_mesh=True,
                     order="block") as ctx:
        scheds = ctx.record_schedules(n_steps)
    dev = np.array([spark_stats(s, first) for s in scheds])
    del scheds
    print(f"spark box in distribution: {n_ref} oracle runs with glibc gating, 16 processes", flush=True)
    seeds = [93843 + 38376427 + 83 * k for k in range(n_ref)]
    init = (grid_dict(g, pr.aij, pr.nij), np.array(p.reals), pr.X0)
    with mp.get_context("spawn").Pool(16, initializer=oracle_spark_init, initargs=init) as pool:
        ref = np.array(pool.map(oracle_glibc_spark, [(s, n_steps, first) for s in seeds]))
    q = 1.0 - np.exp(-c["k_minus"] * p.reals[S.R["T_STEP"]])
    names, pv, ok = spark_dist_check(dev, ref, q, n_steps)
    for nm, v in zip(names, pv):
        print(f"  {nm}: p {v:.4f}", flush=True)
    print(f"  means device {dev.mean(axis=0)}, reference {ref.mean(axis=0)}", flush=True)
    assert ok, dict(zip(names, pv))
    assert ref[:, 1].mean() > 1 and dev[:, 2].mean() > 1.5  # the cluster does recruit
