"""Adjoint of the frozen-channel step (needs an MI355X): gradients of a line
scan loss against central finite differences of the same device forward
run, checkpoint-interval independence, the torch.autograd wrapper of the
fitting harness, and conservation of total Ca in the device step."""
import numpy as np
import pytest

import srs_amd as S
from srs_amd import workloads as W
from srs_amd.fit import LinescanProblem, linescan_loss_fn
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

N_STEPS, N_SAMPLES = 600, 20


def _setup(perturb=0.1, seed=3, var=(0.001, 0.003, 0.001), lift=0.0):
    spec = W.tiny_box_spec(nx=8, ny=8, nz=5, tt=True, ryr=2, lcc=1)
    pr = W.prepare(spec, W.spark_params())
    g = pr.grid
    p = pr.params.copy()
    dt = p.reals[S.R["T_STEP"]]
    p.setr(RYR_K_MINUS=1 / (150 * dt), RYR_K_PLUS=1 / (200 * dt * 0.2 ** 2.1), V_RYR=20 * p.reals[S.R["V_RYR"]])
    ctx = S.Context(0)
    mid = ctx.add_mesh_grid(g, pr.aij, pr.nij, perm=W.kuhn_block_order(g))
    ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=5, stream=0)
    ctx.finalize()
    rs = np.zeros(g.n_ryr, np.int32)
    rs[0] = 1
    R = np.log(np.random.default_rng(seed).random(g.n_ryr + g.n_lcc))
    ctx.state_set(0, pr.X0)
    ctx.channels_set(0, rs, np.ones(g.n_lcc, np.int32), np.full(g.n_lcc, 2, np.int32), R)
    ctx.globals_set(0, 0.0, p.reals[S.R["DEFAULT_CA_NSR"]])
    schedule = ctx.record_schedule(0, N_STEPS)
    assert schedule.sum(axis=1).max() >= 1 and schedule[0, 0] == 1
    # LCCs back to their initial states; they stay there in the frozen runs
    ctx.channels_set(0, schedule[0], np.ones(g.n_lcc, np.int32), np.full(g.n_lcc, 2, np.int32), R)
    # a scan line along x through the cluster, in the centred coordinates
    # the linescan program assumes
    cen = g.centroids()
    lo, hi = cen.min(axis=0), cen.max(axis=0)
    mid_xyz = (lo + hi) / 2
    d = S.read_scan_parameters(None)
    d.axis = 0
    d.x_min, d.x_max, d.res_x_f, d.res_x_ca = float(lo[0] - mid_xyz[0]), float(hi[0] - mid_xyz[0]), 0.02, 0.01
    rc = cen[g.ryr_ele].mean(axis=0) - mid_xyz
    d.offset1, d.offset2 = float(rc[1]), float(rc[2]) + lift
    for k in range(3):
        d.domain[k] = float(hi[k] - lo[k]) + 0.08
        d.var[k] = var[k]
    d.pad_min, d.pad_max, d.pad_res = -0.6, 0.6, 0.05
    sid, nf, _ = ctx.add_scan(mid, cen - mid_xyz, d)
    samples = np.linspace(N_STEPS // N_SAMPLES, N_STEPS, N_SAMPLES).astype(np.int64)
    # target: the same frozen-channel run with every theta moved by +-perturb
    idx = S.adjoint_theta()
    sgn = np.where(np.random.default_rng(seed + 1).random(len(idx)) < 0.5, -1.0, 1.0)
    P_true = p.reals.copy()
    P_true[idx] *= 1 + perturb * sgn
    target = _scan_run(ctx, sid, schedule, samples, P_true, pr.X0, d.f0)
    ctx.params_set(0, p.reals)
    ctx.state_set(0, pr.X0)
    prob = LinescanProblem(ctx, sid, schedule, samples, target, p, pr.X0, ckpt=64)
    return prob, P_true[idx], nf


def _scan_run(ctx, sid, schedule, samples, P, X0, f0):
    """dF/F0 at the sample steps by chunked replays (the same kernel sequence
    as the loss's forward run)."""
    ctx.params_set(0, P)
    ctx.state_set(0, X0)
    out, prev = [], 0
    for s in samples:
        ctx.replay(0, schedule[prev:s])
        prev = int(s)
        out.append(S.dff0(ctx.scan_sample(0, sid)[2], f0))
    return np.array(out)


def test_loss_vanishes_at_the_generating_parameters():
    prob, theta_true, _ = _setup()
    assert prob.loss(theta_true) == 0.0
    assert prob.loss(prob.theta()) > 0.0
    prob.ctx.close()


@pytest.mark.parametrize("var,lift", [((0.001, 0.003, 0.001), 0.0), ((5e-5, 5e-5, 5e-5), 0.0),
                                      ((0.1, 0.1, 0.1), 1.05)])
def test_adjoint_gradient_matches_central_differences(var, lift):
    """Every component: |theta g - theta g_fd| <= 1e-5 relative (components
    three orders below the largest are held to 1e-5 of the largest).  The
    second PSF (s = 7 nm) is too narrow for the blur's walk along the line:
    every weight of the blur and of the loss source is the closed form.  The
    third (s = 0.32 um) has its line 1.05 um above the cluster, where 980 of
    the 1,776 cells are within the 1 um cut-off: most wavefronts hold cells
    on both sides of it.  There the blurred signal is mostly the padding's
    constant, and the loss's rounding over a step of 1e-5 swamps the
    differences (about 1e-19 against sensitivities up to 2e-13; the error
    falls as 1/step), so its differences take steps of 1e-2 and 5e-3 and
    Richardson's extrapolation (4 D(h/2) - D(h)) / 3, which cancels the
    central difference's h^2 term."""
    prob, _, _ = _setup(var=var, lift=lift)
    th = prob.theta()
    L0, g = prob.loss_grad(th)
    assert L0 > 0

    def central(eps):
        d = np.zeros_like(g)
        for k in range(len(th)):
            tp, tm = th.copy(), th.copy()
            tp[k] *= 1 + eps
            tm[k] *= 1 - eps
            d[k] = (prob.loss(tp) - prob.loss(tm)) / (2 * eps * th[k])
        return d

    fd = (4 * central(5e-3) - central(1e-2)) / 3 if lift else central(1e-5)
    sens, sens_fd = th * g, th * fd
    scale = np.abs(sens_fd).max()
    assert scale > 0
    err = np.abs(sens - sens_fd) / np.maximum(np.abs(sens_fd), 1e-3 * scale)
    assert err.max() <= 1e-5, (err, sens, sens_fd)
    prob.ctx.close()


def test_checkpoint_interval_does_not_change_the_gradient():
    prob, _, _ = _setup()
    th = prob.theta()
    out = []
    for ck in (7, 64, 600):
        prob.ckpt = ck
        out.append(prob.loss_grad(th))
    for L, g in out[1:]:
        assert L == out[0][0] and np.array_equal(g, out[0][1])
    prob.ctx.close()


def test_autograd_wrapper_returns_the_adjoint_gradient():
    import torch

    prob, _, _ = _setup()
    th = torch.tensor(prob.theta(), dtype=torch.float64, requires_grad=True)
    loss = linescan_loss_fn().apply(th, prob)
    loss.backward()
    L, g = prob.loss_grad(prob.theta())
    assert loss.item() == L
    np.testing.assert_array_equal(th.grad.numpy(), g)
    prob.ctx.close()


def test_total_calcium_conserved_on_device():
    """No SERCA, refill, LCC or SL buffer, no-flux walls: total Ca (free plus
    every bound form, volume-weighted) is conserved by the device step while
    RyRs move Ca from the jSR to the cleft."""
    pr = W.prepare(W.tiny_box_spec(nx=6, ny=6, nz=4, tt=False, ryr=2, lcc=1), W.spark_params())
    p = pr.params.copy()
    dt = p.reals[S.R["T_STEP"]]
    p.setr(A_P=0.0, V_REFILL=0.0, B_TOT_SL_HI=0.0, RYR_K_MINUS=1 / (300 * dt), V_RYR=50 * p.reals[S.R["V_RYR"]])
    g = pr.grid
    with S.Context(0) as ctx:
        mid = ctx.add_mesh_grid(g, pr.aij, pr.nij)
        ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=2, stream=0)
        ctx.finalize()
        rs = np.ones(g.n_ryr, np.int32)
        ctx.state_set(0, pr.X0)
        ctx.channels_set(0, rs, np.ones(g.n_lcc, np.int32), np.ones(g.n_lcc, np.int32),
                         np.log(np.random.default_rng(1).random(g.n_ryr + g.n_lcc)))
        ctx.globals_set(0, 0.0, p.reals[S.R["DEFAULT_CA_NSR"]])
        n = g.n_ele

        def total(X):
            return float(np.sum(g.V0 * X.reshape(6, n).sum(axis=0)))

        T0 = total(pr.X0)
        ctx.explicit_step(500, S.STEP_GRAPH)
        X = ctx.state_get(0)
    assert abs(total(X) - T0) <= 1e-12 * T0
    assert np.abs(X - pr.X0).max() > 1.0  # Ca did move


def test_schedule_recorded_mid_run_follows_the_gated_trajectory():
    """Recording after earlier steps: row 0 holds the RyR states at the call,
    row k the states after k more gated steps (the step counter of the
    realization does not start at zero), and a realization that has finished
    its protocol is refused."""
    from tests.helpers import fast_gating, gpu_context, initial_fields

    pr = W.prepare(W.tiny_box_spec(nx=6, ny=6, nz=4, tt=True, ryr=3, lcc=1), W.spark_params())
    p = fast_gating(pr.params.copy())
    F = initial_fields(pr, seed=4)
    n0, n = 37, 120
    a = gpu_context([pr], [F], params=[p], seeds=[9])
    b = gpu_context([pr], [F], params=[p], seeds=[9])
    a.explicit_step(n0, S.STEP_PLAIN)
    b.explicit_step(n0, S.STEP_PLAIN)
    sched = a.record_schedule(0, n)
    rows = [b.channels_get(0)[0]]
    for _ in range(n - 1):
        b.explicit_step(1, S.STEP_PLAIN)
        rows.append(b.channels_get(0)[0])
    rows = np.array(rows)
    assert np.array_equal(sched, rows)
    assert len({r.tobytes() for r in rows}) > 1  # the RyRs did move
    ls = a.loop_get(0)
    a.loop_set(0, t=ls["t"], dt_last=ls["dt_last"], step=ls["step"], done=1)
    with pytest.raises(S.SrsError, match="finished"):
        a.record_schedule(0, 4)
    a.close()
    b.close()


# ---------------------------------------------------------------- batched adjoint

B_STEPS, B_SAMPLES = 500, 10
B_SUBSETS = ([0, 1, 2, 3, 4, 5, 6, 7, 8], [0, 4, 8], [1, 3, 4, 5, 7], [4], [2, 4, 6, 8])


def _subset_channels(g, idx):
    """RyRs idx of the grid's lattice with the neighbour table renumbered."""
    idx = np.asarray(idx)
    new = -np.ones(g.n_ryr, np.int64)
    new[idx] = np.arange(len(idx))
    nb = g.ryr_nb.reshape(-1, 4)[idx]
    nb = np.where(nb >= 0, new[np.maximum(nb, 0)], -1).astype(np.int32)
    return g.ryr_ele[idx], g.ryr_jsr[idx], nb.ravel()


def _scan_desc(g):
    cen = g.centroids()
    lo, hi = cen.min(axis=0), cen.max(axis=0)
    mid = (lo + hi) / 2
    d = S.read_scan_parameters(None)
    d.axis = 0
    d.x_min, d.x_max, d.res_x_f, d.res_x_ca = float(lo[0] - mid[0]), float(hi[0] - mid[0]), 0.03, 0.01
    rc = cen[g.ryr_ele].mean(axis=0) - mid
    d.offset1, d.offset2 = float(rc[1]), float(rc[2])
    for k in range(3):
        d.domain[k] = float(hi[k] - lo[k]) + 0.08
        d.var[k] = (0.001, 0.003, 0.001)[k]
    d.pad_min, d.pad_max, d.pad_res = -0.6, 0.6, 0.05
    return cen - mid, d


def _member(pr, m):
    """Parameters, channels, schedule and initial state of member m (seeded)."""
    g = pr.grid
    rng = np.random.default_rng(100 + m)
    p = pr.params.copy()
    p.setr(V_RYR=20 * p.reals[S.R["V_RYR"]], DEFAULT_CA_NSR=p.reals[S.R["DEFAULT_CA_NSR"]] * (0.8 + 0.1 * m))
    sub = B_SUBSETS[m]
    # a frozen trajectory: each RyR opens for one random window
    sched = np.zeros((B_STEPS, len(sub)), np.int32)
    for i in range(len(sub)):
        a = int(rng.integers(0, B_STEPS // 2))
        sched[a:a + int(rng.integers(50, 250)), i] = 1
    sched[0, 0] = 1
    R = np.log(rng.random(len(sub) + g.n_lcc))
    return p, sub, sched, R


def _batch_ctx(pr, members, group=1):
    g = pr.grid
    ctx = S.Context(0)
    ctx.tune(group=group)
    mid = ctx.add_mesh_grid(g, pr.aij, pr.nij, perm=W.kuhn_block_order(g))
    for p, sub, _, _ in members:
        ctx.add_realization(mid, *_subset_channels(g, sub), g.lcc_ele, p, seed=5, stream=0)
    ctx.finalize()
    for k, (p, sub, sched, R) in enumerate(members):
        ctx.state_set(k, pr.X0)
        ctx.channels_set(k, sched[0], np.ones(g.n_lcc, np.int32), np.full(g.n_lcc, 2, np.int32), R)
        ctx.globals_set(k, 0.0, p.reals[S.R["DEFAULT_CA_NSR"]])
    cen, d = _scan_desc(g)
    sid, nf, _ = ctx.add_scan(mid, cen, d)
    return ctx, sid, nf, d


def _targets(pr, members, samples, perturb=0.1):
    """Each member's dF/F0 at its theta moved by +-perturb, from a context of
    its own through replay and scan sampling (a different code path from the
    batched loss)."""
    idx = S.adjoint_theta()
    out, P_true = [], []
    for m, mem in enumerate(members):
        p, sub, sched, R = mem
        ctx, sid, nf, d = _batch_ctx(pr, [mem])
        sgn = np.where(np.random.default_rng(200 + m).random(len(idx)) < 0.5, -1.0, 1.0)
        P = p.reals.copy()
        P[idx] *= 1 + perturb * sgn
        out.append(_scan_run(ctx, sid, sched, samples, P, pr.X0, d.f0))
        P_true.append(P[idx])
        ctx.close()
    return np.array(out), np.array(P_true)


def _batch_setup():
    pr = W.prepare(W.tiny_box_spec(nx=8, ny=8, nz=5, tt=True, ryr=3, lcc=1), W.spark_params())
    members = [_member(pr, m) for m in range(len(B_SUBSETS))]
    samples = np.linspace(B_STEPS // B_SAMPLES, B_STEPS, B_SAMPLES).astype(np.int64)
    return pr, members, samples


def _set_theta(ctx, members, theta):
    idx = S.adjoint_theta()
    for k, (p, _, _, _) in enumerate(members):
        P = p.reals.copy()
        P[idx] = theta[k] if np.ndim(theta) == 2 else theta
        ctx.params_set(k, P)


def test_batched_adjoint_equals_single_realization_runs():
    """Per-member losses and gradients of the batched call equal the
    one-realization call on each member bit for bit, for any sub-batch size;
    the losses vanish at each member's generating parameters; every
    realization, member or not, is left as it was."""
    pr, members, samples = _batch_setup()
    tg, P_true = _targets(pr, members, samples)
    ctx, sid, nf, _ = _batch_ctx(pr, members)
    n = len(members)
    scheds = [m[2] for m in members]
    assert tg.shape == (n, B_SAMPLES, nf)
    _set_theta(ctx, members, P_true)
    L_true, sc = ctx.loss_grad_batch(range(n), sid, scheds, samples, tg, ckpt=64, grad=False, scans=True)
    assert np.all(L_true == 0.0) and np.array_equal(sc, tg)
    _set_theta(ctx, members, np.array([m[0].reals[S.adjoint_theta()] for m in members]))
    before = [(ctx.state_get(k), ctx.channels_get(k), ctx.loop_get(k)) for k in range(n)]
    L_all, g_all = ctx.loss_grad_batch(range(n), sid, scheds, samples, tg, ckpt=64)
    assert np.all(L_all > 0) and np.all(np.isfinite(g_all))
    for mb in (1, 2, 3):
        L2, g2 = ctx.loss_grad_batch(range(n), sid, scheds, samples, tg, ckpt=64, max_batch=mb)
        assert np.array_equal(L2, L_all) and np.array_equal(g2, g_all)
    for k in range(n):
        L1, g1 = ctx.loss_grad(k, sid, scheds[k], samples, tg[k], 64)
        assert L1 == L_all[k] and np.array_equal(g1, g_all[k])
    # a subset of the members, in another order
    sub = [3, 0, 2]
    L3, g3 = ctx.loss_grad_batch(sub, sid, [scheds[k] for k in sub], samples, tg[sub], ckpt=64)
    assert np.array_equal(L3, L_all[sub]) and np.array_equal(g3, g_all[sub])
    for k in range(n):
        X, ch, lp = before[k]
        assert np.array_equal(ctx.state_get(k), X)
        assert all(np.array_equal(a, b) for a, b in zip(ctx.channels_get(k), ch))
        assert ctx.loop_get(k) == lp
    ctx.close()


def test_batched_adjoint_gradients_match_central_differences():
    """Four sampled members: |theta g - theta g_fd| <= 1e-5 relative, the FD
    losses of all members coming from batched loss-only calls."""
    pr, members, samples = _batch_setup()
    tg, _ = _targets(pr, members, samples)
    ctx, sid, nf, _ = _batch_ctx(pr, members)
    n = len(members)
    scheds = [m[2] for m in members]
    th = np.array([m[0].reals[S.adjoint_theta()] for m in members])
    _set_theta(ctx, members, th)
    L0, g = ctx.loss_grad_batch(range(n), sid, scheds, samples, tg)  # automatic checkpoint interval
    fd = np.zeros_like(g)
    eps = 1e-5
    for k in range(th.shape[1]):
        tp, tm = th.copy(), th.copy()
        tp[:, k] *= 1 + eps
        tm[:, k] *= 1 - eps
        _set_theta(ctx, members, tp)
        Lp = ctx.loss_grad_batch(range(n), sid, scheds, samples, tg, grad=False)
        _set_theta(ctx, members, tm)
        Lm = ctx.loss_grad_batch(range(n), sid, scheds, samples, tg, grad=False)
        fd[:, k] = (Lp - Lm) / (2 * eps * th[:, k])
    for m in (0, 1, 2, 4):
        sens, sens_fd = th[m] * g[m], th[m] * fd[m]
        scale = np.abs(sens_fd).max()
        assert scale > 0
        err = np.abs(sens - sens_fd) / np.maximum(np.abs(sens_fd), 1e-3 * scale)
        assert err.max() <= 1e-5, (m, err, sens, sens_fd)
    ctx.close()


def test_batched_adjoint_argument_errors():
    pr, members, samples = _batch_setup()
    ctx, sid, nf, _ = _batch_ctx(pr, members[:2])
    scheds = [m[2] for m in members[:2]]
    tg = np.zeros((2, B_SAMPLES, nf))
    with pytest.raises(S.SrsError, match="twice"):
        ctx.loss_grad_batch([0, 0], sid, [scheds[0], scheds[0]], samples, tg)
    with pytest.raises(S.SrsError, match="out of range"):
        ctx.loss_grad_batch([0, 5], sid, scheds, samples, tg)
    with pytest.raises(S.SrsError, match="sample steps"):
        ctx.loss_grad_batch([0, 1], sid, scheds, samples[::-1].copy(), tg)
    ctx.close()
    ctx, sid, nf, _ = _batch_ctx(pr, members[:2], group=2)
    with pytest.raises(S.SrsError, match="groups of size 1"):
        ctx.loss_grad_batch([0, 1], sid, scheds, samples, tg)
    ctx.close()


def test_schedules_recorded_together_equal_separate_recordings():
    """srs_record_schedules over a context of three realizations (own seeds
    and RyR subsets) gives each the schedule it records alone."""
    pr = W.prepare(W.tiny_box_spec(nx=6, ny=6, nz=4, tt=True, ryr=3, lcc=1), W.spark_params())
    g = pr.grid
    p = pr.params.copy()
    dt = p.reals[S.R["T_STEP"]]
    p.setr(RYR_K_MINUS=1 / (40 * dt), RYR_K_PLUS=1 / (40 * dt * 0.2 ** 2.1))
    subs = (B_SUBSETS[0], B_SUBSETS[2], B_SUBSETS[1])
    n = 300

    def build(ks):
        ctx = S.Context(0)
        mid = ctx.add_mesh_grid(g, pr.aij, pr.nij)
        for k in ks:
            ctx.add_realization(mid, *_subset_channels(g, subs[k]), g.lcc_ele, p, seed=31 + k, stream=k)
        ctx.finalize()
        for q, k in enumerate(ks):
            rs = np.zeros(len(subs[k]), np.int32)
            rs[0] = 1
            ctx.state_set(q, pr.X0)
            ctx.channels_set(q, rs, np.ones(g.n_lcc, np.int32), np.full(g.n_lcc, 2, np.int32),
                             np.log(np.random.default_rng(k).random(len(subs[k]) + g.n_lcc)))
            ctx.globals_set(q, 0.0, p.reals[S.R["DEFAULT_CA_NSR"]])
        return ctx

    ctx = build([0, 1, 2])
    ctx.explicit_step(11, S.STEP_PLAIN)
    both = ctx.record_schedules(n)
    ctx.close()
    for k in range(3):
        one = build([k])
        one.explicit_step(11, S.STEP_PLAIN)
        sk = one.record_schedule(0, n)
        one.close()
        assert both[k].shape == (n, len(subs[k])) and np.array_equal(both[k], sk)
        assert len({r.tobytes() for r in sk}) > 1


def test_batch_fit_reduces_the_mean_loss():
    """The batch-mean problem through the torch.autograd wrapper: its
    gradient is the mean of the members' adjoint gradients, and a few Adam
    steps on log theta lower the mean loss."""
    import torch

    from srs_amd.fit import BatchLinescanProblem, fit, linescan_loss_fn

    pr, members, samples = _batch_setup()
    tg, _ = _targets(pr, members, samples)
    ctx, sid, nf, _ = _batch_ctx(pr, members)
    n = len(members)
    prob = BatchLinescanProblem(ctx, sid, range(n), [m[2] for m in members], samples, tg, [m[0] for m in members],
                                [pr.X0] * n, ckpt=32)
    th0 = prob.theta()
    L, g = prob.loss_grad(th0)
    Ls, gs = ctx.loss_grad_batch(range(n), sid, [m[2] for m in members], samples, tg, ckpt=32)
    assert L == float(np.mean(Ls)) and np.array_equal(g, gs.mean(axis=0))
    t = torch.tensor(th0, dtype=torch.float64, requires_grad=True)
    linescan_loss_fn().apply(t, prob).backward()
    np.testing.assert_array_equal(t.grad.numpy(), g)
    th = fit(prob, steps=6, lr=0.02)
    assert prob.loss(th) < L
    ctx.close()


@pytest.mark.timeout(900)
def test_spark_box_fitting_workload_gradient_matches_central_differences():
    """The fitting workload at its own size: the 56,250-element spark box,
    5 ms (83,519 steps) with the RyR trajectory frozen from a gated run, a
    64-point line scan x 50 samples, checkpoints every 256 steps, all 13
    parameters: adjoint against central differences of the same device
    forward run, |theta g - theta g_fd| <= 1e-5 relative (components three
    orders below the largest held to 1e-5 of the largest); the loss vanishes
    at the generating parameters."""
    print("fitting workload: setup (schedule, target)", flush=True)
    prob, info = W.spark_adjoint_problem()
    assert info["n_steps"] == 83519 and info["n_f"] == 64 and info["n_openings"] > 0
    assert info["loss_at_target"] == 0.0
    th = prob.theta()
    L0, g = prob.loss_grad(th)
    assert L0 > 0
    fd = np.zeros_like(g)
    eps = 1e-5
    for k in range(len(th)):
        tp, tm = th.copy(), th.copy()
        tp[k] *= 1 + eps
        tm[k] *= 1 - eps
        fd[k] = (prob.loss(tp) - prob.loss(tm)) / (2 * eps * th[k])
        print(f"  theta {k}: adjoint {th[k] * g[k]:.6e}, central difference {th[k] * fd[k]:.6e}", flush=True)
    prob.ctx.close()
    sens, sens_fd = th * g, th * fd
    scale = np.abs(sens_fd).max()
    assert scale > 0
    err = np.abs(sens - sens_fd) / np.maximum(np.abs(sens_fd), 1e-3 * scale)
    assert err.max() <= 1e-5, (err, sens, sens_fd)


def _bound_ctx(g, aij, nij, p, bnd, sched, R):
    """A one-realization context whose mesh carries boundary rates built at
    p's D (cpu/InitFuncs.cpp:1012-1026) from the boundary values bnd."""
    br = S.boundary_rates(g, p.reals, bnd)
    assert np.any(br != 0)
    ctx = S.Context(0)
    mid = ctx.add_mesh_grid(g, aij, nij, br=br, perm=W.kuhn_block_order(g))
    ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=5, stream=0)
    ctx.finalize()
    ctx.channels_set(0, sched[0], np.ones(g.n_lcc, np.int32), np.full(g.n_lcc, 2, np.int32), R)
    ctx.globals_set(0, 0.0, p.reals[S.R["DEFAULT_CA_NSR"]])
    cen, d = _scan_desc(g)
    sid, nf, _ = ctx.add_scan(mid, cen, d)
    return ctx, sid, nf


def test_boundary_rates_follow_d_in_step_and_gradient():
    """Bound faces (flag_noflux off) with boundary values above the initial
    state: the boundary rates D Face_C b / V0 move with D.  A run whose D_Ca
    and D_Fluo are changed by srs_params_set equals the oracle run with the
    rates rebuilt at the new D (rel <= 1e-12: the planes are rescaled, not
    rebuilt), and so does its loss; the adjoint's D gradients match central
    differences of losses from contexts whose rates are rebuilt at each
    perturbed D."""
    from oracle.oracle import Fields, Oracle, grid_dict

    spec = W.tiny_box_spec(nx=8, ny=8, nz=5, tt=True, ryr=2, lcc=1)
    g = S.box_grid(spec, no_flux=False)
    aij, nij = S.assemble_operator(g)
    p = W.spark_params()
    dt = W.stable_dt(aij, 0.25)
    p.setr(T_STEP=dt, INIT_RYR_HOLD=dt, V_RYR=20 * p.reals[S.R["V_RYR"]])
    p.seti(INIT_RYR=-1)
    X0, bnd = S.equilibrium_state(g, p.reals)
    bnd = bnd * np.array([3.0, 1.0, 1.0, 1.0, 1.0, 1.2])  # Ca and Fluo flow in through the bound faces
    n = 400
    sched = np.zeros((n, g.n_ryr), np.int32)
    sched[: n // 2, 0] = 1
    R = np.log(np.random.default_rng(3).random(g.n_ryr + g.n_lcc))
    samples = np.linspace(n // 8, n, 8).astype(np.int64)
    # the step: D moved after the mesh was added, against the oracle with the rates rebuilt
    p1 = p.copy().setr(D_CA=0.3, D_DYE=0.05)
    ctx, sid, nf = _bound_ctx(g, aij, nij, p, bnd, sched, R)
    ctx.params_set(0, p1.reals)
    ctx.state_set(0, X0)
    ctx.replay(0, sched)
    Xd = ctx.state_get(0)
    F = Fields(X0, sched[0], np.ones(g.n_lcc), np.full(g.n_lcc, 2), R, [0.0, p.reals[S.R["DEFAULT_CA_NSR"]]])
    Fo = Oracle().steps(grid_dict(g, aij, nij, S.boundary_rates(g, p1.reals, bnd)), p1.reals, F, n, dt, gating=False,
                        schedule=sched)
    assert rel_err(Xd, Fo.X) <= 1e-12
    assert rel_err(Xd, X0) > 1e-3
    # the loss (target zero: the squared dF/F0) and its D gradients
    tgt = np.zeros((len(samples), nf))

    def loss(c, s, P):
        c.params_set(0, P)
        c.state_set(0, X0)
        return c.loss_grad(0, s, sched, samples, tgt, 64, grad=False)

    ctx.params_set(0, p.reals)
    ctx.state_set(0, X0)
    L, gr = ctx.loss_grad(0, sid, sched, samples, tgt, 64)
    L1 = loss(ctx, sid, p1.reals)
    c1, s1, _ = _bound_ctx(g, aij, nij, p1, bnd, sched, R)
    assert abs(loss(c1, s1, p1.reals) - L1) <= 1e-12 * L1
    c1.close()
    eps = 1e-5
    for k, name in ((0, "D_CA"), (1, "D_DYE")):
        th = p.reals[S.R[name]]
        Ls = []
        for sgn in (1, -1):
            q = p.copy().setr(**{name: th * (1 + sgn * eps)})
            c2, s2, _ = _bound_ctx(g, aij, nij, q, bnd, sched, R)
            Ls.append(loss(c2, s2, q.reals))
            c2.close()
        fd = (Ls[0] - Ls[1]) / (2 * eps * th)
        assert abs(gr[k] - fd) <= 1e-5 * abs(fd), (name, gr[k], fd)
    ctx.close()


@pytest.mark.parametrize("levels", [2, 3])
def test_checkpoint_levels_do_not_change_the_batched_gradient(levels):
    """Two- and three-level checkpoint plans (each level's segments
    recomputed from its checkpoints, the last level's with every state kept)
    give the one-level plan's losses and gradients bit for bit, in one
    sub-batch and in sub-batches of two; the plan in use is reported."""
    pr, members, samples = _batch_setup()
    tg, _ = _targets(pr, members, samples)
    ctx, sid, nf, _ = _batch_ctx(pr, members)
    n = len(members)
    scheds = [m[2] for m in members]
    ctx.tune(adj_levels=1)
    L1, g1 = ctx.loss_grad_batch(range(n), sid, scheds, samples, tg, ckpt=23)
    assert ctx.adjoint_plan() == dict(levels=1, strides=[23], sub_batch=n)
    ctx.tune(adj_levels=levels)
    for ck, mb in ((0, 0), (5, 2)):
        L2, g2 = ctx.loss_grad_batch(range(n), sid, scheds, samples, tg, ckpt=ck, max_batch=mb)
        plan = ctx.adjoint_plan()
        assert plan["levels"] == levels and plan["sub_batch"] == (mb or n)
        assert plan["strides"] == sorted(plan["strides"], reverse=True) and plan["strides"][0] < B_STEPS
        if ck:
            assert plan["strides"][-1] == ck
        assert np.array_equal(L2, L1) and np.array_equal(g2, g1)
    ctx.close()


@pytest.mark.timeout(1100)
def test_dyad30_batch_gradients_at_full_batch_match_central_differences():
    """The batched fitting workload at its batch size: 256 members on the
    h = 30 nm dyad geometry (927,816 elements each), own recorded
    trajectories, 2,000 steps, a 64-point line scan x 8 samples, with the
    two-level checkpoint plan (outer checkpoints, inner checkpoints per outer
    segment, recomputed segments of every state).  The gradients of four
    sampled members against central differences of their losses (loss-only
    calls with only that member's theta moved), |theta g - theta g_fd| <=
    1e-5 relative (components three orders below the largest held to 1e-5 of
    the largest); losses vanish at each member's generating parameters."""
    n_real, n_steps = 256, 2000
    print("dyad30 batch: setup", flush=True)
    ctx, sid, scheds, samples, tgt, p, info = W.dyad_adjoint_batch(n_real, n_steps, log=print)
    assert info["n_ele"] == 927816 and info["n_f"] == 64
    idx = S.adjoint_theta()
    ctx.tune(adj_levels=2)
    print("dyad30 batch: losses and gradients of all members", flush=True)
    L, g = ctx.loss_grad_batch(range(n_real), sid, scheds, samples, tgt)
    plan = ctx.adjoint_plan()
    print(f"  plan {plan}", flush=True)
    assert plan["levels"] == 2 and np.all(L > 0) and np.all(np.isfinite(g))
    picked = [0, 85, 170, 255]
    th = p.reals[idx].copy()
    eps = 1e-5
    # a loss-only call starts from each member's current state, which every call restores
    Pt = p.reals.copy()
    Pt[idx] = info["theta_true"]
    for m in picked:
        ctx.params_set(m, Pt)
    Lt = ctx.loss_grad_batch(picked, sid, [scheds[m] for m in picked], samples, tgt[picked], grad=False)
    assert np.all(Lt == 0.0)
    fd = np.zeros((len(picked), len(idx)))
    for k in range(len(idx)):
        Ls = []
        for sgn in (1, -1):
            P = p.reals.copy()
            P[idx[k]] *= 1 + sgn * eps
            for m in picked:
                ctx.params_set(m, P)
            Ls.append(ctx.loss_grad_batch(picked, sid, [scheds[m] for m in picked], samples, tgt[picked], grad=False))
        fd[:, k] = (Ls[0] - Ls[1]) / (2 * eps * th[k])
        print(f"  theta {k}: member 0 adjoint {th[k] * g[0, k]:.6e}, central difference {th[k] * fd[0, k]:.6e}",
              flush=True)
    ctx.close()
    for q, m in enumerate(picked):
        sens, sens_fd = th * g[m], th * fd[q]
        scale = np.abs(sens_fd).max()
        assert scale > 0
        err = np.abs(sens - sens_fd) / np.maximum(np.abs(sens_fd), 1e-3 * scale)
        assert err.max() <= 1e-5, (m, err, sens, sens_fd)


def test_adjoint_register_budget_does_not_change_the_gradient():
    """SRS_TUNE_ADJ_WAVES 4 (the reverse step limited to 128 VGPRs) gives
    the compiler's-budget losses and gradients bit for bit, batched and
    single."""
    pr, members, samples = _batch_setup()
    tg, _ = _targets(pr, members, samples)
    ctx, sid, nf, _ = _batch_ctx(pr, members)
    n = len(members)
    scheds = [m[2] for m in members]
    out = []
    for waves in (0, 4):
        ctx.tune(adj_waves=waves)
        out.append(ctx.loss_grad_batch(range(n), sid, scheds, samples, tg, ckpt=31))
        out.append(ctx.loss_grad(2, sid, scheds[2], samples, tg[2], 64))
    ctx.close()
    assert np.array_equal(out[0][0], out[2][0]) and np.array_equal(out[0][1], out[2][1])
    assert out[1][0] == out[3][0] and np.array_equal(out[1][1], out[3][1])


def test_adjoint_w_forms_give_the_same_gradient():
    """SRS_TUNE_ADJ_W: w written by each reverse step for the next (1) and w
    formed inside the kernel from lambda, staged in LDS for the tile (2),
    give the same losses and gradients bit for bit, batched and single."""
    pr, members, samples = _batch_setup()
    tg, _ = _targets(pr, members, samples)
    ctx, sid, nf, _ = _batch_ctx(pr, members)
    n = len(members)
    scheds = [m[2] for m in members]
    with pytest.raises(S.SrsError, match="w form"):
        ctx.tune(adj_w=3)
    out = []
    for form in (1, 2):
        ctx.tune(adj_w=form)
        out.append(ctx.loss_grad_batch(range(n), sid, scheds, samples, tg, ckpt=31))
        out.append(ctx.loss_grad(2, sid, scheds[2], samples, tg[2], 64))
    ctx.tune(adj_w=0)
    ctx.close()
    assert np.array_equal(out[0][0], out[2][0]) and np.array_equal(out[0][1], out[2][1])
    assert out[1][0] == out[3][0] and np.array_equal(out[1][1], out[3][1])
    assert np.all(np.isfinite(out[0][1])) and np.any(out[0][1] != 0.0)


@pytest.mark.parametrize("levels", [2, 3])
def test_checkpoint_levels_do_not_change_the_single_gradient(levels):
    """A single realization with a multi-level checkpoint plan (its leaf
    segments recomputed through the gated sweep into the segment's slots, the
    levels above through separate sweep and replay launches in the
    realization's own buffers) gives the one-level plan's loss and gradient
    bit for bit."""
    prob, _, _ = _setup()
    th = prob.theta()
    L1, g1 = prob.loss_grad(th)
    prob.ctx.tune(adj_levels=levels)
    Ln, gn = prob.loss_grad(th)
    assert prob.ctx.adjoint_plan()["levels"] == levels
    prob.ctx.tune(adj_levels=0)
    prob.ctx.close()
    assert Ln == L1 and np.array_equal(gn, g1)
