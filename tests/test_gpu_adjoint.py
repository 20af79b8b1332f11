"""Adjoint of the frozen-channel step (needs an MI355X): gradients of a line
scan loss against central finite differences of the same device forward
run, checkpoint-interval independence, the torch.autograd wrapper of the
fitting harness, and conservation of total Ca in the device step."""
import numpy as np
import pytest

import srs_amd as S
from srs_amd import workloads as W
from srs_amd.fit import LinescanProblem, linescan_loss_fn

pytestmark = pytest.mark.gpu

N_STEPS, N_SAMPLES = 600, 20


def _setup(perturb=0.1, seed=3):
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
    d.offset1, d.offset2 = float(rc[1]), float(rc[2])
    for k in range(3):
        d.domain[k] = float(hi[k] - lo[k]) + 0.08
        d.var[k] = (0.001, 0.003, 0.001)[k]
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


def test_adjoint_gradient_matches_central_differences():
    """Every component: |theta g - theta g_fd| <= 1e-5 relative (components
    three orders below the largest are held to 1e-5 of the largest)."""
    prob, _, _ = _setup()
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
