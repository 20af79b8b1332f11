"""Spark time-step benchmark on one MI355X.

Default workload (the plain GPU drop-in): the dyad box -- 2 x 2 x 1 um
cytosol, 200 nm T-tubule, 1.0 x 1.0 x 0.04 um jSR, h = 20 nm Kuhn tets
(2,952,000 elements), 10x10 RyR lattice at 70% occupancy (70 RyRs), 8 LCCs,
FP64, one realization.  A step is one ExplicitStep (fused sweep + gating) of
that realization with its inputs resident in HBM.

Printed as one JSON line:
  value        simulated spark time per wall second (spark-ms/s), plain path
  roofline     the fused sweep kernel against MI355X HBM bandwidth: algorithmic
               bytes per launch (161 B per element: 6 FP64 species read and
               written, a 4 x int32 + 5 x FP64 operator row, 9 B of per-element
               constants) / mean kernel time from HIP events on the context's
               stream
  cpu_baseline the reference's own cpu/RunSimulation.cpp ExplicitStep
               (oracle/_ref) on the same mesh, 1 host core, bounded sample
  batched      the ensemble workload: many independent spark realizations on
               h = 40/30/25 nm meshes in one launch per step, against the
               plain path run one realization at a time over the same mix of
               meshes
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "spark-ms simulated/s & HBM GB/s % peak; batched vs plain, 1 MI355X"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


BYTES_PER_ELEMENT = 161


def host_cores():
    """Cores this process may run on (what nproc prints)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def algorithmic_bytes(n_ele):
    """HBM bytes one explicit step must move per element: 6 FP64 species read
    and written (96 B), the ELL row of 4 int32 + 5 FP64 (56 B) and 9 B of
    per-element constants (the SL buffer capacity and a domain/troponin flag
    byte).  Neighbour gathers are counted as cache hits."""
    return BYTES_PER_ELEMENT * int(n_ele)


def live_bytes(domain, tropc):
    """The bytes the sweep actually has to move: species that are
    identically zero (TnC outside troponin-bearing cytosol, CSQ outside the
    jSR) are neither read nor written."""
    jsr = np.asarray(domain) != 0
    live = 4 + ((~jsr) & (np.asarray(tropc) != 0)) + jsr
    return int(np.sum(16 * live + 65))


def profiled_traffic(path):
    """HBM bytes per sweep launch from a committed rocprofv3 summary: the line
    'traffic_bytes_per_launch: N' (FETCH_SIZE + WRITE_SIZE, calibrated)."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return int(float(line.split(":", 1)[1]))
    return None


def setup_plain(pr, order="block", tune=None, ext=None):
    import srs_amd as S
    from srs_amd.workloads import initial_channels, kuhn_block_order

    ctx = S.Context(0)
    tune = dict(tune or {})
    ctx.tune(**{k: tune[k] for k in ("tile", "plane_pad", "compact", "group") if k in tune})
    if order == "block":
        mid = ctx.add_mesh_grid(pr.grid, pr.aij, pr.nij, perm=kuhn_block_order(pr.grid))
    else:
        mid = ctx.add_mesh_grid(pr.grid, pr.aij, pr.nij, order=S.ORDER_RCM if order == "rcm" else S.ORDER_AS_GIVEN)
    g = pr.grid
    ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, pr.params, seed=38376427, stream=0)
    if ext:
        ctx.extensions(0, **ext)
    ctx.finalize()
    rs, ls, lvs, R = initial_channels(pr)
    ctx.state_set(0, pr.X0)
    ctx.channels_set(0, rs, ls, lvs, R)
    ctx.globals_set(0, 0.0, pr.params.reals[16])
    ctx.tune(**{k: tune[k] for k in ("nt_geometry", "nt_state", "occupancy", "lds", "gated_sweep") if k in tune})
    return ctx


def bench_plain(pr, steps, warmup, order, tune, ext=None):
    import srs_amd as S

    ctx = setup_plain(pr, order, tune, ext)
    ctx.explicit_step(warmup, S.STEP_GRAPH)
    ctx.prepare(steps, S.STEP_GRAPH)  # graphs captured outside the timed region
    ctx.sync()
    t0 = time.perf_counter()
    ctx.explicit_step(steps, S.STEP_GRAPH)
    ctx.sync()
    wall = time.perf_counter() - t0
    total_ms, _ = ctx.last_timing()
    sweep_ms, gate_ms = ctx.time_sweep(50)
    ctx.close()
    return wall, total_ms, sweep_ms, gate_ms


def bench_cpu(pr, budget_s=15.0):
    from oracle.oracle import Fields, Oracle, RefStep, grid_dict
    from srs_amd.workloads import initial_channels

    G = grid_dict(pr.grid, pr.aij, pr.nij)
    rs, ls, lvs, R = initial_channels(pr)
    F = Fields(pr.X0, rs, ls, lvs, R, [0.0, pr.params.reals[16]])
    p = pr.params
    if RefStep.available():
        runner, kind = RefStep(), "reference"
        run = lambda n: runner.steps(G, p.ints, p.reals, F, n, p.reals[1])  # noqa: E731
    else:
        runner, kind = Oracle(), "port"
        run = lambda n: runner.steps(G, p.reals, F, n, p.reals[1], gating=True)  # noqa: E731
    t0 = time.perf_counter()
    run(2)
    per = (time.perf_counter() - t0) / 2
    n = max(3, int(budget_s / per))
    t0 = time.perf_counter()
    run(n)
    wall = time.perf_counter() - t0
    return n, wall, kind


def _cpu_ensemble_worker(args):
    """One host process of the CPU ensemble: the reference's own step (or the
    oracle when it is not built) with gating on one ensemble mesh for about
    `budget` seconds; returns (simulated ms, wall s, kind)."""
    m, budget = args
    from oracle.oracle import Fields, Oracle, RefStep, grid_dict
    from srs_amd import workloads as W
    import srs_amd as S

    g = S.box_grid(W.ensemble_mesh_spec(m))
    aij, nij = S.assemble_operator(g)
    p = W.spark_params()
    dt = W.stable_dt(aij, 0.22)
    p.setr(T_STEP=dt, INIT_RYR_HOLD=dt)
    X0, _ = S.equilibrium_state(g, p.reals)
    rs = np.zeros(g.n_ryr, np.int32)
    rs[g.n_ryr // 2] = 1
    R = np.log(np.random.default_rng(m).random(g.n_ryr + g.n_lcc) * (1 - 2**-53) + 2**-54)
    F = Fields(X0, rs, np.ones(g.n_lcc, np.int32), np.full(g.n_lcc, 2, np.int32), R, [0.0, p.reals[16]])
    G = grid_dict(g, aij, nij)
    if RefStep.available():
        runner, kind = RefStep(), "reference"
        run = lambda n: runner.steps(G, p.ints, p.reals, F, n, dt)  # noqa: E731
    else:
        runner, kind = Oracle(), "port"
        run = lambda n: runner.steps(G, p.reals, F, n, dt, gating=True)  # noqa: E731
    t0 = time.perf_counter()
    run(2)
    n = max(3, int(budget / ((time.perf_counter() - t0) / 2)))
    t0 = time.perf_counter()
    run(n)
    return n * dt, time.perf_counter() - t0, kind


def bench_cpu_ensemble(procs=16, budget_s=8.0):
    """The batched workload's CPU reference: `procs` host processes at once
    (one realization each, the three ensemble meshes in turn, as an MPI run
    with one rank per core would), each timing the reference's own step with
    gating.  Aggregate simulated spark-ms per wall second."""
    import multiprocessing as mp

    procs = max(1, min(procs, os.cpu_count() or 1))
    with mp.get_context("spawn").Pool(procs) as pool:  # fresh interpreters: no GPU state is copied
        res = pool.map(_cpu_ensemble_worker, [(w % 3, budget_s) for w in range(procs)])
    return dict(value=sum(ms / wall for ms, wall, _ in res), unit="spark-ms simulated/s", cores=procs,
                kind=res[0][2], sample=f"{procs} processes, one ensemble realization each (h=40/30/25 nm meshes in "
                                      f"turn), about {budget_s:.0f} s of ExplicitStep + gating each")


def bench_batched(n_real, steps, warmup, order, tune=None, n_tune=512):
    """The batched leg and its plain-sequential reference.  Unless `tune` is
    given, the batched sweep is autotuned (tile, realization group, cache
    hints, LDS staging) on the first `n_tune` members of the same ensemble,
    and the plain path on each of the three meshes on its own."""
    import srs_amd as S
    from srs_amd import workloads as W
    from srs_amd.autotune import autotune

    trials = None
    if tune is None:
        sample = list(range(min(n_real, n_tune)))
        tune, trials = autotune(lambda cfg: W.ensemble_context(n_real, order=order, tune=cfg, only=sample)[0],
                                tiles=(128, 256), compact=(1,), occupancy=(0, 6), groups=(1, 2, 4))
        log(f"batched autotune: {tune}")
    ctx, dts, n_ele, nbytes = W.ensemble_context(n_real, order=order, tune=tune, log=log)
    log(f"batched: {n_real} realizations, {n_ele} elements, group {tune.get('group', 1)}")
    ctx.explicit_step(warmup, S.STEP_GRAPH)
    ctx.prepare(steps, S.STEP_GRAPH)  # graphs captured outside the timed region
    ctx.sync()
    t0 = time.perf_counter()
    ctx.explicit_step(steps, S.STEP_GRAPH)
    ctx.sync()
    wall = time.perf_counter() - t0
    sweep_ms, gate_ms = ctx.time_sweep(5)
    ctx.close()
    # The plain path, one realization at a time: one member on each of the
    # three meshes, then the batch's mix of meshes run back to back.
    members = W.ensemble_members(n_real)
    rate, ptunes = {}, {}
    for m in range(3):
        r = next(k for k, mem in enumerate(members) if mem["mesh"] == m)
        ptune, _ = autotune(lambda cfg: W.ensemble_context(n_real, order=order, tune=cfg, only=[r])[0],
                            tiles=(128, 256), compact=(1,), occupancy=(0, 6))
        ptunes[m] = ptune
        pctx, pdts, _, _ = W.ensemble_context(n_real, order=order, tune=ptune, only=[r])
        pctx.explicit_step(200, S.STEP_GRAPH)
        pctx.prepare(2000, S.STEP_GRAPH)
        pctx.sync()
        t0 = time.perf_counter()
        pctx.explicit_step(2000, S.STEP_GRAPH)
        pctx.sync()
        rate[m] = 2000 * float(pdts[0]) / (time.perf_counter() - t0)
        pctx.close()
    seq = n_real / sum(1.0 / rate[mem["mesh"]] for mem in members)
    return dict(wall=wall, sim_ms=float(np.sum(dts)) * steps, sweep_ms=sweep_ms, gate_ms=gate_ms, n_ele=n_ele,
                nbytes=nbytes, plain_rate=seq, plain_per_mesh=[rate[m] for m in range(3)], tune=tune,
                plain_tunes=[ptunes[m] for m in range(3)], trials=trials)


def bench_adjoint(t_sim=5.0, n_points=64, n_samples=50, ckpt=256, adj_waves=0, tune=None):
    """Adjoint workload: the h = 40 nm spark box, t_sim ms with the RyR
    schedule frozen from one gated run, a 64-point line scan along x through
    the cluster centre sampled 50 times, target from theta moved by +-10%.
    Times one loss evaluation and one loss + gradient (checkpoints every
    `ckpt` steps, recomputation)."""
    from srs_amd import workloads as W

    t0 = time.perf_counter()
    prob, info = W.spark_adjoint_problem(t_sim, n_points, n_samples, ckpt, tune=tune)
    prob.ctx.tune(adj_waves=adj_waves)
    t_setup = time.perf_counter() - t0
    th = prob.theta()
    t0 = time.perf_counter()
    Lf = prob.loss(th)
    t_fwd = time.perf_counter() - t0
    t0 = time.perf_counter()
    L, grad = prob.loss_grad(th)
    t_grad = time.perf_counter() - t0
    prob.ctx.close()
    n_steps = info["n_steps"]
    return dict(workload=f"spark_box h=40 nm ({info['n_ele']} elements), {t_sim} ms = {n_steps} steps, RyR schedule "
                         f"frozen, {info['n_f']}-point line scan x {n_samples} samples, checkpoint every {ckpt} steps, "
                         f"{len(th)} parameters",
                n_steps=n_steps, forward_s=t_fwd, loss_and_gradient_s=t_grad, setup_s=t_setup,
                gradient_over_forward=t_grad / t_fwd, us_per_step_forward=1e6 * t_fwd / n_steps,
                us_per_step_adjoint=1e6 * (t_grad - t_fwd) / n_steps, loss=L, loss_check=abs(L - Lf) <= 1e-12 * L,
                grad_norm=float(np.linalg.norm(grad * th)), loss_at_target_parameters=info["loss_at_target"])


def bench_adjoint_batch(n_real=256, n_steps=256, n_points=64, n_samples=8, max_batch=128, adj_waves=0,
                        sweep_tune=None, adj_w=0):
    """Batched adjoint workload: n_real realizations on the dyad-box geometry
    at h = 30 nm, each with its own RyR trajectory (Philox stream r) recorded
    by one gated run of all of them, frozen for the loss; a 64-point line
    scan along x through the cluster centre sampled n_samples times; targets
    from theta moved by +-10%.  Times the batched loss alone and the loss with
    every member's gradient (the library's checkpoint plan, at most max_batch
    members stepped at once).  sweep_tune: sweep tunables set after the set-up
    (nt_geometry, nt_state, occupancy, lds), for the forward and recomputed
    steps; adj_w: the reverse step's w form (0 by launch size, 1 written by
    each step, 2 formed in the kernel)."""
    from srs_amd import workloads as W

    t0 = time.perf_counter()
    ctx, sid, scheds, samples, tgt, p, info = W.dyad_adjoint_batch(n_real, n_steps, n_points, n_samples, max_batch,
                                                                   log=log)
    ctx.tune(adj_waves=adj_waves, adj_w=adj_w, **(sweep_tune or {}))
    t_setup = time.perf_counter() - t0
    idx = __import__("srs_amd").adjoint_theta()
    t0 = time.perf_counter()
    Lf = ctx.loss_grad_batch(range(n_real), sid, scheds, samples, tgt, grad=False, max_batch=max_batch)
    t_fwd = time.perf_counter() - t0
    log(f"adjoint batch: loss of {n_real} members in {t_fwd:.2f} s")
    t0 = time.perf_counter()
    L, grad = ctx.loss_grad_batch(range(n_real), sid, scheds, samples, tgt, max_batch=max_batch)
    t_grad = time.perf_counter() - t0
    plan = ctx.adjoint_plan()
    log(f"adjoint batch: loss and gradients in {t_grad:.2f} s, plan {plan}")
    ctx.close()
    n_ele = info["n_ele"]
    elem_steps = float(n_real) * n_ele * n_steps
    full_steps = int(round(20.0 / info["dt"]))
    return dict(workload=f"dyad_box geometry at h=30 nm ({n_ele} elements), {n_real} realizations with their own "
                         f"recorded RyR trajectories, {n_steps} steps, {info['n_f']}-point line scan x {n_samples} "
                         f"samples, at most {max_batch} realizations at once, {len(idx)} parameters",
                n_real=n_real, n_elements=n_ele, n_steps=n_steps, forward_s=t_fwd, loss_and_gradient_s=t_grad,
                setup_s=t_setup, checkpoint_plan=plan,
                ps_per_element_step_forward=1e12 * t_fwd / elem_steps,
                ps_per_element_step_loss_and_gradient=1e12 * t_grad / elem_steps,
                full_20ms_steps=full_steps,
                full_20ms_loss_and_gradient_s_extrapolated=t_grad * full_steps / n_steps,
                loss_check=bool(np.all(np.abs(L - Lf) <= 1e-12 * np.abs(L))), mean_loss=float(L.mean()),
                mean_grad_norm=float(np.linalg.norm(grad.mean(axis=0) * p.reals[idx])))


def bench_adjoint_full(n_real=8, t_sim=20.0, n_points=64, n_samples=50, per_call=2, adj_waves=0):
    """The batched fitting workload at its full length: n_real members of the
    h = 30 nm dyad geometry, t_sim ms (392,418 steps at 20 ms), each with its
    own recorded trajectory, loss = squared dF/F0 of a 64-point scan sampled
    n_samples times (zero target), loss and all gradients in one call with
    the library's checkpoint plan, per_call members per call.  The cost of
    more members is that of as many more calls."""
    from srs_amd import workloads as W

    dt = float(W.prepare_dyad(h=0.03).params.reals[1])
    n_steps = int(round(t_sim / dt))
    t0 = time.perf_counter()
    ctx, sid, scheds, samples, tgt, p, info = W.dyad_adjoint_batch(n_real, n_steps, n_points, n_samples, 0, log=log,
                                                                   zero_target=True)
    ctx.tune(adj_waves=adj_waves)
    t_setup = time.perf_counter() - t0
    log(f"adjoint full: {n_real} members x {n_steps} steps set up in {t_setup:.1f} s")
    # one call per `per_call` members (each member's result does not depend on
    # which others share its call), a line after each
    t_grad, grads, plans = 0.0, [], []
    for m0 in range(0, n_real, per_call):
        ids = list(range(m0, min(n_real, m0 + per_call)))
        t0 = time.perf_counter()
        _, g = ctx.loss_grad_batch(ids, sid, [scheds[m] for m in ids], samples, tgt[ids])
        dt_call = time.perf_counter() - t0
        t_grad += dt_call
        grads.append(g)
        plans.append(ctx.adjoint_plan())
        log(f"adjoint full: members {ids[0]}-{ids[-1]} in {dt_call:.1f} s, plan {plans[-1]}")
    grad = np.concatenate(grads)
    plan = plans[0]
    ctx.close()
    elem_steps = float(n_real) * info["n_ele"] * n_steps
    log(f"adjoint full: loss and gradients in {t_grad:.1f} s, plan {plan}")
    return dict(workload=f"dyad_box geometry at h=30 nm ({info['n_ele']} elements), {n_real} realizations, "
                         f"{t_sim} ms = {n_steps} steps, {info['n_f']}-point line scan x {n_samples} samples",
                n_real=n_real, n_steps=n_steps, loss_and_gradient_s=t_grad, setup_s=t_setup, checkpoint_plan=plan,
                ps_per_element_step_loss_and_gradient=1e12 * t_grad / elem_steps,
                s_for_256_members=t_grad * 256 / n_real, finite=bool(np.all(np.isfinite(grad))))


def bench_size_scan(steps=400):
    """Cost per element-step against working-set size: the plain path on the
    dyad box at h = 40/30/25/20 nm (tuned sweep, block order), the ensemble
    with 8..256 members stepped together, and the 2048-member ensemble with
    realization subsets of 0/64/128/192 MB stepped block by block."""
    import srs_amd as S
    from srs_amd import workloads as W

    out = []
    tune = dict(tile=512, compact=1, nt_geometry=1, nt_state=0, occupancy=6, lds=1)
    for h in (0.04, 0.03, 0.025, 0.02):
        pr = W.prepare_dyad(h=h)
        ctx = setup_plain(pr, "block", tune)
        ctx.explicit_step(50, S.STEP_GRAPH)
        ctx.prepare(steps, S.STEP_GRAPH)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.explicit_step(steps, S.STEP_GRAPH)
        ctx.sync()
        wall = time.perf_counter() - t0
        ctx.close()
        n = pr.grid.n_ele
        out.append(dict(workload=f"dyad h={h}", n_elements=n, state_mb=96 * n / 1e6, us_per_step=1e6 * wall / steps,
                        ps_per_element_step=1e12 * wall / steps / n))
        log(json.dumps(out[-1]))
    btune = dict(tile=256, compact=1, group=1, nt_geometry=0, nt_state=1, occupancy=6, lds=1)
    for k, mb in ((8, 0), (16, 0), (64, 0), (256, 0), (2048, 0), (2048, 64), (2048, 128), (2048, 192)):
        ctx, _, n_ele, _ = W.ensemble_context(k, tune=btune)
        ctx.tune(subset_mb=mb)
        m = 128
        ctx.explicit_step(16, S.STEP_GRAPH)
        ctx.prepare(m, S.STEP_GRAPH)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.explicit_step(m, S.STEP_GRAPH)
        ctx.sync()
        wall = time.perf_counter() - t0
        ctx.close()
        out.append(dict(workload=f"ensemble {k} members, subsets {mb} MB", n_elements=n_ele, state_mb=96 * n_ele / 1e6,
                        us_per_step=1e6 * wall / m, ps_per_element_step=1e12 * wall / m / n_ele))
        log(json.dumps(out[-1]))
    return out


def bench_fidelity(n_real=512, t_final=20.0, chunk=2000, group=4):
    """Spark-fidelity ensemble with the reference's Fidelity protocol
    (param/walker2014/Fidelity.xml: flag_fidelity, a run ends once 4 RyRs are
    open or once no channel is open and the cleft Ca is below c_min = 1 uM;
    luminal RyR regulation n = 4; states every 0.01 ms), T_FINAL capped at
    t_final: n_real realizations on the h = 40 nm spark box, each with one
    RyR (drawn per realization) open at t = 0 and its own Philox stream,
    stepped in chunks of `chunk` steps until every run has ended.  Timed with
    finished groups left out of the sweep (the default) and with every tile
    swept, on the same seeds."""
    import srs_amd as S
    from srs_amd import workloads as W

    pr = W.prepare_spark()
    g = pr.grid
    p = pr.params.copy()
    p.setr(T_FINAL=t_final, C_MIN=1.0, RYR_K_PLUS=1.107e-4, RYR_PHI_M=4.0, OUTPUT_STATES_INTERVAL=0.01,
           INIT_RYR_HOLD=0.0)
    p.seti(FLAG_FIDELITY_SIM=1, N_RYR_FIDELITY=4, INIT_RYR=-1, FLAG_V_CLAMP=0)
    dt = float(p.reals[S.R["T_STEP"]])
    first = np.random.default_rng(11).integers(0, g.n_ryr, n_real)
    out = {}
    for drop in (1, 0):
        ctx = S.Context(0)
        ctx.tune(group=group)
        mid = ctx.add_mesh_grid(g, pr.aij, pr.nij, perm=W.kuhn_block_order(g))
        for r in range(n_real):
            ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=38376427, stream=r)
        ctx.finalize()
        ctx.tune(drop_finished=drop)
        for r in range(n_real):
            rs, ls, lvs, R = W.initial_channels(pr, open_ryr=int(first[r]), seed=r)
            ctx.state_set(r, pr.X0)
            ctx.channels_set(r, rs, ls, lvs, R)
            ctx.globals_set(r, p.reals[S.R["DEFAULT_VM"]], p.reals[S.R["DEFAULT_CA_NSR"]])
            ctx.loop_set(r, n_open=1)
        ctx.explicit_step(128, S.STEP_PROTOCOL | S.STEP_GRAPH)  # graph capture outside the timed region
        ctx.sync()
        t0 = time.perf_counter()
        steps = 128
        while True:
            ls = [ctx.loop_get(r) for r in range(n_real)]
            n_done = sum(x["done"] for x in ls)
            if n_done == n_real:
                break
            if (steps // chunk) % 20 == 0:
                log(f"  fidelity: {steps} steps, {n_done} of {n_real} runs ended")
            ctx.explicit_step(chunk, S.STEP_PROTOCOL | S.STEP_GRAPH)
            steps += chunk
        wall = time.perf_counter() - t0
        ctx.close()
        t_end = np.array([x["t"] for x in ls])
        sparks = np.array([x["n_ryr_open_last"] >= 4 for x in ls])
        out["drop" if drop else "all"] = dict(wall_s=wall, steps=steps)
        log(f"fidelity ({'finished groups dropped' if drop else 'every tile swept'}): {n_real} runs in {wall:.2f} s")
    return dict(workload=f"spark_box h=40 nm ({g.n_ele} elements), {n_real} realizations, Fidelity protocol "
                         f"(4 open RyRs or no open channel and cleft Ca < 1 uM ends a run), T_FINAL {t_final} ms, "
                         f"groups of {group}, chunks of {chunk} steps",
                n_real=n_real, dt_ms=dt, fidelity=float(sparks.mean()), mean_run_ms=float(t_end.mean()),
                max_run_ms=float(t_end.max()), sim_ms_total=float(t_end.sum()),
                wall_s=out["drop"]["wall_s"], wall_s_every_tile=out["all"]["wall_s"],
                sparks_per_s=n_real / out["drop"]["wall_s"], sparks_per_s_every_tile=n_real / out["all"]["wall_s"],
                spark_ms_per_s=float(t_end.sum()) / out["drop"]["wall_s"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--batch", type=int, default=2048, help="realizations in the batched leg (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--batch-group", type=int, default=0, choices=(0, 1, 2, 4),
                    help="realizations per realization group in the batched leg (0: autotune the batched sweep)")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of CPU baseline work (0: skip)")
    ap.add_argument("--adjoint", type=int, default=1, help="1: time the adjoint workload (loss + gradient)")
    ap.add_argument("--adjoint-batch", type=int, default=256,
                    help="realizations in the batched adjoint leg on the h=30 nm dyad geometry (0: skip)")
    ap.add_argument("--adjoint-batch-steps", type=int, default=256)
    ap.add_argument("--config-physics", type=int, default=1,
                    help="1: also time the plain workload with the configuration's jSR diffusion and SR sites")
    ap.add_argument("--size-scan", action="store_true",
                    help="also scan the sweep's cost per element-step against working-set size (minutes)")
    ap.add_argument("--adj-waves", type=int, default=0, choices=(0, 4),
                    help="register budget of the adjoint reverse step (SRS_TUNE_ADJ_WAVES) in the adjoint legs")
    ap.add_argument("--adjoint-full", type=int, default=0,
                    help="members of a full-length (20 ms) batched loss-and-gradient run (0: skip; minutes)")
    ap.add_argument("--fidelity", type=int, default=0,
                    help="realizations in a spark-fidelity ensemble run to completion (0: skip)")
    ap.add_argument("--order", choices=("block", "rcm", "given"), default="block",
                    help="device element order: Kuhn cube blocks, reverse Cuthill-McKee, or the reference BFS order")
    ap.add_argument("--tune", default="auto",
                    help="'auto' (autotune the sweep on scratch contexts), 'default', or a JSON object of "
                         "tile/plane_pad/nt_geometry/nt_state")
    ap.add_argument("--workload", choices=("dyad", "spark"), default="dyad",
                    help="plain-path workload: the 2.95M-element dyad box (default) or the 56,250-element spark box")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run one device copy of 2 GiB (HBM byte-counter calibration under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the spark step is not sharded across GPUs",
              file=sys.stderr)
        sys.exit(2)

    from srs_amd import workloads as W

    if args.calib_copy:
        import torch

        a = torch.ones(256 << 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 2 GiB read + 2 GiB written")
        del a, b
    t0 = time.perf_counter()
    pr = W.prepare_dyad() if args.workload == "dyad" else W.prepare_spark()
    log(f"{args.workload} box: {pr.grid.n_ele} elements, {pr.grid.n_ryr} RyR, {pr.grid.n_lcc} LCC, "
        f"dt {pr.params.reals[1]:.4e} ms (setup {time.perf_counter() - t0:.1f} s)")
    dt = float(pr.params.reals[1])
    N = pr.grid.n_ele
    trials = None
    if args.tune == "auto":
        from srs_amd.autotune import autotune

        tune, trials = autotune(lambda cfg: setup_plain(pr, args.order, cfg), log=log)
        log(f"autotuned: {tune}")
    elif args.tune == "default":
        tune = {}
    else:
        tune = json.loads(args.tune)
    wall, total_ms, sweep_ms, gate_ms = bench_plain(pr, args.steps, args.warmup, args.order, tune)
    log(f"plain: {args.steps} steps in {wall:.3f} s ({1e3 * wall / args.steps:.4f} ms/step), sweep {sweep_ms * 1e3:.2f} us,"
        f" gating {gate_ms * 1e3:.2f} us")
    value = args.steps * dt / wall
    # The library folds each step's gating into the next step's sweep launch
    # (k_sweep_gated) for a single realization with the LDS sweep and a
    # register budget of 0 or 6; time_sweep then times that kernel.
    eff = {"lds": 1, "occupancy": 0, "gated_sweep": 1, **tune}
    kernel = "k_sweep_gated" if eff["gated_sweep"] and eff["lds"] and eff["occupancy"] in (0, 6) else "k_sweep"
    nbytes = algorithmic_bytes(N)
    achieved = nbytes / (sweep_ms * 1e-3) / 1e9
    traffic = profiled_traffic(os.path.join(ROOT, "profiles", f"k_sweep_{args.workload}_traffic.md"))
    out = {
        "metric": METRIC, "value": value, "unit": "spark-ms simulated/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": (f"dyad_box 2x2x1 um, T-tubule r=100 nm, jSR 1.0x1.0x0.04 um, h=20 nm Kuhn tets "
                                f"({N} elements), {pr.grid.n_ryr} RyR + {pr.grid.n_lcc} LCC, 1 realization, FP64"
                                if args.workload == "dyad" else
                                f"spark_box 1.0x1.0x0.6 um, jSR 0.4x0.4x0.04 um, h=40 nm Kuhn tets ({N} elements), "
                                f"{pr.grid.n_ryr} RyR + {pr.grid.n_lcc} LCC, 1 realization, FP64"),
                   "n_elements": N, "dt_ms": dt, "element_order": args.order, "tuning": tune,
                   "parallelism": "single realization, one GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": kernel, "kernel_ms": sweep_ms, "bytes_per_launch": nbytes,
                     "bytes_per_element": BYTES_PER_ELEMENT,
                     "live_bytes_per_launch": live_bytes(pr.grid.domain, pr.grid.tropc)},
    }
    if args.config_physics:
        # the same workload with the configuration's jSR diffusion (0.35 um^2/ms)
        # and 47 uM SR sites, which the reference cannot express: dt from the
        # per-domain bound
        prc = W.with_config_physics(pr)
        cwall_g, _, csweep_ms, _ = bench_plain(prc, args.steps, args.warmup, args.order, tune, W.CONFIG_SR_SITES)
        cdt = float(prc.params.reals[1])
        out["config_physics"] = {
            "value": args.steps * cdt / cwall_g, "unit": "spark-ms simulated/s", "dt_ms": cdt,
            "ms_per_step": 1e3 * cwall_g / args.steps, "kernel_ms": csweep_ms,
            "physics": "Ca D 0.22 um^2/ms in the cytosol and 0.35 in the jSR, 47 uM SR sites (K_D 0.6 uM) as a "
                       "rapid buffer; model extensions (the reference uses D_Ca in both domains and has no SR "
                       "sites), so no reference timing exists for it"}
        log(f"config physics: {1e3 * cwall_g / args.steps:.4f} ms/step at dt {cdt:.4e} ms")
    if args.cpu_budget > 0:
        n, cwall, kind = bench_cpu(pr, args.cpu_budget)
        out["cpu_baseline"] = {"value": n * dt / cwall, "unit": "spark-ms simulated/s", "cores": 1, "kind": kind,
                               "host_cores": host_cores(),
                               "sample": f"{n} ExplicitStep calls with gating on the same {N}-element mesh "
                                         f"(the reference's step is single-threaded; one core of the host's "
                                         f"{host_cores()})"}
        log(f"cpu {kind}: {n} steps in {cwall:.2f} s")
    if args.batch > 0:
        b = bench_batched(args.batch, args.batch_steps, args.batch_steps, args.order,
                          {**tune, "group": args.batch_group} if args.batch_group else None)
        brate = b["sim_ms"] / b["wall"]
        bach = b["nbytes"] / (b["sweep_ms"] * 1e-3) / 1e9
        out["batched"] = {
            "workload": f"{args.batch} realizations on h=40/30/25 nm spark boxes (56,250/130,680/230,400 elements), "
                        "9-100 RyR each, SR load 0.5-1.25 mM, FP64",
            "value": brate, "unit": "spark-ms simulated/s", "sparks_per_s": brate / 50.0,
            "steps": args.batch_steps, "ms_per_step": 1e3 * b["wall"] / args.batch_steps,
            "group": b["tune"].get("group", 1), "tuning": b["tune"],
            "sweep_ms": b["sweep_ms"], "sweep_gbs": bach, "sweep_frac": bach / HBM_PEAK_GBS,
            "sweep_traffic": {"bytes": profiled_traffic(os.path.join(ROOT, "profiles", "k_sweep_batched_traffic.md")),
                              "profiled_tuning": {"tile": 256, "compact": 1, "group": 1, "nt_geometry": 0,
                                                  "nt_state": 1, "occupancy": 6, "lds": 1}},
            "plain_sequential_value": b["plain_rate"], "speedup_vs_plain_sequential": brate / b["plain_rate"],
            "plain_per_mesh": {"h40": b["plain_per_mesh"][0], "h30": b["plain_per_mesh"][1],
                               "h25": b["plain_per_mesh"][2]},
            "plain_tuning": {"h40": b["plain_tunes"][0], "h30": b["plain_tunes"][1], "h25": b["plain_tunes"][2]},
            "plain_sequential": "the plain path, one realization per context, over the batch's mix of meshes "
                                "(spark-ms/s = n / sum of 1/rate of each member's mesh)",
        }
        log(f"batched: {brate:.3f} spark-ms/s vs plain {b['plain_rate']:.4f} ({brate / b['plain_rate']:.1f}x)")
        if args.cpu_budget > 0:
            ce = bench_cpu_ensemble(16, min(args.cpu_budget, 8.0))
            out["batched"]["cpu_ensemble"] = ce
            out["batched"]["speedup_vs_cpu_ensemble"] = brate / ce["value"]
            log(f"cpu ensemble ({ce['kind']}, {ce['cores']} processes): {ce['value']:.4f} spark-ms/s")
    if args.adjoint:
        a = bench_adjoint(adj_waves=args.adj_waves)
        out["adjoint"] = a
        log(f"adjoint: {a['n_steps']} steps, forward {a['forward_s']:.2f} s, loss+gradient {a['loss_and_gradient_s']:.2f} s")
    if args.adjoint_batch > 0:
        out["adjoint_batch"] = bench_adjoint_batch(args.adjoint_batch, args.adjoint_batch_steps,
                                                   adj_waves=args.adj_waves)
    if args.size_scan:
        out["size_scan"] = bench_size_scan()
    if args.adjoint_full > 0:
        out["adjoint_full"] = bench_adjoint_full(args.adjoint_full, adj_waves=args.adj_waves)
    if args.fidelity > 0:
        out["fidelity"] = bench_fidelity(args.fidelity)
    if trials:
        out["autotune"] = [{**cfg, "step_us": round(ms * 1e3, 2)} for cfg, ms in trials]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
