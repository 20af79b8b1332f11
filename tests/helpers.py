"""Shared helpers for the parity tests: building small workloads, running the
HIP path through the C ABI, and an oracle-driven replica of the reference's
RunSimulation loop (cpu/RunSimulation.cpp:402-589)."""
from __future__ import annotations

import math

import numpy as np

import srs_amd as S
from oracle.oracle import RNG_GLIBC, RNG_PHILOX, Fields, Oracle, grid_dict
from srs_amd import workloads as W


def fast_gating(p: S.Params) -> S.Params:
    """Rates high enough that short runs see many RyR and LCC transitions."""
    q = p.copy()
    q.setr(RYR_K_MINUS=40.0, RYR_K_PLUS=0.5)
    return q


def initial_fields(pr, vm=0.0, ryr_open=(0,), lcc_state=6, lccv_state=2, seed=0):
    g = pr.grid
    nr, nl = g.n_ryr, g.n_lcc
    rs = np.zeros(nr, np.int32)
    for i in ryr_open:
        if i < nr:
            rs[i] = 1
    rng = np.random.default_rng(seed)
    R0 = np.log(rng.random(nr + nl))
    return Fields(pr.X0, rs, np.full(nl, lcc_state), np.full(nl, lccv_state), R0,
                  [vm, pr.params.reals[S.R["DEFAULT_CA_NSR"]]])


def gpu_context(preps, fields, params=None, seeds=None, order=S.ORDER_RCM, record_capacity=0, share_mesh=False,
                group=1):
    """A context with one realization per prepared workload."""
    ctx = S.Context(0)
    ctx.tune(group=group)
    mids = []
    for k, pr in enumerate(preps):
        if share_mesh and k > 0:
            mids.append(mids[0])
        elif order == "block":
            mids.append(ctx.add_mesh_grid(pr.grid, pr.aij, pr.nij, perm=W.kuhn_block_order(pr.grid)))
        else:
            mids.append(ctx.add_mesh_grid(pr.grid, pr.aij, pr.nij, order=order))
    for k, pr in enumerate(preps):
        g = pr.grid
        p = params[k] if params is not None else pr.params
        seed = seeds[k] if seeds is not None else 1000 + k
        ctx.add_realization(mids[k], g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=seed, stream=k)
    ctx.finalize(record_capacity)
    for k, pr in enumerate(preps):
        F = fields[k]
        ctx.state_set(k, F.X)
        ctx.channels_set(k, F.ryr_state, F.lcc_state, F.lccv_state, F.R)
        ctx.globals_set(k, F.gv[0], F.gv[1])
    return ctx


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), 1e-300)
    d = np.abs(a - b) / den
    d[(a == b)] = 0.0
    return float(d.max()) if d.size else 0.0


def output_due(t, dt_last, iv):
    return math.floor(t / iv) > math.floor((t - dt_last) / iv)


def oracle_loop(O: Oracle, pr, p: S.Params, F: Fields, key: int, stream: int, max_steps=10 ** 7, rng_kind=RNG_PHILOX,
                t_lcc_close=None, gpu_layout=False, on_top=None):
    """RunSimulation's main loop with the oracle's step (Philox gating).
    Returns the records written inside and after the loop (the t = 0 record
    before the loop is the caller's).  t_lcc_close: T_LCC_Close of an LCC
    triggered at t = 0 (cpu/RunSimulation.cpp:459-463, 525-528).  on_top(t,
    dt_last, F): called at the top of every main-loop iteration after the
    protocols, where the reference writes its grid files (:545-550)."""
    G = grid_dict(pr.grid, pr.aij, pr.nij)
    P = p.reals
    Iv = p.ints
    dt = float(P[S.R["T_STEP"]])
    t, dt_last = 0.0, 0.0
    recs = []
    n_open = int(F.ryr_state.sum())
    ca_max = 0.0
    lcc_open = int(Iv[S.I["FLAG_LCC"]])
    fid = bool(Iv[S.I["FLAG_FIDELITY_SIM"]])
    step = 0
    ir = int(Iv[S.I["INIT_RYR"]])
    ryr_out = bool(Iv[S.I["FLAG_RYR_OUT"]])
    if Iv[S.I["GATING_FFWD"]]:  # cpu/RunSimulation.cpp:469-519
        while True:
            O.gating(G, P, F, dt, rng_kind=rng_kind, key=key, stream=stream, step=step)
            step += 1
            t += dt
            dt_last = dt
            n_open = int(F.ryr_state.sum())
            n_lcc = int(np.sum((F.lccv_state == 2) & ((F.lcc_state == 6) | (F.lcc_state == 12))))
            stop = n_open > 0 or n_lcc > 0 or t > P[S.R["T_FINAL"]]
            if not stop and output_due(t, dt_last, P[S.R["OUTPUT_STATES_INTERVAL"]]):
                recs.append(O.record(t, G, F, P, ryr_out=ryr_out, gpu_layout=gpu_layout))
            if Iv[S.I["FLAG_V_CLAMP"]]:
                F.gv[0] = P[S.R["V_CLAMP"]] if t <= P[S.R["T_CLAMP"]] else P[S.R["DEFAULT_VM"]]
            if stop:
                break
            if step > max_steps:
                raise RuntimeError("oracle ffwd loop did not terminate")
    while t <= P[S.R["T_FINAL"]] and (not fid or ((lcc_open or n_open > 0 or ca_max > P[S.R["C_MIN"]])
                                                  and n_open < Iv[S.I["N_RYR_FIDELITY"]])):
        if Iv[S.I["FLAG_LCC"]] and lcc_open and t_lcc_close is not None and t >= t_lcc_close:
            lcc_open = 0
            F.lcc_state[int(Iv[S.I["LCC_INDEX"]])] = 1
        if Iv[S.I["FLAG_V_CLAMP"]]:
            F.gv[0] = P[S.R["V_CLAMP"]] if t <= P[S.R["T_CLAMP"]] else P[S.R["DEFAULT_VM"]]
        if ir != -1 and t < P[S.R["INIT_RYR_HOLD"]] and not F.ryr_state[ir]:
            F.ryr_state[ir] = 1
        if on_top is not None:
            on_top(t, dt_last, F)
        if output_due(t, dt_last, P[S.R["OUTPUT_STATES_INTERVAL"]]):
            n_open = int(F.ryr_state.sum())
            Ca = F.X[: G["n_ele"]]
            ca_max = 0.0
            for e in G["ryr_ele"]:
                if Ca[e] > ca_max:
                    ca_max = Ca[e]
            recs.append(O.record(t, G, F, P, ryr_out=ryr_out, gpu_layout=gpu_layout))
        O.steps(G, P, F, 1, dt, gating=not Iv[S.I["FLAG_NO_GATING"]], rng_kind=rng_kind, key=key, stream=stream,
                step0=step)
        step += 1
        t += dt
        dt_last = dt
        if step > max_steps:
            raise RuntimeError("oracle loop did not terminate")
    recs.append(O.record(t, G, F, P, ryr_out=ryr_out, gpu_layout=gpu_layout))
    return recs, t, step


def spark_stats(states: np.ndarray, first: int):
    """Statistics of one realization's RyR trajectory (rows = steps, columns =
    RyRs): the step at which RyR `first` (open at the start) first closes
    (rows + 1 if it never does), the number of RyR openings, the number of
    distinct RyRs that were ever open, and the largest number open at once."""
    st = states > 0
    closed = np.nonzero(~st[:, first])[0]
    t_close = int(closed[0]) if closed.size else st.shape[0] + 1
    openings = int(np.sum(st[1:] & ~st[:-1]))
    return t_close, openings, int(st.any(axis=0).sum()), int(st.sum(axis=1).max())


# The spark-box distribution test's workload: the 56,250-element spark box,
# its centre RyR (24 of the 7x7 cluster) open at t = 0, 1,500 steps, RyR
# kinetics fast enough that the cluster recruits within that window.
SPARK_DIST = dict(n_steps=1500, first=24, k_plus=1.0, k_minus=20.0, n_dev=1024, n_ref=128, key=38376427)


def spark_dist_fields(pr, k):
    """Initial fields of realization k: equilibrium state, RyR `first` open,
    LCCs closed, next-transition thresholds log(u) from seed 1000 + k."""
    g = pr.grid
    rs = np.zeros(g.n_ryr, np.int32)
    rs[SPARK_DIST["first"]] = 1
    R = np.log(np.random.default_rng(1000 + k).random(g.n_ryr + g.n_lcc))
    return Fields(pr.X0, rs, np.ones(g.n_lcc, np.int32), np.full(g.n_lcc, 2, np.int32), R,
                  [0.0, pr.params.reals[S.R["DEFAULT_CA_NSR"]]])


def spark_dist_check(dev, ref, q, n_steps, alpha=0.01):
    """The distribution test's verdicts: two-sample KS and Welch t p-values of
    each statistic (device vs reference), and the one-sample KS p-values of
    both sets' first closing times against their exact law, a geometric
    distribution of parameter q censored at n_steps (the per-integer
    distance, with the continuous KS tail, which is conservative for a
    discrete law).  Returns (names, p-values, all above alpha)."""
    from scipy import stats

    names, pv = [], []
    for c, name in enumerate(("first closing", "openings", "RyRs recruited", "peak open")):
        names += [f"{name} KS", f"{name} t"]
        pv.append(stats.ks_2samp(dev[:, c], ref[:, c]).pvalue)
        a, b = dev[:, c].astype(float), ref[:, c].astype(float)
        pv.append(1.0 if a.var() == 0 and b.var() == 0 and a.mean() == b.mean() else
                  stats.ttest_ind(a, b, equal_var=False).pvalue)
    k = np.arange(1, n_steps)  # a closing is seen at rows 1 .. n_steps - 1
    law = 1.0 - (1.0 - q) ** k
    for nm, arr in (("device", dev), ("reference", ref)):
        t = arr[:, 0]
        emp = np.searchsorted(np.sort(t), k, side="right") / t.size
        names.append(f"{nm} first closing law")
        pv.append(float(stats.kstwo.sf(np.abs(emp - law).max(), t.size)))
    return names, pv, min(pv) > alpha


_SPARK = {}


def oracle_spark_init(G, reals, X0):
    """Pool initializer for oracle_glibc_spark: the spark box's step arrays
    (grid_dict), parameter vector and equilibrium state, handed over once per
    worker so the workers never load the HIP library."""
    _SPARK.update(G=G, reals=reals, X0=X0)


def oracle_glibc_spark(args):
    """One spark-box run with the oracle's step and the reference's own
    glibc rand() gating (a process-pool worker, after oracle_spark_init):
    n_steps steps from the equilibrium state with RyR `first` open, srand(seed)
    first; returns spark_stats of the RyR trajectory."""
    import ctypes

    seed, n_steps, first = args
    G, P = _SPARK["G"], _SPARK["reals"]
    nr, nl = len(G["ryr_ele"]), len(G["lcc_ele"])
    rs = np.zeros(nr, np.int32)
    rs[first] = 1
    R0 = np.log(np.random.default_rng(seed).random(nr + nl))
    F = Fields(_SPARK["X0"], rs, np.ones(nl, np.int32), np.full(nl, 2, np.int32), R0,
               [0.0, P[S.R["DEFAULT_CA_NSR"]]])
    ctypes.CDLL(None).srand(ctypes.c_uint(seed))
    O = Oracle()
    dt = float(P[S.R["T_STEP"]])
    rows = np.zeros((n_steps, nr), np.int32)
    rows[0] = F.ryr_state
    for k in range(1, n_steps):
        O.steps(G, P, F, 1, dt, gating=True, rng_kind=RNG_GLIBC)
        rows[k] = F.ryr_state
    return spark_stats(rows, first)
