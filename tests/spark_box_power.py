"""Power of the spark-box distribution test (a manual check on the CPU,
not collected by pytest): python -m tests.spark_box_power [n_dev] [procs]

The device's gating is bit-exact with the oracle's Philox restatement (the
gating trajectory tests), so oracle Philox runs with the device test's keys,
streams and initial thresholds are the device test's realizations.  This runs
the device test's check (spark_dist_check) on them against the same 128 glibc
oracle runs, at the test's kinetics (must pass) and with a mistake planted in
the oracle: k_open or k_close 1.5x (must fail).  About 4.5 s of one core per
run at the test's 1,500 steps; n_dev = 1024 takes about 30 minutes on 8
cores."""
import multiprocessing as mp
import sys
import time

import numpy as np

import srs_amd as S
from oracle.oracle import RNG_PHILOX, Oracle, grid_dict
from srs_amd import workloads as W
from tests import helpers as H


def philox_run(args):
    k, k_plus, k_minus = args
    c = H.SPARK_DIST
    G, P = H._SPARK["G"], H._SPARK["reals"].copy()
    P[S.R["RYR_K_PLUS"]], P[S.R["RYR_K_MINUS"]] = k_plus, k_minus
    F = H._SPARK["fields"](k)
    O = Oracle()
    dt = float(P[S.R["T_STEP"]])
    rows = np.zeros((c["n_steps"], len(G["ryr_ele"])), np.int32)
    rows[0] = F.ryr_state
    for s in range(1, c["n_steps"]):
        O.steps(G, P, F, 1, dt, gating=True, rng_kind=RNG_PHILOX, key=c["key"], stream=k, step0=s - 1)
        rows[s] = F.ryr_state
    return H.spark_stats(rows, c["first"])


def _init(G, reals, X0, pr):
    H.oracle_spark_init(G, reals, X0)
    H._SPARK["fields"] = lambda k: H.spark_dist_fields(pr, k)


def main():
    n_dev = int(sys.argv[1]) if len(sys.argv) > 1 else H.SPARK_DIST["n_dev"]
    procs = int(sys.argv[2]) if len(sys.argv) > 2 else 8
    c = H.SPARK_DIST
    pr = W.prepare_spark()
    p = pr.params.copy().setr(RYR_K_PLUS=c["k_plus"], RYR_K_MINUS=c["k_minus"])
    dt = p.reals[S.R["T_STEP"]]
    q = 1.0 - np.exp(-c["k_minus"] * dt)
    init = (grid_dict(pr.grid, pr.aij, pr.nij), np.array(p.reals), pr.X0, pr)
    with mp.get_context("spawn").Pool(procs, initializer=_init, initargs=init) as pool:
        t0 = time.time()
        seeds = [93843 + 38376427 + 83 * k for k in range(c["n_ref"])]
        ref = np.array(pool.map(H.oracle_glibc_spark, [(s, c["n_steps"], c["first"]) for s in seeds]))
        print(f"reference (glibc): {c['n_ref']} runs in {time.time() - t0:.0f} s, means {ref.mean(axis=0)}",
              flush=True)
        for label, kp, km, want in (("as the device", c["k_plus"], c["k_minus"], True),
                                    ("k_close 1.5x", c["k_plus"], 1.5 * c["k_minus"], False),
                                    ("k_open 1.5x", 1.5 * c["k_plus"], c["k_minus"], False)):
            t0 = time.time()
            dev = np.array(pool.map(philox_run, [(k, kp, km) for k in range(n_dev)]))
            names, pv, ok = H.spark_dist_check(dev, ref, q, c["n_steps"])
            worst = int(np.argmin(pv))
            print(f"{label}: {'passes' if ok else 'fails'} (expected {'pass' if want else 'fail'}); smallest p "
                  f"{pv[worst]:.2e} ({names[worst]}); means {dev.mean(axis=0)}; {time.time() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
