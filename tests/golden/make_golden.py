"""Writes tests/golden/ref_step_*.npz: inputs and outputs of the reference's
own ExplicitStep + UpdateChannelGating (cpu/RunSimulation.cpp:591-674, built
unmodified from the reference sources into oracle/_ref by oracle/build.py).

Cases (small Kuhn boxes from the product's mesher; every array the step reads
is stored, so the fixture does not depend on the mesher):
  replay_tt     T-tubule box, RyR states replayed from a random schedule,
                gating off, no-flux walls; fields after 100, 200 and 400 steps
  replay_bound  plain box with flag_noflux off (Boundary_Rates and BOUND-face
                diagonal terms), replayed schedule; fields after 150 and 300
  gating_glibc  T-tubule box, fast coupled RyR gating and LCC gating with
                glibc rand() after srand(4321); fields, channel states and
                residuals after 1500 steps (pins the oracle's glibc mode; the
                product draws from Philox, so only the replay cases are
                device fixtures)

Run from the repository root after `python oracle/build.py`:
    python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import srs_amd as S  # noqa: E402
from oracle.oracle import RefStep, grid_dict  # noqa: E402
from srs_amd import workloads as W  # noqa: E402
from tests.helpers import fast_gating, initial_fields  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def _save(name, G, ints, reals, F0, dt, **outs):
    arrs = {k: v for k, v in G.items() if k != "n_ele" and v is not None}
    np.savez_compressed(os.path.join(OUT, f"ref_step_{name}.npz"), n_ele=G["n_ele"], ints=ints, reals=reals, dt=dt,
                        X0=F0.X, ryr_state0=F0.ryr_state, lcc_state0=F0.lcc_state, lccv_state0=F0.lccv_state,
                        R0=F0.R, gv0=np.asarray(F0.gv, np.float64), **arrs, **outs)


def main():
    if not RefStep.available():
        sys.exit("oracle/_ref is not built (python oracle/build.py with the reference sources present)")
    ref = RefStep()

    # replay, T-tubule box
    pr = W.prepare(W.tiny_box_spec(tt=True), W.spark_params())
    G = grid_dict(pr.grid, pr.aij, pr.nij)
    p = pr.params.copy().seti(FLAG_NO_GATING=1)
    F0 = initial_fields(pr, seed=11)
    sch = (np.random.default_rng(12).random((400, pr.grid.n_ryr)) < 0.4).astype(np.int32)
    dt = float(p.reals[1])
    outs, F, done = {}, F0.copy(), 0
    for n in (100, 200, 400):
        F = ref.steps(G, p.ints, p.reals, F, n - done, dt, schedule=sch[done:n])
        done = n
        outs[f"X_{n}"] = F.X.copy()
    _save("replay_tt", G, p.ints, p.reals, F0, dt, schedule=sch, **outs)

    # replay with boundary rates (flag_noflux off)
    spec = W.tiny_box_spec(nx=5, ny=4, nz=3, tt=False)
    g = S.box_grid(spec, no_flux=False)
    aij, nij = S.assemble_operator(g)
    p = W.spark_params().setr(T_STEP=1e-6)
    X0, bnd = S.equilibrium_state(g, p.reals)
    br = S.boundary_rates(g, p.reals, bnd)
    G = grid_dict(g, aij, nij, br)
    p.seti(FLAG_NO_GATING=1, FLAG_NO_FLUX_BOUNDARY=0, INIT_RYR=0)  # the initial RyR must exist on this mesh
    pr2 = W.Prepared(g, aij, nij, X0, bnd, p)
    F0 = initial_fields(pr2, seed=13)
    sch = (np.random.default_rng(14).random((300, g.n_ryr)) < 0.5).astype(np.int32)
    outs, F, done = {}, F0.copy(), 0
    for n in (150, 300):
        F = ref.steps(G, p.ints, p.reals, F, n - done, 1e-6, schedule=sch[done:n])
        done = n
        outs[f"X_{n}"] = F.X.copy()
    _save("replay_bound", G, p.ints, p.reals, F0, 1e-6, schedule=sch, **outs)

    # gating with glibc rand()
    G = grid_dict(pr.grid, pr.aij, pr.nij)
    p = fast_gating(pr.params)
    p.setr(RYR_A_STAR=0.7, RYR_PHI_M=4.0, V_CLAMP=0.0)
    F0 = initial_fields(pr, seed=15)
    dt = float(p.reals[1])
    ctypes.CDLL(None).srand(4321)
    F = ref.steps(G, p.ints, p.reals, F0.copy(), 1500, dt)
    _save("gating_glibc", G, p.ints, p.reals, F0, dt, srand_seed=np.int64(4321), n_steps=np.int64(1500),
          X_1500=F.X, ryr_state=F.ryr_state, lcc_state=F.lcc_state, lccv_state=F.lccv_state, R=F.R)
    print("wrote", sorted(f for f in os.listdir(OUT) if f.endswith(".npz")))


if __name__ == "__main__":
    main()
