"""Golden vectors of the reference's own step (tests/golden/ref_step_*.npz,
written by tests/golden/make_golden.py from cpu/RunSimulation.cpp compiled
unmodified): the oracle reproduces every case bit for bit (CPU, no reference
sources needed), and the device reproduces the replay cases bit for bit
through the C ABI (-m gpu), so the HIP path is checked against the
reference's outputs directly, not only through the oracle."""
import ctypes
import os

import numpy as np
import pytest

import srs_amd as S
from oracle.oracle import RNG_GLIBC, Fields, Oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REPLAY = ("replay_tt", "replay_bound")


def _load(name):
    with np.load(os.path.join(HERE, f"ref_step_{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _grid(d):
    G = {k: d[k] for k in ("aij", "nij", "domain", "tropc", "V0", "tt_area", "ryr_ele", "ryr_jsr", "ryr_nb",
                           "lcc_ele")}
    G["n_ele"] = int(d["n_ele"])
    G["br"] = d.get("br")
    return G


def _fields(d):
    return Fields(d["X0"].copy(), d["ryr_state0"].copy(), d["lcc_state0"].copy(), d["lccv_state0"].copy(),
                  d["R0"].copy(), list(d["gv0"]))


def _checkpoints(d):
    return sorted(int(k[2:]) for k in d if k.startswith("X_"))


@pytest.mark.parametrize("name", REPLAY)
def test_oracle_reproduces_reference_replay(name):
    d = _load(name)
    G, F = _grid(d), _fields(d)
    done = 0
    for n in _checkpoints(d):
        F = Oracle().steps(G, d["reals"], F, n - done, float(d["dt"]), gating=False, schedule=d["schedule"][done:n])
        done = n
        assert np.array_equal(F.X, d[f"X_{n}"]), (name, n)


def test_oracle_reproduces_reference_glibc_gating():
    d = _load("gating_glibc")
    G, F = _grid(d), _fields(d)
    ctypes.CDLL(None).srand(int(d["srand_seed"]))
    F = Oracle().steps(G, d["reals"], F, int(d["n_steps"]), float(d["dt"]), gating=True, rng_kind=RNG_GLIBC)
    assert np.array_equal(F.ryr_state, d["ryr_state"]) and np.array_equal(F.lcc_state, d["lcc_state"])
    assert np.array_equal(F.lccv_state, d["lccv_state"]) and np.array_equal(F.R, d["R"])
    assert np.array_equal(F.X, d["X_1500"])
    assert np.any(F.ryr_state != d["ryr_state0"]) or np.any(F.lcc_state != d["lcc_state0"])


@pytest.mark.gpu
@pytest.mark.parametrize("order", [S.ORDER_AS_GIVEN, S.ORDER_RCM])
@pytest.mark.parametrize("name", REPLAY)
def test_device_reproduces_reference_replay(name, order):
    d = _load(name)
    p = S.Params(d["ints"].copy(), d["reals"].copy())
    print(f"golden {name} order={order}", flush=True)
    with S.Context(0) as ctx:
        mid = ctx.add_mesh(d["aij"], d["nij"], d["domain"], d["tropc"], d["V0"], d["tt_area"], br=d.get("br"),
                           order=order)
        ctx.add_realization(mid, d["ryr_ele"], d["ryr_jsr"], d["ryr_nb"], d["lcc_ele"], p, seed=1, stream=0)
        ctx.finalize()
        ctx.state_set(0, d["X0"])
        ctx.channels_set(0, d["ryr_state0"], d["lcc_state0"], d["lccv_state0"], d["R0"])
        ctx.globals_set(0, float(d["gv0"][0]), float(d["gv0"][1]))
        done = 0
        for n in _checkpoints(d):
            ctx.replay(0, d["schedule"][done:n])
            done = n
            assert np.array_equal(ctx.state_get(0), d[f"X_{n}"]), (name, n)
