"""The allocation-free replay entry point (needs an MI355X):
srs_explicit_step_replay_static, called through the C ABI, against
srs_explicit_step_replay on the same inputs -- bit-identical fields, channel
states and loop state on two calls in a row, the second with a smaller
schedule -- and its size limit."""
import ctypes as C

import numpy as np
import pytest

import srs_amd as S
from srs_amd import _abi
from srs_amd import workloads as W
from tests.helpers import gpu_context, initial_fields

pytestmark = pytest.mark.gpu

STATIC_STATES = 1 << 21


def _static_replay(ctx, r, sch):
    sch = np.ascontiguousarray(sch, np.int32)
    return _abi.lib().srs_explicit_step_replay_static(ctx._h, r, int(sch.shape[0]), sch.ctypes.data_as(
        C.POINTER(C.c_int32)))


@pytest.mark.parametrize("case", ["tiny", "spark"])
def test_static_replay_equals_replay(case):
    pr = (W.prepare(W.tiny_box_spec(tt=True), W.spark_params()) if case == "tiny" else W.prepare_spark())
    nr = pr.grid.n_ryr
    p = pr.params.copy().seti(FLAG_NO_GATING=1)
    F = initial_fields(pr, ryr_open=(0,), seed=3)
    rng = np.random.default_rng(8)
    first = (rng.random((300, nr)) < 0.3).astype(np.int32)
    second = (rng.random((70, nr)) < 0.3).astype(np.int32)  # smaller than the first
    out = []
    for static in (False, True):
        with gpu_context([pr], [F], params=[p], order="block") as ctx:
            got = []
            for sch in (first, second):
                if static:
                    assert _static_replay(ctx, 0, sch) == 0
                else:
                    ctx.replay(0, sch)
                got.append((ctx.state_get(0), ctx.channels_get(0), ctx.loop_get(0)))
            out.append(got)
    for (Xa, cha, la), (Xb, chb, lb) in zip(*out):
        assert np.array_equal(Xa, Xb)
        assert all(np.array_equal(u, v) for u, v in zip(cha, chb))
        assert la == lb
    assert out[1][1][2]["step"] == 370 and not np.array_equal(out[1][0][0], out[1][1][0])


def test_static_replay_size_limit():
    pr = W.prepare(W.tiny_box_spec(tt=True), W.spark_params())
    nr = pr.grid.n_ryr
    F = initial_fields(pr)
    p = pr.params.copy().seti(FLAG_NO_GATING=1)
    with gpu_context([pr], [F], params=[p]) as ctx:
        X0 = ctx.state_get(0)
        too_many = STATIC_STATES // nr + 1
        sch = np.zeros((too_many, nr), np.int32)
        assert _static_replay(ctx, 0, sch) != 0
        assert "static schedule buffer" in _abi.lib().srs_last_error().decode()
        assert np.array_equal(ctx.state_get(0), X0)  # nothing ran
        ok = np.zeros((STATIC_STATES // nr, nr), np.int32)[:50]
        assert _static_replay(ctx, 0, ok) == 0
