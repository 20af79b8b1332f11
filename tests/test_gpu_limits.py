"""Limits and error paths of the device layer (needs an MI355X for a
context; every case is rejected by host-side checks before any launch),
plus the smallest mesh and the largest channel counts stepped against the
oracle."""
import numpy as np
import pytest

import srs_amd as S
from oracle.oracle import RNG_PHILOX, Fields, Oracle, grid_dict
from srs_amd import workloads as W
from tests.helpers import fast_gating, initial_fields, rel_err

pytestmark = pytest.mark.gpu


def _tiny():
    return W.prepare(W.tiny_box_spec(), W.spark_params())


def _ctx_with_mesh(pr):
    ctx = S.Context(0)
    mid = ctx.add_mesh_grid(pr.grid, pr.aij, pr.nij)
    return ctx, mid


def test_call_order_and_argument_errors():
    pr = _tiny()
    g = pr.grid
    ctx, mid = _ctx_with_mesh(pr)
    with pytest.raises(S.SrsError, match="not finalized"):
        ctx.explicit_step(1)
    with pytest.raises(S.SrsError, match="no realizations"):
        ctx.finalize()
    with pytest.raises(S.SrsError, match="mesh id"):
        ctx.add_realization(7, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, pr.params, seed=1, stream=0)
    bad = g.ryr_ele.copy()
    bad[0] = g.n_ele
    with pytest.raises(S.SrsError, match="out of range"):
        ctx.add_realization(mid, bad, g.ryr_jsr, g.ryr_nb, g.lcc_ele, pr.params, seed=1, stream=0)
    nb = g.ryr_nb.copy()
    nb[0] = g.n_ryr
    with pytest.raises(S.SrsError, match="neighbour"):
        ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, nb, g.lcc_ele, pr.params, seed=1, stream=0)
    p = pr.params.copy().setr(T_STEP=0.0)
    with pytest.raises(S.SrsError, match="t_step"):
        ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=1, stream=0)
    ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, pr.params, seed=1, stream=0)
    ctx.finalize()
    with pytest.raises(S.SrsError, match="already finalized"):
        ctx.finalize()
    with pytest.raises(S.SrsError, match="tile"):
        ctx.tune(tile=256)
    with pytest.raises(S.SrsError, match="occupancy"):
        ctx.tune(occupancy=7)
    with pytest.raises(S.SrsError, match="out of range"):
        ctx.state_get(3)
    with pytest.raises(S.SrsError, match="negative"):
        ctx.explicit_step(-1)
    with pytest.raises(S.SrsError, match="negative"):
        ctx.prepare(-5)
    ctx.prepare(300, S.STEP_GRAPH | S.STEP_PROTOCOL)  # 128-step and 44-step blocks, nothing stepped
    assert ctx.loop_get(0)["step"] == 0
    with pytest.raises(S.SrsError, match="out of range"):
        ctx.checkpoint(2)
    ctx.loop_set(0, done=1)
    with pytest.raises(S.SrsError, match="finished"):
        ctx.record_schedules(8)
    ctx.close()


def test_channel_count_limits():
    """1024 RyRs (the limit) step like the oracle; 1025 are refused."""
    pr = _tiny()
    g = pr.grid
    n = 1024
    rng = np.random.default_rng(4)
    k = rng.integers(0, g.n_ryr, n)
    ryr_ele, ryr_jsr = g.ryr_ele[k].astype(np.int32), g.ryr_jsr[k].astype(np.int32)
    nb = np.full(4 * n, -1, np.int32)
    nb[0::4] = np.r_[-1, np.arange(n - 1)]
    nb[1::4] = np.r_[np.arange(1, n), -1]
    ctx, mid = _ctx_with_mesh(pr)
    with pytest.raises(S.SrsError, match="exceed"):
        ctx.add_realization(mid, np.r_[ryr_ele, ryr_ele[:1]], np.r_[ryr_jsr, ryr_jsr[:1]],
                            np.r_[nb, -np.ones(4, np.int32)], g.lcc_ele, pr.params, seed=1, stream=0)
    p = fast_gating(pr.params).setr(RYR_A_STAR=0.3)
    ctx.add_realization(mid, ryr_ele, ryr_jsr, nb, g.lcc_ele, p, seed=9, stream=0)
    ctx.finalize()
    rs = (rng.random(n) < 0.05).astype(np.int32)
    R = np.log(rng.random(n + g.n_lcc))
    F = Fields(pr.X0, rs, np.full(g.n_lcc, 6, np.int32), np.full(g.n_lcc, 2, np.int32), R, [0.0, 1000.0])
    ctx.state_set(0, F.X)
    ctx.channels_set(0, F.ryr_state, F.lcc_state, F.lccv_state, F.R)
    ctx.globals_set(0, 0.0, 1000.0)
    ctx.explicit_step(60)
    X = ctx.state_get(0)
    st = ctx.channels_get(0)[0]
    ctx.close()
    G = grid_dict(g, pr.aij, pr.nij)
    G["ryr_ele"], G["ryr_jsr"], G["ryr_nb"] = ryr_ele, ryr_jsr, nb
    Fo = Oracle().steps(G, p.reals, F.copy(), 60, p.reals[1], gating=True, rng_kind=RNG_PHILOX, key=9, stream=0)
    assert np.array_equal(st, Fo.ryr_state)
    assert rel_err(X, Fo.X) <= 1e-10


def test_smallest_mesh_one_cube():
    """Six tets (one Kuhn cube), no jSR and no channels, against the oracle."""
    spec = W.tiny_box_spec(nx=1, ny=1, nz=1, jsr=False, ryr=0, lcc=0, sl=False)
    pr = W.prepare(spec, W.spark_params())
    g = pr.grid
    assert g.n_ele == 6
    F = initial_fields(pr, ryr_open=())
    X = F.X.copy()
    X[:g.n_ele] *= np.linspace(1, 3, g.n_ele)
    F = Fields(X, F.ryr_state, F.lcc_state, F.lccv_state, F.R, F.gv)
    ctx, mid = _ctx_with_mesh(pr)
    ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, pr.params, seed=1, stream=0)
    ctx.finalize()
    ctx.state_set(0, F.X)
    ctx.globals_set(0, 0.0, 1000.0)
    ctx.explicit_step(100)
    Xg = ctx.state_get(0)
    ctx.close()
    Fo = Oracle().steps(grid_dict(g, pr.aij, pr.nij), pr.params.reals, F.copy(), 100, pr.params.reals[1], gating=True)
    assert rel_err(Xg, Fo.X) <= 1e-10


def test_record_buffer_overflow_is_reported():
    pr = _tiny()
    g = pr.grid
    p = pr.params.copy()
    dt = p.reals[1]
    p.setr(T_FINAL=100 * dt, OUTPUT_STATES_INTERVAL=dt * (1 - 1e-9))
    ctx, mid = _ctx_with_mesh(pr)
    ctx.add_realization(mid, g.ryr_ele, g.ryr_jsr, g.ryr_nb, g.lcc_ele, p, seed=1, stream=0)
    ctx.finalize(record_capacity=10)
    ctx.state_set(0, pr.X0)
    ctx.explicit_step(50, S.STEP_PROTOCOL)
    with pytest.raises(S.SrsError, match="overflowed"):
        ctx.records(0)
    ctx.close()
