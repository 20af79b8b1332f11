"""The model extensions the reference cannot express (needs an MI355X):
Ca diffusion with D_CAJSR in the jSR and the SR-site rapid buffer.  Both
reduce to the reference model bit for bit when off; on, each is pinned by a
reference-model run that the mathematics makes equal to it."""
import numpy as np
import pytest

import srs_amd as S
from oracle.oracle import Fields, Oracle, grid_dict
from srs_amd import workloads as W
from tests.helpers import gpu_context, initial_fields, rel_err

pytestmark = pytest.mark.gpu


def _perturbed(pr, seed):
    """A non-equilibrium state: every species scaled by 1 + U(0, 0.5) per
    element (zero entries stay zero), RyRs and LCCs closed."""
    rng = np.random.default_rng(seed)
    F = initial_fields(pr, ryr_open=(), lcc_state=1, seed=seed)
    F.X = F.X * (1 + 0.5 * rng.random(F.X.size))
    return F


def _run(pr, p, F, n, ext=None):
    sch = np.zeros((n, pr.grid.n_ryr), np.int32)
    with gpu_context([pr], [F], params=[p]) as ctx:
        if ext is not None:
            ctx.extensions(0, **ext)
        ctx.replay(0, sch)
        return ctx.state_get(0)


def test_jsr_diffusion_equals_per_domain_reference_runs():
    """jSR and cytosol meet only at no-flux faces, so with the RyRs closed
    the two domains evolve apart: a run with D_Ca = 0.22 in the cytosol and
    D_CAJSR = 0.35 in the jSR equals, element for element and bit for bit,
    the reference step at D_Ca = 0.22 in the cytosol and at D_Ca = 0.35 in the
    jSR.  With D_CAJSR = D_Ca the extension changes no bit."""
    pr = W.prepare(W.tiny_box_spec(nx=6, ny=6, nz=4, tt=True, ryr=2, lcc=1), W.spark_params())
    g = pr.grid
    n = 300
    p = pr.params.copy().setr(D_CA=0.22, D_CA_JSR=0.35, V_RYR=0.0)
    F = _perturbed(pr, 3)
    X_ext = _run(pr, p, F.copy(), n, dict(jsr_diffusion=1))
    X_off = _run(pr, p, F.copy(), n)
    jsr = g.domain == 1
    assert jsr.any() and (~jsr).any()
    O = Oracle()
    G = grid_dict(g, pr.aij, pr.nij)
    out = {}
    for D in (0.22, 0.35):
        q = p.copy().setr(D_CA=D)
        out[D] = O.steps(G, q.reals, F.copy(), n, q.reals[1], gating=False,
                         schedule=np.zeros((n, g.n_ryr), np.int32)).X.reshape(6, -1)
    X = X_ext.reshape(6, -1)
    np.testing.assert_array_equal(X[:, ~jsr], out[0.22][:, ~jsr])
    np.testing.assert_array_equal(X[:, jsr], out[0.35][:, jsr])
    np.testing.assert_array_equal(X_off.reshape(6, -1), out[0.22])  # the reference ignores D_CAJSR
    assert not np.array_equal(X[:, jsr], out[0.22][:, jsr])
    p_eq = p.copy().setr(D_CA_JSR=0.22)
    assert np.array_equal(_run(pr, p_eq, F.copy(), n, dict(jsr_diffusion=1)), X_off)


def test_sr_sites_equal_the_sarcolemmal_rapid_buffer_form():
    """The SR-site rapid buffer divides the cytosolic Ca rate by
    1 + b K / (K + c)^2.  The reference's own low-affinity sarcolemmal buffer
    has that form with b = TT area x B_TOT_SL_LO / (V0 1e-15): the reference
    step on a grid whose every cytosol element carries the TT area that makes
    it 47 uM (and no high-affinity buffer) matches the extension's run to
    rel 1e-12 (b is then exact to an ulp).  With b = 0 the extension changes
    no bit."""
    pr = W.prepare(W.tiny_box_spec(nx=6, ny=6, nz=4, tt=False, ryr=2, lcc=1, sl=False), W.spark_params())
    g = pr.grid
    n = 300
    p = pr.params.copy().setr(V_RYR=0.0, B_TOT_SL_HI=0.0, B_TOT_SL_LO=1.0, K_D_SL_LO=0.6)
    F = _perturbed(pr, 5)
    assert np.all(g.tt_area == 0)
    X_ext = _run(pr, p, F.copy(), n, dict(b_sr_sites=47.0, k_d_sr_sites=0.6))
    X_off = _run(pr, p, F.copy(), n)
    X_zero = _run(pr, p, F.copy(), n, dict(b_sr_sites=0.0))
    cyt = g.domain == 0
    G = grid_dict(g, pr.aij, pr.nij)
    G["tt_area"] = np.where(cyt, 47.0 * g.V0 * 1e-15, 0.0)
    O = Oracle()
    sch = np.zeros((n, g.n_ryr), np.int32)
    Fo = O.steps(G, p.reals, F.copy(), n, p.reals[1], gating=False, schedule=sch)
    assert rel_err(X_ext, Fo.X) <= 1e-12
    G0 = grid_dict(g, pr.aij, pr.nij)
    Fo0 = O.steps(G0, p.reals, F.copy(), n, p.reals[1], gating=False, schedule=sch)
    np.testing.assert_array_equal(X_off, Fo0.X)
    np.testing.assert_array_equal(X_zero, X_off)
    assert rel_err(X_ext, X_off) > 1e-6  # the buffer does slow the Ca change


def test_adjoint_refuses_a_realization_with_an_extension():
    pr = W.prepare(W.tiny_box_spec(nx=4, ny=4, nz=3, tt=True, ryr=2, lcc=1), W.spark_params())
    g = pr.grid
    F = initial_fields(pr)
    with gpu_context([pr], [F]) as ctx:
        ctx.extensions(0, b_sr_sites=10.0)
        cen = g.centroids()
        d = S.read_scan_parameters(None)
        d.axis = 0
        d.x_min, d.x_max, d.res_x_f = -0.05, 0.05, 0.01
        sid, nf, _ = ctx.add_scan(0, cen - cen.mean(axis=0), d)
        sch = np.zeros((20, g.n_ryr), np.int32)
        with pytest.raises(S.SrsError, match="extension"):
            ctx.loss_grad(0, sid, sch, [10, 20], np.zeros((2, nf)), 8)
        with pytest.raises(S.SrsError, match="extension"):
            ctx.loss_grad_batch([0], sid, [sch], [10, 20], np.zeros((1, 2, nf)))
        with pytest.raises(S.SrsError, match="dissociation"):
            ctx.extensions(0, b_sr_sites=10.0, k_d_sr_sites=0.0)
