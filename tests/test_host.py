"""Host-side parity (no GPU): the parameter reader, the mesh writer/reader,
the element renumbering, the operator assembly, boundary rates, equilibrium
state and states records of the product library, each checked against the
oracle and against what the mathematics fixes independently of both."""
from __future__ import annotations

import glob
import math
import os
import re
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import srs_amd as S
from oracle.oracle import Fields, grid_dict
from srs_amd import workloads as W

REF_PARAM = "/root/reference/param"

# Appendix of defaults: cpu/InitFuncs.cpp:28-108 (values restated from the
# reference's defaults block; t_step is stored in ms).
DEFAULT_REALS = dict(T_FINAL=500.0, T_STEP=12e-6, OUTPUT_STATES_INTERVAL=0.01, OUTPUT_GRID_INTERVAL=1.0, C_MIN=1.0,
                     INIT_RYR_HOLD=0.0, V_CLAMP=0.0, T_CLAMP=200.0, C_0_CA=0.1, C_0_CA_JSR=1000.0,
                     DEFAULT_CA_NSR=1000.0, CA_0=2000.0, DEFAULT_VM=-80.0, D_CA=0.25, D_CA_JSR=0.25, D_ATP=0.14,
                     D_CMDN=0.025, D_DYE=0.042, B_TOT_ATP=455.0, K_OFF_ATP=45.0, K_ON_ATP=0.225, B_TOT_CMDN=24.0,
                     K_OFF_CMDN=0.238, K_ON_CMDN=0.034, B_TOT_TRPN=70.0, K_OFF_TRPN=0.02, K_ON_TRPN=0.039,
                     B_TOT_SL_HI=1.6e-13, K_D_SL_HI=13.0, B_TOT_SL_LO=0.0, K_D_SL_LO=1100.0, B_TOT_DYE=50.0,
                     K_OFF_DYE=0.110, K_ON_DYE=0.10, B_T_JSR=30000.0, K_OFF_CSQN=63.8, K_ON_CSQN=0.10,
                     V_REFILL=0.095, V_RYR=3e-8, RYR_ETA=2.1, RYR_A_STAR=0.0, RYR_EPS_CC=-0.92, RYR_EPS_OO=-0.85,
                     RYR_K_MINUS=0.5, RYR_K_PLUS=0.1107e-3, RYR_PHI_M=4.0, RYR_PHI_B=1500.0, K_D_I=910.0,
                     K_D_SR=2240.0, A_P=150.0, V_CELL=25.84e3, LCC_DURATION=0.5, K_M_JSR=638.0)
DEFAULT_INTS = dict(FLAG_NO_FLUX_BOUNDARY=1, FLAG_NO_GATING=0, GATING_FFWD=0, FLAG_CLOCK_SEED=1, OUTPUT_GRID_FLAG=0,
                    FLAG_RYR_OUT=0, N_SIMS=1, INIT_RYR=-1, N_RYR_FIDELITY=1000, RAND_SEED_0=38376427,
                    RAND_ADD_A=12783, RAND_ADD_B=83, FLAG_V_CLAMP=0, FLAG_FIDELITY_SIM=0)

# XML symbol -> (kind, target(s)); cpu/InitFuncs.cpp:191-349
FLAGS = dict(flag_noflux="FLAG_NO_FLUX_BOUNDARY", flag_nogating="FLAG_NO_GATING", flag_ffwd="GATING_FFWD",
             flag_use_clock="FLAG_CLOCK_SEED", flag_output_grid="OUTPUT_GRID_FLAG",
             flag_output_ryr_states="FLAG_RYR_OUT")
INTS = dict(sims_per_proc="N_SIMS", ryr_open_max="N_RYR_FIDELITY", seed_base="RAND_SEED_0", seed_A="RAND_ADD_A",
            seed_B="RAND_ADD_B")
REALS = dict(c_min="C_MIN", ryr_open_time="INIT_RYR_HOLD", v_clamp="V_CLAMP", t_clamp="T_CLAMP", t_final="T_FINAL",
             states_interval="OUTPUT_STATES_INTERVAL", grid_interval="OUTPUT_GRID_INTERVAL", CA_I="C_0_CA",
             CA_O="CA_0", V_M="DEFAULT_VM", D_CA="D_CA", D_CAJSR="D_CA_JSR", D_ATP="D_ATP", D_CMDN="D_CMDN",
             D_DYE="D_DYE", B_TOT_ATP="B_TOT_ATP", K_OFF_ATP="K_OFF_ATP", K_ON_ATP="K_ON_ATP",
             B_TOT_CMDN="B_TOT_CMDN", K_OFF_CMDN="K_OFF_CMDN", K_ON_CMDN="K_ON_CMDN", B_TOT_TRPN="B_TOT_TRPN",
             K_OFF_TRPN="K_OFF_TRPN", K_ON_TRPN="K_ON_TRPN", B_TOT_SL="B_TOT_SL_HI", K_D_SL="K_D_SL_HI",
             B_TOT_DYE="B_TOT_DYE", K_OFF_DYE="K_OFF_DYE", K_ON_DYE="K_ON_DYE", B_TOT_CSQN="B_T_JSR",
             K_OFF_CSQN="K_OFF_CSQN", K_ON_CSQN="K_ON_CSQN", V_REFILL="V_REFILL", V_RYR="V_RYR", RYR_ETA="RYR_ETA",
             RYR_A_STAR="RYR_A_STAR", RYR_EPS_CC="RYR_EPS_CC", RYR_EPS_OO="RYR_EPS_OO", RYR_K_MINUS="RYR_K_MINUS",
             RYR_K_PLUS="RYR_K_PLUS", RYR_PHI_N="RYR_PHI_M", RYR_PHI_KD="RYR_PHI_B", K_D_I="K_D_I", K_D_SR="K_D_SR",
             A_P="A_P")

_FLOAT = re.compile(r"\s*([+-]?(\d+\.?\d*([eE][+-]?\d+)?|\.\d+([eE][+-]?\d+)?))")
_INT = re.compile(r"\s*([+-]?\d+)")


def c_atof(s: str) -> float:
    m = _FLOAT.match(s)
    return float(m.group(1)) if m else 0.0


def c_atoi(s: str) -> int:
    m = _INT.match(s)
    return int(m.group(1)) if m else 0


def expected_params(path):
    """The parameter vectors an XML file must produce, from a general XML
    parser and the reference's symbol table."""
    p = S.read_parameters(None)
    ints, reals, title = p.ints.copy(), p.reals.copy(), p.title
    root = ET.parse(path).getroot()
    for par in root.iter("parameter"):
        sym = val = None
        for ch in par:
            if ch.tag == "symbol":
                sym = "".join(ch.itertext())
            elif ch.tag == "value":
                val = "".join(ch.itertext())
        if sym is None or val is None:
            continue
        if sym in FLAGS:
            ints[S.I[FLAGS[sym]]] = 1 if val == "true" else 0
        elif sym == "protocol":
            ints[S.I["FLAG_V_CLAMP"]] = 1 if val == "gain" else 0
            ints[S.I["FLAG_FIDELITY_SIM"]] = 1 if val == "fidelity" else 0
        elif sym == "ryr_open_init":
            if val != "":
                ints[S.I["INIT_RYR"]] = c_atoi(val)
        elif sym in INTS:
            ints[S.I[INTS[sym]]] = c_atoi(val)
        elif sym == "t_step":
            reals[S.R["T_STEP"]] = c_atof(val) * 1e-6
        elif sym == "CA_SR":
            reals[S.R["C_0_CA_JSR"]] = reals[S.R["DEFAULT_CA_NSR"]] = c_atof(val)
        elif sym in REALS:
            reals[S.R[REALS[sym]]] = c_atof(val)
        elif sym == "paramset_title":
            title = val
    return ints, reals, title


def test_parameter_defaults():
    p = S.read_parameters(None)
    for k, v in DEFAULT_REALS.items():
        assert p.reals[S.R[k]] == v, k
    for k, v in DEFAULT_INTS.items():
        assert p.ints[S.I[k]] == v, k
    assert p.title == "Untitled"


@pytest.mark.skipif(not os.path.isdir(REF_PARAM), reason="reference parameter files not present")
def test_every_shipped_parameter_file():
    files = sorted(glob.glob(os.path.join(REF_PARAM, "*.xml")) + glob.glob(os.path.join(REF_PARAM, "*", "*.xml")))
    assert len(files) >= 100
    for f in files:
        p = S.read_parameters(f)
        ints, reals, title = expected_params(f)
        np.testing.assert_array_equal(p.ints, ints, err_msg=f)
        np.testing.assert_array_equal(p.reals, reals, err_msg=f)
        assert p.title == title[: len(p.title)] and len(p.title) == min(len(title), 255), f


def test_model_extensions_are_off_in_every_shipped_parameter_file(tmp_path):
    """The extension symbols (flag_jsr_diffusion, B_TOT_SRSITES, K_D_SRSITES)
    appear in no shipped file, so every shipped file runs the reference
    model; a file that sets them is read with the parameter reader's value
    rules, and the reference's reader ignores them (unknown symbols)."""
    files = sorted(glob.glob(os.path.join(REF_PARAM, "*.xml")) + glob.glob(os.path.join(REF_PARAM, "*", "*.xml")))
    assert len(files) >= 100
    off = dict(jsr_diffusion=0, b_sr_sites=0.0, k_d_sr_sites=0.6)
    assert S.read_extensions(None) == off
    for f in files:
        assert S.read_extensions(f) == off, f
    body = "".join(f"<parameter><symbol>{k}</symbol><value>{v}</value></parameter>" for k, v in [
        ("D_CA", "0.22"), ("D_CAJSR", "0.35"), ("flag_jsr_diffusion", "true"), ("B_TOT_SRSITES", "47uM"),
        ("K_D_SRSITES", "0.6")])
    x = tmp_path / "ext.xml"
    x.write_text(f'<?xml version="1.0"?>\n<simulation>{body}</simulation>\n')
    assert S.read_extensions(str(x)) == dict(jsr_diffusion=1, b_sr_sites=47.0, k_d_sr_sites=0.6)
    p = S.read_parameters(str(x))
    assert p.reals[S.R["D_CA_JSR"]] == 0.35 and p.reals[S.R["D_CA"]] == 0.22
    q = S.read_parameters(None)
    q.setr(D_CA=0.22, D_CA_JSR=0.35)
    np.testing.assert_array_equal(p.reals, q.reals)  # the extension symbols touch no reference parameter
    np.testing.assert_array_equal(p.ints, q.ints)


def test_parameter_value_quirks(tmp_path):
    """atoi/atof prefix parsing, 'true' as the only true flag, ns -> ms."""
    body = "".join(f"<parameter><symbol>{s}</symbol><value>{v}</value></parameter>" for s, v in [
        ("flag_nogating", "True"), ("flag_use_clock", "false"), ("flag_output_grid", "true"),
        ("sims_per_proc", "3 sims"), ("ryr_open_init", ""), ("ryr_open_max", "true"),
        ("t_step", "25.5"), ("D_CA", "2e-1um"), ("CA_SR", "750"), ("protocol", "gain"), ("mesh_file", "x.xml")])
    f = tmp_path / "p.xml"
    f.write_text(f'<?xml version="1.0"?>\n<simulation><!-- c -->{body}</simulation>\n')
    p = S.read_parameters(str(f))
    assert p.ints[S.I["FLAG_NO_GATING"]] == 0 and p.ints[S.I["FLAG_CLOCK_SEED"]] == 0
    assert p.ints[S.I["OUTPUT_GRID_FLAG"]] == 1 and p.ints[S.I["N_SIMS"]] == 3
    assert p.ints[S.I["INIT_RYR"]] == -1 and p.ints[S.I["N_RYR_FIDELITY"]] == 0
    assert p.reals[S.R["T_STEP"]] == 25.5 * 1e-6 and p.reals[S.R["D_CA"]] == 0.2
    assert p.reals[S.R["C_0_CA_JSR"]] == 750 and p.reals[S.R["DEFAULT_CA_NSR"]] == 750
    assert p.ints[S.I["FLAG_V_CLAMP"]] == 1 and p.ints[S.I["FLAG_FIDELITY_SIM"]] == 0


def test_parameter_file_errors(tmp_path):
    with pytest.raises(S.SrsError):
        S.read_parameters(str(tmp_path / "missing.xml"))
    f = tmp_path / "empty.xml"
    f.write_text("<simulation></simulation>")
    with pytest.raises(S.SrsError, match="No parameters"):
        S.read_parameters(str(f))


# ------------------------------------------------------------------ meshes
def small_spec():
    return W.tiny_box_spec(nx=5, ny=4, nz=3)


def test_mesh_files_round_trip(tmp_path):
    """Writing the tetgen/property files and reading them back gives the
    in-memory mesh exactly (the writer prints %.17g)."""
    spec = small_spec()
    base = str(tmp_path / "m")
    S.box_mesh_write(spec, base)
    for suf in (".node", ".ele", ".face", "_tet_properties.txt", "_face_properties.txt", "_ryr.txt", "_lcc.txt"):
        assert os.path.exists(base + suf), suf
    g, _ = S.grid_read(base)
    h = S.box_grid(spec)
    for k in ("V0", "tt_area", "sr_area", "domain", "tropc", "ele_nodes", "node_xyz", "ele1", "ele2", "face_type",
              "face_c", "ryr_ele", "ryr_jsr", "ryr_nb", "lcc_ele", "perm"):
        np.testing.assert_array_equal(getattr(g, k), getattr(h, k), err_msg=k)


def _tet_volume(xyz):
    a, b, c, d = xyz
    return abs(np.dot(b - a, np.cross(c - a, d - a))) / 6.0


def test_element_volumes_and_face_coefficients():
    """V0 = |det|/6 from the node coordinates; for interior faces,
    Face_C = |A_f|^2 / (dXi (A_f . e_xi)) with A_f the face area vector and
    e_xi the unit vector between the two element centroids."""
    g = S.box_grid(small_spec())
    X = g.node_xyz.reshape(-1, 3)
    tets = g.ele_nodes.reshape(-1, 4)
    vol = np.array([_tet_volume(X[t]) for t in tets])
    np.testing.assert_allclose(g.V0, vol, rtol=1e-12)
    cen = X[tets].mean(axis=1)
    n_int = 0
    for f in range(g.n_face):
        e1, e2 = g.ele1[f], g.ele2[f]
        if e2 < 0:
            continue
        shared = sorted(set(tets[e1]) & set(tets[e2]))
        assert len(shared) == 3
        a, b, c = X[shared]
        A = 0.5 * np.cross(b - a, c - a)
        if np.dot(A, cen[e2] - cen[e1]) < 0:
            A = -A
        d = cen[e2] - cen[e1]
        dxi = np.linalg.norm(d)
        fc = np.dot(A, A) / (dxi * np.dot(A, d / dxi))
        assert g.face_c[f] == pytest.approx(fc, rel=1e-12)
        n_int += 1
    assert n_int > 0
    # the box minus nothing: volumes tile the box exactly
    assert vol.sum() == pytest.approx(0.2 * 0.16 * (0.015 + 0.08), rel=1e-12)


def _bfs_properties(order, n, e1, e2):
    """BFS from element 0 over faces with a second element, children in the
    order of the first face (in file order) that reaches them."""
    assert order[0] == 0 and sorted(order) == list(range(n))
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    first_face = {}
    for f, (a, b) in enumerate(zip(e1, e2)):
        if b < 0:
            continue
        first_face.setdefault((a, b), f)
        first_face.setdefault((b, a), f)
    parent = np.full(n, -1)
    for (a, b), _ in first_face.items():
        if pos[a] < pos[b] and (parent[b] < 0 or pos[a] < pos[parent[b]]):
            parent[b] = a
    # FIFO: parents are non-decreasing along the order
    pp = [pos[parent[v]] for v in order[1:]]
    assert all(x <= y for x, y in zip(pp, pp[1:]))
    # siblings ordered by the first face joining them to the parent
    for k in range(1, n - 1):
        u, v = order[k], order[k + 1]
        if parent[u] == parent[v]:
            assert first_face[(parent[u], u)] < first_face[(parent[v], v)]


@pytest.mark.parametrize("nx,ny,nz", [(3, 3, 2), (5, 4, 3)])
def test_bfs_renumbering(oracle, nx, ny, nz):
    spec = W.tiny_box_spec(nx=nx, ny=ny, nz=nz)
    g = S.box_grid(spec)
    # undo the renumbering to get the mesh in generator order
    perm = g.perm
    e1 = perm[g.ele1]
    e2 = np.where(g.ele2 >= 0, perm[np.maximum(g.ele2, 0)], -1)
    ref = oracle.bfs(g.n_ele, e1, e2)
    assert len(ref) == g.n_ele
    _bfs_properties(list(ref), g.n_ele, list(e1), list(e2))
    np.testing.assert_array_equal(S.bfs_order(g.n_ele, e1, e2), ref)
    np.testing.assert_array_equal(perm, ref)


def test_bfs_rejects_a_disconnected_mesh():
    with pytest.raises(S.SrsError, match="not connected"):
        S.bfs_order(4, [0, 2], [1, 3])


def _dense(n, aij, nij):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] += aij[5 * i]
        for k in range(4):
            if aij[5 * i + 1 + k] != 0:
                A[i, nij[4 * i + k]] += aij[5 * i + 1 + k]
    return A


@pytest.mark.parametrize("no_flux", [True, False])
def test_operator_assembly(oracle, no_flux):
    g = S.box_grid(W.tiny_box_spec(nx=4, ny=3, nz=3), no_flux=no_flux)
    n = g.n_ele
    aij, nij = oracle.assemble(n, g.ele1, g.ele2, g.face_type, g.face_c, g.V0)
    # brute force: every face adds its coefficient in face order
    B = np.zeros((n, n))
    for f in range(g.n_face):
        t1, t2, ft, fc = g.ele1[f], g.ele2[f], g.face_type[f], g.face_c[f]
        k1 = fc / g.V0[t1]
        if ft == 0:
            k2 = fc / g.V0[t2]
            B[t1, t2] += k1
            B[t1, t1] -= k1
            B[t2, t1] += k2
            B[t2, t2] -= k2
        elif ft == 1:
            B[t1, t1] -= k1
    np.testing.assert_array_equal(_dense(n, aij, nij), B)
    # ELL packing: columns ascending, missing neighbours are self with 0
    for i in range(n):
        cols = [nij[4 * i + k] for k in range(4) if aij[5 * i + 1 + k] != 0]
        assert cols == sorted(cols) and all(c != i for c in cols)
        for k in range(len(cols), 4):
            assert nij[4 * i + k] == i and aij[5 * i + 1 + k] == 0
    # V-weighted symmetry (both sides equal Face_C) and zero row sums without
    # boundary sinks
    VA = g.V0[:, None] * B
    off = ~np.eye(n, dtype=bool)
    np.testing.assert_allclose(VA[off], VA.T[off], rtol=1e-14, atol=0)
    rs = B.sum(axis=1)
    if no_flux:
        assert np.max(np.abs(rs)) <= 1e-12 * np.max(np.abs(B))
    else:
        assert np.all(rs <= 1e-12 * np.max(np.abs(B))) and np.min(rs) < -1e-3 * np.max(np.abs(B))
    # product library: the same bits
    a2, n2 = S.assemble_operator(g)
    np.testing.assert_array_equal(a2, aij)
    np.testing.assert_array_equal(n2, nij)


def test_operator_decay_matches_eigenmodes(oracle):
    """Pure diffusion of one species with no reactions: n explicit steps
    multiply each eigen-mode of A by (1 + dt D lambda)^n exactly."""
    g = S.box_grid(W.tiny_box_spec(nx=3, ny=3, nz=2, jsr=False, ryr=0, lcc=0, sl=False))
    n = g.n_ele
    aij, nij = oracle.assemble(n, g.ele1, g.ele2, g.face_type, g.face_c, g.V0)
    A = _dense(n, aij, nij)
    lam, Vr = np.linalg.eig(A)
    p = S.read_parameters(None)
    P = p.reals.copy()
    for k in ("B_TOT_ATP", "B_TOT_CMDN", "B_TOT_TRPN", "B_TOT_DYE", "B_T_JSR", "A_P", "B_TOT_SL_HI", "V_REFILL"):
        P[S.R[k]] = 0.0
    D = P[S.R["D_CA"]]
    dt = 0.5 / (D * np.max(np.abs(np.diag(A))))
    rng = np.random.default_rng(3)
    c0 = 0.1 + rng.random(n)
    X = np.zeros(6 * n)
    X[:n] = c0
    G = grid_dict(g, aij, nij)
    G["ryr_ele"] = G["ryr_jsr"] = G["lcc_ele"] = np.zeros(0, np.int32)
    G["ryr_nb"] = np.zeros(0, np.int32)
    F = Fields(X, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0), [0.0, 0.0])
    nst = 40
    oracle.steps(G, P, F, nst, dt, gating=False)
    coef = np.linalg.solve(Vr, c0)
    expect = (Vr @ (coef * (1 + dt * D * lam) ** nst)).real
    np.testing.assert_allclose(F.X[:n], expect, rtol=1e-10, atol=0)


def test_boundary_rates_and_equilibrium(oracle):
    g = S.box_grid(W.tiny_box_spec(nx=4, ny=3, nz=3), no_flux=False)
    P = S.read_parameters(None).reals
    X, bnd = S.equilibrium_state(g, P)
    Xo, bo = oracle.equilibrium(g.n_ele, g.domain, g.tropc, P)
    np.testing.assert_array_equal(X, Xo)
    np.testing.assert_array_equal(bnd, bo)
    diff = np.array([P[S.R[k]] for k in ("D_CA", "D_ATP", "D_CMDN", "D_TRPN", "D_CSQN", "D_DYE")])
    br = S.boundary_rates(g, P, bnd)
    np.testing.assert_array_equal(br, oracle.boundary_rates(g.n_ele, g.ele1, g.face_type, g.face_c, g.V0, diff, bnd))
    # closed form: BR_s[i] = sum over the element's BOUND faces of D_s Face_C b_s / V0
    n = g.n_ele
    exp_br = np.zeros(6 * n)
    for f in np.nonzero(g.face_type == 1)[0]:
        e = g.ele1[f]
        for s in range(6):
            exp_br[s * n + e] += diff[s] * g.face_c[f] * bnd[s] / g.V0[e]
    np.testing.assert_allclose(br, exp_br, rtol=1e-14, atol=0)
    assert np.count_nonzero(br) > 0



@pytest.mark.parametrize("dims", [(4, 3, 3), (5, 4, 2)])
def test_boundary_rates_balance_the_bound_face_sink(oracle, dims):
    """What the mathematics fixes about the boundary rates, independent of
    their formula: a bound face (cpu/InitFuncs.cpp:1019-1024, 1068-1112) is a
    Dirichlet contact with the boundary value b, so a field that equals b
    everywhere has zero diffusive rate, D (A c)_i + BR_i = 0, in every element
    and every diffusing species (the operator's diagonal sink -D Face_C / V0 c
    cancels the source D Face_C b / V0); and the rate of a uniform field c is
    BR (1 - c / b), linear in c and zero in elements without bound faces.
    Checked for the oracle's boundary rates and the product's, through the
    oracle's own step with every reaction switched off."""
    nx, ny, nz = dims
    g = S.box_grid(W.tiny_box_spec(nx=nx, ny=ny, nz=nz, jsr=False, ryr=0, lcc=0, sl=False), no_flux=False)
    P = S.read_parameters(None).reals.copy()
    P[S.R["D_CA"]], P[S.R["D_DYE"]], P[S.R["D_ATP"]], P[S.R["D_CMDN"]] = 0.22, 0.042, 0.14, 0.025
    for k in ("K_ON_TRPN", "K_ON_DYE", "K_ON_ATP", "K_ON_CMDN", "K_OFF_TRPN", "K_OFF_DYE", "K_OFF_ATP",
              "K_OFF_CMDN", "A_P", "V_REFILL"):
        P[S.R[k]] = 0.0
    n = g.n_ele
    aij, nij = S.assemble_operator(g)
    b = np.array([0.7, 310.0, 11.0, 0.0, 0.0, 3.3])  # boundary values of Ca, ATP, CaM, -, -, Fluo
    diff = np.array([P[S.R[k]] for k in ("D_CA", "D_ATP", "D_CMDN", "D_TRPN", "D_CSQN", "D_DYE")])
    br_o = oracle.boundary_rates(n, g.ele1, g.face_type, g.face_c, g.V0, diff, b)
    br_p = S.boundary_rates(g, P, b)
    bound = np.zeros(n, bool)
    bound[g.ele1[g.face_type == 1]] = True
    assert bound.any() and not bound.all()
    dt = 1e-7
    for br in (br_o, br_p):
        G = grid_dict(g, aij, nij, br)
        G["ryr_ele"] = G["ryr_jsr"] = G["lcc_ele"] = np.zeros(0, np.int32)
        G["ryr_nb"] = np.zeros(0, np.int32)
        for scale in (1.0, 0.25):
            X = np.repeat(b * scale, n)
            F = Fields(X, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0),
                       [0.0, 0.0])
            oracle.steps(G, P, F, 1, dt, gating=False)
            rate = (F.X - X) / dt
            for s in (0, 1, 2, 5):
                r = rate[s * n:(s + 1) * n]
                B = br[s * n:(s + 1) * n]
                assert np.all(B[~bound] == 0) and np.all(B[bound] > 0), s
                # zero at c = b, BR (1 - c/b) below it: to round-off of the diagonal sum
                tol = 1e-9 * np.abs(B).max()
                np.testing.assert_allclose(r, B * (1 - scale), rtol=0, atol=tol)


# ----------------------------------------------------------- states records
def test_states_record_hand_case():
    """Two RyRs (one open) and two LCCs (one open) with chosen concentrations:
    every field of the record from its definition."""
    P = S.read_parameters(None).reals.copy()
    ca_cyt, ca_jsr = np.array([3.0, 5.0]), np.array([900.0, 800.0])
    vm = -10.0
    rec = S.states_record(1.25, [1, 0], ca_cyt, ca_jsr, [6, 6], [2, 1], [4.0, 7.0], vm, P)
    assert len(rec) == 7
    vryr, vcell = P[S.R["V_RYR"]], P[S.R["V_CELL"]]
    i_ryr = vryr * (900.0 - 3.0) * vcell * 96485 * 2e-6
    FRT = 96.5 / (310 * 8.314)
    e = math.exp(2 * vm * FRT)
    i_lcc = -9.13e-4 * vm * 2 * FRT * (4.0 * e - 0.34 * P[S.R["CA_0"]]) / (e - 1) * 96485 * 2e-6
    exp = np.array([1.25, i_ryr, 1, 1, i_lcc, 850.0, 4.0], np.float32)
    np.testing.assert_allclose(rec, exp, rtol=2e-7)
    rec2 = S.states_record(1.25, [1, 0], ca_cyt, ca_jsr, [6, 6], [2, 1], [4.0, 7.0], vm, P, ryr_out=True,
                           gpu_layout=True)
    np.testing.assert_array_equal(rec2[7:], np.array([3.0, 5.0, 1, 0], np.float32))


def test_states_record_matches_oracle(oracle):
    pr = W.prepare(W.tiny_box_spec(nx=4, ny=4, nz=3, ryr=3, lcc=2), W.spark_params())
    g = pr.grid
    rng = np.random.default_rng(7)
    X = pr.X0 * (1 + 0.1 * rng.random(pr.X0.size))
    rs = (rng.random(g.n_ryr) < 0.5).astype(np.int32)
    ls = np.array([6, 3][: g.n_lcc], np.int32)
    lv = np.array([2, 2][: g.n_lcc], np.int32)
    F = Fields(X, rs, ls, lv, np.zeros(g.n_ryr + g.n_lcc), [-5.0, 1000.0])
    G = grid_dict(g, pr.aij, pr.nij)
    n = g.n_ele
    for ryr_out in (False, True):
        for gl in (False, True):
            a = oracle.record(0.5, G, F, pr.params.reals, ryr_out=ryr_out, gpu_layout=gl)
            b = S.states_record(0.5, rs, X[g.ryr_ele], X[g.ryr_jsr], ls, lv, X[g.lcc_ele], -5.0, pr.params.reals,
                                ryr_out=ryr_out, gpu_layout=gl)
            np.testing.assert_array_equal(a, b)
    assert n > 0


def test_states_record_rejects_overflow():
    P = S.read_parameters(None).reals
    n = 125
    with pytest.raises(S.SrsError, match="255"):
        S.states_record(0.0, np.zeros(n), np.ones(n), np.ones(n), [], [], [], 0.0, P, ryr_out=True, gpu_layout=True)


def test_open_random_ryr_selection_law(oracle):
    """OpenRandomRyR (cpu/InitFuncs.cpp:1420-1475) picks RyR i with
    probability rate_i / sum(rates) where a closed RyR's rate uses the LINEAR
    phi = PHI_M c_jsr + PHI_B.  With eta = 0 and PHI_B = 0 the rate is
    proportional to the jSR Ca alone; with PHI_M = 0 and eta = 1 to the
    cytosolic Ca alone."""
    import ctypes as C

    from oracle.oracle import OrcChan

    n_ryr = 4
    n_ele = 8
    X = np.zeros(6 * n_ele)
    X[:4] = [1.0, 2.0, 3.0, 4.0]       # cytosolic elements 0..3
    X[4:8] = [10.0, 40.0, 20.0, 30.0]  # jSR elements 4..7
    ryr_ele = np.arange(4, dtype=np.int32)
    ryr_jsr = np.arange(4, 8, dtype=np.int32)
    nb = np.full(16, -1, np.int32)
    P = S.read_parameters(None).reals.copy()
    grid_u = (np.arange(20000) + 0.5) / 20000

    def freq(P):
        counts = np.zeros(n_ryr)
        for u in grid_u:
            st = np.zeros(n_ryr, np.int32)
            c = OrcChan()
            c.n_ryr, c.n_lcc = n_ryr, 0
            c.ryr_state = st.ctypes.data_as(C.POINTER(C.c_int))
            c.ryr_ele = ryr_ele.ctypes.data_as(C.POINTER(C.c_int))
            c.ryr_jsr = ryr_jsr.ctypes.data_as(C.POINTER(C.c_int))
            c.ryr_nb = nb.ctypes.data_as(C.POINTER(C.c_int))
            k = oracle.L.orc_open_random_ryr(X.ctypes.data_as(C.POINTER(C.c_double)), n_ele,
                                             P.ctypes.data_as(C.POINTER(C.c_double)), C.byref(c), float(u))
            assert st[k] == 1 and st.sum() == 1
            counts[k] += 1
        return counts / len(grid_u)

    P[S.R["RYR_ETA"]], P[S.R["RYR_PHI_B"]], P[S.R["RYR_PHI_M"]] = 0.0, 0.0, 2.0
    np.testing.assert_allclose(freq(P), np.array([10, 40, 20, 30]) / 100, atol=1e-4)
    P[S.R["RYR_ETA"]], P[S.R["RYR_PHI_B"]], P[S.R["RYR_PHI_M"]] = 1.0, 5.0, 0.0
    np.testing.assert_allclose(freq(P), np.array([1, 2, 3, 4]) / 10, atol=1e-4)


def test_host_code_under_sanitizers(tmp_path):
    """The host C++ (parameter and scan readers, box mesher, mesh reader on
    whole and truncated files, assembly, BFS) built with AddressSanitizer and
    UndefinedBehaviorSanitizer runs clean (tests/host_sanitize_driver.cpp)."""
    import shutil
    import subprocess

    if shutil.which("g++") is None:
        pytest.skip("no host compiler")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    csrc = os.path.join(root, "srs_amd", "csrc")
    exe = tmp_path / "driver"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-I", os.path.join(root, "include"), "-I", csrc,
           os.path.join(root, "tests", "host_sanitize_driver.cpp")] + \
          [os.path.join(csrc, f) for f in ("host_params.cpp", "host_mesh.cpp", "host_common.cpp")] + ["-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    args = [str(exe), str(tmp_path)]
    if os.path.isdir(REF_PARAM):
        args.append(REF_PARAM)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert "truncation sweep done" in r.stdout
