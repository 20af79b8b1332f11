"""ctypes front end of the oracle (test infrastructure only).

``Oracle`` wraps liboracle.so, the C restatement of the reference step in
srs_oracle.c.  ``RefStep`` wraps oracle/_ref/libsrsref.so, the reference's own
cpu/RunSimulation.cpp compiled by oracle/build.py; it is absent on machines
that never had the reference tree, and tests that need it skip then.

Only tests/, bench.py's cpu_baseline leg and __graft_entry__.smoke() use this.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsrsref.so")

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
u64p = C.POINTER(C.c_uint64)


def _i32(a):
    return a.ctypes.data_as(i32p) if a is not None else None


def _f64(a):
    return a.ctypes.data_as(f64p) if a is not None else None


class OrcGrid(C.Structure):
    _fields_ = [("n_ele", C.c_int), ("aij", f64p), ("nij", i32p), ("br", f64p), ("domain", i32p),
                ("tropc", i32p), ("V0", f64p), ("tt_area", f64p)]


class OrcChan(C.Structure):
    _fields_ = [("n_ryr", C.c_int), ("n_lcc", C.c_int), ("ryr_state", i32p), ("ryr_ele", i32p),
                ("ryr_jsr", i32p), ("ryr_nb", i32p), ("lcc_state", i32p), ("lccv_state", i32p),
                ("lcc_ele", i32p), ("R", f64p)]


class OrcRng(C.Structure):
    _fields_ = [("kind", C.c_int), ("key0", C.c_uint64), ("key1", C.c_uint64), ("realization", C.c_uint64),
                ("step", C.c_uint64)]


RNG_GLIBC = 0
RNG_PHILOX = 1
PHILOX_KEY1 = 0x5352535F414D4431  # the key word the product uses


def _load(path):
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built (oracle/build.py)")
    return C.CDLL(path)


class OrcScan(C.Structure):
    _fields_ = [("axis", C.c_int), ("x_min", C.c_double), ("x_max", C.c_double), ("res_f", C.c_double),
                ("res_ca", C.c_double), ("off1", C.c_double), ("off2", C.c_double), ("dom", C.c_double * 3),
                ("pad_min", C.c_double), ("pad_max", C.c_double), ("pad_res", C.c_double), ("var", C.c_double * 3),
                ("f0", C.c_double), ("c0", C.c_double)]


class Fields:
    """Mutable simulation state shared by the oracle and the reference harness:
    fields X[6*N], channel states, residuals, globals."""

    def __init__(self, X, ryr_state, lcc_state, lccv_state, R, gv):
        self.X = np.ascontiguousarray(X, np.float64).copy()
        self.ryr_state = np.ascontiguousarray(ryr_state, np.int32).copy()
        self.lcc_state = np.ascontiguousarray(lcc_state, np.int32).copy()
        self.lccv_state = np.ascontiguousarray(lccv_state, np.int32).copy()
        self.R = np.ascontiguousarray(R, np.float64).copy()
        self.gv = np.ascontiguousarray(gv, np.float64).copy()

    def copy(self):
        return Fields(self.X, self.ryr_state, self.lcc_state, self.lccv_state, self.R, self.gv)


class Oracle:
    def __init__(self):
        L = self.L = _load(ORACLE_SO)
        L.orc_philox4x64.argtypes = [u64p, u64p, u64p]
        L.orc_u64_to_open01.restype = C.c_double
        L.orc_u64_to_open01.argtypes = [C.c_uint64]
        L.orc_explicit_step.argtypes = [C.c_double, C.POINTER(OrcGrid), f64p, f64p, f64p, f64p, C.POINTER(OrcChan)]
        L.orc_gating.argtypes = [C.c_double, C.POINTER(OrcGrid), f64p, f64p, f64p, C.POINTER(OrcChan),
                                 C.POINTER(OrcRng)]
        L.orc_assemble_operator.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p, f64p, f64p, f64p, i32p]
        L.orc_boundary_rates.argtypes = [C.c_int, C.c_int, i32p, i32p, f64p, f64p, f64p, f64p, f64p]
        L.orc_bfs_order.argtypes = [C.c_int, C.c_int, i32p, i32p, i32p]
        L.orc_equilibrium_state.argtypes = [C.c_int, i32p, i32p, f64p, f64p, f64p]
        L.orc_states_record.argtypes = [C.c_double, f64p, C.c_int, f64p, C.POINTER(OrcChan), f64p, C.c_int, C.c_int,
                                        C.POINTER(C.c_float)]
        L.orc_open_random_ryr.argtypes = [f64p, C.c_int, f64p, C.POINTER(OrcChan), C.c_double]

    # -- primitives
    def philox(self, ctr, key):
        c = np.array(ctr, np.uint64)
        k = np.array(key, np.uint64)
        o = np.zeros(4, np.uint64)
        self.L.orc_philox4x64(c.ctypes.data_as(u64p), k.ctypes.data_as(u64p), o.ctypes.data_as(u64p))
        return o

    def open01(self, x):
        return self.L.orc_u64_to_open01(C.c_uint64(int(x)))

    def assemble(self, n_ele, ele1, ele2, ftype, face_c, V0):
        aij = np.zeros(5 * n_ele)
        nij = np.zeros(4 * n_ele, np.int32)
        e1, e2, ft = (np.ascontiguousarray(a, np.int32) for a in (ele1, ele2, ftype))
        fc, v = np.ascontiguousarray(face_c, np.float64), np.ascontiguousarray(V0, np.float64)
        rc = self.L.orc_assemble_operator(n_ele, len(e1), _i32(e1), _i32(e2), _i32(ft), _f64(fc), _f64(v),
                                          _f64(aij), _i32(nij))
        if rc != 0:
            raise ValueError("row with more than 4 neighbours")
        return aij, nij

    def boundary_rates(self, n_ele, ele1, ftype, face_c, V0, diff, bnd):
        br = np.zeros(6 * n_ele)
        e1, ft = np.ascontiguousarray(ele1, np.int32), np.ascontiguousarray(ftype, np.int32)
        args = [np.ascontiguousarray(a, np.float64) for a in (face_c, V0, diff, bnd)]
        self.L.orc_boundary_rates(n_ele, len(e1), _i32(e1), _i32(ft), *[_f64(a) for a in args], _f64(br))
        return br

    def bfs(self, n_ele, ele1, ele2):
        e1, e2 = np.ascontiguousarray(ele1, np.int32), np.ascontiguousarray(ele2, np.int32)
        v = np.zeros(n_ele, np.int32)
        n = self.L.orc_bfs_order(n_ele, len(e1), _i32(e1), _i32(e2), _i32(v))
        return v[:n]

    def equilibrium(self, n_ele, domain, tropc, reals):
        X = np.zeros(6 * n_ele)
        b = np.zeros(6)
        d, t = np.ascontiguousarray(domain, np.int32), np.ascontiguousarray(tropc, np.int32)
        self.L.orc_equilibrium_state(n_ele, _i32(d), _i32(t), _f64(np.ascontiguousarray(reals, np.float64)),
                                     _f64(X), _f64(b))
        return X, b

    # -- step
    def _grid(self, G):
        self._gk = G
        g = OrcGrid()
        g.n_ele = G["n_ele"]
        g.aij, g.nij = _f64(G["aij"]), _i32(G["nij"])
        g.br = _f64(G.get("br"))
        g.domain, g.tropc, g.V0, g.tt_area = _i32(G["domain"]), _i32(G["tropc"]), _f64(G["V0"]), _f64(G["tt_area"])
        return g

    @staticmethod
    def _chan(G, F):
        c = OrcChan()
        c.n_ryr, c.n_lcc = len(G["ryr_ele"]), len(G["lcc_ele"])
        c.ryr_state, c.ryr_ele, c.ryr_jsr, c.ryr_nb = (_i32(F.ryr_state), _i32(G["ryr_ele"]), _i32(G["ryr_jsr"]),
                                                       _i32(G["ryr_nb"]))
        c.lcc_state, c.lccv_state, c.lcc_ele, c.R = (_i32(F.lcc_state), _i32(F.lccv_state), _i32(G["lcc_ele"]),
                                                     _f64(F.R))
        return c

    def steps(self, G, reals, F: Fields, n_steps, dt, gating=True, rng_kind=RNG_PHILOX, key=0, stream=0,
              step0=0, schedule=None):
        """n_steps of ExplicitStep (+ gating).  Philox counters: step0 + k."""
        g = self._grid(G)
        c = self._chan(G, F)
        P = np.ascontiguousarray(reals, np.float64)
        buf = np.zeros_like(F.X)
        rng = OrcRng(rng_kind, key, PHILOX_KEY1, stream, 0)
        for k in range(n_steps):
            if schedule is not None:
                F.ryr_state[:] = schedule[k]
            self.L.orc_explicit_step(dt, C.byref(g), _f64(P), _f64(F.X), _f64(buf), _f64(F.gv), C.byref(c))
            if gating:
                rng.step = step0 + k
                self.L.orc_gating(dt, C.byref(g), _f64(P), _f64(F.X), _f64(F.gv), C.byref(c), C.byref(rng))
        return F

    def gating(self, G, reals, F: Fields, dt, rng_kind=RNG_PHILOX, key=0, stream=0, step=0):
        """UpdateChannelGating alone (the flag_ffwd loop), Philox counter `step`."""
        g = self._grid(G)
        c = self._chan(G, F)
        P = np.ascontiguousarray(reals, np.float64)
        rng = OrcRng(rng_kind, key, PHILOX_KEY1, stream, step)
        self.L.orc_gating(dt, C.byref(g), _f64(P), _f64(F.X), _f64(F.gv), C.byref(c), C.byref(rng))
        return F

    def linescan(self, scan, centers, V0, ca, caf):
        """linescan/main.cpp:96-249 on one state; scan is any object with the
        srs_scan_desc field names.  Returns (ca[n_ca], caf[n_f], cafb[n_f])."""
        s = OrcScan(int(scan.axis), scan.x_min, scan.x_max, scan.res_x_f, scan.res_x_ca, scan.offset1, scan.offset2,
                    (C.c_double * 3)(*scan.domain), scan.pad_min, scan.pad_max, scan.pad_res,
                    (C.c_double * 3)(*scan.var), scan.f0, scan.c0)
        nf, nca = C.c_int(), C.c_int()
        self.L.orc_linescan_size(C.byref(s), C.byref(nf), C.byref(nca))
        oca, ocaf, ocafb = np.zeros(max(nca.value, 1)), np.zeros(max(nf.value, 1)), np.zeros(max(nf.value, 1))
        cen = np.ascontiguousarray(centers, np.float64).reshape(-1)
        v, a, b = (np.ascontiguousarray(x, np.float64) for x in (V0, ca, caf))
        self.L.orc_linescan(C.byref(s), len(v), _f64(cen), _f64(v), _f64(a), _f64(b), _f64(oca), _f64(ocaf),
                            _f64(ocafb))
        return oca[:nca.value], ocaf[:nf.value], ocafb[:nf.value]

    def record(self, t, G, F: Fields, reals, ryr_out=False, gpu_layout=False):
        c = self._chan(G, F)
        rec = np.zeros(512, np.float32)
        n = self.L.orc_states_record(t, _f64(F.X), G["n_ele"], _f64(F.gv), C.byref(c),
                                     _f64(np.ascontiguousarray(reals, np.float64)), int(ryr_out), int(gpu_layout),
                                     rec.ctypes.data_as(C.POINTER(C.c_float)))
        return rec[:n].copy()


class RefStep:
    """The reference's own ExplicitStep / UpdateChannelGating (oracle/_ref)."""

    def __init__(self):
        L = self.L = _load(REF_SO)
        L.ref_srand.argtypes = [C.c_uint]
        L.ref_explicit_steps.argtypes = [C.c_int, C.c_double, C.c_int, f64p, i32p, f64p, i32p, i32p, f64p, f64p,
                                         C.c_int, C.c_int, i32p, i32p, i32p, i32p, i32p, i32p, i32p, f64p, i32p, f64p,
                                         f64p, f64p, i32p]

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_SO)

    def steps(self, G, ints, reals, F: Fields, n_steps, dt, schedule=None):
        ints = np.ascontiguousarray(ints, np.int32)
        reals = np.ascontiguousarray(reals, np.float64)
        sch = np.ascontiguousarray(schedule, np.int32) if schedule is not None else None
        self.L.ref_explicit_steps(n_steps, dt, G["n_ele"], _f64(G["aij"]), _i32(G["nij"]), _f64(G.get("br")),
                                  _i32(G["domain"]), _i32(G["tropc"]), _f64(G["V0"]), _f64(G["tt_area"]),
                                  len(G["ryr_ele"]), len(G["lcc_ele"]), _i32(F.ryr_state), _i32(G["ryr_ele"]),
                                  _i32(G["ryr_jsr"]), _i32(G["ryr_nb"]), _i32(F.lcc_state), _i32(F.lccv_state),
                                  _i32(G["lcc_ele"]), _f64(F.R), _i32(ints), _f64(reals), _f64(F.gv), _f64(F.X),
                                  _i32(sch))
        return F

    def srand(self, seed):
        self.L.ref_srand(seed)


def grid_dict(g, aij, nij, br=None):
    """The arrays the step needs, from a srs_amd Grid plus its operator."""
    return dict(n_ele=g.n_ele, aij=np.ascontiguousarray(aij, np.float64), nij=np.ascontiguousarray(nij, np.int32),
                br=None if br is None else np.ascontiguousarray(br, np.float64),
                domain=np.ascontiguousarray(g.domain, np.int32), tropc=np.ascontiguousarray(g.tropc, np.int32),
                V0=np.ascontiguousarray(g.V0, np.float64), tt_area=np.ascontiguousarray(g.tt_area, np.float64),
                ryr_ele=np.ascontiguousarray(g.ryr_ele, np.int32), ryr_jsr=np.ascontiguousarray(g.ryr_jsr, np.int32),
                ryr_nb=np.ascontiguousarray(g.ryr_nb, np.int32), lcc_ele=np.ascontiguousarray(g.lcc_ele, np.int32))
