"""The linescan operator (linescan/main.cpp:96-249): the scan parameter reader
of the product library, and the oracle's restatement pinned by closed forms
(one cell, the PSF integral over a fine uniform grid, brute-force nearest
cells).  The device operator is compared with the oracle in
test_gpu_parity.py."""
from __future__ import annotations

import math
import os
import re
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import srs_amd as S

REF_SCAN = "/root/reference/linescan/DefaultScan.xml"
PI_REF = 3.14159  # linescan/parameters.h:45

DEFAULTS = dict(n_scans=26, n_scans_skip=2, axis=2, res_x_f=0.15, res_x_ca=0.01, res_t=2.0, x_min=-4.0, x_max=4.0,
                offset1=0.1075, offset2=0.0, domain=(4.0, 4.0, 4.0), pad_min=-5.0, pad_max=5.0, pad_res=0.3,
                var=(0.0289, 0.1154, 0.0289), f0=4.1667, c0=0.10)

# XML symbol -> field (linescan/main.cpp:461-501)
SYMS = dict(scan_n=("n_scans", int), scan_n_per_line=("n_scans_skip", int), scan_x_res_f=("res_x_f", float),
            flag_x_res_ca=("n_scans_skip", int), scan_x_min=("x_min", float), scan_x_max=("x_max", float),
            scan_offset_1=("offset1", float), scan_offset_2=("offset2", float), scan_domain_x=("domain", 0),
            scan_domain_y=("domain", 1), scan_domain_z=("domain", 2), scan_padding_min=("pad_min", float),
            scan_padding_max=("pad_max", float), scan_padding_res=("pad_res", float), scan_var_x=("var", 0),
            scan_var_y=("var", 1), scan_var_z=("var", 2), scan_f0=("f0", float), scan_c0=("c0", float))


def _as_dict(d):
    out = {}
    for k, v in DEFAULTS.items():
        x = getattr(d, k)
        out[k] = tuple(x) if isinstance(v, tuple) else x
    return out


def test_scan_parameter_defaults():
    assert _as_dict(S.read_scan_parameters(None)) == DEFAULTS


@pytest.mark.skipif(not os.path.exists(REF_SCAN), reason="reference linescan parameters not present")
def test_scan_parameters_of_the_shipped_file():
    exp = dict(DEFAULTS)
    exp["domain"], exp["var"] = list(exp["domain"]), list(exp["var"])
    root = ET.parse(REF_SCAN).getroot()
    assert root.tag == "analysis"
    for par in root.findall("parameter"):
        sym = "".join(par.find("symbol").itertext())
        val = "".join(par.find("value").itertext())
        if sym == "scan_axis":
            exp["axis"] = "XYZ".index(val)
        elif sym in SYMS:
            k, kind = SYMS[sym]
            if kind is int:
                exp[k] = int(re.match(r"\s*[+-]?\d+", val).group(0))
            elif kind is float:
                exp[k] = float(val)
            else:
                exp[k][kind] = float(val)
    exp["domain"], exp["var"] = tuple(exp["domain"]), tuple(exp["var"])
    got = _as_dict(S.read_scan_parameters(REF_SCAN))
    assert got == exp
    assert got["axis"] == 2 and got["x_min"] == -3.0 and got["f0"] == 4.166666666666667


def test_scan_parameter_errors(tmp_path):
    f = tmp_path / "bad.xml"
    f.write_text("<simulation><parameter><symbol>scan_n</symbol><value>3</value></parameter></simulation>")
    with pytest.raises(S.SrsError, match="root"):
        S.read_scan_parameters(str(f))
    f.write_text("<analysis><parameter><symbol>scan_axis</symbol><value>W</value></parameter></analysis>")
    with pytest.raises(S.SrsError, match="axis"):
        S.read_scan_parameters(str(f))


def test_scan_output_file_format(tmp_path):
    """srs_scan_write gives the linescan program's output files byte for byte:
    WriteOutputFile (linescan/main.cpp:316-336) prints every value of a row
    with "%g " and ends the row with a newline; Python's %g is C's printf %g
    for finite doubles (exponent forms, six significant digits, -0)."""
    rng = np.random.default_rng(5)
    rows = rng.random((4, 7)) * 10.0 ** rng.integers(-7, 8, (4, 7))
    rows[0, :4] = [0.0, -0.0, 1e-5, 123456789.0]
    rows[1, :3] = [4.1667, 0.0999999999, -2.5e-300]
    f = tmp_path / "scan_output.txt"
    S.scan_write(str(f), rows)
    expect = "".join("".join("%g " % v for v in row) + "\n" for row in rows)
    assert f.read_bytes() == expect.encode()
    # one row, and no rows at all (an empty file, as the reference writes for 0 scans)
    S.scan_write(str(f), rows[2])
    assert f.read_text() == "".join("%g " % v for v in rows[2]) + "\n"
    S.scan_write(str(f), np.zeros((0, 5)))
    assert f.read_bytes() == b""
    with pytest.raises(S.SrsError, match="cannot open"):
        S.scan_write(str(tmp_path / "no_such_dir" / "x.txt"), rows)


def _scan(**kw):
    d = S.read_scan_parameters(None)
    for k, v in kw.items():
        if k in ("domain", "var"):
            for i, x in enumerate(v):
                getattr(d, k)[i] = x
        else:
            setattr(d, k, v)
    return d


def _no_padding(d, half=2.0):
    """Padding grids collapse to nothing when pad_min/pad_max equal the domain faces."""
    d.domain[0] = d.domain[1] = d.domain[2] = 2 * half
    d.pad_min, d.pad_max = -half, half
    return d


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_single_cell_closed_form(oracle, axis):
    """One cell at a known offset from the line: the blurred value is
    V0 CaF times the normal density with the reference's pi and its axis-to-
    variance assignment (for a Z scan the x offset meets var_z)."""
    var = (0.01, 0.04, 0.09)
    d = _no_padding(_scan(axis=axis, x_min=-0.3, x_max=0.3, res_x_f=0.1, res_x_ca=0.1, offset1=0.05, offset2=-0.02,
                          var=var))
    c = np.array([[0.11, -0.07, 0.13]])
    V0, ca, caf = np.array([2e-4]), np.array([3.0]), np.array([7.0])
    oca, ocaf, ocafb = oracle.linescan(d, c, V0, ca, caf)
    n_pts = int((0.3 - -0.3) / 0.1)  # C truncation: 0.6 / 0.1 = 5.999..., so 5 points
    assert len(ocafb) == n_pts == 5 and len(oca) == n_pts
    sig = [math.sqrt(v) for v in var]
    if axis == 2:
        sig = [sig[2], sig[0], sig[1]]
    for i in range(n_pts):
        p = [0.0, 0.0, 0.0]
        others = [k for k in range(3) if k != axis]
        p[others[0]], p[others[1]] = 0.05, -0.02
        p[axis] = -0.3 + i * 0.1
        r = c[0] - p
        q = sum((r[k] / sig[k]) ** 2 for k in range(3))
        w = math.exp(-0.5 * q) / ((2 * PI_REF) ** 1.5 * sig[0] * sig[1] * sig[2]) if r @ r < 1 else 0.0
        assert ocafb[i] == pytest.approx(2e-4 * 7.0 * w, rel=1e-13, abs=0)
        assert ocaf[i] == 7.0 and oca[i] == 3.0


def test_psf_integrates_to_the_padding_value(oracle):
    """A line inside a fine uniform padding grid of value F0 reads F0 times
    the normalisation error of pi = 3.14159, less the part of the PSF that
    falls in the small box the padding leaves out (the 'mesh domain', here
    a 0.05 um cube of 8 grid cells, computed cell by cell)."""
    h, var = 0.025, 0.01
    d = _scan(axis=0, x_min=-0.1, x_max=0.1, res_x_f=0.1, res_x_ca=0.1, offset1=0.0, offset2=0.0,
              var=(var, var, var), f0=4.0, c0=0.1)
    d.domain[0] = d.domain[1] = d.domain[2] = 2 * h
    d.pad_min, d.pad_max, d.pad_res = -0.8, 0.8, h
    far = np.array([[50.0, 50.0, 50.0]])
    oca, ocaf, ocafb = oracle.linescan(d, far, np.array([1.0]), np.array([9.0]), np.array([9.0]))
    assert len(ocafb) == 2
    norm = (math.pi / PI_REF) ** 1.5
    cube = [np.array([sx, sy, sz]) * h / 2 for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]
    for i, x in enumerate((-0.1, 0.0)):
        p = np.array([x, 0.0, 0.0])
        hole = sum(h ** 3 * 4.0 * math.exp(-0.5 * ((c - p) @ (c - p)) / var) / ((2 * PI_REF) ** 1.5 * var ** 1.5)
                   for c in cube)
        assert ocafb[i] == pytest.approx(4.0 * norm - hole, rel=1e-9)
    assert np.all(ocaf == 4.0) and np.all(oca == 0.1)


def test_nearest_cell_brute_force(oracle):
    rng = np.random.default_rng(5)
    n = 400
    c = rng.uniform(-1, 1, size=(n, 3))
    c[17] = c[3]  # an exact tie: the first index wins
    V0 = rng.uniform(1e-4, 2e-4, n)
    ca, caf = rng.uniform(0, 5, n), rng.uniform(0, 50, n)
    d = _scan(axis=1, x_min=-1.5, x_max=1.5, res_x_f=0.05, res_x_ca=0.02, offset1=0.2, offset2=-0.1,
              var=(0.02, 0.05, 0.03), f0=4.0, c0=0.1, pad_min=-1.6, pad_max=1.6, pad_res=0.4)
    d.domain[0] = d.domain[1] = d.domain[2] = 2.0
    oca, ocaf, ocafb = oracle.linescan(d, c, V0, ca, caf)
    # the padding cells, in the reference's order (linescan/main.cpp:119-141, 268-306)
    pads = []

    def grid(x0, x1, y0, y1, z0, z1):
        dx = dy = dz = 0.4
        dx, dy, dz = (-dx if x0 > x1 else dx), (-dy if y0 > y1 else dy), (-dz if z0 > z1 else dz)
        for i in range(int((x1 - x0) / dx)):
            for j in range(int((y1 - y0) / dy)):
                for k in range(int((z1 - z0) / dz)):
                    pads.append((x0 + i * dx + 0.5 * dx, y0 + j * dy + 0.5 * dy, z0 + k * dz + 0.5 * dz))

    grid(-1.6, 1.6, -1.6, 1.6, 1.0, 1.6)
    grid(-1.6, 1.6, -1.6, 1.6, -1.0, -1.6)
    grid(-1.0, -1.6, -1.6, 1.6, -1.0, 1.0)
    grid(1.0, 1.6, -1.6, 1.6, -1.0, 1.0)
    grid(-1.0, 1.0, -1.0, -1.6, -1.0, 1.0)
    grid(-1.0, 1.0, 1.0, 1.6, -1.0, 1.0)
    allc = np.vstack([c, np.array(pads)])
    allca = np.concatenate([ca, np.full(len(pads), 0.1)])
    allcaf = np.concatenate([caf, np.full(len(pads), 4.0)])
    for res, vals, out in ((0.05, allcaf, ocaf), (0.02, allca, oca)):
        for i in range(len(out)):
            p = np.array([0.2, -1.5 + i * res, -0.1])
            dist = np.sum((allc - p) ** 2, axis=1)
            assert out[i] == vals[int(np.argmin(dist))]
    assert np.any(ocaf == 4.0) and np.any(ocaf != 4.0)


def test_scan_point_range_keeps_every_contributing_point():
    """The forward blur and the adjoint's scan source visit only the points
    a cell can reach (dev_common.h scan_point_range): those whose axis
    coordinate lies within sqrt(1 + 1e-9 - t^2) (+- one point) of the
    cell's, t the cell's distance to the line, and none when
    t^2 >= 1 + 1e-9.  Restated here with the same double arithmetic: over
    random cells near and far from lines on each axis (including cells at
    distance ~1 um and beyond the line's ends), the visited range holds every
    point that passes the full loop's 1 um test, so the sum is the full
    loop's term for term."""
    import math

    rng = np.random.default_rng(5)
    for axis in (0, 1, 2):
        base = rng.uniform(-0.3, 0.3, 3)
        x_min, res, n_f = -0.96, 0.0297, 64
        cells = np.concatenate([rng.uniform(-2.0, 2.0, (4000, 3)),
                                base + rng.normal(0, 1, (4000, 3)) / np.sqrt(3.0)])
        # cells at distance 1 um from the line, give or take rounding
        u = rng.normal(0, 1, (2000, 3))
        u[:, axis] = 0
        u /= np.linalg.norm(u, axis=1)[:, None]
        ring = base + u * (1.0 + rng.uniform(-1e-15, 1e-15, (2000, 1)))
        ring[:, axis] = rng.uniform(-1.5, 1.5, 2000)
        cells = np.concatenate([cells, ring])
        for c in cells:
            p_axis = [x_min + i * res for i in range(n_f)]
            full = []
            for i in range(n_f):
                p = list(base)
                p[axis] = p_axis[i]
                rx, ry, rz = c[0] - p[0], c[1] - p[1], c[2] - p[2]
                if rx * rx + ry * ry + rz * rz < 1.0:
                    full.append(i)
            t2 = 0.0
            for a in range(3):
                if a != axis:
                    t2 += (c[a] - base[a]) * (c[a] - base[a])
            if not (t2 < 1.0 + 1e-9):
                assert full == []
                continue
            half = math.sqrt(1.0 + 1e-9 - t2)
            lo = max(0, int(math.floor((c[axis] - half - x_min) / res)) - 1)
            hi = min(n_f - 1, int(math.ceil((c[axis] + half - x_min) / res)) + 1)
            assert all(lo <= i <= hi for i in full), (c, full, lo, hi)
