"""Mesh and channel file edge cases (no GPU): every malformed input named in
the reader is rejected with a clear error, never read past its end, and the
operator assembly refuses rows the ELL format cannot hold."""
from __future__ import annotations

import os

import numpy as np
import pytest

import srs_amd as S
from srs_amd import workloads as W


@pytest.fixture()
def mesh(tmp_path):
    base = str(tmp_path / "m")
    S.box_mesh_write(W.tiny_box_spec(nx=3, ny=3, nz=2), base)
    return base


def _edit(path, fn):
    with open(path) as f:
        lines = f.read().splitlines()
    lines = fn(lines)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def test_reads_the_written_mesh(mesh):
    g, warn = S.grid_read(mesh)
    assert g.n_ele == 3 * 3 * 2 * 6 and warn == 0


@pytest.mark.parametrize("suffix", [".node", ".ele", "_tet_properties.txt", "_face_properties.txt", "_ryr.txt",
                                    "_lcc.txt"])
def test_missing_file(mesh, suffix):
    os.remove(mesh + suffix)
    with pytest.raises(S.SrsError, match="cannot open"):
        S.grid_read(mesh)


@pytest.mark.parametrize("suffix,match", [("_tet_properties.txt", "truncated"), ("_face_properties.txt", "truncated"),
                                          (".node", "truncated"), (".ele", "truncated"), ("_ryr.txt", "truncated")])
def test_truncated_file(mesh, suffix, match):
    _edit(mesh + suffix, lambda L: L[:-2])
    with pytest.raises(S.SrsError, match=match):
        S.grid_read(mesh)


def test_malformed_face_line(mesh):
    _edit(mesh + "_face_properties.txt", lambda L: L[:5] + ["1 2 x"] + L[6:])
    with pytest.raises(S.SrsError, match="malformed"):
        S.grid_read(mesh)


def test_face_type_out_of_range(mesh):
    def bad(L):
        parts = L[3].split()
        parts[-1] = "7"
        return L[:3] + [" ".join(parts)] + L[4:]
    _edit(mesh + "_face_properties.txt", bad)
    with pytest.raises(S.SrsError, match="face type"):
        S.grid_read(mesh)


def test_face_element_out_of_range(mesh):
    def bad(L):
        parts = L[2].split()
        parts[0] = "100000"
        return L[:2] + [" ".join(parts)] + L[3:]
    _edit(mesh + "_face_properties.txt", bad)
    with pytest.raises(S.SrsError, match="range"):
        S.grid_read(mesh)


def test_element_node_out_of_range(mesh):
    def bad(L):
        parts = L[1].split()
        parts[2] = "999999"
        return L[:1] + [" ".join(parts)] + L[2:]
    _edit(mesh + ".ele", bad)
    with pytest.raises(S.SrsError, match="node"):
        S.grid_read(mesh)


def test_channel_indices_out_of_range(mesh):
    _edit(mesh + "_lcc.txt", lambda L: L[:1] + ["123456"] + L[2:])
    with pytest.raises(S.SrsError, match="LCC"):
        S.grid_read(mesh)


def test_ryr_neighbour_out_of_range(mesh):
    def bad(L):
        parts = L[1].split()
        parts[2] = "77"
        return L[:1] + [" ".join(parts)] + L[2:]
    _edit(mesh + "_ryr.txt", bad)
    with pytest.raises(S.SrsError, match="neighbour"):
        S.grid_read(mesh)


def test_shared_channel_elements_are_counted(mesh):
    # two LCCs in one element: read, with one warning (cpu/InitFuncs.cpp:657-671)
    def dup(L):
        return ["2", L[1], L[1]]
    _edit(mesh + "_lcc.txt", dup)
    g, warn = S.grid_read(mesh)
    assert g.n_lcc == 2 and warn == 1


def test_assembly_rejects_more_than_four_neighbours():
    # element 0 shares interior faces with five others
    e1 = np.zeros(5, np.int32)
    e2 = np.arange(1, 6, dtype=np.int32)
    g = S.Grid(n_ele=6, n_face=5, n_nodes=0, V0=np.ones(6), tt_area=np.zeros(6), sr_area=np.zeros(6),
               domain=np.zeros(6, np.int32), tropc=np.zeros(6, np.int32), ele_nodes=np.zeros(24, np.int32),
               node_xyz=np.zeros(0), ele1=e1, ele2=e2, face_type=np.zeros(5, np.int32), face_c=np.ones(5),
               ryr_ele=np.zeros(0, np.int32), ryr_jsr=np.zeros(0, np.int32), ryr_nb=np.zeros(0, np.int32),
               lcc_ele=np.zeros(0, np.int32), perm=np.arange(6, dtype=np.int32))
    with pytest.raises(S.SrsError, match="more than 4"):
        S.assemble_operator(g)


def test_assembly_rejects_a_boundary_face_without_element():
    g = S.Grid(n_ele=2, n_face=1, n_nodes=0, V0=np.ones(2), tt_area=np.zeros(2), sr_area=np.zeros(2),
               domain=np.zeros(2, np.int32), tropc=np.zeros(2, np.int32), ele_nodes=np.zeros(8, np.int32),
               node_xyz=np.zeros(0), ele1=np.array([5], np.int32), ele2=np.array([-1], np.int32),
               face_type=np.zeros(1, np.int32), face_c=np.ones(1), ryr_ele=np.zeros(0, np.int32),
               ryr_jsr=np.zeros(0, np.int32), ryr_nb=np.zeros(0, np.int32), lcc_ele=np.zeros(0, np.int32),
               perm=np.arange(2, dtype=np.int32))
    with pytest.raises(S.SrsError, match="range"):
        S.assemble_operator(g)
