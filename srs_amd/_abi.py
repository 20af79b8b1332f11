"""ctypes binding of libsrs_amd.so (declarations in include/srs_abi.h).

The library is loaded from this package directory only; a missing library is
an error (there is no fallback path).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsrs_amd.so")

N_INT = 19
N_REAL = 56
TITLE_LEN = 256

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)
u8p = C.POINTER(C.c_uint8)


class SrsGrid(C.Structure):
    _fields_ = [
        ("n_ele", C.c_int32), ("n_face", C.c_int32), ("n_nodes", C.c_int32),
        ("V0", f64p), ("tt_area", f64p), ("sr_area", f64p),
        ("domain", i32p), ("tropc", i32p),
        ("ele_nodes", i32p), ("node_xyz", f64p),
        ("ele1", i32p), ("ele2", i32p), ("face_type", i32p), ("face_c", f64p),
        ("n_ryr", C.c_int32), ("n_lcc", C.c_int32),
        ("ryr_ele", i32p), ("ryr_jsr", i32p), ("ryr_nb", i32p), ("lcc_ele", i32p),
        ("perm", i32p),
    ]


class SrsBoxSpec(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("xs", f64p), ("ys", f64p), ("zs", f64p),
        ("jsr_lo", C.c_double * 3), ("jsr_hi", C.c_double * 3),
        ("sl_face_z0", C.c_int32),
        ("tt_axis_x", C.c_double), ("tt_axis_z", C.c_double), ("tt_radius", C.c_double),
        ("trpn_radius", C.c_double),
        ("ryr_rows", C.c_int32), ("ryr_cols", C.c_int32),
        ("ryr_spacing", C.c_double), ("ryr_center", C.c_double * 2), ("ryr_z_cyt", C.c_double),
        ("ryr_z_jsr", C.c_double),
        ("ryr_mask", u8p),
        ("n_lcc", C.c_int32),
        ("lcc_xyz", f64p),
    ]


class SrsMeshDesc(C.Structure):
    _fields_ = [
        ("n_ele", C.c_int32), ("aij", f64p), ("nij", i32p), ("br", f64p),
        ("domain", i32p), ("tropc", i32p), ("V0", f64p), ("tt_area", f64p), ("order", C.c_int32),
        ("perm", i32p),
    ]


class SrsScanDesc(C.Structure):
    _fields_ = [
        ("n_scans", C.c_int32), ("n_scans_skip", C.c_int32), ("axis", C.c_int32),
        ("res_x_f", C.c_double), ("res_x_ca", C.c_double), ("res_t", C.c_double),
        ("x_min", C.c_double), ("x_max", C.c_double), ("offset1", C.c_double), ("offset2", C.c_double),
        ("domain", C.c_double * 3), ("pad_min", C.c_double), ("pad_max", C.c_double), ("pad_res", C.c_double),
        ("var", C.c_double * 3), ("f0", C.c_double), ("c0", C.c_double),
    ]


class SrsChanDesc(C.Structure):
    _fields_ = [
        ("n_ryr", C.c_int32), ("ryr_ele", i32p), ("ryr_jsr", i32p), ("ryr_nb", i32p),
        ("n_lcc", C.c_int32), ("lcc_ele", i32p),
    ]


class SrsExtensions(C.Structure):
    _fields_ = [("jsr_diffusion", C.c_int32), ("b_sr_sites", C.c_double), ("k_d_sr_sites", C.c_double)]


class SrsLoopState(C.Structure):
    _fields_ = [
        ("t", C.c_double), ("dt_last", C.c_double), ("step", C.c_int64), ("done", C.c_int32),
        ("lcc_open", C.c_int32), ("t_lcc_close", C.c_double), ("n_ryr_open_last", C.c_int32),
        ("ca_ss_max_last", C.c_double), ("ffwd", C.c_int32),
    ]


_PROTOS = {
    "srs_last_error": (C.c_char_p, []),
    "srs_version": (C.c_char_p, []),
    "srs_read_parameters": (C.c_int, [C.c_char_p, i32p, f64p, C.c_char_p]),
    "srs_grid_read": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.POINTER(SrsGrid)), i32p]),
    "srs_grid_free": (None, [C.POINTER(SrsGrid)]),
    "srs_bfs_order": (C.c_int, [C.c_int32, C.c_int32, i32p, i32p, i32p]),
    "srs_assemble_operator": (C.c_int, [C.c_int32, C.c_int32, i32p, i32p, i32p, f64p, f64p, f64p, i32p]),
    "srs_boundary_rates": (C.c_int, [C.c_int32, C.c_int32, i32p, i32p, f64p, f64p, f64p, f64p, f64p]),
    "srs_equilibrium_state": (C.c_int, [C.c_int32, i32p, i32p, f64p, f64p, f64p]),
    "srs_box_mesh_write": (C.c_int, [C.POINTER(SrsBoxSpec), C.c_char_p]),
    "srs_box_grid": (C.c_int, [C.POINTER(SrsBoxSpec), C.c_int, C.POINTER(C.POINTER(SrsGrid))]),
    "srs_morton_order": (C.c_int, [C.c_int32, i32p, f64p, i32p]),
    "srs_states_record": (C.c_int, [C.c_double, C.c_int32, i32p, f64p, f64p, C.c_int32, i32p, i32p, f64p,
                                    C.c_double, f64p, C.c_int32, C.c_int32, f32p]),
    "srs_run_cru3d": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32,
                                C.POINTER(C.c_double)]),
    "srs_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "srs_ctx_destroy": (C.c_int, [C.c_void_p]),
    "srs_ctx_record_layout": (C.c_int, [C.c_void_p, C.c_int32]),
    "srs_ctx_extensions": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(SrsExtensions)]),
    "srs_read_extensions": (C.c_int, [C.c_char_p, C.POINTER(SrsExtensions)]),
    "srs_ctx_tune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "srs_ctx_add_mesh": (C.c_int, [C.c_void_p, C.POINTER(SrsMeshDesc), i32p]),
    "srs_ctx_add_realization": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(SrsChanDesc), i32p, f64p,
                                          C.c_uint64, C.c_uint64, i32p]),
    "srs_ctx_finalize": (C.c_int, [C.c_void_p, C.c_int32]),
    "srs_state_set": (C.c_int, [C.c_void_p, C.c_int32, f64p]),
    "srs_state_get": (C.c_int, [C.c_void_p, C.c_int32, f64p]),
    "srs_channels_set": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, i32p, f64p]),
    "srs_channels_get": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, i32p, f64p]),
    "srs_globals_set": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double]),
    "srs_globals_get": (C.c_int, [C.c_void_p, C.c_int32, f64p, f64p]),
    "srs_loop_set": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(SrsLoopState)]),
    "srs_loop_get": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(SrsLoopState)]),
    "srs_explicit_step": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint32]),
    "srs_channel_gating": (C.c_int, [C.c_void_p]),
    "srs_explicit_step_replay": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, i32p]),
    "srs_explicit_step_replay_static": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, i32p]),
    "srs_channel_ca": (C.c_int, [C.c_void_p, C.c_int32, f64p, f64p, f64p]),
    "srs_records_read": (C.c_int, [C.c_void_p, C.c_int32, f32p, C.c_int32, i32p, i32p]),
    "srs_ctx_sync": (C.c_int, [C.c_void_p]),
    "srs_real_info": (C.c_int, [C.c_void_p, C.c_int32, i32p, i32p, i32p]),
    "srs_last_timing": (C.c_int, [C.c_void_p, f64p, f64p]),
    "srs_time_sweep": (C.c_int, [C.c_void_p, C.c_int32, f64p, f64p]),
    "srs_scan_read_parameters": (C.c_int, [C.c_char_p, C.POINTER(SrsScanDesc)]),
    "srs_ctx_add_scan": (C.c_int, [C.c_void_p, C.c_int32, f64p, C.POINTER(SrsScanDesc), i32p, i32p, i32p]),
    "srs_scan_sample": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, f64p, f64p, f64p]),
    "srs_scan_write": (C.c_int, [C.c_char_p, f64p, C.c_int32, C.c_int32]),
    "srs_adjoint_theta": (C.c_int, [i32p]),
    "srs_params_set": (C.c_int, [C.c_void_p, C.c_int32, f64p]),
    "srs_record_schedule": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, i32p]),
    "srs_linescan_loss_grad": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int64, i32p, C.c_int32,
                                         C.POINTER(C.c_int64), f64p, C.c_int32, f64p, f64p]),
    "srs_linescan_loss_grad_batch": (C.c_int, [C.c_void_p, C.c_int32, i32p, C.c_int32, C.c_int64, i32p, C.c_int32,
                                               C.POINTER(C.c_int64), f64p, C.c_int32, C.c_int32, f64p, f64p, f64p]),
    "srs_record_schedules": (C.c_int, [C.c_void_p, C.c_int64, i32p]),
    "srs_adjoint_last_plan": (C.c_int, [C.c_void_p, i32p, C.POINTER(C.c_int64), i32p]),
    "srs_explicit_step_prepare": (C.c_int, [C.c_void_p, C.c_int64, C.c_uint32]),
    "srs_checkpoint_size": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_size_t)]),
    "srs_checkpoint_save": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
    "srs_checkpoint_load": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_size_t]),
}

EXPORTED = tuple(_PROTOS)

_lib = None


class SrsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"srs error {code}: {msg}")
        self.code = code


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() or srs_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise SrsError(rc, lib().srs_last_error().decode(errors="replace"))
