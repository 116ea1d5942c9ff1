"""AdaptiveMesh for MI355X: Python bindings of the gfx950 C ABI.

The product is ``libadaptivemesh_amd.so`` (HIP kernels for gfx950 behind
``include/adaptivemesh.h`` plus the native host runtime behind
``include/adaptivemesh_host.h``).  This module loads it with ctypes and
mirrors its structs; it raises if the library is missing -- there is no
CPU fallback.  PyTorch is used only to own device memory: kernel entry
points take ``tensor.data_ptr()`` and the current HIP stream.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libadaptivemesh_amd.so"

AM_MAX_VARS = 32
AM_INTERP_LAGRANGE, AM_INTERP_TRILINEAR, AM_INTERP_TRICUBIC, AM_INTERP_LEARNED = 0, 1, 2, 3
AM_WORD_CF3, AM_WORD_F32, AM_WORD_F16 = 0, 1, 2
AM_EQUAL_FORMULA, AM_EQUAL_COPY, AM_EQUAL_SKIP = 0, 1, 2
AM_LAYOUT_PADDED, AM_LAYOUT_INTERIOR_FIRST = 0, 1

ERRORS = {
    -1: "AM_ERR_INVALID", -2: "AM_ERR_UNSUPPORTED_GEOMETRY", -3: "AM_ERR_SLOTS_EXHAUSTED",
    -4: "AM_ERR_DEPTH_LIMIT", -5: "AM_ERR_OCCUPIED", -6: "AM_ERR_MISALIGNED",
    -7: "AM_ERR_TOO_MANY_VARS", -8: "AM_ERR_ALLOC", -9: "AM_ERR_LAUNCH", -10: "AM_ERR_IO", -11: "AM_ERR_FORMAT",
}


class AMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Geometry(C.Structure):
    """am_geometry: S, P, FD order, outer_size, outer_dx, outer_dt."""

    _fields_ = [("size_domain", C.c_int32), ("padding_domain", C.c_int32), ("fd_order", C.c_int32),
                ("outer_size", C.c_float), ("outer_dx", C.c_float), ("outer_dt", C.c_float)]

    @property
    def S(self):
        return self.size_domain

    @property
    def P(self):
        return self.padding_domain

    @property
    def T(self):
        return self.size_domain + 2 * self.padding_domain


class Mesh(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("children", C.c_void_p), ("boundaries", C.c_void_p),
                ("tiles", C.c_void_p), ("n_tiles", C.c_int32), ("substep", C.c_uint32),
                ("max_depth", C.c_uint32), ("grid_limit", C.c_int32)]


class FdBatch(C.Structure):
    _fields_ = [("field", C.c_void_p * AM_MAX_VARS), ("d1", C.c_void_p * AM_MAX_VARS),
                ("d2diag", C.c_void_p * AM_MAX_VARS), ("d2off", C.c_void_p * AM_MAX_VARS),
                ("seed_d", C.c_int32 * AM_MAX_VARS), ("seed_m", C.c_int32 * AM_MAX_VARS),
                ("n_vars", C.c_int32), ("z_block", C.c_int32), ("seed_table", C.c_void_p),
                ("word_format", C.c_int32), ("prefetch_planes", C.c_int32),
                ("direct_faces", C.c_int32), ("tile_layout", C.c_int32), ("adv", C.c_void_p * AM_MAX_VARS),
                ("shift", C.c_void_p * 3), ("advect_mask", C.c_uint32), ("ko_sigma", C.c_float)]


class FieldBatch(C.Structure):
    _fields_ = [("old_", C.c_void_p * AM_MAX_VARS), ("new_", C.c_void_p * AM_MAX_VARS),
                ("older", C.c_void_p * AM_MAX_VARS), ("n_vars", C.c_int32), ("tile_layout", C.c_int32)]


class BssnParams(C.Structure):
    _fields_ = [("eta", C.c_float), ("shift_coeff", C.c_float), ("chi_floor", C.c_float),
                ("advect_shift", C.c_int32), ("slicing", C.c_int32)]


class BssnFields(C.Structure):
    _fields_ = [("state", C.c_void_p * 24), ("old_", C.c_void_p * 24), ("new_", C.c_void_p * 24),
                ("acc", C.c_void_p * 24), ("d1", C.c_void_p * 24), ("d2diag", C.c_void_p * 11),
                ("d2off", C.c_void_p * 11), ("word_format", C.c_int32), ("push_layers", C.c_int32),
                ("push_axes", C.c_int32), ("tile_layout", C.c_int32), ("adv", C.c_void_p * 24)]


class BssnSimOptions(C.Structure):
    _fields_ = [("params", BssnParams), ("interp", C.c_int32), ("z_block", C.c_int32), ("periodic", C.c_int32),
                ("seed", C.c_int32), ("word_format", C.c_int32)]


class _Trial(C.Structure):
    _fields_ = [("kind", C.c_int32), ("value", C.c_int32), ("ms", C.c_float)]


class TuneReport(C.Structure):
    _fields_ = [("z_block", C.c_int32), ("fd_grid_limit", C.c_int32), ("rhs_grid_limit", C.c_int32),
                ("n_trials", C.c_int32), ("trial", _Trial * 32), ("fd_prefetch", C.c_int32),
                ("direct_faces", C.c_int32), ("overlap_depths", C.c_int32)]


class MlpWork(C.Structure):
    _fields_ = [(k, C.c_void_p) for k in ("z1", "h1", "z2", "h2", "y", "dy", "dz2", "dz1")]


class SwaveFields(C.Structure):
    _fields_ = [("state", C.c_void_p * 2), ("old_", C.c_void_p * 2), ("new_", C.c_void_p * 2),
                ("acc", C.c_void_p * 2), ("d1_u", C.c_void_p), ("d2diag_u", C.c_void_p), ("push_layers", C.c_int32),
                ("push_axes", C.c_int32), ("tile_layout", C.c_int32)]


class SwaveSimOptions(C.Structure):
    _fields_ = [("interp", C.c_int32), ("z_block", C.c_int32), ("periodic", C.c_int32), ("seed", C.c_int32)]


BSSN_VARS = ["chi", "gt_xx", "gt_xy", "gt_xz", "gt_yy", "gt_yz", "gt_zz", "K", "At_xx", "At_xy", "At_xz", "At_yy",
             "At_yz", "At_zz", "Gt_x", "Gt_y", "Gt_z", "alpha", "beta_x", "beta_y", "beta_z", "B_x", "B_y", "B_z"]
BSSN_HESS_VARS = (0, 1, 2, 3, 4, 5, 6, 17, 18, 19, 20)

_lib = None

_P = C.c_void_p
_SIGS = {
    "am_last_error_string": (C.c_char_p, []),
    "am_version": (C.c_int, []),
    "am_geometry_init": (C.c_int, [C.POINTER(Geometry), C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float]),
    "am_geometry_supported": (C.c_int, [C.POINTER(Geometry)]),
    "am_device_alloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "am_device_free": (C.c_int, [_P]),
    "am_memcpy_h2d": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "am_memcpy_d2h": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "am_memset_async": (C.c_int, [_P, C.c_int, C.c_size_t, _P]),
    "am_memcpy_d2d": (C.c_int, [_P, _P, C.c_size_t, _P]),
    "am_stream_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "am_stream_destroy": (C.c_int, [_P]),
    "am_device_cu_count": (C.c_int32, []),
    "am_word_bytes": (C.c_int32, [C.c_int32]),
    "am_graph_begin": (C.c_int, [_P]),
    "am_graph_end": (C.c_int, [_P, C.POINTER(C.c_void_p)]),
    "am_graph_launch": (C.c_int, [_P, _P]),
    "am_graph_destroy": (C.c_int, [_P]),
    "am_seed_table_fill": (C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    "am_mlp_param_count": (C.c_int64, []),
    "am_gemm_f32": (C.c_int, [_P, C.c_int64, C.c_int64, _P, C.c_int64, C.c_int64, _P, C.c_int64, C.c_int32,
                              C.c_int32, C.c_int32, _P, C.c_int32, _P, _P, _P]),
    "am_mlp_forward": (C.c_int, [_P, _P, C.c_int32, _P, _P]),
    "am_mlp_backward": (C.c_int, [_P, _P, C.c_int32, _P, _P, _P]),
    "am_mlp_mse": (C.c_int, [_P, _P, C.c_int32, _P, _P, _P]),
    "am_mlp_mse_blocks": (C.c_int64, [C.c_int32]),
    "am_mlp_adam": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float,
                              _P]),
    "am_mlp_pairs_per_tile": (C.c_int64, [C.POINTER(Geometry)]),
    "am_ghost_fill_learned": (C.c_int, [C.POINTER(FieldBatch), C.POINTER(Mesh), C.POINTER(Geometry), C.c_float, _P,
                                        C.c_float, _P, C.c_int32, _P]),
    "am_mlp_pairs": (C.c_int, [C.POINTER(C.c_void_p), _P, C.c_int32, C.POINTER(Geometry), _P, _P, _P, _P, _P]),
    "am_stream_sync": (C.c_int, [_P]),
    "am_device_sync": (C.c_int, []),
    "am_fd_d1": (C.c_int, [_P, _P, C.POINTER(Mesh), C.POINTER(Geometry), C.c_int32, _P]),
    "am_fd_d1_d2diag": (C.c_int, [_P, _P, _P, C.POINTER(Mesh), C.POINTER(Geometry), C.c_int32, _P]),
    "am_fd_mixed": (C.c_int, [_P, _P, C.POINTER(Mesh), C.POINTER(Geometry), C.c_int32, _P]),
    "am_fd_all": (C.c_int, [_P, _P, _P, _P, C.POINTER(Mesh), C.POINTER(Geometry), C.c_int32, C.c_int32, _P]),
    "am_fd_batched": (C.c_int, [C.POINTER(FdBatch), C.POINTER(Mesh), C.POINTER(Geometry), _P]),
    "am_ghost_fill": (C.c_int, [_P, _P, _P, C.POINTER(Mesh), C.POINTER(Geometry), C.c_int32, C.c_int32, _P]),
    "am_ghost_fill_batched": (C.c_int, [C.POINTER(FieldBatch), C.POINTER(Mesh), C.POINTER(Geometry), C.c_int32,
                                        C.c_int32, _P]),
    "am_copy_new_to_old": (C.c_int, [_P, _P, C.POINTER(Mesh), C.POINTER(Geometry), _P]),
    "am_copy_new_to_old_batched": (C.c_int, [C.POINTER(FieldBatch), C.POINTER(Mesh), C.POINTER(Geometry), _P]),
    "am_copy_new_to_old_interior": (C.c_int, [C.POINTER(FieldBatch), C.POINTER(Mesh), _P, C.POINTER(Geometry), _P]),
    "am_amr_criterion": (C.c_int, [_P, _P, C.POINTER(Mesh), C.POINTER(Geometry), _P, _P]),
    "am_copy_to_child": (C.c_int, [_P, C.c_uint32, C.c_int32, C.c_int32, C.POINTER(Geometry), C.c_int32, _P]),
    "am_copy_to_child_batched": (C.c_int, [C.POINTER(FieldBatch), C.c_uint32, C.c_int32, C.c_int32,
                                           C.POINTER(Geometry), C.c_int32, _P]),
    "am_wave_predictor": (C.c_int, [_P, _P, _P, _P, C.POINTER(Mesh), C.POINTER(Geometry), _P]),
    "am_wave_corrector": (C.c_int, [_P, _P, _P, _P, C.POINTER(Mesh), C.POINTER(Geometry), _P]),
    "am_deriv_trace": (C.c_int, [_P, _P, C.POINTER(C.c_void_p), _P, C.POINTER(Mesh), C.POINTER(Geometry), _P]),
    "am_christoffel": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int32,
                                 C.POINTER(Mesh), C.POINTER(Geometry), C.c_int32, _P]),
    "am_tile_lists_workspace": (C.c_int32, [C.c_int32, C.c_int32]),
    "am_tile_lists": (C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, C.c_int32, _P, _P, _P, _P, _P, _P]),
    # host runtime
    "am_tree_create": (_P, [C.c_uint32, C.c_int32]),
    "am_tree_destroy": (None, [_P]),
    "am_tree_add_node": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "am_tree_remove_node": (C.c_int, [_P, C.c_int32]),
    "am_tree_increment_timer": (C.c_int, [_P]),
    "am_tree_find_node_idx": (C.c_uint32, [_P, C.c_uint32, C.c_uint32]),
    "am_tree_export": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "am_tree_mesh": (C.c_int, [_P, C.c_int32, C.POINTER(Mesh)]),
    "am_wave_sim_create": (_P, [C.POINTER(Geometry), C.c_uint32, C.c_int32]),
    "am_wave_sim_destroy": (None, [_P]),
    "am_wave_sim_add_node": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "am_wave_sim_export": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "am_wave_sim_set_field": (C.c_int, [_P, C.c_int32, _P]),
    "am_wave_sim_get_field": (C.c_int, [_P, C.c_int32, _P]),
    "am_wave_sim_copy_bounds": (C.c_int, [_P]),
    "am_wave_sim_timestep": (C.c_int, [_P]),
    "am_wave_sim_rng_state": (C.c_int32, [_P]),
    "am_wave_sim_set_regrid": (C.c_int, [_P, C.c_int32]),
    "am_wave_sim_pressure": (C.c_int, [_P, _P]),
    "am_wave_sim_save": (C.c_int, [_P, C.c_char_p]),
    "am_checkpoint_inspect": (C.c_int, [C.c_char_p, _P]),
    "am_wave_sim_raster": (C.c_int, [_P, C.c_int32, C.c_uint32, C.c_uint32, C.c_float, C.c_int32, _P]),
    "am_rasterise_slice": (C.c_int, [_P, C.POINTER(Mesh), C.POINTER(Geometry), C.c_uint32, C.c_uint32, C.c_float,
                                     C.c_int32, _P, _P]),
    "am_encode_rgba8": (C.c_int, [_P, _P, C.c_int64, _P]),
    "am_encode_rgba8_f32": (C.c_int, [_P, _P, C.c_int64, _P]),
    "am_encode_rgba8_cf3": (C.c_int, [_P, _P, C.c_int64, _P]),
    "am_init_temp": (C.c_int, [_P, _P, C.POINTER(Mesh), C.POINTER(Geometry), C.c_float, _P]),
    "am_wave_sim_init_temp": (C.c_int, [_P, C.c_float]),
    "am_write_png": (C.c_int, [C.c_char_p, _P, C.c_uint32, C.c_uint32]),
    "am_wave_sim_load": (C.c_int, [_P, C.c_char_p]),
    "am_ghost_fill_stage_batched": (C.c_int, [C.POINTER(FieldBatch), C.POINTER(Mesh), C.POINTER(Geometry), C.c_float,
                                              C.c_int32, C.c_int32, _P]),
    "am_ghost_fill_stage_faces": (C.c_int, [C.POINTER(FieldBatch), C.POINTER(Mesh), _P, C.c_int32,
                                            C.POINTER(Geometry), C.c_float, C.c_int32, C.c_int32, _P]),
    "am_ghost_fill_learned_faces": (C.c_int, [C.POINTER(FieldBatch), C.POINTER(Mesh), _P, C.c_int32,
                                              C.POINTER(Geometry), C.c_float, _P, C.c_float, _P, _P, C.c_int32,
                                              _P]),
    "am_bssn_rhs_stage": (C.c_int, [C.POINTER(BssnFields), C.c_int32, C.POINTER(BssnParams), C.POINTER(Mesh),
                                    C.POINTER(Geometry), C.c_int32, _P]),
    "am_bssn_constraints": (C.c_int, [C.POINTER(BssnFields), C.POINTER(BssnParams), C.POINTER(Mesh),
                                      C.POINTER(Geometry), _P, _P]),
    "am_bssn_constraint_blocks": (C.c_int64, [C.POINTER(Geometry), C.c_int32]),
    "am_tile_stride": (C.c_int64, [C.POINTER(Geometry), C.c_int32]),
    "am_bssn_sim_set_dissipation": (C.c_int, [_P, C.c_float, C.c_int32]),
    "am_bssn_sim_get_dissipation": (C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_int32)]),
    "am_tile_cell_offsets": (C.c_int, [C.POINTER(Geometry), C.c_int32, _P]),
    "am_tile_layout_convert": (C.c_int, [_P, _P, C.c_int32, C.POINTER(Geometry), C.c_int32, C.c_int32, _P]),
    "am_bssn_constraints_noalloc": (C.c_int, [C.POINTER(BssnFields), C.POINTER(BssnParams), C.POINTER(Mesh),
                                              C.POINTER(Geometry), _P, _P]),
    "am_event_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "am_event_record": (C.c_int, [_P, _P]),
    "am_event_create_ordering": (C.c_int, [C.POINTER(C.c_void_p)]),
    "am_stream_wait_event": (C.c_int, [_P, _P]),
    "am_event_elapsed_ms": (C.c_int, [_P, _P, C.POINTER(C.c_float)]),
    "am_event_destroy": (C.c_int, [_P]),
    "am_bssn_sim_create": (_P, [C.POINTER(Geometry), C.c_uint32, C.POINTER(BssnSimOptions)]),
    "am_bssn_sim_constraints": (C.c_int, [_P, _P]),
    "am_swave_sim_create": (_P, [C.POINTER(Geometry), C.c_uint32, C.POINTER(SwaveSimOptions)]),
    "am_swave_sim_energy": (C.c_int, [_P, _P]),
    "am_swave_rhs_stage": (C.c_int, [C.POINTER(SwaveFields), C.c_int32, C.POINTER(Mesh), C.POINTER(Geometry),
                                     C.c_int32, _P]),
    "am_swave_energy": (C.c_int, [C.POINTER(SwaveFields), C.POINTER(Mesh), C.POINTER(Geometry), _P, _P]),
}

# the RK4 driver families share one call set (adaptivemesh_host.h)
_RK4_SIM_SIGS = {
    "destroy": (None, [_P]),
    "add_node": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "export": (C.c_int, [_P, _P, _P, _P, _P, _P, _P]),
    "set_field": (C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    "get_field": (C.c_int, [_P, C.c_int32, C.c_int32, _P]),
    "device_ptr": (_P, [_P, C.c_int32]),
    "fill_ghosts": (C.c_int, [_P]),
    "copy_back": (C.c_int, [_P]),
    "step_depth": (C.c_int, [_P, C.c_int32]),
    "stage_depth": (C.c_int, [_P, C.c_int32, C.c_int32]),
    "substep": (C.c_int, [_P]),
    "timestep": (C.c_int, [_P]),
    "sync": (C.c_int, [_P]),
    "zone_cycles": (C.c_uint64, [_P]),
    "rng_state": (C.c_int32, [_P]),
    "set_rng_state": (C.c_int, [_P, C.c_int32]),
    "counter": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_int32)]),
    "profile": (C.c_int, [_P, C.c_int32]),
    "kernel_times": (C.c_int, [_P, _P, _P]),
    "set_launch": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32]),
    "get_launch": (C.c_int, [_P, _P]),
    "set_fd_prefetch": (C.c_int, [_P, C.c_int32]),
    "set_push_layers": (C.c_int, [_P, C.c_int32]),
    "get_push_layers": (C.c_int, [_P]),
    "set_direct_faces": (C.c_int, [_P, C.c_int32]),
    "get_direct_faces": (C.c_int, [_P]),
    "set_overlap_depths": (C.c_int, [_P, C.c_int32]),
    "get_overlap_depths": (C.c_int, [_P]),
    "use_graph": (C.c_int, [_P, C.c_int32]),
    "autotune": (C.c_int, [_P, C.c_int32, _P]),
    "set_prolongation": (C.c_int, [_P, _P, C.c_float]),
    "prolongation_stats": (C.c_int, [_P, _P]),
    "save": (C.c_int, [_P, C.c_char_p]),
    "load": (C.c_int, [_P, C.c_char_p]),
    "raster": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_uint32, C.c_uint32, C.c_float, C.c_int32, _P]),
}
for _prefix in ("am_bssn_sim", "am_swave_sim"):
    for _k, _v in _RK4_SIM_SIGS.items():
        _SIGS[f"{_prefix}_{_k}"] = _v


def exported_symbols() -> list[str]:
    return list(_SIGS)


def lib() -> C.CDLL:
    """Load the gfx950 library (built by adaptivemesh_amd.build).  Raises if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m adaptivemesh_amd.build` "
                               "(there is no CPU fallback)")
        # One HIP runtime per process: PyTorch ships its own libamdhip64.so.7.
        # Loading torch first makes this library bind to that copy (same
        # SONAME) instead of starting a second runtime that would find no
        # device once torch holds it.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> int:
    if rc != 0:
        raise AMError(rc, f"{what}: {lib().am_last_error_string().decode()}")
    return rc


class CheckpointHeader(C.Structure):
    """am_checkpoint_header (adaptivemesh_host.h)."""
    _fields_ = [("magic", C.c_char * 8), ("kind", C.c_uint32), ("geometry", Geometry), ("max_slots", C.c_uint32),
                ("periodic", C.c_uint32), ("max_depth", C.c_int32), ("first_free_slot", C.c_uint32),
                ("final_used_slot", C.c_uint32), ("timer", C.c_uint32), ("seed", C.c_int32), ("n_vars", C.c_uint32),
                ("flags", C.c_uint32), ("n_sections", C.c_uint32)]


def write_png(path, pixels: np.ndarray):
    """am_write_png: a (height, width) array of RGBA8 words (R in the low byte) to a PNG file."""
    px = np.ascontiguousarray(pixels, dtype=np.uint32)
    if px.ndim != 2:
        raise ValueError("pixels must be (height, width)")
    check(lib().am_write_png(os.fsencode(path), px.ctypes.data, px.shape[1], px.shape[0]), "am_write_png")


def inspect_checkpoint(path) -> CheckpointHeader:
    """Check a checkpoint file natively (no GPU) and return its header."""
    h = CheckpointHeader()
    check(lib().am_checkpoint_inspect(os.fsencode(path), C.byref(h)), "am_checkpoint_inspect")
    return h


def geometry(S: int = 16, P: int = 3, fd_order: int = 0, outer_size: float = 100.0, cfl: float = 0.4) -> Geometry:
    g = Geometry()
    check(lib().am_geometry_init(C.byref(g), S, P, fd_order, outer_size, cfl), "am_geometry_init")
    return g


def interior_first_index(S: int, P: int) -> np.ndarray:
    """Offset within its slot, in the interior-first tile layout
    (am_tile_layout), of every cell of a padded tile, in padded order
    (index (z * T + y) * T + x): the interior block, the six face slabs
    (face f = +x +y +z -x -y -z, layer counted outward, in-face cell (a, b)
    = the other two axes in x, y, z order), the edges, the corners."""
    T = S + 2 * P
    z, y, x = np.meshgrid(np.arange(T), np.arange(T), np.arange(T), indexing="ij")
    c = np.stack([x.ravel(), y.ravel(), z.ravel()])
    inner = c - P
    inside = (inner >= 0) & (inner < S)
    code = np.where(c < P, c, c - S)
    n_out = 3 - inside.sum(0)
    out = np.empty(T ** 3, np.int64)
    S2, G = S * S, 2 * P
    F0, E0 = S ** 3, S ** 3 + 6 * P * S2
    C0 = E0 + 3 * S * G * G
    m = n_out == 0
    out[m] = (inner[2, m] * S + inner[1, m]) * S + inner[0, m]
    for k in range(3):
        a_ax, b_ax = (1, 2) if k == 0 else ((0, 2) if k == 1 else (0, 1))
        m = (n_out == 1) & ~inside[k]
        high = c[k, m] >= P + S
        f = np.where(high, k, k + 3)
        l = np.where(high, c[k, m] - P - S, P - 1 - c[k, m])
        out[m] = F0 + ((f * P + l) * S + inner[b_ax, m]) * S + inner[a_ax, m]
        o0, o1 = (1, 2) if k == 0 else ((0, 2) if k == 1 else (0, 1))
        m = (n_out == 2) & inside[k]
        out[m] = E0 + ((k * S + inner[k, m]) * G + code[o0, m]) * G + code[o1, m]
    m = n_out == 3
    out[m] = C0 + (code[0, m] * G + code[1, m]) * G + code[2, m]
    return out


def ptr(x) -> int | None:
    """Device address of a torch tensor (or None)."""
    if x is None:
        return None
    return x.data_ptr()


def stream_handle():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Tree:
    """The native CPU-managed octree (AMR.h:118-294) through the C ABI."""

    def __init__(self, max_slots: int, periodic: bool = False):
        self.max_slots = max_slots
        self.h = lib().am_tree_create(max_slots, 1 if periodic else 0)
        if not self.h:
            raise AMError(-8, lib().am_last_error_string().decode())

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.am_tree_destroy(h)
            self.h = None

    def add_node(self, parent: int, offset: int) -> int:
        return lib().am_tree_add_node(self.h, parent, offset)

    def remove_node(self, node: int) -> int:
        return lib().am_tree_remove_node(self.h, node)

    def increment_timer(self):
        check(lib().am_tree_increment_timer(self.h), "increment_timer")

    def find_node_idx(self, packed: int, max_depth: int = 0xFFFFFFFF) -> int:
        return lib().am_tree_find_node_idx(self.h, packed, max_depth)

    def export(self) -> dict:
        n = self.max_slots
        out = dict(parent=np.empty(n, np.int32), children=np.empty(8 * n, np.int32),
                   positions=np.empty(n, np.uint32), boundaries=np.empty(6 * n, np.uint64),
                   lifetime=np.empty(n, np.uint32), scalars=np.empty(4, np.int32))
        check(lib().am_tree_export(self.h, *(a.ctypes.data for a in out.values())), "am_tree_export")
        s = out["scalars"]
        out.update(max_depth=int(s[0]), first_free=int(s[1]), final_used=int(s[2]), timer=int(s[3]))
        return out

    def mesh(self, depth: int = -1) -> Mesh:
        m = Mesh()
        check(lib().am_tree_mesh(self.h, depth, C.byref(m)), "am_tree_mesh")
        return m


class WaveSim:
    """AMR<wave_AMR_data>: the reference SimpleWave driver (SimpleWave.h:191-279)."""

    def __init__(self, geom: Geometry, max_slots: int, periodic: bool = False):
        self.geom = geom
        self.max_slots = max_slots
        self.h = lib().am_wave_sim_create(C.byref(geom), max_slots, 1 if periodic else 0)
        if not self.h:
            raise AMError(-8, lib().am_last_error_string().decode())

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            _lib.am_wave_sim_destroy(h)
            self.h = None

    def add_node(self, parent: int, offset: int) -> int:
        r = lib().am_wave_sim_add_node(self.h, parent, offset)
        if r < 0:
            raise AMError(r, lib().am_last_error_string().decode())
        return r

    def export(self) -> dict:
        n = self.max_slots
        out = dict(parent=np.empty(n, np.int32), children=np.empty(8 * n, np.int32),
                   positions=np.empty(n, np.uint32), boundaries=np.empty(6 * n, np.uint64),
                   lifetime=np.empty(n, np.uint32), scalars=np.empty(4, np.int32))
        check(lib().am_wave_sim_export(self.h, *(a.ctypes.data for a in out.values())), "am_wave_sim_export")
        return out

    def set_field(self, which: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        assert a.size == self.max_slots * self.geom.T ** 3
        check(lib().am_wave_sim_set_field(self.h, which, a.ctypes.data), "set_field")

    def get_field(self, which: int) -> np.ndarray:
        a = np.empty(self.max_slots * self.geom.T ** 3, np.float32)
        check(lib().am_wave_sim_get_field(self.h, which, a.ctypes.data), "get_field")
        return a

    def copy_bounds(self):
        check(lib().am_wave_sim_copy_bounds(self.h), "copy_bounds")

    def timestep(self):
        check(lib().am_wave_sim_timestep(self.h), "timestep")

    def set_regrid(self, on: bool):
        """modify_nodes on (the default, as in the reference) or off."""
        check(lib().am_wave_sim_set_regrid(self.h, 1 if on else 0), "set_regrid")

    def pressure(self) -> tuple:
        """The two average refinement pressures of the last modify_nodes."""
        out = np.zeros(2, np.float32)
        check(lib().am_wave_sim_pressure(self.h, out.ctypes.data), "pressure")
        return float(out[0]), float(out[1])

    def save(self, path):
        """Checkpoint: octree, counter, seed, fields and derivative words."""
        check(lib().am_wave_sim_save(self.h, os.fsencode(path)), "save")

    def load(self, path):
        """Resume from a checkpoint of a driver built alike."""
        check(lib().am_wave_sim_load(self.h, os.fsencode(path)), "load")

    def init_temp(self, mul: float = 1.0):
        """The reference wave test's initial data (kernel.cu:9-43) in both fields."""
        check(lib().am_wave_sim_init_temp(self.h, mul), "init_temp")

    def raster(self, which: int = 1, width: int = 1024, height: int = 1024, z_slice: float = 1e-9,
               boxes: bool = True) -> np.ndarray:
        """rasterise_to_image of old (0) or new (1) as (height, width) RGBA8 words."""
        px = np.empty((height, width), np.uint32)
        check(lib().am_wave_sim_raster(self.h, which, width, height, z_slice, 1 if boxes else 0, px.ctypes.data),
              "raster")
        return px

    def save_image(self, path, which: int = 1, width: int = 1024, height: int = 1024, z_slice: float = 1e-9,
                   boxes: bool = True):
        """rasterise_to_image + save_image (AMR.h:1355, image_process.h:55)."""
        write_png(path, self.raster(which, width, height, z_slice, boxes))

    @property
    def rng_state(self) -> int:
        return lib().am_wave_sim_rng_state(self.h)


def bssn_options(eta=2.0, shift_coeff=0.75, chi_floor=1e-4, advect_shift=1, slicing=0,
                 interp=AM_INTERP_LAGRANGE, z_block=0, periodic=False, seed=12894712,
                 word_format=AM_WORD_CF3) -> BssnSimOptions:
    o = BssnSimOptions()
    o.params.eta, o.params.shift_coeff, o.params.chi_floor = eta, shift_coeff, chi_floor
    o.params.advect_shift, o.params.slicing = advect_shift, slicing
    o.interp, o.z_block, o.periodic, o.seed = interp, z_block, 1 if periodic else 0, seed
    o.word_format = word_format
    return o


KERNEL_CLASSES = ("fd", "christoffel", "rhs_stage", "ghost", "copy_back")


class _Rk4Sim:
    """Shared binding of the RK4 driver calls (adaptivemesh_host.h); PREFIX
    names the exported family."""
    PREFIX = ""
    NV = 0

    def _f(self, name):
        return getattr(lib(), f"{self.PREFIX}_{name}")

    def _create(self, geom: Geometry, max_slots: int, options):
        self.geom = geom
        self.max_slots = max_slots
        self.options = options
        self.h = self._f("create")(C.byref(geom), max_slots, C.byref(options))
        if not self.h:
            raise AMError(-8, lib().am_last_error_string().decode())

    def __del__(self):
        h = getattr(self, "h", None)
        if h and _lib is not None:
            getattr(_lib, f"{self.PREFIX}_destroy")(h)
            self.h = None

    def add_node(self, parent: int, offset: int) -> int:
        r = self._f("add_node")(self.h, parent, offset)
        if r < 0:
            raise AMError(r, lib().am_last_error_string().decode())
        return r

    def export(self) -> dict:
        n = self.max_slots
        out = dict(parent=np.empty(n, np.int32), children=np.empty(8 * n, np.int32),
                   positions=np.empty(n, np.uint32), boundaries=np.empty(6 * n, np.uint64),
                   lifetime=np.empty(n, np.uint32), scalars=np.empty(4, np.int32))
        check(self._f("export")(self.h, *(a.ctypes.data for a in out.values())), "export")
        s = out["scalars"]
        out.update(max_depth=int(s[0]), first_free=int(s[1]), final_used=int(s[2]), timer=int(s[3]))
        return out

    def set_field(self, which: int, var: int, arr: np.ndarray):
        a = np.ascontiguousarray(arr, dtype=np.float32)
        assert a.size == self.max_slots * self.geom.T ** 3
        check(self._f("set_field")(self.h, which, var, a.ctypes.data), "set_field")

    def get_field(self, which: int, var: int) -> np.ndarray:
        a = np.empty(self.max_slots * self.geom.T ** 3, np.float32)
        check(self._f("get_field")(self.h, which, var, a.ctypes.data), "get_field")
        return a

    def save(self, path):
        """Checkpoint: octree, counter, seed and the old/new(/older) fields."""
        check(self._f("save")(self.h, os.fsencode(path)), "save")

    def load(self, path):
        """Resume from a checkpoint of a driver built alike."""
        check(self._f("load")(self.h, os.fsencode(path)), "load")

    def raster(self, var: int, which: int = 1, width: int = 1024, height: int = 1024, z_slice: float = 1e-9,
               boxes: bool = True) -> np.ndarray:
        """A z-slice of variable `var` of old (0) or new (1) as (height, width) RGBA8 words."""
        px = np.empty((height, width), np.uint32)
        check(self._f("raster")(self.h, which, var, width, height, z_slice, 1 if boxes else 0, px.ctypes.data),
              "raster")
        return px

    def save_image(self, path, var: int, which: int = 1, width: int = 1024, height: int = 1024,
                   z_slice: float = 1e-9, boxes: bool = True):
        write_png(path, self.raster(var, which, width, height, z_slice, boxes))

    def device_ptr(self, which: int) -> int:
        """Raw device pointer of a driver buffer: 0 old, 1 new, 2 acc, 3
        older (fields in the interior-first layout, field v at v *
        max_slots * tile_stride floats; acc interior-only), 4-7 the words, 8
        the upwinded advection + dissipation (interior-only, after
        set_dissipation)."""
        return self._f("device_ptr")(self.h, which)

    @property
    def tile_stride(self) -> int:
        """Floats per slot of the driver's state fields (interior-first layout)."""
        return int(lib().am_tile_stride(C.byref(self.geom), AM_LAYOUT_INTERIOR_FIRST))

    def field_ptr(self, which: int, var: int) -> int:
        """Device pointer of field `var` of buffer `which` (0 old, 1 new, 3 older)."""
        return self.device_ptr(which) + var * self.max_slots * self.tile_stride * 4

    def padded_fields(self, which: int, out=None):
        """Every field of buffer `which` (0 old, 1 new, 3 older) converted on the
        device to the reference's padded layout: a (NV, max_slots * T^3)
        float32 torch tensor on the GPU (`out` if given)."""
        import torch
        T = self.geom.T
        if out is None:
            out = torch.empty((self.NV, self.max_slots * T ** 3), dtype=torch.float32, device="cuda")
        self.sync()
        for v in range(self.NV):
            check(lib().am_tile_layout_convert(self.field_ptr(which, v), out[v].data_ptr(), self.max_slots,
                                               C.byref(self.geom), AM_LAYOUT_INTERIOR_FIRST, AM_LAYOUT_PADDED,
                                               stream_handle()), "am_tile_layout_convert")
        torch.cuda.synchronize()
        return out

    def fill_ghosts(self):
        check(self._f("fill_ghosts")(self.h), "fill_ghosts")

    def copy_back(self):
        check(self._f("copy_back")(self.h), "copy_back")

    def step_depth(self, d: int):
        check(self._f("step_depth")(self.h, d), "step_depth")

    def stage_depth(self, d: int, stage: int):
        check(self._f("stage_depth")(self.h, d, stage), "stage_depth")

    def substep(self):
        check(self._f("substep")(self.h), "substep")

    def timestep(self):
        check(self._f("timestep")(self.h), "timestep")

    def sync(self):
        check(self._f("sync")(self.h), "sync")

    @property
    def zone_cycles(self) -> int:
        return self._f("zone_cycles")(self.h)

    @property
    def rng_state(self) -> int:
        return self._f("rng_state")(self.h)

    @rng_state.setter
    def rng_state(self, v: int):
        check(self._f("set_rng_state")(self.h, v), "rng")

    def counter(self):
        t, d = C.c_uint32(), C.c_int32()
        check(self._f("counter")(self.h, C.byref(t), C.byref(d)), "counter")
        return t.value, d.value

    def profile(self, on: bool):
        check(self._f("profile")(self.h, 1 if on else 0), "profile")

    def set_launch(self, z_block: int = 0, fd_grid_limit: int = 0, rhs_grid_limit: int = 0):
        check(self._f("set_launch")(self.h, z_block, fd_grid_limit, rhs_grid_limit), "set_launch")

    def set_fd_prefetch(self, planes: int):
        check(self._f("set_fd_prefetch")(self.h, planes), "set_fd_prefetch")

    def set_push_layers(self, layers: int):
        """Equal-depth ghost layers refreshed per stage (0 = fd_order / 2, P = all)."""
        check(self._f("set_push_layers")(self.h, layers), "set_push_layers")

    @property
    def push_layers(self) -> int:
        return self._f("get_push_layers")(self.h)

    def set_direct_faces(self, axes: int):
        """Axes (bits 1 x, 2 y, 4 z; default 4) whose equal-depth faces
        stages 0-2 do not push: the next stage's FD reads them from the
        neighbours.  Stage 3 pushes every face."""
        check(self._f("set_direct_faces")(self.h, int(axes)), "set_direct_faces")

    @property
    def direct_faces(self) -> int:
        return self._f("get_direct_faces")(self.h)

    def set_overlap_depths(self, on: bool):
        """Run the coarser active depths of a sub-step on a second stream
        beside the finest (when the finest has no coarser face); the
        results do not change."""
        check(self._f("set_overlap_depths")(self.h, 1 if on else 0), "set_overlap_depths")

    @property
    def overlap_depths(self) -> bool:
        return bool(self._f("get_overlap_depths")(self.h))

    def get_launch(self) -> tuple:
        """(z_block, fd_grid_limit, rhs_grid_limit, fd_prefetch)"""
        out = np.zeros(4, np.int32)
        check(self._f("get_launch")(self.h, out.ctypes.data), "get_launch")
        return tuple(int(x) for x in out)

    def use_graph(self, on: bool = True):
        check(self._f("use_graph")(self.h, 1 if on else 0), "use_graph")

    def autotune(self, reps: int = 5) -> dict:
        r = TuneReport()
        check(self._f("autotune")(self.h, reps, C.byref(r)), "autotune")
        kinds = ("fd_z_block", "fd_grid_limit", "rhs_grid_limit", "fd_prefetch", "direct_faces", "overlap_depths")
        return {"z_block": r.z_block, "fd_grid_limit": r.fd_grid_limit, "rhs_grid_limit": r.rhs_grid_limit,
                "fd_prefetch": r.fd_prefetch, "direct_faces": r.direct_faces,
                "overlap_depths": bool(r.overlap_depths),
                "trials": [(kinds[r.trial[i].kind], r.trial[i].value, round(r.trial[i].ms, 4))
                           for i in range(r.n_trials)]}

    def set_prolongation(self, params, threshold: float = 1e-5):
        """params: a device tensor of am_mlp_param_count() floats (kept
        referenced here while the driver uses it) or an object with a
        `params` tensor (learned.Prolongation)."""
        t = getattr(params, "params", params)
        self._mlp_params = t
        check(self._f("set_prolongation")(self.h, t.data_ptr(), threshold), "set_prolongation")

    def prolongation_stats(self) -> dict:
        out = np.zeros(4, np.uint64)
        check(self._f("prolongation_stats")(self.h, out.ctypes.data), "prolongation_stats")
        return dict(learned_faces=int(out[0]), fallback_faces=int(out[1]), tricubic_cells=int(out[2]),
                    cells=int(out[3]))

    def kernel_times(self) -> dict:
        ms = np.zeros(len(KERNEL_CLASSES))
        cnt = np.zeros(len(KERNEL_CLASSES), np.int64)
        check(self._f("kernel_times")(self.h, ms.ctypes.data, cnt.ctypes.data), "kernel_times")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNEL_CLASSES)}


class BssnSim(_Rk4Sim):
    """AMR<BSSN_rk4_data>: 24-field BSSN with RK4 and sub-cycling, native."""
    PREFIX = "am_bssn_sim"
    NV = 24

    def __init__(self, geom: Geometry, max_slots: int, options: BssnSimOptions | None = None):
        self._create(geom, max_slots, options or bssn_options())

    def constraints(self) -> dict:
        out = np.zeros(4)
        check(lib().am_bssn_sim_constraints(self.h, out.ctypes.data), "constraints")
        return dict(H_l2=out[0], M_l2=out[1], H_max=out[2], cells=out[3])

    def set_dissipation(self, ko_sigma: float, upwind: bool = True):
        """Upwinded advection with Kreiss-Oliger dissipation of strength
        ko_sigma (upwind=False, ko_sigma=0: the centred scheme)."""
        check(lib().am_bssn_sim_set_dissipation(self.h, float(ko_sigma), 1 if upwind else 0), "set_dissipation")

    @property
    def dissipation(self):
        sg, up = C.c_float(), C.c_int32()
        check(lib().am_bssn_sim_get_dissipation(self.h, C.byref(sg), C.byref(up)), "get_dissipation")
        return sg.value, bool(up.value)


def swave_options(interp=AM_INTERP_LAGRANGE, z_block=0, periodic=True, seed=12894712) -> SwaveSimOptions:
    o = SwaveSimOptions()
    o.interp, o.z_block, o.periodic, o.seed = interp, z_block, 1 if periodic else 0, seed
    return o


class SwaveSim(_Rk4Sim):
    """AMR<swave_rk4_data>: u_tt = lap u as (u, v) with RK4 and sub-cycling."""
    PREFIX = "am_swave_sim"
    NV = 2

    def __init__(self, geom: Geometry, max_slots: int, options: SwaveSimOptions | None = None):
        self._create(geom, max_slots, options or swave_options())

    def energy(self) -> dict:
        out = np.zeros(4)
        check(lib().am_swave_sim_energy(self.h, out.ctypes.data), "energy")
        return dict(E=out[0], u2=out[1], u_max=out[2], cells=out[3])
