"""ctypes binding of the C ABI in include/atdf.h (libatdf.so).

Pointers are device addresses (``tensor.data_ptr()``); the stream is torch's
current HIP stream unless one is given.  Every call raises ``AtdfError`` on a
non-zero status, with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

from . import LIB_PATH

OK = 0
ERR_INVALID_ARGUMENT = 1
ERR_CAPACITY = 2
ERR_TOO_MANY_INTERSECTIONS = 3
ERR_ORDER = 4
ERR_HIP = 5
ERR_UNSUPPORTED = 6

GRID_F32 = 0
GRID_BF16 = 1
DECODER_F32 = 0
DECODER_BF16 = 1


class AtdfError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"atdf status {status}: {msg}")
        self.status = status


class Tree(C.Structure):
    _fields_ = [("num_nodes", C.c_int64), ("node_position_min", C.c_void_p), ("node_position_max", C.c_void_p),
                ("node_children", C.c_void_p), ("node_parent", C.c_void_p), ("node_active", C.c_void_p),
                ("children_per_node", C.c_int32)]


class Sampling(C.Structure):
    _fields_ = [("max_samples_per_node", C.c_int32), ("max_intersections", C.c_int32), ("jitter_mode", C.c_int32),
                ("reserved", C.c_int32), ("jitter_seed", C.c_uint64), ("jitter", C.c_void_p),
                ("ray_select", C.c_void_p), ("ray_list", C.c_void_p)]


class GridAdam(C.Structure):
    _fields_ = [("param", C.c_void_p), ("param_bf16", C.c_void_p), ("exp_avg", C.c_void_p),
                ("exp_avg_sq", C.c_void_p), ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("step", C.c_int64)]


class Lattice(C.Structure):
    _fields_ = [("size", C.c_int32 * 3), ("mode", C.c_int32), ("swap_xy", C.c_int32), ("roi_min", C.c_float * 3),
                ("step", C.c_float * 3)]


class CameraRig(C.Structure):
    _fields_ = [("camera_model", C.c_int32), ("w", C.c_int32), ("h", C.c_int32), ("align_corners", C.c_int32),
                ("num_cameras", C.c_int64), ("num_images", C.c_int64), ("intrinsics", C.c_void_p),
                ("rotation", C.c_void_p), ("translation", C.c_void_p), ("volume_translation", C.c_float * 3)]


class ProjectionSet(C.Structure):
    _fields_ = [("projection", C.c_void_p), ("mask", C.c_void_p), ("w", C.c_int32), ("h", C.c_int32),
                ("num_images", C.c_int64), ("align_corners", C.c_int32)]


class TVShape(C.Structure):
    _fields_ = [("B", C.c_int64), ("C", C.c_int64), ("D", C.c_int32), ("n", C.c_int32 * 3), ("stride_b", C.c_int64),
                ("stride_c", C.c_int64), ("stride_s", C.c_int64 * 3)]


class NeighborSpec(C.Structure):
    _fields_ = [("size", C.c_int32 * 3), ("epsilon", C.c_float), ("use_roi", C.c_int32), ("roi_min", C.c_float * 3),
                ("roi_max", C.c_float * 3), ("variant", C.c_int32), ("plane", C.c_int32)]


class Intervals(C.Structure):
    _fields_ = [("node", C.c_void_p), ("tmin", C.c_void_p), ("tmax", C.c_void_p), ("count", C.c_void_p),
                ("first", C.c_void_p), ("ray", C.c_void_p), ("total", C.c_int64)]


class Proposal(C.Structure):
    _fields_ = [("leaf_table", C.c_void_p), ("table_depth", C.c_int32), ("mix", C.c_float),
                ("kl_threshold", C.c_float), ("jitter_mode", C.c_int32), ("jitter_seed", C.c_uint64),
                ("coarse", C.c_void_p), ("params", C.c_void_p)]


class Samples(C.Structure):
    _fields_ = [("global_coordinate", C.c_void_p), ("weight", C.c_void_p), ("ray_t", C.c_void_p),
                ("ray_index", C.c_void_p), ("local_index_in_ray", C.c_void_p), ("node_id", C.c_void_p)]


class VM(C.Structure):
    """atdf_vm: the tensorial (plane x line) field of the in-ROI nodes."""
    _fields_ = [("planes", C.c_void_p), ("lines", C.c_void_p), ("num_planes", C.c_int64),
                ("node_to_plane", C.c_void_p), ("Gp", C.c_int32), ("Gl", C.c_int32), ("roi_min_x", C.c_float),
                ("roi_max_x", C.c_float), ("grad_planes", C.c_void_p), ("grad_lines", C.c_void_p)]


class Field(C.Structure):
    _fields_ = [("C", C.c_int32), ("H", C.c_int32), ("G", C.c_int32), ("grid_dtype", C.c_int32),
                ("softplus_beta", C.c_float), ("rays_per_wave", C.c_int32), ("grid", C.c_void_p),
                ("num_grids", C.c_int64), ("node_to_grid", C.c_void_p), ("mlp", C.c_void_p),
                ("decoder_dtype", C.c_int32), ("reserved", C.c_int32), ("vm", C.POINTER(VM))]


_P, _I32, _I64, _F, _SZ, _U64 = C.c_void_p, C.c_int32, C.c_int64, C.c_float, C.c_size_t, C.c_uint64

# name -> (restype, argtypes)
_SIGNATURES = {
    "atdf_last_error": (C.c_char_p, []),
    "atdf_version": (C.c_char_p, []),
    "atdf_tree_packed_bytes": (_SZ, [_I64]),
    "atdf_tree_pack": (_I32, [C.POINTER(Tree), _P, _P]),
    "atdf_tree_pack_box": (_I32, [C.POINTER(C.c_float), C.POINTER(C.c_float), _P, _P]),
    "atdf_traverse_count": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, _P, _P, _P]),
    "atdf_scan_workspace_bytes": (_SZ, [_I64]),
    "atdf_exclusive_scan_i32": (_I32, [_P, _I64, _P, _P, _SZ, _P]),
    "atdf_traverse_write": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, _P, C.POINTER(Intervals),
                                   C.POINTER(Samples), _P]),
    "atdf_traverse_cache_bytes": (_SZ, [_I64, _I32]),
    "atdf_traverse_count_cached": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, _P, _P, _P, _I32, _P]),
    "atdf_traverse_write_cached": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, _P,
                                          C.POINTER(Intervals), C.POINTER(Samples), _P, _I32, _P]),
    "atdf_local_samples": (_I32, [_P, _P, _P, _P, _I64, _P, _P]),
    "atdf_grid_sample3d_fwd": (_I32, [_P, _I64, _I32, _I32, _I32, _I32, _P, _P, _I64, _P, _P]),
    "atdf_grid_sample3d_bwd": (_I32, [_P, _I64, _I32, _I32, _I32, _I32, _P, _P, _I64, _P, _P, _P, _P]),
    "atdf_grid_sample2d_fwd": (_I32, [_P, _I64, _I32, _I32, _I32, _P, _P, _I64, _P, _P]),
    "atdf_grid_sample2d_bwd": (_I32, [_P, _I64, _I32, _I32, _I32, _P, _P, _I64, _P, _P, _P, _P]),
    "atdf_grid_sample1d_fwd": (_I32, [_P, _I64, _I32, _I32, _P, _P, _I64, _P, _P]),
    "atdf_grid_sample1d_bwd": (_I32, [_P, _I64, _I32, _I32, _P, _P, _I64, _P, _P, _P, _P]),
    "atdf_gradient_correction_bwd": (_I32, [_P, _P, _I64, _P, _P, _P]),
    "atdf_compute_bound_weight": (_I32, [_P, _I64, _P, _P, _P]),
    "atdf_integrate_xray": (_I32, [_P, _P, _P, _I64, _I32, _I64, _P, _P]),
    "atdf_accumulate_loss_per_node": (_I32, [_P, _I32, _I64, _P, _P, _P, _I64, _I64, _P, _P, _P]),
    "atdf_project_fwd": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, C.POINTER(Intervals),
                                C.POINTER(Field), _P, _P]),
    "atdf_project_bwd_workspace_bytes": (_SZ, [_I64, _I64, C.POINTER(Field)]),
    "atdf_project_bwd_staging_bytes": (_SZ, [_I64, _I64, C.POINTER(Field)]),
    "atdf_project_bwd": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, C.POINTER(Intervals),
                                C.POINTER(Field), _P, _P, _P, _P, _SZ, _P]),
    "atdf_project_bwd_noalloc_max_bytes": (_SZ, []),
    "atdf_project_bwd_noalloc": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, C.POINTER(Intervals),
                                        C.POINTER(Field), _P, _P, _P, _P]),
    "atdf_node_density_stats": (_I32, [_P, _I64, _P, _I64, C.POINTER(Field), _I32, _I32, _I32, _U64, _P, _P, _P,
                                        _P]),
    "atdf_mse_workspace_bytes": (_SZ, [_I64]),
    "atdf_mse_loss": (_I32, [_P, _P, _I64, _F, _P, _P, _P, _SZ, _P]),
    "atdf_adam_step": (_I32, [_P, _P, _P, _P, _I64, _F, _F, _F, _F, _I64, _I32, _P]),
    "atdf_adam_step_bf16": (_I32, [_P, _P, _P, _P, _P, _I64, _F, _F, _F, _F, _I64, _I32, _P]),
    "atdf_project_bwd_adam": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, C.POINTER(Intervals),
                                     C.POINTER(Field), _P, _P, _P, C.POINTER(GridAdam), _P, _SZ, _P]),
    "atdf_project_prepare": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, C.POINTER(Intervals),
                                    C.POINTER(Field), _P, _SZ, _P]),
    "atdf_project_fwd_prepared": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, C.POINTER(Intervals),
                                         C.POINTER(Field), _P, _P, _SZ, _P]),
    "atdf_project_bwd_prepared": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, C.POINTER(Intervals),
                                         C.POINTER(Field), _P, _P, _P, C.POINTER(GridAdam), _P, _SZ, _P]),
    "atdf_project_fwd_bins": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Sampling), _P, C.POINTER(Intervals),
                                     C.POINTER(Field), _P, _P, _P]),
    "atdf_proposal_num_params": (_I64, []),
    "atdf_leaf_table_numel": (_I64, [_I32]),
    "atdf_leaf_table": (_I32, [_P, _I64, _I32, _P, _P]),
    "atdf_proposal_features": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Proposal), _P, _P, _P]),
    "atdf_proposal_sample": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Proposal), _P, _P, _P, _P, _P, _P, _P]),
    "atdf_proposal_allocate": (_I32, [_P, _I64, _P, _P, _I64, C.POINTER(Proposal), _I32, _P, C.POINTER(Intervals),
                                      _P, _P, _P, _P]),
    "atdf_proposal_merge_counts": (_I32, [_I64, _P, _P, _P, _P, _P]),
    "atdf_proposal_write": (_I32, [_I64, _P, _P, _P, _P, _P, _P, C.POINTER(Intervals), _P]),
    "atdf_proposal_grad_workspace_bytes": (_SZ, [_I64]),
    "atdf_proposal_grad": (_I32, [_P, _P, _P, _I64, _F, _P, _P, _P, _SZ, _P]),
    "atdf_sample_volume": (_I32, [_P, _I64, _P, C.POINTER(Field), C.POINTER(Lattice), _P, _P, _P, _P, _P]),
    "atdf_compact_rays": (_I32, [_P, _I64, _P, _P, _P]),
    "atdf_node_id_for_position": (_I32, [_P, _I64, _P, _P, _I64, _P, _P, _P]),
    "atdf_grid_transfer": (_I32, [_P, _I64, _P, _P, _I64, _P, _I64, _P, _P, _I64, _I32, _I32, _I32, _P]),
    "atdf_generate_rays": (_I32, [C.POINTER(CameraRig), _P, _P, _P, _I64, _P, _P, _P]),
    "atdf_log_convert": (_I32, [_P, _P, _I64, C.c_double, C.c_double, C.c_double, _I32, _I32, _P, _P]),
    "atdf_sample_projection": (_I32, [C.POINTER(ProjectionSet), _P, _P, _I64, _P, _P, _P]),
    "atdf_random_ray_batch": (_I32, [C.POINTER(CameraRig), C.POINTER(ProjectionSet), _P, _I64, _P, _I32, _I32,
                                     _I32, _I32, C.c_uint64, _I64, _P, _P, _P, _P, _P, _P, _P, _P]),
    "atdf_point_in_any_image": (_I32, [C.POINTER(CameraRig), _P, _P, _P, _I64, _P, _P]),
    "atdf_node_lattice_points": (_I32, [_P, _I64, _P, _I64, _I32, _P, _P]),
    "atdf_node_density_stats_masked": (_I32, [_P, _I64, _P, _I64, C.POINTER(Field), _I32, _I32, _I32, C.c_uint64,
                                              _P, _P, _P, _P, _P]),
    "atdf_tv_workspace_bytes": (_SZ, [_I64]),
    "atdf_tv_loss": (_I32, [_P, C.POINTER(TVShape), _P, C.c_double, _P, _P, _P, _SZ, _P]),
    "atdf_grid_sample_bwd_workspace_bytes": (_SZ, [_I64, _I64]),
    "atdf_grid_sample_bwd_max_slot_floats": (_I64, []),
    "atdf_grid_sample_bwd_slots": (_I32, [_I32, _P, _I64, _I32, _P, _P, _P, _I64, _P, _P, _P, _P, _SZ, _P]),
    "atdf_fourier_ring_workspace_bytes": (_SZ, [_I32, _I32]),
    "atdf_fourier_ring_means": (_I32, [_P, _I64, _I32, _I32, _P, _P, _SZ, _P]),
    "atdf_decode_features": (_I32, [_P, _I64, C.POINTER(Field), _P, _P]),
    "atdf_decode_features_bwd_workspace_bytes": (_SZ, [_I64, C.POINTER(Field)]),
    "atdf_decode_features_bwd": (_I32, [_P, _I64, C.POINTER(Field), _P, _P, _P, _P, _SZ, _P]),
    "atdf_neighbor_candidates": (_I64, [_I64, C.POINTER(NeighborSpec)]),
    "atdf_neighbor_flags": (_I32, [_P, _I64, _P, _I64, C.POINTER(NeighborSpec), _P, _P]),
    "atdf_neighbor_write": (_I32, [_P, _I64, _P, _I64, C.POINTER(NeighborSpec), _P, _P, _P, _P]),
    "atdf_grid_layout": (_I32, [_P, _P, _I64, _I32, _I32, _I32, _P]),
    "atdf_f32_to_bf16": (_I32, [_P, _P, _I64, _P]),
}

_lib = None


def lib():
    """Load libatdf.so and declare every exported entry point."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m adaptivetensordf_amd.build`")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def stream_ptr(stream=None):
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def ptr(t):
    """Device pointer of a tensor (None stays NULL)."""
    return None if t is None else C.c_void_p(t.data_ptr())


def call(name, *args):
    st = getattr(lib(), name)(*args)
    if st != OK:
        raise AtdfError(st, lib().atdf_last_error().decode())
    return st
