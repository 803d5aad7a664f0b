"""ctypes bindings of liblwamd.so (include/livox_warp_amd.h).

The library is built in-tree (`make -C livox-warp_amd`, or `__graft_entry__.build()`); loading fails loudly
when it is missing -- there is no other implementation of these kernels in the product.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblwamd.so")

LW_EINVAL = -1000
LW_ENOTBOUND = -1001
LW_ELOOP = -1002
LW_ENOLOOP = -1003
LW_ERANGE = -1004
LW_PLANE_NOALLOC_MAX = 256  # poses lw_plane_icp_noalloc takes at once

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


class RegParams(ctypes.Structure):
    _fields_ = [("damping", _f64), ("degen_thr", _f64), ("prior_info", _f64), ("reg_voxel", _f32),
                ("min_range", _f32), ("max_range", _f32), ("noise", _f32), ("alpha", _f32), ("ret_mask", _i32),
                ("max_iter", _i32), ("min_corr", _i32), ("wide_iters", _i32), ("acc_mode", _i32)]


class RegResult(ctypes.Structure):
    _fields_ = [("T", _f64 * 16), ("rms", _f64), ("cond", _f64), ("ok", _i32), ("iters", _i32), ("corr", _i32),
                ("n_ds", _i32), ("degen", _i32), ("weak", _i32), ("rms0", _f64), ("m", _i32), ("pad", _i32)]


class FrameIn(ctypes.Structure):
    _fields_ = [("T_pred", _f64 * 16), ("T_prev_inv", _f64 * 16), ("T_key_inv", _f64 * 16), ("kf_dist", _f64),
                ("kf_angle", _f64), ("redeskew_rot", _f64), ("redeskew_trans", _f64), ("xi", _f32 * 6),
                ("t_mid", _f32), ("inv_dt", _f32), ("min_points", _i32), ("has_prev", _i32), ("has_key", _i32),
                ("redeskew", _i32)]


class FrameOut(ctypes.Structure):
    _fields_ = [("reg", RegResult * 3), ("T", _f64 * 16), ("d", _f64 * 6), ("xi", _f32 * 6), ("m", _i32),
                ("skipped", _i32), ("ok", _i32), ("registrations", _i32), ("keyframe", _i32), ("map_voxels", _i32),
                ("map_overflow", _i32), ("guess", _i32), ("fallback", _i32), ("pad", _i32)]


class FramePrior(ctypes.Structure):
    _fields_ = [("y", _vp), ("T", _f64 * 16), ("hist0", _f64 * 6), ("sd_t", _f64 * 6), ("mu_t", _f64 * 6),
                ("rms_gate", _f64)]


class BatchSrc(ctypes.Structure):
    # device addresses as 64-bit integers (the same layout as the C pointers), so numpy can view arrays of these
    _fields_ = [("xyz", ctypes.c_uint64), ("attr", ctypes.c_uint64), ("t", ctypes.c_uint64), ("m", _i32),
                ("warm", _i32)]


class LocParams(ctypes.Structure):
    _fields_ = [("damping", _f64), ("reg_voxel", _f32), ("noise", _f32), ("alpha", _f32), ("iters", _i32),
                ("min_count", _i32), ("block", _i32), ("grid", _i32), ("use_graph", _i32), ("chunk", _i32)]


class MapView(ctypes.Structure):
    _fields_ = [("min_range", _f32), ("max_range", _f32), ("min_refl", _i32), ("noise_mask", _i32),
                ("ret_mask", _i32)]


# name -> (restype, argtypes); every symbol include/livox_warp_amd.h declares
SIGNATURES = {
    "lw_version": (_i32, []),
    "lw_create": (_i32, [ctypes.POINTER(_vp), _i64, _i64, _vp]),
    "lw_destroy": (_i32, [_vp]),
    "lw_set_stream": (_i32, [_vp, _vp]),
    "lw_synchronize": (_i32, [_vp]),
    "lw_bind_frame": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "lw_map_reset": (_i32, [_vp]),
    "lw_frame_reset": (_i32, [_vp]),
    "lw_frame_append": (_i32, [_vp, _vp, _vp, _vp, _i64, _f32, _i32, _i32]),
    "lw_frame_count": (_i32, [_vp, _i32p]),
    "lw_deskew": (_i32, [_vp, _i32, _f32, _f32, _f32p]),
    "lw_thin": (_i32, [_vp, _i32, _f32, _f32, _f32, _i32, _i32p]),
    "lw_thin_read": (_i32, [_vp, _f32p, _i32]),
    "lw_thin_load": (_i32, [_vp, _f32p, _i32]),
    "lw_map_insert": (_i32, [_vp, _vp, _vp, _i64, _f32p, _f32, _f32, _f32, _i32]),
    "lw_map_stats": (_i32, [_vp, _i32p]),
    "lw_map_read": (_i32, [_vp, _i64p, _f32p, _f32p, _i32p]),
    "lw_map_find": (_i32, [_vp, _i64p, _i32, _i32p]),
    "lw_map_read_voxels": (_i32, [_vp, _i32, _i64p, _f32p, _f32p, _i32p, _i32p]),
    "lw_batch_create": (_i32, [_vp, _vp, _i32, _vp]),
    "lw_batch_destroy": (_i32, [_vp]),
    "lw_batch_frame": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "lw_batch_prof_enable": (_i32, [_vp, _i32]),
    "lw_batch_prof_reset": (_i32, [_vp]),
    "lw_batch_prof_read": (_i32, [_vp, _f64p]),
    "lw_icp_accum": (_i32, [_vp, _i32, _f32p, _f32, _f32, _i32, _f32, _f32, _f32, _f32, _f32, _i32, _f64p, _i32p]),
    "lw_register": (_i32, [_vp, _i32, ctypes.POINTER(RegParams), _f64p, ctypes.POINTER(RegResult)]),
    "lw_register_thinned": (_i32, [_vp, _i32, ctypes.POINTER(RegParams), _f64p, ctypes.POINTER(RegResult)]),
    "lw_set_persistent": (_i32, [_vp, _i32]),
    "lw_odom_frame": (_i32, [_vp, ctypes.POINTER(RegParams), ctypes.POINTER(FrameIn), ctypes.POINTER(FrameOut)]),
    "lw_odom_frame_prior": (_i32, [_vp, ctypes.POINTER(RegParams), ctypes.POINTER(FrameIn),
                                   ctypes.POINTER(FramePrior), ctypes.POINTER(FrameOut)]),
    "lw_loc_batch": (_i32, [_vp, _vp, _vp, _vp, _i32, _vp, _i32, ctypes.POINTER(LocParams), _vp, _vp, _vp, _vp]),
    "lw_loc_tile_info": (_i32, [_i32, _i32, _i32p]),
    "lw_vox_create": (_i32, [ctypes.POINTER(_vp), _i32, _vp]),
    "lw_vox_destroy": (_i32, [_vp]),
    "lw_vox_clear": (_i32, [_vp]),
    "lw_vox_insert": (_i32, [_vp, _vp, _vp, _vp, _i64, _f64]),
    "lw_vox_stats": (_i32, [_vp, _i32p]),
    "lw_vox_emit": (_i32, [_vp, _f64, _vp, _vp, _vp, _i64, ctypes.POINTER(_i64)]),
    "lw_vox_read": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "lw_mlp_param_count": (_i32, []),
    "lw_mlp_create": (_i32, [ctypes.POINTER(_vp), _i32, _vp]),
    "lw_mlp_destroy": (_i32, [_vp]),
    "lw_mlp_set_stream": (_i32, [_vp, _vp]),
    "lw_mlp_set_params": (_i32, [_vp, _f32p]),
    "lw_mlp_get_params": (_i32, [_vp, _f32p]),
    "lw_mlp_get_grads": (_i32, [_vp, _f32p]),
    "lw_mlp_forward": (_i32, [_vp, _vp, _i64, _i32, _vp]),
    "lw_mlp_loss": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "lw_mlp_backward": (_i32, [_vp, _i32]),
    "lw_mlp_adam": (_i32, [_vp, _f32, _f32, _f32, _f32, _i32]),
    "lw_mlp_train_step": (_i32, [_vp, _vp, _i64, _vp, _i32, _f32, _f32, _f32, _f32, _i32, _vp, _i32]),
    "lw_frame_features": (_i32, [_vp, _i32, _f32, _f32, _f32p, _vp]),
    "lw_prof_enable": (_i32, [_vp, _i32]),
    "lw_prof_reset": (_i32, [_vp]),
    "lw_prof_read": (_i32, [_vp, _f64p]),
    "lw_prof_phases": (_i32, [_vp, _f64p]),
    "lw_host_gn_step": (_i32, [_f64p, _f64p, _f64p, _f64, _f64, _f64, _i32, _f64p, _f64p, _i32p]),
    "lw_host_voxel_precision": (_i32, [_f32p, _f32p, _i32, _f32, _f32, _f32, _f32, _f32p, _f32p, _f32p]),
    "lw_gn_step_check": (_i32, [_f64p, _f64p, _f64p, _f64, _f64, _f64, _i32, _f64p]),
    "lw_host_apply_twist": (_i32, [_f64p, _f64p]),
    "lw_host_twist_between": (_i32, [_f64p, _f64p, _f64p]),
    "lw_host_cond": (_f64, [_f64p]),
    "lw_imap_create": (_i32, [ctypes.POINTER(_vp), _i64, _vp]),
    "lw_imap_destroy": (_i32, [_vp]),
    "lw_imap_clear": (_i32, [_vp]),
    "lw_imap_set_poses": (_i32, [_vp, _i32p, _f32p, _i32]),
    "lw_imap_insert": (_i32, [_vp, _vp, _vp, _vp, _vp, _i32, ctypes.POINTER(MapView), _f32]),
    "lw_imap_carve": (_i32, [_vp, _vp, _vp, _vp, _i32, _f32, _i32, _f32, _i32]),
    "lw_imap_rebase_clock": (_i32, [_vp, _f32]),
    "lw_imap_stats": (_i32, [_vp, _i32p]),
    "lw_imap_read": (_i32, [_vp, _vp, _i64]),
    "lw_imap_load": (_i32, [_vp, _vp, _i64]),
    "lw_lvxr_index": (_i32, [_vp, _i64, _i64, _vp, _vp, _i64, _i64p]),
    "lw_lvxr_clock": (_i32, [_vp, _vp, _i64, _i64, _vp]),
    "lw_lvxr_decode": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _i32, _i32, _i64, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lw_grid_work_bytes": (_i64, [_i32p]),
    "lw_grid_build": (_i32, [_vp, _i32, _f32p, _i32p, _i32p, _i32, _vp, _vp, _i64, _vp, _vp]),
    "lw_plane_work_bytes": (_i64, [_i32]),
    "lw_plane_icp_noalloc": (_i32, [_vp, _i32, _vp, _i32, _f32p, _i32p, _vp, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _vp,
                                    _vp]),
    "lw_plane_icp": (_i32, [_vp, _i32, _vp, _i32, _f32p, _i32p, _vp, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _vp, _i64, _vp,
                             _vp]),
    "lw_plane_eval": (_i32, [_vp, _i32, _vp, _i32, _f32p, _i32p, _vp, _vp, _vp, _f32p, _vp, _vp]),
    "lw_grid_distances": (_i32, [_vp, _i32, _f32p, _f32p, _i32p, _vp, _vp, _vp, _vp, _vp]),
    "lw_score_grid": (_i32, [_vp, _i32, _vp, _i32, _i32, _f32, _f32, _f32p, _i32p, _vp, _f32, _f32, _vp, _vp, _vp]),
    "lw_score_list": (_i32, [_vp, _i32, _vp, _i32, _f32p, _i32p, _vp, _f32, _vp, _vp]),
}

_lib = None


class LwError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LwError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        if rc == LW_EINVAL:
            raise LwError(f"{what}: invalid argument")
        if rc == LW_ENOTBOUND:
            raise LwError(f"{what}: frame buffers not bound")
        if rc == LW_ELOOP:
            raise LwError(f"{what}: the persistent registration loop stopped waiting")
        if rc == LW_ENOLOOP:
            raise LwError(f"{what}: no cooperative launch on this device")
        if rc == LW_ERANGE:
            raise LwError(f"{what}: size beyond the fixed scratch of an allocation-free entry point")
        raise LwError(f"{what}: HIP error {-rc}")


def ptr(a, t):
    """ctypes pointer of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)
