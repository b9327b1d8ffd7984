"""ctypes binding of libasv.so, the C ABI declared in include/asv.h.

The library is built in-tree (``anisotropic_supervoxel_amd/libasv.so``) by
``__graft_entry__.build()``.  Loading fails loudly when it is missing: there
is no CPU fallback for any entry point.

A process that also uses torch's GPU support imports torch first: torch's
bundled HIP runtime and /opt/rocm's share a soname, the first one loaded
serves the process, and torch only initialises on its own.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libasv.so")

ASV_OK = 0
ERRORS = {-1: "ASV_ERR_ARG", -2: "ASV_ERR_HIP", -3: "ASV_ERR_STATE", -4: "ASV_ERR_NOMEM", -5: "ASV_ERR_ADMIT"}

# Every symbol include/asv.h declares, with (restype, argtypes).
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
SIGNATURES = {
    "asv_default_params": (None, [_vp, _i32, _i32, _i32, _i32]),
    "asv_create": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "asv_destroy": (ctypes.c_int, [_vp]),
    "asv_seeds_placed": (ctypes.c_int, [_vp]),
    "asv_grid_step": (ctypes.c_int, [_vp]),
    "asv_device_bytes": (_i64, [_vp]),
    "asv_footprint": (_i64, [_vp]),
    "asv_schedule": (ctypes.c_int, [_vp, _vp, _vp, _i32]),
    "asv_plan": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _i32]),
    "asv_grid_seeds": (ctypes.c_int, [_vp, _vp, _i32]),
    "asv_upload_frames": (ctypes.c_int, [_vp, _vp, _vp]),
    "asv_upload_flow": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "asv_prepare": (ctypes.c_int, [_vp, _vp]),
    "asv_seed_stage": (ctypes.c_int, [_vp, _i32, _vp]),
    "asv_flood": (ctypes.c_int, [_vp, _i32, _vp]),
    "asv_flood_passes": (ctypes.c_int, [_vp, _i32, _i32, _vp]),
    "asv_centroids": (ctypes.c_int, [_vp, _i32, _vp]),
    "asv_run": (ctypes.c_int, [_vp, _vp]),
    "asv_run_loop": (ctypes.c_int, [_vp, _vp]),
    "asv_download_labels": (ctypes.c_int, [_vp, _vp, _vp]),
    "asv_download_sums": (ctypes.c_int, [_vp, _vp, _vp]),
    "asv_enable_trace": (ctypes.c_int, [_vp, _i32]),
    "asv_force_exact_clamp": (ctypes.c_int, [_vp, _i32]),
    "asv_enable_march_stats": (ctypes.c_int, [_vp, _i32]),
    "asv_download_march_stats": (ctypes.c_int, [_vp, _vp, _vp]),
    "asv_debug_fail_alloc": (ctypes.c_int, [_i64]),
    "asv_download_trace": (ctypes.c_int, [_vp, _vp, _vp]),
    "asv_download_lab": (ctypes.c_int, [_vp, _vp, _vp]),
    "asv_download_metric": (ctypes.c_int, [_vp, _vp, _vp]),
    "asv_run_timed": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "asv_synth_video": (ctypes.c_int, [ctypes.c_uint64, _i32, _i32, _i32, _i32, _f32, _f32, _vp, _vp, _vp, _vp]),
    "asv_postprocess": (ctypes.c_int, [_vp, _i32, _vp, _i64, _vp, _vp]),
    "asv_write_label_txt": (ctypes.c_int, [_vp, _i32, _i32, _i32, ctypes.c_char_p]),
    "asv_write_render_ppm": (ctypes.c_int, [_vp, _i32, _i32, _i32, _i32, ctypes.c_char_p]),
    "asv_eval_metrics": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _vp, _vp]),
    "asv_plan_seeds": (ctypes.c_int, [_vp, _vp, _i32]),
    "asv_connectivity": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "asv_initialization": (ctypes.c_int, [_vp, _i32]),
    "asv_connectivity_labels": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _i32, _vp, _vp]),
    "asv_download_labels_mapped": (ctypes.c_int, [_vp, _vp, _i32, _i32, _vp, _i64, _vp, _vp]),
    "asv_stream_block_into": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "asv_stream_checksum": (ctypes.c_int, [_vp, _vp]),
    "asv_postprocess_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "asv_postprocess_labels": (ctypes.c_int, [_vp, _i32, _vp, _vp, _i64, _vp, _vp, _vp]),
    "asv_postprocess_labels_noalloc": (ctypes.c_int, [_vp, _i32, _vp, _vp, _i64, _vp, _vp, _vp]),
    "asv_host_alloc": (ctypes.c_int, [_i64, ctypes.POINTER(_vp)]),
    "asv_host_free": (ctypes.c_int, [_vp]),
    "asv_reseed": (ctypes.c_int, [_vp, _vp, _i32, _i32, _i32, _vp]),
    "asv_shift_frames": (ctypes.c_int, [_vp, _i32, _i32, _vp]),
    "asv_upload_frames_range": (ctypes.c_int, [_vp, _i32, _i32, _vp, _vp, _vp, _vp]),
    "asv_create_seeded": (ctypes.c_int, [_vp, _vp, _i32, _i32, ctypes.POINTER(_vp)]),
    "asv_stream_default_params": (None, [_vp, _i32, _i32, _i32, _i32]),
    "asv_stream_plan": (ctypes.c_int, [_vp, _vp, _i32]),
    "asv_stream_create": (ctypes.c_int, [_vp, ctypes.POINTER(_vp)]),
    "asv_stream_destroy": (ctypes.c_int, [_vp]),
    "asv_stream_block_count": (ctypes.c_int, [_vp]),
    "asv_stream_seeds_placed": (ctypes.c_int, [_vp]),
    "asv_stream_block_info": (ctypes.c_int, [_vp, _vp]),
    "asv_stream_block": (ctypes.c_int, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "asv_stream_run_video": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "asv_stream_last_error": (ctypes.c_char_p, []),
    "asv_serve_default_params": (None, [_vp]),
    "asv_serve_footprint": (_i64, [_vp, _vp]),
    "asv_serve_run": (ctypes.c_int, [_vp, _vp, _i32, _vp, _vp, _i32, _vp, _vp]),
    "asv_synth_video_range": (ctypes.c_int, [ctypes.c_uint64, _i32, _i32, _i32, _i32, _f32, _f32, _i32, _i32,
                                             _vp, _vp, _vp, _vp]),
    "asv_batch_create": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_vp)]),
    "asv_batch_destroy": (ctypes.c_int, [_vp]),
    "asv_batch_size": (ctypes.c_int, [_vp]),
    "asv_batch_clip": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(_vp)]),
    "asv_batch_run": (ctypes.c_int, [_vp, _vp]),
    "asv_batch_set_autotune": (ctypes.c_int, [_vp, _i32]),
    "asv_batch_launch_count": (ctypes.c_int, [_vp, _vp]),
    "asv_batch_tuning": (ctypes.c_int, [_vp, _vp, _i32, _vp]),
    "asv_batch_last_error": (ctypes.c_char_p, []),
    "asv_last_error": (ctypes.c_char_p, []),
    "asv_device_count": (ctypes.c_int, []),
}


class Params(ctypes.Structure):
    """Mirror of struct asv_params."""

    _fields_ = [
        ("width", _i32), ("height", _i32), ("depth", _i32),
        ("seeds_num", _i32), ("lloyd_iters", _i32), ("compactness", _f32),
        ("tie_rule", _i32), ("angle_gate", _i32), ("schedule", _i32),
        ("frames_prepared", _i32), ("seed_init", _i32),
    ]


class StreamParams(ctypes.Structure):
    """Mirror of struct asv_stream_params."""

    _fields_ = [
        ("width", _i32), ("height", _i32), ("total_frames", _i32), ("seeds_num", _i32),
        ("seeds_per_100_frames", _i32), ("block_size", _i32), ("overlap", _i32),
        ("lloyd_iters", _i32), ("compactness", _f32), ("angle_gate", _i32), ("tie_rule", _i32),
        ("frames_prepared", _i32),
    ]


class ServeParams(ctypes.Structure):
    _fields_ = [("memory_budget", _i64), ("max_active", _i32), ("block_size", _i32), ("overlap", _i32),
                ("lloyd_iters", _i32), ("compactness", _f32), ("angle_gate", _i32), ("honor_arrivals", _i32),
                ("batch_windows", _i32)]


class ServeRequest(ctypes.Structure):
    _fields_ = [("stream_id", _i32), ("width", _i32), ("height", _i32), ("frames", _i32), ("density", _f32),
                ("n_ellipses", _i32), ("video_seed", ctypes.c_uint64), ("vmax", _f32), ("noise_sigma", _f32),
                ("arrival_ms", ctypes.c_double), ("rgb", _vp), ("fwd", _vp), ("bwd", _vp), ("cancel", _vp),
                ("cancel_after_windows", _i32)]


class ServeWindow(ctypes.Structure):
    _fields_ = [("stream_id", _i32), ("window", _i32), ("frames", _i32), ("seeds", _i32),
                ("ready_ms", ctypes.c_double), ("start_ms", ctypes.c_double), ("end_ms", ctypes.c_double)]


class ServeResult(ctypes.Structure):
    _fields_ = [("stream_id", _i32), ("status", _i32), ("labels_hash", ctypes.c_uint64), ("footprint", _i64),
                ("voxel_iters", _i64), ("admitted_ms", ctypes.c_double), ("done_ms", ctypes.c_double)]


_lib = None


def load():
    """Load libasv.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class AsvError(RuntimeError):
    pass


def check(rc):
    if rc != ASV_OK:
        msg = load().asv_last_error().decode(errors="replace")
        raise AsvError(f"{ERRORS.get(rc, rc)}: {msg}")
    return rc


def ptr(a):
    """Data pointer of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)
