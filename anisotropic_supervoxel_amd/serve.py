"""Served streaming form: many concurrent video streams as sliding windows
with seed carry, continuous batching on one GPU (asv_serve_run).

Workload of the served benchmark: Poisson arrivals at 40 streams/s; each
stream is a synthetic video (seed = stream id) with resolution uniform from
{320x240, 640x480, 1280x720}, 60-300 frames and supervoxel density uniform
in [1, 4] per 1000 pixels per frame; windows of 10 frames, stride 5,
5 Lloyd iterations per window.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import ServeParams, ServeRequest, ServeResult, ServeWindow

RESOLUTIONS = [(320, 240), (640, 480), (1280, 720)]


def default_params(**kw):
    p = ServeParams()
    _lib.load().asv_serve_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown serve parameter {k}")
        setattr(p, k, v)
    return p


def make_requests(n, rate_per_s=40.0, seed=0, resolutions=None, frames=(60, 300), density=(1.0, 4.0),
                  ellipses=(8, 32)):
    """n stream requests with Poisson arrival times (ms)."""
    rng = np.random.default_rng(seed)
    res = resolutions or RESOLUTIONS
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1000.0 / rate_per_s)
        W, H = res[int(rng.integers(0, len(res)))]
        out.append(dict(stream_id=i, width=W, height=H, frames=int(rng.integers(frames[0], frames[1] + 1)),
                        density=float(rng.uniform(*density)), n_ellipses=int(rng.integers(ellipses[0], ellipses[1] + 1)),
                        video_seed=i, vmax=2.0, noise_sigma=4.0, arrival_ms=t))
    return out


_FIELDS = [f[0] for f in ServeRequest._fields_ if f[0] not in ("rgb", "fwd", "bwd", "cancel",
                                                                  "cancel_after_windows")]


def _request(r):
    """A ServeRequest from a request dict; optional keys: frames_data (rgb,
    fwd, bwd arrays), cancel (a one-element int32 array the caller may set
    while the loop runs), cancel_after_windows."""
    q = ServeRequest(**{k: r[k] for k in _FIELDS})
    frames = r.get("frames_data")
    if frames is not None:
        rgb, fwd, bwd = frames
        q.rgb, q.fwd, q.bwd = rgb.ctypes.data, fwd.ctypes.data, bwd.ctypes.data
    flag = r.get("cancel")
    if flag is not None:
        assert flag.dtype == np.int32 and flag.size >= 1
        q.cancel = flag.ctypes.data
    q.cancel_after_windows = int(r.get("cancel_after_windows", 0))
    return q


class PinnedArray:
    """A numpy view of page-locked host memory from asv_host_alloc."""

    def __init__(self, shape, dtype):
        self.lib = _lib.load()
        n = int(np.prod(shape)) * np.dtype(dtype).itemsize
        p = ctypes.c_void_p()
        if self.lib.asv_host_alloc(n, ctypes.byref(p)) != 0:
            raise _lib.AsvError(self.lib.asv_last_error().decode())
        self.ptr = p
        buf = (ctypes.c_char * n).from_address(p.value)
        self.array = np.frombuffer(buf, dtype=dtype).reshape(shape)

    def __del__(self):
        try:
            self.lib.asv_host_free(self.ptr)
        except Exception:
            pass


def attach_frames(requests, pinned=True):
    """Generates every stream's frames and flows up front into host memory
    (page-locked by default), so the serving loop reads them as arrived
    input instead of synthesising them."""
    lib = _lib.load()
    for r in requests:
        D, H, W = r["frames"], r["height"], r["width"]
        if pinned:
            bufs = [PinnedArray((D, H, W, 3), np.uint8), PinnedArray((D, H, W, 2), np.float32),
                    PinnedArray((D, H, W, 2), np.float32)]
            rgb, fwd, bwd = (b.array for b in bufs)
            r["_pinned"] = bufs
        else:
            rgb = np.empty((D, H, W, 3), np.uint8)
            fwd = np.empty((D, H, W, 2), np.float32)
            bwd = np.empty((D, H, W, 2), np.float32)
        rc = lib.asv_synth_video(r["video_seed"], W, H, D, r["n_ellipses"], r["vmax"], r["noise_sigma"],
                                 rgb.ctypes.data, fwd.ctypes.data, bwd.ctypes.data, None)
        if rc != 0:
            raise _lib.AsvError("synthetic video failed")
        r["frames_data"] = (rgb, fwd, bwd)
    return requests


def footprint(params, req):
    r = _request(req)
    return _lib.load().asv_serve_footprint(ctypes.byref(params), ctypes.byref(r))


def run(params, requests):
    """Runs the serving loop.  Returns (results list of dicts, windows as a
    structured numpy array, wall ms)."""
    lib = _lib.load()
    n = len(requests)
    reqs = (ServeRequest * max(n, 1))()
    for i, r in enumerate(requests):
        reqs[i] = _request(r)
    res = (ServeResult * max(n, 1))()
    cap = sum(max(1, r["frames"]) for r in requests) + 16
    wins = (ServeWindow * cap)()
    nw = ctypes.c_int32(0)
    wall = ctypes.c_double(0.0)
    rc = lib.asv_serve_run(ctypes.byref(params), reqs, n, res, wins, cap, ctypes.byref(nw), ctypes.byref(wall))
    if rc != 0:
        raise _lib.AsvError(f"asv_serve_run failed ({rc}): {lib.asv_last_error().decode()} "
                            f"{lib.asv_stream_last_error().decode()}")
    results = [dict(stream_id=res[i].stream_id, status=res[i].status, labels_hash=res[i].labels_hash,
                    footprint=res[i].footprint, voxel_iters=res[i].voxel_iters, admitted_ms=res[i].admitted_ms,
                    done_ms=res[i].done_ms) for i in range(n)]
    dt = np.dtype([("stream_id", "i4"), ("window", "i4"), ("frames", "i4"), ("seeds", "i4"),
                   ("ready_ms", "f8"), ("start_ms", "f8"), ("end_ms", "f8")])
    w = np.frombuffer(bytes(wins), dtype=dt, count=cap)[: min(nw.value, cap)].copy()
    return results, w, wall.value


def summarize(results, windows, wall_ms):
    lat = windows["end_ms"] - windows["ready_ms"]
    vox = sum(r["voxel_iters"] for r in results if r["status"] == 0)
    return {
        "streams": len(results), "windows": int(len(windows)),
        "p50_window_ms": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_window_ms": float(np.percentile(lat, 99)) if len(lat) else None,
        "wall_ms": wall_ms, "voxel_iters": int(vox),
        "mvoxel_iters_per_s": vox / (wall_ms * 1e-3) / 1e6 if wall_ms > 0 else None,
    }


def labels_checksum(labels):
    """The checksum asv_serve_run reports per stream (asv_stream_checksum):
    sum over voxels i of splitmix64((i << 32) | uint32(label_i)) mod 2^64,
    with i the voxel's index in the whole video."""
    lab = np.ascontiguousarray(labels, np.int32).reshape(-1)
    x = (np.arange(lab.size, dtype=np.uint64) << np.uint64(32)) | lab.view(np.uint32).astype(np.uint64)
    with np.errstate(over="ignore"):
        x = x + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        x = x ^ (x >> np.uint64(31))
    return int(x.sum(dtype=np.uint64))
