"""Batched clips (asv_batch_*): many clips of uneven sizes running their
Lloyd loops together on one GPU, every JFA pass index one persistent launch
over the (clip, column) work items of the clips sharing its step, each Lloyd
iteration replayed as a HIP graph, launch shapes picked by an autotuner.

Batched workload: n clips from the synthetic generator with video seeds
seed0 .. seed0 + n - 1; per clip, drawn in order from
numpy.random.default_rng(rng_seed): frame size uniform from {160x120,
320x240, 480x360, 640x480}, frames uniform in [20, 120], K uniform in
[200, 4000], ellipses uniform in [8, 32]; |v| <= 2 px/frame, noise sigma 4,
10 Lloyd iterations.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import AsvError, check
from .api import Clip, default_params, synth_video

SIZES = [(160, 120), (320, 240), (480, 360), (640, 480)]


def make_clips(n=64, seed0=100, rng_seed=163, iters=10):
    rng = np.random.default_rng(rng_seed)
    out = []
    for i in range(n):
        W, H = SIZES[int(rng.integers(0, len(SIZES)))]
        out.append(dict(seed=seed0 + i, W=W, H=H, D=int(rng.integers(20, 121)), K=int(rng.integers(200, 4001)),
                        ellipses=int(rng.integers(8, 33)), vmax=2.0, sigma=4.0, iters=iters))
    return out


def clip_params(c, **kw):
    return default_params(c["W"], c["H"], c["D"], c["K"], lloyd_iters=c["iters"], **kw)


def clip_video(c):
    return synth_video(c["seed"], c["W"], c["H"], c["D"], c["ellipses"], c["vmax"], c["sigma"])


class Batch:
    """asv_batch over a list of asv_params."""

    def __init__(self, params, autotune=True):
        self.lib = _lib.load()
        self.params = list(params)
        arr = (_lib.Params * len(self.params))(*self.params)
        h = ctypes.c_void_p()
        rc = self.lib.asv_batch_create(arr, len(self.params), ctypes.byref(h))
        if rc != _lib.ASV_OK:
            raise AsvError(f"{_lib.ERRORS.get(rc, rc)}: {self.lib.asv_batch_last_error().decode()}")
        self.h = h
        if not autotune:
            check(self.lib.asv_batch_set_autotune(h, 0))
        self.clips = []
        for i, p in enumerate(self.params):
            ch = ctypes.c_void_p()
            check(self.lib.asv_batch_clip(h, i, ctypes.byref(ch)))
            self.clips.append(Clip.borrowed(ch, p, self))

    def upload(self, i, rgb, fwd, bwd):
        self.clips[i].upload(rgb, fwd, bwd)

    def run(self):
        rc = self.lib.asv_batch_run(self.h, None)
        if rc != _lib.ASV_OK:
            raise AsvError(f"{_lib.ERRORS.get(rc, rc)}: {self.lib.asv_batch_last_error().decode()}")

    def labels(self, i):
        return self.clips[i].labels()

    def sums(self, i):
        return self.clips[i].sums()

    def tuning(self):
        """Launch groups per iteration: (pass, K, clips or -1, grid, items per
        block), and the autotuner's GPU time in ms."""
        cap = 256
        shapes = np.zeros((cap, 5), np.int32)
        t = ctypes.c_float(0.0)
        n = self.lib.asv_batch_tuning(self.h, shapes.ctypes.data, cap, ctypes.byref(t))
        return [tuple(int(v) for v in r) for r in shapes[:min(n, cap)]], t.value

    def close(self):
        if getattr(self, "h", None):
            for c in self.clips:
                c.close()
            self.lib.asv_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
