"""Python host over the C ABI: one `Clip` per video clip.

The stage methods mirror the reference's Image interface
(GPU-AS/Image.h:71-88) in the order Render::run calls them
(GPU-AS/Render.cpp:254-319):

    init_gpu           -> Clip(params)                  (asv_create)
    RGB2LAB            -> Clip.prepare()                (asv_prepare)
    Ld_first_init /
    Ld_next_seeds_matrix -> Clip.seed_stage(it)         (asv_seed_stage;
                          matrix_slover evaluated at the seeds and rings)
    flood              -> Clip.flood(it)                (asv_flood)
    centroids_next_seeds -> Clip.centroids(it)          (asv_centroids)
    label readback     -> Clip.labels()                 (asv_download_labels)

Errors surface as AsvError carrying the library's message, the way the
reference prints and returns on failure.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import AsvError, Params, check, ptr  # noqa: F401


def default_params(width, height, depth, seeds_num, **kw):
    """asv_default_params, then keyword overrides (lloyd_iters, compactness,
    tie_rule, angle_gate, schedule, frames_prepared, seed_init)."""
    lib = _lib.load()
    p = Params()
    lib.asv_default_params(ctypes.byref(p), width, height, depth, seeds_num)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p


def plan(params):
    """Host-only plan for params: (A seeds placed, grid step s, steps, mult)."""
    A = ctypes.c_int32(0)
    s = ctypes.c_int32(0)
    steps = np.zeros(64, np.int32)
    mult = np.zeros(64, np.float32)
    n = _lib.load().asv_plan(ctypes.byref(params), ctypes.byref(A), ctypes.byref(s), ptr(steps), ptr(mult), 64)
    if n < 0:
        check(n)
    return A.value, s.value, steps[:n].tolist(), mult[:n].tolist()


def seeds_placed_for(width, height, depth, seeds_num):
    """Grid seeds placed for these sizes (host-side rule, no GPU needed)."""
    return plan(default_params(width, height, depth, seeds_num))[0]


def synth_video(seed, width, height, depth, n_ellipses, vmax=2.0, noise_sigma=4.0):
    """The benchmark's synthetic video: returns (rgb u8 [D,H,W,3],
    fwd f32 [D,H,W,2], bwd f32 [D,H,W,2], region i32 [D,H,W]), image rows."""
    lib = _lib.load()
    rgb = np.empty((depth, height, width, 3), np.uint8)
    fwd = np.empty((depth, height, width, 2), np.float32)
    bwd = np.empty((depth, height, width, 2), np.float32)
    reg = np.empty((depth, height, width), np.int32)
    check(lib.asv_synth_video(seed, width, height, depth, n_ellipses, vmax, noise_sigma,
                              ptr(rgb), ptr(fwd), ptr(bwd), ptr(reg)))
    return rgb, fwd, bwd, reg


def postprocess(sums, A, labels):
    """Image::check_dead_seeds on host arrays (asv_postprocess).  Returns
    (labels after post-processing, dead seed count, final label count)."""
    sums8 = np.zeros((A + 1, 8), np.int64)
    sums8[:, :sums.shape[1]] = sums[:, :8]
    out = np.ascontiguousarray(labels, np.int32).copy()
    dead = ctypes.c_int32(0)
    nfin = ctypes.c_int32(0)
    check(_lib.load().asv_postprocess(ptr(sums8), A, ptr(out), out.size, ctypes.byref(dead), ctypes.byref(nfin)))
    return out, dead.value, nfin.value


def write_label_txt(labels, path):
    D, H, W = labels.shape
    lab = np.ascontiguousarray(labels, np.int32)
    check(_lib.load().asv_write_label_txt(ptr(lab), W, H, D, str(path).encode()))


def write_render_ppm(labels, n_labels, prefix):
    D, H, W = labels.shape
    lab = np.ascontiguousarray(labels, np.int32)
    check(_lib.load().asv_write_render_ppm(ptr(lab), W, H, D, n_labels, str(prefix).encode()))


def eval_metrics(labels, regions):
    """(3D undersegmentation error, 3D boundary recall) of labels against
    ground-truth regions of the same shape and row order."""
    D, H, W = labels.shape
    a = np.ascontiguousarray(labels, np.int32)
    r = np.ascontiguousarray(regions, np.int32)
    ue = ctypes.c_double(0.0)
    br = ctypes.c_double(0.0)
    check(_lib.load().asv_eval_metrics(ptr(a), ptr(r), W, H, D, ctypes.byref(ue), ctypes.byref(br)))
    return ue.value, br.value


def plan_seeds(params):
    """Grid seed voxels (A x 3, GL orientation) asv_create would place."""
    A = plan(params)[0]
    xyz = np.zeros((A, 3), np.int32)
    n = _lib.load().asv_plan_seeds(ctypes.byref(params), ptr(xyz), xyz.size)
    if n < 0:
        check(n)
    return xyz


def stream_params(width, height, total_frames, seeds_num, **kw):
    """asv_stream_default_params, then keyword overrides."""
    p = _lib.StreamParams()
    _lib.load().asv_stream_default_params(ctypes.byref(p), width, height, total_frames, seeds_num)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(f"unknown stream parameter {k}")
        setattr(p, k, v)
    return p


def _check_stream(rc):
    if rc != _lib.ASV_OK:
        lib = _lib.load()
        raise AsvError(f"{_lib.ERRORS.get(rc, rc)}: {lib.asv_stream_last_error().decode()} "
                       f"{lib.asv_last_error().decode()}")


def stream_plan(sp):
    """Block intervals [(start, end)] of the temporal-block run."""
    iv = np.zeros(2 * 4096, np.int32)
    n = _lib.load().asv_stream_plan(ctypes.byref(sp), ptr(iv), 4096)
    if n < 0:
        _check_stream(n)
    return [(int(iv[2 * i]), int(iv[2 * i + 1])) for i in range(n)]


def stream_run_video(sp, rgb, fwd, bwd, stream=None):
    """Whole video through the temporal blocks; returns (labels [D,H,W] global
    ids in GL rows, seeds placed)."""
    D, H, W = sp.total_frames, sp.height, sp.width
    rgb = np.ascontiguousarray(rgb, np.uint8)
    fwd = np.ascontiguousarray(fwd, np.float32)
    bwd = np.ascontiguousarray(bwd, np.float32)
    assert rgb.shape == (D, H, W, 3) and fwd.shape == (D, H, W, 2) and bwd.shape == (D, H, W, 2)
    labels = np.full((D, H, W), -7, np.int32)
    placed = ctypes.c_int32(0)
    _check_stream(_lib.load().asv_stream_run_video(ctypes.byref(sp), ptr(rgb), ptr(fwd), ptr(bwd), ptr(labels),
                                                   ctypes.byref(placed), ctypes.c_void_p(stream) if stream else None))
    return labels, placed.value


class Clip:
    """One clip's device context (asv_ctx).  With seeds_xyz (n x 3) and
    grid_step, the initial seeds are given instead of the grid rule."""

    def __init__(self, params, stream=None, seeds_xyz=None, grid_step=None):
        self.lib = _lib.load()
        self.p = params
        self.stream = ctypes.c_void_p(stream) if stream else None
        h = ctypes.c_void_p()
        if seeds_xyz is None:
            check(self.lib.asv_create(ctypes.byref(params), ctypes.byref(h)))
        else:
            xyz = np.ascontiguousarray(seeds_xyz, np.int32)
            check(self.lib.asv_create_seeded(ctypes.byref(params), ptr(xyz), len(xyz), int(grid_step),
                                             ctypes.byref(h)))
        self.h = h
        self.A = self.lib.asv_seeds_placed(h)
        self.s = self.lib.asv_grid_step(h)
        self.shape = (params.depth, params.height, params.width)

    @classmethod
    def borrowed(cls, handle, params, owner):
        """A view of a context owned by someone else (a batch's clip); close()
        leaves the context alone."""
        self = cls.__new__(cls)
        self.lib = _lib.load()
        self.p = params
        self.stream = None
        self.h = handle
        self._owner = owner
        self.A = self.lib.asv_seeds_placed(handle)
        self.s = self.lib.asv_grid_step(handle)
        self.shape = (params.depth, params.height, params.width)
        return self

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_owner", None) is not None:
            self.h = None
            return
        if getattr(self, "h", None):
            self.lib.asv_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- host-side tables ----------------------------------------------------
    def schedule(self):
        steps = np.zeros(64, np.int32)
        mult = np.zeros(64, np.float32)
        n = self.lib.asv_schedule(self.h, ptr(steps), ptr(mult), 64)
        return steps[:n].tolist(), mult[:n].tolist()

    def grid_seeds(self):
        xyz = np.zeros((self.A, 3), np.int32)
        self.lib.asv_grid_seeds(self.h, ptr(xyz), xyz.size)
        return xyz

    def device_bytes(self):
        return self.lib.asv_device_bytes(self.h)

    # -- inputs --------------------------------------------------------------
    def upload(self, rgb, fwd, bwd):
        D, H, W = self.shape
        rgb = np.ascontiguousarray(rgb, np.uint8)
        fwd = np.ascontiguousarray(fwd, np.float32)
        bwd = np.ascontiguousarray(bwd, np.float32)
        if rgb.shape != (D, H, W, 3) or fwd.shape != (D, H, W, 2) or bwd.shape != (D, H, W, 2):
            raise AsvError("ASV_ERR_ARG: input shapes do not match the clip")
        check(self.lib.asv_upload_frames(self.h, ptr(rgb), self.stream))
        check(self.lib.asv_upload_flow(self.h, ptr(fwd), ptr(bwd), self.stream))
        self._keep = (rgb, fwd, bwd)   # async copies read host memory

    # -- stages --------------------------------------------------------------
    def prepare(self):
        check(self.lib.asv_prepare(self.h, self.stream))

    def seed_stage(self, it):
        check(self.lib.asv_seed_stage(self.h, it, self.stream))

    def flood(self, it):
        check(self.lib.asv_flood(self.h, it, self.stream))

    def flood_passes(self, it, npasses):
        """The first npasses passes of iteration it only (tests and tools)."""
        check(self.lib.asv_flood_passes(self.h, it, npasses, self.stream))

    def centroids(self, it):
        check(self.lib.asv_centroids(self.h, it, self.stream))

    def run(self):
        check(self.lib.asv_run(self.h, self.stream))

    def initialization(self, density=True):
        """Image::initialization (asv_initialization): density-based seed
        re-placement at the next prepare (or back to the plain grid)."""
        check(self.lib.asv_initialization(self.h, 1 if density else 0))

    def connectivity(self, supsz=0):
        """Image::enforce_connectivity_3d on the clip's current labels
        (asv_connectivity); the clip's labels are rewritten.  Returns the
        number of segments kept."""
        n = ctypes.c_int32(0)
        check(self.lib.asv_connectivity(self.h, int(supsz), ctypes.byref(n), self.stream))
        return n.value

    def run_timed(self):
        ms = np.zeros(6, np.float32)
        n = ctypes.c_int32(0)
        check(self.lib.asv_run_timed(self.h, self.stream, ptr(ms), ctypes.byref(n)))
        return ms, n.value

    def enable_trace(self, on=True):
        check(self.lib.asv_enable_trace(self.h, 1 if on else 0))

    def force_exact_clamp(self, on=True):
        """Every later run takes the marching pass's exact per-tap path."""
        check(self.lib.asv_force_exact_clamp(self.h, 1 if on else 0))

    def enable_march_stats(self, on=True):
        """Count the marching passes' steps in every later run (on zeroes
        the counters)."""
        check(self.lib.asv_enable_march_stats(self.h, 1 if on else 0))

    def march_stats(self, by_step=False):
        """Marching-step counters summed since they were enabled: steps,
        wide steps (candidate lists evaluated in chunks), their chunks and
        per-tap steps; for the steps within one candidate mask, candidate
        evaluations, wave rounds (each 64-voxel wave-slice's busiest voxel)
        and wave-slices; for the steps that extend a table generation, the
        waves' cycles per phase (claims, conversion, masks, load issue,
        evaluation) and those wave-steps; by_step adds the same per JFA step
        size."""
        out = np.zeros((16, 16), np.uint64)
        check(self.lib.asv_download_march_stats(self.h, ptr(out), self.stream))
        names = ("steps", "wide", "chunks", "per_tap", "candidates", "wave_rounds", "wave_slices",
                 "cyc_claim", "cyc_convert", "cyc_mask", "cyc_prefetch", "cyc_eval", "timed_wave_steps")

        def row(r):
            return {n: int(r[i]) for i, n in enumerate(names)}
        total = row(out.sum(axis=0))
        if by_step:
            total["by_step"] = {k + 1: row(out[k]) for k in range(16) if out[k, 0]}
        return total

    # -- outputs -------------------------------------------------------------
    def labels(self):
        out = np.empty(self.shape, np.int32)
        check(self.lib.asv_download_labels(self.h, ptr(out), self.stream))
        return out

    def sums(self):
        out = np.empty((self.A + 1, 8), np.int64)
        check(self.lib.asv_download_sums(self.h, ptr(out), self.stream))
        return out

    def trace(self):
        out = np.empty((self.p.lloyd_iters, self.A, 12), np.float32)
        check(self.lib.asv_download_trace(self.h, ptr(out), self.stream))
        return out

    def lab(self):
        out = np.empty(self.shape, np.uint32)
        check(self.lib.asv_download_lab(self.h, ptr(out), self.stream))
        return out

    def metric(self):
        """The per-voxel metric (m0..m5, psd flag, 0) of every voxel,
        computed on request (the Lloyd loop evaluates it only at seeds)."""
        out = np.empty(self.shape + (8,), np.float32)
        check(self.lib.asv_download_metric(self.h, ptr(out), self.stream))
        return out
