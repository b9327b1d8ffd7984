"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package never
imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("W", ctypes.c_int), ("H", ctypes.c_int), ("D", ctypes.c_int), ("K", ctypes.c_int),
        ("iters", ctypes.c_int), ("compactness", ctypes.c_float), ("tie_rule", ctypes.c_int),
        ("angle_gate", ctypes.c_int), ("schedule", ctypes.c_int), ("prepared", ctypes.c_int),
    ]


class OrcSeed(ctypes.Structure):
    _fields_ = [("x", ctypes.c_int), ("y", ctypes.c_int), ("z", ctypes.c_int),
                ("lab", ctypes.c_float * 3), ("m", ctypes.c_float * 6)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        i = ctypes.c_int
        f = ctypes.c_float
        sig = {
            "orc_lin_table": (None, [vp]),
            "orc_lab_of": (None, [i, i, i, vp]),
            "orc_blur_flip": (None, [vp, i, i, i, vp]),
            "orc_lab_volume": (None, [vp, ctypes.c_int64, vp]),
            "orc_metric_at": (None, [vp, vp, i, i, i, i, i, i, vp, vp]),
            "orc_metric_field": (None, [vp, vp, i, i, i, vp, vp]),
            "orc_psd3": (i, [vp]),
            "orc_nearest_psd": (None, [vp, vp]),
            "orc_grid_seeds": (i, [i, i, i, i, vp, vp]),
            "orc_schedule": (i, [i, i, i, vp, vp, i]),
            "orc_jfa_pass": (None, [vp, vp, vp, vp, i, i, i, i, f, i]),
            "orc_reduce": (None, [vp, vp, i, i, i, i, vp]),
            "orc_finalize": (None, [vp, vp, i, i, i, i, vp]),
            "orc_run": (i, [vp, vp, vp, vp, vp, vp, vp]),
            "orc_seed_record_floats": (i, []),
            "orc_connectivity": (i, [vp, i, i, i, i, vp]),
            "orc_half_round": (f, [f]),
            "orc_density_seeds": (i, [vp, i, i, i, i, vp, vp]),
            "orc_seed_metric": (None, [vp, vp, vp, i, i, i, i, i, i, i, vp]),
            "orc_angle_gate": (i, [f, f, f, f]),
            "orc_run_seeded": (i, [vp, vp, vp, vp, vp, i, i, vp, vp, vp]),
            "orc_set_threads": (None, [i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


def lin_table():
    out = np.zeros(256, np.float32)
    load().orc_lin_table(_p(out))
    return out


def lab_of(r, g, b):
    out = np.zeros(3, np.uint8)
    load().orc_lab_of(int(r), int(g), int(b), _p(out))
    return tuple(int(v) for v in out)


def lab_volume(rgb):
    rgb = np.ascontiguousarray(rgb, np.uint8)
    out = np.zeros(rgb.shape[:-1], np.uint32)
    load().orc_lab_volume(_p(rgb), out.size, _p(out))
    return out


def blur_flip(raw):
    raw = np.ascontiguousarray(raw, np.uint8)
    D, H, W, _ = raw.shape
    out = np.zeros_like(raw)
    load().orc_blur_flip(_p(raw), W, H, D, _p(out))
    return out


def metric_at(fwd, bwd, x, y, t):
    D, H, W, _ = fwd.shape
    fwd = np.ascontiguousarray(fwd, np.float32)
    bwd = np.ascontiguousarray(bwd, np.float32)
    out = np.zeros(6, np.float32)
    psd = ctypes.c_int(0)
    load().orc_metric_at(_p(fwd), _p(bwd), W, H, D, x, y, t, _p(out), ctypes.byref(psd))
    return out, psd.value


def metric_field(fwd, bwd):
    D, H, W, _ = fwd.shape
    fwd = np.ascontiguousarray(fwd, np.float32)
    bwd = np.ascontiguousarray(bwd, np.float32)
    m6 = np.zeros((D, H, W, 6), np.float32)
    psd = np.zeros((D, H, W), np.uint8)
    load().orc_metric_field(_p(fwd), _p(bwd), W, H, D, _p(m6), _p(psd))
    return m6, psd


def psd3(m):
    m = np.ascontiguousarray(m, np.float32)
    return load().orc_psd3(_p(m))


def nearest_psd(m):
    m = np.ascontiguousarray(m, np.float32)
    out = np.zeros(6, np.float32)
    load().orc_nearest_psd(_p(m), _p(out))
    return out


def grid_seeds(W, H, D, K):
    n = load().orc_grid_seeds(W, H, D, K, None, None)
    if n <= 0:
        return np.zeros((0, 3), np.int32), 0
    xyz = np.zeros((n, 3), np.int32)
    s = ctypes.c_int(0)
    load().orc_grid_seeds(W, H, D, K, _p(xyz), ctypes.byref(s))
    return xyz, s.value


def schedule(mode, s, W):
    steps = np.zeros(64, np.int32)
    mult = np.zeros(64, np.float32)
    n = load().orc_schedule(mode, s, W, _p(steps), _p(mult), 64)
    return steps[:n].tolist(), mult[:n].tolist()


def seeds_array(recs):
    """list of (x, y, z, (L,a,b), (m0..m5)) -> ctypes array of OrcSeed."""
    arr = (OrcSeed * len(recs))()
    for i, (x, y, z, lab, m) in enumerate(recs):
        arr[i].x, arr[i].y, arr[i].z = int(x), int(y), int(z)
        for k in range(3):
            arr[i].lab[k] = float(lab[k])
        for k in range(6):
            arr[i].m[k] = float(m[k])
    return arr


def jfa_pass(labels_in, lab, seeds, k, ratio, tie_rule=0):
    labels_in = np.ascontiguousarray(labels_in, np.int32)
    lab = np.ascontiguousarray(lab, np.uint32)
    D, H, W = labels_in.shape
    out = np.zeros_like(labels_in)
    load().orc_jfa_pass(_p(labels_in), _p(out), _p(lab), ctypes.cast(seeds, ctypes.c_void_p),
                        W, H, D, k, ratio, tie_rule)
    return out


def reduce(labels, lab, A):
    labels = np.ascontiguousarray(labels, np.int32)
    lab = np.ascontiguousarray(lab, np.uint32)
    D, H, W = labels.shape
    sums = np.zeros((A + 1, 7), np.int64)
    load().orc_reduce(_p(labels), _p(lab), W, H, D, A, _p(sums))
    return sums


def finalize(sums, lab, W, H, D, A):
    """orc_finalize on (A+1) x 7 sums; returns [(x, y, z, (L, a, b))] * A."""
    sums = np.ascontiguousarray(sums, np.int64)
    lab = np.ascontiguousarray(lab, np.uint32)
    arr = (OrcSeed * A)()
    load().orc_finalize(_p(sums), _p(lab), W, H, D, A, ctypes.cast(arr, ctypes.c_void_p))
    return [(arr[i].x, arr[i].y, arr[i].z, tuple(arr[i].lab)) for i in range(A)]


def run(rgb, fwd, bwd, K, iters, compactness=8.0, tie_rule=0, angle_gate=1, schedule_mode=0,
        prepared=0, want_trace=False):
    """Whole plain run.  Returns (labels [D,H,W], sums [A+1,7], trace or None, A)."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    fwd = np.ascontiguousarray(fwd, np.float32)
    bwd = np.ascontiguousarray(bwd, np.float32)
    D, H, W, _ = rgb.shape
    A = load().orc_grid_seeds(W, H, D, K, None, None)
    if A <= 0:
        raise ValueError("grid seeding places no seed")
    P = OrcParams(W, H, D, K, iters, compactness, tie_rule, angle_gate, schedule_mode, prepared)
    labels = np.zeros((D, H, W), np.int32)
    sums = np.zeros((A + 1, 7), np.int64)
    trace = np.zeros((iters, A, 12), np.float32) if want_trace else None
    r = load().orc_run(ctypes.byref(P), _p(rgb), _p(fwd), _p(bwd), _p(labels), _p(sums),
                       _p(trace) if want_trace else None)
    if r < 0:
        raise ValueError(f"orc_run failed: {r}")
    return labels, sums, trace, A


def connectivity(labels, supsz):
    """GPU-AS/Image.cpp:1938-2062 restated: (connected labels [D,H,W], count)."""
    labels = np.ascontiguousarray(labels, np.int32)
    D, H, W = labels.shape
    out = np.zeros_like(labels)
    n = load().orc_connectivity(_p(labels), W, H, D, int(supsz), _p(out))
    if n < 0:
        raise MemoryError("orc_connectivity")
    return out, n


def seed_metric(m6, psd, fwd, x, y, z, gate):
    """seed_matrix's selection (GPU-AS/Image.cpp:689-872, :1351-1534) over a
    given metric field m6 [D,H,W,6], PSD flags psd [D,H,W] and image-oriented
    forward flow fwd [D,H,W,2] (read by the angle gate)."""
    m6 = np.ascontiguousarray(m6, np.float32)
    psd = np.ascontiguousarray(psd, np.uint8)
    fwd = np.ascontiguousarray(fwd, np.float32)
    D, H, W = psd.shape
    out = np.zeros(6, np.float32)
    load().orc_seed_metric(_p(m6), _p(psd), _p(fwd), W, H, D, int(x), int(y), int(z), int(gate), _p(out))
    return out


def angle_gate(su, sv, nu, nv):
    return bool(load().orc_angle_gate(float(su), float(sv), float(nu), float(nv)))


def run_seeded(rgb, fwd, bwd, xyz, s, iters, compactness=8.0, tie_rule=0, angle_gate_on=1, prepared=0,
               want_trace=False):
    """Plain run from explicit seed voxels xyz (A x 3) and grid step s.
    Returns (labels [D,H,W], sums [A+1,7], trace or None)."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    fwd = np.ascontiguousarray(fwd, np.float32)
    bwd = np.ascontiguousarray(bwd, np.float32)
    xyz = np.ascontiguousarray(xyz, np.int32)
    D, H, W, _ = rgb.shape
    A = len(xyz)
    P = OrcParams(W, H, D, A, iters, compactness, tie_rule, angle_gate_on, 0, prepared)
    labels = np.zeros((D, H, W), np.int32)
    sums = np.zeros((A + 1, 7), np.int64)
    trace = np.zeros((iters, A, 12), np.float32) if want_trace else None
    r = load().orc_run_seeded(ctypes.byref(P), _p(rgb), _p(fwd), _p(bwd), _p(xyz), A, int(s), _p(labels), _p(sums),
                              _p(trace) if want_trace else None)
    if r < 0:
        raise ValueError(f"orc_run_seeded failed: {r}")
    return labels, sums, trace


def half_round(x):
    return load().orc_half_round(float(x))


def density_seeds(rgb, K, prepared=0):
    """Image::initialization (GPU-AS/Image.cpp:350-620) restated: the grid
    seeds re-placed by colour-gradient density.  rgb: D x H x W x 3 raw
    frames (blurred and flipped here unless prepared).  Returns (A x 3 voxel
    positions, A slot densities)."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    img = rgb if prepared else blur_flip(rgb)
    D, H, W, _ = img.shape
    A = load().orc_grid_seeds(W, H, D, K, None, None)
    xyz = np.zeros((max(A, 1), 3), np.int32)
    dens = np.zeros(max(A, 1), np.float32)
    r = load().orc_density_seeds(_p(img), W, H, D, K, _p(xyz), _p(dens))
    if r < 0:
        raise ValueError(f"orc_density_seeds failed: {r}")
    return xyz[:r], dens[:r]
