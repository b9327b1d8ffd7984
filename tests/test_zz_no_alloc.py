"""asv_postprocess_labels_noalloc (the post-processing entry point that
allocates nothing: its scratch is a fixed per-device array) against
asv_postprocess_labels on the same inputs, bit for bit, over two calls in a
row -- the second with a smaller volume and fewer seeds, so stale scratch
from the first call would show -- and its refusal of sizes beyond the
array."""
import ctypes

import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv

pytestmark = pytest.mark.gpu


def _case(rng, A, V, dead, unlabelled_runs):
    alive = np.array([i for i in range(A) if i not in dead])
    labels = np.repeat(rng.choice(alive, V // 37 + 1), 37)[:V].astype(np.int32)
    for _ in range(unlabelled_runs):
        a = int(rng.integers(0, max(1, V - 9000)))
        labels[a:a + int(rng.integers(1, 9000))] = -1
    counts = np.array([(labels == i).sum() for i in range(A)] + [(labels < 0).sum()], np.int64)
    sums = np.zeros((A + 1, 8), np.int64)
    sums[:, 3] = counts
    return labels, sums


def _run(fn, sums, A, labels):
    import torch
    d_in = torch.from_numpy(labels).cuda()
    d_out = torch.full_like(d_in, -7)
    torch.cuda.synchronize()
    dead, fin = ctypes.c_int32(-1), ctypes.c_int32(-1)
    rc = fn(sums.ctypes.data, A, d_in.data_ptr(), d_out.data_ptr(), labels.size, ctypes.byref(dead),
            ctypes.byref(fin), None)
    asv._lib.check(rc)
    return d_out.cpu().numpy(), dead.value, fin.value


def test_noalloc_equals_allocating_entry_point(have_gpu):
    lib = asv._lib.load()
    rng = np.random.default_rng(5)
    cases = [(3000, 4096 * 301 + 777, list(range(0, 3000, 97)), 8),      # remap, ~1.2 M voxels
             (300, 4096 * 29 + 1234, [7, 299], 6),                        # replace + fill, smaller
             (12, 400, [], 1)]                                            # fill only, smallest
    for A, V, dead, runs in cases:
        labels, sums = _case(rng, A, V, dead, runs)
        want = _run(lib.asv_postprocess_labels, sums, A, labels)
        got = _run(lib.asv_postprocess_labels_noalloc, sums, A, labels)
        assert got[1:] == want[1:]
        assert np.array_equal(got[0], want[0]), (A, V)


def test_noalloc_refuses_sizes_beyond_its_scratch(have_gpu):
    import torch
    lib = asv._lib.load()
    d = torch.zeros(16, dtype=torch.int32, device="cuda")
    e = torch.zeros(16, dtype=torch.int32, device="cuda")
    sums = np.zeros((65538, 8), np.int64)
    dead, fin = ctypes.c_int32(0), ctypes.c_int32(0)
    # the sizes are checked before any device memory is touched
    assert lib.asv_postprocess_labels_noalloc(sums.ctypes.data, 65537, d.data_ptr(), e.data_ptr(), 16,
                                              ctypes.byref(dead), ctypes.byref(fin), None) == -1
    assert lib.asv_postprocess_labels_noalloc(sums.ctypes.data, 12, d.data_ptr(), e.data_ptr(), (1 << 29) + 1,
                                              ctypes.byref(dead), ctypes.byref(fin), None) == -1
