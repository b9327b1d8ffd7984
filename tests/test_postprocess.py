"""CPU tier: the library's host post-processing and output formats against
the pure-Python restatement in oracle/post.py, on constructed cases that
hit every branch of check_dead_seeds (GPU-AS/Image.cpp:1579-1720)."""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
from anisotropic_supervoxel_amd import api
import post as ref


def _case(rng, A, V, dead, unlabelled, lead_unlabelled=False):
    alive = [i for i in range(A) if i not in dead]
    labels = rng.choice(alive, V).astype(np.int32)
    # runs, as real label volumes have
    labels = np.repeat(labels[: V // 4 + 1], 4)[:V].copy()
    if unlabelled:
        idx = rng.choice(V, unlabelled, replace=False)
        labels[idx] = -1
    if lead_unlabelled:
        labels[0] = -1
    counts = np.array([(labels == i).sum() for i in range(A)] + [(labels < 0).sum()], np.int64)
    sums = np.zeros((A + 1, 8), np.int64)
    sums[:, 3] = counts
    return labels, sums, counts


@pytest.mark.parametrize("A,dead,unl,lead", [
    (12, [], 0, False),              # clean
    (12, [], 7, False),              # -1 voxels only
    (12, [], 7, True),               # leading -1 stays -1
    (12, [3], 0, False),             # one dead: last id -> dead id
    (12, [0, 11], 5, False),         # two dead, one is the last id
    (12, [2, 5, 9], 3, False),       # three dead
    (12, [1, 2, 4, 8, 10], 0, False),  # more than three: remap
    (12, [0, 3, 4, 6], 6, True),     # remap with -1 and the stale-first-label quirk
])
def test_postprocess_matches_restatement(A, dead, unl, lead):
    rng = np.random.default_rng(A + len(dead) * 7 + unl)
    labels, sums, counts = _case(rng, A, 400, dead, unl, lead)
    got, ndead, nfin = api.postprocess(sums, A, labels)
    want, wdead, wfin = ref.check_dead_seeds(list(counts), A, list(labels))
    assert ndead == wdead == len(dead)
    assert nfin == wfin == A - len(dead)
    assert got.tolist() == want


def test_label_txt_format(tmp_path):
    rng = np.random.default_rng(0)
    labels = rng.integers(-1, 3000, (3, 4, 5)).astype(np.int32)
    p = tmp_path / "500.txt"
    api.write_label_txt(labels, p)
    assert p.read_text() == ref.label_txt(labels.tolist())


def test_render_ppm(tmp_path):
    labels = np.array([[[0, 1, -1], [2, 2, 0]]], np.int32)        # D=1, H=2, W=3
    api.write_render_ppm(labels, 3, tmp_path / "f")
    data = (tmp_path / "f00000.ppm").read_bytes()
    head = b"P6\n3 2\n255\n"
    assert data.startswith(head)
    img = np.frombuffer(data[len(head):], np.uint8).reshape(2, 3, 3)
    # the top image row is GL row 1 = labels (2, 2, 0); the bottom is (0, 1, -1)
    assert (img[0, 0] == img[0, 1]).all()                          # label 2 twice
    assert (img[1, 2] == 0).all()                                  # label -1 is black
    assert (img[1, 0] == img[0, 2]).all()                          # label 0 in both rows


def test_eval_metrics_closed_forms():
    D, H, W = 2, 8, 8
    gt = np.zeros((D, H, W), np.int32)
    gt[:, :, 4:] = 1
    # perfect segmentation: UE 0, BR 1
    ue, br = api.eval_metrics(gt, gt)
    assert ue == 0.0 and br == 1.0
    # one supervoxel over everything: each region leaks the other half
    one = np.zeros_like(gt)
    ue, br = api.eval_metrics(one, gt)
    assert ue == pytest.approx(1.0) and br == 0.0
    # boundary shifted by one column: still recalled within tolerance 1
    sh = np.zeros_like(gt)
    sh[:, :, 5:] = 1
    ue, br = api.eval_metrics(sh, gt)
    assert br == 1.0
    # region 0 (64 voxels) meets supervoxel 0 (80): 16/64; region 1 meets
    # supervoxels 0 and 1 (80 + 48): 64/64
    assert ue == pytest.approx((16 / 64 + 64 / 64) / 2)


def _device_post(sums, A, labels):
    import ctypes
    import torch
    lib = asv._lib.load()
    d_in = torch.from_numpy(np.ascontiguousarray(labels, np.int32)).cuda()
    d_out = torch.full_like(d_in, -7)
    torch.cuda.synchronize()
    dead = ctypes.c_int32(-1)
    fin = ctypes.c_int32(-1)
    s = np.ascontiguousarray(sums, np.int64)
    asv._lib.check(lib.asv_postprocess_labels(s.ctypes.data, A, d_in.data_ptr(), d_out.data_ptr(), labels.size,
                                              ctypes.byref(dead), ctypes.byref(fin), None))
    return d_out.cpu().numpy(), dead.value, fin.value


@pytest.mark.gpu
@pytest.mark.parametrize("A,dead,unl,lead", [
    (12, [], 7, False), (12, [], 7, True), (12, [3], 0, False), (12, [0, 11], 5, False),
    (12, [2, 5, 9], 3, False), (12, [1, 2, 4, 8, 10], 0, False), (12, [0, 3, 4, 6], 6, True),
])
def test_device_postprocess_matches_restatement(have_gpu, A, dead, unl, lead):
    rng = np.random.default_rng(A + len(dead) * 7 + unl)
    labels, sums, counts = _case(rng, A, 400, dead, unl, lead)
    got, ndead, nfin = _device_post(sums, A, labels)
    want, wdead, wfin = ref.check_dead_seeds(list(counts), A, list(labels))
    assert (ndead, nfin) == (wdead, wfin)
    assert got.tolist() == want


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["fill", "fill_lead_run", "replace_fill", "remap", "remap_lead_run", "remap_l0_runs"])
def test_device_postprocess_large_matches_host(have_gpu, case):
    """Volumes spanning many 4096-voxel scan chunks, with runs of -1 (and of
    the first voxel's label) that cross chunk boundaries."""
    rng = np.random.default_rng(sum(map(ord, case)))
    A, V = 300, 4096 * 29 + 1234
    dead = {"fill": [], "fill_lead_run": [], "replace_fill": [7, 299], "remap": [1, 5, 9, 13, 200],
            "remap_lead_run": [2, 3, 4, 5], "remap_l0_runs": [10, 11, 12, 13, 14]}[case]
    alive = np.array([i for i in range(A) if i not in dead])
    labels = np.repeat(rng.choice(alive, V // 37 + 1), 37)[:V].astype(np.int32)
    if case != "remap":
        for _ in range(6):                     # long -1 runs across chunk edges
            a = int(rng.integers(0, V - 9000))
            labels[a:a + int(rng.integers(1, 9000))] = -1
    if case.endswith("lead_run"):
        labels[:10000] = -1
    if case == "remap_l0_runs":
        l0 = labels[0]
        for _ in range(5):
            a = int(rng.integers(1, V - 6000))
            labels[a:a + 6000] = l0
    counts = np.array([(labels == i).sum() for i in range(A)] + [(labels < 0).sum()], np.int64)
    sums = np.zeros((A + 1, 8), np.int64)
    sums[:, 3] = counts
    got, ndead, nfin = _device_post(sums, A, labels)
    want, wdead, wfin = api.postprocess(sums, A, labels)
    nd = int((counts[:A] < 1).sum())          # planted runs can empty more ids
    assert nd >= len(dead)
    assert (ndead, nfin) == (wdead, wfin) == (nd, A - nd)
    assert np.array_equal(got, want)
