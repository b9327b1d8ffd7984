"""3D connectivity enforcement (GPU-AS/Image.cpp:1938-2062, enforce_connectivity_3d
:1808-1837): the oracle's restatement pinned by independent component labelling
and constructed cases, then the device version against the oracle bit for bit."""
import ctypes

import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
import oracle as orc

# the reference's neighbourhood: dz in -1..1 times the 8 in-plane offsets,
# (0, 0, +-1) excluded
STRUCT24 = np.ones((3, 3, 3), bool)
STRUCT24[0, 1, 1] = STRUCT24[2, 1, 1] = False


def components(labels):
    """Components of equal labels under STRUCT24 (scipy), numbered in raster
    order of their first voxel: (ids [D,H,W], sizes in that order)."""
    from scipy import ndimage
    comp = np.full(labels.shape, -1, np.int64)
    nxt = 0
    for val in np.unique(labels):
        lab, n = ndimage.label(labels == val, structure=STRUCT24)
        m = lab > 0
        comp[m] = lab[m] - 1 + nxt
        nxt += n
    flat = comp.ravel()
    first = np.full(nxt, flat.size, np.int64)
    np.minimum.at(first, flat, np.arange(flat.size))
    order = np.argsort(first)
    rank = np.empty(nxt, np.int64)
    rank[order] = np.arange(nxt)
    ids = rank[comp]
    sizes = np.bincount(ids.ravel(), minlength=nxt)
    return ids, sizes


def blocky(rng, shape, cell, n_labels, flip=0.0):
    """Piecewise-constant labels on a coarse grid, upsampled, plus noise."""
    D, H, W = shape
    coarse = rng.integers(0, n_labels, (D // cell + 1, H // cell + 1, W // cell + 1))
    lab = coarse.repeat(cell, 0).repeat(cell, 1).repeat(cell, 2)[:D, :H, :W].astype(np.int32)
    if flip:
        m = rng.random(shape) < flip
        lab[m] = rng.integers(-1, n_labels, int(m.sum()))
    return np.ascontiguousarray(lab)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_keeps_every_large_component(seed):
    rng = np.random.default_rng(seed)
    lab = blocky(rng, (9, 14, 17), 3, 4, flip=0.05)
    supsz = 40                                       # segments under 10 voxels merge
    out, n = orc.connectivity(lab, supsz)
    ids, sizes = components(lab)
    big = np.flatnonzero(sizes >= supsz // 4)
    assert n == len(big)
    # large components keep their voxels and are numbered in raster order
    for k, c in enumerate(big):
        assert np.all(out[ids == c] == k)
    # every small component is uniform and takes one of the kept numbers
    for c in np.flatnonzero(sizes < supsz // 4):
        v = np.unique(out[ids == c])
        assert len(v) == 1 and 0 <= v[0] < max(n, 1)


def test_oracle_no_merging_is_plain_component_labelling():
    rng = np.random.default_rng(5)
    lab = blocky(rng, (6, 10, 12), 2, 3, flip=0.2)
    out, n = orc.connectivity(lab, 0)                # threshold 0: nothing is small
    ids, sizes = components(lab)
    assert n == len(sizes)
    assert np.array_equal(out, ids)


def test_oracle_everything_small_collapses_to_zero():
    rng = np.random.default_rng(6)
    lab = blocky(rng, (5, 7, 9), 2, 5, flip=0.1)
    out, n = orc.connectivity(lab, 4 * lab.size + 4)
    assert n == 0 and np.all(out == 0)


def test_oracle_pure_z_neighbours_do_not_connect():
    lab = np.zeros((2, 3, 3), np.int32)
    lab[0, 1, 1] = 7
    lab[1, 1, 1] = 7                                 # stacked only along z
    out, n = orc.connectivity(lab, 0)
    assert out[0, 1, 1] != out[1, 1, 1]
    lab[1, 0, 1] = 7                                 # a diagonal link joins them
    out, n = orc.connectivity(lab, 0)
    assert out[0, 1, 1] == out[1, 1, 1]


def test_oracle_adjacent_label_is_last_visited_neighbour_in_scan_order():
    # A one-voxel segment of label 3 at (z=1, y=0, x=0) is the first voxel of
    # slice 1, so only slice 0 holds visited neighbours; in scan order they
    # are (1,0), (1,1), (0,1) (the in-plane offsets (1,0), (1,1), (0,1)).
    W, H = 8, 6
    lab = np.full((2, H, W), 9, np.int32)
    lab[1, 0, 0] = 3
    lab[0, :, 0] = 1                  # segment 0: column x = 0
    lab[0, :, 1:] = 2                 # segment 1: the rest of slice 0
    out, n = orc.connectivity(lab, 8) # segments under 2 voxels merge
    assert out[0, 0, 0] == 0 and out[0, 0, 1] == 1
    assert out[1, 0, 0] == out[0, 1, 0] == 0     # last neighbour (0,1) is in column 0
    lab[0] = 2
    lab[0, 0, :] = 1                  # segment 0: row y = 0; segment 1: rows y >= 1
    out, n = orc.connectivity(lab, 8)
    assert out[0, 0, 0] == 0 and out[0, 1, 0] == 1
    assert out[1, 0, 0] == 1                     # last neighbour (0,1) is in row 1, not (1,0) in row 0


def _device_connect(labels, supsz):
    import torch
    lib = asv._lib.load()
    D, H, W = labels.shape
    d_in = torch.from_numpy(np.ascontiguousarray(labels, np.int32)).cuda()
    d_out = torch.full_like(d_in, -7)
    torch.cuda.synchronize()
    n = ctypes.c_int32(-1)
    asv._lib.check(lib.asv_connectivity_labels(d_in.data_ptr(), d_out.data_ptr(), W, H, D, int(supsz),
                                               ctypes.byref(n), None))
    return d_out.cpu().numpy(), n.value


@pytest.mark.gpu
@pytest.mark.parametrize("shape,cell,nl,flip,supsz", [
    ((9, 14, 17), 3, 4, 0.05, 40),
    ((1, 33, 47), 2, 3, 0.1, 16),
    ((20, 1, 30), 2, 2, 0.2, 12),
    ((12, 40, 52), 4, 6, 0.02, 200),
    ((16, 48, 64), 5, 3, 0.3, 0),
    ((8, 30, 30), 3, 5, 0.1, 10**6),
    ((40, 120, 160), 6, 8, 0.01, 1000),
])
def test_device_matches_oracle(have_gpu, shape, cell, nl, flip, supsz):
    rng = np.random.default_rng(hash(shape) % 1000)
    lab = blocky(rng, shape, cell, nl, flip)
    print("connectivity case", shape, supsz, flush=True)
    want, nw = orc.connectivity(lab, supsz)
    got, ng = _device_connect(lab, supsz)
    assert ng == nw
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_device_on_a_run_matches_oracle(have_gpu):
    W, H, D, K = 64, 48, 16, 80
    rgb, fwd, bwd, _ = asv.synth_video(21, W, H, D, 6, 2.0, 4.0)
    p = asv.default_params(W, H, D, K, lloyd_iters=4)
    with asv.Clip(p) as c:
        c.upload(rgb, fwd, bwd)
        c.run()
        raw = c.labels()
        n = c.connectivity()
        got = c.labels()
    want, nw = orc.connectivity(raw, W * H * D // K)
    assert n == nw
    assert np.array_equal(got, want)


@pytest.mark.gpu
def test_cli_connect_reports_oracle_count(have_gpu, tmp_path):
    """The Render host's 6.2 path (GPU-AS/Render.cpp:324-341) through the CLI."""
    import os
    import re
    import subprocess
    seed, W, H, D, E, K, iters = 4, 48, 40, 12, 5, 60, 3
    cli = os.path.join(os.path.dirname(asv.__file__), "asv")
    r = subprocess.run([cli, "--synth", f"{seed},{W},{H},{D},{E},2,4", "--K", str(K), "--iters", str(iters),
                        "--connect", "--no-render", "--no-txt", "--save", str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = int(re.search(r"label_num=(\d+)", r.stdout).group(1))
    rgb, fwd, bwd, _ = asv.synth_video(seed, W, H, D, E, 2.0, 4.0)
    raw, _, _, _ = orc.run(rgb, fwd, bwd, K, iters)
    _, nw = orc.connectivity(raw, W * H * D // K)
    assert n == nw
