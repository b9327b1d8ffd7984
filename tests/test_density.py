"""Density-based seed re-placement (Image::initialization, GPU-AS/Image.cpp:350-620,
shaders/derivative_t.gs, shaders/density_grid.vs): the oracle's restatement
pinned by closed forms and a brute-force painting of the reference's grid
volume, then the device version against the oracle bit for bit."""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
import oracle as orc
import stream as ostream


def test_half_rounding_is_ieee_float16():
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.random(4000) * 800, np.arange(0, 40000, 1.0), rng.random(200) * 1e-4,
                         np.float32(0.33) * np.arange(0, 2300)]).astype(np.float32)
    ours = np.array([orc.half_round(x) for x in xs], np.float32)
    assert np.array_equal(ours, xs.astype(np.float16).astype(np.float32))


def test_flat_video_has_only_the_clear_density():
    flat = np.full((12, 40, 48, 3), 77, np.uint8)
    xyz, dens = orc.density_seeds(flat, 60)
    assert np.all(dens == 3.0)
    assert len(xyz) == len(orc.grid_seeds(48, 40, 12, 60)[0])


def test_unrepresentable_float16_ids_keep_the_clear_density():
    # ids pass through an R16F texture: in [2048, 4096) only even ids exist,
    # so every odd id's slot receives nothing and keeps the clear value 3
    W, H, D, K = 320, 240, 20, 3000
    rgb, _, _, _ = asv.synth_video(3, W, H, D, 10, 2.0, 4.0)
    xyz, dens = orc.density_seeds(rgb, K)
    A = len(xyz)
    assert A > 2100
    odd = np.arange(2049, min(A, 4096), 2)
    assert np.all(dens[odd] == 3.0)
    assert np.all(dens[odd - 1] > 3.0)                   # their even neighbours absorb both boxes


def grid_dims(W, H, D, K):
    """dNum, hNum, wNum (GPU-AS/Image.cpp:37-45)."""
    dnum = int(float(np.float32(K * D * D // W // H)) ** (1.0 / 3.0))
    dnum = max(dnum, 1)
    fs = K // dnum
    wnum = int(np.sqrt(np.float32(fs * W // H)))
    return dnum, fs // wnum, wnum


def painted_grid(W, H, D, K):
    """The reference's grid volume, painted box by box in seed order."""
    dn, hn, wn = grid_dims(W, H, D, K)
    g, _ = orc.grid_seeds(W, H, D, K)
    seeds = g.astype(np.float64) + 0.5
    vol = np.full((D, H, W), -1, np.int64)
    k = 0
    for t in range(dn):
        for j in range(hn):
            for i in range(wn):
                x, y, z = seeds[k]
                x1 = int(np.floor(0.5 * (seeds[k - 1][0] + x))) if i > 0 else 0
                x2 = int(np.floor(0.5 * (seeds[k + 1][0] + x))) if i < wn - 1 else W - 1
                y1 = int(np.floor(0.5 * (seeds[k - wn][1] + y))) if j > 0 else 0
                y2 = int(np.floor(0.5 * (seeds[k + wn][1] + y))) if j < hn - 1 else H - 1
                z1 = int(np.floor(0.5 * (seeds[k - wn * hn][2] + z))) if t > 0 else 0
                z2 = int(np.floor(0.5 * (seeds[k + wn * hn][2] + z))) if t < dn - 1 else D - 1
                vol[z1:z2 + 1, y1:y2 + 1, x1:x2 + 1] = k
                k += 1
    return vol


@pytest.mark.parametrize("spot", [(5, 17, 23), (0, 0, 0), (11, 39, 47), (6, 20, 0)])
def test_one_changed_voxel_feeds_the_boxes_of_its_four_taps(spot):
    W, H, D, K = 48, 40, 12, 60
    img = np.full((D, H, W, 3), 100, np.uint8)
    t0, y0, x0 = spot
    img[t0, y0, x0] = (130, 90, 100)                     # |delta| sum 40 against every neighbour
    _, dens = orc.density_seeds(img, K, prepared=1)
    grid = painted_grid(W, H, D, K)
    assert grid.min() >= 0
    expect = np.zeros(len(dens), np.float64)
    for (t, y, x) in [(t0, y0, x0), (t0, y0, x0 + 1), (t0, y0 + 1, x0), (t0 + 1, y0, x0)]:
        if t >= D or y >= H or x >= W:
            continue
        c = img[t, y, x].astype(int)
        nb = [img[t, y, max(x - 1, 0)], img[t, max(y - 1, 0), x], img[max(t - 1, 0), y, x]]
        g = np.float32(0.0)
        for n in nb:
            g = np.float32(g + np.float32(np.float32(0.33) * np.float32(np.abs(c - n.astype(int)).sum())))
        expect[grid[t, y, x]] += float(np.float16(g))
    want = (3.0 + expect).astype(np.float32)
    assert np.array_equal(dens, want)


def brute_force_densities(img, K):
    """Every slot's density from the painted grid volume, in numpy: the
    gradient of derivative_t.gs:24-49 (float32 sums, float16 storage), the
    slot of density_grid.vs:23-29 for the float16 id of the voxel's box, an
    exact sum per slot plus the clear colour 3 (Image.cpp:484-486), over all
    A slots (Image.cpp:524-552 reads A texels)."""
    D, H, W, _ = img.shape
    grid = painted_grid(W, H, D, K)
    A = int(grid.max()) + 1
    c = img.astype(np.int64)
    left = np.concatenate([c[:, :, :1], c[:, :, :-1]], axis=2)
    up = np.concatenate([c[:, :1], c[:, :-1]], axis=1)
    prev = np.concatenate([c[:1], c[:-1]], axis=0)
    g = np.zeros((D, H, W), np.float32)
    for nb in (left, up, prev):
        s = np.abs(nb - c).sum(axis=-1).astype(np.float32)
        g = (g + np.float32(0.33) * s).astype(np.float32)
    pg = g.astype(np.float16).astype(np.float32)
    kr = grid.astype(np.float32).astype(np.float16).astype(np.float32)
    wr = np.float32(1.0) / np.float32(W)
    ty = np.floor(kr * wr).astype(np.float32)
    tx = np.floor(kr - ty / wr).astype(np.float32)
    slot = (ty.astype(np.int64) * W + tx.astype(np.int64)).ravel()
    ok = (tx.ravel() >= 0) & (tx.ravel() < W) & (slot >= 0)
    acc = np.zeros(max(A, int(slot[ok].max()) + 1), np.int64)
    np.add.at(acc, slot[ok], (pg.ravel()[ok] * 4096.0).astype(np.int64))
    return (3.0 + acc[:A] / 4096.0).astype(np.float32)


def test_densities_cover_every_slot_when_float16_ids_stop_short():
    """320x240x20, K=2149: A = 2146 seeds, but the last id 2145 is odd and
    rounds to 2144 in float16, so no voxel addresses slot 2145.  The slot
    table must still cover all A slots (the last one at the clear density
    3), and every density equals the brute-force painting."""
    W, H, D, K = 320, 240, 20, 2149
    A = len(orc.grid_seeds(W, H, D, K)[0])
    assert A == 2146 and int(np.float16(A - 1)) + 1 < A
    rgb, _, _, _ = asv.synth_video(37, W, H, D, 10, 2.0, 4.0)
    img = orc.blur_flip(rgb)
    xyz, dens = orc.density_seeds(img, K, prepared=1)
    assert len(xyz) == A
    assert dens[A - 1] == 3.0
    assert np.array_equal(dens, brute_force_densities(img, K))


@pytest.mark.gpu
@pytest.mark.parametrize("seed,W,H,D,K,prepared", [
    (31, 48, 40, 12, 60, 0), (32, 160, 120, 30, 500, 0), (33, 96, 64, 20, 300, 1), (34, 320, 240, 20, 3000, 0),
    # the float16 rounding of the last id lies below it (A = 2146, 2145 -> 2144)
    (37, 320, 240, 20, 2149, 0),
])
def test_device_density_seeds_match_oracle(have_gpu, seed, W, H, D, K, prepared):
    rgb, fwd, bwd, _ = asv.synth_video(seed, W, H, D, 8, 2.0, 4.0)
    if prepared:
        rgb = orc.blur_flip(rgb)
    want, _ = orc.density_seeds(rgb, K, prepared=prepared)
    p = asv.default_params(W, H, D, K, lloyd_iters=2, seed_init=1, frames_prepared=prepared)
    with asv.Clip(p) as c:
        c.upload(rgb, fwd, bwd)
        grid = c.grid_seeds().copy()
        c.prepare()
        got = c.grid_seeds()
        c.prepare()                                      # a second prepare starts from the grid again
        again = c.grid_seeds()
    assert np.array_equal(grid, orc.grid_seeds(W, H, D, K)[0])
    assert np.array_equal(got, want)
    assert np.array_equal(again, want)


@pytest.mark.gpu
def test_density_run_matches_oracle(have_gpu):
    W, H, D, K, iters = 64, 48, 16, 90, 4
    rgb, fwd, bwd, _ = asv.synth_video(35, W, H, D, 6, 2.0, 4.0)
    xyz, _ = orc.density_seeds(rgb, K)
    s = orc.grid_seeds(W, H, D, K)[1]
    want, wsums = ostream.run_seeded(rgb, fwd, bwd, xyz, s, iters, 8.0, 1)
    p = asv.default_params(W, H, D, K, lloyd_iters=iters, seed_init=1)
    with asv.Clip(p) as c:
        c.upload(rgb, fwd, bwd)
        c.run()
        got = c.labels()
        sums = c.sums()
    assert np.array_equal(got, want)
    assert np.array_equal(sums[:, :7], wsums)


@pytest.mark.gpu
def test_initialization_switches_placement(have_gpu):
    W, H, D, K = 48, 40, 12, 60
    rgb, fwd, bwd, _ = asv.synth_video(36, W, H, D, 6, 2.0, 4.0)
    want, _ = orc.density_seeds(rgb, K)
    grid = orc.grid_seeds(W, H, D, K)[0]
    with asv.Clip(asv.default_params(W, H, D, K, lloyd_iters=2)) as c:
        c.upload(rgb, fwd, bwd)
        c.initialization(True)
        c.prepare()
        assert np.array_equal(c.grid_seeds(), want)
        c.initialization(False)
        assert np.array_equal(c.grid_seeds(), grid)
        c.prepare()
        assert np.array_equal(c.grid_seeds(), grid)


@pytest.mark.gpu
def test_cli_density_runs(have_gpu, tmp_path):
    """The Render host's step 2 (GPU-AS/Render.cpp:257-258) through the CLI."""
    import os
    import subprocess
    cli = os.path.join(os.path.dirname(asv.__file__), "asv")
    r = subprocess.run([cli, "--synth", "6,48,40,12,5,2,4", "--K", "60", "--iters", "3", "--density", "--no-render",
                        "--no-txt", "--save", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "final label num" in r.stdout
