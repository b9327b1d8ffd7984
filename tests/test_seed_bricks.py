"""The fused first two JFA passes (lowest-id rule, sparse seeds): one launch
sequence over bricks of the volume that reads the seed positions instead of
a seed volume (anisotropic_supervoxel_amd/csrc/asv_seedpass.hip).

Each case checks the second pass's output itself, right after the seed
stage (asv_flood_passes), against two oracle passes over the seed volume
the reference's draw_seeds leaves (every voxel -1, the higher id on a voxel
two seeds share), and then whole runs against the oracle: the later passes
could otherwise hide a wrong label of the fused passes.
"""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
import oracle as orc

pytestmark = pytest.mark.gpu


def _seed_volume(xyz, D, H, W):
    vol = np.full((D, H, W), -1, np.int32)
    for i, (x, y, z) in enumerate(xyz):
        vol[z, y, x] = max(vol[z, y, x], i)
    return vol


def _records(tr):
    """Trace rows (x, y, z, L, a, b, m0..m5) -> oracle seed records."""
    return orc.seeds_array([(r[0], r[1], r[2], r[3:6], r[6:12]) for r in tr])


def _check_fused_passes(rgb, fwd, bwd, xyz, s, compactness=8.0):
    D, H, W, _ = rgb.shape
    A = len(xyz)
    p = asv.default_params(W, H, D, A, lloyd_iters=2, compactness=compactness)
    with asv.Clip(p, seeds_xyz=xyz, grid_step=s) as c:
        steps, mult = c.schedule()
        # the fused form's conditions: lowest id, first step <= 2, a sparse
        # second pass of step <= 16
        assert steps[0] <= 2 and steps[1] <= 16 and A * 729 * 3 <= W * H * D, "case must take the fused passes"
        c.enable_trace()
        c.upload(rgb, fwd, bwd)
        c.prepare()
        c.seed_stage(1)
        with pytest.raises(asv.api.AsvError, match="ASV_ERR_ARG"):
            c.flood_passes(1, 1)
        c.flood_passes(1, 2)
        got = c.labels()
        lab = c.lab()
        tr = c.trace()[0]
    ratio = np.float32(compactness) / np.float32(s)
    ratio = np.float32(ratio * ratio)
    recs = _records(tr)
    vol = _seed_volume(xyz, D, H, W)
    out1 = orc.jfa_pass(vol, lab, recs, steps[0], float(np.float32(mult[0]) * ratio))
    want = orc.jfa_pass(out1, lab, recs, steps[1], float(np.float32(mult[1]) * ratio))
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} mismatches after two passes, first at (z, y, x) {bad[:5].tolist()}"
    assert (want >= 0).any()
    return p


def _check_run(rgb, fwd, bwd, xyz, s, iters=2):
    D, H, W, _ = rgb.shape
    p = asv.default_params(W, H, D, len(xyz), lloyd_iters=iters)
    with asv.Clip(p, seeds_xyz=xyz, grid_step=s) as c:
        c.upload(rgb, fwd, bwd)
        c.run()
        labels, sums = c.labels(), c.sums()
    want, wsums, _ = orc.run_seeded(rgb, fwd, bwd, xyz, s, iters)
    assert int((labels != want).sum()) == 0
    assert np.array_equal(sums[:, :7], wsums)


def _clustered_seeds(rng, W, H, D):
    """300 seeds in a 12^3 box (lists past the 256 a block stages in LDS,
    many chunks), 20 of them put on voxels other seeds hold (up to three a
    voxel), seeds on every face, edge and corner, and a few spread out."""
    c = rng.integers(0, 12, (300, 3)) + np.array([W // 2 - 6, H // 2 - 6, D // 2 - 6])
    c[280:] = c[rng.integers(0, 280, 20)]                      # collisions
    corners = np.array([[x, y, z] for x in (0, W - 1) for y in (0, H - 1) for z in (0, D - 1)])
    faces = np.stack([rng.integers(0, W, 24), rng.integers(0, H, 24), rng.integers(0, D, 24)], axis=1)
    faces[0:4, 0] = 0
    faces[4:8, 0] = W - 1
    faces[8:12, 1] = 0
    faces[12:16, 1] = H - 1
    faces[16:20, 2] = 0
    faces[20:24, 2] = D - 1
    spread = np.stack([rng.integers(0, W, 40), rng.integers(0, H, 40), rng.integers(0, D, 40)], axis=1)
    xyz = np.concatenate([c, corners, faces, spread]).astype(np.int32)
    return xyz[rng.permutation(len(xyz))]


@pytest.mark.parametrize("s", [3, 18, 33])
def test_fused_passes_clustered_collisions_faces(have_gpu, s):
    """Steps 1 then 1, 9 or 16: a brick list far past LDS staging, seeds
    sharing voxels (the lower ids hidden), taps clamped at every face."""
    W, H, D = 200, 160, 48
    rng = np.random.default_rng(100 + s)
    xyz = _clustered_seeds(rng, W, H, D)
    rgb, fwd, bwd, _ = asv.synth_video(31, W, H, D, 8, 2.0, 4.0)
    _check_fused_passes(rgb, fwd, bwd, xyz, s)
    _check_run(rgb, fwd, bwd, xyz, s)


def test_fused_passes_thin_clip(have_gpu):
    """Six frames and a step of 9: every tap along t is clamped at both ends,
    and bricks are partial along x, y and t."""
    W, H, D = 200, 100, 6
    rng = np.random.default_rng(7)
    xyz = np.stack([rng.integers(0, W, 26), rng.integers(0, H, 26), rng.integers(0, D, 26)], axis=1)
    xyz = np.concatenate([xyz, [[0, 0, 0], [W - 1, H - 1, D - 1], [W - 1, 0, 0], [3, 3, 3], [3, 3, 3]]])
    xyz = xyz.astype(np.int32)
    rgb, fwd, bwd, _ = asv.synth_video(32, W, H, D, 5, 2.0, 4.0)
    _check_fused_passes(rgb, fwd, bwd, xyz, 18)
    _check_run(rgb, fwd, bwd, xyz, 18, iters=3)


def test_fused_passes_benchmark_clip_first_iteration(have_gpu):
    """The 640x480x200, K = 10000 clip's grid seeds: the fused passes'
    output equals two oracle passes over the seed volume."""
    rgb, fwd, bwd, _ = asv.synth_video(2, 640, 480, 200, 32, 4.0, 6.0)
    xyz, s = orc.grid_seeds(640, 480, 200, 10000)
    _check_fused_passes(rgb, fwd, bwd, xyz, s)


@pytest.mark.parametrize("W,H,D,K,vseed", [
    (200, 160, 48, 120, 33),     # steps 1, 11, 6, 3, 2, 1, 2, 1
    (128, 96, 30, 15, 34),       # steps 1, 14, 7, 4, 2, 1, 2, 1: the step-7 pass reads 32-bit labels
    (200, 160, 48, 250, 35),     # steps 1, 9, 5, 3, 2, 1, 2, 1: the fused passes write 16-bit labels
    (200, 160, 48, 300, 36),     # steps 1, 8, 4, 2, 1, 2, 1
])
def test_partial_floods_end_in_32bit_labels(have_gpu, W, H, D, K, vseed):
    """A flood stopped after n passes (asv_flood_passes) for every n the
    fused form allows: the fused passes and marching passes that feed a
    marching pass of step 1-5 in the same flood hand it 16-bit labels and
    the flood's last pass writes 32-bit labels, so each stopping point sees
    the first, middle and last label widths in another combination.  Each partial flood equals
    the oracle's first n passes over the seed volume (-1 voxels included)."""
    rgb, fwd, bwd, _ = asv.synth_video(vseed, W, H, D, 8, 2.0, 4.0)
    xyz, s = orc.grid_seeds(W, H, D, K)
    A = len(xyz)
    p = asv.default_params(W, H, D, A, lloyd_iters=2)
    vol = _seed_volume(xyz, D, H, W)
    ratio = np.float32(np.float32(p.compactness) / np.float32(s))
    ratio = np.float32(ratio * ratio)
    got = {}
    with asv.Clip(p, seeds_xyz=xyz, grid_step=s) as c:
        steps, mult = c.schedule()
        assert steps[0] <= 2 and A * 729 * 3 <= W * H * D, "case must take the fused passes"
        c.enable_trace()
        c.upload(rgb, fwd, bwd)
        c.prepare()
        for n in range(2, len(steps) + 1):
            c.seed_stage(1)
            c.flood_passes(1, n)
            got[n] = c.labels()        # (iteration 1's seed stage may run again)
        lab = c.lab()
        recs = _records(c.trace()[0])
    want = vol
    for i, k in enumerate(steps):
        want = orc.jfa_pass(want, lab, recs, k, float(np.float32(mult[i]) * ratio))
        if i + 1 >= 2:
            bad = np.argwhere(got[i + 1] != want)
            assert bad.size == 0, f"{len(bad)} mismatches after {i + 1} passes, first at {bad[:5].tolist()}"
