"""Wide marching steps: a step whose blocks' columns reach more seeds than
one candidate mask holds (31 at steps 1-4 of a launch with at least 1500
voxels per seed, 63 otherwise) evaluates its list in chunks of that many
entries; a list past 253 entries, or any overflowing step at step 2,
evaluates per tap.  The marching-step counters show which of these paths
each case took; labels, sums and every iteration's seed records are
bit-exact against the oracle either way."""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
import oracle as orc

pytestmark = pytest.mark.gpu


def _run_counted(rgb, fwd, bwd, K, iters):
    D, H, W, _ = rgb.shape
    p = asv.default_params(W, H, D, K, lloyd_iters=iters)
    with asv.Clip(p) as c:
        c.enable_trace()
        c.enable_march_stats()
        c.upload(rgb, fwd, bwd)
        c.run()
        got = dict(labels=c.labels(), sums=c.sums(), A=c.A, trace=c.trace())
        stats = c.march_stats(by_step=True)
    return got, stats


def _assert_parity(got, rgb, fwd, bwd, K, iters):
    want_labels, want_sums, want_trace, A = orc.run(rgb, fwd, bwd, K, iters, want_trace=True)
    assert got["A"] == A
    tr_diff = np.argwhere(np.any(got["trace"] != want_trace, axis=2))
    assert tr_diff.size == 0, f"seed records differ first at (iter, seed) {tr_diff[:5].tolist()}"
    assert int((got["labels"] != want_labels).sum()) == 0
    assert np.array_equal(got["sums"][:, :7], want_sums)


@pytest.mark.parametrize("seed,shape,K,iters,expect", [
    # 14 voxels per seed (dense: 63-entry masks), marching steps 1, 2, 1:
    # step-1 lists of several 63-entry chunks, step-2 overflows per tap
    (8, (12, 48, 48), 2000, 3, {(1, "wide"), (2, "per_tap")}),
    # 1778 voxels per seed (31-entry masks at steps 1-4), marching steps 3, 2,
    # 1, 2, 1: wide step-3 lists of two 31-entry chunks
    (44, (60, 160, 200), 1100, 3, {(3, "wide")}),
    # 1016 voxels per seed (dense), grid step 10: every list in one mask
    (30, (60, 160, 200), 1900, 3, set()),
    # 4 voxels per seed: step-1 lists past 253 entries (per tap) and wide ones
    (31, (8, 32, 32), 2000, 2, {(1, "wide"), (1, "per_tap"), (2, "per_tap")}),
    # 19890 seeds, 16 voxels per seed: the hashed table's wide steps at step 1
    (47, (16, 160, 128), 20000, 2, {(1, "wide")}),
])
def test_wide_steps_bit_exact(have_gpu, seed, shape, K, iters, expect):
    D, H, W = shape
    rgb, fwd, bwd, _ = asv.synth_video(seed, W, H, D, 6, 2.0, 4.0)
    got, st = _run_counted(rgb, fwd, bwd, K, iters)
    by = st["by_step"]
    print(f"case {seed} {shape} K={K}: " + "; ".join(
        f"step {k}: wide {r['wide']} chunks {r['chunks']} per_tap {r['per_tap']} of {r['steps']}"
        for k, r in by.items()), flush=True)
    assert st["steps"] > 0
    assert st["chunks"] >= 2 * st["wide"]
    for k, path in expect:
        assert by[k][path] > 0, (k, path)
    if not expect:
        assert st["wide"] == 0 and st["per_tap"] == 0
    _assert_parity(got, rgb, fwd, bwd, K, iters)


def test_counting_kernel_matches_plain(have_gpu):
    """The counting form of the marching kernel gives the labels of the
    ordinary one, and the counters sum over runs until re-enabled."""
    rgb, fwd, bwd, _ = asv.synth_video(30, 200, 160, 60, 6, 2.0, 4.0)
    p = asv.default_params(200, 160, 60, 1900, lloyd_iters=3)
    with asv.Clip(p) as c:
        c.upload(rgb, fwd, bwd)
        c.run()
        plain = c.labels()
        c.enable_march_stats()
        c.run()
        counted = c.labels()
        one = c.march_stats()
        c.run()
        two = c.march_stats()
        c.enable_march_stats()
        zero = c.march_stats()
    assert np.array_equal(plain, counted)
    # the counts double over a second identical run (cycle counts are timings)
    counts = [k for k in one if not k.startswith("cyc_")]
    assert one["steps"] > 0 and all(two[k] == 2 * one[k] for k in counts)
    assert one["cyc_eval"] > 0 and one["timed_wave_steps"] > 0
    assert all(v == 0 for v in zero.values())
    # a wave's rounds cover each of its voxels' candidates
    assert one["candidates"] <= 64 * one["wave_rounds"] and one["wave_slices"] > 0


def test_benchmark_clip_step_paths(have_gpu):
    """The 640x480x200 benchmark clip (K=10000, marching steps 5, 3, 2, 1,
    2, 1): every step of each step size is counted (two slices per step,
    1200 columns of 16 x 16 below step 4, 600 of 32 x 16 from step 4), none
    needs more than one candidate mask, and
    the counting kernel's labels are the ordinary kernel's (whose labels the
    full-size parity test checks against the oracle)."""
    rgb, fwd, bwd, _ = asv.synth_video(2, 640, 480, 200, 32, 4.0, 6.0)
    p = asv.default_params(640, 480, 200, 10000, lloyd_iters=2)
    with asv.Clip(p) as c:
        c.upload(rgb, fwd, bwd)
        c.run()
        plain = c.labels()
        c.enable_march_stats()
        c.run()
        counted = c.labels()
        st = c.march_stats(by_step=True)
    for k, r in st["by_step"].items():
        print(f"benchmark clip, 2 iterations, step {k}: {r}, candidates per voxel "
              f"{r['candidates'] / (64 * r['wave_slices']):.2f}, lane use "
              f"{r['candidates'] / (64 * r['wave_rounds']):.3f}", flush=True)
    assert np.array_equal(plain, counted)
    assert sorted(st["by_step"]) == [1, 2, 3, 5]
    for k, r in st["by_step"].items():
        passes = 2 if k in (1, 2) else 1
        columns = (640 // (32 if k >= 4 else 16)) * (480 // 16)
        assert r["steps"] >= 2 * passes * columns * 100, k
        assert r["wide"] == 0 and r["per_tap"] == 0
        assert r["candidates"] <= 64 * r["wave_rounds"]
