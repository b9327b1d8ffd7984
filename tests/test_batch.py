"""Batched clips: every clip's labels and sums bit-identical to its own
single-clip run (and, for small clips, to the oracle), through the batched
marching launches, the per-clip fallback (a step above 16, the scan-order
tie), graph replays across runs and with the autotuner on or off."""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
from anisotropic_supervoxel_amd import batch, api
import oracle as orc

# (seed, W, H, D, K, tie_rule): uneven sizes and steps; (96, 80, 40, K=6)
# has s = 37, so its first halving step is 18 (no marching kernel)
CLIPS = [(11, 48, 40, 12, 40, 0), (12, 64, 48, 21, 120, 0), (13, 40, 32, 30, 30, 0), (14, 96, 80, 40, 6, 0),
         (15, 33, 29, 17, 25, 0), (16, 56, 40, 26, 200, 1), (17, 80, 64, 9, 300, 0)]


def test_workload_is_deterministic():
    a, b = batch.make_clips(), batch.make_clips()
    assert a == b and len(a) == 64
    assert [c["seed"] for c in a] == list(range(100, 164))
    assert {(c["W"], c["H"]) for c in a} <= set(batch.SIZES)
    assert all(20 <= c["D"] <= 120 and 200 <= c["K"] <= 4000 and 8 <= c["ellipses"] <= 32 for c in a)


def test_batch_rejects_mixed_iterations():
    p1 = asv.default_params(32, 32, 8, 20, lloyd_iters=3)
    p2 = asv.default_params(32, 32, 8, 20, lloyd_iters=4)
    with pytest.raises(asv._lib.AsvError):
        batch.Batch([p1, p2])


def _inputs():
    out = []
    for seed, W, H, D, K, tie in CLIPS:
        rgb, fwd, bwd, _ = asv.synth_video(seed, W, H, D, 6, 2.0, 4.0)
        p = asv.default_params(W, H, D, K, lloyd_iters=4, tie_rule=tie)
        out.append((p, rgb, fwd, bwd))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("autotune", [True, False])
def test_batch_equals_single_clip_runs(have_gpu, autotune):
    ins = _inputs()
    want = []
    for p, rgb, fwd, bwd in ins:
        with asv.Clip(p) as c:
            c.upload(rgb, fwd, bwd)
            c.run()
            want.append((c.labels(), c.sums()))
    b = batch.Batch([x[0] for x in ins], autotune=autotune)
    try:
        for i, (p, rgb, fwd, bwd) in enumerate(ins):
            b.upload(i, rgb, fwd, bwd)
        for rep in range(2):                                  # second run replays the captured graphs
            b.run()
            shapes, _ = b.tuning()
            if not autotune:
                assert any(s[2] > 1 for s in shapes)          # launches covering several clips
            assert any(s[2] == -1 for s in shapes)            # the per-clip fallback (step 18)
            for i in range(len(ins)):
                assert np.array_equal(b.labels(i), want[i][0]), (rep, i)
                assert np.array_equal(b.sums(i), want[i][1]), (rep, i)
    finally:
        b.close()


@pytest.mark.gpu
def test_batch_matches_oracle(have_gpu):
    ins = _inputs()[:3]
    b = batch.Batch([x[0] for x in ins])
    try:
        for i, (p, rgb, fwd, bwd) in enumerate(ins):
            b.upload(i, rgb, fwd, bwd)
        b.run()
        for i, (p, rgb, fwd, bwd) in enumerate(ins):
            labels, sums, _, _ = orc.run(rgb, fwd, bwd, p.seeds_num, p.lloyd_iters)
            assert np.array_equal(b.labels(i), labels)
            assert np.array_equal(b.sums(i)[:, :7], sums[:, :7])
    finally:
        b.close()


def _stratified(clips):
    """Indices of the config-4 clips checked against the oracle: the
    smallest clip of each frame size, the smallest and largest K, the
    longest clip and the largest clip."""
    pick = set()
    for W, H in batch.SIZES:
        same = [i for i, c in enumerate(clips) if (c["W"], c["H"]) == (W, H)]
        if same:
            pick.add(min(same, key=lambda i: clips[i]["D"]))
    vol = [c["W"] * c["H"] * c["D"] for c in clips]
    pick.add(min(range(len(clips)), key=lambda i: clips[i]["K"]))
    pick.add(max(range(len(clips)), key=lambda i: clips[i]["K"]))
    pick.add(max(range(len(clips)), key=lambda i: clips[i]["D"]))
    pick.add(max(range(len(clips)), key=lambda i: vol[i]))
    return sorted(pick)


@pytest.mark.gpu
@pytest.mark.slow
@pytest.mark.timeout(900)
def test_config4_64_clips(have_gpu):
    """The batched benchmark workload itself: 64 clips (video seeds 100-163,
    160x120 to 640x480, 20-120 frames, K 200-4000, 10 iterations, 644 M
    voxels) in one batch, labels and sums of every clip bit-identical to its
    own single-clip run, over two batched runs (the second replays the
    captured graphs), and a stratified subset of eight clips equal to the
    oracle."""
    clips = batch.make_clips()
    params = [batch.clip_params(c) for c in clips]
    single = []
    b = batch.Batch(params)
    try:
        for i, (c, p) in enumerate(zip(clips, params)):
            rgb, fwd, bwd, _ = batch.clip_video(c)
            with asv.Clip(p) as cl:
                cl.upload(rgb, fwd, bwd)
                cl.run()
                single.append((cl.labels(), cl.sums()))
            b.upload(i, rgb, fwd, bwd)
            print(f"config 4: clip {i} ({c['W']}x{c['H']}x{c['D']}, K={c['K']}) run alone", flush=True)
        for rep in range(2):
            b.run()
            for i in range(len(clips)):
                assert np.array_equal(b.labels(i), single[i][0]), (rep, i)
                assert np.array_equal(b.sums(i), single[i][1]), (rep, i)
    finally:
        b.close()
    picks = _stratified(clips)
    assert len(picks) >= 6
    for i in picks:
        c = clips[i]
        rgb, fwd, bwd, _ = batch.clip_video(c)
        labels, sums, _, _ = orc.run(rgb, fwd, bwd, c["K"], c["iters"])
        print(f"config 4: clip {i} checked against the oracle", flush=True)
        assert np.array_equal(single[i][0], labels), c
        assert np.array_equal(single[i][1][:, :7], sums), c
