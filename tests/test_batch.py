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
