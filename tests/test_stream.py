"""Temporal blocks (the streamGPU-AS program): block plan on the CPU, and
the GPU block-by-block run against the oracle restatement in
oracle/stream.py, bit-exact in the stitched global labels."""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
from anisotropic_supervoxel_amd import api
import stream as ostream


@pytest.mark.parametrize("D,B,O", [(25, 30, 10), (39, 30, 10), (40, 30, 10), (69, 30, 10), (70, 30, 10),
                                   (85, 30, 10), (200, 30, 10), (10, 10, 5), (14, 10, 5), (60, 10, 5),
                                   (63, 10, 5), (300, 10, 5), (97, 12, 4)])
def test_stream_plan_matches_restatement(D, B, O):
    sp = api.stream_params(64, 48, D, 500, block_size=B, overlap=O)
    got = api.stream_plan(sp)
    want = ostream.plan_blocks(D, B, O)
    assert got == want
    # blocks cover the video, consecutive blocks overlap by O frames
    assert got[0][0] == 0 and got[-1][1] == D
    for (a0, a1), (b0, b1) in zip(got, got[1:]):
        assert b0 == a1 - O


def test_stream_rejects_bad_params():
    with pytest.raises(api.AsvError):
        api.stream_plan(api.stream_params(64, 48, 50, 500, block_size=10, overlap=10))


@pytest.mark.gpu
@pytest.mark.parametrize("seed,W,H,D,K,B,O,iters", [
    (21, 48, 40, 70, 240, 30, 10, 4),     # the reference's 30/10 blocks, 3 blocks
    (22, 40, 32, 33, 150, 10, 5, 3),      # the served 10-frame windows, stride 5
    (23, 56, 40, 25, 200, 30, 10, 3),     # a single block
    (24, 128, 96, 40, 30, 30, 10, 3),     # sparse blocks (s 25-29, steps 1 then 12-14): the fused first passes
])
def test_stream_video_bit_exact(have_gpu, seed, W, H, D, K, B, O, iters):
    rgb, fwd, bwd, _ = asv.synth_video(seed, W, H, D, 6, 2.0, 4.0)
    sp = api.stream_params(W, H, D, K, block_size=B, overlap=O, lloyd_iters=iters)
    got, placed = api.stream_run_video(sp, rgb, fwd, bwd)
    want, wplaced = ostream.run_video(rgb, fwd, bwd, K, B, O, iters)
    assert placed == wplaced
    bad = int((got != want).sum())
    assert bad == 0, f"{bad} stitched labels differ"
    assert (got >= 0).all()
