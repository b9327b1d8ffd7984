"""Temporal blocks (the streamGPU-AS program): block plan on the CPU, and
the GPU block-by-block run against the oracle restatement in
oracle/stream.py, bit-exact in the stitched global labels."""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
from anisotropic_supervoxel_amd import api
import oracle as orc
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


@pytest.mark.gpu
def test_shifted_frames_keep_their_lab(have_gpu):
    """asv_shift_frames moves the overlap frames' LAB with them and the next
    prepare converts only the frames after them; frames re-uploaded inside
    the moved range are converted again.  Each prepare's LAB volume equals
    the oracle's LAB of the frames then on the device."""
    from anisotropic_supervoxel_amd import _lib
    W, H, D, O = 96, 64, 10, 5
    a = asv.synth_video(40, W, H, D, 6, 2.0, 4.0)
    b = asv.synth_video(41, W, H, D, 6, 2.0, 4.0)
    rgb_a, fwd_a, bwd_a = (np.ascontiguousarray(x) for x in a[:3])
    rgb_b, fwd_b, bwd_b = (np.ascontiguousarray(x) for x in b[:3])

    def want(rgb):
        return orc.lab_volume(orc.blur_flip(rgb))

    p = asv.default_params(W, H, D, 40, lloyd_iters=1)
    with asv.Clip(p) as c:
        lib = c.lib
        c.upload(rgb_a, fwd_a, bwd_a)
        c.prepare()
        assert np.array_equal(c.lab(), want(rgb_a))
        # the last O frames become the first O; new frames after them
        assert lib.asv_shift_frames(c.h, D - O, O, None) == 0
        first = O
        assert lib.asv_upload_frames_range(c.h, first, D - first, _lib.ptr(rgb_b[first:]), _lib.ptr(fwd_b[first:]),
                                           _lib.ptr(bwd_b[first:]), None) == 0
        c.prepare()
        now = np.concatenate([rgb_a[D - O:], rgb_b[first:]])
        assert np.array_equal(c.lab(), want(now))
        # again, with frames 3 .. D re-uploaded: frames 3 and 4 of the moved
        # range change and must be converted again
        assert lib.asv_shift_frames(c.h, D - O, O, None) == 0
        first = 3
        assert lib.asv_upload_frames_range(c.h, first, D - first, _lib.ptr(rgb_a[first:]), _lib.ptr(fwd_a[first:]),
                                           _lib.ptr(bwd_a[first:]), None) == 0
        c.prepare()
        now = np.concatenate([now[D - O:D - O + first], rgb_a[first:]])
        assert np.array_equal(c.lab(), want(now))
