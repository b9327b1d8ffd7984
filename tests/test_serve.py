"""Served streaming: admission by footprint, and each stream's result the
same alone and in the mix (bit-exact stitched labels)."""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
from anisotropic_supervoxel_amd import api, serve
import stream as ostream


def _small_requests():
    sizes = [(48, 40, 33, 1.5), (64, 48, 21, 2.0), (40, 32, 47, 3.0), (56, 40, 26, 1.0), (32, 32, 40, 4.0),
             (48, 48, 15, 2.5)]
    return [dict(stream_id=i, width=W, height=H, frames=D, density=d, n_ellipses=5, video_seed=100 + i,
                 vmax=2.0, noise_sigma=4.0, arrival_ms=3.0 * i) for i, (W, H, D, d) in enumerate(sizes)]


def test_requests_distribution():
    reqs = serve.make_requests(200, rate_per_s=40.0, seed=1)
    t = np.array([r["arrival_ms"] for r in reqs])
    assert (np.diff(t) > 0).all()
    assert 3500 < t[-1] < 6500                     # ~200 / 40 s
    assert {(r["width"], r["height"]) for r in reqs} == set(serve.RESOLUTIONS)
    assert all(60 <= r["frames"] <= 300 and 1.0 <= r["density"] <= 4.0 for r in reqs)


def test_footprint_and_refusal_without_device():
    p = serve.default_params(memory_budget=1 << 20, honor_arrivals=0)
    reqs = _small_requests()
    fps = [serve.footprint(p, r) for r in reqs]
    assert all(f > (1 << 20) for f in fps)
    big = dict(reqs[0], width=1280, height=720)
    assert serve.footprint(p, big) > serve.footprint(p, reqs[0])
    # nothing fits a 1 MiB budget: every request is refused, no GPU work runs
    results, windows, wall = serve.run(p, reqs)
    assert all(r["status"] == -5 for r in results) and len(windows) == 0


@pytest.mark.gpu
def test_served_mix_equals_alone_and_oracle(have_gpu):
    reqs = _small_requests()
    p = serve.default_params(max_active=3, memory_budget=8 << 30)
    mix, windows, wall = serve.run(p, reqs)
    assert all(r["status"] == 0 for r in mix)
    by_id = {r["stream_id"]: r for r in mix}
    # alone: one stream at a time through the same loop
    for r in reqs:
        alone, _, _ = serve.run(serve.default_params(max_active=1, honor_arrivals=0), [r])
        assert alone[0]["labels_hash"] == by_id[r["stream_id"]]["labels_hash"], r
    # and equal to the oracle's block-by-block restatement for two streams
    for r in reqs[:2]:
        rgb, fwd, bwd, _ = asv.synth_video(r["video_seed"], r["width"], r["height"], r["frames"], r["n_ellipses"],
                                           r["vmax"], r["noise_sigma"])
        seeds100 = int(r["density"] * r["width"] * r["height"] / 10.0)
        want, _ = ostream.run_video(rgb, fwd, bwd, 1, block_size=10, overlap=5, iters=5, seeds_per_100=seeds100)
        assert serve.labels_checksum(want) == by_id[r["stream_id"]]["labels_hash"]
    s = serve.summarize(mix, windows, wall)
    assert s["windows"] == len(windows) > 0 and s["p99_window_ms"] >= s["p50_window_ms"] > 0


@pytest.mark.gpu
def test_batched_windows_equal_unbatched(have_gpu):
    """The windows of concurrent streams run through shared batched launches
    (batch_windows = 1) give every stream the checksum it gets when each
    stream runs its own launches (the default), at several concurrencies."""
    reqs = _small_requests()
    want = {r["stream_id"]: r["labels_hash"] for r in serve.run(
        serve.default_params(max_active=3, batch_windows=0, honor_arrivals=0), reqs)[0]}
    for active in (1, 3, 6):
        res, windows, _ = serve.run(serve.default_params(max_active=active, batch_windows=1, honor_arrivals=0), reqs)
        assert all(r["status"] == 0 for r in res)
        assert {r["stream_id"]: r["labels_hash"] for r in res} == want, active


@pytest.mark.gpu
def test_failures_end_only_their_own_request(have_gpu):
    """A malformed request (width 1), a request cancelled after two windows,
    one cancelled by its flag before it starts, and one whose device
    allocation fails (every allocation of 1 MiB or more fails as
    hipErrorOutOfMemory, which only the 160 x 120 stream's flow buffers
    reach): each ends with its own status while every other stream completes
    with the checksum it has alone."""
    lib = asv._lib.load()
    reqs = _small_requests()
    alone = {}
    for r in reqs:
        res, _, _ = serve.run(serve.default_params(max_active=1, honor_arrivals=0), [r])
        assert res[0]["status"] == 0
        alone[r["stream_id"]] = res[0]["labels_hash"]
    flag = np.zeros(1, np.int32)
    mix = [dict(r) for r in reqs]
    mix.append(dict(reqs[0], stream_id=90, width=1))                        # malformed
    mix.append(dict(reqs[2], stream_id=91, cancel_after_windows=2))         # 47 frames: 8 windows
    flag[0] = 1
    mix.append(dict(reqs[1], stream_id=92, cancel=flag))                    # cancelled before it starts
    mix.append(dict(reqs[3], stream_id=93, width=160, height=120))          # allocation fails
    try:
        asv._lib.check(lib.asv_debug_fail_alloc(1 << 20))
        res, windows, _ = serve.run(serve.default_params(max_active=3, memory_budget=8 << 30), mix)
    finally:
        lib.asv_debug_fail_alloc(0)
    by_id = {r["stream_id"]: r for r in res}
    assert by_id[90]["status"] == -1
    assert by_id[91]["status"] == -6 and sum(w["stream_id"] == 91 for w in windows) == 2
    assert by_id[92]["status"] == -6 and not any(w["stream_id"] == 92 for w in windows)
    assert by_id[93]["status"] == -4
    for r in reqs:
        assert by_id[r["stream_id"]]["status"] == 0
        assert by_id[r["stream_id"]]["labels_hash"] == alone[r["stream_id"]], r


@pytest.mark.gpu
def test_admission_limits_concurrency(have_gpu):
    reqs = _small_requests()
    p0 = serve.default_params()
    fp = max(serve.footprint(p0, r) for r in reqs)
    # a budget for one stream at a time: windows of different streams never overlap
    p = serve.default_params(max_active=4, memory_budget=fp, honor_arrivals=0)
    res, w, _ = serve.run(p, reqs)
    assert all(r["status"] == 0 for r in res)
    spans = {}
    for row in w:
        a, b = spans.get(row["stream_id"], (np.inf, -np.inf))
        spans[row["stream_id"]] = (min(a, row["start_ms"]), max(b, row["end_ms"]))
    iv = sorted(spans.values())
    for (a0, a1), (b0, b1) in zip(iv, iv[1:]):
        assert b0 >= a1 - 1e-6


@pytest.mark.gpu
def test_reused_context_equals_fresh(have_gpu):
    """Streams of one frame size through one worker: each later stream adopts
    the previous one's resident context when its seed and frame capacity
    suffice (different lengths and densities exercise both outcomes)."""
    spec = [(33, 2.0), (21, 1.0), (40, 3.0), (26, 2.0), (17, 2.5)]
    reqs = [dict(stream_id=i, width=48, height=40, frames=D, density=d, n_ellipses=5, video_seed=200 + i,
                 vmax=2.0, noise_sigma=4.0, arrival_ms=0.0) for i, (D, d) in enumerate(spec)]
    seq, _, _ = serve.run(serve.default_params(max_active=1, honor_arrivals=0), reqs)
    assert all(r["status"] == 0 for r in seq)
    for i, r in enumerate(reqs):
        alone, _, _ = serve.run(serve.default_params(max_active=1, honor_arrivals=0), [r])
        assert alone[0]["labels_hash"] == seq[i]["labels_hash"], r


def _splitmix64(x):
    m = (1 << 64) - 1
    x = (x + 0x9E3779B97F4A7C15) & m
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & m
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & m
    return x ^ (x >> 31)


def test_labels_checksum_restatement():
    """The numpy checksum against scalar SplitMix64 (Steele, Lea and Flood,
    2014; first output for seed 0 is 0xE220A8397B1DCDAF)."""
    assert _splitmix64(0) == 0xE220A8397B1DCDAF
    lab = np.array([[3, -1, 0], [7, 7, 2 ** 31 - 1]], np.int32)
    want = sum(_splitmix64((i << 32) | (int(v) & 0xFFFFFFFF)) for i, v in enumerate(lab.reshape(-1))) % (1 << 64)
    assert serve.labels_checksum(lab) == want


@pytest.mark.gpu
def test_stream_checksum_matches_stitched_labels(have_gpu):
    """asv_stream_block_into window by window: the device checksum equals the
    checksum of the stitched labels, which equal the oracle's."""
    import ctypes
    W, H, D, K = 48, 40, 37, 120
    rgb, fwd, bwd, _ = asv.synth_video(321, W, H, D, 6, 2.0, 4.0)
    sp = api.stream_params(W, H, D, K, block_size=10, overlap=5, lloyd_iters=4)
    lib = asv._lib.load()
    s = ctypes.c_void_p()
    asv._lib.check(lib.asv_stream_create(ctypes.byref(sp), ctypes.byref(s)))
    nb = lib.asv_stream_block_count(s)
    iv = np.zeros(2 * nb, np.int32)
    lib.asv_stream_plan(ctypes.byref(sp), iv.ctypes.data, nb)
    out = np.full((D, H, W), -9, np.int32)
    pos = 0
    try:
        for b in range(nb):
            f0 = int(iv[2 * b])
            win = np.zeros((int(iv[2 * b + 1]) - f0, H, W), np.int32)
            done = ctypes.c_int32(0)
            rc = lib.asv_stream_block_into(s, b, rgb[f0:].ctypes.data, fwd[f0:].ctypes.data, bwd[f0:].ctypes.data,
                                           win.ctypes.data, ctypes.byref(done), None)
            assert rc == 0, lib.asv_stream_last_error()
            assert f0 == pos or b == 0
            out[pos:pos + done.value] = win[:done.value]
            pos += done.value
        ck = ctypes.c_uint64(0)
        asv._lib.check(lib.asv_stream_checksum(s, ctypes.byref(ck)))
    finally:
        lib.asv_stream_destroy(s)
    assert pos == D
    want, _ = ostream.run_video(rgb, fwd, bwd, K, 10, 5, 4)
    assert np.array_equal(out, want)
    assert ck.value == serve.labels_checksum(want)


@pytest.mark.gpu
@pytest.mark.slow
@pytest.mark.timeout(900)
def test_config5_resolutions_alone_mix_oracle(have_gpu):
    """Served streams at the workload's own resolutions: a 640x480 and a
    1280x720 stream at density 4 per 1000 pixels per frame (1280x720
    windows hold up to ~37 k seeds), a 320x240 stream at density 1, 20-25
    frames each, windows of 10 with stride 5, 5 iterations.  Each stream's
    stitched labels are the same alone and in the mix, and equal to the
    oracle's block-by-block restatement."""
    spec = [(640, 480, 20, 4.0), (1280, 720, 20, 4.0), (320, 240, 25, 1.0)]
    reqs = [dict(stream_id=i, width=W, height=H, frames=D, density=d, n_ellipses=16, video_seed=700 + i,
                 vmax=2.0, noise_sigma=4.0, arrival_ms=0.0) for i, (W, H, D, d) in enumerate(spec)]
    mix, windows, _ = serve.run(serve.default_params(max_active=3, honor_arrivals=0), reqs)
    assert all(r["status"] == 0 for r in mix)
    assert max(w["seeds"] for w in windows if w["stream_id"] == 1) > 16384
    by_id = {r["stream_id"]: r for r in mix}
    # the same mix with the windows' loops batched into shared launches
    mixb, _, _ = serve.run(serve.default_params(max_active=3, honor_arrivals=0, batch_windows=1), reqs)
    assert {r["stream_id"]: r["labels_hash"] for r in mixb} == {r["stream_id"]: r["labels_hash"] for r in mix}
    for r in reqs:
        alone, _, _ = serve.run(serve.default_params(max_active=1, honor_arrivals=0), [r])
        assert alone[0]["labels_hash"] == by_id[r["stream_id"]]["labels_hash"], r
        rgb, fwd, bwd, _ = asv.synth_video(r["video_seed"], r["width"], r["height"], r["frames"], r["n_ellipses"],
                                           r["vmax"], r["noise_sigma"])
        seeds100 = int(r["density"] * r["width"] * r["height"] / 10.0)
        want, _ = ostream.run_video(rgb, fwd, bwd, 1, block_size=10, overlap=5, iters=5, seeds_per_100=seeds100)
        assert serve.labels_checksum(want) == by_id[r["stream_id"]]["labels_hash"], r
