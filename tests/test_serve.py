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
        assert serve.fnv1a_labels(want) == by_id[r["stream_id"]]["labels_hash"]
    s = serve.summarize(mix, windows, wall)
    assert s["windows"] == len(windows) > 0 and s["p99_window_ms"] >= s["p50_window_ms"] > 0


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
