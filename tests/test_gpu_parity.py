"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar: bit-exact for everything (labels, integer sums, seed records, LAB
volume, metric field) -- the kernels and the oracle implement one arithmetic
specification.
"""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
import oracle as orc

pytestmark = pytest.mark.gpu


def _clip_run(rgb, fwd, bwd, K, iters, trace=True, **kw):
    D, H, W, _ = rgb.shape
    p = asv.default_params(W, H, D, K, lloyd_iters=iters, **kw)
    with asv.Clip(p) as c:
        if trace:
            c.enable_trace()
        c.upload(rgb, fwd, bwd)
        c.run()
        out = dict(labels=c.labels(), sums=c.sums(), A=c.A)
        if trace:
            out["trace"] = c.trace()
        out["lab"] = c.lab()
        out["metric"] = c.metric()
    return out


def _assert_run_equal(got, rgb, fwd, bwd, K, iters, compactness=8.0, tie_rule=0, angle_gate=1,
                      schedule=0, frames_prepared=0):
    want_labels, want_sums, want_trace, A = orc.run(rgb, fwd, bwd, K, iters, compactness, tie_rule,
                                                    angle_gate, schedule, frames_prepared, want_trace=True)
    assert got["A"] == A
    tr_diff = np.argwhere(np.any(got["trace"] != want_trace, axis=2))
    assert tr_diff.size == 0, f"seed records differ first at (iter, seed) {tr_diff[:5].tolist()}"
    bad = int((got["labels"] != want_labels).sum())
    assert bad == 0, f"{bad} label mismatches"
    assert np.array_equal(got["sums"][:, :7], want_sums)


def test_lab_exhaustive_all_rgb(have_gpu):
    """All 2^24 RGB triples, GPU LAB kernel vs oracle, bit-exact."""
    n = 1 << 24
    idx = np.arange(n, dtype=np.uint32)
    rgb = np.stack([(idx >> 16) & 255, (idx >> 8) & 255, idx & 255], axis=-1).astype(np.uint8)
    rgb = rgb.reshape(1, 4096, 4096, 3)
    fwd = np.zeros((1, 4096, 4096, 2), np.float32)
    p = asv.default_params(4096, 4096, 1, 1, lloyd_iters=1, frames_prepared=1)
    with asv.Clip(p) as c:
        c.upload(rgb, fwd, fwd)
        c.prepare()
        got = c.lab()
    want = orc.lab_volume(rgb)
    bad = int((got != want).sum())
    assert bad == 0, f"{bad} of 2^24 LAB values differ"


@pytest.mark.parametrize("seed,scale", [(0, 3.0), (1, 3.0), (2, 3.0), (3, 150.0)])
def test_metric_field_random_flows(have_gpu, seed, scale):
    """Per-voxel metric (LSQ + PSD flag) on flows with discontinuities,
    including rank-deficient corners and flows up to 150 px, bit-exact (the
    field is computed on request with the seed kernels' device function)."""
    rng = np.random.default_rng(seed)
    D, H, W = 6, 29, 37
    fwd = rng.uniform(-scale, scale, (D, H, W, 2)).astype(np.float32)
    bwd = rng.uniform(-scale, scale, (D, H, W, 2)).astype(np.float32)
    # blocks of exactly-integer flow and zero flow
    fwd[:, :10, :12] = np.round(fwd[:, :10, :12])
    bwd[:, 15:, 20:] = 0.0
    rgb = rng.integers(0, 256, (D, H, W, 3), dtype=np.uint8)
    p = asv.default_params(W, H, D, 20, lloyd_iters=1)
    with asv.Clip(p) as c:
        c.upload(rgb, fwd, bwd)
        c.prepare()
        m8 = c.metric()
    m6, psd = orc.metric_field(fwd, bwd)
    assert np.array_equal(m8[..., :6].view(np.uint32), m6.view(np.uint32))
    assert np.array_equal((m8[..., 6] != 0).astype(np.uint8), psd)
    assert (psd == 0).any(), "case must exercise the non-PSD fallbacks"


_RING1 = [(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]
_RING2 = [(x, -2) for x in range(-2, 3)] + [(-2, -1), (2, -1), (-2, 0), (2, 0), (-2, 1), (2, 1)] + \
    [(x, 2) for x in range(-2, 3)]


def _seed_metric_paths(psd, seeds):
    """How many seeds take each branch of seed_matrix (GPU-AS/Image.cpp:
    689-872) without the angle gate: own metric PSD, mean of the PSD 1-ring,
    of the PSD 2-ring, nearest PSD of the own metric."""
    D, H, W = psd.shape
    count = [0, 0, 0, 0]
    for x, y, z in seeds:
        if psd[z, y, x]:
            count[0] += 1
            continue
        for i, ring in enumerate((_RING1, _RING2)):
            if any(psd[z, y + dy, x + dx] for dx, dy in ring if 0 <= x + dx < W and 0 <= y + dy < H):
                count[1 + i] += 1
                break
        else:
            count[3] += 1
    return count


@pytest.mark.parametrize("gate", [0, 1])
def test_seed_metric_fallbacks_bit_exact(have_gpu, gate):
    """Random flows (8 px) make most metrics non-PSD, so the seeds of the
    first iteration take every branch of the seed metric -- own, 1-ring
    mean, 2-ring mean, nearest PSD -- and later iterations run them through
    the flow-angle gate when it is on.  Every iteration's seed records,
    labels and sums bit-exact."""
    rng = np.random.default_rng(23)
    D, H, W, K, iters = 8, 40, 48, 400, 3
    fwd = rng.uniform(-8.0, 8.0, (D, H, W, 2)).astype(np.float32)
    bwd = rng.uniform(-8.0, 8.0, (D, H, W, 2)).astype(np.float32)
    rgb = rng.integers(0, 256, (D, H, W, 3), dtype=np.uint8)
    _, psd = orc.metric_field(fwd, bwd)
    seeds, _ = orc.grid_seeds(W, H, D, K)
    paths = _seed_metric_paths(psd, seeds)
    assert all(n > 0 for n in paths), f"case must take every branch, got {paths}"
    got = _clip_run(rgb, fwd, bwd, K, iters, angle_gate=gate)
    _assert_run_equal(got, rgb, fwd, bwd, K, iters, angle_gate=gate)


@pytest.mark.parametrize("seed,shape,K,iters,kw", [
    (3, (10, 24, 32), 30, 4, {}),
    (4, (9, 31, 23), 25, 5, {"tie_rule": 1}),
    (5, (12, 20, 40), 40, 4, {"angle_gate": 0, "compactness": 6.0}),
    (6, (8, 26, 30), 20, 3, {"frames_prepared": 1}),
    (7, (11, 22, 35), 30, 3, {"schedule": 1}),
    # dense seeds: tile regions hold more than 63 ids (per-tap fallback path)
    (8, (12, 48, 48), 2000, 3, {}),
    # partial tiles on both axes, steps above the tiled range (generic kernel)
    (9, (6, 37, 70), 6, 3, {}),
    # sparse seeds: the second pass evaluates only the seeds' neighbourhoods
    # (seed sets clamped at every face), for both tie rules
    (13, (40, 80, 96), 20, 3, {}),
    (14, (30, 64, 64), 12, 3, {"tie_rule": 1}),
    (15, (16, 40, 200), 30, 3, {}),
    # halving schedule, first steps 12 and 6: the second pass's 13^3 tap
    # cells are read from global memory instead of LDS
    (16, (40, 200, 24), 20, 3, {"schedule": 1}),
])
def test_small_runs_bit_exact(have_gpu, seed, shape, K, iters, kw):
    D, H, W = shape
    if seed >= 13:
        A, s = asv.plan(asv.default_params(W, H, D, K))[:2]
        assert A * 729 * 3 <= W * H * D and 1 <= s // 2 <= 16, "case must take the sparse second pass"
    rgb, fwd, bwd, _ = asv.synth_video(seed, W, H, D, 6, 2.0, 4.0)
    got = _clip_run(rgb, fwd, bwd, K, iters, **kw)
    _assert_run_equal(got, rgb, fwd, bwd, K, iters, **{
        "compactness": kw.get("compactness", 8.0), "tie_rule": kw.get("tie_rule", 0),
        "angle_gate": kw.get("angle_gate", 1), "schedule": kw.get("schedule", 0),
        "frames_prepared": kw.get("frames_prepared", 0)})


def test_nan_flow_bit_exact(have_gpu):
    """x86 default NaNs (0xFFC00000) in the flow reach the least squares of
    the seeds and their rings (the pseudo-inverse then returns zero metrics)
    and the angle gate (a NaN cosine fails it): labels, sums and every seed
    record bit-exact, NaNs compared as NaNs."""
    W, H, D, K, iters = 48, 40, 10, 60, 3
    rgb, fwd, bwd, _ = asv.synth_video(17, W, H, D, 5, 2.0, 4.0)
    fwd, bwd = fwd.copy(), bwd.copy()
    nan = np.array([0xFFC00000], np.uint32).view(np.float32)[0]
    fwd[2:8, 8:30:3, 10:40:4] = nan
    bwd[3:9, 12:34:5, 6:30:3] = nan
    got = _clip_run(rgb, fwd, bwd, K, iters)
    want_labels, want_sums, want_trace, A = orc.run(rgb, fwd, bwd, K, iters, want_trace=True)
    assert got["A"] == A
    tr = got["trace"]
    assert np.array_equal(np.isnan(tr), np.isnan(want_trace))
    assert np.array_equal(np.nan_to_num(tr, nan=0.0), np.nan_to_num(want_trace, nan=0.0))
    assert int((got["labels"] != want_labels).sum()) == 0
    assert np.array_equal(got["sums"][:, :7], want_sums)


@pytest.mark.parametrize("seed,shape,K,iters", [(18, (10, 40, 48), 60, 3), (19, (24, 64, 80), 60, 2)])
def test_exact_clamp_path_bit_exact(have_gpu, seed, shape, K, iters):
    """The marching pass's exact per-tap path, which a run takes when a seed
    metric could make a quadratic form NaN (non-finite or beyond 2^60; flows
    cannot produce one, so the flag is forced here): the same labels, sums
    and records as the oracle."""
    D, H, W = shape
    rgb, fwd, bwd, _ = asv.synth_video(seed, W, H, D, 6, 2.0, 4.0)
    p = asv.default_params(W, H, D, K, lloyd_iters=iters)
    with asv.Clip(p) as c:
        c.enable_trace()
        c.force_exact_clamp(True)
        c.upload(rgb, fwd, bwd)
        c.run()
        got = dict(labels=c.labels(), sums=c.sums(), A=c.A, trace=c.trace())
    _assert_run_equal(got, rgb, fwd, bwd, K, iters)


def test_many_seeds_hash_table_path(have_gpu):
    """More seeds than the marching pass's byte table has slots (A > 16384):
    ids share slots (i mod 16384) and every lookup is checked."""
    rgb, fwd, bwd, _ = asv.synth_video(12, 640, 480, 48, 20, 3.0, 5.0)
    got = _clip_run(rgb, fwd, bwd, 18000, 2)
    assert got["A"] > 16384
    _assert_run_equal(got, rgb, fwd, bwd, 18000, 2)


def test_shared_table_slots_collide_everywhere(have_gpu):
    """32768 seeds on a 32 x 32 x 32 lattice, numbered so that x-neighbours
    are ids m and m + 16384, which share a slot of the marching pass's byte
    table: nearly every block meets a slot held by another id, evaluates
    that step per tap and rebuilds.  Labels and sums bit-exact against the
    oracle."""
    W, H, D = 128, 128, 64
    q = np.arange(32 * 32 * 32)
    i, j, k = q % 32, (q // 32) % 32, q // 1024
    lattice = np.stack([4 * i + 1, 4 * j + 1, 2 * k], axis=1).astype(np.int32)
    ids = np.where(q % 2 == 0, q // 2, q // 2 + 16384)
    xyz = np.empty_like(lattice)
    xyz[ids] = lattice
    rgb, fwd, bwd, _ = asv.synth_video(24, W, H, D, 10, 2.0, 4.0)
    s, iters = 3, 2
    p = asv.default_params(W, H, D, len(xyz), lloyd_iters=iters)
    with asv.Clip(p, seeds_xyz=xyz, grid_step=s) as c:
        c.upload(rgb, fwd, bwd)
        c.run()
        labels, sums = c.labels(), c.sums()
    want, wsums, _ = orc.run_seeded(rgb, fwd, bwd, xyz, s, iters)
    assert int((labels != want).sum()) == 0
    assert np.array_equal(sums[:, :7], wsums)


def test_config1_bit_exact(have_gpu):
    """160x120x30, seed 0, 12 ellipses, K=500 (A=468), 10 iterations."""
    rgb, fwd, bwd, _ = asv.synth_video(0, 160, 120, 30, 12, 2.0, 4.0)
    got = _clip_run(rgb, fwd, bwd, 500, 10)
    assert got["A"] == 468
    _assert_run_equal(got, rgb, fwd, bwd, 500, 10)


def test_config1_quality_matches_oracle(have_gpu):
    """Undersegmentation error and boundary recall against the generator's
    ground-truth regions, GPU labels vs the oracle's (the labels are
    bit-exact, so the metrics are equal; the bar is 0.5 %)."""
    from anisotropic_supervoxel_amd import api
    rgb, fwd, bwd, regions = asv.synth_video(0, 160, 120, 30, 12, 2.0, 4.0)
    got = _clip_run(rgb, fwd, bwd, 500, 10, trace=False)
    want, _, _, _ = orc.run(rgb, fwd, bwd, 500, 10)
    ue_g, br_g = api.eval_metrics(np.ascontiguousarray(got["labels"][:, ::-1, :]), regions)
    ue_w, br_w = api.eval_metrics(np.ascontiguousarray(want.reshape(got["labels"].shape)[:, ::-1, :]), regions)
    assert abs(ue_g - ue_w) <= 0.005 * max(ue_w, 1e-12) and abs(br_g - br_w) <= 0.005 * br_w
    assert 0.0 < br_g <= 1.0 and ue_g >= 0.0


def test_config2_bit_exact(have_gpu):
    """320x240x85, seed 1, 16 ellipses, K=2000 (A=1955), 10 iterations."""
    rgb, fwd, bwd, _ = asv.synth_video(1, 320, 240, 85, 16, 2.0, 4.0)
    got = _clip_run(rgb, fwd, bwd, 2000, 10)
    assert got["A"] == 1955
    _assert_run_equal(got, rgb, fwd, bwd, 2000, 10)


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_config3_all_iterations_bit_exact(have_gpu):
    """The benchmark clip at full size and length: 640x480x200, seed 2, 32
    ellipses, velocities up to 4 px/frame, noise 6, K=10000 (A=9720), all 10
    Lloyd iterations (first-iteration and next-iteration seed stages, the
    sparse first and second passes, all four marching steps, reduction and
    finalisation) against the oracle: labels, final sums and every
    iteration's seed records bit-exact.  The oracle takes about a minute on
    16 cores."""
    rgb, fwd, bwd, _ = asv.synth_video(2, 640, 480, 200, 32, 4.0, 6.0)
    p = asv.default_params(640, 480, 200, 10000, lloyd_iters=10)
    with asv.Clip(p) as c:
        c.enable_trace()
        c.upload(rgb, fwd, bwd)
        c.run()
        got = dict(labels=c.labels(), sums=c.sums(), A=c.A, trace=c.trace())
    assert got["A"] == 9720
    print("config 3: GPU run done, oracle running", flush=True)
    _assert_run_equal(got, rgb, fwd, bwd, 10000, 10)
