"""Pins of the oracle functions the GPU-vs-oracle tests cannot check on their
own (a shared misreading of the reference would pass those): the seed-metric
selection with its ring fallbacks and flow-angle gate, the reference's
scan-order tie rule, the clamp of the quadratic form (negative, -0, +-inf,
NaN) and the scatter collision.  Each case is built so that a plausible
mistake -- a ring read in another order, a transposed offset table, a gate
threshold off by one float, a tie kept by <= instead of <, a NaN accepted --
changes the expected value.
"""
import numpy as np
import pytest

import anisotropic_supervoxel_amd as asv
import oracle as orc

f32 = np.float32

# ring tables as the reference lists them (GPU-AS/Image.cpp:701-702, :766-767)
RING1 = list(zip([-1, 0, 1, -1, 1, -1, 0, 1], [-1, -1, -1, 0, 0, 1, 1, 1]))
RING2 = list(zip([-2, -1, 0, 1, 2, -2, 2, -2, 2, -2, 2, -2, -1, 0, 1, 2],
                 [-2, -2, -2, -2, -2, -1, -1, 0, 0, 1, 1, 2, 2, 2, 2, 2]))


# --------------------------------------------------------- a4: angle gate --
PI_REF = 3.141592653        # GPU-AS/Image.cpp:2


def acosf_cr(c):
    """float acos, correctly rounded: the double acos rounded to float."""
    return f32(np.arccos(np.float64(c)))


def reference_gate(su, sv, nu, nv):
    """GPU-AS/Image.cpp:1406-1409: a float numerator, norms from pow(float, 2)
    (double) and sqrt, the quotient stored in a float, float acos, the
    product with 180 / PI in double stored in a float, then theta < 90."""
    su, sv, nu, nv = f32(su), f32(sv), f32(nu), f32(nv)
    num = f32(f32(su * nu) + f32(sv * nv))
    den = np.sqrt(np.float64(su) ** 2 + np.float64(sv) ** 2) * np.sqrt(np.float64(nu) ** 2 + np.float64(nv) ** 2)
    with np.errstate(invalid="ignore", divide="ignore"):
        cs = f32(np.float64(num) / den)
        theta = f32(np.float64(acosf_cr(cs)) * (180.0 / PI_REF))
    return bool(theta < 90)


def test_angle_gate_every_float_cosine_near_the_threshold():
    """Seed flow (1, 0), neighbour flow (c, 1): the cosine runs over every
    float in [1.49e-8, 1.68e-8] around the threshold at 1.589e-8, where the
    correctly rounded acos steps from the float above pi/2 to the one below."""
    bits = np.arange(0x32800000, 0x32900000, 7, dtype=np.uint32)
    cs = bits.view(np.float32)
    got = np.array([orc.angle_gate(1.0, 0.0, c, 1.0) for c in cs])
    want = np.array([reference_gate(1.0, 0.0, c, 1.0) for c in cs])
    assert np.array_equal(got, want)
    assert got.any() and not got.all()                   # the threshold lies inside the sweep


def test_angle_gate_random_and_special_flows():
    rng = np.random.default_rng(3)
    n = 20000
    v = rng.normal(size=(n, 4)) * np.exp(rng.uniform(-8, 6, (n, 1)))
    v[: n // 10, 2] = v[: n // 10, 0] * 1e-7            # nearly perpendicular
    v[n // 10: n // 5, 3] = 0.0
    v = v.astype(np.float32)
    special = [(1, 0, 0, 1), (1, 0, -1e-30, 1), (0, 0, 1, 1), (1, 1, 0, 0), (0, 0, 0, 0), (2, 0, 3, 0),
               (1, 0, -1, 0), (1, 1, 1, -0.9999999), (3e19, 0, 1, 0), (1e-20, 1e-20, 1e-20, -1e-20)]
    for su, sv, nu, nv in list(map(tuple, v)) + special:
        assert orc.angle_gate(su, sv, nu, nv) == reference_gate(su, sv, nu, nv), (su, sv, nu, nv)


# ----------------------------------------------------- a4: seed metric --
class Field:
    """A metric field of W x H x D voxels with chosen values and PSD flags
    (the selection reads only these and, for the gate, the forward flow)."""

    def __init__(self, W=9, H=8, D=3, seed=0):
        rng = np.random.default_rng(seed)
        self.W, self.H, self.D = W, H, D
        self.m6 = rng.uniform(-4, 4, (D, H, W, 6)).astype(np.float32)
        self.psd = np.zeros((D, H, W), np.uint8)
        self.fwd = np.zeros((D, H, W, 2), np.float32)    # image rows: row H-1-y holds GL row y

    def set_flow(self, x, y, z, u, v_gl):
        self.fwd[z, self.H - 1 - y, x] = (u, -v_gl)

    def select(self, x, y, z, gate=0):
        return orc.seed_metric(self.m6, self.psd, self.fwd, x, y, z, gate)


def float_mean(vals):
    s = np.zeros(6, np.float32)
    for v in vals:
        s = (s + v).astype(np.float32)
    return (s / f32(len(vals))).astype(np.float32)


def test_seed_metric_own_psd():
    F = Field()
    F.psd[1, 3, 4] = 1
    F.psd[1, 3, 5] = 1                                   # a PSD neighbour must not matter
    assert np.array_equal(F.select(4, 3, 1), F.m6[1, 3, 4])


@pytest.mark.parametrize("t", range(8))
def test_seed_metric_single_1ring_neighbour(t):
    """Only the t-th 1-ring neighbour is PSD: its metric, exactly.  The
    offsets are (x, y) pairs; a transposed table would read another voxel."""
    F = Field(seed=t)
    dx, dy = RING1[t]
    F.psd[2, 4 + dy, 4 + dx] = 1
    F.psd[2, 4 + 2, 4 + 2] = 1                            # a 2-ring neighbour is ignored
    assert np.array_equal(F.select(4, 4, 2), F.m6[2, 4 + dy, 4 + dx])


def test_seed_metric_1ring_mean_in_ring_order():
    """Three PSD neighbours whose float sum depends on the order (1e8 + -1e8
    + 1 = 1, but 1 + -1e8 + 1e8 = 0): the mean must follow the ring order."""
    F = Field()
    x, y, z = 4, 4, 1
    picks = [0, 4, 6]
    vals = [np.array([1e8, 3, 1, 1e-3, 5, 2], np.float32), np.array([-1e8, 1, 2, 1e7, 5, 3], np.float32),
            np.array([1, 1e-8, 3, -1e7, 5, 4], np.float32)]
    for t, v in zip(picks, vals):
        dx, dy = RING1[t]
        F.m6[z, y + dy, x + dx] = v
        F.psd[z, y + dy, x + dx] = 1
    want = float_mean(vals)
    assert not np.array_equal(want, float_mean(vals[::-1])), "case must be order-sensitive"
    assert np.array_equal(F.select(x, y, z), want)


def test_seed_metric_empty_1ring_takes_2ring_mean_in_order():
    F = Field(seed=5)
    x, y, z = 4, 4, 0
    picks = [1, 7, 8, 15]
    vals = [np.array([3e7, 1, 1, 1, 1, 1], np.float32), np.array([1, 2e7, 1, 1, 1, 1], np.float32),
            np.array([-3e7, -2e7, 7, 1, 1, 1], np.float32), np.array([1, 1, 1, 0.1, 0.2, 0.3], np.float32)]
    for t, v in zip(picks, vals):
        dx, dy = RING2[t]
        F.m6[z, y + dy, x + dx] = v
        F.psd[z, y + dy, x + dx] = 1
    want = float_mean(vals)
    assert not np.array_equal(want, float_mean([vals[i] for i in (2, 0, 1, 3)]))
    assert np.array_equal(F.select(x, y, z), want)


def test_seed_metric_out_of_volume_neighbours_skipped():
    """Seed at the corner (0, 0): only in-volume neighbours count; a PSD
    voxel that an unclamped offset would wrap to must not be read."""
    F = Field(seed=7)
    F.psd[1, 1, 1] = 1                                   # ring-1 offset (+1, +1)
    F.psd[1, 0, F.W - 1] = 1                             # where x = -1 would wrap in a flat index
    F.psd[1, F.H - 1, 0] = 1
    assert np.array_equal(F.select(0, 0, 1), F.m6[1, 1, 1])


def test_seed_metric_both_rings_empty_nearest_psd():
    F = Field(seed=9)
    x, y, z = 4, 3, 2
    got = F.select(x, y, z)
    own = F.m6[z, y, x]
    assert np.array_equal(got, orc.nearest_psd(own))
    M = np.array([[own[0], own[1], own[2]], [own[1], own[3], own[4]], [own[2], own[4], own[5]]], np.float64)
    w, Q = np.linalg.eigh(M)
    R = (Q * np.maximum(w, 0.0)) @ Q.T
    np.testing.assert_allclose(got, [R[0, 0], R[0, 1], R[0, 2], R[1, 1], R[1, 2], R[2, 2]], atol=1e-5)


def test_seed_metric_gate_keeps_neighbours_under_90_degrees():
    """Gate on: the seed moves along +x; PSD neighbours moving at 89.9 deg
    count, at 90.0 (cos 0) and 90.1 deg or with zero flow do not.  Gate off:
    all count."""
    F = Field(seed=11)
    x, y, z = 4, 4, 1
    F.set_flow(x, y, z, 2.0, 0.0)
    angles = {0: 89.9, 1: 90.0, 2: 90.1, 3: None, 4: 45.0, 5: 135.0}
    kept = []
    for t, a in angles.items():
        dx, dy = RING1[t]
        F.psd[z, y + dy, x + dx] = 1
        if a is not None:
            r = np.deg2rad(a)
            u, v = (0.0, 3.0) if a == 90.0 else (3.0 * np.cos(r), 3.0 * np.sin(r))
            F.set_flow(x + dx, y + dy, z, u, v)
        su, sv = 2.0, 0.0
        nu, nv = F.fwd[z, F.H - 1 - (y + dy), x + dx]
        if reference_gate(su, -0.0, nu, -nv):
            kept.append(t)
    assert kept == [0, 4], kept
    want = float_mean([F.m6[z, y + RING1[t][1], x + RING1[t][0]] for t in kept])
    assert np.array_equal(F.select(x, y, z, gate=1), want)
    want_off = float_mean([F.m6[z, y + RING1[t][1], x + RING1[t][0]] for t in angles])
    assert np.array_equal(F.select(x, y, z, gate=0), want_off)


def test_seed_metric_gate_on_last_frame_rejects_every_neighbour():
    """flow_forward of the last frame is a 4 x 4 dummy in the reference
    (GPU-AS/Image.cpp:666-667); read as zero flow, every cosine is NaN and
    the gate rejects all neighbours: nearest PSD of the own metric."""
    F = Field(seed=13)
    x, y, z = 4, 4, F.D - 1
    for t in range(8):
        dx, dy = RING1[t]
        F.psd[z, y + dy, x + dx] = 1
        F.set_flow(x + dx, y + dy, z, 1.0, 0.0)
    F.set_flow(x, y, z, 1.0, 0.0)
    assert np.array_equal(F.select(x, y, z, gate=1), orc.nearest_psd(F.m6[z, y, x]))
    assert np.array_equal(F.select(x, y, z, gate=0), float_mean([F.m6[z, y + dy, x + dx] for dx, dy in RING1]))


# ------------------------------------------- a7: scan order, clamp of q --
SLAB_OFFSETS = [(-1, -1), (0, -1), (1, -1), (-1, 0), (0, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]   # flood.gs:36-46


def brute_pass(labels, lab, seeds, k, ratio, tie_rule):
    """flood.gs:26-182 in Python: slabs z, z-k, z+k (:73, :105, :139), the
    nine offsets of :36-46 in order, clamped taps, skip id < 0 (:77),
    distance in float64 (the data make it exact), a NaN quadratic form loses
    (:90-95), a negative one is 0; strict < from 1e8 (:63, :95), or the
    lowest id among equal distances."""
    D, H, W = labels.shape
    out = np.full_like(labels, -1)
    for z in range(D):
        for y in range(H):
            for x in range(W):
                pk = int(lab[z, y, x])
                lv = np.array([pk & 255, (pk >> 8) & 255, (pk >> 16) & 255], np.float64)
                best, best_id = 1e8, -1
                for zz in (z, min(max(z - k, 0), D - 1), min(max(z + k, 0), D - 1)):
                    for ox, oy in SLAB_OFFSETS:
                        i = int(labels[zz, min(max(y + oy * k, 0), H - 1), min(max(x + ox * k, 0), W - 1)])
                        if i < 0:
                            continue
                        sx, sy, sz, slab, m = seeds[i]
                        dv = np.array([x - sx, y - sy, z - sz], np.float64)
                        M = np.array([[m[0], m[1], m[2]], [m[1], m[3], m[4]], [m[2], m[4], m[5]]], np.float64)
                        with np.errstate(invalid="ignore", over="ignore"):
                            q = float(dv @ M @ dv)
                        if np.isnan(q):
                            continue
                        q = max(q, 0.0)
                        dist = ratio * q + float(((lv - np.array(slab)) ** 2).sum())
                        if dist < best or (tie_rule == 0 and dist == best and i < best_id):
                            best, best_id = dist, i
                out[z, y, x] = best_id
    return out


def test_scan_order_tie_rule_brute_force():
    """Identity metrics, one colour everywhere and seeds on a lattice make
    most voxels equidistant from several seeds: the scan-order rule keeps the
    first tap in the reference's order, the lowest-id rule the lowest id,
    and the two differ on this volume."""
    D, H, W = 5, 9, 11
    rng = np.random.default_rng(21)
    seeds = []
    for i, (x, y, z) in enumerate([(1, 1, 1), (3, 1, 1), (1, 3, 1), (3, 3, 3), (7, 5, 2), (9, 7, 0), (5, 5, 4),
                                   (9, 1, 3)]):
        seeds.append((x, y, z, (10, 10, 10), (1, 0, 0, 1, 0, 1)))
    lab = np.full((D, H, W), 10 | (10 << 8) | (10 << 16), np.uint32)
    labels = rng.integers(-1, len(seeds), (D, H, W)).astype(np.int32)
    arr = orc.seeds_array(seeds)
    differ = 0
    for k in (1, 2, 3):
        for tie in (0, 1):
            got = orc.jfa_pass(labels, lab, arr, k, 1.0, tie)
            assert np.array_equal(got, brute_pass(labels, lab, seeds, k, 1.0, tie)), (k, tie)
        differ += int((orc.jfa_pass(labels, lab, arr, k, 1.0, 0) != orc.jfa_pass(labels, lab, arr, k, 1.0, 1)).sum())
    assert differ > 0, "case must separate the two tie rules"


def test_quadratic_form_clamp_and_nan():
    """A metric with a negative eigenvalue gives negative quadratic forms
    (clamped to 0, so the colour term alone decides), an infinite entry
    gives +inf (rejected), -inf (clamped to 0, accepted) or inf - inf = NaN,
    and a NaN entry makes every form NaN: such candidates lose whatever the
    sign of the NaN."""
    D, H, W = 4, 6, 7
    nan, inf = float("nan"), float("inf")
    seeds = [(1, 1, 1, (20, 20, 20), (1, 0, 0, -3, 0, 1)),      # indefinite
             (5, 4, 2, (21, 20, 20), (1, 0, 0, 1, 0, nan)),     # NaN everywhere
             (3, 2, 0, (25, 20, 20), (1, 0, 0, 1, 0, -inf)),    # -inf off its frame, 0 * -inf = NaN on it
             (6, 0, 3, (30, 30, 30), (1, 0, 0, 1, 0, 1)),
             (2, 5, 3, (20, 20, 20), (1, 0, 0, 1, 0, inf))]     # +inf off its frame, NaN on it
    rng = np.random.default_rng(4)
    lab = rng.integers(15, 35, (D, H, W, 3))
    labp = (lab[..., 0] | (lab[..., 1] << 8) | (lab[..., 2] << 16)).astype(np.uint32)
    labels = rng.integers(-1, len(seeds), (D, H, W)).astype(np.int32)
    arr = orc.seeds_array(seeds)
    for k in (1, 2):
        for tie in (0, 1):
            got = orc.jfa_pass(labels, labp, arr, k, 0.5, tie)
            assert np.array_equal(got, brute_pass(labels, labp, seeds, k, 0.5, tie)), (k, tie)
            assert not (got == 1).any() and not (got == 4).any()   # NaN and +inf forms never win
    # the -inf seed wins off its own frame through a clamped form, never on it
    got = orc.jfa_pass(labels, labp, arr, 1, 0.5, 0)
    assert (got[1:] == 2).any() and not (got[0] == 2).any() and (got == 0).any()


# ------------------------------------------------------ a6/a9: collision --
def test_two_seeds_on_one_voxel():
    """Seeds 3 and 4 placed on the same voxel: the scatter keeps the higher
    id (draw_next_seeds, later primitive wins), so label 3 never appears
    after the first flood, its count is 0, and its next position is the
    neighbour-slot midpoint of seeds 2 and 4 (aver_exception.vs:25-62) with
    the LAB of that voxel."""
    W, H, D = 30, 24, 8
    rgb, fwd, bwd, _ = asv.synth_video(41, W, H, D, 4, 1.5, 3.0)
    xyz, s = orc.grid_seeds(W, H, D, 40)
    xyz = xyz.copy()
    xyz[3] = xyz[4]
    labels1, sums1, _ = orc.run_seeded(rgb, fwd, bwd, xyz, s, 1)
    assert not (labels1 == 3).any() and (labels1 == 4).any()
    assert sums1[3, 3] == 0 and sums1[2, 3] > 0 and sums1[4, 3] > 0
    _, _, trace = orc.run_seeded(rgb, fwd, bwd, xyz, s, 2, want_trace=True)
    assert np.array_equal(trace[0, 3, :3], trace[0, 4, :3])

    def mean(j):
        return [f32(sums1[j, a] / (2.0 * sums1[j, 3])) for a in range(3)]
    p = [int(np.floor(f32(0.5) * f32(a + b))) for a, b in zip(mean(2), mean(4))]
    assert [int(v) for v in trace[1, 3, :3]] == p
    lab = orc.lab_volume(orc.blur_flip(rgb))
    pk = int(lab[p[2], p[1], p[0]])
    assert [int(v) for v in trace[1, 3, 3:6]] == [pk & 255, (pk >> 8) & 255, (pk >> 16) & 255]
