"""CPU tier: pins of the CPU oracle against closed forms, brute force and
independent numpy formulations (the reference ships no tests or golden
vectors for this path, and its GLSL/Eigen/OpenCV pipeline cannot run here).

Each test is chosen so that a plausible slip in the oracle (a dropped term,
a wrong sign or index, a transposed operand) fails it.
"""
import numpy as np
import pytest

import oracle as orc


# ---------------------------------------------------------------- a2: LAB --
def _lab_numpy(rgb):
    """Independent double-precision CIE conversion (sRGB D65, CIE eps/kappa)."""
    c = np.asarray(rgb, np.float64) / 255.0
    lin = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    M = np.array([[0.412453, 0.357580, 0.180423],
                  [0.212671, 0.715160, 0.072169],
                  [0.019334, 0.119193, 0.950227]])
    xyz = lin @ M.T
    r = xyz / np.array([0.950456, 1.0, 1.088754])
    f = np.where(r > 0.008856, np.cbrt(r), (903.3 * r + 16.0) / 116.0)
    L = 116.0 * f[..., 1] - 16.0
    a = 500.0 * (f[..., 0] - f[..., 1])
    b = 200.0 * (f[..., 1] - f[..., 2])
    out = np.stack([L * 2.55, a + 128.0, b + 128.0], -1)
    return np.clip(np.floor(out + 0.5), 0, 255).astype(np.int64)


def test_lab_anchors():
    assert orc.lab_of(255, 255, 255) == (255, 128, 128)
    assert orc.lab_of(0, 0, 0) == (0, 128, 128)
    # greys are neutral: a = b = 128
    for g in (1, 17, 64, 128, 200, 254):
        L, a, b = orc.lab_of(g, g, g)
        assert (a, b) == (128, 128), g


def test_lab_against_numpy_cie():
    rng = np.random.default_rng(0)
    rgb = rng.integers(0, 256, (20000, 3))
    rgb = np.concatenate([rgb, np.eye(3, dtype=np.int64) * 255, 255 - np.eye(3, dtype=np.int64) * 255])
    want = _lab_numpy(rgb)
    got = orc.lab_volume(rgb.astype(np.uint8)[None]).reshape(-1)
    got = np.stack([got & 255, (got >> 8) & 255, (got >> 16) & 255], -1).astype(np.int64)
    diff = np.abs(got - want)
    assert diff.max() <= 1          # float32 vs double only moves .5 boundaries
    assert (diff.sum(-1) == 0).mean() > 0.995
    # primaries exactly
    for i in range(3):
        assert np.array_equal(got[20000 + i], want[20000 + i])


def test_lin_table_against_formula():
    t = orc.lin_table()
    c = np.arange(256) / 255.0
    ref = np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)
    # u = c/255 and (u + 0.055)/1.055 are float32 steps (as in the shader)
    np.testing.assert_allclose(t, ref, rtol=1e-6, atol=1e-9)
    assert t[0] == 0.0 and t[255] == 1.0


# ------------------------------------------------------- a1: blur + flip --
def test_blur_flip_brute_force():
    rng = np.random.default_rng(1)
    D, H, W = 2, 7, 9
    raw = rng.integers(0, 256, (D, H, W, 3), dtype=np.uint8)
    got = orc.blur_flip(raw)
    k = np.array([1, 2, 1])
    want = np.zeros_like(raw)
    for t in range(D):
        for yi in range(H):
            for x in range(W):
                acc = np.zeros(3, np.int64)
                for j in (-1, 0, 1):
                    yy = abs(yi + j) if yi + j >= 0 else -(yi + j)
                    yy = 2 * H - 2 - yy if yy >= H else yy
                    for i in (-1, 0, 1):
                        xx = -(x + i) if x + i < 0 else x + i
                        xx = 2 * W - 2 - xx if xx >= W else xx
                        acc += k[j + 1] * k[i + 1] * raw[t, yy, xx].astype(np.int64)
                want[t, H - 1 - yi, x] = (acc + 8) >> 4
    assert np.array_equal(got, want)


# --------------------------------------------------------- a3: metric ----
def _flows(D, H, W):
    return np.zeros((D, H, W, 2), np.float32), np.zeros((D, H, W, 2), np.float32)


def test_metric_zero_flow_is_identity():
    fwd, bwd = _flows(4, 6, 7)
    for t in range(4):
        m, psd = orc.metric_at(fwd, bwd, 3, 2, t)
        np.testing.assert_allclose(m, [1, 0, 0, 1, 0, 1], atol=1e-6)
        assert psd == 1


@pytest.mark.parametrize("u,v", [(1, 0), (2, -1), (-3, 2), (0, 4)])
def test_metric_uniform_flow_closed_form(u, v):
    """F_fwd = (u, v), F_bwd = -(u, v) in GL orientation gives
    Q = (x - u t)^2 + (y - v t)^2 + t^2: m = (1, 0, -u, 1, -v, 1 + u^2 + v^2),
    on interior, forward-only (t = 0) and backward-only (t = D-1) frames."""
    D, H, W = 5, 9, 11
    fwd, bwd = _flows(D, H, W)
    fwd[..., 0], fwd[..., 1] = u, -v          # image rows: v points down
    bwd[..., 0], bwd[..., 1] = -u, v
    want = [1, 0, -u, 1, -v, 1 + u * u + v * v]
    for t in (0, 2, D - 1):
        m, psd = orc.metric_at(fwd, bwd, 5, 4, t)
        np.testing.assert_allclose(m, want, atol=1e-5)
        assert psd == 1
        M = np.array([[m[0], m[1], m[2]], [m[1], m[3], m[4]], [m[2], m[4], m[5]]])
        assert abs(np.linalg.det(M) - 1.0) < 1e-4


def test_metric_min_norm_on_underdetermined_corner():
    """Corner pixel of the first frame: 4 forward rows for 6 unknowns; the
    solution must be numpy's minimum-norm least squares."""
    rng = np.random.default_rng(3)
    D, H, W = 3, 5, 5
    fwd = rng.uniform(-2, 2, (D, H, W, 2)).astype(np.float32)
    bwd = rng.uniform(-2, 2, (D, H, W, 2)).astype(np.float32)
    x, y, t = 0, 0, 0
    rows, b = [], []
    eb = [3, 2, 3, 2, 1, 2, 3, 2, 3]
    for m in (-1, 0, 1):
        for n in (-1, 0, 1):
            nx, ny = x + n, y + m
            if not (0 <= nx < W and 0 <= ny < H):
                continue
            f = fwd[t, H - 1 - ny, nx]
            dx = n + np.floor(np.float32(f[0]) + np.float32(0.5))
            dy = m + np.floor(np.float32(-f[1]) + np.float32(0.5))
            rows.append([dx * dx, 2 * dx * dy, 2 * dx, dy * dy, 2 * dy, 1])
            b.append(eb[(m + 1) * 3 + n + 1])
    A = np.array(rows, np.float64)
    want = np.linalg.lstsq(A, np.array(b, np.float64), rcond=None)[0]
    got, _ = orc.metric_at(fwd, bwd, x, y, t)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)


def test_metric_full_rank_against_lstsq():
    rng = np.random.default_rng(4)
    D, H, W = 3, 6, 6
    fwd = rng.uniform(-3, 3, (D, H, W, 2)).astype(np.float32)
    bwd = rng.uniform(-3, 3, (D, H, W, 2)).astype(np.float32)
    eb = [3, 2, 3, 2, 1, 2, 3, 2, 3]
    for (x, y, t) in [(2, 3, 1), (4, 1, 1), (1, 1, 1)]:
        rows, b = [], []
        for m in (-1, 0, 1):
            for n in (-1, 0, 1):
                nx, ny = x + n, y + m
                for F, s in ((fwd, 2), (bwd, -2)):
                    f = F[t, H - 1 - ny, nx]
                    dx = n + np.floor(np.float32(f[0]) + np.float32(0.5))
                    dy = m + np.floor(np.float32(-f[1]) + np.float32(0.5))
                    rows.append([dx * dx, 2 * dx * dy, s * dx, dy * dy, s * dy, 1])
                    b.append(eb[(m + 1) * 3 + n + 1])
        want = np.linalg.lstsq(np.array(rows, np.float64), np.array(b, np.float64), rcond=None)[0]
        got, _ = orc.metric_at(fwd, bwd, x, y, t)
        np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-5)


def test_psd_and_nearest_psd_against_numpy():
    rng = np.random.default_rng(5)
    for _ in range(300):
        B = rng.normal(size=(3, 3))
        M = (B + B.T) / 2
        m = np.array([M[0, 0], M[0, 1], M[0, 2], M[1, 1], M[1, 2], M[2, 2]], np.float32)
        Mf = np.array([[m[0], m[1], m[2]], [m[1], m[3], m[4]], [m[2], m[4], m[5]]], np.float64)
        lam, Q = np.linalg.eigh(Mf)
        if abs(lam).min() < 1e-4:
            continue
        assert orc.psd3(m) == int(lam.min() >= 0)
        near = Q @ np.diag(np.maximum(lam, 0)) @ Q.T
        want = np.array([near[0, 0], near[0, 1], near[0, 2], near[1, 1], near[1, 2], near[2, 2]])
        np.testing.assert_allclose(orc.nearest_psd(m), want, atol=2e-5)


# ------------------------------------------------- a5 / a7: grid, schedule --
@pytest.mark.parametrize("W,H,D,K,A,s,steps", [
    (160, 120, 30, 500, 468, 10, [1, 5, 3, 2, 1, 2, 1]),
    (320, 240, 85, 2000, 1955, 14, [1, 7, 4, 2, 1, 2, 1]),
    (640, 480, 200, 10000, 9720, 18, [1, 9, 5, 3, 2, 1, 2, 1]),
    # V div K = 1000 is a perfect cube: double pow truncates to s = 9
    (320, 240, 10, 768, 768, 9, [1, 4, 2, 1, 2, 1]),
])
def test_grid_and_schedule(W, H, D, K, A, s, steps):
    xyz, s_got = orc.grid_seeds(W, H, D, K)
    assert len(xyz) == A and s_got == s
    st, mult = orc.schedule(0, s, W)
    assert st == steps
    assert mult == [1.0] * (len(steps) - 2) + [2.0, 2.0]
    assert len({tuple(p) for p in xyz}) == A              # distinct voxels
    assert (xyz >= 0).all() and (xyz < [W, H, D]).all()


def test_schedule_halving_mode():
    st, mult = orc.schedule(1, 10, 160)
    assert st == [80, 40, 20, 10, 5, 3, 2, 1, 1] and set(mult) == {1.0}


# ------------------------------------------------------------- a7: JFA --
def _jfa_brute(labels, lab, seeds, k, ratio):
    D, H, W = labels.shape
    out = np.full_like(labels, -1)
    for z in range(D):
        for y in range(H):
            for x in range(W):
                pk = int(lab[z, y, x])
                lv = np.array([pk & 255, (pk >> 8) & 255, (pk >> 16) & 255], np.float64)
                best = (1e8, 1 << 40)
                for zz in (z, min(max(z - k, 0), D - 1), min(max(z + k, 0), D - 1)):
                    for yy in (y - k, y, y + k):
                        for xx in (x - k, x, x + k):
                            i = labels[zz, min(max(yy, 0), H - 1), min(max(xx, 0), W - 1)]
                            if i < 0:
                                continue
                            sx, sy, sz, slab, m = seeds[i]
                            d = np.array([x - sx, y - sy, z - sz], np.float64)
                            M = np.array([[m[0], m[1], m[2]], [m[1], m[3], m[4]], [m[2], m[4], m[5]]])
                            q = max(0.0, d @ M @ d)
                            dist = ratio * q + float(((lv - np.array(slab)) ** 2).sum())
                            if (dist, i) < best:
                                best = (dist, i)
                out[z, y, x] = best[1] if best[0] < 1e8 else -1
    return out


def test_jfa_pass_brute_force():
    """Integer metrics and colours keep the float distance exact, so the
    oracle must equal the brute-force 27-tap argmin, lowest id on ties."""
    rng = np.random.default_rng(6)
    D, H, W = 5, 7, 9
    nseeds = 6
    seeds = []
    for i in range(nseeds):
        m = [int(rng.integers(1, 4)), int(rng.integers(-1, 2)), 0, int(rng.integers(1, 4)), 0, int(rng.integers(1, 4))]
        seeds.append((int(rng.integers(0, W)), int(rng.integers(0, H)), int(rng.integers(0, D)),
                      [int(v) for v in rng.integers(0, 40, 3)], m))
    lab = rng.integers(0, 40, (D, H, W, 3))
    labp = (lab[..., 0] | (lab[..., 1] << 8) | (lab[..., 2] << 16)).astype(np.uint32)
    labels = rng.integers(-1, nseeds, (D, H, W)).astype(np.int32)
    arr = orc.seeds_array(seeds)
    for k in (1, 2, 3):
        got = orc.jfa_pass(labels, labp, arr, k, 0.5, 0)
        want = _jfa_brute(labels, labp, seeds, k, 0.5)
        assert np.array_equal(got, want), k


def test_selection_key_min_as_double():
    """The marching pass takes the minimum of two selection keys
    (float bits of d << 32 | id) as one fp64 minimum.  For every distance a
    pass can produce (+0 .. +inf, the canonical NaN 0x7FC00000) and any id,
    the key read as a double is non-negative and not NaN, and the double
    minimum picks the same key as the u64 minimum, denormals and the initial
    key (1e8, 0) included."""
    rng = np.random.default_rng(12)
    n = 200000
    special = np.array([0x0, 0x1, 0x7FFFFF, 0x800000, 0x3F800000, 0x4CBEBC20, 0x4CBEBC21, 0x7F7FFFFF,
                        0x7F800000, 0x7FC00000], np.uint64)
    hi = np.concatenate([rng.integers(0, 0x7F800001, n, dtype=np.uint64),
                         special[rng.integers(0, len(special), n)]])
    lo = np.concatenate([rng.integers(0, 1 << 14, n, dtype=np.uint64),
                         rng.integers(0, 4, n, dtype=np.uint64)])
    keys = (hi << np.uint64(32)) | lo
    other = np.concatenate([keys[1:], keys[:1]])
    other[::7] = np.uint64(0x4CBEBC20) << np.uint64(32)
    a, b = keys.view(np.float64), other.view(np.float64)
    assert not np.isnan(a).any() and (a >= 0).all()
    got = np.minimum(a, b).view(np.uint64)
    assert np.array_equal(got, np.minimum(keys, other))


def test_jfa_single_seed_floods_volume():
    D, H, W = 6, 10, 12
    lab = np.zeros((D, H, W), np.uint32)
    labels = np.full((D, H, W), -1, np.int32)
    labels[2, 4, 5] = 0
    arr = orc.seeds_array([(5, 4, 2, (0, 0, 0), (1, 0, 0, 1, 0, 1))])
    for k in [1, 4, 2, 1, 2, 1]:
        labels = orc.jfa_pass(labels, lab, arr, k, 1.0, 0)
    assert (labels == 0).all()


# ------------------------------------------------- a8 / a9: reduce, finalize --
def test_reduce_exact_sums():
    rng = np.random.default_rng(7)
    D, H, W, A = 4, 5, 6, 3
    labels = rng.integers(-1, A, (D, H, W)).astype(np.int32)
    lab = rng.integers(0, 1 << 24, (D, H, W)).astype(np.uint32)
    sums = orc.reduce(labels, lab, A)
    z, y, x = np.meshgrid(np.arange(D), np.arange(H), np.arange(W), indexing="ij")
    for s in range(A + 1):
        sel = (labels == s) if s < A else (labels < 0)
        want = [(2 * x[sel] + 1).sum(), (2 * y[sel] + 1).sum(), (2 * z[sel] + 1).sum(), sel.sum(),
                (lab[sel] & 255).sum(), ((lab[sel] >> 8) & 255).sum(), ((lab[sel] >> 16) & 255).sum()]
        assert list(sums[s]) == [int(v) for v in want]
    assert sums[:, 3].sum() == D * H * W


def test_run_invariants():
    """Whole run on a small synthetic clip: counts add up to V, labels are
    in range, and the final sums equal the sums recomputed from the labels."""
    import anisotropic_supervoxel_amd as asv
    rgb, fwd, bwd, _ = asv.synth_video(11, 40, 32, 10, 5, 2.0, 4.0)
    labels, sums, trace, A = orc.run(rgb, fwd, bwd, 60, 4, want_trace=True)
    assert sums[:, 3].sum() == labels.size
    assert labels.min() >= -1 and labels.max() < A
    lab = orc.lab_volume(orc.blur_flip(rgb))
    assert np.array_equal(orc.reduce(labels, lab, A), sums)
    assert trace.shape == (4, A, 12)
    # iteration-1 records sit on the grid with the LAB of their voxel
    xyz, _ = orc.grid_seeds(40, 32, 10, 60)
    assert np.array_equal(trace[0, :, :3].astype(np.int64), xyz)
    pk = lab[xyz[:, 2], xyz[:, 1], xyz[:, 0]]
    assert np.array_equal(trace[0, :, 3].astype(np.int64), pk & 255)


def test_finalize_dead_seed_rule():
    """aver_exception.vs:25-70 on constructed sums: an alive seed takes
    floor(mean) and mean LAB; a dead seed takes floor(mid of the id
    neighbours' means) and the LAB image there; seed 0 reads slot W-1 and
    seed A-1 reads the unlabelled slot A."""
    W, H, D, A = 8, 6, 4, 5
    lab = (np.arange(W * H * D, dtype=np.uint32) % 200).reshape(D, H, W)
    sums = np.zeros((A + 1, 7), np.int64)

    def put(i, pts, labsum):
        pts = np.array(pts)
        sums[i, 0:3] = (2 * pts + 1).sum(0)
        sums[i, 3] = len(pts)
        sums[i, 4:7] = labsum
    put(1, [(2, 1, 1), (3, 1, 1), (3, 2, 2)], (30, 31, 32))
    put(3, [(6, 5, 3), (7, 4, 3)], (10, 11, 13))
    put(5, [(0, 0, 0)], (1, 1, 1))                # unlabelled accumulator (slot A)
    seeds = orc.finalize(sums, lab, W, H, D, A)
    f32 = np.float32

    def centre_mean(i):                            # mean of voxel centres x + .5
        return [sums[i, a] / (2.0 * sums[i, 3]) for a in range(3)]

    def mid(a, b):
        return tuple(int(np.floor(f32(0.5) * (f32(u) + f32(v)))) for u, v in zip(a, b))
    # alive seed 1: floor(mean of centres) = floor(19/6, 11/6, 11/6)
    assert seeds[1][:3] == (3, 1, 1)
    np.testing.assert_allclose(seeds[1][3], [10.0, 31 / 3, 32 / 3], rtol=1e-6)
    # dead seed 2: midpoint of the means of id neighbours 1 and 3
    want2 = mid(centre_mean(1), centre_mean(3))
    assert seeds[2][:3] == want2
    assert seeds[2][3][0] == lab[want2[2], want2[1], want2[0]] & 255
    # dead seed 0: slot -1 reads the clamped texel W-1 = 7 > A, cleared to 0
    assert seeds[0][:3] == mid([0, 0, 0], centre_mean(1))
    # dead seed 4 (= A-1): neighbours 3 and the unlabelled slot A
    assert seeds[4][:3] == mid(centre_mean(3), centre_mean(5))
