"""The oracle of the pose tracker's per-frame stage (oracle/pf_pose_ref.c),
pinned against library routines and closed forms:

* det_log / det_sincos_turn against the C library's log, sin and cos, and the
  Box-Muller pairs they make against the textbook transform and a normal
  distribution (Kolmogorov-Smirnov);
* the blur taps against the Gaussian formula, the blur against
  scipy.ndimage.gaussian_filter (mode 'mirror' = BORDER_REFLECT_101) and
  against the exact response to an impulse;
* proposals against their definition from the Philox words;
* association weights and measurements against hand computations of
  pfPose.cpp:244-320;
* the proposals' bound-scaled resampling against an exact integer
  restatement (bisection over big-integer prefix sums), against the
  maximum-relative resampler where the two scales coincide, and its
  all-zero fallback against the uniform closed form;
* get3Dpose against re-projection through K [R | t] and rpy against
  scipy's Rotation.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle


def _fma(a, b, c):
    """a * b + c rounded once (exact rational arithmetic)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def test_det_log_matches_libm():
    rng = np.random.default_rng(1)
    xs = np.concatenate([np.exp(rng.uniform(-23, 0, 4000)), rng.uniform(1e-6, 1, 2000),
                         [1.0, 0.5, 0.25, 2.0 ** -32, 1 - 2.0 ** -53, 0.7071067811865476, 0.7071067811865475]])
    for x in xs:
        ref = math.log(x)
        assert abs(oracle.det_log(x) - ref) <= 4.5e-16 * max(1.0, abs(ref)), x
    assert oracle.det_log(1.0) == 0.0


def test_det_sincos_matches_libm():
    rng = np.random.default_rng(2)
    ws = np.concatenate([rng.integers(0, 2 ** 32, 5000, dtype=np.uint64),
                         [0, 1 << 29, 1 << 30, 3 << 30, (1 << 32) - 1, (1 << 29) + 1, (1 << 30) - 1]])
    for w in ws:
        a = 2.0 * math.pi * float(w) / 2.0 ** 32
        s, c = oracle.det_sincos_turn(int(w))
        assert abs(s - math.sin(a)) <= 1.2e-15 and abs(c - math.cos(a)) <= 1.2e-15, w
    assert oracle.det_sincos_turn(0) == (0.0, 1.0)


def test_normal_pairs_are_box_muller_and_normal():
    from scipy import stats
    words = oracle.philox_u32(7, 16, 0, 2 * 40000).reshape(-1, 2)
    z = np.array([oracle.normal_pair(int(a), int(b)) for a, b in words])
    u1 = (words[:, 0].astype(np.float64) + 1.0) * 2.0 ** -32
    th = 2.0 * np.pi * words[:, 1].astype(np.float64) * 2.0 ** -32
    r = np.sqrt(-2.0 * np.log(u1))
    assert np.allclose(z[:, 0], r * np.cos(th), rtol=0, atol=1e-13)
    assert np.allclose(z[:, 1], r * np.sin(th), rtol=0, atol=1e-13)
    for k in range(2):
        assert stats.kstest(z[:, k], "norm").pvalue > 1e-3
    assert abs(np.corrcoef(z[:, 0], z[:, 1])[0, 1]) < 0.02


def test_gauss_taps():
    t = oracle.gauss_taps(15, 3.0)
    x = np.arange(-7, 8)
    g = np.exp(-x ** 2 / 18.0)
    assert np.allclose(t, g / g.sum(), rtol=2e-7, atol=0)
    assert np.array_equal(t, t[::-1])
    assert abs(float(t.astype(np.float64).sum()) - 1.0) < 1e-6


def test_blur_matches_scipy_mirror():
    from scipy import ndimage
    rng = np.random.default_rng(3)
    for rows, cols in [(480, 640), (37, 53), (8, 8), (9, 200)]:
        img = rng.integers(0, 256, (rows, cols), dtype=np.uint8)
        img[: rows // 3] = 0                                       # flat regions and edges too
        out = oracle.blur_u8(img)
        ref = ndimage.gaussian_filter(img.astype(np.float64), sigma=3.0, mode="mirror", truncate=7.0 / 3.0)
        err = np.abs(out.astype(np.float64) - ref)
        assert err.max() <= 0.5 + 1e-3, (rows, cols, err.max())
        near_half = np.abs(ref - np.floor(ref) - 0.5) < 1e-3
        assert np.array_equal(out[~near_half], np.rint(ref[~near_half]).astype(np.uint8))


def test_blur_impulse_and_constant():
    t = oracle.gauss_taps(15, 3.0).astype(np.float64)
    img = np.zeros((40, 50), np.uint8)
    img[20, 25] = 255
    out = oracle.blur_u8(img)
    ref = 255.0 * np.outer(t, t)
    assert np.array_equal(out[13:28, 18:33], np.rint(ref).astype(np.uint8))
    assert out.sum() == out[13:28, 18:33].sum()
    assert np.array_equal(oracle.blur_u8(np.full((30, 30), 77, np.uint8)), np.full((30, 30), 77, np.uint8))
    # reflect-101 at the corner: pixel (0, 0) mirrors without repeating the edge
    img = np.zeros((30, 30), np.uint8)
    img[0, 0] = 200
    out = oracle.blur_u8(img)
    assert out[0, 0] == np.rint(200.0 * t[7] * t[7]) and out[0, 1] == np.rint(200.0 * t[7] * t[6])


def test_blur_rejects_small_images():
    with pytest.raises(ValueError):
        oracle.blur_u8(np.zeros((7, 40), np.uint8))


ROI = np.array([300, 120, 40, 48], np.int32)


def test_uniform_proposals_fill_the_box():
    n = 20000
    p = oracle.tracker_proposals(5, 0, 0, n, True, ROI, 480, 640, [0.0, 0.0], 0.8)
    w = oracle.philox_u32(5, 16, 0, 2 * n).reshape(n, 2).astype(np.float64) * 2.0 ** -32
    xmin, xmax = max(300 - 160, 0), min(300 + 200, 640)
    ymin, ymax = min(120 + 48, 480), min(120 + 7 * 48, 480)
    for c in range(0, n, 37):
        assert p[0, c] == _fma(xmax - xmin, w[c, 0], xmin) and p[1, c] == _fma(ymax - ymin, w[c, 1], ymin)
    assert p[0].min() >= xmin and p[0].max() < xmax and p[1].min() >= ymin and p[1].max() < ymax
    # the box is clamped to the image
    roi = np.array([10, 450, 30, 40], np.int32)
    p = oracle.tracker_proposals(5, 1, 3, 1000, True, roi, 480, 640, [0.0, 0.0], 0.8)
    assert p[0].min() >= 0 and p[0].max() < 10 + 150 and np.all(p[1] == 480.0)


def test_gaussian_proposals_follow_their_words():
    n, g, arm, y = 4000, 17, 1, np.array([211.5, 301.25])
    p = oracle.tracker_proposals(9, arm, g, n, False, ROI, 480, 640, y, 0.8)
    words = oracle.philox_u32(9, 16 + arm, 2 * g * n, 2 * n).reshape(n, 2)
    sd = 0.8 * 40.0
    for c in range(0, n, 97):
        z0, z1 = oracle.normal_pair(int(words[c, 0]), int(words[c, 1]))
        assert p[0, c] == _fma(sd, z0, y[0]) and p[1, c] == _fma(sd, z1, y[1])
    assert abs(p[0].mean() - y[0]) < 4 * sd / math.sqrt(n) and abs(p[1].std() / sd - 1) < 0.05


def test_association_weights_by_hand():
    img = np.zeros((480, 640), np.uint8)
    img[100, 200] = 51
    img[101, 201] = 0
    props = np.array([[200.7, 201.2, -1.0, 200.0, 0.0, 639.999, 640.0],
                      [100.3, 101.9, 100.0, 480.0, 5.0, 479.5, 5.0]])
    p_own = np.array([1e-3, 2e-3, 3e-3, 4e-3, 5e-3, 6e-3, 7e-3])
    p_oth = np.array([4e-4, 0.0, 0.0, 0.0, 0.0, 1e-5, 0.0])
    img[479, 639] = 255
    w = oracle.tracker_weights(props, p_own, p_oth, img, 0.05, 1e-4)
    z0 = 1e-3 * 0.05 + 4e-4 * 0.05 + 1e-4 * 0.9
    z5 = 6e-3 * 0.05 + 1e-5 * 0.05 + 1e-4 * 0.9
    assert w[0] == pytest.approx((51 / 255.0) / z0, rel=1e-15)
    assert w[1] == 0.0                      # L = 0
    assert w[2] == 0.0 and w[3] == 0.0      # outside (x <= 0; y >= rows)
    assert w[4] == 0.0                      # x = 0 is outside (strict inequality)
    assert w[5] == pytest.approx(1.0 / z5, rel=1e-15)
    assert w[6] == 0.0                      # x = cols


def _scaled_brute(w, N, U, scale):
    """a_j = min{i : C_i N 2^32 > (j 2^32 + U) C_L} over C = prefix sums of
    q_i = floor(2^32 w_i scale) (the FP64 product, then exact integers)."""
    import bisect
    x = np.asarray(w, np.float64) * scale
    q = [int(math.floor(Fraction(float(v)) * 2 ** 32)) if v > 0 else 0 for v in x]
    C, acc = [], 0
    for v in q:
        acc += v
        C.append(acc * N * 2 ** 32)
    return np.array([bisect.bisect_right(C, (j * 2 ** 32 + U) * acc) for j in range(N)])


def test_scaled_resampling_exact():
    rng = np.random.default_rng(7)
    for trial in range(60):
        L, N = int(rng.integers(1, 80)), int(rng.integers(1, 120))
        zf = float(rng.uniform(1e-5, 1e-3))
        w = rng.random(L) / zf                 # within the bound 1 / zf
        w[rng.random(L) < 0.3] = 0.0
        if not w.any():
            w[0] = 1.0
        U = int(rng.integers(0, 2 ** 32))
        a, st = oracle.resample_scaled(w, N, U, zf)
        assert st == 0 and np.array_equal(a, _scaled_brute(w, N, U, zf)), trial


def test_scaled_resampling_equals_max_relative_when_scales_coincide():
    rng = np.random.default_rng(8)
    for trial in range(20):
        L, N = int(rng.integers(2, 50)), int(rng.integers(1, 100))
        w = rng.random(L)
        w[int(rng.integers(0, L))] = 64.0       # max w = 1 / scale exactly: w / max == w * scale
        U = int(rng.integers(0, 2 ** 32))
        a, _ = oracle.resample_scaled(w, N, U, 1.0 / 64.0)
        b, _ = oracle.resample_linear(w, N, U)
        assert np.array_equal(a, b)


def test_scaled_resampling_all_zero_is_uniform():
    L, N, U = 37, 11, 3_000_000_000
    a, st = oracle.resample_scaled(np.zeros(L), N, U, 9e-5)
    expect = [((j * 2 ** 32 + U) * L) // (N * 2 ** 32) for j in range(N)]
    assert st == 1 and list(a) == expect


def test_measurement_rows():
    props = np.array([[10.0, 11.0, 12.0], [20.0, 21.0, 22.0]])
    Z = oracle.tracker_measurement([2, 0, 2, 1], props, ROI)
    assert np.array_equal(Z[2], [12.0, 10.0, 12.0, 11.0]) and np.array_equal(Z[3], [22.0, 20.0, 22.0, 21.0])
    assert np.all(Z[0] == 320.0) and np.all(Z[4] == 320.0)
    assert np.all(Z[1] == 144.0) and np.all(Z[5] == 120 + 1.65 * 48)


def test_rpy_is_scipy_zyx():
    from scipy.spatial.transform import Rotation
    rng = np.random.default_rng(4)
    for _ in range(20):
        r, p, y = rng.uniform(-3, 3, 3)
        R = oracle.rpy(r, p, y)
        assert np.allclose(R, Rotation.from_euler("ZYX", [y, p, r]).as_matrix(), atol=1e-14)


def _pose_vector(rng):
    e = np.zeros(22)
    for k in range(5):
        e[3 * k:3 * k + 3] = [rng.uniform(100, 540), rng.uniform(80, 400), rng.uniform(1.2, 3.0)]
    e[15:18] = rng.uniform(-0.4, 0.4, 3)
    e[18:21] = rng.uniform(-0.3, 0.3, 3)
    e[21] = 1.0
    return e


def test_get3Dpose_reprojects():
    rng = np.random.default_rng(5)
    K = np.array([[525.0, 0, 319.5], [0, 525.0, 239.5], [0, 0, 1]])
    for _ in range(30):
        e = _pose_vector(rng)
        X = oracle.get3Dpose(e)
        R = oracle.rpy(e[16], e[17], e[15])
        proj = K @ (R @ X + e[18:21, None])
        for k in range(5):
            u, v, d = e[3 * k:3 * k + 3]
            assert np.allclose(proj[:, k], [u * d, v * d, d], rtol=1e-12, atol=1e-9)
    # identity pose: X = (d (u - cx) / fx, d (v - cy) / fy, d)
    e = _pose_vector(rng)
    e[15:21] = 0.0
    X = oracle.get3Dpose(e)
    for k in range(5):
        u, v, d = e[3 * k:3 * k + 3]
        assert np.allclose(X[:, k], [d * (u - 319.5) / 525.0, d * (v - 239.5) / 525.0, d], rtol=1e-14)


def test_project_order():
    rng = np.random.default_rng(6)
    Hp, M, x = rng.standard_normal((22, 12)), rng.standard_normal(22), rng.standard_normal(12)
    assert np.allclose(oracle.project(Hp, M, x), Hp @ x + M, rtol=1e-13, atol=1e-13)


def test_tf_rows_telescope():
    """The chain sums of publishTFtree's offsets equal the map (x, z, -y) of
    each joint minus the mean head, and Neck / Head rows as published."""
    rng = np.random.default_rng(7)
    p1, p2 = rng.normal(size=(3, 5)), rng.normal(size=(3, 5))
    tf = oracle.tf_rows(p1, p2)
    head = 0.5 * (p1[:, 3] + p2[:, 3])
    neck = 0.5 * (p1[:, 4] + p2[:, 4])
    T = lambda v: np.array([v[0], v[2], -v[1]])
    ref = [np.zeros(3), T(p2[:, 0] - head), T(p1[:, 0] - head), T(p2[:, 1] - head), T(p1[:, 1] - head),
           T(p1[:, 2] - head), T(p2[:, 2] - head), T(neck - head)]
    assert np.allclose(tf, np.array(ref), rtol=0, atol=1e-14)
    assert np.array_equal(tf[0], np.zeros(3))


def test_pose2d_points():
    e1, e2 = np.arange(22.0), 100.0 + np.arange(22.0)
    p = oracle.pose2d(e1, e2)
    assert np.array_equal(p[0], e1[0:2]) and np.array_equal(p[1], e2[0:2])
    assert np.array_equal(p[2], 0.5 * (e1[9:11] + e2[9:11])) and np.array_equal(p[3], 0.5 * (e1[12:14] + e2[12:14]))
    assert np.array_equal(p[4], e1[3:5]) and np.array_equal(p[5], e2[3:5])
    assert np.array_equal(p[6], e1[6:8]) and np.array_equal(p[7], e2[6:8])
