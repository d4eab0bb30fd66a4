"""Pins the FP64 oracle (oracle/pf_oracle.c) on the CPU.

Each check is independent of the oracle's own code: golden vectors written by
tests/golden/make_golden.py (scipy, extended-precision conditioning, exact
rationals), the reference's shipped model files, closed forms and invariants
chosen so that a dropped term, a wrong sign/index or a transposed operand
fails at least one of them.
"""
import math
import os

import numpy as np
import pytest
import scipy.stats

import oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = np.load(os.path.join(GOLD, "golden_vectors.npz"))


# ---------------------------------------------------------------- Philox, exp
def test_philox_known_answer_vectors():
    """Random123 known-answer vectors of Philox4x32-10."""
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, out in kat:
        assert list(oracle.philox4x32_10(ctr, key)) == out


def test_det_exp_accuracy_and_special_values():
    xs = np.concatenate([-np.logspace(-12, np.log10(740), 4000), [0.0, -1e-300]])
    rel = max(abs(oracle.det_exp(x) / math.exp(x) - 1.0) for x in xs if math.exp(x) > 1e-300)
    assert rel < 1e-14
    assert oracle.det_exp(0.0) == 1.0
    assert oracle.det_exp(-800.0) == 0.0 and oracle.det_exp(float("nan")) == 0.0


# ---------------------------------------------------------------- resampling
@pytest.mark.parametrize("L,N", [(7, 20), (40, 40), (100, 33)])
def test_resample_matches_exact_rationals(L, N):
    w = G[f"sys{L}_{N}_w"]; U = int(G[f"sys{L}_{N}_U"][0])
    a, st = oracle.resample_linear(w, N, U)
    assert st == 0 and np.array_equal(a, G[f"sys{L}_{N}_anc"])
    # the same weights as log-weights give the same ancestors
    with np.errstate(divide="ignore"):
        b, _ = oracle.resample_log(np.log(w), N, U)
    assert np.array_equal(a, b)


def test_resample_properties():
    rng = np.random.default_rng(0)
    for trial in range(200):
        L = int(rng.integers(1, 60)); N = int(rng.integers(1, 300))
        w = rng.random(L) ** 4
        w[rng.random(L) < 0.3] = 0.0
        if not w.any():
            w[0] = 0.5
        U = int(rng.integers(0, 2 ** 32))
        a, st = oracle.resample_linear(w, N, U)
        assert st == 0
        assert np.all(np.diff(a) >= 0)                      # systematic: non-decreasing
        assert np.all(w[a] > 0)                             # zero weights never selected
        cnt = np.bincount(a, minlength=L)
        expect = N * w / w.sum()
        assert np.all(np.abs(cnt - expect) < 1.0 + 1e-6)    # n_a in {floor, ceil} of N w_a


def test_resample_degenerate_and_u_extremes():
    a, st = oracle.resample_linear(np.zeros(5), 10, 123)
    assert st == 1 and np.array_equal(a, (np.arange(10) * 5 + 123 / 2 ** 32 * 5) // 10)
    w = np.array([0.0, 2.0, 0.0, 2.0])
    a0, _ = oracle.resample_linear(w, 4, 0)
    a1, _ = oracle.resample_linear(w, 4, 2 ** 32 - 1)
    assert list(a0) == [1, 1, 3, 3] and list(a1) == [1, 1, 3, 3]


def _block_scaled_brute(lw, N, U):
    """The block-scaled resampling of log-weights restated with Python
    integers: blocks of 256, e_b = ceil(m_b log2 e), q_i = floor(2^32
    det_exp(lw_i - e_b ln 2)), C_k = sum of the earlier blocks' sums >> (E - e_b)
    plus the block-local prefix >> (E - e_b), a_j by bisection."""
    import bisect
    L = len(lw)
    blocks = [list(range(b, min(b + 256, L))) for b in range(0, L, 256)]
    es = []
    for ids in blocks:
        fin = [lw[i] for i in ids if math.isfinite(lw[i])]
        es.append(None if not fin else max(-2 ** 30, min(2 ** 30, math.ceil(max(fin) * 1.4426950408889634))))
    E = max((e for e in es if e is not None), default=None)
    C, pre = [], 0
    for ids, e in zip(blocks, es):
        s = 64 if (e is None or E - e >= 64) else E - e
        loc = 0
        for i in ids:
            if e is not None and math.isfinite(lw[i]):
                v = oracle.det_exp(lw[i] - e * 0.6931471805599453)
                loc += min(2 ** 32, math.floor(math.ldexp(v, 32))) if v > 0 else 0
            C.append(pre + (loc >> s if s < 64 else 0))
        pre += loc >> s if s < 64 else 0
    T = pre
    if T == 0:
        C, T = list(range(1, L + 1)), L
    scaled = [c * N * 2 ** 32 for c in C]
    return np.array([bisect.bisect_right(scaled, (j * 2 ** 32 + U) * T) for j in range(N)]), int(pre == 0)


def test_resample_log_block_scaled_exact():
    rng = np.random.default_rng(5)
    for trial in range(12):
        L, N = int(rng.integers(1, 1800)), int(rng.integers(1, 700))
        nb = (L + 255) // 256
        offs = rng.choice([0.0, -30.0, -60.0, 25.0, -500.0], size=nb)     # blocks 2^40 .. 2^700 apart
        lw = rng.normal(-3.0, 2.0, L) + np.repeat(offs, 256)[:L]
        lw[rng.random(L) < 0.1] = -np.inf
        if nb > 2 and trial % 3 == 0:
            lw[256:512] = -np.inf                                      # an empty block
        U = int(rng.integers(0, 2 ** 32))
        a, st = oracle.resample_log(lw, N, U)
        b, stb = _block_scaled_brute(list(lw), N, U)
        assert st == stb == 0 and np.array_equal(a, b), trial
    # every weight zero: the uniform resampling
    a, st = oracle.resample_log(np.full(700, -np.inf), 9, 77)
    b, stb = _block_scaled_brute([-math.inf] * 700, 9, 77)
    assert st == stb == 1 and np.array_equal(a, b)
    assert np.array_equal(a, (np.arange(9) * 2 ** 32 + 77) * 700 // (9 * 2 ** 32))


def test_block_pow_closed_form():
    for m in (-1000.5, -1.0, -1e-300, 0.0, 1e-300, 0.69, 0.7, 1.0, 100.0, 3e9, -3e9):
        assert oracle.block_pow(m) == max(-2 ** 30, min(2 ** 30, math.ceil(m * 1.4426950408889634)))
    # the block's largest weight keeps at least 2^31 quanta: exp(m - e ln 2) in (1/2, 1]
    for m in np.linspace(-50, 50, 1001):
        e = oracle.block_pow(float(m))
        v = oracle.det_exp(float(m) - e * 0.6931471805599453)
        assert 0.5 < v <= 1.0 + 1e-15


def test_wheel_quirk_agrees_with_spec_off_ties():
    """The reference's literal FP64 wheel (pf2DRao.cpp:193-208) and the spec
    pick the same indices for generic weights (they differ only on exact ties
    and at the FP total)."""
    rng = np.random.default_rng(1)
    agree = 0
    for _ in range(100):
        L, N = int(rng.integers(2, 50)), int(rng.integers(1, 200))
        w = rng.random(L); w /= w.sum()
        U = int(rng.integers(1, 2 ** 32 - 1))
        a, _ = oracle.resample_linear(w, N, U)
        b, _ = oracle.resample_wheel(w, N, U / 2 ** 32)
        agree += np.array_equal(a, b)
    assert agree >= 95


# ------------------------------------------------------------------ densities
@pytest.mark.parametrize("m", [2, 6])
def test_mvn_logpdf_golden(m):
    S, u, X = G[f"mvn{m}_S"], G[f"mvn{m}_u"], G[f"mvn{m}_X"]
    got = np.array([oracle.mvn_logpdf(x, u, S) for x in X])
    assert np.allclose(got, G[f"mvn{m}_logpdf"], rtol=1e-12, atol=1e-10)


def test_mvn_logpdf_not_spd():
    assert oracle.mvn_logpdf([0, 0], [0, 0], np.array([[1.0, 2.0], [2.0, 1.0]])) == -np.inf


# ------------------------------------------------------------------ KF_model
@pytest.mark.parametrize("m", [6, 2])
def test_kf_update_equals_gaussian_conditioning(m):
    x, P, H, BH, R, z = (G[f"kf{m}_{k}"] for k in ("x", "P", "H", "BH", "R", "z"))
    xo, Po = oracle.kf_update(H, BH, R, z, x, P)
    assert np.allclose(xo, G[f"kf{m}_xpost"], rtol=1e-10, atol=1e-10)
    assert np.allclose(Po, G[f"kf{m}_Ppost"], rtol=1e-9, atol=1e-10)
    # information form: P+^-1 = P^-1 + H^T R^-1 H
    assert np.allclose(np.linalg.inv(Po), np.linalg.inv(P) + H.T @ np.linalg.inv(R) @ H, rtol=1e-8)


def test_kf_predict_closed_form_and_fixed_points():
    rng = np.random.default_rng(2)
    d, g = 12, 0.9
    mu = rng.standard_normal(d)
    A = rng.standard_normal((d, d)); Sig = A @ A.T + np.eye(d)
    F, B, Q = g * np.eye(d), (1 - g) * mu, (1 - g * g) * Sig
    x = rng.standard_normal(d); P = np.eye(d) * 3
    xo, Po = oracle.kf_predict(F, B, Q, x, P)
    assert np.allclose(xo, g * x + (1 - g) * mu) and np.allclose(Po, g * g * P + (1 - g * g) * Sig)
    xf, Pf = oracle.kf_predict(F, B, Q, mu, Sig)            # (mu, Sigma) is the fixed point
    assert np.allclose(xf, mu) and np.allclose(Pf, Sig)
    Fn = rng.standard_normal((d, d))                         # generic F: F P F^T, not F^T P F
    xo, Po = oracle.kf_predict(Fn, B, Q, x, P)
    assert np.allclose(Po, Fn @ P @ Fn.T + Q) and np.allclose(xo, Fn @ x + B)


# ------------------------------------------------------------- my_gmm loader
def test_load_gaussian_on_shipped_model():
    """my_gmm::loadGaussian (my_gmm.cpp:45-75) on the reference's left-arm model."""
    from mkfbodytracker_pdaf_amd import capi
    p = os.path.join(GOLD, "data13D_PCA_100000_15_12.yml")
    means = capi.yml_read_matrix(p, "means"); covs = capi.yml_read_matrix(p, "covs")
    proj = capi.yml_read_matrix(p, "pca_proj"); pmean = capi.yml_read_matrix(p, "pca_mean")[0]
    gam = capi.yml_read_matrix(p, "gamma").ravel()
    sel = [9, 10, 0, 1, 12, 13]
    for k in (0, 7, 14):
        out = oracle.load_gaussian(means[k], covs[12 * k:12 * (k + 1)], proj, pmean, gam[k], sel, 100 * np.eye(6))
        assert np.allclose(out["F"], gam[k] * np.eye(12))
        assert np.allclose(out["Q"], (1 - gam[k] ** 2) * covs[12 * k:12 * (k + 1)])
        assert np.allclose(out["B"], (1 - gam[k]) * means[k])
        # H = H1 pca_proj^T: row r picks pose coordinate sel[r] (head u,v; hand u,v; neck u,v)
        assert np.array_equal(out["H"], proj[:, sel].T) and np.array_equal(out["BH"], pmean[sel])
        assert out["H"][2, 0] == proj[0, 0] and out["BH"][0] == pmean[9]


# ------------------------------------------------------------------ RBPF step
def _model(seed=0, K=3, d=4, m=2):
    rng = np.random.default_rng(seed)
    mu = rng.standard_normal((K, d)) * 3
    Sig = np.stack([(lambda A: A @ A.T + np.eye(d))(rng.standard_normal((d, d))) for _ in range(K)])
    H = rng.standard_normal((m, d)); BH = rng.standard_normal(m); R = 2.0 * np.eye(m)
    pi = rng.dirichlet(np.ones(K)); g = rng.uniform(0.7, 0.95, K)
    return oracle.Model(pi, g, mu, Sig, H, BH, R)


def test_rbpf_step_single_component_is_kalman_filter():
    """K = 1: every particle runs predict + KF update; weights are the KF
    predictive densities; identical particles give lml = that density."""
    mdl = _model(K=1)
    N, d = 7, mdl.d
    x0 = np.tile(mdl.mu[0] + 0.5, (N, 1)); P0 = np.tile(np.eye(d), (N, 1, 1))
    z = np.array([1.0, -2.0])
    Z = np.tile(z[:, None], (1, N))
    r = mdl.update_ref(x0, P0, Z, 5, 6)
    g = mdl.gamma[0]
    xp = g * x0[0] + (1 - g) * mdl.mu[0]; Pp = g * g * np.eye(d) + (1 - g * g) * mdl.Sigma[0]
    S = mdl.H @ Pp @ mdl.H.T + mdl.R
    lw = scipy.stats.multivariate_normal(mdl.H @ xp + mdl.BH, S).logpdf(z)
    assert np.allclose(r["lw"], lw, rtol=1e-12) and abs(r["lml"] - lw) < 1e-10
    K = Pp @ mdl.H.T @ np.linalg.inv(S)
    assert np.allclose(r["x"], xp + K @ (z - mdl.H @ xp - mdl.BH)) and np.allclose(r["P"], Pp - K @ S @ K.T)


def test_rbpf_components_follow_systematic_draw_of_pi():
    mdl = _model(K=3)
    N = 300
    x0 = np.tile(mdl.mu[0], (N, 1)); P0 = np.tile(np.eye(mdl.d), (N, 1, 1))
    r = mdl.update_ref(x0, P0, np.zeros((2, N)), 77, 1)
    a, _ = oracle.resample_linear(mdl.pi, N, 77)
    assert np.array_equal(r["comp"], a)
    anc, _ = oracle.resample_log(r["lw"], N, 1)
    assert np.array_equal(r["anc"], anc) and np.allclose(r["x"], r["xu"][anc])


def test_pdaf_single_certain_detection_reduces_to_kf():
    """PDAF with one detection, P_D = 1 and an infinite gate is the KF update;
    its log-likelihood is log(N(z) / lambda)."""
    mdl = _model(K=1)
    N, d = 3, mdl.d
    x0 = np.tile(mdl.mu[0], (N, 1)); P0 = np.tile(np.eye(d), (N, 1, 1))
    cand = np.array([[[0.7, -1.1]]])
    lam = 1e-3
    a = mdl.update_pdaf(x0, P0, [1], [1], cand, 1, 2, 1.0, 1e300, lam)
    b = mdl.update_pdaf(x0, P0, [0], [1], cand, 1, 2, 1.0, 1e300, lam)   # certain joint
    assert np.allclose(a["xu"], b["xu"], rtol=1e-12) and np.allclose(a["Pu"], b["Pu"], rtol=1e-10)
    assert np.allclose(a["lw"], b["lw"] - math.log(lam), rtol=1e-12)


def test_pdaf_equals_moment_matched_mixture():
    """PDAF posterior = moment match of {beta_0: predicted, beta_c: KF with z_c}
    (the association mixture), computed here explicitly."""
    mdl = _model(K=1, d=5)
    d = mdl.d
    x0 = mdl.mu[0][None] + 0.3; P0 = np.eye(d)[None] * 2
    cand = np.array([[[0.5, 0.2], [3.0, -1.0], [40.0, 40.0]]])
    pd, gate, lam = 0.9, 9.21, 0.01
    r = mdl.update_pdaf(x0, P0, [1], [3], cand, 3, 4, pd, gate, lam)
    g = mdl.gamma[0]
    xp = g * x0[0] + (1 - g) * mdl.mu[0]; Pp = g * g * P0[0] + (1 - g * g) * mdl.Sigma[0]
    y = mdl.H @ xp + mdl.BH; S = mdl.H @ Pp @ mdl.H.T + mdl.R
    K = Pp @ mdl.H.T @ np.linalg.inv(S)
    pg = 1 - math.exp(-gate / 2)
    comps = [((1 - pd * pg), xp, Pp)]
    for c in cand[0]:
        nu = c - y
        d2 = nu @ np.linalg.solve(S, nu)
        e = pd / lam * scipy.stats.multivariate_normal(y, S).pdf(c) if d2 <= gate else 0.0
        comps.append((e, xp + K @ nu, Pp - K @ S @ K.T))
    wsum = sum(c[0] for c in comps)
    mean = sum(c[0] * c[1] for c in comps) / wsum
    cov = sum(c[0] * (c[2] + np.outer(c[1] - mean, c[1] - mean)) for c in comps) / wsum
    assert np.allclose(r["xu"][0], mean, rtol=1e-10) and np.allclose(r["Pu"][0], cov, rtol=1e-9)
    assert abs(r["lw"][0] - math.log(wsum)) < 1e-10


def test_pdaf_no_detection_keeps_prediction():
    mdl = _model(K=1)
    x0 = mdl.mu[0][None] + 1; P0 = np.eye(mdl.d)[None]
    r = mdl.update_pdaf(x0, P0, [1], [0], np.zeros((1, 4, 2)), 1, 2, 0.95, 9.21, 1e-3)
    g = mdl.gamma[0]
    assert np.allclose(r["xu"][0], g * x0[0] + (1 - g) * mdl.mu[0])
    assert abs(r["lw"][0] - math.log(1 - 0.95 * (1 - math.exp(-4.605)))) < 1e-12


def test_estimate_and_proposals():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((50, 12))
    assert np.allclose(oracle.rbpf_estimate(x), x.mean(0), rtol=1e-13)
    Hp = rng.standard_normal((22, 12)); M = rng.standard_normal(22); xbar = x.mean(0)
    n = rng.standard_normal((2, 100))
    s = oracle.get_samples(Hp, M, xbar, 100, 30.0, n)
    y = Hp[:2] @ xbar + M[:2]
    assert np.allclose(s, y[:, None] + 0.8 * 30.0 * n)          # 0.8 scale is a std-dev (Q14)
    w = oracle.get_sample_prob(Hp, M, xbar, s, 30.0)
    assert np.allclose(w, scipy.stats.multivariate_normal(y, 0.8 * 30.0 * np.eye(2)).pdf(s.T))  # a variance


# ------------------------------------------------------------------ pf2D
def test_pf2d_prior_and_likelihood_golden():
    mus, Ss, pis, X = G["gmm_mu"], G["gmm_S"], G["gmm_pi"], G["gmm_X"]
    Sinv, cn = zip(*[oracle.pf2d_load_gaussian(S) for S in Ss])
    gm = oracle.Pf2dGmm(pis, mus, np.stack(Sinv), np.array(cn))
    meas = np.array([[300.0, 200.0], [100.0, 50.0]])
    N = X.shape[0]
    r = gm.update(X, meas, 1, np.zeros((N, 8)), np.zeros((N, 8)), side=0)
    lik = (scipy.stats.multivariate_normal(meas[0], 15 * np.eye(2)).logpdf(X[:, 6:8])
           + scipy.stats.multivariate_normal(meas[1], 15 * np.eye(2)).logpdf(X[:, 0:2]))
    assert np.allclose(r["lw"], G["gmm_logpdf"] + lik, rtol=1e-11)
    w = np.exp(r["lw"] - r["lw"].max()); w /= w.sum()
    assert np.allclose(r["weights"], w, rtol=1e-10)


def test_pf2d_predict_draw_order_and_estimator():
    gm = oracle.Pf2dGmm([1.0], np.zeros((1, 8)), np.eye(8)[None], np.array([1.0]))
    N = 4
    p = np.arange(N * 8, dtype=float).reshape(N, 8) + 100
    normals = np.tile(np.arange(8, dtype=float), (N, 1))
    r = gm.update(p, np.zeros((2, 2)), 0, normals, np.zeros((N, 8)), side=0)
    moved = r["particles"] - p[r["anc"]]
    # normals consumed for columns 2,3,4,5 then 6,7 then 0,1 (pf2D.cpp:96-98)
    assert np.allclose(moved[0], 5.0 * np.array([6, 7, 0, 1, 2, 3, 4, 5]))
    e = oracle.pf2d_estimate(r["particles"], r["weights"])
    assert np.allclose(e, (r["weights"][:, None] * r["particles"]).sum(0))


def test_pf2d_init_bounds():
    N = 2000
    u = np.random.default_rng(4).random((N, 8))
    for side in (0, 1):
        p = oracle.pf2d_init(N, 8, side, u)
        assert np.all(p[:, 0::2][:, [0, 1, 2]] >= 1) and np.all(p[:, [0, 2, 4]] < 640)
        assert np.all(p[:, [1, 3, 5, 7]] < 480)
        assert np.all(p[:, 6] >= 320 * side + 1) and np.all(p[:, 6] < 320 + 320 * side)
