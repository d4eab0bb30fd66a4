"""GPU parity of the plain particle filter pf2D (src/pf2D.cpp) against the
FP64 oracle.  Particles: relative error <= 1e-6 (FP32 pixel state);
ancestors bit-exact on the GPU's own log-weights; log-weights |d| <= 1e-3
(FP32 Mahalanobis terms up to ~1e3)."""
import numpy as np
import pytest

import oracle
from mkfbodytracker_pdaf_amd.pf2d import ParticleFilter, Pf2d

pytestmark = pytest.mark.gpu


def _gmm(seed, K=15, d=8):
    rng = np.random.default_rng(seed)
    mu = np.stack([rng.uniform([0, 0] * 4, [640, 480] * 4) for _ in range(K)])
    S = []
    for _ in range(K):
        A = rng.standard_normal((d, d)) * 20
        S.append(A @ A.T + np.diag(rng.uniform(400, 4000, d)))
    w = rng.dirichlet(np.ones(K))
    return mu, np.stack(S), w


def _oracle_gmm(mu, S, w):
    Sinv = []; c = []
    for k in range(len(w)):
        si, ck = oracle.pf2d_load_gaussian(S[k])
        Sinv.append(si); c.append(ck)
    return oracle.Pf2dGmm(w, mu, np.stack(Sinv), np.array(c))


@pytest.mark.parametrize("N", [1, 100, 5000])
def test_pf2d_update_teacher_forced(gpu, N):
    mu, S, w = _gmm(1)
    g = Pf2d(N, 8, side=1)
    g.load_gmm(mu, S, w)
    og = _oracle_gmm(mu, S, w)
    rng = np.random.default_rng(N)
    unif = rng.random((N, 8))
    g.init(unif)
    p0 = oracle.pf2d_init(N, 8, 1, unif)
    pg, wg = g.get_particles()
    assert np.allclose(pg, p0, rtol=1e-6) and np.allclose(wg, 1.0 / N)
    for t in range(3):
        meas = np.array([[rng.uniform(300, 600), rng.uniform(50, 400)], [rng.uniform(0, 640), rng.uniform(0, 480)]])
        U = int(rng.integers(0, 2 ** 32))
        normals = rng.standard_normal((N, 8))
        reinit = rng.random((N, 8))
        pin, _ = g.get_particles()
        st, out = g.update(meas, U, normals, reinit)
        ref = og.update(pin, meas, U, normals, reinit, side=1)
        assert st == ref["status"] == 0
        assert np.all(np.abs(out["lw"] - ref["lw"]) <= 1e-3 * np.maximum(1.0, np.abs(ref["lw"]) / 100))
        anc_o, _ = oracle.resample_log(out["lw"].astype(np.float64), N, U)
        assert np.array_equal(out["anc"], anc_o)
        pg, wg = g.get_particles()
        expect = pin[out["anc"]].copy()
        for tt, c in enumerate([2, 3, 4, 5, 6, 7, 0, 1]):
            expect[:, c] += 5.0 * normals[:, tt]
        assert np.allclose(pg, expect, rtol=1e-6, atol=1e-4)
        # estimator: pre-resample normalised weights on the new particles (Q9)
        assert np.allclose(wg, ref["weights"], rtol=1e-3, atol=1e-9)
        e = g.estimate()
        assert np.allclose(e, oracle.pf2d_estimate(pg, wg), rtol=1e-9, atol=1e-6)
    g.close()


def test_pf2d_degenerate_reinitialises(gpu):
    """No usable weight (every log-weight non-finite; in the reference's exp
    domain, every weight zero or NaN, pf2D.cpp:232-250): the particles are
    re-drawn uniformly and the weights set to 1/N.  In the log domain equal
    far-away particles are not degenerate, so NaN particles force the case."""
    N = 256
    mu, S, w = _gmm(2)
    g = Pf2d(N, 8, side=0)
    g.load_gmm(mu, S, w)
    g.set_particles(np.full((N, 8), np.nan))
    reinit = np.random.default_rng(0).random((N, 8))
    normals = np.zeros((N, 8))
    st, out = g.update(np.array([[0.0, 0.0], [0.0, 0.0]]), 7, normals, reinit)
    assert st == 1
    og = _oracle_gmm(mu, S, w)
    ref = og.update(np.full((N, 8), np.nan), np.zeros((2, 2)), 7, normals, reinit, side=0)
    assert ref["status"] == 1
    p, wts = g.get_particles()
    assert np.allclose(p, oracle.pf2d_init(N, 8, 0, reinit), rtol=1e-6)
    assert np.allclose(wts, 1.0 / N)
    g.close()


def test_pf2d_reference_class_runs(gpu):
    pf = ParticleFilter(2000, 8, True, seed=3)
    mu, S, w = _gmm(3)
    for k in range(len(w)):
        pf.gmm.loadGaussian(mu[k], S[k], w[k])
    for _ in range(5):
        pf.update(np.array([[400.0, 200.0], [100.0, 300.0]]))
    e = pf.getEstimator()
    assert e.shape == (1, 8) and np.all(np.isfinite(e))


def test_pf2d_predict_is_the_reference_random_walk(gpu):
    """pf2d_predict on its own (ParticleFilter::predict, pf2D.cpp:90-102):
    x_i += 5 n_i with the normals consumed for columns 2..5, 6, 7, 0, 1; the
    weights are kept."""
    N = 3000
    mu, S, w = _gmm(4)
    g = Pf2d(N, 8, side=0)
    g.load_gmm(mu, S, w)
    rng = np.random.default_rng(4)
    g.init(rng.random((N, 8)))
    st, _ = g.update(np.array([[400.0, 200.0], [100.0, 300.0]]), 99, rng.standard_normal((N, 8)), rng.random((N, 8)))
    pin, win = g.get_particles()
    normals = rng.standard_normal((N, 8))
    g.predict(normals)
    pg, wg = g.get_particles()
    expect = pin.copy()
    for tt, c in enumerate([2, 3, 4, 5, 6, 7, 0, 1]):
        expect[:, c] += 5.0 * normals[:, tt]
    assert np.allclose(pg, expect, rtol=1e-6, atol=1e-4)
    assert np.array_equal(wg, win)
    g.close()


def test_pf2d_precision_form_loads_the_same_prior(gpu):
    """The GMM given as the members the reference's update reads (mean,
    sigma_i, det_s, weight; pf2D.cpp:159-163) weighs the particles as the
    covariance form does, and both match the oracle."""
    N = 2000
    mu, S, w = _gmm(5)
    og = _oracle_gmm(mu, S, w)
    rng = np.random.default_rng(5)
    unif = rng.random((N, 8))
    meas = np.array([[420.0, 180.0], [200.0, 260.0]])
    normals = rng.standard_normal((N, 8)); reinit = rng.random((N, 8))
    outs = []
    for form in ("cov", "precision"):
        g = Pf2d(N, 8, side=1)
        if form == "cov":
            g.load_gmm(mu, S, w)
        else:
            g.load_gmm_precision(mu, og.Sinv, og.cnorm, w)
        g.init(unif)
        st, out = g.update(meas, 1234, normals, reinit)
        outs.append(out["lw"].astype(np.float64))
        g.close()
    ref = og.update(oracle.pf2d_init(N, 8, 1, unif), meas, 1234, normals, reinit, side=1)
    for lw in outs:
        assert np.all(np.abs(lw - ref["lw"]) <= 1e-3 * np.maximum(1.0, np.abs(ref["lw"]) / 100))
    assert np.all(np.abs(outs[0] - outs[1]) <= 1e-5 * np.maximum(1.0, np.abs(outs[0])))


def test_pf2d_default_constructor_and_predict(gpu):
    """ParticleFilter() (pf2D.h:29) has no particles; the sized filter's
    predict() and update() read the gmm members as they are at the call."""
    from mkfbodytracker_pdaf_amd import capi
    pf0 = ParticleFilter()
    with pytest.raises(capi.PfError) as e:
        pf0.update(np.zeros((2, 2)))
    assert e.value.code == capi.PF_ERR_STATE
    pf = ParticleFilter(1000, 8, False, seed=8)
    mu, S, w = _gmm(6, K=3)
    for k in range(3):
        pf.gmm.loadGaussian(mu[k], S[k], w[k])
    pf.update(np.array([[300.0, 200.0], [100.0, 300.0]]))
    p0, _ = pf._h.get_particles()
    pf.predict()
    p1, _ = pf._h.get_particles()
    dlt = p1 - p0
    assert np.all(np.abs(dlt.mean(axis=0)) < 0.8) and np.all(np.abs(dlt.std(axis=0) - 5.0) < 0.6)
    pf.gmm.det_s[0] *= 2.0
    pf.update(np.array([[300.0, 200.0], [100.0, 300.0]]))
    assert np.all(np.isfinite(pf.getEstimator()))
