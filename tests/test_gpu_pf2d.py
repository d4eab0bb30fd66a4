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
