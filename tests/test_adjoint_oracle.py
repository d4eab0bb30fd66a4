"""The FP64 reference gradient of the log-marginal-likelihood (oracle
``orc_adj_*``) pinned on the CPU:

* its value reproduces the recorded forward run of the oracle's own
  ``orc_rbpf_update_pdaf`` (same frozen choices, same per-frame log mean weights);
* its forward-mode gradient agrees with central finite differences
  (h = 1e-4, FP64) on every parameter group;
* its logit gradient equals the closed form sum_t (sum_{j: k_tj = k} wbar_tj - pi_k)
  (the score of the component draws), computed from the run's weights alone.
"""
import numpy as np
import pytest

import oracle
from mkfbodytracker_pdaf_amd import synthetic

PD, GATE = 0.95, 9.210340371976184
LAM = 3.0 / (640 * 480)


def _recorded_run(K=3, N=40, T=5, seed=5):
    mdl = synthetic.make_model(seed=seed, K=K)
    seq = synthetic.simulate(mdl, T, seed=seed + 1, cand_per_joint=4, cap=6)
    om = oracle.Model(mdl.weights, mdl.gamma, mdl.means, mdl.Sigma, mdl.H(), mdl.BH(), mdl.R)
    bins, _ = oracle.resample_linear(mdl.weights, N, 77)
    x, P = om.reset(bins)
    comp, anc, lml, lw = [], [], [], []
    for t in range(T):
        r = om.update_pdaf(x, P, mdl.clutter, seq.count[t], seq.cand[t], int(seq.uniforms[t, 0]),
                           int(seq.uniforms[t, 1]), PD, GATE, LAM)
        comp.append(r["comp"]); anc.append(r["anc"]); lml.append(r["lml"]); lw.append(r["lw"])
        x, P = r["x"], r["P"]
    prob = oracle.AdjProblem(mdl.H(), mdl.BH(), mdl.R, mdl.clutter, seq.count, seq.cand, bins,
                             np.array(comp), np.array(anc), PD, GATE, LAM, K=K)
    theta = oracle.AdjProblem.theta(mdl.means, mdl.weights, mdl.gamma, mdl.covs)
    return mdl, prob, theta, np.array(lml), np.array(lw), np.array(comp)


@pytest.fixture(scope="module")
def run():
    return _recorded_run()


def test_value_reproduces_forward_run(run):
    mdl, prob, theta, lml, lw, comp = run
    F, lml_r = prob.value(theta)
    assert np.allclose(lml_r, lml, rtol=1e-11, atol=1e-9)
    assert abs(F - lml.sum()) < 1e-8 * max(1.0, abs(F))


def _fd(prob, theta, idx, h=1e-4):
    lg = theta[prob.K * prob.d:prob.K * prob.d + prob.K]
    logpi0 = lg - np.logaddexp.reduce(lg)
    out = np.empty(len(idx))
    for n, i in enumerate(idx):
        tp = theta.copy(); tp[i] += h
        tm = theta.copy(); tm[i] -= h
        out[n] = (prob.value(tp, logpi0)[0] - prob.value(tm, logpi0)[0]) / (2 * h)
    return out


def _groups(prob):
    K, d = prob.K, prob.d
    a = K * d
    rng = np.random.default_rng(0)
    chol_lower = [a + 2 * K + k * d * d + i * d + j for k in range(K) for i in range(d) for j in range(i + 1)]
    return {"means": rng.choice(a, 12, replace=False), "logits": np.arange(a, a + K),
            "rates": np.arange(a + K, a + 2 * K), "chol": rng.choice(chol_lower, 16, replace=False)}


def test_gradient_matches_central_differences(run):
    mdl, prob, theta, *_ = run
    F, g = prob.grad(theta)
    for name, idx in _groups(prob).items():
        fd = _fd(prob, theta, idx)
        scale = max(np.abs(fd).max(), 1e-3)
        assert np.abs(g[idx] - fd).max() <= 1e-5 * scale, (name, g[idx], fd)
    # the upper triangle of each factor is not used: zero gradient
    K, d = prob.K, prob.d
    L = g[K * d + 2 * K:].reshape(K, d, d)
    assert np.all(L[:, np.triu_indices(d, 1)[0], np.triu_indices(d, 1)[1]] == 0.0)


def test_logit_gradient_closed_form(run):
    mdl, prob, theta, lml, lw, comp = run
    _, g = prob.grad(theta)
    K, d, N = prob.K, prob.d, prob.N
    pi = mdl.weights / mdl.weights.sum()
    expect = np.zeros(K)
    for t in range(prob.T):
        wb = np.exp(lw[t] - lml[t] - np.log(N))
        for k in range(K):
            expect[k] += wb[comp[t] == k].sum() - pi[k] * wb.sum()
    assert np.allclose(g[K * d:K * d + K], expect, rtol=1e-9, atol=1e-10)


def test_directional_derivatives_are_gradient_projections(run):
    """orc_adj_dirderiv (one forward pass over up to 8 directions) equals the
    full forward-mode gradient projected on those directions."""
    mdl, prob, theta, lml, lw, comp = run
    F, g = prob.grad(theta)
    V = np.random.default_rng(11).standard_normal((8, prob.n_theta))
    F2, dd = prob.dirderiv(theta, V)
    assert F2 == F
    assert np.allclose(dd, V @ g, rtol=1e-10, atol=1e-10 * np.abs(V @ g).max())
