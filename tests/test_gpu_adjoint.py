"""GPU adjoint of the log-marginal-likelihood (pf_batch_lml_grad) against the
FP64 oracle gradient of the same recorded run.

The oracle differentiates, by forward-mode dual numbers in FP64, the
log-marginal-likelihood of the GPU run's own recorded choices (component
draws, ancestors); that reference is itself pinned by central finite
differences (tests/test_adjoint_oracle.py).

Tolerances: per parameter group (means, logits, rates, Cholesky factors) and
sequence, ||g_gpu - g_oracle|| <= 1e-3 ||g_oracle||; per-frame log mean
weights within 1e-4 max(1, |lml|).
"""
import numpy as np
import pytest

import oracle
from mkfbodytracker_pdaf_amd import synthetic
from mkfbodytracker_pdaf_amd.batch import RbpfBatch

pytestmark = pytest.mark.gpu

CAP = 8
K = 3
PD, GATE = 0.95, 9.210340371976184
LAM = 3.0 / (640 * 480)
NS = [96, 160, 64]
TS = [6, 9, 4]


def _problem():
    models, seqs, bins = [], [], []
    for s, (N, T) in enumerate(zip(NS, TS)):
        mdl = synthetic.make_model(seed=40 + s, K=K)
        models.append(mdl)
        seqs.append(synthetic.simulate(mdl, T, seed=60 + s, cand_per_joint=4, cap=CAP))
        bins.append(oracle.resample_linear(mdl.weights, N, 500 + s)[0])
    return models, seqs, bins


def _inputs(seqs):
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray
    cnt = np.concatenate([q.count for q in seqs]).astype(np.int32)
    cand = np.concatenate([q.cand for q in seqs]).astype(np.float32)
    uni = np.concatenate([q.uniforms for q in seqs]).astype(np.uint32)
    return DeviceArray(cnt), DeviceArray(cand), DeviceArray(uni), DeviceArray.empty(cnt.shape[0], np.float64)


def _batch(models, bins):
    b = RbpfBatch(NS, TS, K=K, m=6, cand_capacity=CAP, clutter_density=LAM)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    for s in range(len(NS)):
        b.reset(s, bins[s])
    return b


@pytest.fixture(scope="module")
def run():
    models, seqs, bins = _problem()
    b = _batch(models, bins)
    d = _inputs(seqs)
    g = b.lml_grad(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, checkpoint_every=4)
    lml = d[3].get()
    choices = [b.get_choices(s) for s in range(len(NS))]
    b.close()
    return models, seqs, bins, g, lml, choices


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def test_adjoint_matches_oracle_gradient(gpu, run):
    models, seqs, bins, g, lml, choices = run
    offs = np.concatenate([[0], np.cumsum(TS)[:-1]])
    for s, mdl in enumerate(models):
        comp, anc = choices[s]
        prob = oracle.AdjProblem(mdl.H(), mdl.BH(), mdl.R, mdl.clutter, seqs[s].count, seqs[s].cand, bins[s],
                                 comp, anc, PD, GATE, LAM, K=K)
        theta = oracle.AdjProblem.theta(mdl.means, mdl.weights, mdl.gamma, mdl.covs)
        F, go = prob.grad(theta)
        _, lml_o = prob.value(theta)
        lg = lml[offs[s]:offs[s] + TS[s]]
        assert np.all(np.abs(lg - lml_o) <= 1e-4 * np.maximum(1.0, np.abs(lml_o))), (s, lg, lml_o)
        gm, gl, gr, gc = prob.split(go)
        for name, gpu_g, ref in (("means", g["means"][s], gm), ("logits", g["logits"][s], gl),
                                 ("rates", g["rates"][s], gr), ("chol", g["chol"][s], gc)):
            assert _rel(gpu_g, ref) <= 1e-3, (s, name, _rel(gpu_g, ref))


def test_adjoint_forward_is_the_plain_run(gpu, run):
    """The recorded forward of the adjoint gives bit-identical log mean
    weights to pf_batch_run_dev on the same inputs."""
    models, seqs, bins, g, lml, choices = run
    b = _batch(models, bins)
    d = _inputs(seqs)
    b.run_dev(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr)
    assert np.array_equal(d[3].get(), lml)
    b.close()


@pytest.mark.parametrize("C", [1, 3, 100])
def test_adjoint_independent_of_checkpoint_interval(gpu, run, C):
    models, seqs, bins, g, lml, choices = run
    b = _batch(models, bins)
    d = _inputs(seqs)
    g2 = b.lml_grad(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, checkpoint_every=C)
    # the replayed states are bit-identical to the recorded ones and the
    # component terms are summed in an order fixed by the tile: the
    # gradients are bit-identical
    for k in g:
        assert np.array_equal(g2[k], g[k]), (C, k, _rel(g2[k], g[k]))
    fwd, bwd = b.last_adjoint_ms()
    assert fwd > 0 and bwd > 0
    b.close()


def test_torch_function_backward(gpu, run):
    torch = pytest.importorskip("torch")
    from mkfbodytracker_pdaf_amd.torch_adjoint import BatchProblem, lml as lml_fn
    models, seqs, bins, g, lml, choices = run
    m0 = models[0]
    cnt = np.concatenate([q.count for q in seqs])
    cand = np.concatenate([q.cand for q in seqs])
    uni = np.concatenate([q.uniforms for q in seqs])
    prob = BatchProblem(NS, TS, m0.pca_proj, m0.pca_mean, m0.sel, m0.R, m0.clutter, cnt, cand, uni, bins,
                        checkpoint_every=4, cand_capacity=CAP, clutter_density=LAM)
    Ls = np.linalg.cholesky(m0.Sigma)
    params = [torch.tensor(v, dtype=torch.float64, requires_grad=True)
              for v in (m0.means, np.log(m0.weights), 1.0 - m0.gamma, Ls)]
    out = lml_fn(*params, prob)
    out.backward()
    # the same problem through the C ABI directly: every sequence shares model 0
    b = RbpfBatch(NS, TS, K=K, m=6, cand_capacity=CAP, clutter_density=LAM)
    for s in range(len(NS)):
        b.load_model(s, m0.means, m0.covs, m0.weights, m0.gamma, m0.pca_proj, m0.pca_mean, m0.sel, m0.R)
    b.set_joint_clutter(m0.clutter)
    for s in range(len(NS)):
        b.reset(s, bins[s])
    d = _inputs(seqs)
    gd = b.lml_grad(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, checkpoint_every=4)
    assert abs(out.item() - d[3].get().sum()) < 1e-7 * max(1.0, abs(out.item()))
    # the torch path rebuilds weights and covariances from logits and factors
    # (exp / log, L L^T), so its model differs from model 0 by rounding only
    for p, k in zip(params, ("means", "logits", "rates", "chol")):
        assert _rel(p.grad.numpy(), gd[k].sum(axis=0)) <= 1e-6, k
    b.close()
    # inputs on a torch GPU device in FP32: the value and every gradient come
    # back on that device in the input's dtype (the gradient of the
    # FP32-rounded parameters)
    dev = torch.device("cuda", 0)
    params32 = [p.detach().to(device=dev, dtype=torch.float32).requires_grad_(True) for p in params]
    out32 = lml_fn(*params32, prob)
    out32.backward()
    assert out32.device == dev
    for p32, p, k in zip(params32, params, ("means", "logits", "rates", "chol")):
        assert p32.grad.device == dev and p32.grad.dtype == torch.float32, k
        assert _rel(p32.grad.double().cpu().numpy(), p.grad.numpy()) <= 1e-3, k


def test_adjoint_workload_shape(gpu):
    """The adjoint workload's own shape: 64 sequences x 16,384 particles x 500
    frames (own model per sequence, seeded as the bench draws them),
    checkpoint every 50 frames.  Size-independent properties: the recorded
    forward gives the plain run's log mean weights bit for bit, the gradient
    does not depend on the checkpoint interval (50 vs 125, bit for bit: the
    component sums are added in an order fixed by the tiles) and every entry
    is finite."""
    from mkfbodytracker_pdaf_amd.batch import RbpfBatch
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray
    S, N, T, seed, cap, cpj = 64, 16_384, 500, 11, 16, 8
    lam = (cpj - 1) / (640.0 * 480.0)
    models = [synthetic.make_model(seed=seed * 1000 + s) for s in range(S)]
    seqs = [synthetic.simulate_fast(models[s], T, seed=seed * 1000 + s, cand_per_joint=cpj, cap=cap)
            for s in range(S)]
    bins = [oracle.resample_linear(models[s].weights, N, 700 + s)[0] for s in range(S)]
    b = RbpfBatch([N] * S, [T] * S, K=models[0].K, m=models[0].m, cand_capacity=cap, clutter_density=lam)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    d_cnt = DeviceArray(np.concatenate([q.count for q in seqs]).astype(np.int32))
    d_cand = DeviceArray(np.concatenate([q.cand for q in seqs]).astype(np.float32))
    d_uni = DeviceArray(np.concatenate([q.uniforms for q in seqs]).astype(np.uint32))
    d_lml = DeviceArray.empty(S * T, np.float64)

    def reset_all():
        for s in range(S):
            b.reset(s, bins[s])

    reset_all()
    b.run_dev(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr)
    lml_plain = d_lml.get()
    grads = []
    for C in (50, 125):
        reset_all()
        grads.append(b.lml_grad(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr, checkpoint_every=C))
        assert np.array_equal(d_lml.get(), lml_plain), C
    for k in grads[0]:
        assert np.all(np.isfinite(grads[0][k])), k
        assert np.array_equal(grads[1][k], grads[0][k]), (k, _rel(grads[1][k], grads[0][k]))
    b.close()


def test_adjoint_k15_real_size_directional(gpu):
    """K = 15 at a real size: sequence 0 with N = 4096 particles and T = 50
    frames (batched with a second sequence of 1000 x 30).  The oracle's
    forward-mode pass gives the directional derivatives of the same recorded
    run along 8 unit directions: one random direction inside each parameter
    group (means, logits, rates, the lower Cholesky entries) and four over all
    parameters.  The GPU gradient projected on each matches to 1e-3 of the
    GPU gradient's norm over the direction's support (the norm tolerance of
    the small-problem check bounds exactly this); the per-frame log mean
    weights match the oracle's to 1e-4."""
    K15, Ns, Ts = 15, [4096, 1000], [50, 30]
    models = [synthetic.make_model(seed=70 + s, K=K15) for s in range(2)]
    seqs = [synthetic.simulate(models[s], Ts[s], seed=80 + s, cand_per_joint=4, cap=CAP) for s in range(2)]
    bins = [oracle.resample_linear(models[s].weights, Ns[s], 900 + s)[0] for s in range(2)]
    b = RbpfBatch(Ns, Ts, K=K15, m=6, cand_capacity=CAP, clutter_density=LAM)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    for s in range(2):
        b.reset(s, bins[s])
    d = _inputs(seqs)
    g = b.lml_grad(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, checkpoint_every=10)
    lml = d[3].get()[:Ts[0]]
    comp, anc = b.get_choices(0)
    b.close()
    mdl = models[0]
    prob = oracle.AdjProblem(mdl.H(), mdl.BH(), mdl.R, mdl.clutter, seqs[0].count, seqs[0].cand, bins[0],
                             comp, anc, PD, GATE, LAM, K=K15)
    theta = oracle.AdjProblem.theta(mdl.means, mdl.weights, mdl.gamma, mdl.covs)
    _, lml_o = prob.value(theta)
    assert np.all(np.abs(lml - lml_o) <= 1e-4 * np.maximum(1.0, np.abs(lml_o)))
    n, dd_ = prob.n_theta, mdl.d
    a = K15 * dd_
    low = np.zeros((K15, dd_, dd_), bool)
    low[:, np.tril_indices(dd_)[0], np.tril_indices(dd_)[1]] = True
    support = [np.r_[np.ones(a, bool), np.zeros(n - a, bool)],
               np.r_[np.zeros(a, bool), np.ones(K15, bool), np.zeros(n - a - K15, bool)],
               np.r_[np.zeros(a + K15, bool), np.ones(K15, bool), np.zeros(n - a - 2 * K15, bool)],
               np.r_[np.zeros(a + 2 * K15, bool), low.ravel()]]
    every = np.r_[np.ones(a + 2 * K15, bool), low.ravel()]
    support += [every] * 4
    rng = np.random.default_rng(31)
    V = np.zeros((8, n))
    for c, sup in enumerate(support):
        v = np.where(sup, rng.standard_normal(n), 0.0)
        V[c] = v / np.linalg.norm(v)
    _, dd = prob.dirderiv(theta, V)
    gflat = np.concatenate([g["means"][0].ravel(), g["logits"][0].ravel(), g["rates"][0].ravel(),
                            g["chol"][0].ravel()])
    assert gflat.shape == (n,)
    for c, sup in enumerate(support):
        err = abs(gflat @ V[c] - dd[c])
        scale = np.linalg.norm(np.where(sup, gflat, 0.0))
        print(f"direction {c}: gpu {gflat @ V[c]:.6e} oracle {dd[c]:.6e} |err|/|g| {err / scale:.2e}", flush=True)
        assert err <= 1e-3 * scale, (c, err, scale)
