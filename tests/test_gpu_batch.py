"""Batched multi-sequence engine (include/pf_batch.h) against independent
single-filter runs of the same sequences.

Every sequence of a batch must reproduce, bit for bit, the tracks, the final
ancestors and the per-frame log mean weights of its own pf_run_candidates_dev
run (same model, uniforms and detections), for every schedule: the batch
shares kernels' per-particle arithmetic with the single filter and its
resampler sums are in an order fixed by the particle count alone.
"""
import numpy as np
import pytest

import oracle
from mkfbodytracker_pdaf_amd import capi, synthetic
from mkfbodytracker_pdaf_amd.batch import RbpfBatch
from mkfbodytracker_pdaf_amd.rbpf import Rbpf

pytestmark = pytest.mark.gpu

CAP = 8
LAM = 3.0 / (640 * 480)


def _problem(Ns, Ts, seed0=11):
    models, seqs, bins = [], [], []
    for s, (N, T) in enumerate(zip(Ns, Ts)):
        mdl = synthetic.make_model(seed=seed0 + s)
        models.append(mdl)
        seqs.append(synthetic.simulate(mdl, T, seed=100 + s, cand_per_joint=4, cap=CAP))
        bins.append(oracle.resample_linear(mdl.weights, N, 1000 + s)[0])
    return models, seqs, bins


def _single(mdl, seq, bins, N, T):
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray
    rb = Rbpf(N, cand_capacity=CAP, clutter_density=LAM)
    rb.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, sel=mdl.sel, R=mdl.R)
    rb.set_joint_clutter(mdl.clutter)
    rb.reset(bins)
    d_uni = DeviceArray(seq.uniforms.astype(np.uint32))
    d_lml = DeviceArray.empty(T, np.float64)
    d_cnt = DeviceArray(seq.count.astype(np.int32))
    d_cand = DeviceArray(seq.cand.astype(np.float32))
    rb.run_candidates_dev(T, d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr, 1)
    rb.synchronize()
    x, P = rb.get_tracks()
    est = rb.estimate()
    lml = d_lml.get()
    rb.close()
    return x, P, est, lml


def _batch_inputs(seqs):
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray
    cnt = np.concatenate([q.count for q in seqs]).astype(np.int32)
    cand = np.concatenate([q.cand for q in seqs]).astype(np.float32)
    uni = np.concatenate([q.uniforms for q in seqs]).astype(np.uint32)
    F = cnt.shape[0]
    return DeviceArray(cnt), DeviceArray(cand), DeviceArray(uni), DeviceArray.empty(F, np.float64)


def _make_batch(models, bins, Ns, Ts):
    b = RbpfBatch(Ns, Ts, K=models[0].K, m=models[0].m, cand_capacity=CAP, clutter_density=LAM)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    for s in range(len(Ns)):
        b.reset(s, bins[s])
    return b


NS = [300, 1000, 2500, 777, 2049, 64]
TS = [5, 17, 12, 1, 9, 17]


@pytest.fixture(scope="module")
def problem():
    models, seqs, bins = _problem(NS, TS)
    ref = [_single(models[s], seqs[s], bins[s], NS[s], TS[s]) for s in range(len(NS))]
    return models, seqs, bins, ref


def _check(b, ref, lml_all):
    offs = b.frame_offsets
    for s in range(b.S):
        x, P = b.get_tracks(s)
        xr, Pr, er, lr = ref[s]
        assert np.array_equal(x, xr), s
        assert np.array_equal(P, Pr), s
        assert np.array_equal(b.estimate(s), er), s
        lb = lml_all[offs[s]:offs[s] + TS[s]]
        assert np.array_equal(lb, lr), (s, lb, lr)


@pytest.mark.parametrize("schedule,graph", [((256, 1, 2), True), ((256, 1, 2), False), ((64, 2, 1), True),
                                            ((128, 2, 3), True), ((128, 1, 4), False)])
def test_batch_matches_single_runs(gpu, problem, schedule, graph):
    models, seqs, bins, ref = problem
    b = _make_batch(models, bins, NS, TS)
    b.schedule = schedule
    assert b.schedule == schedule
    d_cnt, d_cand, d_uni, d_lml = _batch_inputs(seqs)
    b.run_dev(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr, graph=graph)
    assert b.last_run_ms() > 0
    _check(b, ref, d_lml.get())
    b.close()


def test_batch_autotune_then_run(gpu, problem):
    models, seqs, bins, ref = problem
    b = _make_batch(models, bins, NS, TS)
    d_cnt, d_cand, d_uni, d_lml = _batch_inputs(seqs)
    best, best_ms, def_ms = b.autotune(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr, frames=8)
    assert best == b.schedule
    assert best[0] in (64, 128, 256) and best[1] in (1, 2)
    assert 0 < best_ms <= def_ms
    for s in range(len(NS)):                  # tracks were consumed by the timing runs
        b.reset(s, bins[s])
    b.run_dev(d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr)
    _check(b, ref, d_lml.get())
    b.close()


def test_batch_errors(gpu):
    from mkfbodytracker_pdaf_amd.capi import PfError
    with pytest.raises(PfError):
        RbpfBatch([100], [5], K=15, m=6, d=8)           # the batched engine runs d = 12 or 24
    with pytest.raises(PfError):
        RbpfBatch([0], [5], K=15, m=6)
    b = RbpfBatch([100, 50], [5, 3], K=15, m=6)
    mdl = synthetic.make_model(seed=1)
    with pytest.raises(PfError):                        # reset before every model is loaded
        b.reset(0, np.zeros(100, np.int32))
    with pytest.raises(PfError):
        b.schedule = (96, 1, 2)
    b.load_model(0, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.load_model(1, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    with pytest.raises(PfError):
        b.reset(1, np.full(50, 15, np.int32))           # component index out of range
    b.close()


def test_batch_d24_matches_single_runs(gpu):
    """The 24-dimensional state (5 joints, K = 64: component rows read from
    global memory) through the batched engine equals independent runs."""
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray
    joints = ("head", "left_hand", "right_hand", "left_elbow", "right_elbow")
    Ns, Ts, cap, lam = [700, 1300], [6, 4], 16, 10.0 / (640 * 480)
    models = [synthetic.make_model(seed=90 + s, K=64, d=24, joints=joints, clutter=(1, 1, 1, 1, 1))
              for s in range(2)]
    seqs = [synthetic.simulate(models[s], Ts[s], seed=70 + s, cand_per_joint=11, cap=cap) for s in range(2)]
    bins = [oracle.resample_linear(models[s].weights, Ns[s], 40 + s)[0] for s in range(2)]
    ref = []
    for s, mdl in enumerate(models):
        rb = Rbpf(Ns[s], cand_capacity=cap, clutter_density=lam)
        rb.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, sel=mdl.sel, R=mdl.R)
        rb.set_joint_clutter(mdl.clutter)
        rb.reset(bins[s])
        d = [DeviceArray(seqs[s].count.astype(np.int32)), DeviceArray(seqs[s].cand.astype(np.float32)),
             DeviceArray(seqs[s].uniforms.astype(np.uint32)), DeviceArray.empty(Ts[s], np.float64)]
        rb.run_candidates_dev(Ts[s], d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, 1)
        rb.synchronize()
        ref.append((*rb.get_tracks(), d[3].get()))
        rb.close()
    b = RbpfBatch(Ns, Ts, K=64, m=10, d=24, cand_capacity=cap, clutter_density=lam)
    b.variant = capi.PF_BATCH_VALU          # the single filter's kernel (bit-equal runs)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    for s in range(2):
        b.reset(s, bins[s])
    d = _batch_inputs(seqs)
    b.run_dev(d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr)
    lml = d[3].get()
    offs = b.frame_offsets
    for s in range(2):
        x, P = b.get_tracks(s)
        assert np.array_equal(x, ref[s][0]) and np.array_equal(P, ref[s][1]), s
        assert np.array_equal(lml[offs[s]:offs[s] + Ts[s]], ref[s][2]), s
    b.close()


def test_scaled_shape_batch_against_oracle_and_single_runs(gpu):
    """The scaled workload's own shape: 32 sequences x 31,250 particles in one
    batch (d = 24, K = 64, m = 10, 5 PDAF joints, 16 candidates per joint),
    two frames.  Sequence 0's first frame against the FP64 oracle (components
    bit-exact, log-weights and states 1e-4), and sequences 0 and 31 equal bit
    for bit to independent single-sequence runs."""
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray
    S, N, T, K, d, cap, cpj = 32, 31_250, 2, 64, 24, 16, 11
    lam = (cpj - 1) / (640.0 * 480.0)
    joints = ("head", "left_hand", "right_hand", "left_elbow", "right_elbow")
    models = [synthetic.make_model(seed=13_000 + s, K=K, d=d, joints=joints, clutter=(1, 1, 1, 1, 1))
              for s in range(S)]
    seqs = [synthetic.simulate_fast(models[s], T, seed=13_000 + s, cand_per_joint=cpj, cap=cap) for s in range(S)]
    bins = [oracle.resample_linear(models[s].weights, N, 500 + s)[0] for s in range(S)]

    def single(s):
        mdl = models[s]
        rb = Rbpf(N, cand_capacity=cap, clutter_density=lam)
        rb.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, sel=mdl.sel, R=mdl.R)
        rb.set_joint_clutter(mdl.clutter)
        rb.reset(bins[s])
        return rb

    # sequence 0, frame 0 against the oracle, from the reset tracks
    mdl = models[0]
    rb = single(0)
    om = oracle.Model(mdl.weights, mdl.gamma, mdl.means, mdl.Sigma, mdl.H(), mdl.BH(), mdl.R)
    x0, P0 = rb.get_tracks()
    U1, U2 = int(seqs[0].uniforms[0, 0]), int(seqs[0].uniforms[0, 1])
    st, out = rb.step_candidates(seqs[0].count[0], seqs[0].cand[0], U1, U2)
    ref = om.update_pdaf(x0, P0, mdl.clutter, seqs[0].count[0], seqs[0].cand[0], U1, U2, 0.95, 9.210340371976184,
                         lam)
    assert np.array_equal(out["comp"], ref["comp"])
    assert np.all(np.abs(out["lw"] - ref["lw"]) <= 1e-4 * np.maximum(1.0, np.abs(ref["lw"])))
    anc_o, _ = oracle.resample_log(out["lw"].astype(np.float64), N, U2)
    assert np.array_equal(out["anc"], anc_o)
    x, P = rb.get_tracks()
    n = x.shape[0]
    rel = lambda a, b: (np.linalg.norm((a - b).reshape(n, -1), axis=1)
                        / np.maximum(np.linalg.norm(b.reshape(n, -1), axis=1), 1e-30))
    assert rel(x, ref["xu"][out["anc"]]).max() <= 1e-4
    assert rel(P, ref["Pu"][out["anc"]]).max() <= 1e-4
    rb.close()

    # the whole batch against single runs of its first and last sequences
    singles = {}
    for s in (0, S - 1):
        rb = single(s)
        dd = [DeviceArray(seqs[s].count.astype(np.int32)), DeviceArray(seqs[s].cand.astype(np.float32)),
              DeviceArray(seqs[s].uniforms.astype(np.uint32)), DeviceArray.empty(T, np.float64)]
        rb.run_candidates_dev(T, dd[0].ptr, dd[1].ptr, dd[2].ptr, dd[3].ptr, 1)
        rb.synchronize()
        singles[s] = (*rb.get_tracks(), dd[3].get())
        rb.close()
    b = RbpfBatch([N] * S, [T] * S, K=K, m=10, d=d, cand_capacity=cap, clutter_density=lam)
    b.variant = capi.PF_BATCH_VALU          # the single filter's kernel (bit-equal runs)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    for s in range(S):
        b.reset(s, bins[s])
    dd = _batch_inputs(seqs)
    b.run_dev(dd[0].ptr, dd[1].ptr, dd[2].ptr, dd[3].ptr)
    lml = dd[3].get()
    assert np.all(np.isfinite(lml))
    offs = b.frame_offsets
    for s, (xs, Ps, ls) in singles.items():
        x, P = b.get_tracks(s)
        assert np.array_equal(x, xs) and np.array_equal(P, Ps), s
        assert np.array_equal(lml[offs[s]:offs[s] + T], ls), s
    b.close()


def test_batched_workload_shape_matches_single_runs(gpu):
    """The batched workload's own shape: 256 sequences with lengths U{100..2000}
    frames and particle counts log-uniform in [1k, 64k] (seeded as the bench
    draws them), each with its own model, in one batch; the sequences with
    the most particles, the most frames and the fewest particles equal bit for
    bit to independent single-sequence runs, every frame's log mean weight
    finite."""
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray
    seed, S, cap, cpj = 7, 256, 16, 8
    lam = (cpj - 1) / (640.0 * 480.0)
    rng = np.random.default_rng(seed)
    Ts = rng.integers(100, 2001, S).astype(np.int32)
    Ns = np.exp(rng.uniform(np.log(1000), np.log(64000), S)).astype(np.int32)
    models = [synthetic.make_model(seed=seed * 1000 + s) for s in range(S)]
    seqs = [synthetic.simulate_fast(models[s], int(Ts[s]), seed=seed * 1000 + s, cand_per_joint=cpj, cap=cap)
            for s in range(S)]
    bins = [oracle.resample_linear(models[s].weights, int(Ns[s]), 300 + s)[0] for s in range(S)]
    picks = sorted({int(np.argmax(Ns)), int(np.argmax(Ts)), int(np.argmin(Ns))})
    singles = {}
    for s in picks:
        mdl, N, T = models[s], int(Ns[s]), int(Ts[s])
        rb = Rbpf(N, cand_capacity=cap, clutter_density=lam)
        rb.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, sel=mdl.sel, R=mdl.R)
        rb.set_joint_clutter(mdl.clutter)
        rb.reset(bins[s])
        dd = [DeviceArray(seqs[s].count.astype(np.int32)), DeviceArray(seqs[s].cand.astype(np.float32)),
              DeviceArray(seqs[s].uniforms.astype(np.uint32)), DeviceArray.empty(T, np.float64)]
        rb.run_candidates_dev(T, dd[0].ptr, dd[1].ptr, dd[2].ptr, dd[3].ptr, 1)
        rb.synchronize()
        singles[s] = (*rb.get_tracks(), dd[3].get())
        rb.close()
    b = RbpfBatch(Ns, Ts, K=models[0].K, m=models[0].m, cand_capacity=cap, clutter_density=lam)
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    for s in range(S):
        b.reset(s, bins[s])
    dd = _batch_inputs(seqs)
    b.run_dev(dd[0].ptr, dd[1].ptr, dd[2].ptr, dd[3].ptr)
    lml = dd[3].get()
    assert lml.shape[0] == int(Ts.sum()) and np.all(np.isfinite(lml))
    offs = b.frame_offsets
    for s, (xs, Ps, ls) in singles.items():
        x, P = b.get_tracks(s)
        assert np.array_equal(x, xs) and np.array_equal(P, Ps), s
        assert np.array_equal(lml[offs[s]:offs[s] + int(Ts[s])], ls), s
    b.close()
