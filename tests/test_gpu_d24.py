"""The scaled state (d = 24, 5 joints in 2-D, PDAF on every joint) against the
FP64 oracle.

K = 15 keeps the model image in LDS; K = 64 exceeds the LDS budget, so the
component rows are read from global memory.  Tolerances as for d = 12:
component draws and ancestors bit-exact given the GPU's weights, particle
means and covariances within 1e-4 normwise relative error, log-weights
within 1e-4 max(1, |lw|).
"""
import numpy as np
import pytest

import oracle
from mkfbodytracker_pdaf_amd import synthetic
from mkfbodytracker_pdaf_amd.rbpf import Rbpf

pytestmark = pytest.mark.gpu

JOINTS5 = ("head", "left_hand", "right_hand", "left_elbow", "right_elbow")
CAP = 16
LAM = 10.0 / (640 * 480)


def _setup(N, K, seed=7):
    mdl = synthetic.make_model(seed=seed, K=K, d=24, joints=JOINTS5, clutter=(1, 1, 1, 1, 1))
    rb = Rbpf(N, cand_capacity=CAP, clutter_density=LAM)
    rb.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, sel=mdl.sel, R=mdl.R)
    rb.set_joint_clutter(mdl.clutter)
    om = oracle.Model(mdl.weights, mdl.gamma, mdl.means, mdl.Sigma, mdl.H(), mdl.BH(), mdl.R)
    return mdl, rb, om


def _relerr(a, b):
    n = a.shape[0]
    a = a.reshape(n, -1); b = b.reshape(n, -1)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)


@pytest.mark.parametrize("K", [15, 64])
def test_d24_candidate_step_teacher_forced(gpu, K):
    N = 512
    mdl, rb, om = _setup(N, K)
    bins, _ = oracle.resample_linear(mdl.weights, N, 3)
    x0, P0 = om.reset(bins)
    rb.set_tracks(x0, P0)
    x1, P1 = rb.get_tracks()
    assert _relerr(x1, x0).max() <= 1e-6 and _relerr(P1, P0).max() <= 1e-5     # FP32 round trip of the basis
    seq = synthetic.simulate(mdl, 3, seed=21, cand_per_joint=11, cap=CAP)
    for t in range(3):
        U1, U2 = int(seq.uniforms[t, 0]), int(seq.uniforms[t, 1])
        st, out = rb.step_candidates(seq.count[t], seq.cand[t], U1, U2)
        ref = om.update_pdaf(x0, P0, mdl.clutter, seq.count[t], seq.cand[t], U1, U2, 0.95, 9.210340371976184, LAM)
        assert np.array_equal(out["comp"], ref["comp"])
        assert np.all(np.abs(out["lw"] - ref["lw"]) <= 1e-4 * np.maximum(1.0, np.abs(ref["lw"])))
        anc_o, _ = oracle.resample_log(out["lw"].astype(np.float64), N, U2)
        assert np.array_equal(out["anc"], anc_o)
        x, P = rb.get_tracks()
        assert _relerr(x, ref["xu"][out["anc"]]).max() <= 1e-4, t
        assert _relerr(P, ref["Pu"][out["anc"]]).max() <= 1e-4, t
        x0, P0 = x, P
    rb.close()


def test_d24_device_run_matches_host_steps(gpu):
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray
    N, T, K = 2000, 9, 64
    mdl, rb, _ = _setup(N, K)
    _, rb2, _ = _setup(N, K)
    seq = synthetic.simulate(mdl, T, seed=5, cand_per_joint=11, cap=CAP)
    bins, _ = oracle.resample_linear(mdl.weights, N, 8)
    rb.reset(bins)
    rb2.reset(bins)
    lml = []
    for t in range(T):
        st, out = rb.step_candidates(seq.count[t], seq.cand[t], int(seq.uniforms[t, 0]), int(seq.uniforms[t, 1]))
        lml.append(out["lml"])
    x1, P1 = rb.get_tracks()
    d_uni = DeviceArray(seq.uniforms.astype(np.uint32))
    d_lml = DeviceArray.empty(T, np.float64)
    d_cnt = DeviceArray(seq.count.astype(np.int32))
    d_cand = DeviceArray(seq.cand.astype(np.float32))
    rb2.run_candidates_dev(T, d_cnt.ptr, d_cand.ptr, d_uni.ptr, d_lml.ptr, 1)
    rb2.synchronize()
    x2, P2 = rb2.get_tracks()
    assert np.array_equal(x1, x2) and np.array_equal(P1, P2)
    assert np.array_equal(d_lml.get(), np.array(lml))
    assert np.allclose(rb.estimate(), x1.mean(axis=0), rtol=1e-9, atol=1e-9)
    rb.close(); rb2.close()
