"""The four-lanes-per-particle step of the 24-dimensional state against the
FP64 oracle: PF_BATCH_MFMA (covariance tiles updated on the matrix cores) and
PF_BATCH_QUAD_VALU (the same layout on the vector ALU).

Frames are teacher forced: every frame starts each checked sequence from the
GPU's own tracks of the previous frame, and the oracle runs the same frame
from those tracks.  Tolerances as for the other d = 24 tests: log-weights
within 1e-4 max(1, |lw|), ancestors bit-exact given the GPU's log-weights,
particle means and covariances within 1e-4 normwise relative error per
particle (the resampled tracks against the oracle's updated tracks at the
GPU's ancestors), the frame's log mean weight within 1e-4.
"""
import numpy as np
import pytest

import oracle
from mkfbodytracker_pdaf_amd import capi, synthetic
from mkfbodytracker_pdaf_amd.batch import RbpfBatch
from mkfbodytracker_pdaf_amd.devmem import DeviceArray

pytestmark = pytest.mark.gpu

JOINTS5 = ("head", "left_hand", "right_hand", "left_elbow", "right_elbow")
JOINTS3 = ("head", "left_hand", "right_hand")
CAP = 16
GATE = 9.210340371976184


def _relerr(a, b):
    n = a.shape[0]
    a = a.reshape(n, -1); b = b.reshape(n, -1)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)


def _one_frame_inputs(seqs, t):
    """Inputs of frame t of every sequence as a batch of one-frame sequences."""
    cnt = np.stack([q.count[t] for q in seqs]).astype(np.int32)
    cand = np.stack([q.cand[t] for q in seqs]).astype(np.float32)
    uni = np.stack([q.uniforms[t] for q in seqs]).astype(np.uint32)
    return [DeviceArray(cnt), DeviceArray(cand), DeviceArray(uni), DeviceArray.empty(len(seqs), np.float64)]


def _run_teacher_forced(models, Ns, seqs, variant, frames, checked, lam, threads=8):
    """Runs `frames` one-frame batches; sequences in `checked` start every
    frame from their previous GPU tracks and are compared with the oracle,
    the others restart from their reset tracks.  Returns the per-frame log
    mean weights of every sequence."""
    S = len(models)
    d = models[0].d
    b = RbpfBatch(Ns, [1] * S, K=models[0].K, m=models[0].m, d=d, cand_capacity=CAP, clutter_density=lam)
    b.variant = variant
    assert b.variant == variant
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter(models[0].clutter)
    bins = [oracle.resample_linear(models[s].weights, int(Ns[s]), 900 + s)[0] for s in range(S)]
    oms = {s: oracle.Model(models[s].weights, models[s].gamma, models[s].means, models[s].Sigma, models[s].H(),
                           models[s].BH(), models[s].R) for s in checked}
    oracle.set_threads(threads)
    state = {}
    lmls = []
    try:
        for t in range(frames):
            for s in range(S):
                if t > 0 and s in state:
                    b.set_tracks(s, *state[s])
                else:
                    b.reset(s, bins[s])
            d_in = _one_frame_inputs(seqs, t)
            starts = {}
            for s in checked:
                if t == 0:
                    starts[s] = oms[s].reset(bins[s])
                else:
                    starts[s] = state[s]
            print(f"variant {variant} frame {t}: S={S} N={list(Ns[:2])}..", flush=True)
            b.run_dev(d_in[0].ptr, d_in[1].ptr, d_in[2].ptr, d_in[3].ptr, graph=False)
            lml = d_in[3].get()
            assert np.all(np.isfinite(lml)), t
            lmls.append(lml)
            for s in checked:
                mdl, q = models[s], seqs[s]
                U1, U2 = int(q.uniforms[t, 0]), int(q.uniforms[t, 1])
                x0, P0 = starts[s]
                ref = oms[s].update_pdaf(x0, P0, mdl.clutter, q.count[t], q.cand[t], U1, U2, 0.95, GATE, lam)
                lw, anc = b.last_frame(s)
                assert np.all(np.abs(lw - ref["lw"]) <= 1e-4 * np.maximum(1.0, np.abs(ref["lw"]))), (s, t)
                anc_o, _ = oracle.resample_log(lw.astype(np.float64), int(Ns[s]), U2)
                assert np.array_equal(anc, anc_o), (s, t)
                x, P = b.get_tracks(s)
                ex = _relerr(x, ref["xu"][anc]).max()
                eP = _relerr(P, ref["Pu"][anc]).max()
                print(f"  seq {s} frame {t}: max rel err x {ex:.2e} P {eP:.2e}", flush=True)
                assert ex <= 1e-4 and eP <= 1e-4, (s, t, ex, eP)
                assert abs(lml[s] - ref["lml"]) <= 1e-4 * max(1.0, abs(ref["lml"])), (s, t)
                state[s] = (x, P)
            for a in d_in:
                a.close()
    finally:
        oracle.set_threads(1)
        b.close()
    return np.array(lmls)


@pytest.mark.parametrize("variant", [capi.PF_BATCH_MFMA, capi.PF_BATCH_QUAD_VALU])
@pytest.mark.parametrize("K,Ns,joints,clutter", [
    (64, (1000, 31), JOINTS5, (1, 1, 1, 1, 1)),      # component rows in global memory; 31: a partial wave
    (15, (777, 256), JOINTS5, (0, 1, 1, 0, 1)),      # image in LDS; certain and PDAF joints mixed
    (64, (300, 4099), JOINTS3, (1, 0, 1)),           # three joints (rows 0..5 of the selection basis)
])
def test_q24_teacher_forced(gpu, variant, K, Ns, joints, clutter):
    cpj = 11
    lam = (cpj - 1) / (640.0 * 480.0)
    models = [synthetic.make_model(seed=600 + 7 * s + K, K=K, d=24, joints=joints, clutter=clutter)
              for s in range(len(Ns))]
    seqs = [synthetic.simulate(models[s], 4, seed=610 + s, cand_per_joint=cpj, cap=CAP) for s in range(len(Ns))]
    _run_teacher_forced(models, np.array(Ns, np.int32), seqs, variant, 4, checked=range(len(Ns)), lam=lam)


def test_q24_mfma_scaled_shape_teacher_forced(gpu):
    """The scaled workload's own shape: 32 sequences x 31,250 particles, K = 64,
    5 PDAF joints, 16 candidates per joint with Poisson(10) clutter, on the
    matrix-core variant: sequence 0 teacher forced for three frames and
    sequence 31 for one against the oracle, every sequence's log mean weight
    finite and within 1e-4 of the vector-ALU variant's on the first frame."""
    S, N, K, cpj = 32, 31_250, 64, 11
    lam = (cpj - 1) / (640.0 * 480.0)
    models = [synthetic.make_model(seed=13_000 + s, K=K, d=24, joints=JOINTS5, clutter=(1, 1, 1, 1, 1))
              for s in range(S)]
    seqs = [synthetic.simulate_fast(models[s], 3, seed=13_000 + s, cand_per_joint=cpj, cap=CAP) for s in range(S)]
    Ns = np.full(S, N, np.int32)
    lml_m = _run_teacher_forced(models, Ns, seqs, capi.PF_BATCH_MFMA, 3, checked=(0,), lam=lam, threads=16)
    lml_m31 = _run_teacher_forced(models, Ns, seqs, capi.PF_BATCH_MFMA, 1, checked=(S - 1,), lam=lam, threads=16)
    lml_v = _run_teacher_forced(models, Ns, seqs, capi.PF_BATCH_VALU, 1, checked=(), lam=lam)
    assert np.array_equal(lml_m[0], lml_m31[0])          # frame 0 from the reset tracks: same run
    assert np.all(np.abs(lml_m[0] - lml_v[0]) <= 1e-4 * np.maximum(1.0, np.abs(lml_v[0])))


def test_q24_variant_api(gpu):
    b = RbpfBatch([100], [2], K=15, m=10, d=24)
    assert b.variant == capi.PF_BATCH_MFMA                   # the default for d = 24
    b.variant = capi.PF_BATCH_QUAD_VALU
    assert b.variant == capi.PF_BATCH_QUAD_VALU
    with pytest.raises(capi.PfError):
        b.variant = 7
    with pytest.raises(capi.PfError):
        b.schedule = (128, 1, 2)                              # four-lane variants run 256-thread tiles
    b.variant = capi.PF_BATCH_VALU
    b.schedule = (128, 1, 2)
    b.close()
    b12 = RbpfBatch([100], [2], K=15, m=6)
    assert b12.variant == capi.PF_BATCH_VALU
    with pytest.raises(capi.PfError):
        b12.variant = capi.PF_BATCH_MFMA
    b12.close()
