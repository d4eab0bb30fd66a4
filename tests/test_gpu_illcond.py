"""Ill-conditioned innovation covariances against the FP64 oracle.

A joint whose 2 x 2 innovation covariance S = H_b P H_b^T + R_b has
cond(S) > ~1e6 is factored in FP64 (the Cholesky, W = L^-1 H_b P and the
whitened innovations), the rest of the frame in FP32.  The inputs: gamma = 1
(the predict is the identity), for one joint R_b = 100 [[1, rho], [rho, 1]]
with rho = 1 - 1e-7, and tracks whose block H_b P H_b^T = [[1, 1], [1, 1]] +
1e-5 I (P = u u^T + 1e-5 (w1 w1^T + w2 w2^T) + 0.1 N N^T, H_b u = (1, 1),
H_b w_i = e_i, N spanning the null space of H_b): both are nearly singular
along (1, -1), so cond(S) ~ 1e7 while the update only halves the variance
there (no collapse of P that FP32 storage could not represent).  Each test
first checks on the CPU that an FP32 factorization of these S misses the
oracle's log-density by far more than the tolerance, so a kernel without
the FP64 path fails it.

Tolerances as elsewhere: log-weights 1e-4 max(1, |lw|), ancestors bit-exact
given the GPU's log-weights, means and covariances 1e-4 normwise per particle,
when the ill-conditioned joint is the first joint of the frame.  A later
joint sees P after the earlier joints' FP32 updates: their rounding
(~6e-8 |P|) perturbs the thin direction of its S (~2e-5 here) by ~1e-3
relative, which moves the log-weight by up to ~1e-3 relative (measured:
5.6e-4 for the second joint, 9.3e-4 for the third; the means and
covariances stay within 1e-4).  Those cases are checked at 2e-3.
A non-positive-definite S (an indefinite imported P) gives the particle zero
weight and PF_ERR_NOT_SPD (pf_step) / status bit 4 (batched runs).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from mkfbodytracker_pdaf_amd import capi, synthetic
from mkfbodytracker_pdaf_amd.batch import RbpfBatch
from mkfbodytracker_pdaf_amd.devmem import DeviceArray
from mkfbodytracker_pdaf_amd.rbpf import Rbpf

pytestmark = pytest.mark.gpu

JOINTS5 = ("head", "left_hand", "right_hand", "left_elbow", "right_elbow")
GATE = 9.210340371976184
RHO = 1.0 - 1e-7


def _relerr(a, b):
    n = a.shape[0]
    a = a.reshape(n, -1); b = b.reshape(n, -1)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)


def _ill_model(seed, jb, d=12, K=15, joints=("head", "left_hand", "right_hand")):
    mdl = synthetic.make_model(seed=seed, K=K, d=d, joints=joints, clutter=(0,) * len(joints))
    mdl.gamma = np.ones(K)
    R = mdl.R.copy()
    R[2 * jb:2 * jb + 2, 2 * jb:2 * jb + 2] = 100.0 * np.array([[1.0, RHO], [RHO, 1.0]])
    mdl.R = R
    return mdl


def _ill_tracks(mdl, jb, N, seed):
    """Means near the GMM means; P with H_b P H_b^T = [[1, 1], [1, 1]] + 1e-5 I."""
    rng = np.random.default_rng(seed)
    H = mdl.H()
    Hb = H[2 * jb:2 * jb + 2]
    d = H.shape[1]
    x = mdl.means[rng.integers(0, mdl.K, N)] + 0.1 * rng.standard_normal((N, d))
    Hp = np.linalg.pinv(Hb)                           # H_b Hp = I
    null = np.linalg.svd(Hb)[2][2:]                   # d - 2 directions with H_b n = 0
    P = np.empty((N, d, d))
    for j in range(N):
        u = Hp @ np.ones(2) + null.T @ (0.3 * rng.standard_normal(d - 2))
        Nm = null.T @ rng.standard_normal((d - 2, d - 2))
        P[j] = np.outer(u, u) + 1e-5 * (Hp @ Hp.T) + 0.1 * (Nm @ Nm.T) / (d - 2)
    return x, P


def _fp32_factor_error(S64):
    """|log(l00 l11)| error of the FP32 Cholesky of each 2 x 2 S (numpy float32)."""
    S = S64.astype(np.float32)
    l00 = np.sqrt(S[:, 0, 0])
    l10 = S[:, 0, 1] / l00
    with np.errstate(invalid="ignore"):
        l11 = np.sqrt(S[:, 1, 1] - l10 * l10)
    ref = 0.5 * np.log(np.linalg.det(S64))
    with np.errstate(invalid="ignore", divide="ignore"):
        err = np.abs(np.log(l00.astype(np.float64) * l11.astype(np.float64)) - ref)
    return np.where(np.isfinite(err), err, np.inf)


def _joint_S(mdl, jb, P):
    Hb = mdl.H()[2 * jb:2 * jb + 2]
    return np.einsum("ik,nkl,jl->nij", Hb, P, Hb) + mdl.R[2 * jb:2 * jb + 2, 2 * jb:2 * jb + 2]


def _check_against(ref, lw, anc, x, P, N, U2, lw_tol=1e-4):
    err = np.max(np.abs(lw - ref["lw"]) / np.maximum(1.0, np.abs(ref["lw"])))
    print(f"  max rel err lw {err:.2e}", flush=True)
    assert err <= lw_tol
    anc_o, _ = oracle.resample_log(lw.astype(np.float64), N, U2)
    assert np.array_equal(anc, anc_o)
    ex = _relerr(x, ref["xu"][anc]).max()
    eP = _relerr(P, ref["Pu"][anc]).max()
    assert ex <= 1e-4 and eP <= 1e-4, (ex, eP)


@pytest.mark.parametrize("jb,lw_tol", [(0, 1e-4), (1, 2e-3), (2, 2e-3)])
def test_illcond_joint_reference_mode_against_oracle(gpu, jb, lw_tol):
    N = 2000
    mdl = _ill_model(41, jb)
    rb = Rbpf(N)
    rb.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, sel=mdl.sel, R=mdl.R)
    x0, P0 = _ill_tracks(mdl, jb, N, 5)
    rb.set_tracks(x0, P0)
    x0, P0 = rb.get_tracks()                             # the GPU's FP32 state, exactly
    S = _joint_S(mdl, jb, P0)
    cond = np.linalg.cond(S)
    assert np.median(cond) > 1e6
    assert np.mean(_fp32_factor_error(S) > 1e-3) > 0.5   # FP32 alone would fail the tolerance
    # measurements: y + L e with the joint's own S (well inside the gate)
    rng = np.random.default_rng(9)
    H, BH = mdl.H(), mdl.BH()
    Sfull = np.einsum("ik,nkl,jl->nij", H, P0, H) + mdl.R
    Lf = np.linalg.cholesky(Sfull)
    y = x0 @ H.T + BH
    Z = (y + np.einsum("nij,nj->ni", Lf, rng.standard_normal((N, 6)))).astype(np.float32).astype(np.float64).T
    U1, U2 = 12345, 987654321
    st, out = rb.step(Z, U1, U2)
    assert st & capi.PF_WARN_ILLCOND
    om = oracle.Model(mdl.weights, mdl.gamma, mdl.means, mdl.Sigma, mdl.H(), mdl.BH(), mdl.R)
    ref = om.update_ref(x0, P0, Z, U1, U2)
    assert np.array_equal(out["comp"], ref["comp"])
    x, P = rb.get_tracks()
    _check_against(ref, out["lw"], out["anc"], x, P, N, U2, lw_tol)
    assert abs(out["lml"] - ref["lml"]) <= lw_tol * max(1.0, abs(ref["lml"]))
    rb.close()


def test_illcond_device_run_reports_through_synchronize(gpu):
    """A device-resident run with the same ill-conditioned joint: same frame
    as the host-driven step bit for bit, and the warning comes back from
    pf_synchronize."""
    N, jb = 2000, 1
    mdl = _ill_model(41, jb)
    x0, P0 = _ill_tracks(mdl, jb, N, 5)
    rng = np.random.default_rng(9)
    res = []
    for run in (False, True):
        rb = Rbpf(N)
        rb.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, sel=mdl.sel, R=mdl.R)
        rb.set_tracks(x0, P0)
        xs, Ps = rb.get_tracks()
        H, BH = mdl.H(), mdl.BH()
        Z = (xs @ H.T + BH + rng.standard_normal((N, 6)) * 0.0).astype(np.float32).T
        if not run:
            st, out = rb.step(Z.astype(np.float64), 77, 88)
            lml = out["lml"]
        else:
            d_Z = DeviceArray(Z[None].copy())
            d_uni = DeviceArray(np.array([[77, 88]], np.uint32))
            d_lml = DeviceArray.empty(1, np.float64)
            rb.run_ref_dev(1, d_Z.ptr, d_uni.ptr, d_lml.ptr, 1)
            st = rb.synchronize()
            lml = float(d_lml.get()[0])
        assert st & capi.PF_WARN_ILLCOND
        res.append((*rb.get_tracks(), lml))
        rb.close()
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]


def test_not_spd_innovation_gets_zero_weight(gpu):
    """Particles with an indefinite covariance (imported tracks): S is not
    positive definite even in FP64, their weight is zero and pf_step returns
    PF_ERR_NOT_SPD with the frame's outputs written; the other particles
    match the oracle."""
    N = 1000
    mdl = synthetic.make_model(seed=43)
    rb = Rbpf(N)
    rb.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, sel=mdl.sel, R=mdl.R)
    om = oracle.Model(mdl.weights, mdl.gamma, mdl.means, mdl.Sigma, mdl.H(), mdl.BH(), mdl.R)
    bins, _ = oracle.resample_linear(mdl.weights, N, 3)
    x0, P0 = om.reset(bins)
    bad = np.arange(N) % 3 == 0
    P0[bad] = -1e4 * np.eye(mdl.d)
    rb.set_tracks(x0, P0)
    x0, P0 = rb.get_tracks()
    seq = synthetic.simulate(mdl, 1, seed=3)
    Z = synthetic.per_particle_measurements(seq.z[0], N, seed=3).astype(np.float32).astype(np.float64)
    L = capi.lib()
    lw = np.empty(N, np.float32); anc = np.empty(N, np.int32); comp = np.empty(N, np.int32); lml = C.c_double()
    code = L.pf_step(rb.h, capi.ptr(np.ascontiguousarray(Z)), 5, 6, capi.ptr(lw), capi.ptr(anc), capi.ptr(comp),
                     C.byref(lml))
    assert code == capi.PF_ERR_NOT_SPD and b"not positive definite" in L.pf_last_error()
    assert np.all(np.isneginf(lw[bad])) and np.all(np.isfinite(lw[~bad]))
    ref = om.update_ref(x0, P0, Z, 5, 6)
    assert np.all(np.isneginf(ref["lw"][bad]))
    good = ~bad
    assert np.all(np.abs(lw[good] - ref["lw"][good]) <= 1e-4 * np.maximum(1.0, np.abs(ref["lw"][good])))
    anc_o, _ = oracle.resample_log(lw.astype(np.float64), N, 6)
    assert np.array_equal(anc, anc_o) and not np.any(bad[anc])    # zero weights are never drawn
    rb.close()


@pytest.mark.parametrize("d,variant", [(12, capi.PF_BATCH_VALU), (24, capi.PF_BATCH_MFMA),
                                       (24, capi.PF_BATCH_QUAD_VALU)])
def test_illcond_batched_against_oracle(gpu, d, variant):
    """Two sequences in one batch, the second with the ill-conditioned joint
    as its first joint (certain detections, candidate mode): the second
    matches the oracle and alone carries PF_WARN_ILLCOND.  One detection per
    joint is shared by every particle, so the whitened innovations along the
    thin direction reach ~1e3 (log-weights ~ -1e7): FP32 rounding of P in an
    earlier joint's update would show at ~1e-3 relative there, hence the
    first joint."""
    N, jb = 777, 0
    joints = ("head", "left_hand", "right_hand") if d == 12 else JOINTS5
    m = 2 * len(joints)
    models = [synthetic.make_model(seed=60, K=15, d=d, joints=joints, clutter=(0,) * len(joints)),
              _ill_model(61, jb, d=d, K=15, joints=joints)]
    b = RbpfBatch([N, N], [1, 1], K=15, m=m, d=d, cand_capacity=4, clutter_density=1e-5)
    if d == 24:
        b.variant = variant
    for s, mdl in enumerate(models):
        b.load_model(s, mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, mdl.sel, mdl.R)
    b.set_joint_clutter((0,) * len(joints))
    bins0, _ = oracle.resample_linear(models[0].weights, N, 8)
    b.reset(0, bins0)
    x1, P1 = _ill_tracks(models[1], jb, N, 12)
    b.set_tracks(1, x1, P1)
    # the GPU's FP32 state of sequence 1, read back through a one-sequence filter of the same model
    mdl = models[1]
    probe = Rbpf(N)
    probe.load_model(mdl.means, mdl.covs, mdl.weights, mdl.gamma, mdl.pca_proj, mdl.pca_mean, sel=mdl.sel, R=mdl.R)
    probe.set_tracks(x1, P1)
    xs, Ps = probe.get_tracks()
    probe.close()
    assert np.mean(_fp32_factor_error(_joint_S(mdl, jb, Ps)) > 1e-3) > 0.5
    # one certain detection per joint at the mean projection of sequence 1
    cnt = np.ones((2, len(joints)), np.int32)
    cand = np.zeros((2, len(joints), 4, 2), np.float32)
    y = (xs @ mdl.H().T + mdl.BH()).mean(axis=0)
    cand[1, :, 0, :] = y.reshape(len(joints), 2)
    cand[0, :, 0, :] = y.reshape(len(joints), 2)
    uni = np.array([[11, 22], [33, 44]], np.uint32)
    d_in = [DeviceArray(cnt), DeviceArray(cand), DeviceArray(uni), DeviceArray.empty(2, np.float64)]
    st = b.run_dev(d_in[0].ptr, d_in[1].ptr, d_in[2].ptr, d_in[3].ptr, graph=False)
    assert st & capi.PF_WARN_ILLCOND
    stat = b.status()
    assert not (stat[0] & capi.PF_WARN_ILLCOND) and (stat[1] & capi.PF_WARN_ILLCOND)
    om = oracle.Model(mdl.weights, mdl.gamma, mdl.means, mdl.Sigma, mdl.H(), mdl.BH(), mdl.R)
    ref = om.update_pdaf(xs, Ps, np.zeros(len(joints), np.int32), cnt[1], cand[1].astype(np.float64), 33, 44,
                         0.95, GATE, 1e-5)
    lw, anc = b.last_frame(1)
    x, P = b.get_tracks(1)
    _check_against(ref, lw, anc, x, P, N, 44)
    b.close()
