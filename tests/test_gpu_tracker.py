"""The pose tracker's per-frame pipeline (include/pf_tracker.h) against the
FP64 oracle (oracle/pf_pose_ref.c) and against the filter's own C ABI.

Per frame, with both arms built from the shipped GMM files:
* blurred image: bit-exact;
* proposals: bit-exact (Philox words, Box-Muller from IEEE-exact operations,
  positions from the same projected mean);
* association weights: the oracle's from the reference's formula within
  1e-12 relative, same zero pattern;
* proposal indices: bit-exact against the oracle resampler fed the GPU's
  weights;
* measurements: bit-exact (FP32 of the oracle's FP64 rows);
* arm updates: bit-identical to pf_step on twin handles with the same
  measurements and uniforms, whose component draws, ancestors, states and
  log mean weights are checked against the oracle's update (tolerances of
  the filter's own tests: 1e-4);
* pose output: e within 1e-12 and get3Dpose within 1e-9 of the oracle's
  literal LU form, both from the GPU's estimate.
Device-resident runs (graph and eager) equal frame-by-frame calls bit for bit.
"""
import numpy as np
import pytest

import oracle
from mkfbodytracker_pdaf_amd import capi, synthetic
from mkfbodytracker_pdaf_amd.rbpf import Rbpf
from mkfbodytracker_pdaf_amd.tracker import PoseTracker

pytestmark = pytest.mark.gpu

LEFT = "tests/golden/data13D_PCA_100000_15_12.yml"
RIGHT = "tests/golden/data23D_PCA_100000_15_12.yml"
SEED = 11


def _arm(path, N, a):
    rb = Rbpf(N)
    rb.load_model_yml(path)
    w = capi.yml_read_matrix(path, "weights").ravel()
    bins, _ = rb.resample(w, N, int(capi.philox_u32(SEED, 32 + a, 0, 1)[0]))
    rb.reset(bins)
    return rb, bins


def _pca(path):
    return capi.yml_read_matrix(path, "pca_proj"), capi.yml_read_matrix(path, "pca_mean").ravel()


def _tracker(N, **cfg):
    arms = [_arm(p, N, a)[0] for a, p in enumerate((LEFT, RIGHT))]
    (P1, M1), (P2, M2) = _pca(LEFT), _pca(RIGHT)
    tr = PoseTracker(arms[0], arms[1], P1, M1, P2, M2, seed=SEED, **cfg)
    return tr, arms


def _oracle_model(path):
    means = capi.yml_read_matrix(path, "means")
    covs = capi.yml_read_matrix(path, "covs")
    w = capi.yml_read_matrix(path, "weights").ravel()
    g = capi.yml_read_matrix(path, "gamma").ravel()
    P, M = _pca(path)
    K, d = means.shape
    sel = np.array([9, 10, 0, 1, 12, 13], np.int32)
    kf = oracle.load_gaussian(means[0], covs[:d], P, M, g[0], sel, 100.0 * np.eye(6))
    return oracle.Model(w, g, means, covs.reshape(K, d, d), kf["H"], kf["BH"], kf["R"]), w


def _rel(a, b):
    n = a.shape[0]
    a = a.reshape(n, -1); b = b.reshape(n, -1)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-30)


def test_blur_bit_exact(gpu):
    tr, arms = _tracker(200)
    rng = np.random.default_rng(1)
    imgs = [rng.integers(0, 256, (480, 640), dtype=np.uint8), synthetic.likelihood_image(synthetic.HANDS, seed=4),
            np.zeros((480, 640), np.uint8), np.full((480, 640), 255, np.uint8)]
    for img in imgs:
        assert np.array_equal(tr.blur(img), oracle.blur_u8(img))
    tr.close()
    small, arms2 = _tracker(200, rows=37, cols=70, blur_ksize=9, blur_sigma=1.7)
    img = rng.integers(0, 256, (37, 70), dtype=np.uint8)
    assert np.array_equal(small.blur(img), oracle.blur_u8(img, 9, 1.7))
    small.close()
    for a in arms + arms2:
        a.close()


@pytest.fixture(scope="module")
def frames():
    return synthetic.tracker_sequence(4, seed=5)


def _against_oracle(tr, arms, twins, arms_o, pcas, N, imgs, rois, frames=3, twin_exact=True):
    """Runs frames through tr and checks every stage against the oracle (the
    module docstring's list); twins are handles with the same models, started
    from the arms' tracks through get_tracks / set_tracks.  That round trip
    goes through the FP64 basis change of the selection basis and back, which
    can move an FP32 covariance entry by one unit in the last place for some
    models; twin_exact=False then compares the twin within 1e-6."""
    orc = oracle.PoseTrackerOracle(arms_o, [P.T for P, _ in pcas], [M for _, M in pcas], N, seed=SEED)
    for g in range(frames):
        img, roi = imgs[g], rois[g]
        xbars = [a.estimate() for a in arms]
        before = [a.get_tracks() for a in arms]
        out = tr.callback(img, roi)
        last = tr.get_last()
        assert out["status"] == 0
        ref = orc.proposal_stage(xbars, img, roi, g, uniform=(g == 0), weights_from=last["w"])
        assert np.array_equal(last["blurred"], ref["blurred"])
        words = orc.frame_words(g)
        for a in range(2):
            assert np.array_equal(last["props"][a], ref["props"][a]), (g, a)
            w_g, w_o = last["w"][a], ref["w"][a]
            assert np.array_equal(w_g == 0, w_o == 0), (g, a)
            nz = w_o > 0
            assert nz.sum() > 10, (g, a)                    # the blobs are found
            assert np.max(np.abs(w_g[nz] - w_o[nz]) / w_o[nz]) <= 1e-12, (g, a)
            assert np.array_equal(last["bins"][a], ref["bins"][a]), (g, a)
            assert np.array_equal(last["Z"][a], ref["Z"][a].astype(np.float32)), (g, a)
            # the arm update equals pf_step on a twin with the same inputs ...
            twins[a].set_tracks(*before[a])
            st, o = twins[a].step(last["Z"][a].astype(np.float64), int(words[2 * a]), int(words[2 * a + 1]))
            xa, Pa = arms[a].get_tracks()
            xt, Pt = twins[a].get_tracks()
            if twin_exact:
                assert np.array_equal(xa, xt) and np.array_equal(Pa, Pt), (g, a)
                assert out["lml"][a] == o["lml"], (g, a)
            else:
                assert _rel(xt, xa).max() <= 1e-6 and _rel(Pt, Pa).max() <= 1e-6, (g, a)
                assert abs(out["lml"][a] - o["lml"]) <= 1e-9 * max(1.0, abs(o["lml"])), (g, a)
            # ... whose step the oracle pins
            r = arms_o[a].update_ref(before[a][0], before[a][1], last["Z"][a].astype(np.float64),
                                     int(words[2 * a]), int(words[2 * a + 1]))
            assert np.array_equal(o["comp"], r["comp"]), (g, a)
            assert np.all(np.abs(o["lw"] - r["lw"]) <= 1e-4 * np.maximum(1.0, np.abs(r["lw"]))), (g, a)
            anc_o, _ = oracle.resample_log(o["lw"].astype(np.float64), N, int(words[2 * a + 1]))
            assert np.array_equal(o["anc"], anc_o), (g, a)
            assert _rel(xa, r["xu"][o["anc"]]).max() <= 1e-4 and _rel(Pa, r["Pu"][o["anc"]]).max() <= 1e-4
            assert abs(o["lml"] - r["lml"]) <= 1e-4 * max(1.0, abs(r["lml"])), (g, a)
        # pose output from the updated tracks
        es, p3, tf, p2d = orc.pose([a.estimate() for a in arms])
        for a in range(2):
            assert np.allclose(out["pose"][a], es[a], rtol=1e-12, atol=1e-12), (g, a)
            assert np.allclose(out["pos3d"][a], p3[a], rtol=1e-9, atol=1e-12), (g, a)
            assert np.allclose(out["pos3d"][a], oracle.get3Dpose(out["pose"][a]), rtol=1e-9, atol=1e-12)
        # TF rows and 2-D points from the GPU's own pose: the oracle's literal
        # chain composition (sums of published offsets) and point list
        assert np.allclose(out["tf"], oracle.tf_rows(out["pos3d"][0], out["pos3d"][1]), rtol=0, atol=1e-12)
        assert np.allclose(out["tf"], tf, rtol=0, atol=1e-9)
        assert np.array_equal(out["pose2d"], oracle.pose2d(out["pose"][0], out["pose"][1]))
        assert np.allclose(out["pose2d"], p2d, rtol=1e-12, atol=1e-12)
    po = tr.pose()
    for k in ("pose", "pos3d", "tf", "pose2d"):
        assert np.array_equal(po[k], out[k]), k


def test_frames_match_oracle(gpu, frames):
    N = 1000
    imgs, rois = frames
    tr, arms = _tracker(N)
    twins = [_arm(p, N, a)[0] for a, p in enumerate((LEFT, RIGHT))]
    arms_o = [_oracle_model(LEFT)[0], _oracle_model(RIGHT)[0]]
    _against_oracle(tr, arms, twins, arms_o, [_pca(LEFT), _pca(RIGHT)], N, imgs, rois)
    tr.close()
    for a in arms + twins:
        a.close()


def test_tracker_shape_frames_match_oracle(gpu, frames):
    """The tracker workload's own shape: 100,000 particles per arm (10^6
    proposals per arm), the shipped models, two frames checked stage by stage
    against the oracle.  The twins go through the FP64 basis round trip of
    get_tracks / set_tracks, so they are compared within 1e-6."""
    N = 100_000
    imgs, rois = frames
    tr, arms = _tracker(N)
    twins = [_arm(p, N, a)[0] for a, p in enumerate((LEFT, RIGHT))]
    arms_o = [_oracle_model(LEFT)[0], _oracle_model(RIGHT)[0]]
    _against_oracle(tr, arms, twins, arms_o, [_pca(LEFT), _pca(RIGHT)], N, imgs, rois, frames=2,
                    twin_exact=False)
    tr.close()
    for a in arms + twins:
        a.close()


def _resized_arm(path, D, seed):
    """The shipped arm model with state dimension D: its first D coordinates
    (a marginal of the mixture) for D < 12, or padded with D - 12 coordinates
    of small independent noise and small loadings for D > 12."""
    means = capi.yml_read_matrix(path, "means")
    K, d = means.shape
    covs = capi.yml_read_matrix(path, "covs").reshape(K, d, d)
    w = capi.yml_read_matrix(path, "weights").ravel()
    g = capi.yml_read_matrix(path, "gamma").ravel()
    P, M = _pca(path)
    if D <= d:
        return means[:, :D].copy(), covs[:, :D, :D].copy(), w, g, P[:D].copy(), M
    rng = np.random.default_rng(seed)
    e = D - d
    mu = np.concatenate([means, 0.1 * rng.standard_normal((K, e))], axis=1)
    S = np.zeros((K, D, D))
    S[:, :d, :d] = covs
    S[:, d:, d:] = 0.05 * np.eye(e)
    Pd = np.concatenate([P, 0.02 * rng.standard_normal((e, P.shape[1]))], axis=0)
    return mu, S, w, g, Pd, M


def test_arms_of_other_state_dimensions(gpu, frames):
    """Arms with d = 8 and d = 16 (and 24 and 12): the hand projections of
    both arms are staged together in k_propose, the filter step, ancestors
    and estimate kernels run at each arm's own d."""
    N = 600
    imgs, rois = frames
    sel = np.array([9, 10, 0, 1, 12, 13], np.int32)
    for dims in ((8, 16), (24, 12)):
        models = [_resized_arm(p, D, 40 + D) for p, D in zip((LEFT, RIGHT), dims)]

        def handle(a):
            mu, S, w, g, P, M = models[a]
            rb = Rbpf(N)
            rb.load_model(mu, S.reshape(-1, mu.shape[1]), w, g, P, M)
            bins, _ = rb.resample(w, N, int(capi.philox_u32(SEED, 32 + a, 0, 1)[0]))
            rb.reset(bins)
            return rb

        arms, twins = [handle(0), handle(1)], [handle(0), handle(1)]
        arms_o = []
        for mu, S, w, g, P, M in models:
            kf = oracle.load_gaussian(mu[0], S[0], P, M, g[0], sel, 100.0 * np.eye(6))
            arms_o.append(oracle.Model(w, g, mu, S, kf["H"], kf["BH"], kf["R"]))
        pcas = [(m[4], m[5]) for m in models]
        tr = PoseTracker(arms[0], arms[1], pcas[0][0], pcas[0][1], pcas[1][0], pcas[1][1], seed=SEED)
        _against_oracle(tr, arms, twins, arms_o, pcas, N, imgs, rois, frames=2, twin_exact=False)
        tr.close()
        for a in arms + twins:
            a.close()


@pytest.mark.parametrize("graph,T", [(True, 6), (False, 6), (True, 20)])
def test_device_run_equals_frame_calls(gpu, graph, T):
    from mkfbodytracker_pdaf_amd.devmem import DeviceArray
    N = 700
    imgs, rois = synthetic.tracker_sequence(T, seed=8)
    tr_a, arms_a = _tracker(N)
    outs = [tr_a.callback(imgs[t], rois[t]) for t in range(T)]
    tr_b, arms_b = _tracker(N)
    d_img, d_roi = DeviceArray(imgs), DeviceArray(rois)
    d_lml = DeviceArray.empty(2 * T, np.float64)
    d_pose = DeviceArray.empty(T * 2 * 22, np.float64)
    d_p3 = DeviceArray.empty(T * 2 * 15, np.float64)
    d_tf = DeviceArray.empty(T * 24, np.float64)
    d_p2 = DeviceArray.empty(T * 16, np.float64)
    tr_b.run_dev(T, d_img.ptr, d_roi.ptr, d_lml.ptr, d_pose.ptr, d_p3.ptr, d_tf.ptr, d_p2.ptr, graph=graph)
    tr_b.synchronize()
    assert tr_b.last_run_ms() > 0
    lml = d_lml.get().reshape(2, T)
    pose = d_pose.get().reshape(T, 2, 22)
    p3 = d_p3.get().reshape(T, 2, 3, 5)
    tf = d_tf.get().reshape(T, 8, 3)
    p2 = d_p2.get().reshape(T, 8, 2)
    for t in range(T):
        assert np.array_equal(lml[:, t], outs[t]["lml"]), t
        assert np.array_equal(pose[t], outs[t]["pose"]), t
        assert np.array_equal(p3[t], outs[t]["pos3d"]), t
        assert np.array_equal(tf[t], outs[t]["tf"]), t
        assert np.array_equal(p2[t], outs[t]["pose2d"]), t
    for a in range(2):
        xa, Pa = arms_a[a].get_tracks()
        xb, Pb = arms_b[a].get_tracks()
        assert np.array_equal(xa, xb) and np.array_equal(Pa, Pb)
    # a second run continues the frame count (Gaussian proposals, next words)
    tr_a.callback(imgs[0], rois[0])
    tr_b.run_dev(1, d_img.ptr, d_roi.ptr, d_lml.ptr, d_pose.ptr, d_p3.ptr, graph=graph)
    tr_b.synchronize()
    for a in range(2):
        assert np.array_equal(arms_a[a].get_tracks()[0], arms_b[a].get_tracks()[0])
    for t in (tr_a, tr_b):
        t.close()
    for a in arms_a + arms_b:
        a.close()


def test_frame_without_face(gpu):
    """A frame without a face ROI updates nothing, reports zero 2-D points
    and makes the next frame start again from uniform proposals."""
    N = 400
    imgs, rois = synthetic.tracker_sequence(3, seed=12)
    tr_a, arms_a = _tracker(N)
    tr_b, arms_b = _tracker(N)
    for t in range(2):
        oa, ob = tr_a.callback(imgs[t], rois[t]), tr_b.callback(imgs[t], rois[t])
    before = [a.get_tracks()[0] for a in arms_a]
    none = tr_a.callback(imgs[2], None)
    assert np.all(none["pose2d"] == 0.0) and none["status"] == 0
    assert all(np.array_equal(a.get_tracks()[0], b) for a, b in zip(arms_a, before))
    oa = tr_a.callback(imgs[2], rois[2])
    last = tr_a.get_last()
    x0, y0, w, h = (int(v) for v in rois[2])
    for a in range(2):       # uniform box below the face
        px, py = last["props"][a]
        assert px.min() >= max(x0 - 4 * w, 0) and px.max() < min(x0 + 5 * w, 640)
        assert py.min() >= min(y0 + h, 480) and py.max() < min(y0 + 7 * h, 480)
    # the same frame index and words as a tracker reset before that frame
    ob = tr_b.callback(imgs[2], rois[2])
    assert not np.array_equal(oa["pose"], ob["pose"])    # b used Gaussian proposals
    for t in (tr_a, tr_b):
        t.close()
    for a in arms_a + arms_b:
        a.close()


def test_blank_image_falls_back_to_uniform(gpu):
    N = 300
    tr, arms = _tracker(N)
    out = tr.callback(np.zeros((480, 640), np.uint8), synthetic.FACE_ROI)
    assert out["status"] == 1
    last = tr.get_last()
    n_c = tr.n_c
    for a in range(2):
        assert np.all(last["w"][a] == 0.0)
        U = int(oracle.philox_u32(SEED, 24, 0, 6)[4 + a])
        ref, st = oracle.resample_scaled(np.zeros(n_c), N, U, 1e-4 * 0.9)
        assert st == 1 and np.array_equal(last["bins"][a], ref)
    tr.close()
    for a in arms:
        a.close()


def test_tracker_errors(gpu):
    from mkfbodytracker_pdaf_amd.capi import PfError
    a1, _ = _arm(LEFT, 200, 0)
    a2, _ = _arm(RIGHT, 300, 1)
    (P1, M1), (P2, M2) = _pca(LEFT), _pca(RIGHT)
    with pytest.raises(PfError):
        PoseTracker(a1, a2, P1, M1, P2, M2)                     # different N
    a3, _ = _arm(RIGHT, 200, 1)
    with pytest.raises(PfError):
        PoseTracker(a1, a3, P1[:, :20], M1[:20], P2[:, :20], M2[:20])   # npose < 21
    with pytest.raises(PfError):
        PoseTracker(a1, a3, P1, M1, P2, M2, blur_ksize=14)
    with pytest.raises(PfError):
        PoseTracker(a1, a3, P1, M1, P2, M2, proposals_per_particle=6000)  # above the scan limit
    tr = PoseTracker(a1, a3, P1, M1, P2, M2)
    with pytest.raises(PfError):
        tr.callback(np.zeros((480, 640), np.uint8), (300, 50, 0, 40))     # empty ROI
    tr.close()
    for a in (a1, a2, a3):
        a.close()
