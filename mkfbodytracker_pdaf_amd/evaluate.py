"""Evaluation harness of the pose tracker on synthetic ground truth.

The counterpart of the reference's only experiment protocol:
``extractTFdata`` (src/extractTFdata.cpp:45-97) logs, at every frame, the
tracker's TF joints and the Kinect skeleton, both relative to the head, to
``camdata_<time>.txt`` and ``kinectdata_<time>.txt`` (a ``%<model file>``
header line, then one line of 8 joints x (x, y, z) per frame), and
``bin/runSamplingTest.sh`` repeats a recorded run 20 times.

Here a synthetic upper body stands in for the Kinect skeleton.  Each arm's
pose follows its own GMM's generative model, the one the filter assumes
(my_gmm.cpp:51-74): a component drawn from the mixture weights and held for
a geometric number of frames, x_t = g_k x_{t-1} + (1 - g_k) mu_k + noise of
covariance (1 - g_k^2) Sigma_k, the 22-dim pose e = pca_proj^T x + pca_mean,
and 3-D joints by the pinhole back-projection of get3Dpose (pfPose.cpp:93-127).
The likelihood image carries a blob at each arm's projected hand plus clutter
and noise; the face ROI is centred on the mean projected head.  The tracker
runs the whole sequence on the GPU (include/pf_tracker.h) and its TF rows are
compared with the truth's rows in the same frame convention.

    python -m mkfbodytracker_pdaf_amd.evaluate --frames 300 --particles 500 --repeats 20 --out DIR
"""
from __future__ import annotations

import argparse
import json
import os
import time

import numpy as np

JOINTS = ("Head", "Left_Hand", "Right_Hand", "Left_Elbow", "Right_Elbow", "Right_Shoulder", "Left_Shoulder", "Neck")
FACE_W, FACE_H = 40, 48


def _read_arm(path):
    from . import capi
    r = {k: capi.yml_read_matrix(path, k) for k in ("means", "covs", "weights", "gamma", "pca_proj", "pca_mean")}
    K, d = r["means"].shape
    r["covs"] = r["covs"].reshape(K, d, d)
    r["weights"] = r["weights"].ravel() / r["weights"].sum()
    r["gamma"] = r["gamma"].ravel()
    r["pca_mean"] = r["pca_mean"].ravel()
    return r


def rpy(roll, pitch, yaw):
    """R3(yaw) R2(pitch) R1(roll) (pfPose.cpp:84-91)."""
    cr, sr, cp, sp, cy, sy = np.cos(roll), np.sin(roll), np.cos(pitch), np.sin(pitch), np.cos(yaw), np.sin(yaw)
    R1 = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    R2 = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    R3 = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return R3 @ R2 @ R1


def backproject(e, camera=(525.0, 525.0, 319.5, 239.5)):
    """3 x 5 joints (hand, elbow, shoulder, head, neck) of a 22-dim pose:
    R^T ((d (u - cx)/fx, d (v - cy)/fy, d) - t), R = rpy(e16, e17, e15)."""
    fx, fy, cx, cy = camera
    R = rpy(e[16], e[17], e[15])
    out = np.empty((3, 5))
    for k in range(5):
        u, v, d = e[3 * k:3 * k + 3]
        out[:, k] = R.T @ (np.array([d * (u - cx) / fx, d * (v - cy) / fy, d]) - e[18:21])
    return out


def tf_rows(p1, p2):
    """Joints relative to the head in the TF convention of publishTFtree
    (pfPose.cpp:129-171): a camera-frame difference v is published as
    (v.x, v.z, -v.y); head and neck are the means of the arms' columns;
    arm 1 is Right_*, arm 2 Left_*.  8 x 3, JOINTS order."""
    head = 0.5 * (p1[:, 3] + p2[:, 3])
    neck = 0.5 * (p1[:, 4] + p2[:, 4])
    m = lambda v: np.array([v[0], v[2], -v[1]])
    return np.array([np.zeros(3), m(p2[:, 0] - head), m(p1[:, 0] - head), m(p2[:, 1] - head),
                     m(p1[:, 1] - head), m(p1[:, 2] - head), m(p2[:, 2] - head), m(neck - head)])


def simulate_truth(arms, T, seed, hold=30.0):
    """Pose sequences of both arms from their GMMs: e (T x 2 x 22) and 3-D
    joints (T x 2 x 3 x 5)."""
    rng = np.random.default_rng(seed)
    E = np.empty((T, 2, arms[0]["pca_mean"].size))
    P3 = np.empty((T, 2, 3, 5))
    for a, m in enumerate(arms):
        K, d = m["means"].shape
        L = [np.linalg.cholesky(m["covs"][k] + 1e-12 * np.eye(d)) for k in range(K)]
        k = rng.choice(K, p=m["weights"])
        x = m["means"][k].copy()
        for t in range(T):
            if rng.random() < 1.0 / hold:
                k = rng.choice(K, p=m["weights"])
            g = m["gamma"][k]
            x = g * x + (1 - g) * m["means"][k] + np.sqrt(max(1 - g * g, 0.0)) * (L[k] @ rng.standard_normal(d))
            E[t, a] = m["pca_proj"].T @ x + m["pca_mean"]
            P3[t, a] = backproject(E[t, a])
    return E, P3


def render(E, seed, rows=480, cols=640):
    """Likelihood images with blobs at both projected hands and the face
    ROIs centred on the mean projected head (T x rows x cols uint8, T x 4)."""
    from .synthetic import likelihood_image
    T = E.shape[0]
    imgs = np.empty((T, rows, cols), np.uint8)
    rois = np.empty((T, 4), np.int32)
    for t in range(T):
        hands = [(E[t, a, 0], E[t, a, 1]) for a in range(2)]
        imgs[t] = likelihood_image(hands, rows, cols, seed=seed * 100003 + t)
        hu = 0.5 * (E[t, 0, 9] + E[t, 1, 9])
        hv = 0.5 * (E[t, 0, 10] + E[t, 1, 10])
        rois[t] = (int(round(hu - FACE_W / 2)), int(round(hv - FACE_H / 2)), FACE_W, FACE_H)
    return imgs, rois


def write_logs(out_dir, model_name, cam_rows, kinect_rows, stamp=None):
    """The two files of extractTFdata (extractTFdata.cpp:24-37, 87-121):
    '%' + model name, then per frame ' x y z' for the 8 joints."""
    stamp = stamp or time.strftime("%Y-%m-%d-%H-%M-%S")
    paths = []
    for name, rows in (("kinectdata", kinect_rows), ("camdata", cam_rows)):
        p = os.path.join(out_dir, f"{name}_{stamp}.txt")
        with open(p, "w") as f:
            f.write("%" + model_name + "\n")
            for r in np.asarray(rows).reshape(len(rows), 24):
                f.write("".join(" %.17g" % v for v in r) + "\n")
        paths.append(p)
    return paths


def read_log(path):
    with open(path) as f:
        head = f.readline().rstrip("\n")
        rows = np.array([[float(v) for v in line.split()] for line in f if line.strip()])
    return head[1:], rows.reshape(-1, 8, 3)


def joint_errors(cam_rows, kinect_rows):
    """Mean Euclidean error per joint (m) over the frames, Head excluded."""
    d = np.linalg.norm(np.asarray(cam_rows) - np.asarray(kinect_rows), axis=2)
    return {JOINTS[j]: float(d[:, j].mean()) for j in range(1, 8)}


def run_once(left, right, N, T, seed, graph=True):
    """One tracked sequence; returns dict(cam, kinect, pose2d, lml, ms)."""
    from .devmem import DeviceArray
    from .tracker import PoseTracker
    arms = [_read_arm(left), _read_arm(right)]
    E, P3 = simulate_truth(arms, T, seed)
    imgs, rois = render(E, seed)
    kinect = np.array([tf_rows(P3[t, 0], P3[t, 1]) for t in range(T)])
    tr = PoseTracker.from_yml(left, right, n_particles=N, seed=seed)
    d_img, d_roi = DeviceArray(imgs), DeviceArray(rois)
    d_lml = DeviceArray.empty(2 * T, np.float64)
    d_tf = DeviceArray.empty(T * 24, np.float64)
    d_p2 = DeviceArray.empty(T * 16, np.float64)
    tr.run_dev(T, d_img.ptr, d_roi.ptr, d_lml.ptr, None, None, d_tf.ptr, d_p2.ptr, graph=graph)
    tr.synchronize()
    ms = tr.last_run_ms()
    tr.close()
    cam = d_tf.get().reshape(T, 8, 3)
    p2 = d_p2.get().reshape(T, 8, 2)
    hand_px = float(np.mean([np.linalg.norm(p2[:, a] - E[:, a, 0:2], axis=1).mean() for a in range(2)]))
    return dict(cam=cam, kinect=kinect, pose2d=p2, lml=d_lml.get().reshape(2, T), ms=ms, hand_px=hand_px)


def main(argv=None):
    here = os.path.dirname(os.path.abspath(__file__))
    golden = os.path.join(here, "..", "tests", "golden")
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--left", default=os.path.join(golden, "data13D_PCA_100000_15_12.yml"))
    ap.add_argument("--right", default=os.path.join(golden, "data23D_PCA_100000_15_12.yml"))
    ap.add_argument("--particles", type=int, default=500)
    ap.add_argument("--frames", type=int, default=300)
    ap.add_argument("--repeats", type=int, default=20)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=None, help="directory for the camdata/kinectdata logs (none if omitted)")
    args = ap.parse_args(argv)
    summary = []
    for r in range(args.repeats):
        res = run_once(args.left, args.right, args.particles, args.frames, args.seed + r)
        err = joint_errors(res["cam"], res["kinect"])
        if args.out:
            os.makedirs(args.out, exist_ok=True)
            write_logs(args.out, os.path.basename(args.left), res["cam"], res["kinect"],
                       stamp=f"run{r:02d}")
        summary.append(dict(run=r, joint_error_m=err, mean_error_m=float(np.mean(list(err.values()))),
                            hand_error_px=res["hand_px"], ms_per_frame=res["ms"] / args.frames))
        print(json.dumps(summary[-1]), flush=True)
    print(json.dumps({"runs": args.repeats, "frames": args.frames, "particles": args.particles,
                      "mean_error_m": float(np.mean([s["mean_error_m"] for s in summary])),
                      "hand_error_px": float(np.mean([s["hand_error_px"] for s in summary]))}))


if __name__ == "__main__":
    main()
