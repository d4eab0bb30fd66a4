"""Seeded synthetic workloads for the particle-filter hot path.

Model (the reference's structure, src/my_gmm.cpp:45-75, with synthetic
parameters): K components over a d-dim PCA pose state; means ~ N(0, 1);
covariances diag(a_k) + u u^T + v v^T (component 0 has condition number
>= 1e6, like the shipped models' 4.9e6); weights ~ Dirichlet(1); mean
reversion rates r_k ~ U(0.05, 0.3) with gamma_k = 1 - r_k.

Measurement geometry: the pose vector stores (u, v, depth) per joint, the
layout of the reference's PCA space (src/my_gmm.cpp:59-60,
src/pfPose.cpp:114-120).  pca_mean is a canonical upper-body skeleton in
metres projected by the pinhole camera fx = fy = 525, cx = 319.5, cy = 239.5
(the Kinect intrinsics of src/pfPose.cpp:101) on a 640 x 480 image; pca_proj
is a random 3-D displacement basis mapped through the pinhole Jacobian at the
canonical pose, so one latent unit moves a joint by ~15 px.  The filter's
measurement matrix H selects (u, v) of the measured joints; detections carry
sigma = 5 px noise; each cluttered joint has 1 true detection plus
Poisson(c - 1) clutter uniform over the image (truncated to the capacity).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

FX = FY = 525.0
CX, CY = 319.5, 239.5
WIDTH, HEIGHT = 640, 480

# canonical joints (metres, camera frame)
_JOINTS = {
    "head": (0.0, -0.45, 2.0),
    "left_hand": (-0.35, 0.15, 1.8),
    "right_hand": (0.35, 0.15, 1.8),
    "left_elbow": (-0.30, -0.10, 1.9),
    "right_elbow": (0.30, -0.10, 1.9),
    "neck": (0.0, -0.25, 2.0),
}


def project(X):
    """Pinhole projection of (..., 3) points -> (..., 3) (u, v, depth)."""
    X = np.asarray(X, np.float64)
    return np.stack([FX * X[..., 0] / X[..., 2] + CX, FY * X[..., 1] / X[..., 2] + CY, X[..., 2]], -1)


@dataclass
class Model:
    means: np.ndarray      # K x d
    covs: np.ndarray       # (K*d) x d
    weights: np.ndarray    # K
    gamma: np.ndarray      # K
    pca_proj: np.ndarray   # d x npose
    pca_mean: np.ndarray   # npose
    sel: np.ndarray        # m indices into the pose vector
    clutter: np.ndarray    # J flags: 1 = PDAF joint
    R: np.ndarray          # m x m
    joints: list = field(default_factory=list)

    @property
    def K(self):
        return self.means.shape[0]

    @property
    def d(self):
        return self.means.shape[1]

    @property
    def m(self):
        return self.sel.size

    @property
    def Sigma(self):
        return self.covs.reshape(self.K, self.d, self.d)

    def H(self):
        return self.pca_proj[:, self.sel].T.copy()

    def BH(self):
        return self.pca_mean[self.sel].copy()


def make_model(seed: int, K: int = 15, d: int = 12, joints=("head", "left_hand", "right_hand"),
               clutter=(0, 1, 1), r_scale: float = 100.0, cond: float = 2e6) -> Model:
    rng = np.random.default_rng(seed)
    means = rng.standard_normal((K, d))
    covs = np.empty((K, d, d))
    for k in range(K):
        if k == 0:
            a = np.logspace(np.log10(4.0 / cond), np.log10(4.0), d)
            rng.shuffle(a)
        else:
            a = rng.uniform(0.05, 2.0, d)
        u = 0.5 * rng.standard_normal(d)
        v = 0.5 * rng.standard_normal(d)
        S = np.diag(a)
        if k != 0:
            S = S + np.outer(u, u) + np.outer(v, v)
        covs[k] = 0.5 * (S + S.T)
    weights = rng.dirichlet(np.ones(K))
    gamma = 1.0 - rng.uniform(0.05, 0.3, K)
    names = list(joints)
    J = len(names)
    X0 = np.array([_JOINTS[n] for n in names])                 # J x 3
    pose_mean = project(X0).reshape(-1)                         # (u, v, depth) per joint
    B3 = 0.03 * rng.standard_normal((d, J, 3))                  # metres per latent unit
    proj = np.empty((d, J, 3))
    for j in range(J):
        X, Y, Z = X0[j]
        Jac = np.array([[FX / Z, 0.0, -FX * X / Z ** 2],
                        [0.0, FY / Z, -FY * Y / Z ** 2],
                        [0.0, 0.0, 1.0]])
        proj[:, j, :] = B3[:, j, :] @ Jac.T
    pca_proj = proj.reshape(d, 3 * J)
    sel = np.array([3 * j + c for j in range(J) for c in (0, 1)], np.int32)
    R = r_scale * np.eye(2 * J)
    return Model(means, covs.reshape(K * d, d), weights, gamma, pca_proj, pose_mean, sel,
                 np.asarray(clutter, np.int32), R, names)


@dataclass
class Sequence:
    x: np.ndarray          # T x d true states
    comp: np.ndarray       # T true components
    z: np.ndarray          # T x m true noisy measurements
    count: np.ndarray      # T x J detections per joint
    cand: np.ndarray       # T x J x cap x 2
    uniforms: np.ndarray   # T x 2 uint32 (component draw, track resampling)


def simulate(model: Model, T: int, seed: int, cand_per_joint: int = 4, cap: int = 8,
             noise_px: float = 5.0) -> Sequence:
    rng = np.random.default_rng(seed + 1000)
    K, d, m = model.K, model.d, model.m
    J = m // 2
    Sig = model.Sigma
    L = np.linalg.cholesky(Sig)
    k = rng.choice(K, p=model.weights)
    x = model.means[k] + L[k] @ rng.standard_normal(d)
    H, BH = model.H(), model.BH()
    xs = np.empty((T, d)); comps = np.empty(T, np.int32); zs = np.empty((T, m))
    count = np.zeros((T, J), np.int32)
    cand = np.zeros((T, J, cap, 2))
    for t in range(T):
        k = rng.choice(K, p=model.weights)
        g = model.gamma[k]
        x = g * x + (1 - g) * model.means[k] + np.sqrt(1 - g * g) * (L[k] @ rng.standard_normal(d))
        z = H @ x + BH + noise_px * rng.standard_normal(m)
        xs[t], comps[t], zs[t] = x, k, z
        for b in range(J):
            if model.clutter[b]:
                nclut = rng.poisson(cand_per_joint - 1)
                pts = [z[2 * b:2 * b + 2]]
                for _ in range(nclut):
                    pts.append(np.array([rng.uniform(0, WIDTH), rng.uniform(0, HEIGHT)]))
                pts = np.array(pts)[:cap]
                rng.shuffle(pts)
                count[t, b] = len(pts)
                cand[t, b, :len(pts)] = pts
            else:
                count[t, b] = 1
                cand[t, b, 0] = z[2 * b:2 * b + 2]
    from .capi import philox_u32
    uni = philox_u32(seed, 1, 0, 2 * T).reshape(T, 2)
    return Sequence(xs, comps, zs, count, cand, uni)


def simulate_fast(model: Model, T: int, seed: int, cand_per_joint: int = 4, cap: int = 8,
                  noise_px: float = 5.0) -> Sequence:
    """Same generative model as ``simulate`` with all random draws made in
    bulk (a different random stream): for long or many sequences."""
    rng = np.random.default_rng(seed + 2000)
    K, d, m = model.K, model.d, model.m
    J = m // 2
    L = np.linalg.cholesky(model.Sigma)
    comps = rng.choice(K, size=T, p=model.weights).astype(np.int32)
    k0 = rng.choice(K, p=model.weights)
    x = model.means[k0] + L[k0] @ rng.standard_normal(d)
    g = model.gamma[comps]
    eps = rng.standard_normal((T, d))
    drift = (1 - g)[:, None] * model.means[comps] + np.sqrt(1 - g * g)[:, None] * np.einsum("tij,tj->ti", L[comps], eps)
    xs = np.empty((T, d))
    for t in range(T):
        x = g[t] * x + drift[t]
        xs[t] = x
    H, BH = model.H(), model.BH()
    zs = xs @ H.T + BH + noise_px * rng.standard_normal((T, m))
    count = np.zeros((T, J), np.int32)
    cand = np.zeros((T, J, cap, 2))
    for b in range(J):
        zb = zs[:, 2 * b:2 * b + 2]
        if model.clutter[b]:
            n = np.minimum(1 + rng.poisson(cand_per_joint - 1, T), cap)
            pts = np.stack([rng.uniform(0, WIDTH, (T, cap)), rng.uniform(0, HEIGHT, (T, cap))], -1)
            slot = (rng.random(T) * n).astype(np.int64)
            pts[np.arange(T), slot] = zb
            pts[np.arange(cap)[None, :] >= n[:, None]] = 0.0
            count[:, b] = n
            cand[:, b] = pts
        else:
            count[:, b] = 1
            cand[:, b, 0] = zb
    from .capi import philox_u32
    uni = philox_u32(seed, 1, 0, 2 * T).reshape(T, 2)
    return Sequence(xs, comps, zs, count, cand, uni)


def init_bins(model: Model, N: int, seed: int, resample):
    """Initial components: resample(gmm.weight, N) then resetTracker
    (src/pfPose.cpp:68-71).  ``resample(w, N, U)`` is the resampler to use."""
    from .capi import philox_u32
    U = int(philox_u32(seed, 2, 0, 1)[0])
    return resample(model.weights, N, U)


def per_particle_measurements(z, N: int, seed: int, jitter_px: float = 5.0):
    """Reference-mode m x N measurement matrix: column j = z + per-particle
    jitter (the reference fills each column from a resampled proposal,
    src/pfPose.cpp:303-323)."""
    rng = np.random.default_rng(seed)
    return z[:, None] + jitter_px * rng.standard_normal((z.size, N))


# ------------------------------------------------- pose tracker inputs
# Face ROI of the shipped models' mean pose (head at about (324, 75) px).
FACE_ROI = (304, 50, 40, 48)
HANDS = ((388.0, 281.0), (264.0, 268.0))


def likelihood_image(points, rows: int = 480, cols: int = 640, seed: int = 0, amp: float = 230.0,
                     sigma: float = 7.0, clutter: int = 2) -> np.ndarray:
    """uint8 hand-likelihood image: Gaussian blobs at the (x, y) points and at
    `clutter` random places, plus sparse low-level noise."""
    rng = np.random.default_rng(seed)
    pts = [tuple(p) for p in points] + [(rng.uniform(0, cols), rng.uniform(0, rows)) for _ in range(clutter)]
    yy = np.arange(rows)[:, None].astype(np.float64)
    xx = np.arange(cols)[None, :].astype(np.float64)
    img = np.zeros((rows, cols))
    for x, y in pts:
        img += amp * np.exp(-((xx - x) ** 2 + (yy - y) ** 2) / (2.0 * sigma ** 2))
    noise = rng.random((rows, cols)) < 0.01
    img[noise] += rng.uniform(0, 60, int(noise.sum()))
    return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def tracker_sequence(T: int, seed: int, rows: int = 480, cols: int = 640):
    """T frames of (likelihood image, face ROI): both hands move on smooth
    random paths around their mean positions, the face ROI jitters by a pixel
    or two.  Returns (images T x rows x cols uint8, rois T x 4 int32)."""
    rng = np.random.default_rng(seed)
    imgs = np.empty((T, rows, cols), np.uint8)
    rois = np.empty((T, 4), np.int32)
    ph = rng.uniform(0, 2 * np.pi, 4)
    for t in range(T):
        hands = [(hx + 25.0 * np.sin(0.07 * t + ph[2 * i]), hy + 20.0 * np.cos(0.05 * t + ph[2 * i + 1]))
                 for i, (hx, hy) in enumerate(HANDS)]
        imgs[t] = likelihood_image(hands, rows, cols, seed=seed * 1000 + t)
        rois[t] = (FACE_ROI[0] + int(rng.integers(-2, 3)), FACE_ROI[1] + int(rng.integers(-2, 3)),
                   FACE_ROI[2], FACE_ROI[3])
    return imgs, rois
