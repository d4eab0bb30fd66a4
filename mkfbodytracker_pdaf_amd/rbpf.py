"""Python mirror of the reference's RBPF classes on top of the C ABI.

Reference interface (src/pf2DRao.h:13-31, src/my_gmm.h:9-33,
src/KF_model.h:8-16): ``ParticleFilter(nParticles)`` with ``update``,
``getEstimator``, ``getSamples``, ``getSampleProb``, ``resample`` and a public
``gmm`` (``loadGaussian``, ``resetTracker``, ``tracks``, ``nParticles``), and
``KF_model`` with ``predict`` / ``update``.  Same names, same argument meaning;
matrices are numpy arrays laid out like the reference's cv::Mat.  Every
numeric call runs on the GPU through libmkfbt.so.  Randomness: the reference
draws from cv::RNG inside each call (src/pf2DRao.cpp:96,179-197); here each
filter owns a seeded Philox4x32-10 stream, and every call also accepts the
uniforms explicitly.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi
from .capi import check, f64, i32, ptr

REFERENCE_SELECTION = (9, 10, 0, 1, 12, 13)  # my_gmm.cpp:61-67 (head u,v; hand u,v; neck u,v)


class Rbpf:
    """Owner of one C-ABI handle (one filter resident on one GPU)."""

    def __init__(self, n_particles: int, device: int = 0, **cfg):
        self.lib = capi.lib()
        c = capi.default_config(n_particles=int(n_particles), device=int(device), **cfg)
        h = C.c_void_p()
        check(self.lib.pf_create(C.byref(h), C.byref(c)))
        self.h = h
        self.cfg = c
        self.N = int(n_particles)
        self.d = self.m = self.K = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.pf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- model
    def load_model(self, means, covs, weights, gamma, pca_proj, pca_mean, sel=None, R=None):
        means = f64(means); covs = f64(covs); pca_proj = f64(pca_proj)
        K, d = means.shape
        npose = pca_proj.shape[1]
        sel_a = None if sel is None else i32(sel)
        m = 6 if sel is None else sel_a.size
        R_a = None if R is None else f64(R)
        check(self.lib.pf_load_model(self.h, K, d, npose, m, ptr(sel_a), ptr(means), ptr(covs),
                                     ptr(f64(weights)), ptr(f64(gamma)), ptr(pca_proj),
                                     ptr(f64(pca_mean)), ptr(R_a)))
        self.K, self.d, self.m = K, d, m

    def load_model_yml(self, path, gamma_path=None, sel=None, R=None):
        sel_a = None if sel is None else i32(sel)
        m = 6 if sel is None else sel_a.size
        R_a = None if R is None else f64(R)
        check(self.lib.pf_load_model_yml(self.h, path.encode(),
                                         None if gamma_path is None else gamma_path.encode(),
                                         m, ptr(sel_a), ptr(R_a)))
        self.m = m
        self.d = capi.yml_read_matrix(path, "means").shape[1]
        self.K = capi.yml_read_matrix(path, "means").shape[0]

    def set_joint_clutter(self, clutter):
        check(self.lib.pf_set_joint_clutter(self.h, ptr(i32(clutter))))

    # ------------------------------------------------------------ tracks
    def reset(self, bins):
        check(self.lib.pf_reset(self.h, ptr(i32(bins))))

    def set_tracks(self, x, P):
        check(self.lib.pf_set_tracks(self.h, ptr(f64(x)), ptr(f64(P))))

    def get_tracks(self):
        x = np.empty((self.N, self.d)); P = np.empty((self.N, self.d, self.d))
        check(self.lib.pf_get_tracks(self.h, ptr(x), ptr(P)))
        return x, P

    # -------------------------------------------------------------- step
    def step(self, Z, u_comp: int, u_res: int):
        """Reference-mode ParticleFilter::update; returns (status, dict of outputs)."""
        lw = np.empty(self.N, np.float32); anc = np.empty(self.N, np.int32)
        comp = np.empty(self.N, np.int32); lml = C.c_double(0.0)
        st = check(self.lib.pf_step(self.h, ptr(f64(Z)), int(u_comp), int(u_res), ptr(lw), ptr(anc),
                                    ptr(comp), C.byref(lml)))
        return st, dict(lw=lw, anc=anc, comp=comp, lml=lml.value)

    def step_candidates(self, count, cand, u_comp: int, u_res: int):
        lw = np.empty(self.N, np.float32); anc = np.empty(self.N, np.int32)
        comp = np.empty(self.N, np.int32); lml = C.c_double(0.0)
        cand = f64(cand)
        st = check(self.lib.pf_step_candidates(self.h, ptr(i32(count)), ptr(cand), int(u_comp),
                                               int(u_res), ptr(lw), ptr(anc), ptr(comp), C.byref(lml)))
        return st, dict(lw=lw, anc=anc, comp=comp, lml=lml.value)

    def estimate(self):
        out = np.empty(self.d)
        check(self.lib.pf_estimate(self.h, ptr(out)))
        return out

    def resample(self, w, N: int, u: int, log: bool = False):
        w = f64(w)
        out = np.empty(int(N), np.int32)
        fn = self.lib.pf_resample_log if log else self.lib.pf_resample
        st = check(fn(self.h, ptr(w), w.size, int(N), int(u), ptr(out)))
        return out, st

    def get_samples(self, Hp, M, n: int, scale: float, normals):
        Hp = f64(Hp); out = np.empty((2, int(n)))
        check(self.lib.pf_get_samples(self.h, ptr(Hp), ptr(f64(M).ravel()), Hp.shape[0], int(n),
                                      float(scale), ptr(f64(normals)), ptr(out)))
        return out

    def get_sample_prob(self, Hp, M, in1, in2, scale: float):
        Hp = f64(Hp); in1 = f64(in1); in2 = f64(in2)
        n = in1.shape[1]
        if in2.shape[1] != n:
            raise ValueError("in1 and in2 must have the same number of columns")
        w1 = np.empty(n); w2 = np.empty(n)
        check(self.lib.pf_get_sample_prob(self.h, ptr(Hp), ptr(f64(M).ravel()), Hp.shape[0], ptr(in1),
                                          ptr(in2), n, float(scale), ptr(w1), ptr(w2)))
        return w1, w2

    def kf_predict(self, F, B, Q, x, P):
        x = f64(x).copy(); P = f64(P).copy()
        n = 1 if x.ndim == 1 else x.shape[0]
        d = x.shape[-1]
        check(self.lib.pf_kf_predict(self.h, n, d, ptr(f64(F)), ptr(f64(B).ravel()), ptr(f64(Q)), ptr(x), ptr(P)))
        return x, P

    def kf_update(self, H, BH, R, z, x, P):
        x = f64(x).copy(); P = f64(P).copy(); H = f64(H)
        n = 1 if x.ndim == 1 else x.shape[0]
        check(self.lib.pf_kf_update(self.h, n, x.shape[-1], H.shape[0], ptr(H), ptr(f64(BH).ravel()),
                                    ptr(f64(R)), ptr(f64(z)), ptr(x), ptr(P)))
        return x, P

    # -------------------------------------------- device-resident sequences
    def run_candidates_dev(self, T, count_ptr, cand_ptr, uni_ptr, lml_ptr, flags=1):
        """T device-resident frames (pf_run_candidates_dev).  flags: 1 replays
        captured graphs of frames, 2 times every launch (no graph), 4 times
        the whole run; last_kernel_time() reads the timing."""
        check(self.lib.pf_run_candidates_dev(self.h, int(T), count_ptr, cand_ptr, uni_ptr, lml_ptr, int(flags)))

    def run_ref_dev(self, T, Z_ptr, uni_ptr, lml_ptr, flags=1):
        """As run_candidates_dev with one measurement column per particle."""
        check(self.lib.pf_run_ref_dev(self.h, int(T), Z_ptr, uni_ptr, lml_ptr, int(flags)))

    def synchronize(self) -> int:
        """Waits for the handle's work; returns its warning bits (capi.PF_WARN_*)."""
        return check(self.lib.pf_synchronize(self.h))

    def stream(self):
        s = C.c_void_p()
        check(self.lib.pf_get_stream(self.h, C.byref(s)))
        return s.value

    def last_kernel_time(self):
        a, b, n = C.c_double(0), C.c_double(0), C.c_int32(0)
        check(self.lib.pf_last_run_kernel_time(self.h, C.byref(a), C.byref(b), C.byref(n)))
        return a.value, b.value, n.value

    def last_graph_frames(self) -> int:
        """Frames of the last device-resident run replayed from graphs."""
        n = C.c_int32(0)
        check(self.lib.pf_last_run_graph_frames(self.h, C.byref(n)))
        return n.value

    @property
    def state_layout(self) -> int:
        """Device layout of the state in use: capi.PF_LAYOUT_ROWS or
        PF_LAYOUT_RECORDS (pf_get_state_layout)."""
        n = C.c_int32(0)
        check(self.lib.pf_get_state_layout(self.h, C.byref(n)))
        return n.value

    @state_layout.setter
    def state_layout(self, layout: int) -> None:
        """PF_LAYOUT_AUTO, PF_LAYOUT_ROWS or PF_LAYOUT_RECORDS; the current
        tracks are converted (pf_set_state_layout)."""
        check(self.lib.pf_set_state_layout(self.h, int(layout)))


# ============================================ reference-named mirror classes
class KF_model:
    """KF_model (src/KF_model.h:8-16): public Q, R, F, B, H, BH; predict/update on the GPU."""

    _owner: Rbpf | None = None

    def __init__(self):
        self.Q = self.R = self.F = self.B = self.H = self.BH = None

    @classmethod
    def _handle(cls):
        if cls._owner is None:
            cls._owner = Rbpf(1)
        return cls._owner

    def predict(self, state, cov):
        """Returns (state, cov) after x = F x + B, P = F P F^T + Q (KF_model.cpp:11-15)."""
        return self._handle().kf_predict(self.F, self.B, self.Q, np.ravel(state), cov)

    def update(self, measurement, state, cov):
        """Returns (state, cov) after the Kalman update (KF_model.cpp:17-25)."""
        return self._handle().kf_update(self.H, self.BH, self.R, np.ravel(measurement), np.ravel(state), cov)


@dataclass
class state_params:
    """One track {state d x 1, cov d x d, weight} (src/my_gmm.h:9-18)."""
    state: np.ndarray
    cov: np.ndarray
    weight: float


class my_gmm:
    """my_gmm (src/my_gmm.h:20-33).  Components accumulate on the host and are
    uploaded to the filter's device tables on first use."""

    def __init__(self, pf: "ParticleFilter"):
        self._pf = pf
        self.mean, self.cov, self.weight, self.KFtracker = [], [], [], []
        self._gamma = []
        self._proj = self._pmean = None
        self.sel = REFERENCE_SELECTION
        self.R = None
        self._dirty = True

    @property
    def nParticles(self):
        return self._pf.N

    def loadGaussian(self, u, s, H, m, w, g):
        """my_gmm::loadGaussian (my_gmm.cpp:45-75): u 1 x d mean, s d x d cov,
        H = pca_proj (d x npose), m = pca_mean (1 x npose), weight w, gamma g."""
        u = f64(u).ravel(); s = f64(s)
        self.mean.append(u); self.cov.append(s); self.weight.append(float(w)); self._gamma.append(float(g))
        self._proj = f64(H); self._pmean = f64(m).ravel()
        kf = KF_model()
        d = u.size
        R = self.R if self.R is not None else 100.0 * np.eye(len(self.sel))
        kf.Q = (1.0 - g * g) * s
        kf.R = R
        kf.F = g * np.eye(d)
        kf.B = (1.0 - g) * u
        H1 = np.zeros((len(self.sel), self._proj.shape[1]))
        for r, c in enumerate(self.sel):
            H1[r, c] = 1.0
        kf.H = H1 @ self._proj.T
        kf.BH = H1 @ self._pmean
        self.KFtracker.append(kf)
        self._dirty = True

    def _upload(self):
        if not self._dirty:
            return
        self._pf._rb.load_model(np.stack(self.mean), np.concatenate(self.cov, axis=0), self.weight,
                                self._gamma, self._proj, self._pmean, sel=self.sel, R=self.R)
        self._dirty = False

    def resetTracker(self, bins):
        """my_gmm::resetTracker (my_gmm.cpp:30-42)."""
        self._upload()
        self._pf._rb.reset(bins)

    @property
    def tracks(self):
        x, P = self._pf._rb.get_tracks()
        w = 1.0 / self.nParticles
        return [state_params(x[j].reshape(-1, 1), P[j], w) for j in range(self.nParticles)]


class ParticleFilter:
    """ParticleFilter (src/pf2DRao.h:13-31) on the GPU."""

    def __init__(self, nParticles: int, seed: int = 0, device: int = 0, **cfg):
        self.N = int(nParticles)
        self._rb = Rbpf(self.N, device=device, **cfg)
        self.gmm = my_gmm(self)
        self._seed = int(seed)
        self._draws = 0

    def _u32(self, n=1):
        out = capi.philox_u32(self._seed, 0, self._draws, n)
        self._draws += n
        return out

    def update(self, measurement, uniforms=None):
        """ParticleFilter::update (pf2DRao.cpp:125-158); measurement m x N.
        Returns the resampler status (1 = degenerate uniform fallback)."""
        self.gmm._upload()
        u = self._u32(2) if uniforms is None else uniforms
        st, out = self._rb.step(measurement, int(u[0]), int(u[1]))
        self.last = out
        return st

    def getEstimator(self):
        """ParticleFilter::getEstimator (pf2DRao.cpp:23-31): d x 1."""
        return self._rb.estimate().reshape(-1, 1)

    def getSamples(self, H, M, N, scale, normals=None):
        """ParticleFilter::getSamples (pf2DRao.cpp:85-103): 2 x N proposals."""
        if normals is None:
            normals = normals_from_u32(self._u32(4 * int(N)))[: 2 * int(N)].reshape(2, int(N))
        return self._rb.get_samples(H, M, int(N), float(scale), normals)

    def getSampleProb(self, H, M, input1, input2, scale):
        """ParticleFilter::getSampleProb (pf2DRao.cpp:105-122) -> (weight1, weight2)."""
        w1, w2 = self._rb.get_sample_prob(H, M, input1, input2, float(scale))
        return list(w1), list(w2)

    def resample(self, weights, N, u=None):
        """ParticleFilter::resample (pf2DRao.cpp:175-210) -> list of N indices."""
        uu = int(self._u32(1)[0]) if u is None else int(u)
        out, _ = self._rb.resample(weights, int(N), uu)
        return list(out)


def normals_from_u32(u):
    """Box-Muller normals from pairs of 32-bit uniforms (host helper)."""
    u = np.asarray(u, np.uint64)
    a = (u[0::2].astype(np.float64) + 1.0) * 2.0 ** -32
    b = u[1::2].astype(np.float64) * 2.0 ** -32
    r = np.sqrt(-2.0 * np.log(a))
    return np.concatenate([r * np.cos(2 * np.pi * b), r * np.sin(2 * np.pi * b)])
