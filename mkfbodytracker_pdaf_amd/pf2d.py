"""Python mirror of the reference's plain particle filter pf2D (src/pf2D.h:12-51)
on the C ABI of include/pf2d.h.  Same class and method names; randomness is
drawn from a seeded Philox4x32-10 stream (the reference uses rand() and
cv::randu / cv::randn, src/pf2D.cpp:64-68,96-98,255) or passed explicitly."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, f64, ptr
from .rbpf import normals_from_u32


class Pf2d:
    """Owner of one pf2D handle."""

    def __init__(self, n_particles: int, n_dims: int = 8, side: int = 0, device: int = 0):
        self.lib = capi.lib()
        h = C.c_void_p()
        check(self.lib.pf2d_create(C.byref(h), int(n_particles), int(n_dims), int(side), int(device)))
        self.h, self.N, self.d, self.side = h, int(n_particles), int(n_dims), int(side)

    def close(self):
        if getattr(self, "h", None):
            self.lib.pf2d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_gmm(self, means, covs, weights):
        means = f64(means)
        check(self.lib.pf2d_load_gmm(self.h, means.shape[0], ptr(means), ptr(f64(covs)), ptr(f64(weights))))

    def load_gmm_precision(self, means, sigma_i, det_s, weights):
        means = f64(means)
        check(self.lib.pf2d_load_gmm_precision(self.h, means.shape[0], ptr(means), ptr(f64(sigma_i)),
                                               ptr(f64(det_s)), ptr(f64(weights))))

    def init(self, unif):
        check(self.lib.pf2d_init(self.h, ptr(f64(unif))))

    def set_particles(self, particles, weights=None):
        check(self.lib.pf2d_set_particles(self.h, ptr(f64(particles)), ptr(None if weights is None else f64(weights))))

    def get_particles(self):
        p = np.empty((self.N, self.d)); w = np.empty(self.N)
        check(self.lib.pf2d_get_particles(self.h, ptr(p), ptr(w)))
        return p, w

    def update(self, meas, u: int, normals, reinit_unif):
        lw = np.empty(self.N, np.float32); anc = np.empty(self.N, np.int32)
        st = check(self.lib.pf2d_update(self.h, ptr(f64(meas)), int(u), ptr(f64(normals)), ptr(f64(reinit_unif)),
                                        ptr(lw), ptr(anc)))
        return st, dict(lw=lw, anc=anc)

    def predict(self, normals):
        check(self.lib.pf2d_predict(self.h, ptr(f64(normals))))

    def estimate(self):
        out = np.empty(self.d)
        check(self.lib.pf2d_estimate(self.h, ptr(out)))
        return out


class my_gmm:
    """pf2D my_gmm (src/pf2D.h:12-23): loadGaussian(mean, sigma, weight) keeps
    mean, sigma_i = sigma^-1 (Cholesky) and det_s = 1 / ((2 pi)^(d/2)
    sqrt(det sigma)) (pf2D.cpp:28-37); the filter's update reads these members
    as they are at the call."""

    def __init__(self):
        self.mean, self.sigma_i, self.det_s, self.weight = [], [], [], []
        self.N = 0

    def loadGaussian(self, mean, sigma, weight):
        sigma = f64(sigma)
        d = sigma.shape[0]
        try:
            L = np.linalg.cholesky(0.5 * (sigma + sigma.T))
        except np.linalg.LinAlgError:
            raise capi.PfError(capi.PF_ERR_NOT_SPD, "pf2D loadGaussian: sigma is not positive definite") from None
        Li = np.linalg.inv(L)
        self.N += 1
        self.mean.append(f64(mean).ravel())
        self.sigma_i.append(Li.T @ Li)
        self.det_s.append(float(np.exp(-0.5 * d * np.log(2.0 * np.pi) - np.log(np.diag(L)).sum())))
        self.weight.append(float(weight))

    def _arrays(self):
        K = self.N
        return (np.stack(self.mean[:K]), np.stack(self.sigma_i[:K]), np.array(self.det_s[:K], np.float64),
                np.array(self.weight[:K], np.float64))


class ParticleFilter:
    """pf2D ParticleFilter (src/pf2D.h:25-51) on the GPU.  ParticleFilter()
    has no particles (update, predict and getEstimator raise PF_ERR_STATE);
    ParticleFilter(numParticles, numDims, side1) is pf2D.cpp:44-71."""

    def __init__(self, numParticles: int | None = None, numDims: int = 8, side1: bool = False, seed: int = 0,
                 device: int = 0):
        self.gmm = my_gmm()
        self._seed, self._draws = int(seed), 0
        self._sent = None
        self._h = None
        if numParticles is None:
            self.N = self.d = 0
            return
        self._h = Pf2d(numParticles, numDims, int(bool(side1)), device)
        self.N, self.d = int(numParticles), int(numDims)
        self._h.init(self._uniforms(self.N * self.d).reshape(self.N, self.d))

    def _need(self, what):
        if self._h is None:
            raise capi.PfError(capi.PF_ERR_STATE, f"pf2D {what}: the filter has no particles")

    def _sync_gmm(self):
        if self.gmm.N <= 0:
            raise capi.PfError(capi.PF_ERR_STATE, "pf2D update: no GMM components loaded")
        arrs = self.gmm._arrays()
        if self._sent is None or any(a.shape != b.shape or not np.array_equal(a, b) for a, b in zip(arrs, self._sent)):
            self._h.load_gmm_precision(*arrs)
            self._sent = tuple(a.copy() for a in arrs)

    def _u32(self, n):
        out = capi.philox_u32(self._seed, 0, self._draws, n)
        self._draws += n
        return out

    def _uniforms(self, n):
        return self._u32(n).astype(np.float64) * 2.0 ** -32

    def _normals(self, n):
        return normals_from_u32(self._u32(2 * ((n + 1) // 2)))[:n]

    def predict(self):
        """ParticleFilter::predict (pf2D.cpp:90-102): x += 5 n, normals per
        particle for columns 2..5, 6, 7, 0, 1."""
        self._need("predict")
        self._h.predict(self._normals(8 * self.N).reshape(self.N, 8))

    def update(self, measurement):
        """ParticleFilter::update (pf2D.cpp:148-210); measurement is 2 x 2."""
        self._need("update")
        self._sync_gmm()
        u = int(self._u32(1)[0])
        normals = self._normals(8 * self.N).reshape(self.N, 8)
        reinit = self._uniforms(self.N * self.d).reshape(self.N, self.d)
        st, _ = self._h.update(measurement, u, normals, reinit)
        return st

    def getEstimator(self):
        """ParticleFilter::getEstimator (pf2D.cpp:79-87): 1 x d."""
        self._need("getEstimator")
        return self._h.estimate().reshape(1, -1)
