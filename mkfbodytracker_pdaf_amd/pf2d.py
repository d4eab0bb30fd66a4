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

    def estimate(self):
        out = np.empty(self.d)
        check(self.lib.pf2d_estimate(self.h, ptr(out)))
        return out


class my_gmm:
    """pf2D my_gmm (src/pf2D.h:12-23): loadGaussian(mean, sigma, weight)."""

    def __init__(self, pf: "ParticleFilter"):
        self._pf = pf
        self.mean, self.sigma, self.weight = [], [], []
        self.N = 0

    def loadGaussian(self, mean, sigma, weight):
        self.mean.append(f64(mean).ravel()); self.sigma.append(f64(sigma)); self.weight.append(float(weight))
        self.N += 1
        self._pf._h.load_gmm(np.stack(self.mean), np.stack(self.sigma), self.weight)


class ParticleFilter:
    """pf2D ParticleFilter(numParticles, numDims, side1) (src/pf2D.cpp:44-71) on the GPU."""

    def __init__(self, numParticles: int, numDims: int = 8, side1: bool = False, seed: int = 0, device: int = 0):
        self._h = Pf2d(numParticles, numDims, int(bool(side1)), device)
        self.gmm = my_gmm(self)
        self._seed, self._draws = int(seed), 0
        self.N, self.d = int(numParticles), int(numDims)
        self._h.init(self._uniforms(self.N * self.d).reshape(self.N, self.d))

    def _u32(self, n):
        out = capi.philox_u32(self._seed, 0, self._draws, n)
        self._draws += n
        return out

    def _uniforms(self, n):
        return self._u32(n).astype(np.float64) * 2.0 ** -32

    def update(self, measurement):
        """ParticleFilter::update (pf2D.cpp:148-210); measurement is 2 x 2."""
        u = int(self._u32(1)[0])
        normals = normals_from_u32(self._u32(8 * self.N))[: 8 * self.N].reshape(self.N, 8)
        reinit = self._uniforms(self.N * self.d).reshape(self.N, self.d)
        st, _ = self._h.update(measurement, u, normals, reinit)
        return st

    def getEstimator(self):
        """ParticleFilter::getEstimator (pf2D.cpp:79-87): 1 x d."""
        return self._h.estimate().reshape(1, -1)
