"""Many independent RBPF sequences on one GPU (C ABI ``include/pf_batch.h``).

Each sequence runs the candidate-mode step of ``Rbpf.step_candidates``
(src/pf2DRao.cpp:125-158 with the PDAF joint likelihood) with its own GMM,
particle count and length; the reference runs one ParticleFilter per arm,
one after the other (src/pfPose.cpp:325-326).  One persistent step kernel
and one weights kernel per frame serve every sequence still active at that
frame.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, f64, i32, ptr


class RbpfBatch:
    """S sequences, sequence s with ``n_particles[s]`` particles and
    ``n_frames[s]`` frames, all with K components, d = 12 and m rows."""

    def __init__(self, n_particles, n_frames, K: int, m: int, d: int = 12, device: int = 0, **cfg):
        self.lib = capi.lib()
        self.N = i32(n_particles)
        self.T = i32(n_frames)
        assert self.N.shape == self.T.shape
        self.S = int(self.N.size)
        self.K, self.d, self.m = int(K), int(d), int(m)
        self.J = self.m // 2
        c = capi.default_config(device=int(device), **cfg)
        self.cfg = c
        h = C.c_void_p()
        check(self.lib.pf_batch_create(C.byref(h), self.S, ptr(self.N), ptr(self.T), self.K, self.d, self.m,
                                       C.byref(c)))
        self.h = h
        self.frame_offsets = np.concatenate([[0], np.cumsum(self.T)[:-1]]).astype(np.int64)

    def close(self):
        if getattr(self, "h", None):
            self.lib.pf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load_model(self, s, means, covs, weights, gamma, pca_proj, pca_mean, sel, R=None):
        means = f64(means); pca_proj = f64(pca_proj)
        sel_a = i32(sel)
        assert means.shape == (self.K, self.d) and sel_a.size == self.m
        check(self.lib.pf_batch_load_model(self.h, int(s), pca_proj.shape[1], ptr(sel_a), ptr(means),
                                           ptr(f64(covs)), ptr(f64(weights)), ptr(f64(gamma)), ptr(pca_proj),
                                           ptr(f64(pca_mean)), ptr(None if R is None else f64(R))))

    def set_joint_clutter(self, clutter):
        check(self.lib.pf_batch_set_joint_clutter(self.h, ptr(i32(clutter))))

    def reset(self, s, bins):
        check(self.lib.pf_batch_reset(self.h, int(s), ptr(i32(bins))))

    def run_dev(self, count_ptr, cand_ptr, uni_ptr, lml_ptr, graph: bool = True) -> int:
        """Device pointers to the concatenated per-sequence inputs (frame
        offsets ``frame_offsets``): count (F x J int32), cand (F x J x cap x 2
        float32), uniforms (F x 2 uint32); lml (F float64) is written.
        Returns the run's warning bits (capi.PF_WARN_*)."""
        return check(self.lib.pf_batch_run_dev(self.h, count_ptr, cand_ptr, uni_ptr, lml_ptr, 1 if graph else 0))

    def status(self):
        """Per-sequence status bits of the last run (capi.PF_WARN_* and 4 = not SPD)."""
        out = np.zeros(self.S, np.int32)
        check(self.lib.pf_batch_get_status(self.h, ptr(out)))
        return out

    def last_run_ms(self) -> float:
        v = C.c_double(0.0)
        check(self.lib.pf_batch_last_run_ms(self.h, C.byref(v)))
        return v.value

    def set_tracks(self, s, x, P):
        check(self.lib.pf_batch_set_tracks(self.h, int(s), ptr(f64(x)), ptr(f64(P))))

    def last_frame(self, s):
        """(log-weights, ancestors) of sequence s's last frame after a run."""
        n = int(self.N[s])
        lw = np.empty(n, np.float32); anc = np.empty(n, np.int32)
        check(self.lib.pf_batch_get_last_frame(self.h, int(s), ptr(lw), ptr(anc)))
        return lw, anc

    @property
    def variant(self) -> int:
        """Step kernel of the d = 24 state: capi.PF_BATCH_VALU, PF_BATCH_MFMA or PF_BATCH_QUAD_VALU."""
        v = C.c_int32(0)
        check(self.lib.pf_batch_get_variant(self.h, C.byref(v)))
        return v.value

    @variant.setter
    def variant(self, v: int):
        check(self.lib.pf_batch_set_variant(self.h, int(v)))

    def get_tracks(self, s):
        n = int(self.N[s])
        x = np.empty((n, self.d)); P = np.empty((n, self.d, self.d))
        check(self.lib.pf_batch_get_tracks(self.h, int(s), ptr(x), ptr(P)))
        return x, P

    def estimate(self, s):
        out = np.empty(self.d)
        check(self.lib.pf_batch_estimate(self.h, int(s), ptr(out)))
        return out

    @property
    def schedule(self):
        sc = capi.Schedule()
        check(self.lib.pf_batch_get_schedule(self.h, C.byref(sc)))
        return sc.as_tuple()

    @schedule.setter
    def schedule(self, value):
        sc = capi.Schedule(*[int(v) for v in value])
        check(self.lib.pf_batch_set_schedule(self.h, C.byref(sc)))

    def autotune(self, count_ptr, cand_ptr, uni_ptr, lml_ptr, frames: int = 16):
        """Times every candidate schedule over ``frames`` frames and keeps the
        fastest; returns (best schedule, best ms, default-schedule ms).  The
        tracks are consumed (reset before the next run)."""
        sc = capi.Schedule()
        best, dflt = C.c_double(0.0), C.c_double(0.0)
        check(self.lib.pf_batch_autotune(self.h, count_ptr, cand_ptr, uni_ptr, lml_ptr, int(frames), C.byref(sc),
                                         C.byref(best), C.byref(dflt)))
        return sc.as_tuple(), best.value, dflt.value

    # ------------------------------------------------------------- adjoint
    def lml_grad(self, count_ptr, cand_ptr, uni_ptr, lml_ptr, checkpoint_every: int = 50):
        """Log-marginal-likelihood gradient of every sequence (the adjoint of
        the run, ``pf_batch_lml_grad``).  Returns a dict of FP64 arrays:
        means (S, K, d), logits (S, K), rates (S, K), chol (S, K, d, d)."""
        S, K, d = self.S, self.K, self.d
        g = dict(means=np.zeros((S, K, d)), logits=np.zeros((S, K)), rates=np.zeros((S, K)),
                 chol=np.zeros((S, K, d, d)))
        check(self.lib.pf_batch_lml_grad(self.h, count_ptr, cand_ptr, uni_ptr, lml_ptr, int(checkpoint_every),
                                         ptr(g["means"]), ptr(g["logits"]), ptr(g["rates"]), ptr(g["chol"])))
        return g

    def get_choices(self, s):
        """(comp, anc), each n_frames[s] x n_particles[s], of the last lml_grad run."""
        T, N = int(self.T[s]), int(self.N[s])
        comp = np.empty((T, N), np.int32); anc = np.empty((T, N), np.int32)
        check(self.lib.pf_batch_get_choices(self.h, int(s), ptr(comp), ptr(anc)))
        return comp, anc

    def last_adjoint_ms(self):
        f, b = C.c_double(0.0), C.c_double(0.0)
        check(self.lib.pf_batch_last_adjoint_ms(self.h, C.byref(f), C.byref(b)))
        return f.value, b.value


def plan(n_particles, n_frames, block: int = 256, ppl: int = 1) -> dict:
    """The validated layout of a batch of these sequences under a schedule
    (pf_batch_plan; host only, no GPU needed)."""
    N = capi.i32(n_particles)
    T = capi.i32(n_frames)
    out = np.zeros(6, np.int64)
    check(capi.lib().pf_batch_plan(len(N), ptr(N), ptr(T), int(block), int(ppl), ptr(out)))
    return dict(zip(("ld", "blocks", "tiles", "weight_blocks", "nb_max", "Tmax"), (int(v) for v in out)))
