"""The pose tracker's per-frame pipeline (``include/pf_tracker.h``) from Python.

``PoseTracker`` keeps the non-ROS core of the reference's ``PFTracker``
(src/pfPose.h:30-78): two arm filters, their PCA projections, and per frame
the likelihood blur, hand proposals, association weights, proposal
resampling, the 6 x N measurements, both arm updates and the pose output
(``callback``, pfPose.cpp:332-349, without the ROS publishing and drawing).
Every step runs in the gfx950 kernels of ``libmkfbt.so``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, ptr
from .rbpf import Rbpf


class PoseTracker:
    def __init__(self, arm1: Rbpf, arm2: Rbpf, pca_proj1, pca_mean1, pca_proj2, pca_mean2, **cfg):
        self.lib = capi.lib()
        c = capi.TrackerConfig()
        self.lib.pf_tracker_config_default(C.byref(c))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.cfg = c
        self.arms = (arm1, arm2)
        self.pca_proj = (np.ascontiguousarray(pca_proj1, np.float64), np.ascontiguousarray(pca_proj2, np.float64))
        self.pca_mean = (np.ascontiguousarray(pca_mean1, np.float64).ravel(),
                         np.ascontiguousarray(pca_mean2, np.float64).ravel())
        self.npose = self.pca_mean[0].size
        h = C.c_void_p()
        check(self.lib.pf_tracker_create(C.byref(h), C.byref(c), arm1.h, arm2.h, self.npose,
                                         ptr(self.pca_proj[0]), ptr(self.pca_mean[0]),
                                         ptr(self.pca_proj[1]), ptr(self.pca_mean[1])))
        self.h = h
        n = C.c_int32(0)
        check(self.lib.pf_tracker_num_proposals(self.h, C.byref(n)))
        self.n_c = n.value
        self.N = arm1.N

    @classmethod
    def from_yml(cls, left_path: str, right_path: str, n_particles: int = 500, seed: int = 1, device: int = 0,
                 **cfg):
        """Both arms from the GMM yml files (pfPose.cpp:34-65; each file's own
        gamma) and initialised as pfPose.cpp:68-71: tracks from a systematic
        resample of the component weights."""
        arms = []
        for a, path in enumerate((left_path, right_path)):
            rb = Rbpf(n_particles, device=device)
            rb.load_model_yml(path)
            w = capi.yml_read_matrix(path, "weights").ravel()
            u = int(capi.philox_u32(seed, 32 + a, 0, 1)[0])
            bins, _ = rb.resample(w, n_particles, u)
            rb.reset(bins)
            arms.append(rb)
        proj = [capi.yml_read_matrix(p, "pca_proj") for p in (left_path, right_path)]
        mean = [capi.yml_read_matrix(p, "pca_mean").ravel() for p in (left_path, right_path)]
        tr = cls(arms[0], arms[1], proj[0], mean[0], proj[1], mean[1], seed=seed, **cfg)
        tr._owns_arms = True
        return tr

    _owns_arms = False

    def close(self):
        if getattr(self, "h", None):
            self.lib.pf_tracker_destroy(self.h)
            self.h = None
        if self._owns_arms:
            for a in self.arms:
                a.close()
            self._owns_arms = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        check(self.lib.pf_tracker_reset(self.h))

    # ---------------------------------------------------------- per frame
    def _host_outputs(self):
        o = dict(lml=np.zeros(2), pose=np.zeros((2, self.npose)), pos3d=np.zeros((2, 3, 5)),
                 tf=np.zeros((8, 3)), pose2d=np.zeros((8, 2)))
        st = capi.TrackerOutputs(ptr(o["lml"]), ptr(o["pose"]), ptr(o["pos3d"]), ptr(o["tf"]), ptr(o["pose2d"]))
        return o, st

    def callback(self, likelihood, roi):
        """One frame: likelihood (rows x cols uint8), roi = (x_offset,
        y_offset, width, height) of the face, or None for a frame without a
        face.  Returns dict(lml (2,), pose (2 x npose), pos3d (2 x 3 x 5),
        tf (8 x 3, joints relative to Head), pose2d (8 x 2), status)."""
        img = np.ascontiguousarray(likelihood, np.uint8)
        if img.shape != (self.cfg.rows, self.cfg.cols):
            raise ValueError(f"likelihood image must be {self.cfg.rows} x {self.cfg.cols}")
        o, st_out = self._host_outputs()
        r = None if roi is None else np.ascontiguousarray(roi, np.int32)
        st = check(self.lib.pf_tracker_frame(self.h, ptr(img), ptr(r), C.byref(st_out)))
        o["status"] = st
        return o

    frame = callback

    def pose(self):
        """Pose outputs of both arms' current tracks (pfPose.cpp:347-348,
        93-208): dict(pose, pos3d, tf, pose2d)."""
        o, st_out = self._host_outputs()
        check(self.lib.pf_tracker_pose(self.h, C.byref(st_out)))
        del o["lml"]
        return o

    def get_last(self):
        rows, cols = self.cfg.rows, self.cfg.cols
        blurred = np.zeros((rows, cols), np.uint8)
        props = np.zeros((2, 2, self.n_c))
        w = np.zeros((2, self.n_c))
        bins = np.zeros((2, self.N), np.int32)
        Z = np.zeros((2, 6, self.N), np.float32)
        check(self.lib.pf_tracker_get_last(self.h, ptr(blurred), ptr(props), ptr(w), ptr(bins), ptr(Z)))
        return dict(blurred=blurred, props=props, w=w, bins=bins, Z=Z)

    def blur(self, img):
        img = np.ascontiguousarray(img, np.uint8)
        out = np.zeros_like(img)
        check(self.lib.pf_tracker_blur(self.h, ptr(img), ptr(out)))
        return out

    # ------------------------------------------------ device-resident run
    def run_dev(self, T, lik_ptr, roi_ptr, lml_ptr=None, pose_ptr=None, pos3d_ptr=None, tf_ptr=None,
                pose2d_ptr=None, graph=True):
        out = capi.TrackerOutputs(lml_ptr, pose_ptr, pos3d_ptr, tf_ptr, pose2d_ptr)
        check(self.lib.pf_tracker_run_dev(self.h, int(T), lik_ptr, roi_ptr, C.byref(out), 1 if graph else 0))

    def synchronize(self):
        check(self.lib.pf_tracker_synchronize(self.h))

    def last_run_ms(self) -> float:
        v = C.c_double(0.0)
        check(self.lib.pf_tracker_last_run_ms(self.h, C.byref(v)))
        return v.value
