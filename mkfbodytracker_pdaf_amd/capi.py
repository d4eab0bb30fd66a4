"""ctypes binding of the C ABI in ``include/pf_rbpf.h`` (``libmkfbt.so``).

The library is the product: every call below runs the gfx950 kernels.  There
is no CPU fallback; loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._build import LIB

PF_OK = 0
PF_WARN_DEGENERATE = 1
PF_WARN_ILLCOND = 2
PF_ERR_ARG = -1
PF_ERR_HIP = -2
PF_ERR_NOT_SPD = -3
PF_ERR_STATE = -4
PF_ERR_UNSUPPORTED = -5
PF_ERR_IO = -6
PF_BATCH_VALU, PF_BATCH_MFMA, PF_BATCH_QUAD_VALU = 0, 1, 2   # include/pf_batch.h
PF_LAYOUT_AUTO, PF_LAYOUT_ROWS, PF_LAYOUT_RECORDS = 0, 1, 2   # include/pf_rbpf.h
PF_NOALLOC_MAX_WEIGHTS = 524288   # include/pf_rbpf.h

# Every symbol declared in include/pf_rbpf.h, pf2d.h, pf_batch.h and pf_tracker.h.
EXPORTS = [
    "pf_config_default", "pf_last_error", "pf_create", "pf_destroy", "pf_load_model",
    "pf_set_joint_clutter", "pf_reset", "pf_set_tracks", "pf_get_tracks", "pf_step",
    "pf_step_candidates", "pf_estimate", "pf_resample", "pf_resample_log", "pf_get_samples",
    "pf_resample_log_noalloc", "pf_get_sample_prob", "pf_run_ref_dev", "pf_run_candidates_dev", "pf_synchronize",
    "pf_get_stream", "pf_last_run_kernel_time", "pf_last_run_graph_frames", "pf_set_state_layout",
    "pf_get_state_layout", "pf_kf_predict", "pf_kf_update",
    "pf_yml_read_matrix", "pf_read_camera_info", "pf_load_model_yml", "pf_philox_u32",
    "pf2d_create", "pf2d_destroy", "pf2d_load_gmm", "pf2d_load_gmm_precision", "pf2d_init", "pf2d_set_particles",
    "pf2d_get_particles", "pf2d_update", "pf2d_predict", "pf2d_estimate",
    "pf_batch_create", "pf_batch_destroy", "pf_batch_plan", "pf_batch_load_model", "pf_batch_set_joint_clutter",
    "pf_batch_reset", "pf_batch_run_dev", "pf_batch_get_tracks", "pf_batch_estimate",
    "pf_batch_set_tracks", "pf_batch_get_last_frame", "pf_batch_set_variant", "pf_batch_get_variant",
    "pf_batch_get_status",
    "pf_batch_get_schedule", "pf_batch_set_schedule", "pf_batch_autotune", "pf_batch_last_run_ms",
    "pf_batch_lml_grad", "pf_batch_get_choices", "pf_batch_last_adjoint_ms",
    "pf_tracker_config_default", "pf_tracker_create", "pf_tracker_destroy", "pf_tracker_reset",
    "pf_tracker_frame", "pf_tracker_run_dev", "pf_tracker_synchronize", "pf_tracker_last_run_ms",
    "pf_tracker_get_last", "pf_tracker_pose", "pf_tracker_blur", "pf_tracker_num_proposals",
]


class Schedule(C.Structure):
    _fields_ = [("block", C.c_int32), ("ppl", C.c_int32), ("blocks_per_cu", C.c_int32)]

    def as_tuple(self):
        return (self.block, self.ppl, self.blocks_per_cu)


class PfError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("n_particles", C.c_int32), ("device", C.c_int32), ("cand_capacity", C.c_int32),
                ("block_size", C.c_int32), ("p_detect", C.c_double), ("gate_chi2", C.c_double),
                ("clutter_density", C.c_double)]


class TrackerConfig(C.Structure):
    _fields_ = [("rows", C.c_int32), ("cols", C.c_int32), ("proposals_per_particle", C.c_int32),
                ("p_assoc", C.c_double), ("z_floor", C.c_double), ("spread", C.c_double),
                ("blur_ksize", C.c_int32), ("blur_sigma", C.c_double), ("fx", C.c_double),
                ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double), ("seed", C.c_uint64)]


class TrackerOutputs(C.Structure):
    _fields_ = [("lml", C.c_void_p), ("pose", C.c_void_p), ("pos3d", C.c_void_p), ("tf", C.c_void_p),
                ("pose2d", C.c_void_p)]


_lib = None
_vp = C.c_void_p


def _sig(L, name, res, *args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = list(args)


def lib():
    """Load libmkfbt.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise ImportError(f"{LIB} is missing: build it with __graft_entry__.build() "
                          "or python -m mkfbodytracker_pdaf_amd._build")
    L = C.CDLL(LIB)
    i32, u32, u64, i64, dbl = C.c_int32, C.c_uint32, C.c_uint64, C.c_int64, C.c_double
    _sig(L, "pf_config_default", None, C.POINTER(Config))
    _sig(L, "pf_last_error", C.c_char_p)
    _sig(L, "pf_create", C.c_int, C.POINTER(_vp), C.POINTER(Config))
    _sig(L, "pf_destroy", C.c_int, _vp)
    _sig(L, "pf_load_model", C.c_int, _vp, i32, i32, i32, i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
    _sig(L, "pf_set_joint_clutter", C.c_int, _vp, _vp)
    _sig(L, "pf_reset", C.c_int, _vp, _vp)
    _sig(L, "pf_set_tracks", C.c_int, _vp, _vp, _vp)
    _sig(L, "pf_get_tracks", C.c_int, _vp, _vp, _vp)
    _sig(L, "pf_step", C.c_int, _vp, _vp, u32, u32, _vp, _vp, _vp, _vp)
    _sig(L, "pf_step_candidates", C.c_int, _vp, _vp, _vp, u32, u32, _vp, _vp, _vp, _vp)
    _sig(L, "pf_estimate", C.c_int, _vp, _vp)
    _sig(L, "pf_resample", C.c_int, _vp, _vp, i32, i32, u32, _vp)
    _sig(L, "pf_resample_log", C.c_int, _vp, _vp, i32, i32, u32, _vp)
    _sig(L, "pf_resample_log_noalloc", C.c_int, _vp, i32, i32, u32, _vp, _vp, _vp)
    _sig(L, "pf_get_samples", C.c_int, _vp, _vp, _vp, i32, i32, dbl, _vp, _vp)
    _sig(L, "pf_get_sample_prob", C.c_int, _vp, _vp, _vp, i32, _vp, _vp, i32, dbl, _vp, _vp)
    _sig(L, "pf_run_ref_dev", C.c_int, _vp, i32, _vp, _vp, _vp, i32)
    _sig(L, "pf_run_candidates_dev", C.c_int, _vp, i32, _vp, _vp, _vp, _vp, i32)
    _sig(L, "pf_synchronize", C.c_int, _vp)
    _sig(L, "pf_get_stream", C.c_int, _vp, C.POINTER(_vp))
    _sig(L, "pf_last_run_kernel_time", C.c_int, _vp, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(i32))
    _sig(L, "pf_last_run_graph_frames", C.c_int, _vp, C.POINTER(i32))
    _sig(L, "pf_set_state_layout", C.c_int, _vp, i32)
    _sig(L, "pf_get_state_layout", C.c_int, _vp, C.POINTER(i32))
    _sig(L, "pf_kf_predict", C.c_int, _vp, i32, i32, _vp, _vp, _vp, _vp, _vp)
    _sig(L, "pf_kf_update", C.c_int, _vp, i32, i32, i32, _vp, _vp, _vp, _vp, _vp, _vp)
    _sig(L, "pf_yml_read_matrix", C.c_int, C.c_char_p, C.c_char_p, C.POINTER(i32), C.POINTER(i32), _vp, i64)
    _sig(L, "pf_read_camera_info", C.c_int, C.c_char_p, C.POINTER(i32), C.POINTER(i32), _vp, _vp, _vp)
    _sig(L, "pf_load_model_yml", C.c_int, _vp, C.c_char_p, C.c_char_p, i32, _vp, _vp)
    _sig(L, "pf_philox_u32", None, u64, u32, u64, i64, _vp)
    _sig(L, "pf2d_create", C.c_int, C.POINTER(_vp), i32, i32, i32, i32)
    _sig(L, "pf2d_destroy", C.c_int, _vp)
    _sig(L, "pf2d_load_gmm", C.c_int, _vp, i32, _vp, _vp, _vp)
    _sig(L, "pf2d_load_gmm_precision", C.c_int, _vp, i32, _vp, _vp, _vp, _vp)
    _sig(L, "pf2d_init", C.c_int, _vp, _vp)
    _sig(L, "pf2d_set_particles", C.c_int, _vp, _vp, _vp)
    _sig(L, "pf2d_get_particles", C.c_int, _vp, _vp, _vp)
    _sig(L, "pf2d_update", C.c_int, _vp, _vp, u32, _vp, _vp, _vp, _vp)
    _sig(L, "pf2d_predict", C.c_int, _vp, _vp)
    _sig(L, "pf2d_estimate", C.c_int, _vp, _vp)
    _sig(L, "pf_batch_create", C.c_int, C.POINTER(_vp), i32, _vp, _vp, i32, i32, i32, C.POINTER(Config))
    _sig(L, "pf_batch_destroy", C.c_int, _vp)
    _sig(L, "pf_batch_plan", C.c_int, i32, _vp, _vp, i32, i32, _vp)
    _sig(L, "pf_batch_load_model", C.c_int, _vp, i32, i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp)
    _sig(L, "pf_batch_set_joint_clutter", C.c_int, _vp, _vp)
    _sig(L, "pf_batch_reset", C.c_int, _vp, i32, _vp)
    _sig(L, "pf_batch_run_dev", C.c_int, _vp, _vp, _vp, _vp, _vp, i32)
    _sig(L, "pf_batch_get_tracks", C.c_int, _vp, i32, _vp, _vp)
    _sig(L, "pf_batch_estimate", C.c_int, _vp, i32, _vp)
    _sig(L, "pf_batch_set_tracks", C.c_int, _vp, i32, _vp, _vp)
    _sig(L, "pf_batch_get_last_frame", C.c_int, _vp, i32, _vp, _vp)
    _sig(L, "pf_batch_set_variant", C.c_int, _vp, i32)
    _sig(L, "pf_batch_get_variant", C.c_int, _vp, C.POINTER(i32))
    _sig(L, "pf_batch_get_status", C.c_int, _vp, _vp)
    _sig(L, "pf_batch_get_schedule", C.c_int, _vp, C.POINTER(Schedule))
    _sig(L, "pf_batch_set_schedule", C.c_int, _vp, C.POINTER(Schedule))
    _sig(L, "pf_batch_autotune", C.c_int, _vp, _vp, _vp, _vp, _vp, i32, C.POINTER(Schedule), C.POINTER(dbl),
         C.POINTER(dbl))
    _sig(L, "pf_batch_last_run_ms", C.c_int, _vp, C.POINTER(dbl))
    _sig(L, "pf_batch_lml_grad", C.c_int, _vp, _vp, _vp, _vp, _vp, i32, _vp, _vp, _vp, _vp)
    _sig(L, "pf_batch_get_choices", C.c_int, _vp, i32, _vp, _vp)
    _sig(L, "pf_batch_last_adjoint_ms", C.c_int, _vp, C.POINTER(dbl), C.POINTER(dbl))
    _sig(L, "pf_tracker_config_default", None, C.POINTER(TrackerConfig))
    _sig(L, "pf_tracker_create", C.c_int, C.POINTER(_vp), C.POINTER(TrackerConfig), _vp, _vp, i32, _vp, _vp,
         _vp, _vp)
    _sig(L, "pf_tracker_destroy", C.c_int, _vp)
    _sig(L, "pf_tracker_reset", C.c_int, _vp)
    _sig(L, "pf_tracker_frame", C.c_int, _vp, _vp, _vp, C.POINTER(TrackerOutputs))
    _sig(L, "pf_tracker_run_dev", C.c_int, _vp, i32, _vp, _vp, C.POINTER(TrackerOutputs), i32)
    _sig(L, "pf_tracker_synchronize", C.c_int, _vp)
    _sig(L, "pf_tracker_last_run_ms", C.c_int, _vp, C.POINTER(dbl))
    _sig(L, "pf_tracker_get_last", C.c_int, _vp, _vp, _vp, _vp, _vp, _vp)
    _sig(L, "pf_tracker_pose", C.c_int, _vp, C.POINTER(TrackerOutputs))
    _sig(L, "pf_tracker_blur", C.c_int, _vp, _vp, _vp)
    _sig(L, "pf_tracker_num_proposals", C.c_int, _vp, C.POINTER(i32))
    _lib = L
    return L


def check(code: int) -> int:
    """Raise PfError for negative status codes; return non-negative ones."""
    if code < 0:
        raise PfError(code, lib().pf_last_error().decode())
    return code


def ptr(a):
    """Pointer argument for a numpy array (or None).  The returned ctypes
    pointer holds a reference to the array, so a temporary such as
    ptr(f64(x)) stays alive for the whole foreign call."""
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def philox_u32(seed: int, stream: int, first: int, count: int) -> np.ndarray:
    out = np.empty(count, np.uint32)
    lib().pf_philox_u32(int(seed), int(stream), int(first), int(count), ptr(out))
    return out


def default_config(**kw) -> Config:
    c = Config()
    lib().pf_config_default(C.byref(c))
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def yml_read_matrix(path: str, key: str) -> np.ndarray:
    r, c = C.c_int32(0), C.c_int32(0)
    check(lib().pf_yml_read_matrix(path.encode(), key.encode(), C.byref(r), C.byref(c), None, 0))
    out = np.empty((r.value, c.value), np.float64)
    check(lib().pf_yml_read_matrix(path.encode(), key.encode(), C.byref(r), C.byref(c), ptr(out), out.size))
    return out


def read_camera_info(path: str):
    w, h = C.c_int32(0), C.c_int32(0)
    K = np.zeros(9); D = np.zeros(5); P = np.zeros(12)
    check(lib().pf_read_camera_info(path.encode(), C.byref(w), C.byref(h), ptr(K), ptr(D), ptr(P)))
    return dict(width=w.value, height=h.value, K=K.reshape(3, 3), D=D, P=P.reshape(3, 4))
