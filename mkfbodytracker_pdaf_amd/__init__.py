"""MI355X (gfx950) particle-filter hot path of mkfbodytracker_pdaf.

The reference's pf2DRao / KF_model / my_gmm (and pf2D) classes, re-built as
hand-written HIP kernels behind a C ABI (include/pf_rbpf.h, include/pf2d.h),
with Python mirrors of the reference class interfaces on top.
"""
from . import capi  # noqa: F401
from ._build import LIB, build  # noqa: F401

__all__ = ["capi", "build", "LIB"]
