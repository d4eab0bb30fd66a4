"""pf_resample_log_noalloc (the resampler that allocates nothing, static
device scratch) against pf_resample_log (handle scratch) through the C ABI:
identical ancestor indices and status on consecutive calls, the second
smaller than the first, so the shared scratch's leftovers from the first call
must not leak into the second."""
import ctypes as C

import numpy as np
import pytest

from mkfbodytracker_pdaf_amd import capi
from mkfbodytracker_pdaf_amd.devmem import DeviceArray, synchronize
from mkfbodytracker_pdaf_amd.rbpf import Rbpf


def _noalloc(lw, N, U):
    L = capi.lib()
    d_lw = DeviceArray(np.ascontiguousarray(lw, np.float64))
    d_anc = DeviceArray(np.full(N, -1, np.int32))
    d_st = DeviceArray(np.full(1, -7, np.int32))
    capi.check(L.pf_resample_log_noalloc(d_lw.ptr, lw.size, N, U, d_anc.ptr, d_st.ptr, None))
    synchronize()
    out, st = d_anc.get(), int(d_st.get()[0])
    for a in (d_lw, d_anc, d_st):
        a.close()
    return out, st


@pytest.mark.gpu
def test_noalloc_matches_handle_resampler(gpu):
    rng = np.random.default_rng(20261203)
    cases = [(400_000, 250_000), (5_000, 3_000), (300, 1_000)]
    rb = Rbpf(max(n for _, n in cases))
    for L, N in cases:                            # each call smaller than the one before
        lw = rng.normal(-40.0, 8.0, L)
        lw[rng.random(L) < 0.2] = -np.inf
        U = int(rng.integers(0, 2 ** 32))
        print(f"noalloc resample L={L} N={N} U={U}", flush=True)
        got, st = _noalloc(lw, N, U)
        ref, st_ref = rb.resample(lw, N, U, log=True)
        assert st == st_ref == 0
        assert np.array_equal(got, ref), (L, N)
    # every weight non-finite: the uniform fallback and its status, after the
    # non-degenerate calls above left their status in the scratch
    lw = np.full(777, -np.inf)
    got, st = _noalloc(lw, 500, 99)
    ref, st_ref = rb.resample(lw, 500, 99, log=True)
    assert st == st_ref == 1 and np.array_equal(got, ref)
    # and a normal call after the degenerate one clears the status again
    lw = rng.normal(0.0, 1.0, 1000)
    got, st = _noalloc(lw, 1000, 5)
    ref, st_ref = rb.resample(lw, 1000, 5, log=True)
    assert st == st_ref == 0 and np.array_equal(got, ref)
    rb.close()


@pytest.mark.gpu
def test_noalloc_rejects_oversized_input(gpu):
    L = capi.lib()
    d_lw = DeviceArray(np.zeros(16, np.float64))
    d_anc = DeviceArray(np.zeros(16, np.int32))
    code = L.pf_resample_log_noalloc(d_lw.ptr, capi.PF_NOALLOC_MAX_WEIGHTS + 1, 16, 0, d_anc.ptr, None, None)
    assert code == capi.PF_ERR_ARG
    assert b"exceeds" in L.pf_last_error()
    assert L.pf_resample_log_noalloc(None, 16, 16, 0, d_anc.ptr, None, None) == capi.PF_ERR_ARG
    d_lw.close()
    d_anc.close()
