"""CPU-side checks of the product library: it loads, exports every symbol the
public headers declare, and its host-side input readers / RNG behave (no
device compute is called here)."""
import os
import re

import numpy as np
import pytest

from mkfbodytracker_pdaf_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def _declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(pf\w*)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    declared = _declared_symbols()
    assert declared, "no declarations found"
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert set(capi.EXPORTS) <= declared


def test_yml_reader_shipped_models():
    """cv::FileStorage reads of the shipped GMM/PCA files (pfPose.cpp:34-53)."""
    for f in ("data13D_PCA_100000_15_12.yml", "data23D_PCA_100000_15_12.yml"):
        p = os.path.join(GOLD, f)
        means = capi.yml_read_matrix(p, "means")
        covs = capi.yml_read_matrix(p, "covs")
        w = capi.yml_read_matrix(p, "weights")
        proj = capi.yml_read_matrix(p, "pca_proj")
        mean = capi.yml_read_matrix(p, "pca_mean")
        g = capi.yml_read_matrix(p, "gamma")
        assert means.shape == (15, 12) and covs.shape == (180, 12) and w.shape == (1, 15)
        assert proj.shape == (12, 22) and mean.shape == (1, 22) and g.shape == (15, 1)
        # dt f values are FP32 in the file
        assert np.array_equal(proj, proj.astype(np.float32).astype(np.float64))
        S = covs.reshape(15, 12, 12)
        assert np.abs(S - S.transpose(0, 2, 1)).max() <= 1e-9 * np.abs(S).max()
        assert abs(w.sum() - 1.0) < 1e-12
        assert np.all((g > 0.8) & (g < 1.0))
        ev = np.linalg.eigvalsh(S)
        assert ev.min() > 0 and (ev[:, -1] / ev[:, 0]).max() > 1e6   # cond > 1e6 is exercised
    # spot values from the files' first entries
    p = os.path.join(GOLD, "data13D_PCA_100000_15_12.yml")
    assert capi.yml_read_matrix(p, "means")[0, 0] == -8.5242692850189997e+01
    assert capi.yml_read_matrix(p, "pca_mean")[0, 0] == np.float32(3.88015259e+02)


def test_camera_info_reader():
    ci = capi.read_camera_info(os.path.join(GOLD, "cal.yml"))
    assert ci["width"] == 640 and ci["height"] == 480
    assert ci["K"][0, 0] == 660.326889 and ci["K"][1, 2] == 240.784699 and ci["K"][2, 2] == 1
    assert ci["D"][1] == -0.33931
    assert ci["P"][1, 1] == 661.85144


def test_yml_errors_are_reported():
    with pytest.raises(capi.PfError) as e:
        capi.yml_read_matrix(os.path.join(GOLD, "cal.yml"), "no_such_key")
    assert e.value.code == capi.PF_ERR_IO
    with pytest.raises(capi.PfError):
        capi.yml_read_matrix("/nonexistent.yml", "means")


def test_philox_known_answers():
    """Random123 known-answer vectors of Philox4x32-10 through the product's stream."""
    import oracle
    # stream word i of (seed, stream) = philox(ctr={i/4, 0, stream, 0}, key=seed)[i % 4]
    a = capi.philox_u32(0, 0, 0, 4)
    assert [hex(v) for v in a] == ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    b = capi.philox_u32(123456789012345, 7, 1000, 4096)
    c = oracle.philox_u32(123456789012345, 7, 1000, 4096)
    assert np.array_equal(b, c)


def test_ptr_keeps_a_temporary_array_alive():
    """ptr(f64(x)) of an array that f64 must copy: the copy stays alive as long
    as the pointer object (the foreign call's argument), so the call never
    reads freed memory."""
    import ctypes as C
    import gc
    src = np.asfortranarray(np.arange(600.0).reshape(6, 100))
    p = capi.ptr(capi.f64(src))
    gc.collect()
    junk = [np.full(600, -1.0) for _ in range(200)]
    vals = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_double)), shape=(600,))
    assert np.array_equal(vals, np.ascontiguousarray(src).ravel())
    del junk
