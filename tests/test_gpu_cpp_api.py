"""Runs the C++ drop-in check (tests/cpp/test_host_api.cpp, built by
__graft_entry__.build()): the reference's PFTracker call sequence on the
C++ mirror classes, against the shipped GMM/PCA model."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_host_api")


def test_cpp_mirror_classes(gpu):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/test_host_api is not built (run __graft_entry__.build())")
    r = subprocess.run([BIN, os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "OK" in r.stdout
