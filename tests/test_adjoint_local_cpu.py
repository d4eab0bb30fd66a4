"""The hand-derived reverse of a joint's local function (the 2-row KF / PDAF
update as a function of y and S, used by the GPU adjoint) against its
forward-mode dual-number derivatives, on 4000 random certain and PDAF joints
(host build of csrc/pf_adjoint_local.hpp; tolerance 2e-3 relative, FP32)."""
import os
import subprocess

import pytest

BIN = os.path.join(os.path.dirname(__file__), "cpp", "test_adjoint_local")


def test_local_reverse_matches_dual_numbers():
    if not os.path.exists(BIN):
        pytest.skip("tests/cpp/test_adjoint_local is built by __graft_entry__.build()")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "max relative adjoint error" in r.stdout
