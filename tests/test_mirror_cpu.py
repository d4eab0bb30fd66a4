"""Host-side members of the reference-class mirrors (no GPU): pf2D
my_gmm::loadGaussian keeps sigma_i and det_s (src/pf2D.cpp:28-37), checked
against the oracle's restatement, which tests/test_oracle.py pins against
numpy's inverse and determinant."""
import numpy as np
import pytest

import oracle
from mkfbodytracker_pdaf_amd import capi
from mkfbodytracker_pdaf_amd.pf2d import my_gmm


def test_pf2d_my_gmm_members_match_oracle():
    rng = np.random.default_rng(21)
    g = my_gmm()
    for k in range(4):
        A = rng.standard_normal((8, 8)) * 30
        S = A @ A.T + np.diag(rng.uniform(100, 3000, 8))
        mu = rng.uniform(0, 640, 8)
        g.loadGaussian(mu.reshape(1, 8), S, 0.25)
        si, c = oracle.pf2d_load_gaussian(S)
        assert g.N == k + 1
        assert np.linalg.norm(g.sigma_i[k] - si) <= 1e-12 * np.linalg.norm(si)
        assert abs(g.det_s[k] - c) <= 1e-12 * c
        assert np.array_equal(g.mean[k], mu) and g.weight[k] == 0.25
    with pytest.raises(capi.PfError) as e:
        g.loadGaussian(np.zeros(8), -np.eye(8), 1.0)
    assert e.value.code == capi.PF_ERR_NOT_SPD and g.N == 4
