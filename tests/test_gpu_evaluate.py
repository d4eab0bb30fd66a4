"""The evaluation harness end to end on the GPU: a synthetic sequence tracked
by the device pipeline, logged and scored like extractTFdata's recordings."""
import os

import numpy as np
import pytest

from mkfbodytracker_pdaf_amd import evaluate

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
LEFT = os.path.join(GOLDEN, "data13D_PCA_100000_15_12.yml")
RIGHT = os.path.join(GOLDEN, "data23D_PCA_100000_15_12.yml")


def test_harness_tracks_synthetic_body(gpu, tmp_path):
    res = evaluate.run_once(LEFT, RIGHT, N=500, T=40, seed=4)
    assert res["cam"].shape == (40, 8, 3) and np.all(np.isfinite(res["cam"]))
    assert np.all(res["cam"][:, 0] == 0.0)                   # Head relative to itself
    assert np.all(np.isfinite(res["lml"])) and res["ms"] > 0
    err = evaluate.joint_errors(res["cam"], res["kinect"])
    assert all(np.isfinite(v) for v in err.values())
    # the hands are found: mean image error of the tracked hands well below
    # the proposal spread (0.8 x 40 px)
    assert res["hand_px"] < 32.0, res["hand_px"]
    pk, pc = evaluate.write_logs(str(tmp_path), "data13D", res["cam"], res["kinect"], stamp="t")
    assert np.array_equal(evaluate.read_log(pc)[1], res["cam"])
