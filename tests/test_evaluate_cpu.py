"""Host logic of the evaluation harness (mkfbodytracker_pdaf_amd/evaluate.py):
the ground-truth TF rows and back-projection against the oracle's literal
restatements, the extractTFdata log layout, and the error summary."""
import os

import numpy as np

import oracle
from mkfbodytracker_pdaf_amd import evaluate

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
LEFT = os.path.join(GOLDEN, "data13D_PCA_100000_15_12.yml")
RIGHT = os.path.join(GOLDEN, "data23D_PCA_100000_15_12.yml")


def test_truth_geometry_matches_oracle():
    arms = [evaluate._read_arm(LEFT), evaluate._read_arm(RIGHT)]
    E, P3 = evaluate.simulate_truth(arms, 12, seed=3)
    assert E.shape == (12, 2, 22) and P3.shape == (12, 2, 3, 5)
    assert np.all(np.isfinite(E)) and np.all(np.isfinite(P3))
    for t in range(12):
        for a in range(2):
            assert np.allclose(P3[t, a], oracle.get3Dpose(E[t, a]), rtol=1e-12, atol=1e-12)
        assert np.allclose(evaluate.tf_rows(P3[t, 0], P3[t, 1]), oracle.tf_rows(P3[t, 0], P3[t, 1]),
                           rtol=0, atol=1e-14)
    # the poses stay near the models' mean pose (hands within the image)
    assert np.all((E[:, :, 0] > 0) & (E[:, :, 0] < 640) & (E[:, :, 1] > 0) & (E[:, :, 1] < 480))


def test_render_centres_the_face():
    arms = [evaluate._read_arm(LEFT), evaluate._read_arm(RIGHT)]
    E, _ = evaluate.simulate_truth(arms, 4, seed=5)
    imgs, rois = evaluate.render(E, seed=5)
    assert imgs.shape == (4, 480, 640) and imgs.dtype == np.uint8
    for t in range(4):
        hu = 0.5 * (E[t, 0, 9] + E[t, 1, 9])
        hv = 0.5 * (E[t, 0, 10] + E[t, 1, 10])
        assert abs(rois[t, 0] + rois[t, 2] / 2.0 - hu) <= 0.5 and abs(rois[t, 1] + 0.5 * rois[t, 3] - hv) <= 0.5
        for a in range(2):           # a blob at each hand
            x, y = int(E[t, a, 0]), int(E[t, a, 1])
            assert imgs[t, y, x] > 150


def test_log_layout_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    cam, kin = rng.normal(size=(5, 8, 3)), rng.normal(size=(5, 8, 3))
    pk, pc = evaluate.write_logs(str(tmp_path), "/data13D.yml", cam, kin, stamp="x")
    assert os.path.basename(pk) == "kinectdata_x.txt" and os.path.basename(pc) == "camdata_x.txt"
    lines = open(pc).read().splitlines()
    assert lines[0] == "%/data13D.yml" and len(lines) == 6
    assert all(l.startswith(" ") and len(l.split()) == 24 for l in lines[1:])
    name, rows = evaluate.read_log(pc)
    assert name == "/data13D.yml" and np.array_equal(rows, cam)
    assert np.array_equal(evaluate.read_log(pk)[1], kin)


def test_joint_errors():
    kin = np.zeros((3, 8, 3))
    cam = kin.copy()
    cam[:, 1, 0] = 0.03               # Left_Hand 3 cm off in x
    cam[:, 7] = [0.0, 0.04, 0.03]     # Neck 5 cm off
    err = evaluate.joint_errors(cam, kin)
    assert set(err) == set(evaluate.JOINTS[1:])
    assert abs(err["Left_Hand"] - 0.03) < 1e-15 and abs(err["Neck"] - 0.05) < 1e-15
    assert err["Right_Hand"] == 0.0
