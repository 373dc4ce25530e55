"""Accuracy scoring against dense ground truth (sgm_amd.accuracy), the
reference's rule (scripts/make_panel.py:34-50): bad>1px, bad>2px and MAE over
the valid pixels outside the leftmost D columns.  Known answers only."""
import numpy as np
import pytest

from sgm_amd import accuracy


def test_perfect_and_offset_maps():
    gt = np.random.default_rng(0).integers(0, 60, (20, 90)).astype(np.uint8)
    assert accuracy.score(gt, gt, 16) == {"bad1_pct": 0.0, "bad2_pct": 0.0, "mae": 0.0, "scored_px": 20 * 74}
    s = accuracy.score(gt + 1, gt, 16)       # off by exactly 1: not bad at >1
    assert (s["bad1_pct"], s["bad2_pct"], s["mae"]) == (0.0, 0.0, 1.0)
    s = accuracy.score(gt + 2, gt, 16)       # off by 2: bad at >1, not at >2
    assert (s["bad1_pct"], s["bad2_pct"], s["mae"]) == (100.0, 0.0, 2.0)
    d = gt.copy()
    d[:, 16:][:, ::2] += 3                   # half the scored columns off by 3
    s = accuracy.score(d, gt, 16)
    assert (s["bad1_pct"], s["bad2_pct"], s["mae"]) == (50.0, 50.0, 1.5)


def test_leftmost_D_columns_and_invalid_pixels_are_not_scored():
    gt = np.full((8, 40), 10, np.uint8)
    d = gt.copy()
    d[:, :32] = 200                          # errors only where x < D
    assert accuracy.score(d, gt, 32)["bad1_pct"] == 0.0
    valid = np.ones_like(gt, bool)
    d[:, 35] = 99
    valid[:, 35] = False
    s = accuracy.score(d, gt, 32, valid)
    assert s["bad1_pct"] == 0.0 and s["scored_px"] == 8 * 7


def test_decimated_maps_use_the_even_pixels():
    gt = np.arange(12 * 80, dtype=np.int64).reshape(12, 80) % 50
    gt = gt.astype(np.uint8)
    disp = gt[0::2, 0::2].copy()
    s = accuracy.score(disp, gt, 16, decimated=True)
    assert s["bad1_pct"] == 0.0 and s["scored_px"] == 6 * (40 - 8)  # 2x >= 16 from x = 8
    disp[:, 7] += 50                          # 2x = 14 < D: not scored
    assert accuracy.score(disp, gt, 16, decimated=True)["mae"] == 0.0


def test_left_ground_truth_of_a_right_field():
    # right(x) = left(x + g(x)); left pixels reached by exactly one right pixel are valid
    g = np.array([[0, 0, 2, 2, 0, 0, 0, 0]], np.uint8)
    gt, valid = accuracy.left_gt_from_right(g)
    # right x=2,3 land on left 4,5, which right x=4,5 (g=0) also reach: ambiguous;
    # left 2,3 are reached by nobody: occluded
    np.testing.assert_array_equal(valid, [[1, 1, 0, 0, 0, 0, 1, 1]])
    g = np.array([[1, 1, 1, 0, 0]], np.uint8)
    gt, valid = accuracy.left_gt_from_right(g)
    np.testing.assert_array_equal(valid, [[0, 1, 1, 0, 1]])
    np.testing.assert_array_equal(gt[valid], [1, 1, 0])


def test_left_ground_truth_matches_a_loop():
    rng = np.random.default_rng(3)
    g = rng.integers(0, 6, (5, 30)).astype(np.uint8)
    gt, valid = accuracy.left_gt_from_right(g)
    for y in range(5):
        hits = {}
        for x in range(30):
            if x + g[y, x] < 30:
                hits.setdefault(x + g[y, x], []).append(int(g[y, x]))
        for xl in range(30):
            h = hits.get(xl, [])
            assert valid[y, xl] == (len(h) == 1)
            if len(h) == 1:
                assert gt[y, xl] == h[0]


def test_shape_mismatch_is_an_error():
    with pytest.raises(ValueError):
        accuracy.score(np.zeros((4, 40), np.uint8), np.zeros((5, 40), np.uint8), 16)
