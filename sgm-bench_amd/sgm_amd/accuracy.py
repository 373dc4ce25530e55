"""Accuracy of a disparity map against dense ground truth, the way the reference
scores its real scenes (scripts/make_panel.py:34-50, README.md:372-378):

  err = |disp - gt|;  scoreable = valid pixels outside the leftmost D columns
  bad>1px = share of scoreable pixels with err > 1
  bad>2px = share with err > 2
  MAE     = mean err over the scoreable pixels

The leftmost D columns are excluded because x - d < 0 has no correspondent in
the right image there (make_panel.py:41-46).

Ground truth of the synthetic scenes:
  * the reference generator (sgm_scene_reference) gives left-image disparities
    directly, every pixel valid (common/gen_synthetic.c:100-112);
  * the planes generator gives the disparity field in RIGHT-image coordinates,
    right(x) = left(x + g(x)).  Its left-image ground truth is the forward warp
    gtL(x + g(x)) = g(x).  A left pixel is valid where exactly one right pixel
    reaches it: where none does it is occluded in the right image, and where
    two do (at plane edges) both disparities match it exactly, so neither is
    the one answer.  Right pixels the generator replaced by fresh texture stay
    in the valid set: they are part of what the scene asks the matcher to get
    right.
  * the multiscale output is the decimated (W/2)x(H/2) argmin at (2y, 2x) in
    full-resolution disparities (multiscale/sgm_ms_ref.c:194-199), so it is
    scored against gtL[2y, 2x], with the columns 2x < D excluded.
"""
from __future__ import annotations

import numpy as np


def left_gt_from_right(g: np.ndarray):
    """(gtL, valid) from a right-image disparity field g (planes generator)."""
    g = np.asarray(g).astype(np.int64)
    H, W = g.shape
    xs = np.arange(W)[None, :] + g
    rows = np.broadcast_to(np.arange(H)[:, None], (H, W))
    inside = xs < W
    flat = rows[inside] * W + xs[inside]
    count = np.bincount(flat, minlength=H * W).reshape(H, W)
    gt = np.zeros(H * W, np.int64)
    gt[flat] = g[inside]  # where count == 1 the one writer is the value
    valid = count == 1
    return np.where(valid, gt.reshape(H, W), 0).astype(np.uint8), valid


def score(disp: np.ndarray, gt: np.ndarray, D: int, valid: np.ndarray | None = None,
          decimated: bool = False) -> dict:
    """bad>1px %, bad>2px %, MAE and the scored pixel count (make_panel.py:34-50).

    disp: the engine's output; gt, valid: full-resolution left-image ground
    truth; decimated: disp is the multiscale (W/2)x(H/2) output."""
    gt = np.asarray(gt)
    if valid is None:
        valid = np.ones(gt.shape, bool)
    if decimated:
        gt = gt[0::2, 0::2][:disp.shape[0], :disp.shape[1]]
        valid = valid[0::2, 0::2][:disp.shape[0], :disp.shape[1]]
        xfull = 2 * np.arange(disp.shape[1])
    else:
        xfull = np.arange(disp.shape[1])
    if disp.shape != gt.shape:
        raise ValueError(f"disparity {disp.shape} and ground truth {gt.shape} differ in shape")
    scoreable = valid & (xfull >= D)[None, :]
    n = int(scoreable.sum())
    if n == 0:
        raise ValueError("no scoreable pixels (image narrower than D?)")
    err = np.abs(disp.astype(np.float64) - gt.astype(np.float64))[scoreable]
    return {"bad1_pct": round(100.0 * float((err > 1.0).sum()) / n, 3),
            "bad2_pct": round(100.0 * float((err > 2.0).sum()) / n, 3),
            "mae": round(float(err.mean()), 4),
            "scored_px": n}


def score_planes(disp: np.ndarray, g_right: np.ndarray, D: int, decimated: bool = False) -> dict:
    """score() against a planes scene's right-image disparity field."""
    gt, valid = left_gt_from_right(g_right)
    return score(disp, gt, D, valid, decimated)
