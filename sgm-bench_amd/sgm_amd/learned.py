"""The learned disparity-window path: training data, training loop, proposal
and evaluation around the engine's network (sgmh_cnn) and windowed SGM
(sgmh_window_dev).

Pipeline for one frame (1920x1080, D = 128, 4 paths):
  census -> network proposal (per 16x16 tile: 32-disparity window + confidence;
  below the threshold the tile keeps all 128 disparities) -> windowed SGM.
Training pairs are seeded slanted-plane scenes (disparities in [0, 127],
10-30 planes) labelled by the engine's own plain full-D SGM.
"""
from __future__ import annotations

import os
import time
from concurrent.futures import ThreadPoolExecutor
from typing import Iterable, List

import numpy as np

from . import Cnn, Context, scene_planes

D = 128
MODEL_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "models")
MODEL_PATH = os.path.join(MODEL_DIR, "window_cnn_d128.npy")
TRAIN_SEED0 = 100000   # training scenes: seeds TRAIN_SEED0 + i
EVAL_SEED0 = 900000    # disjoint evaluation scenes


def scene(seed: int, W: int = 1920, H: int = 1080):
    L, R, _ = scene_planes(W, H, D - 1, seed, planes=(10, 30), occl_permille=0, noise=2, slanted=True)
    return L, R


def scenes(seeds: Iterable[int], W: int, H: int, workers: int = 16):
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(lambda s: scene(s, W, H), seeds))


def init_params(seed: int = 1) -> np.ndarray:
    """He-normal weights, zero biases, in the engine's parameter layout."""
    rng = np.random.default_rng(seed)
    parts = []
    for _ in range(6):
        parts.append(rng.normal(0, np.sqrt(2.0 / (32 * 9)), 32 * 32 * 9).astype(np.float32))
        parts.append(np.zeros(32, np.float32))
    return np.concatenate(parts)


def load_params() -> np.ndarray:
    return np.load(MODEL_PATH, allow_pickle=False)


def train(ctx: Context, net: Cnn, n_pairs: int = 2000, epochs: int = 1, lr: float = 2e-3,
          W: int = 1920, H: int = 1080, chunk: int = 100, log=print, first_pair: int = 0) -> List[float]:
    """Adam on the per-pixel cross-entropy; every pair (seeds TRAIN_SEED0 +
    first_pair ...) is generated, labelled by the plain full-D SGM on the GPU
    and used once per epoch."""
    import torch
    s = torch.cuda.current_stream().cuda_stream
    dl = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    dr = torch.empty_like(dl)
    dt = torch.empty_like(dl)
    losses = []
    t0 = time.time()
    step = 0
    total = n_pairs * epochs
    for ep in range(epochs):
        for c0 in range(0, n_pairs, chunk):
            batch = scenes(range(TRAIN_SEED0 + first_pair + c0, TRAIN_SEED0 + first_pair + min(n_pairs, c0 + chunk)),
                           W, H)
            for L, R in batch:
                dl.copy_(torch.from_numpy(L))
                dr.copy_(torch.from_numpy(R))
                ctx.plain_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, 4, dt.data_ptr(), s)
                # cosine decay to 10%
                cur = lr * (0.1 + 0.9 * 0.5 * (1 + np.cos(np.pi * step / max(1, total - 1))))
                losses.append(net.train_step_dev(dl.data_ptr(), dr.data_ptr(), dt.data_ptr(), W, H,
                                                 float(cur), s))
                step += 1
            log(f"epoch {ep} pairs {min(n_pairs, c0 + chunk)}/{n_pairs} loss "
                f"{np.mean(losses[-chunk:]):.4f} ({time.time() - t0:.0f} s)")
    return losses


def heuristic_blk(L: np.ndarray, R: np.ndarray, thresh: float = 0.9, with_conf: bool = False):
    """The same proposal rule without the network: per tile, the histogram of
    the per-pixel argmin of the network's input features (the low-resolution
    3x3 SAD over 32 bins, best of the 4 sub-block phases) as the
    distribution.  The baseline the network must beat."""
    H, W = L.shape

    def bsum(img, ph):
        a = img.astype(np.int64)[:, np.maximum(np.arange(W) - ph, 0)]
        return a.reshape(H // 4, 4, W // 4, 4).sum(axis=(1, 3))

    Ld = bsum(L, 0)
    Rds = [bsum(R, ph) for ph in range(4)]
    h, w = Ld.shape
    f = np.empty((h, w, 32), np.int64)
    xs = np.arange(w)
    for k in range(32):
        best = None
        for Rd in Rds:
            diff = np.full((h, w), 4080, np.int64)
            ok = xs >= k
            diff[:, ok] = np.abs(Ld[:, ok] - Rd[:, xs[ok] - k])
            pad = np.pad(diff, 1, mode="edge")
            acc = sum(pad[dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3))
            best = acc if best is None else np.minimum(best, acc)
        f[:, :, k] = best
    am = f.argmin(-1)
    th, tw = (h + 3) // 4, (w + 3) // 4
    p = np.zeros((th, tw, 32))
    for y in range(h):
        np.add.at(p[y // 4], (xs // 4, am[y]), 1)
    p /= p.sum(-1, keepdims=True)
    win = np.stack([p[..., 4 * b:4 * b + 8].sum(-1) for b in range(7)], -1)
    b0, conf = win.argmax(-1), win.max(-1)
    blk = np.where(conf >= thresh, b0 | ((b0 + 1) << 4), 7 << 4).astype(np.uint8)
    return (blk, b0, conf) if with_conf else blk


def blk_at_fallback(b0: np.ndarray, conf: np.ndarray, fallback: float) -> np.ndarray:
    """Block table with the given fraction of tiles at full D: the least
    confident ones (ties broken by tile order).  Compares proposals at equal
    cost."""
    n = conf.size
    k = int(round(fallback * n))
    order = np.argsort(conf.ravel(), kind="stable")
    full = np.zeros(n, bool)
    full[order[:k]] = True
    b = b0.ravel()
    return np.where(full, 7 << 4, b | ((b + 1) << 4)).astype(np.uint8).reshape(conf.shape)


def oracle_blk(plain: np.ndarray, cover: float = 1.0) -> np.ndarray:
    """The best windows in hindsight, from the plain D=128 disparities: per
    tile the 32-disparity window holding the most of its pixels, kept where it
    holds at least `cover` of them (else all blocks).  No proposal can do
    better at the same fallback rate; it bounds what the learned path can
    reach on a given scene."""
    H, W = plain.shape
    th, tw = (H + 15) // 16, (W + 15) // 16
    hist = np.zeros((th, tw, 8))
    blkid = plain.astype(np.int64) // 16
    for y in range(H):
        np.add.at(hist[y // 16], (np.arange(W) // 16, blkid[y]), 1)
    frac = hist / hist.sum(-1, keepdims=True)
    win = np.stack([frac[..., b] + frac[..., b + 1] for b in range(7)], -1)
    b0, got = win.argmax(-1), win.max(-1)
    return np.where(got >= cover, b0 | ((b0 + 1) << 4), 7 << 4).astype(np.uint8)


def evaluate(ctx: Context, net: Cnn, seeds: Iterable[int], thresh: float = 0.9, W: int = 1920,
             H: int = 1080, reps: int = 20) -> dict:
    """Per frame: the learned path (proposal + windowed SGM) against plain D=128:
    percent of pixels that differ, mean |d - d_plain|, tiles that fell back,
    and the time of both paths (median over reps, HIP events)."""
    import torch
    s = torch.cuda.current_stream().cuda_stream
    dl = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    dr = torch.empty_like(dl)
    dp = torch.empty_like(dl)
    dw = torch.empty_like(dl)
    T = ((H + 15) // 16) * ((W + 15) // 16)
    db = torch.empty(T, dtype=torch.uint8, device="cuda")
    dh = torch.empty(T, dtype=torch.uint8, device="cuda")
    diffs, mads, falls, hdiffs, hfalls = [], [], [], [], []
    odiffs, ofalls = [], []
    for seed in seeds:
        L, R = scene(seed, W, H)
        dl.copy_(torch.from_numpy(L))
        dr.copy_(torch.from_numpy(R))
        ctx.plain_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, 4, dp.data_ptr(), s)
        net.propose_dev(dl.data_ptr(), dr.data_ptr(), W, H, thresh, db.data_ptr(), 0, s)
        ctx.window_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, db.data_ptr(), dw.data_ptr(), s)
        torch.cuda.synchronize()
        p, w, b = dp.cpu().numpy(), dw.cpu().numpy(), db.cpu().numpy()
        diffs.append(float((p != w).mean() * 100))
        mads.append(float(np.abs(p.astype(int) - w.astype(int)).mean()))
        falls.append(float(((b >> 4) - (b & 15) == 7).mean() * 100))
        hb = heuristic_blk(L, R, thresh).ravel()
        dh.copy_(torch.from_numpy(hb))
        ctx.window_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, dh.data_ptr(), dw.data_ptr(), s)
        torch.cuda.synchronize()
        hdiffs.append(float((dw.cpu().numpy() != p).mean() * 100))
        hfalls.append(float(((hb >> 4) - (hb & 15) == 7).mean() * 100))
        ob = oracle_blk(p).ravel()
        dh.copy_(torch.from_numpy(ob))
        ctx.window_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, dh.data_ptr(), dw.data_ptr(), s)
        torch.cuda.synchronize()
        odiffs.append(float((dw.cpu().numpy() != p).mean() * 100))
        ofalls.append(float(((ob >> 4) - (ob & 15) == 7).mean() * 100))

    def timed(fn):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))

    t_plain = timed(lambda: ctx.plain_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, 4, dp.data_ptr(), s))
    t_prop = timed(lambda: net.propose_dev(dl.data_ptr(), dr.data_ptr(), W, H, thresh, db.data_ptr(), 0, s))
    t_win = timed(lambda: ctx.window_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, db.data_ptr(),
                                         dw.data_ptr(), s))

    def both():
        net.propose_dev(dl.data_ptr(), dr.data_ptr(), W, H, thresh, db.data_ptr(), 0, s)
        ctx.window_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, db.data_ptr(), dw.data_ptr(), s)

    t_learned = timed(both)
    return {"frames": len(diffs), "thresh": thresh,
            "pct_pixels_differ": float(np.mean(diffs)), "mean_abs_diff": float(np.mean(mads)),
            "pct_tiles_full_D": float(np.mean(falls)),
            "heuristic_pct_pixels_differ": float(np.mean(hdiffs)),
            "heuristic_pct_tiles_full_D": float(np.mean(hfalls)),
            "hindsight_pct_pixels_differ": float(np.mean(odiffs)),
            "hindsight_pct_tiles_full_D": float(np.mean(ofalls)),
            "ms_plain_d128": t_plain, "ms_proposal": t_prop, "ms_window_sgm": t_win,
            "ms_learned_total": t_learned, "speedup_vs_plain": t_plain / t_learned}


def main():
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train / evaluate the window-proposal network")
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=1)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval", type=int, default=8, help="evaluation frames")
    ap.add_argument("--out", default=MODEL_PATH)
    ap.add_argument("--eval-only", action="store_true")
    ap.add_argument("--resume", action="store_true", help="train on from the committed model")
    ap.add_argument("--first-pair", type=int, default=0, help="training seeds start at TRAIN_SEED0 + this")
    ap.add_argument("--thresh", default="0.9,0.8,0.95", help="comma-separated confidence thresholds to evaluate")
    a = ap.parse_args()
    ctx = Context(0)
    net = Cnn(ctx, load_params() if (a.eval_only or a.resume) else init_params())
    if not a.eval_only:
        train(ctx, net, a.pairs, a.epochs, a.lr, first_pair=a.first_pair)
        os.makedirs(os.path.dirname(a.out), exist_ok=True)
        np.save(a.out, net.params(), allow_pickle=False)
        print("saved", a.out)
    for th in (float(t) for t in a.thresh.split(",")):
        print(json.dumps(evaluate(ctx, net, range(EVAL_SEED0, EVAL_SEED0 + a.eval), th)), flush=True)


if __name__ == "__main__":
    main()
