"""cnn_oracle.py -- float reference of the window-proposal network.

TEST INFRASTRUCTURE ONLY (the checker for the HIP CNN kernels; see
oracle/oracle.py for what may use the oracle).  The network has no reference
counterpart: it is defined here and by its HIP implementation together.

Definition (D = 128, frames with W and H multiples of 4):
  features   Ld = 4x4 block sums of the left image (h = H/4, w = W/4);
             Rd_p = 4x4 block sums of the right image at phase p = 0..3
             (block x covers columns 4x-p .. 4x-p+3, clamped at x = 0).
             For bin k in 0..31 (full-resolution disparities 4k + p):
             f_k(x, y) = min over p of the sum over the clamped 3x3
             neighbourhood (x', y') of |Ld(x', y') - Rd_p(x' - k, y')|, the
             difference taken as 4080 (= 16*255) where x' - k < 0, divided by
             9*4080; rounded to bf16.  Layout [h][w][32].
  layers     6 x (3x3 conv, 32 -> 32 channels, zero padding, fp32 bias);
             layers 1-5 ReLU and bf16 output, layer 6 linear fp32 logits,
             to which the fixed residual -SKIP * features is added (a low
             matching cost is a high logit before the network's correction).
             Inputs and weights enter the products as bf16, sums in fp32.
  head       per low-resolution pixel p = softmax(logits); per 16x16-pixel
             tile (4x4 low-resolution pixels, fewer at the bottom / right
             edge) P_t = the mean of its pixels' p.  Window start b0
             (16-disparity blocks) maximises conf = sum of P_t over bins
             4*b0 .. 4*b0+7; below the confidence threshold the tile falls
             back to all blocks.
  loss       per-pixel cross-entropy -sum_k q_k log p_k, averaged over the
             low-resolution pixels, with q the histogram of the target
             disparities of the pixel's 4x4 block in bins of 4.
Parameters are stored flat in fp32: for each layer W[co][ci][tap] (tap =
3*ky + kx) then b[co].
"""
from __future__ import annotations

import numpy as np

C = 32          # channels / bins
SKIP = 32.0     # logits = last conv - SKIP * features (a fixed residual from the matching costs)
ROUND = True    # bf16 roundings on; off only for finite-difference checks
LAYERS = 6
NPARAM = LAYERS * (C * C * 9 + C)


def ft():
    """Working float type: float32 (the device's), float64 with roundings off."""
    return np.float32 if ROUND else np.float64


def bf16(a):
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    if not ROUND:
        return np.ascontiguousarray(a, np.float64)
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def split_params(theta):
    out, o = [], 0
    for _ in range(LAYERS):
        w = theta[o:o + C * C * 9].reshape(C, C, 9)
        o += C * C * 9
        b = theta[o:o + C]
        o += C
        out.append((w, b))
    return out


def init_params(seed: int) -> np.ndarray:
    """He-normal weights, zero biases (the same stream is not required on the
    GPU; weights are set explicitly there)."""
    rng = np.random.default_rng(seed)
    parts = []
    for _ in range(LAYERS):
        parts.append(rng.normal(0, np.sqrt(2.0 / (C * 9)), C * C * 9).astype(np.float32))
        parts.append(np.zeros(C, np.float32))
    return np.concatenate(parts)


def blocksum(img, phase=0):
    """4x4 block sums; block x covers columns 4x - phase .. 4x - phase + 3,
    clamped at the left edge."""
    a = np.asarray(img, np.int64)
    H, W = a.shape
    cols = np.maximum(np.arange(W) - phase, 0)
    a = a[:, cols]
    return a.reshape(H // 4, 4, W // 4, 4).sum(axis=(1, 3))


def features(left, right, phases=4):
    Ld = blocksum(left)
    h, w = Ld.shape
    f = np.empty((h, w, C), ft())
    xs = np.arange(w)
    for k in range(C):
        best = None
        for ph in range(phases):
            Rd = blocksum(right, ph)
            diff = np.full((h, w), 4080, np.int64)
            ok = xs >= k
            diff[:, ok] = np.abs(Ld[:, ok] - Rd[:, xs[ok] - k])
            pad = np.pad(diff, 1, mode="edge")
            acc = np.zeros((h, w), np.int64)
            for dy in range(3):
                for dx in range(3):
                    acc += pad[dy:dy + h, dx:dx + w]
            best = acc if best is None else np.minimum(best, acc)
        f[:, :, k] = best.astype(ft()) / ft()(9 * 4080)
    return bf16(f)


def conv(x, w, b):
    """3x3 'same' conv, x [h][w][ci] (bf16 values), w [co][ci][9]; fp32 out."""
    h, wd, _ = x.shape
    xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
    wb = bf16(w)
    y = np.zeros((h, wd, w.shape[0]), ft())
    for t in range(9):
        ky, kx = divmod(t, 3)
        y += xp[ky:ky + h, kx:kx + wd, :] @ wb[:, :, t].T
    return y + b


def forward(theta, x0):
    acts = [x0]
    x = x0
    pre = []
    for i, (w, b) in enumerate(split_params(theta)):
        z = conv(x, w, b)
        pre.append(z)
        if i < LAYERS - 1:
            x = bf16(np.maximum(z, 0))
            acts.append(x)
        else:
            x = z - ft()(SKIP) * x0
    return acts, x   # acts[i] = input of layer i; x = logits


def softmax(z):
    e = np.exp(z - z.max(axis=-1, keepdims=True))
    return e / e.sum(axis=-1, keepdims=True)


def tile_probs(logits):
    """Per 16x16 tile (4x4 low-resolution pixels, fewer at the edges): the
    mean of its pixels' softmax distributions over the 32 bins."""
    p = softmax(logits)
    h, w, _ = p.shape
    th, tw = (h + 3) // 4, (w + 3) // 4
    P = np.zeros((th, tw, C), ft())
    n = np.zeros((th, tw, 1), ft())
    for y in range(h):
        for x in range(w):
            P[y // 4, x // 4] += p[y, x]
            n[y // 4, x // 4] += 1
    return P / n


def targets(disp):
    """Per low-resolution pixel: the histogram of the target disparities of
    its 4x4 full-resolution block, in bins of 4 (D = 128)."""
    d = np.minimum(np.asarray(disp, np.int64) // 4, C - 1)
    H, W = d.shape
    h, w = H // 4, W // 4
    q = np.zeros((h, w, C), ft())
    blk = d[:4 * h, :4 * w].reshape(h, 4, w, 4).transpose(0, 2, 1, 3).reshape(h, w, 16)
    for k in range(C):
        q[..., k] = (blk == k).sum(-1)
    return q / 16


def loss_and_grads(theta, left, right, disp):
    """Mean per-pixel cross-entropy against the block histograms and its
    gradient w.r.t. theta (fp32 backprop; the bf16 roundings of the forward are
    treated as identity)."""
    theta = np.asarray(theta, ft())
    x0 = features(left, right)
    acts, logits = forward(theta, x0)
    p = softmax(logits)
    q = targets(disp)
    h, w, _ = logits.shape
    npix = h * w
    loss = float(-(q * np.log(np.maximum(p, 1e-30))).sum() / npix)
    dy = ((p - q) / npix).astype(ft())
    params = split_params(theta)
    grads = [None] * LAYERS
    for i in range(LAYERS - 1, -1, -1):
        w, b = params[i]
        x = acts[i]
        hh, ww, _ = x.shape
        xp = np.pad(x, ((1, 1), (1, 1), (0, 0)))
        gw = np.zeros_like(w)
        for t in range(9):
            ky, kx = divmod(t, 3)
            gw[:, :, t] = dy.reshape(-1, C).T @ xp[ky:ky + hh, kx:kx + ww, :].reshape(-1, C)
        gb = dy.sum(axis=(0, 1))
        grads[i] = (gw, gb)
        if i > 0:
            dx = np.zeros((hh + 2, ww + 2, C), ft())
            wb = bf16(w)
            for t in range(9):
                ky, kx = divmod(t, 3)
                dx[ky:ky + hh, kx:kx + ww, :] += dy @ wb[:, :, t]
            dx = dx[1:-1, 1:-1, :]
            dy = dx * (acts[i] > 0)
    flat = []
    for gw, gb in grads:
        flat += [gw.ravel(), gb]
    return loss, np.concatenate(flat).astype(ft())


def propose(theta, left, right, thresh=0.9):
    """Per-tile block ranges (b0 | b1 << 4) and confidences, D = 128."""
    _, logits = forward(theta, features(left, right))
    p = tile_probs(logits)
    win = np.stack([p[..., 4 * b:4 * b + 8].sum(-1) for b in range(7)], -1)
    b0 = win.argmax(-1)
    conf = win.max(-1)
    blk = np.where(conf >= thresh, b0 | ((b0 + 1) << 4), 7 << 4).astype(np.uint8)
    return blk, conf.astype(np.float32)
