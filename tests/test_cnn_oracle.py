"""The network's float oracle is pinned by finite differences (bf16 roundings
off) and by closed forms of its features and head."""
import numpy as np
import pytest

import cnn_oracle as co
import sgm_amd


@pytest.fixture
def exact():
    co.ROUND = False
    yield
    co.ROUND = True


def small_case(seed=0, W=48, H=32):
    L, R, _ = sgm_amd.scene_planes(W, H, 40, seed, planes=(2, 3))
    disp = np.random.default_rng(seed).integers(0, 128, (H, W)).astype(np.uint8)
    return L, R, disp


def test_gradients_match_finite_differences(exact):
    L, R, disp = small_case()
    theta = co.init_params(3).astype(np.float64)
    theta[co.C * co.C * 9: co.C * co.C * 9 + co.C] = 0.05  # non-zero biases
    loss, g = co.loss_and_grads(theta, L, R, disp)
    rng = np.random.default_rng(1)
    # probe parameters in every layer, weights and biases
    idx = []
    for layer in range(co.LAYERS):
        base = layer * (co.C * co.C * 9 + co.C)
        idx += list(base + rng.integers(0, co.C * co.C * 9, 4)) + [base + co.C * co.C * 9 + int(rng.integers(0, co.C))]
    th64 = theta.astype(np.float64)
    for i in idx:
        h = 2e-3 * max(1e-2, abs(float(th64[i])))
        tp, tm = th64.copy(), th64.copy()
        tp[i] += h
        tm[i] -= h
        lp, _ = co.loss_and_grads(tp, L, R, disp)
        lm, _ = co.loss_and_grads(tm, L, R, disp)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-3 * abs(fd) + 1e-7, (i, fd, g[i])


def test_features_closed_form():
    L = np.full((16, 32), 10, np.uint8)
    R = np.full((16, 32), 13, np.uint8)
    f = co.features(L, R)  # |16*10 - 16*13| = 48 everywhere x-k >= 0
    assert np.allclose(f[:, 5:, 0], co.bf16(np.float32(48 * 9) / np.float32(9 * 4080)))
    # bin k at x < k (and its 3x3 neighbourhood) sees the 4080 out-of-image difference
    assert np.isclose(f[0, 0, 5], 1.0)


def test_head_window_rule(monkeypatch):
    monkeypatch.setattr(co, "SKIP", 0.0)  # the rule alone, without the feature residual
    theta = np.zeros(co.NPARAM, np.float32)
    # last-layer bias puts all mass on bins 12..15 -> blocks 3 and 4 have the
    # window [48, 80), i.e. b0 = 3
    theta[5 * (co.C * co.C * 9 + co.C) + co.C * co.C * 9 + 12: 5 * (co.C * co.C * 9 + co.C) + co.C * co.C * 9 + 16] = 20.0
    L = np.zeros((32, 64), np.uint8)
    blk, conf = co.propose(theta, L, L)
    assert (conf > 0.99).all()
    assert ((blk & 15) == 2).all() or ((blk & 15) == 3).all()


def test_feature_residual_orders_bins():
    """With every weight zero the logits are -SKIP * features: the tile's best
    bin is the one whose matching cost is lowest."""
    L, R, _ = sgm_amd.scene_planes(64, 32, 40, 5, planes=(1, 1))
    theta = np.zeros(co.NPARAM, np.float32)
    f = co.features(L, R)
    _, z = co.forward(theta, f)
    np.testing.assert_allclose(z, -co.SKIP * f, rtol=1e-6)


def test_targets_histogram():
    d = np.zeros((16, 32), np.uint8)
    d[:, 16:] = 9
    d[0, 0] = 127
    q = co.targets(d)
    assert q.shape == (4, 8, 32)
    assert q[0, 0, 0] == 15 / 16 and q[0, 0, 31] == 1 / 16 and q[1, 0, 0] == 1.0
    assert q[0, 4, 2] == 1.0 and np.allclose(q.sum(-1), 1.0)


def test_tile_probs_is_mean_of_pixel_softmax():
    z = np.random.default_rng(0).normal(0, 2, (5, 6, 32)).astype(np.float32)
    P = co.tile_probs(z)
    assert P.shape == (2, 2, 32)
    p = co.softmax(z)
    np.testing.assert_allclose(P[1, 1], p[4:5, 4:6].mean(axis=(0, 1)), rtol=1e-5)
    np.testing.assert_allclose(P.sum(-1), 1.0, rtol=1e-5)
