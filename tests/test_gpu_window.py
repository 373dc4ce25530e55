"""Windowed SGM (the learned path's reduced search) against its oracle."""
import numpy as np
import pytest

import oracle
import sgm_amd

pytestmark = pytest.mark.gpu


def tiles(H, W):
    return (H + 15) // 16, (W + 15) // 16


@pytest.mark.parametrize("H,W,D", [(61, 97, 128), (120, 200, 64), (48, 80, 32), (300, 40, 128)])
def test_full_range_is_plain(ctx, H, W, D):
    L, R, _ = sgm_amd.scene_planes(W, H, min(D, W) - 1, H + W, planes=(3, 6))
    blk = np.full(tiles(H, W), (D // 16 - 1) << 4, np.uint8)
    got = ctx.window(L, R, blk, D)
    np.testing.assert_array_equal(got, ctx.plain(L, R, D, 4))
    np.testing.assert_array_equal(got, oracle.sgm_plain(L, R, D, 4))


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("H,W,D", [(61, 97, 128), (96, 160, 64), (40, 70, 32)])
def test_random_windows_match_oracle(ctx, seed, H, W, D):
    rng = np.random.default_rng(seed * 100 + D)
    L, R, _ = sgm_amd.scene_planes(W, H, min(D, W) - 1, seed + 7, planes=(3, 6))
    nb = D // 16
    b0 = rng.integers(0, nb - 1, tiles(H, W))
    width = rng.choice([1, 2, 2, 2, 3], size=b0.shape)
    b1 = np.minimum(b0 + width - 1, nb - 1)
    full = rng.random(b0.shape) < 0.15           # fallback tiles
    b0 = np.where(full, 0, b0)
    b1 = np.where(full, nb - 1, b1)
    blk = (b0 | (b1 << 4)).astype(np.uint8)
    got = ctx.window(L, R, blk, D)
    want = oracle.sgm_window(L, R, blk, D, 4)
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{bad.shape[0]} pixels differ, first at {bad[0]}"


def test_window_rejects_bad_table(ctx):
    L = np.zeros((32, 32), np.uint8)
    blk = np.full((2, 2), 0x90, np.uint8)  # block 9 of 8
    with pytest.raises(sgm_amd.SgmError, match="block range"):
        ctx.window(L, L, blk, 128)
