"""GPU parity of the multiscale path (Configurations B/C)."""
import numpy as np
import pytest

import oracle
import sgm_amd
from conftest import make_scene, rng_image

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


@pytest.mark.parametrize("H,W", [(16, 16), (61, 97), (800, 1920)])
def test_sobel_half(ctx, H, W):
    img = rng_image(H, W, 3 * W + H)
    d_img = dev(img)
    d_out = torch.empty_like(d_img)
    s = torch.cuda.current_stream().cuda_stream
    ctx.sobel_dev(d_img.data_ptr(), W, H, d_out.data_ptr(), s)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_out.cpu().numpy(), oracle.sobelx(img))
    d_h = torch.empty((H // 2, W // 2), dtype=torch.uint8, device="cuda")
    ctx.half_dev(d_img.data_ptr(), W, H, d_h.data_ptr(), s)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_h.cpu().numpy(), oracle.half(img))


@pytest.mark.parametrize("H,W,D", [(32, 32, 64), (62, 98, 64), (120, 200, 64), (64, 160, 128), (40, 64, 32),
                                   (40, 64, 16), (48, 300, 256), (200, 320, 128)])
def test_ms_matches_oracle(ctx, H, W, D):
    L, R, _ = sgm_amd.scene_planes(W, H, D - 1 if D - 1 < W else W - 1, W + H + D, planes=(3, 6))
    got, got1 = ctx.ms(L, R, D, want_stage1=True)
    want, want1 = oracle.sgm_ms(L, R, D, want_stage1=True)
    np.testing.assert_array_equal(got1, want1)
    np.testing.assert_array_equal(got, want)


def test_reference_goldens_ms(ctx, goldens):
    for g in goldens["reference_goldens"]:
        if g["mode"] != "ms":
            continue
        L, R = make_scene(g["scene"])
        out = ctx.ms(L, R, g["D"])
        assert list(out.shape) == g["shape"]
        assert sgm_amd.fnv1a64(out) == g["fnv"], g["name"]


def test_computed_goldens_ms(ctx, goldens):
    for g in goldens["computed"]:
        if g["mode"] != "ms":
            continue
        L, R = make_scene(g["scene"])
        out = ctx.ms(L, R, g["D"])
        assert sgm_amd.fnv1a64(out) == g["fnv"], g["name"]


def test_ms_rejects_odd(ctx):
    L = np.zeros((33, 64), np.uint8)
    with pytest.raises(sgm_amd.SgmError, match="even"):
        ctx.ms(L, L, 64)
