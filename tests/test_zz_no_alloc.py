"""sgmh_plain_noalloc_dev (the plain path on the library's static scratch array,
no device allocation) against sgmh_plain_dev, called through the C ABI.

Two calls in a row, the second on a smaller frame (and with a different D and
path count), so the second one runs on a scratch array that still holds the
first call's codes and volumes; then a frame too large for the array is
rejected before any GPU work.  The CPU oracle checks the first frame as well.
"""
import numpy as np
import pytest
import torch

import oracle
import sgm_amd

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _run_both(ctx, L, R, D, paths):
    H, W = L.shape
    d_l, d_r = _dev(L), _dev(R)
    want = torch.full_like(d_l, 0xA5)
    got = torch.full_like(d_l, 0x5A)
    s = torch.cuda.current_stream().cuda_stream
    ctx.plain_dev(d_l.data_ptr(), d_r.data_ptr(), W, H, D, paths, want.data_ptr(), s)
    sgm_amd.plain_noalloc_dev(d_l.data_ptr(), d_r.data_ptr(), W, H, D, paths, got.data_ptr(), s)
    torch.cuda.synchronize()
    return got.cpu().numpy(), want.cpu().numpy()


def test_noalloc_matches_plain_dev_twice(ctx):
    cases = [(640, 480, 64, 4, 5), (333, 217, 32, 8, 6)]
    for k, (W, H, D, paths, seed) in enumerate(cases):
        print(f"noalloc case {k}: {W}x{H} D={D} {paths}-path", flush=True)
        assert sgm_amd.noalloc_scratch_bytes(W, H, D, paths) <= sgm_amd.NOALLOC_SCRATCH_BYTES
        L, R, _ = sgm_amd.scene_planes(W, H, D - 1, seed, planes=(8, 12))
        got, want = _run_both(ctx, L, R, D, paths)
        bad = np.argwhere(got != want)
        assert bad.size == 0, f"case {k}: {bad.shape[0]} pixels differ, first at {bad[0]}"
        if k == 0:
            np.testing.assert_array_equal(got, oracle.sgm_plain(L, R, D, paths))


def test_noalloc_rejects_frames_beyond_the_scratch(ctx):
    W, H, D, paths = 1920, 1080, 64, 4
    assert sgm_amd.noalloc_scratch_bytes(W, H, D, paths) > sgm_amd.NOALLOC_SCRATCH_BYTES
    d = torch.zeros((H, W), dtype=torch.uint8, device="cuda")
    with pytest.raises(sgm_amd.SgmError, match="scratch"):
        sgm_amd.plain_noalloc_dev(d.data_ptr(), d.data_ptr(), W, H, D, paths, d.data_ptr())
    with pytest.raises(sgm_amd.SgmError, match="paths"):
        sgm_amd.plain_noalloc_dev(d.data_ptr(), d.data_ptr(), 64, 64, 64, 5, d.data_ptr())
