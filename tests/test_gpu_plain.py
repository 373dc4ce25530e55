"""GPU parity of the plain path: every phase and the whole pipeline against the
CPU oracle, and against the reference's golden hashes, through the C ABI."""
import numpy as np
import pytest

import oracle
import sgm_amd
from conftest import make_scene, rng_image

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def stream():
    return torch.cuda.current_stream().cuda_stream


@pytest.mark.parametrize("H,W", [(16, 16), (61, 97), (480, 640), (1080, 1920)])
def test_census_descriptors(ctx, H, W):
    img = rng_image(H, W, W + H)
    d_img = dev(img)
    d_codes = torch.empty((H, W), dtype=torch.int64, device="cuda")
    ctx.census_dev(d_img.data_ptr(), W, H, d_codes.data_ptr(), stream())
    torch.cuda.synchronize()
    got = d_codes.cpu().numpy().view(np.uint64)
    want = oracle.census(img)
    diff = np.argwhere(got != want)
    assert diff.size == 0, f"first differing descriptor at {diff[0]}: got {got[tuple(diff[0])]:x} want {want[tuple(diff[0])]:x}"


DIRS = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, 1), (1, -1), (-1, -1)]


@pytest.fixture
def layout_ctx(ctx, request):
    """The shared context with its plain path layout set for one test."""
    ctx.set_path_layout(request.param)
    try:
        yield ctx
    finally:
        ctx.set_path_layout(sgm_amd.LAYOUT_AUTO)


LAYOUTS = [pytest.param(sgm_amd.LAYOUT_LINES4, id="lines4"), pytest.param(sgm_amd.LAYOUT_WAVE, id="wave")]
# the placed launch (one workgroup per CU, waves take row or column groups by
# their SIMD, queues reset by the last wave) runs 4-path frames at D = 32, 64
# and 128; the others fall back to the four-line launch
LAYOUTS_ALL = LAYOUTS + [pytest.param(sgm_amd.LAYOUT_PLACED, id="placed")]


def check_path_volumes(ctx, D, H, W):
    L, R = rng_image(H, W, 7 * D + W), rng_image(H, W, 7 * D + W + 1)
    cl, cr = oracle.census(L), oracle.census(R)
    C = oracle.cost(cl, cr, D)
    d_cl, d_cr = dev(cl.view(np.int64)), dev(cr.view(np.int64))
    d_L = torch.empty((H, W, D), dtype=torch.uint8, device="cuda")
    for dx, dy in DIRS:
        d_L.fill_(0xAB)
        ctx.path_plain_dev(d_cl.data_ptr(), d_cr.data_ptr(), W, H, D, dx, dy, d_L.data_ptr(), stream())
        torch.cuda.synchronize()
        got = d_L.cpu().numpy()
        want = oracle.path_L(C, dx, dy)
        bad = np.argwhere(got != want)
        assert bad.size == 0, f"dir ({dx},{dy}) D={D}: first mismatch at {bad[0]}"


@pytest.mark.parametrize("D", [16, 32])
@pytest.mark.parametrize("H,W", [(23, 77), (40, 300)])
def test_path_volumes(ctx, D, H, W):
    check_path_volumes(ctx, D, H, W)


# both scanline layouts of the plain path kernel, every direction
@pytest.mark.parametrize("layout_ctx", LAYOUTS, indirect=True)
@pytest.mark.parametrize("D", [64, 128, 256])
@pytest.mark.parametrize("H,W", [(23, 77), (40, 300), (70, 17)])
def test_path_volumes_layouts(layout_ctx, D, H, W):
    check_path_volumes(layout_ctx, D, H, W)


@pytest.mark.parametrize("D", [16, 32])
def test_wave_layout_rejects_narrow_D(ctx, D):
    L, R, _ = sgm_amd.scene_planes(64, 32, D - 1, 5, planes=(3, 6))
    ctx.set_path_layout(sgm_amd.LAYOUT_WAVE)
    try:
        with pytest.raises(sgm_amd.SgmError, match="wave-line layout"):
            ctx.plain(L, R, D, 4)
    finally:
        ctx.set_path_layout(sgm_amd.LAYOUT_AUTO)
    # the context still works after the error
    np.testing.assert_array_equal(ctx.plain(L, R, D, 4), oracle.sgm_plain(L, R, D, 4))


@pytest.mark.parametrize("H,W,D,paths", [
    (16, 16, 16, 4), (17, 33, 32, 4), (61, 97, 64, 4), (61, 97, 64, 8), (300, 40, 64, 4),
    (120, 200, 128, 4), (120, 200, 128, 8), (20, 1000, 256, 4), (64, 64, 256, 8),
    # D = 64 frames of 512 rows or more take the broadcast row variant
    (515, 77, 64, 4), (513, 130, 64, 8)])
def test_plain_matches_oracle(ctx, H, W, D, paths):
    L, R, _ = sgm_amd.scene_planes(W, H, min(D - 1, W - 1), W * H + D, planes=(3, 6))
    got = ctx.plain(L, R, D, paths)
    want = oracle.sgm_plain(L, R, D, paths)
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{bad.shape[0]} pixels differ, first at {bad[0]}"


@pytest.mark.parametrize("layout_ctx", LAYOUTS_ALL, indirect=True)
@pytest.mark.parametrize("H,W,D", [(61, 97, 64), (300, 40, 64), (120, 200, 128), (20, 1000, 256),
                                   (515, 77, 64), (480, 640, 64)])
def test_plain_layouts_match_oracle(layout_ctx, H, W, D):
    L, R, _ = sgm_amd.scene_planes(W, H, min(D - 1, W - 1), W * H + 3 * D, planes=(3, 6))
    got = layout_ctx.plain(L, R, D, 4)
    want = oracle.sgm_plain(L, R, D, 4)
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{bad.shape[0]} pixels differ, first at {bad[0]}"


def test_plain_dev_matches_host_entry(ctx):
    L, R = make_scene({"kind": "reference", "W": 640, "H": 480, "gen_D": 64, "seed": 1})
    host = ctx.plain(L, R, 64, 4)
    d_l, d_r = dev(L), dev(R)
    d_o = torch.empty_like(d_l)
    ctx.plain_dev(d_l.data_ptr(), d_r.data_ptr(), 640, 480, 64, 4, d_o.data_ptr(), stream())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d_o.cpu().numpy(), host)


def test_reference_goldens_plain(ctx, goldens):
    """The reference's own on-disk goldens, input regenerated on the fly."""
    for g in goldens["reference_goldens"]:
        if g["mode"] != "plain":
            continue
        L, R = make_scene(g["scene"])
        out = ctx.plain(L, R, g["D"], g["paths"])
        assert sgm_amd.fnv1a64(out) == g["fnv"], g["name"]


def test_computed_goldens_plain(ctx, goldens):
    """Hashes of the compiled reference oracle on this repository's scenes."""
    for g in goldens["computed"]:
        if g["mode"] != "plain":
            continue
        L, R = make_scene(g["scene"])
        assert sgm_amd.fnv1a64(np.concatenate([L.ravel(), R.ravel()])) == g["input_fnv"], g["name"]
        out = ctx.plain(L, R, g["D"], g["paths"])
        assert sgm_amd.fnv1a64(out) == g["fnv"], g["name"]


@pytest.mark.parametrize("layout_ctx", LAYOUTS_ALL, indirect=True)
def test_goldens_both_layouts(layout_ctx, goldens):
    """Every plain golden with D = 64, 128 or 256 (the reference's on-disk ones
    and the compiled reference's hashes, 4 and 8 paths, up to 1920x1080) with
    the path kernel's scanline layout forced."""
    gs = [g for g in goldens["reference_goldens"] + goldens["computed"] + goldens["ties"]
          if g["mode"] == "plain" and g["D"] % 64 == 0]
    assert len(gs) >= 20
    for g in gs:
        L, R = make_scene(g["scene"])
        out = layout_ctx.plain(L, R, g["D"], g["paths"])
        assert sgm_amd.fnv1a64(out) == g["fnv"], g["name"]


def test_tie_fixtures(ctx, goldens):
    """Scenes where 8-43 % of pixels have tied path-sum minima (a texture that
    repeats every 8-24 columns, flat patches): bit-exact to the compiled
    reference, which pins the argmin's lowest-d rule (common/sgm_ref.c:120-124)
    on the GPU, through both entry points."""
    s = stream()
    for g in goldens["ties"]:
        L, R = make_scene(g["scene"])
        assert sgm_amd.fnv1a64(np.concatenate([L.ravel(), R.ravel()])) == g["input_fnv"], g["name"]
        out = ctx.plain(L, R, g["D"], g["paths"])
        assert sgm_amd.fnv1a64(out) == g["fnv"], (g["name"], g["tie_pct"])
        d_l, d_r = dev(L), dev(R)
        d_o = torch.empty_like(d_l)
        ctx.plain_dev(d_l.data_ptr(), d_r.data_ptr(), L.shape[1], L.shape[0], g["D"], g["paths"], d_o.data_ptr(), s)
        torch.cuda.synchronize()
        assert sgm_amd.fnv1a64(d_o.cpu().numpy()) == g["fnv"], g["name"]


def test_repeat_and_resize(ctx):
    """Workspaces are reused across calls and regrown on a size change."""
    a = sgm_amd.scene_planes(300, 200, 63, 1)
    b = sgm_amd.scene_planes(700, 500, 63, 2)
    ra = oracle.sgm_plain(a[0], a[1], 64, 4)
    rb = oracle.sgm_plain(b[0], b[1], 64, 4)
    for L, R, want in ((a[0], a[1], ra), (b[0], b[1], rb), (a[0], a[1], ra), (a[0], a[1], ra)):
        np.testing.assert_array_equal(ctx.plain(L, R, 64, 4), want)


def test_errors_are_reported(ctx):
    L = np.zeros((32, 32), np.uint8)
    with pytest.raises(sgm_amd.SgmError, match="D=48"):
        ctx.plain(L, L, 48, 4)
    with pytest.raises(sgm_amd.SgmError, match="paths"):
        ctx.plain(L, L, 64, 6)
    with pytest.raises(sgm_amd.SgmError, match="outside"):
        ctx.plain(np.zeros((8, 8), np.uint8), np.zeros((8, 8), np.uint8), 64, 4)
