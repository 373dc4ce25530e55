"""The window-proposal network's HIP kernels against its float oracle:
features bit-exact, logits and gradients within the bf16 tolerance, the
proposal rule, and that training lowers the loss."""
import numpy as np
import pytest

import cnn_oracle as co
import sgm_amd

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SIZES = [(64, 96), (52, 100), (120, 200)]   # (H, W): ragged conv tiles and partial 16x16 tiles
# frames wide enough for workgroups away from both image edges (the features
# kernel's unclamped interior path) as well as edge ones
WIDE = SIZES + [(64, 512)]


def case(H, W, seed):
    L, R, _ = sgm_amd.scene_planes(W, H, min(127, W - 1), seed, planes=(3, 6))
    disp = np.random.default_rng(seed).integers(0, 128, (H, W)).astype(np.uint8)
    theta = co.init_params(seed + 5)
    for layer in range(co.LAYERS):   # non-zero biases so every bias path matters
        o = layer * (co.C * co.C * 9 + co.C) + co.C * co.C * 9
        theta[o:o + co.C] = np.random.default_rng(seed + layer).normal(0, 0.05, co.C)
    return L, R, disp, theta.astype(np.float32)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def grads_on_gpu(ctx, theta, L, R, disp):
    H, W = L.shape
    net = sgm_amd.Cnn(ctx, theta)
    dl, dr, dt = dev(L), dev(R), dev(disp)
    s = torch.cuda.current_stream().cuda_stream
    loss = net.train_step_dev(dl.data_ptr(), dr.data_ptr(), dt.data_ptr(), W, H, 0.0, s)
    f, z = net.activations(W, H)
    g = net.grads()
    net.close()
    return loss, f, z, g


@pytest.mark.parametrize("H,W", WIDE)
def test_features_bit_exact(ctx, H, W):
    L, R, disp, theta = case(H, W, H + W)
    _, f, _, _ = grads_on_gpu(ctx, theta, L, R, disp)
    want = co.features(L, R)
    np.testing.assert_array_equal(f.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("H,W", SIZES)
def test_logits_match_oracle(ctx, H, W):
    """Tolerance: products of bf16 values summed in fp32 in a different order,
    then re-rounded to bf16 between layers -- a few bf16 ulps at the logits."""
    L, R, disp, theta = case(H, W, H + W + 1)
    _, _, z, _ = grads_on_gpu(ctx, theta, L, R, disp)
    _, want = co.forward(theta, co.features(L, R))
    scale = np.abs(want).max()
    err = np.abs(z - want).max()
    assert err <= 2e-2 * scale, (err, scale)
    rel = np.linalg.norm(z - want) / np.linalg.norm(want)
    assert rel < 5e-3, rel


@pytest.mark.parametrize("H,W", SIZES)
def test_gradients_match_oracle(ctx, H, W):
    """The device backward rounds dlogits and the propagated data gradient to
    bf16; the oracle keeps them in fp32.  Per layer the weight and bias
    gradients agree to 3% in relative L2 norm (cosine > 0.999)."""
    L, R, disp, theta = case(H, W, H + W + 2)
    loss, _, _, g = grads_on_gpu(ctx, theta, L, R, disp)
    want_loss, want = co.loss_and_grads(theta, L, R, disp)
    assert abs(loss - want_loss) <= 1e-3 * abs(want_loss), (loss, want_loss)
    per = co.C * co.C * 9 + co.C
    for layer in range(co.LAYERS):
        for lo, hi in ((0, co.C * co.C * 9), (co.C * co.C * 9, per)):
            a = g[layer * per + lo: layer * per + hi].astype(np.float64)
            b = want[layer * per + lo: layer * per + hi].astype(np.float64)
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
            assert rel < 3e-2 and cos > 0.999, (layer, lo, rel, cos)


@pytest.mark.parametrize("H,W", WIDE)
def test_proposal_matches_oracle(ctx, H, W):
    L, R, disp, theta = case(H, W, H + W + 3)
    # a confident network: sharpen the last layer so both branches of the rule occur
    last = 5 * (co.C * co.C * 9 + co.C)
    theta[last:last + co.C * co.C * 9] *= 8
    net = sgm_amd.Cnn(ctx, theta)
    T = ((H + 15) // 16) * ((W + 15) // 16)
    dl, dr = dev(L), dev(R)
    db = torch.empty(T, dtype=torch.uint8, device="cuda")
    dc = torch.empty(T, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    blk_w, conf_w = co.propose(theta, L, R, 0.5)
    net.propose_dev(dl.data_ptr(), dr.data_ptr(), W, H, 0.5, db.data_ptr(), dc.data_ptr(), s)
    torch.cuda.synchronize()
    blk, conf = db.cpu().numpy().reshape(blk_w.shape), dc.cpu().numpy().reshape(conf_w.shape)
    net.close()
    np.testing.assert_allclose(conf, conf_w, atol=2e-2)
    # tiles whose decision is not on a knife edge must agree exactly
    p = co.tile_probs(co.forward(theta, co.features(L, R))[1])
    win = np.sort(np.stack([p[..., 4 * b:4 * b + 8].sum(-1) for b in range(7)], -1), -1)
    clear = (win[..., -1] - win[..., -2] > 0.05) & (np.abs(conf_w - 0.5) > 0.05)
    assert clear.mean() > 0.3
    np.testing.assert_array_equal(blk[clear], blk_w[clear])
    # the rule itself: a confident tile gets two adjacent blocks, the rest all eight
    conf_tiles = conf >= 0.5
    assert ((blk[conf_tiles] >> 4) == (blk[conf_tiles] & 15) + 1).all()
    assert (blk[~conf_tiles] == 7 << 4).all()


def test_training_lowers_loss(ctx):
    H, W = 96, 160
    net = sgm_amd.Cnn(ctx, co.init_params(11))
    s = torch.cuda.current_stream().cuda_stream
    frames = []
    for seed in range(4):
        L, R, _ = sgm_amd.scene_planes(W, H, 127, 500 + seed, planes=(2, 4))
        dl, dr = dev(L), dev(R)
        dt = torch.empty_like(dl)
        ctx.plain_dev(dl.data_ptr(), dr.data_ptr(), W, H, 128, 4, dt.data_ptr(), s)
        frames.append((dl, dr, dt))
    first = [net.train_step_dev(a.data_ptr(), b.data_ptr(), t.data_ptr(), W, H, 0.0, s) for a, b, t in frames]
    for it in range(60):
        a, b, t = frames[it % 4]
        net.train_step_dev(a.data_ptr(), b.data_ptr(), t.data_ptr(), W, H, 3e-3, s)
    last = [net.train_step_dev(a.data_ptr(), b.data_ptr(), t.data_ptr(), W, H, 0.0, s) for a, b, t in frames]
    net.close()
    assert np.mean(last) < 0.7 * np.mean(first), (first, last)


def test_rejects_bad_frames(ctx):
    net = sgm_amd.Cnn(ctx, co.init_params(0))
    d = torch.zeros((30, 30), dtype=torch.uint8, device="cuda")
    with pytest.raises(sgm_amd.SgmError, match="multiples of 4"):
        net.propose_dev(d.data_ptr(), d.data_ptr(), 30, 30, 0.9, d.data_ptr(), 0, 0)
    net.close()
