"""The learned window path end to end: the windowed SGM runs exactly on the
network's proposal, and training is what makes the windows good -- the
trained network's windows differ from plain D=128 on far fewer pixels than
an untrained network's at equal full-D fallback rates, and (at 30 %
fallback) on fewer than the same proposal rule without a network (the
per-pixel argmin of the matching features), which is measured beside them
and printed."""
import numpy as np
import pytest

import oracle
import sgm_amd
from sgm_amd import learned

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

W, H, D = 1920, 1080, 128


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def proposal(ctx, net, L, R):
    """Per tile: the network's best window start and its confidence."""
    h, w = L.shape
    T = ((h + 15) // 16) * ((w + 15) // 16)
    db = torch.empty(T, dtype=torch.uint8, device="cuda")
    dc = torch.empty(T, dtype=torch.float32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    dl, dr = dev(L), dev(R)   # held until the launch has run: the allocator may reuse a dropped block
    net.propose_dev(dl.data_ptr(), dr.data_ptr(), w, h, 0.0, db.data_ptr(), dc.data_ptr(), s)
    torch.cuda.synchronize()
    shape = ((h + 15) // 16, (w + 15) // 16)
    return (db.cpu().numpy() & 15).reshape(shape), dc.cpu().numpy().reshape(shape)


def test_learned_path_is_windowed_sgm_on_the_proposal(ctx):
    L, R = learned.scene(learned.EVAL_SEED0 + 7, 480, 272)
    net = sgm_amd.Cnn(ctx, learned.load_params())
    b0, conf = proposal(ctx, net, L, R)
    net.close()
    blk = np.where(conf >= 0.9, b0 | ((b0 + 1) << 4), 7 << 4).astype(np.uint8)
    assert 0 < (conf >= 0.9).mean() < 1
    np.testing.assert_array_equal(ctx.window(L, R, blk, D), oracle.sgm_window(L, R, blk, D, 4))


@pytest.mark.parametrize("w,h", [(640, 480), (644, 484)])   # the second: partial 16x16 tiles
def test_learned_call_equals_proposal_then_window(ctx, w, h):
    """The one-call learned frame (census on a side stream beside the network)
    gives the block table and disparities of the two separate calls."""
    L, R = learned.scene(learned.EVAL_SEED0 + 3, w, h)
    h, w = L.shape
    net = sgm_amd.Cnn(ctx, learned.load_params())
    T = ((h + 15) // 16) * ((w + 15) // 16)
    dl, dr = dev(L), dev(R)
    b1 = torch.empty(T, dtype=torch.uint8, device="cuda")
    b2 = torch.empty_like(b1)
    d1 = torch.empty((h, w), dtype=torch.uint8, device="cuda")
    d2 = torch.empty_like(d1)
    s = torch.cuda.current_stream().cuda_stream
    for _ in range(3):   # repeated calls reuse the side stream and events
        net.propose_dev(dl.data_ptr(), dr.data_ptr(), w, h, 0.9, b1.data_ptr(), 0, s)
        ctx.window_dev(dl.data_ptr(), dr.data_ptr(), w, h, D, b1.data_ptr(), d1.data_ptr(), s)
        net.learned_dev(dl.data_ptr(), dr.data_ptr(), w, h, 0.9, d2.data_ptr(), b2.data_ptr(), s)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(b2.cpu().numpy(), b1.cpu().numpy())
        np.testing.assert_array_equal(d2.cpu().numpy(), d1.cpu().numpy())
    net.close()


@pytest.fixture(scope="module")
def frames():
    return [learned.scene(learned.EVAL_SEED0 + i, W, H) for i in range(3)]


def differ_at(ctx, frames, pick, fallback):
    """Mean percent of pixels that differ from plain D=128 when each
    proposal keeps exactly `fallback` of its tiles at full D."""
    out = []
    for L, R in frames:
        plain = ctx.plain(L, R, D, 4)
        b0, conf = pick(L, R)
        blk = learned.blk_at_fallback(b0, conf, fallback)
        out.append((ctx.window(L, R, blk, D) != plain).mean() * 100)
    return float(np.mean(out))


@pytest.mark.parametrize("fallback", [0.2, 0.3])
def test_trained_network_gains(ctx, frames, fallback):
    trained = sgm_amd.Cnn(ctx, learned.load_params())
    untrained = sgm_amd.Cnn(ctx, learned.init_params())
    t = differ_at(ctx, frames, lambda L, R: proposal(ctx, trained, L, R), fallback)
    u = differ_at(ctx, frames, lambda L, R: proposal(ctx, untrained, L, R), fallback)
    hres = differ_at(ctx, frames, lambda L, R: learned.heuristic_blk(L, R, with_conf=True)[1:], fallback)
    trained.close()
    untrained.close()
    print(f"fallback {fallback:.0%}: trained {t:.3f}%, untrained {u:.3f}%, without network {hres:.3f}%")
    assert t < 0.5 * u, (t, u)
    assert t < 3.0, t
    if fallback >= 0.3:   # where the margin is clear, the network also beats the rule without it
        assert t < hres, (t, hres)
