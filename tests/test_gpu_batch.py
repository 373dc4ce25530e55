"""Batched mode: every pair of a mixed submission is bit-exact to its oracle,
through the captured graph, on replays, and through the single-pair loop."""
import numpy as np
import pytest

import oracle
import sgm_amd

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def upload(pairs_np):
    """pairs_np: list of (spec, L, R) -> (pair dicts for the ABI, keep-alive tensors)."""
    keep, pairs = [], []
    for spec, L, R in pairs_np:
        dl, dr = torch.from_numpy(L).cuda(), torch.from_numpy(R).cuda()
        oh, ow = (spec["H"], spec["W"]) if spec["mode"] == sgm_amd.PLAIN else (spec["H"] // 2, spec["W"] // 2)
        do = torch.full((oh, ow), 0xEE, dtype=torch.uint8, device="cuda")
        keep += [dl, dr, do]
        pairs.append(dict(d_left=dl.data_ptr(), d_right=dr.data_ptr(), d_disp=do.data_ptr(),
                          W=spec["W"], H=spec["H"], D=spec["D"], mode=spec["mode"], paths=spec["paths"]))
    return pairs, keep


def want(spec, L, R):
    if spec["mode"] == sgm_amd.PLAIN:
        return oracle.sgm_plain(L, R, spec["D"], spec["paths"])
    return oracle.sgm_ms(L, R, spec["D"])


SMALL = [
    dict(W=160, H=120, D=64, mode=sgm_amd.PLAIN, paths=4, seed=1),
    dict(W=97, H=61, D=64, mode=sgm_amd.PLAIN, paths=8, seed=2),
    dict(W=200, H=90, D=128, mode=sgm_amd.PLAIN, paths=4, seed=3),
    dict(W=64, H=40, D=32, mode=sgm_amd.PLAIN, paths=8, seed=4),
    dict(W=128, H=96, D=64, mode=sgm_amd.MULTISCALE, paths=8, seed=5),
    dict(W=160, H=64, D=128, mode=sgm_amd.MULTISCALE, paths=8, seed=6),
    dict(W=62, H=34, D=32, mode=sgm_amd.MULTISCALE, paths=8, seed=7),
]


@pytest.mark.parametrize("flags", [sgm_amd.BATCH_AUTOTUNE, sgm_amd.BATCH_NO_GRAPH, 0,
                                   # the schedule choices the autotuner may make, pinned
                                   sgm_amd.BATCH_CENSUS_32X32, sgm_amd.BATCH_XCD_SLICES,
                                   sgm_amd.BATCH_CENSUS_32X32 | sgm_amd.BATCH_XCD_SLICES | sgm_amd.BATCH_NO_GRAPH,
                                   sgm_amd.BATCH_FUSED, sgm_amd.BATCH_FUSED | sgm_amd.BATCH_NO_GRAPH,
                                   sgm_amd.BATCH_FUSED | sgm_amd.BATCH_CENSUS_32X32 | sgm_amd.BATCH_XCD_SLICES,
                                   sgm_amd.BATCH_PIPELINED, sgm_amd.BATCH_PIPELINED | sgm_amd.BATCH_NO_GRAPH,
                                   sgm_amd.BATCH_PIPELINED | sgm_amd.BATCH_CENSUS_32X32 | sgm_amd.BATCH_XCD_SLICES])
def test_mixed_batch_matches_oracle(ctx, flags):
    data = []
    for sp in SMALL:
        L, R, _ = sgm_amd.scene_planes(sp["W"], sp["H"], min(sp["D"], sp["W"]) - 1, sp["seed"], planes=(3, 6))
        data.append((sp, L, R))
    pairs, keep = upload(data)
    b = ctx.batch(pairs, flags)
    s = torch.cuda.current_stream().cuda_stream
    wants = [want(sp, L, R) for sp, L, R in data]
    for rep in range(3):  # graph replays, with eager work in between
        for t in keep[2::3]:
            t.fill_(0xEE)
        b.run(s)
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()
        for (sp, _, _), w, d_out in zip(data, wants, keep[2::3]):
            got = d_out.cpu().numpy()
            assert np.array_equal(got, w), f"replay {rep}: {sp}: {(got != w).sum()} pixels differ"
    for t in keep[2::3]:
        t.fill_(0xEE)
    b.run_loop(s)
    torch.cuda.synchronize()
    for (sp, _, _), w, d_out in zip(data, wants, keep[2::3]):
        assert np.array_equal(d_out.cpu().numpy(), w), f"loop: {sp}"
    info = b.describe()
    assert info["pairs"] == len(SMALL)
    sched = info["schedule"]
    if not flags & sgm_amd.BATCH_AUTOTUNE:
        fused = bool(flags & sgm_amd.BATCH_FUSED)
        assert sched["census_tile"] == ("32x32" if flags & sgm_amd.BATCH_CENSUS_32X32 else "64x16")
        # the fused launch has one global queue whatever the flags ask
        assert sched["scanline_grouping"] == ("per-XCD" if flags & sgm_amd.BATCH_XCD_SLICES and not fused
                                              else "global")
        assert sched["fused"] == fused
        assert sched["pipelined"] == bool(flags & sgm_amd.BATCH_PIPELINED)
    else:
        # the autotuner timed both census tile shapes and both groupings, all
        # with separate path and argmin launches
        tried = {(t["census_tile"], t["scanline_grouping"]) for t in info["tuning"]}
        assert tried == {(c, g) for c in ("64x16", "32x32") for g in ("global", "per-XCD")}
        assert not any(t["fused"] for t in info["tuning"]) and not sched["fused"]
        assert {t["lanes"] for t in info["tuning"]} >= {"pipelined-3", "pipelined-4", "pipelined-6"}
    b.close()


@pytest.mark.parametrize("seed,flags", [(11, 0), (12, sgm_amd.BATCH_FUSED), (13, sgm_amd.BATCH_PIPELINED),
                                        (14, sgm_amd.BATCH_AUTOTUNE)])
def test_random_batches(ctx, seed, flags):
    """Random compositions: 5-9 pairs of random (ragged) size, D, configuration
    and path count, through one schedule each, every pair against its oracle
    on two consecutive runs."""
    rng = np.random.default_rng(seed)
    data = []
    for i in range(int(rng.integers(5, 10))):
        mode = sgm_amd.PLAIN if rng.random() < 0.6 else sgm_amd.MULTISCALE
        D = int(rng.choice([32, 64, 128]))
        if mode == sgm_amd.PLAIN:
            W, H, paths = int(rng.integers(17, 180)), int(rng.integers(16, 120)), int(rng.choice([4, 8]))
        else:
            W, H, paths = 2 * int(rng.integers(16, 90)), 2 * int(rng.integers(16, 60)), 8
        sp = dict(W=W, H=H, D=D, mode=mode, paths=paths)
        L, R, _ = sgm_amd.scene_planes(W, H, min(D, W) - 1, seed * 100 + i, planes=(2, 6))
        data.append((sp, L, R))
    pairs, keep = upload(data)
    b = ctx.batch(pairs, flags)
    s = torch.cuda.current_stream().cuda_stream
    wants = [want(sp, L, R) for sp, L, R in data]
    for rep in range(2):
        for t in keep[2::3]:
            t.fill_(0xEE)
        b.run(s)
        torch.cuda.synchronize()
        for (sp, _, _), w, d_out in zip(data, wants, keep[2::3]):
            got = d_out.cpu().numpy()
            assert np.array_equal(got, w), f"run {rep}: {sp}: {(got != w).sum()} pixels differ"
    b.close()


@pytest.mark.parametrize("flags", [sgm_amd.BATCH_AUTOTUNE,
                                   sgm_amd.BATCH_CENSUS_32X32 | sgm_amd.BATCH_XCD_SLICES,
                                   sgm_amd.BATCH_FUSED, sgm_amd.BATCH_PIPELINED])
def test_full_batch_goldens(ctx, goldens, flags):
    data, specs = [], sgm_amd.batch_spec()
    for sp in specs:
        L, R = sgm_amd.make_scene(sp["scene"])
        data.append((sp, L, R))
    pairs, keep = upload(data)
    b = ctx.batch(pairs, flags)
    b.run(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    gold = {g["name"]: g["fnv"] for g in goldens["batch"]}
    for sp, d_out in zip(specs, keep[2::3]):
        assert sgm_amd.fnv1a64(d_out.cpu().numpy()) == gold[sp["name"]], sp["name"]
    print(b.describe())
    b.close()


def test_batch_rejects_unsupported(ctx):
    ok = torch.zeros((64, 64), dtype=torch.uint8, device="cuda")
    pair = dict(d_left=ok.data_ptr(), d_right=ok.data_ptr(), d_disp=ok.data_ptr(), W=64, H=64, D=64,
                mode=sgm_amd.PLAIN, paths=4)
    with pytest.raises(sgm_amd.SgmError, match="exclude each other"):
        ctx.batch([pair], sgm_amd.BATCH_FUSED | sgm_amd.BATCH_PIPELINED)
    L = torch.zeros((64, 64), dtype=torch.uint8, device="cuda")
    bad = dict(d_left=L.data_ptr(), d_right=L.data_ptr(), d_disp=L.data_ptr(), W=64, H=64, D=256,
               mode=sgm_amd.PLAIN, paths=4)
    with pytest.raises(sgm_amd.SgmError, match="batched mode"):
        ctx.batch([bad])
