"""bench.py -- throughput of the MI355X SGM engine on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload A|P|B|C]
                    [--no-cpu-baseline] [--cpu-seconds S]

A step is one frame of the named workload, inputs already resident in HBM:
  A  plain SGM, 1920x1080, D=64, 4 paths, seeded planes scene (the headline)
  P  plain SGM,  640x480,  D=64, 4 paths
  B  multiscale 8-path two-scale SGM, 1920x800,  D=64
  C  multiscale 8-path two-scale SGM, 1920x1080, D=64
  batch  24 pairs of uneven size in one submission (8x640x480 D64, 8x1280x720
     D64, 4x1920x800 D128 multiscale, 4x1920x1080 D128), one graph launch per
     step; frames/s against a loop of single-pair launches of the same pairs
  learned  1920x1080 D=128 4 paths: window proposal by the trained network,
     then SGM over each tile's 32-disparity window (full D below confidence
     0.9); frames/s against plain D=128 on the same frames, and the share of
     pixels whose disparity differs from plain
The last frame of the timed loop is hash-checked against the compiled
reference oracle's golden for that scene; a mismatch exits 2, as the
reference harness does (common/harness.c:236).

Prints ONE JSON line.  `value` is MDE/s = W*H*D*frames/s (README.md:150-153 of
the reference; multiscale counts both scales, (W*H/4 + W*H)*D, the reference's
"work" convention, sgm_ms_ref.c:211-216).  `roofline` is measured live on the
dominant kernel (path aggregation) with HIP events on the stream it runs on;
`cpu_baseline` times the CPU oracle on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sgm-bench_amd"))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E peak


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="A", choices=["A", "P", "B", "C", "batch", "learned"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="approximate CPU time budget per CPU-baseline leg")
    return ap.parse_args()


def golden_for(seed, mode, D, paths, W, H):
    with open(os.path.join(ROOT, "tests", "golden", "goldens.json")) as f:
        g = json.load(f)
    for rec in g["computed"]:
        sp = rec["scene"]
        if (rec["mode"] == mode and rec["D"] == D and rec["paths"] == paths and sp["W"] == W
                and sp["H"] == H and sp["kind"] == "planes" and sp["seed"] == seed):
            return rec
    return None


def kernel_bytes(mode, W, H, D, paths):
    """Algorithmic HBM bytes of one launch of the path-aggregation kernel: each
    (pixel, direction) writes its D path costs once as uint8 and reads the two
    64-bit census codes of that pixel once (plain); multiscale reads the uint8
    cost volume once per direction and writes D bytes per direction only at the
    pixels the decimated argmin reads, at both scales."""
    N = W * H
    if mode == "plain":
        return paths * (N * D + 16 * N)
    Nh = N // 4
    stage1 = 8 * (Nh * D + Nh * D)      # read C1, write all of L1
    stage2 = 8 * (N * D + Nh * D)       # read C2, write L2 at even pixels
    return stage1 + stage2


def frame_model_bytes(mode, W, H, D):
    """The reference's published per-frame traffic model (README.md:97-98 and
    cuda/sgm_cuda_opt.cu:139-151 for the plain path; README.md:215-217 for B/C)."""
    N = W * H
    return 12 * N * D + 83 * N if mode == "plain" else int(45.75 * N * D)


def cpu_baseline(L, R, mode, D, paths, budget_s, mdes_per_frame):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    res = {}
    for threads in (0, 1):
        frames, t0 = 0, time.perf_counter()
        while True:
            if mode == "plain":
                oracle.sgm_plain(L, R, D, paths, threads=threads)
            else:
                oracle.sgm_ms(L, R, D, threads=threads)
            frames += 1
            el = time.perf_counter() - t0
            if el >= budget_s or (threads == 1 and el * (frames + 1) / frames > 2.5 * budget_s):
                break
        res[threads] = (frames, el)
    nthreads = oracle.omp_threads()
    f_mt, t_mt = res[0]
    f_1, t_1 = res[1]
    return {
        "value": mdes_per_frame * f_mt / t_mt,
        "unit": "MDE/s",
        "cores": nthreads,
        "kind": "port",
        "sample": f"{f_mt} full frames of the same scene through the oracle's OpenMP restatement "
                  f"({nthreads} threads observed in the team), {t_mt:.1f} s",
        "ms_per_frame": 1000.0 * t_mt / f_mt,
        "scalar": {"value": mdes_per_frame * f_1 / t_1, "cores": 1,
                   "ms_per_frame": 1000.0 * t_1 / f_1,
                   "sample": f"{f_1} frames through the scalar oracle (-O2, 1 thread), {t_1:.1f} s"},
    }


def bench_batch(a):
    import numpy as np
    import torch
    import sgm_amd

    specs = sgm_amd.batch_spec()
    with open(os.path.join(ROOT, "tests", "golden", "goldens.json")) as f:
        gold = {g["name"]: g["fnv"] for g in json.load(f)["batch"]}
    keep, pairs, scenes = [], [], []
    for sp in specs:
        L, R = sgm_amd.make_scene(sp["scene"])
        scenes.append((L, R))
        dl, dr = torch.from_numpy(L).cuda(), torch.from_numpy(R).cuda()
        oh, ow = (sp["H"], sp["W"]) if sp["mode"] == sgm_amd.PLAIN else (sp["H"] // 2, sp["W"] // 2)
        do = torch.empty((oh, ow), dtype=torch.uint8, device="cuda")
        keep.append((dl, dr, do))
        pairs.append(dict(d_left=dl.data_ptr(), d_right=dr.data_ptr(), d_disp=do.data_ptr(), W=sp["W"],
                          H=sp["H"], D=sp["D"], mode=sp["mode"], paths=sp["paths"]))
    ctx = sgm_amd.Context(0)
    b = ctx.batch(pairs, sgm_amd.BATCH_AUTOTUNE)
    sp_ = torch.cuda.current_stream().cuda_stream

    def timed(fn, steps, warm):
        for _ in range(warm):
            fn(sp_)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn(sp_)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def check():
        bad = [s["name"] for s, (_, _, do) in zip(specs, keep)
               if sgm_amd.fnv1a64(do.cpu().numpy()) != gold[s["name"]]]
        return bad

    t_loop = timed(b.run_loop, a.steps, a.warmup)
    bad_loop = check()
    for _, _, do in keep:
        do.fill_(0)
    t_graph = timed(b.run, a.steps, a.warmup)
    bad = check()
    n = len(specs)
    fps = n * a.steps / t_graph
    fps_loop = n * a.steps / t_loop
    mde = sum((s["W"] * s["H"] if s["mode"] == sgm_amd.PLAIN else s["W"] * s["H"] * 5 // 4) * s["D"]
              for s in specs) / 1e6
    rec = {
        "metric": "batched frames/s, 24 uneven pairs, bit-exact",
        "value": round(fps, 1),
        "unit": "frames/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1000 * t_graph / a.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u8/u16 integer",
        "data": "synthetic (24 seeded planes scenes, generated on the host)",
        "config": {"workload": "batch: 8x640x480 D64 + 8x1280x720 D64 + 4x1920x800 D128 multiscale "
                               "+ 4x1920x1080 D128, one hipGraph launch per step"},
        "loop_frames_per_s": round(fps_loop, 1),
        "speedup_vs_loop": round(fps / fps_loop, 3),
        "MDE_per_s": round(mde * a.steps / t_graph, 1),
        "golden_match": not bad and not bad_loop,
        "golden_failures": bad + ["loop:" + x for x in bad_loop],
        "plan": b.describe(),
    }
    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        # one pair of each kind through the OpenMP oracle, weighted by its count
        kinds, total = {}, 0.0
        for s, (L, R) in zip(specs, scenes):
            key = (s["W"], s["H"], s["D"], s["mode"])
            if key not in kinds:
                t0 = time.perf_counter()
                if s["mode"] == sgm_amd.PLAIN:
                    oracle.sgm_plain(L, R, s["D"], s["paths"])
                else:
                    oracle.sgm_ms(L, R, s["D"])
                kinds[key] = time.perf_counter() - t0
            total += kinds[key]
        rec["cpu_baseline"] = {"value": round(n / total, 3), "unit": "frames/s",
                               "cores": oracle.omp_threads(), "kind": "port",
                               "sample": "one pair of each of the 4 kinds through the oracle's OpenMP "
                                         "restatement, weighted by its count in the batch (%.1f s)"
                                         % sum(kinds.values())}
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)
    b.close()
    ctx.close()
    return 0 if rec["golden_match"] else 2


def bench_learned(a):
    import numpy as np
    import torch
    import sgm_amd
    from sgm_amd import learned

    # 0.92: the lowest threshold at which the held-out frames differ from plain D=128 on <= 1 % of pixels
    W, H, D, thresh = 1920, 1080, 128, 0.92
    nframes = 4
    ctx = sgm_amd.Context(0)
    theta = learned.load_params()
    net = sgm_amd.Cnn(ctx, theta)
    sp = torch.cuda.current_stream().cuda_stream
    T = ((H + 15) // 16) * ((W + 15) // 16)
    frames = []
    for i in range(nframes):
        L, R = learned.scene(learned.EVAL_SEED0 + i, W, H)
        frames.append((L, R, torch.from_numpy(L).cuda(), torch.from_numpy(R).cuda()))
    d_plain = torch.empty((H, W), dtype=torch.uint8, device="cuda")
    d_win = torch.empty_like(d_plain)
    d_blk = torch.empty(T, dtype=torch.uint8, device="cuda")

    def learned_step(k):
        # one call: the census runs on a side stream beside the proposal network
        _, _, dl, dr = frames[k % nframes]
        net.learned_dev(dl.data_ptr(), dr.data_ptr(), W, H, thresh, d_win.data_ptr(), d_blk.data_ptr(), sp)

    def plain_step(k):
        _, _, dl, dr = frames[k % nframes]
        ctx.plain_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, 4, d_plain.data_ptr(), sp)

    def timed(fn):
        for k in range(a.warmup):
            fn(k)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(a.steps):
            fn(k)
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    t_learned = timed(learned_step)
    t_plain = timed(plain_step)
    # quality and the per-frame algorithmic bytes of the windowed path kernel
    diffs, mads, falls, act_bytes = [], [], [], []
    for k in range(nframes):
        plain_step(k)
        learned_step(k)
        torch.cuda.synchronize()
        p, w, b = d_plain.cpu().numpy(), d_win.cpu().numpy(), d_blk.cpu().numpy()
        diffs.append(float((p != w).mean() * 100))
        mads.append(float(np.abs(p.astype(int) - w.astype(int)).mean()))
        falls.append(float(((b >> 4) - (b & 15) == 7).mean() * 100))
        width = ((b >> 4).astype(np.int64) - (b & 15) + 1).reshape((H + 15) // 16, (W + 15) // 16)
        per_px = np.repeat(np.repeat(width, 16, 0), 16, 1)[:H, :W] * 16
        act_bytes.append(4 * int(per_px.sum()) + 4 * 16 * W * H)
    # phase split, HIP events on the launch stream
    ctx.set_timing(True)
    acc = {"census_ms": 0.0, "aggregate_ms": 0.0, "argmin_ms": 0.0}
    nprof = 20
    for k in range(nprof):
        _, _, dl, dr = frames[k % nframes]
        net.propose_dev(dl.data_ptr(), dr.data_ptr(), W, H, thresh, d_blk.data_ptr(), 0, sp)
        ctx.window_dev(dl.data_ptr(), dr.data_ptr(), W, H, D, d_blk.data_ptr(), d_win.data_ptr(), sp)
        t = ctx.times()
        for key in acc:
            acc[key] += max(t[key], 0.0)
    ctx.set_timing(False)
    phase = {k: v / nprof for k, v in acc.items()}
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    _, _, dl, dr = frames[0]
    e0.record()
    for _ in range(nprof):
        net.propose_dev(dl.data_ptr(), dr.data_ptr(), W, H, thresh, d_blk.data_ptr(), 0, sp)
    e1.record()
    torch.cuda.synchronize()
    phase["proposal_ms"] = e0.elapsed_time(e1) / nprof
    kb = int(np.mean(act_bytes))
    agg_s = phase["aggregate_ms"] / 1000.0
    fps = a.steps / t_learned
    fps_plain = a.steps / t_plain
    rec = {
        "metric": "learned-window frames/s, 1080p D=128 4-path, vs plain D=128",
        "value": round(fps, 1),
        "unit": "frames/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": round(1000 * t_learned / a.steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16 network, u8/u16 integer SGM",
        "data": "synthetic (held-out slanted-planes scenes, seeds %d..%d)" % (learned.EVAL_SEED0,
                                                                           learned.EVAL_SEED0 + nframes - 1),
        "config": {"workload": "learned: proposal network + windowed SGM, 1920x1080 D=128 4-path, "
                               "confidence threshold %.2f" % thresh, "W": W, "H": H, "D": D, "thresh": thresh},
        "plain_d128_frames_per_s": round(fps_plain, 1),
        "speedup_vs_plain": round(fps / fps_plain, 3),
        "pct_pixels_differ": round(float(np.mean(diffs)), 3),
        "mean_abs_diff": round(float(np.mean(mads)), 4),
        "pct_tiles_full_D": round(float(np.mean(falls)), 2),
        "phase_ms": {k: round(v, 4) for k, v in phase.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_path_window (all directions, one launch)",
            "achieved": round(kb / agg_s / 1e9, 1) if agg_s > 0 else None,
            "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(kb / agg_s / 1e9 / HBM_PEAK_GBS, 4) if agg_s > 0 else None,
            "traffic": None,
            "algorithmic_bytes": kb,
        },
    }
    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        import cnn_oracle
        L, R = frames[0][0], frames[0][1]
        t0 = time.perf_counter()
        blk, _ = cnn_oracle.propose(theta, L, R, thresh)
        oracle.sgm_window(L, R, blk, D, 4)
        el = time.perf_counter() - t0
        rec["cpu_baseline"] = {"value": round(1.0 / el, 4), "unit": "frames/s", "cores": 1, "kind": "port",
                               "sample": "one 1080p frame: the network's numpy oracle (proposal) and the "
                                         "scalar windowed-SGM oracle, %.1f s" % el}
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)
    net.close()
    ctx.close()
    return 0


def main():
    a = parse()
    if a.workload == "learned":
        if a.gpus != 1:
            print("bench.py: one GPU only", file=sys.stderr)
            return 1
        return bench_learned(a)
    if a.workload == "batch":
        if a.gpus != 1:
            print("bench.py: one GPU only", file=sys.stderr)
            return 1
        return bench_batch(a)
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not supported.", file=sys.stderr)
        return 1
    import numpy as np
    import torch
    import sgm_amd

    wl = sgm_amd.WORKLOADS[a.workload]
    W, H, D = wl["W"], wl["H"], wl["D"]
    mode = "plain" if a.workload in ("A", "P") else "ms"
    paths = 4 if mode == "plain" else 8
    L, R, _ = sgm_amd.workload_scene(a.workload)
    dev = torch.device("cuda:0")
    d_l = torch.from_numpy(L).to(dev)
    d_r = torch.from_numpy(R).to(dev)
    oh, ow = (H, W) if mode == "plain" else (H // 2, W // 2)
    d_o = torch.empty((oh, ow), dtype=torch.uint8, device=dev)
    ctx = sgm_amd.Context(0)
    stream = torch.cuda.current_stream(dev)
    sp = stream.cuda_stream

    def step():
        if mode == "plain":
            ctx.plain_dev(d_l.data_ptr(), d_r.data_ptr(), W, H, D, paths, d_o.data_ptr(), sp)
        else:
            ctx.ms_dev(d_l.data_ptr(), d_r.data_ptr(), W, H, D, d_o.data_ptr(), 0, sp)

    for _ in range(a.warmup):
        step()
    torch.cuda.synchronize()
    # ---- headline: K frames back to back ----
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    ms_frame = 1000.0 * el / a.steps
    out = d_o.cpu().numpy()
    h = sgm_amd.fnv1a64(out)
    gold = golden_for(wl["seed"], mode, D, paths, W, H)
    golden_ok = None if gold is None else (h == gold["fnv"])

    mde = (W * H * D if mode == "plain" else (W * H // 4 + W * H) * D) / 1e6
    value = mde * a.steps / el

    # ---- per-kernel durations, HIP events on the launch stream ----
    ctx.set_timing(True)
    nprof = max(10, min(a.steps, 50))
    acc = {"census_ms": 0.0, "cost_ms": 0.0, "aggregate_ms": 0.0, "argmin_ms": 0.0, "total_ms": 0.0}
    for _ in range(nprof):
        step()
        t = ctx.times()
        for k in acc:
            acc[k] += max(t[k], 0.0)
    ctx.set_timing(False)
    phase = {k: v / nprof for k, v in acc.items()}
    kb = kernel_bytes(mode, W, H, D, paths)
    agg_s = phase["aggregate_ms"] / 1000.0
    achieved = kb / agg_s / 1e9 if agg_s > 0 else None
    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_%s.json" % a.workload)
    if os.path.exists(tfile):
        with open(tfile) as f:
            traffic = json.load(f).get("path_kernel_bytes")

    rec = {
        "metric": "MDE/s @1080p D=64 4-path bit-exact" if a.workload == "A" else
                  f"MDE/s workload {a.workload} bit-exact",
        "value": round(value, 1),
        "unit": "MDE/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_frame, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/u16 integer",
        "data": "synthetic (seeded planes scene, generated on the host)",
        "config": {"workload": f"{a.workload}: {'plain' if mode == 'plain' else 'multiscale 2-scale'} "
                               f"SGM {W}x{H} D={D} {paths}-path",
                   "W": W, "H": H, "D": D, "paths": paths, "scene_seed": wl["seed"]},
        "fps": round(1000.0 / ms_frame, 2),
        "golden_match": golden_ok,
        "hash": h,
        "phase_ms": {k: round(v, 4) for k, v in phase.items()},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_path (all directions, one launch)",
            "achieved": None if achieved is None else round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": None if achieved is None else round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes": kb,
            "frame_model_GBs": round(frame_model_bytes(mode, W, H, D) / (ms_frame / 1000.0) / 1e9, 1),
        },
    }
    if not a.no_cpu_baseline:
        rec["cpu_baseline"] = cpu_baseline(L, R, mode, D, paths, a.cpu_seconds, mde)
    else:
        rec["cpu_baseline"] = None
    print(json.dumps(rec), flush=True)
    ctx.close()
    return 2 if golden_ok is False else 0


if __name__ == "__main__":
    sys.exit(main())
