"""The drop-in on the GPU: the harness binaries (host/harness.c -> the SGM_IMPL
adapter -> sgmh_plain) and the multiscale binary, run on the reference's own
scenes and gated by its goldens, as common/harness.c:182-194,236 and
multiscale/sgm_ms_ref.c:208-223 specify: GOLDEN OK and exit 0 on a match,
GOLDEN MISMATCH and exit 2 on a golden with one byte changed.  The goldens are
written by the oracle and checked against the FNV-1a of the reference's
on-disk golden maps before they gate anything."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import sgm_amd

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "sgm-bench_amd", "bin")
CAL = os.path.join(ROOT, "sgm-bench_amd", "host", "roofline_mi355x.cal")
# scripts/record_ms.sh:25-26, the reference's parser of the multiscale result line
MS_LINE = re.compile(r"^(\S+)\s+in (\d+)x(\d+) out (\d+)x(\d+) D=(\d+) paths=(\d+)(?: th=(\d+))?\s+"
                     r"median ([\d.]+) ms\s+fps ([\d.]+)\s+MDE/s\(work\) (\d+)\s+hash (\w+)(.*)$")
# the batch binary's result line
BATCH_LINE = re.compile(r"^batch_hip\s+pairs (\d+) \((\d+) plain, (\d+) multiscale\)\s+median ([\d.]+) ms\s+"
                        r"frames/s ([\d.]+)\s+loop ([\d.]+) ms\s+speedup ([\d.]+)\s+MDE/s (\d+)(.*)$")
# common/harness.c:217-229, the JSON record's keys in order
RECORD_KEYS = ["impl", "board", "W", "H", "D", "paths", "census", "P1", "P2", "threads", "threads_requested",
               "cpu_mask", "warm", "timed", "median_ms", "p95_ms", "min_ms", "rmsd_ms", "spread_hi", "spread_lo",
               "unstable", "transfer_ms", "threads_used", "fps", "mpix_s", "census_ms", "cost_ms", "aggregate_ms",
               "argmin_ms", "hash", "golden_match", "roofline_ok", "cflags", "git", "ts"]


def run(*args):
    print("run:", " ".join(os.path.basename(a) if os.sep in a else a for a in args), flush=True)
    return subprocess.run(list(args), capture_output=True, text=True, timeout=180)


def write_pair(d, L, R, G=None):
    paths = [os.path.join(d, n) for n in ("left.pgm", "right.pgm", "golden.pgm")]
    sgm_amd.pgm_write(paths[0], L)
    sgm_amd.pgm_write(paths[1], R)
    if G is not None:
        sgm_amd.pgm_write(paths[2], G)
    return paths


@pytest.fixture(scope="module")
def config_a(tmp_path_factory):
    """The reference's Configuration A scene (gen_synthetic 1920 1080, seed 1)
    and data/golden/synthetic.pgm, regenerated."""
    d = str(tmp_path_factory.mktemp("a"))
    L, R, _ = sgm_amd.scene_reference(1920, 1080, 64, 1)
    G = oracle.sgm_plain(L, R, 64, 4)
    assert sgm_amd.fnv1a64(G) == "0bc0102058d1505f"
    return d, write_pair(d, L, R, G), G


def test_harness_golden_ok(config_a):
    d, (lp, rp, gp), _ = config_a
    js = os.path.join(d, "run.json")
    r = run(os.path.join(BIN, "sgm_hip"), lp, rp, "-g", gp, "-w", "3", "-n", "20", "--no-roofline", "-j", js)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "GOLDEN OK" in r.stdout and "hash 0bc0102058d1505f" in r.stdout, r.stdout
    rec = json.loads(open(js).read().splitlines()[-1])
    assert list(rec) == RECORD_KEYS
    assert (rec["impl"], rec["W"], rec["H"], rec["D"], rec["paths"]) == ("hip", 1920, 1080, 64, 4)
    assert (rec["golden_match"], rec["hash"], rec["timed"]) == (1, "0bc0102058d1505f", 20)
    assert rec["aggregate_ms"] >= 0 and rec["cost_ms"] == -1 and rec["transfer_ms"] > 0
    assert rec["min_ms"] <= rec["median_ms"] <= rec["p95_ms"]
    print(r.stdout.strip())


def test_harness_golden_mismatch_exits_2(config_a):
    d, (lp, rp, _), G = config_a
    bad = G.copy()
    bad[700, 1234] ^= 1
    gp = os.path.join(d, "golden_bad.pgm")
    sgm_amd.pgm_write(gp, bad)
    r = run(os.path.join(BIN, "sgm_hip"), lp, rp, "-g", gp, "-w", "0", "-n", "1", "--no-roofline")
    assert r.returncode == 2, r.stdout + r.stderr
    assert "GOLDEN MISMATCH: 1 pixels differ, first at (1234,700)" in r.stderr, r.stderr
    assert "GOLDEN FAIL" in r.stdout


def test_harness_cost_gate_armed(config_a):
    """With the MI355X calibration the gate arms itself and prints each phase's
    ns/op against its calibration.  Exit 3 would mean a phase over 1.15x of it
    (a right answer produced too slowly) and still GOLDEN OK."""
    _, (lp, rp, gp), _ = config_a
    r = run(os.path.join(BIN, "sgm_hip"), lp, rp, "-g", gp, "-w", "10", "-n", "50", "--board", "mi355x",
            "--roofline-cal", CAL)
    print(r.stdout.strip())
    assert r.returncode in (0, 3), r.stdout + r.stderr
    assert "GOLDEN OK" in r.stdout
    ratios = [float(x) for x in re.findall(r"([\d.]+)x", r.stdout.split("roofline:")[1])]
    assert len(ratios) == 3, r.stdout
    assert (r.returncode == 3) == (max(ratios) > 1.15), (ratios, r.stderr)


def test_harness_d128_shared_golden(tmp_path):
    """data/shared/golden_d128.pgm: gen_synthetic -DSGM_D=256 1920 1080 seed 3, D=128."""
    L, R, _ = sgm_amd.scene_reference(1920, 1080, 256, 3)
    G = oracle.sgm_plain(L, R, 128, 4)
    assert sgm_amd.fnv1a64(G) == "760646cc7238afea"
    lp, rp, gp = write_pair(str(tmp_path), L, R, G)
    r = run(os.path.join(BIN, "sgm_hip_d128"), lp, rp, "-g", gp, "-w", "1", "-n", "3", "--no-roofline")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "GOLDEN OK" in r.stdout and "D=128" in r.stdout


def test_harness_p8(tmp_path, goldens):
    g = next(x for x in goldens["computed"] if x["name"] == "p8_640x480_d64")
    L, R = sgm_amd.make_scene(g["scene"])
    G = oracle.sgm_plain(L, R, 64, 8)
    assert sgm_amd.fnv1a64(G) == g["fnv"]
    lp, rp, gp = write_pair(str(tmp_path), L, R, G)
    r = run(os.path.join(BIN, "sgm_hip_p8"), lp, rp, "-g", gp, "-w", "1", "-n", "3", "--no-roofline")
    assert r.returncode == 0, r.stdout + r.stderr
    assert "GOLDEN OK" in r.stdout and "paths=8" in r.stdout


@pytest.mark.parametrize("H,fnv", [(800, "bcb9cb0bd6f49799"), (1080, "0f0961d623009df5")])
def test_ms_binary_on_reference_goldens(tmp_path, H, fnv):
    """data/multiscale/golden_ms.pgm (B) and golden_ms_1080.pgm (C); the result
    line parses with the reference's own recorder pattern."""
    L, R, _ = sgm_amd.scene_reference(1920, H, 64, 1)
    G = oracle.sgm_ms(L, R, 64)
    assert sgm_amd.fnv1a64(G) == fnv
    lp, rp, gp = write_pair(str(tmp_path), L, R, G)
    r = run(os.path.join(BIN, "sgm_ms_hip"), lp, rp, "-g", gp, "-n", "5")
    assert r.returncode == 0, r.stdout + r.stderr
    line = [x for x in r.stdout.splitlines() if x.startswith("ms_hip")][0]
    m = MS_LINE.search(line)
    assert m, line
    impl, W, Hh, ow, oh, D, paths, th, ms, fps, mde, h, tail = m.groups()
    assert (int(W), int(Hh), int(ow), int(oh), int(D), int(paths)) == (1920, H, 960, H // 2, 64, 8)
    assert h == fnv and "GOLDEN OK" in tail
    bad = G.copy()
    bad[3, 5] ^= 2
    sgm_amd.pgm_write(gp, bad)
    r = run(os.path.join(BIN, "sgm_ms_hip"), lp, rp, "-g", gp, "-n", "1")
    assert r.returncode == 2, r.stdout + r.stderr


def test_batch_binary_goldens(tmp_path, goldens):
    """sgm_batch_hip builds the 24-pair batch from its seeds, autotunes it and
    hashes every pair: GOLDEN OK and exit 0 against the fixtures, and exit 2
    once one pair's golden is changed (a pinned schedule, no autotuning)."""
    gold = {g["name"]: g["fnv"] for g in goldens["batch"]}
    gp, js = str(tmp_path / "goldens.txt"), str(tmp_path / "rec.jsonl")

    def write(g):
        with open(gp, "w") as f:
            for k, v in g.items():
                f.write(f"{k} {v}\n")

    exe = os.path.join(BIN, "sgm_batch_hip")
    write(gold)
    r = run(exe, "-g", gp, "-n", "5", "-j", js, "-v")
    assert r.returncode == 0, r.stdout + r.stderr
    line = [x for x in r.stdout.splitlines() if x.startswith("batch_hip")][0]
    m = BATCH_LINE.search(line)
    assert m, line
    n, nplain, nms, med, fps, loop, speedup, mde, tail = m.groups()
    assert (int(n), int(nplain), int(nms)) == (24, 20, 4) and "GOLDEN OK" in tail
    assert abs(float(loop) / float(med) - float(speedup)) < 2e-3
    hashes = dict(x.split()[0::5] for x in r.stdout.splitlines() if " hash " in x and not x.startswith("batch_hip"))
    assert hashes == gold
    rec = json.loads(open(js).read().splitlines()[-1])
    assert rec["impl"] == "batch_hip" and rec["golden_match"] == 1 and rec["plan"]["pairs"] == 24
    assert rec["plan"]["tuning"], "the default plan is autotuned"
    first = next(iter(gold))
    write({**gold, first: "%016x" % (int(gold[first], 16) ^ 1)})
    r = run(exe, "-g", gp, "-n", "1", "-f", "0")
    assert r.returncode == 2 and "GOLDEN FAIL" in r.stdout, r.stdout + r.stderr
    assert first in r.stderr
