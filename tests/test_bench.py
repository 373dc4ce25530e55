"""bench.py's host logic on the CPU: the reference harness's statistics, the
byte models, and the one-GPU refusal."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_harness_stats_ranks():
    """p95 is the ceil(0.95 n) - 1 rank (common/harness.c:164-170): at n = 20
    the 19th value, not the maximum; rmsd is about the median."""
    st = bench.harness_stats([float(x) for x in range(20, 0, -1)])
    assert (st["median_ms"], st["p95_ms"], st["min_ms"], st["max_ms"]) == (11.0, 19.0, 1.0, 20.0)
    want = (sum((x - 11.0) ** 2 for x in range(1, 21)) / 19) ** 0.5
    assert abs(st["rmsd_ms"] - want) < 1e-4
    assert st["spread_hi"] == 1.727 and st["unstable"]  # p95/median above 1.25
    assert bench.harness_stats([1.0] * 9 + [2.0])["p95_ms"] == 2.0  # n = 10: rank 9
    assert bench.harness_stats([1.0, 1.0, 1.0, 1.3] * 5)["unstable"]
    assert not bench.harness_stats([1.0, 1.0, 1.0, 1.2] * 5)["unstable"]


def test_byte_models():
    # the reference's published figures at 1080p D=64 (README.md:97-98): 1.46 GB aggregation
    assert bench.survey_agg_bytes(1920, 1080, 64) == 1_459_814_400
    assert bench.frame_model_bytes("plain", 1920, 1080, 64) == 1_764_633_600
    assert bench.kernel_bytes("plain", 1920, 1080, 64, 4) == 663_552_000


def test_grid_cells_are_the_reference_sweep(goldens):
    cells = {(W, H, D) for W, H, D, _ in bench.GRID_CELLS}
    assert cells == {(W, H, D) for D in (32, 64, 128)
                     for W, H in ((1920, 1080), (1280, 720), (960, 540), (640, 480))}
    names = {r["name"]: r for r in goldens["reference_goldens"] + goldens["computed"]}
    for W, H, D, name in bench.GRID_CELLS:
        r = names[name]
        assert r["scene"] == {"kind": "reference", "W": W, "H": H, "gen_D": D, "seed": 1} and r["D"] == D
        assert r["goldencheck"]["ok"], name


@pytest.mark.parametrize("gpus", ["2", "8"])
def test_more_than_one_gpu_is_refused(gpus):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", gpus],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "one GPU" in r.stderr
