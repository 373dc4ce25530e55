"""The per-phase cost gate (sgm-bench_amd/host/sgm_gate.h), which restates
the reference's common/roofline.h:62-179, driven from a small C program:
a calibration written from one set of stage times arms the gate for the
same implementation, board, resolution and D only; a phase within the limit
passes, one above it fails (the harness turns that into exit code 3); a
fused phase's ops fold into aggregate; an unreported aggregate fails closed;
no calibration leaves the gate unarmed and says so."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <stdlib.h>
#include "sgm_gate.h"
/* gate MODE CAL IMPL BOARD W H census cost aggregate argmin [max] */
int main(int argc, char **argv)
{
    if (argc < 11) return 64;
    sgm_stage_times t = {atof(argv[7]), atof(argv[8]), atof(argv[9]), atof(argv[10]), -1, 0};
    const int W = atoi(argv[5]), H = atoi(argv[6]);
    const char *cal = strcmp(argv[2], "-") ? argv[2] : NULL;
    if (!strcmp(argv[1], "cal"))
        return sgm_gate_calibrate(cal, argv[3], argv[4], W, H, &t);
    int armed = -1;
    const double mx = argc > 11 ? atof(argv[11]) : SGM_GATE_DEFAULT_MAX;
    const int rc = sgm_gate_check(cal, argv[3], argv[4], W, H, &t, mx, &armed);
    printf("\nrc=%d armed=%d\n", rc, armed);
    return 0;
}
"""

W, H = 1920, 1080
D, PATHS, BITS = 64, 4, 62   # include/sgm_params.h defaults
PX = W * H
OPS = {"census": 2 * PX * BITS, "cost": PX * D, "aggregate": PATHS * PX * D * 4, "argmin": PX * D}
TIMES = {"census": 0.028, "cost": -1.0, "aggregate": 0.275, "argmin": 0.080}


@pytest.fixture(scope="module")
def gate(tmp_path_factory):
    cc = shutil.which("gcc")
    if cc is None:
        pytest.skip("gcc not found")
    d = tmp_path_factory.mktemp("gate")
    src = d / "gate.c"
    src.write_text(DRIVER)
    exe = d / "gate"
    subprocess.run([cc, "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "sgm-bench_amd", "host"), str(src), "-o", str(exe)],
                   check=True)
    return str(exe)


def run(gate, mode, cal, times, impl="hip", board="mi355x", w=W, h=H, mx=None):
    args = [gate, mode, str(cal), impl, board, str(w), str(h)]
    args += [repr(times[k]) for k in ("census", "cost", "aggregate", "argmin")]
    if mx is not None:
        args.append(repr(mx))
    p = subprocess.run(args, capture_output=True, text=True, check=True)
    if mode == "cal":
        return p
    last = p.stdout.strip().splitlines()[-1]
    rc, armed = (int(f.split("=")[1]) for f in last.split())
    return rc, armed, p.stderr


def calibration(gate, tmp_path, times=TIMES, **kw):
    cal = tmp_path / "roofline.cal"
    run(gate, "cal", cal, times, **kw)
    return cal


def test_calibration_records_ns_per_op(gate, tmp_path):
    """Each reported phase is written as ms * 1e6 / ops; the fused cost
    phase (-1) is not written and its ops are counted under aggregate."""
    cal = calibration(gate, tmp_path)
    rows = {}
    for line in cal.read_text().splitlines():
        impl, board, res, dd, phase, v = line.split()
        assert (impl, board, res, dd) == ("hip", "mi355x", f"{W}x{H}", f"D{D}")
        rows[phase] = float(v)
    assert set(rows) == {"census", "aggregate", "argmin"}
    want = {"census": TIMES["census"] * 1e6 / OPS["census"],
            "aggregate": TIMES["aggregate"] * 1e6 / (OPS["aggregate"] + OPS["cost"]),
            "argmin": TIMES["argmin"] * 1e6 / OPS["argmin"]}
    for k, v in want.items():
        assert rows[k] == pytest.approx(v, rel=1e-4, abs=2e-6), k


def test_same_times_pass_and_arm(gate, tmp_path):
    cal = calibration(gate, tmp_path)
    rc, armed, _ = run(gate, "check", cal, TIMES)
    assert (rc, armed) == (0, 1)


@pytest.mark.parametrize("phase", ["census", "aggregate", "argmin"])
def test_one_slow_phase_fails(gate, tmp_path, phase):
    """10 % over the calibration passes the default 1.15 limit; 30 % over fails
    and names the phase."""
    cal = calibration(gate, tmp_path)
    ok = dict(TIMES, **{phase: TIMES[phase] * 1.10})
    assert run(gate, "check", cal, ok)[:2] == (0, 1)
    slow = dict(TIMES, **{phase: TIMES[phase] * 1.30})
    rc, armed, err = run(gate, "check", cal, slow)
    assert (rc, armed) == (1, 1)
    assert f"phase '{phase}'" in err and "ROOFLINE FAIL" in err


def test_limit_is_the_callers(gate, tmp_path):
    cal = calibration(gate, tmp_path)
    slow = dict(TIMES, aggregate=TIMES["aggregate"] * 1.30)
    assert run(gate, "check", cal, slow, mx=1.35)[:2] == (0, 1)
    assert run(gate, "check", cal, slow, mx=1.25)[:2] == (1, 1)
    assert run(gate, "check", cal, TIMES, mx=0.9)[:2] == (1, 1)   # every phase at 1.0x


def test_unreported_aggregate_fails_closed(gate, tmp_path):
    """Without aggregate the ops folded into it would go unchecked: fail,
    calibrated or not."""
    times = dict(TIMES, aggregate=-1.0)
    rc, armed, err = run(gate, "check", tmp_path / "absent.cal", times)
    assert (rc, armed) == (1, 0) and "no aggregate time was reported" in err
    cal = calibration(gate, tmp_path)
    assert run(gate, "check", cal, times)[:2] == (1, 0)


def test_no_calibration_is_not_armed(gate, tmp_path):
    rc, armed, err = run(gate, "check", tmp_path / "absent.cal", TIMES)
    assert (rc, armed) == (0, 0) and "NOT ARMED" in err
    rc, armed, err = run(gate, "check", "-", TIMES)
    assert (rc, armed) == (0, 0) and "NOT ARMED" in err


@pytest.mark.parametrize("key", [dict(impl="hip_p8"), dict(board="other"), dict(w=1280, h=720)])
def test_calibration_is_keyed(gate, tmp_path, key):
    """A calibration arms only its own implementation, board and resolution:
    a row for another key leaves the gate unarmed even for a slow run."""
    cal = calibration(gate, tmp_path)
    slow = {k: (v * 3 if v > 0 else v) for k, v in TIMES.items()}
    rc, armed, err = run(gate, "check", cal, slow, **key)
    assert (rc, armed) == (0, 0) and "no calibration for" in err


def test_committed_mi355x_calibration_parses(gate):
    """The committed MI355X calibration arms the default build at 1080p and
    passes the times it was recorded from (to the six printed decimals)."""
    cal = os.path.join(ROOT, "sgm-bench_amd", "host", "roofline_mi355x.cal")
    rows = [l.split() for l in open(cal) if l.strip()]
    mine = {r[4]: float(r[5]) for r in rows if r[:4] == ["hip", "mi355x", f"{W}x{H}", f"D{D}"]}
    assert set(mine) == {"census", "aggregate", "argmin"}
    times = {"census": mine["census"] * OPS["census"] / 1e6, "cost": -1.0,
             "aggregate": mine["aggregate"] * (OPS["aggregate"] + OPS["cost"]) / 1e6,
             "argmin": mine["argmin"] * OPS["argmin"] / 1e6}
    assert run(gate, "check", cal, times)[:2] == (0, 1)
    assert run(gate, "check", cal, dict(times, aggregate=times["aggregate"] * 1.3))[:2] == (1, 1)
