"""The C-ABI libraries load and export every symbol their headers declare
(no GPU work is done here)."""
import ctypes
import os
import re

import numpy as np
import pytest

import sgm_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    src = open(os.path.join(ROOT, "include", header)).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sgm[h]?_[a-z0-9_]+)\s*\(", src)))


@pytest.mark.parametrize("header,lib", [("sgm_hip.h", "libsgm_hip.so"), ("sgm_scene.h", "libsgm_host.so")])
def test_exports(header, lib):
    names = declared(header)
    assert names, header
    so = ctypes.CDLL(os.path.join(sgm_amd.LIBDIR, lib))
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_null_context_is_an_error_not_a_crash():
    h = sgm_amd.hip_lib()
    assert h.sgmh_plain(None, None, None, 64, 64, 64, 4, None, None) == -1
    assert b"NULL" in h.sgmh_last_error()


def test_fnv_known_answers():
    assert sgm_amd.fnv1a64(np.zeros(0, np.uint8)) == "cbf29ce484222325"
    assert sgm_amd.fnv1a64(np.frombuffer(b"a", np.uint8)) == "af63dc4c8601ec8c"
    assert sgm_amd.fnv1a64(np.frombuffer(b"foobar", np.uint8)) == "85944171f73967e8"


def test_pgm_roundtrip(tmp_path):
    img = np.random.default_rng(0).integers(0, 256, (7, 13), dtype=np.uint8)
    p = str(tmp_path / "x.pgm")
    sgm_amd.pgm_write(p, img)
    np.testing.assert_array_equal(sgm_amd.pgm_read(p), img)
    (tmp_path / "bad.pgm").write_bytes(b"P6\n1 1\n255\nx")
    with pytest.raises(sgm_amd.SgmError):
        sgm_amd.pgm_read(str(tmp_path / "bad.pgm"))


def test_planes_scene_deterministic():
    a = sgm_amd.scene_planes(128, 96, 63, 5, planes=(8, 12), occl_permille=50)
    b = sgm_amd.scene_planes(128, 96, 63, 5, planes=(8, 12), occl_permille=50)
    c = sgm_amd.scene_planes(128, 96, 63, 6, planes=(8, 12), occl_permille=50)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert not np.array_equal(a[1], c[1])
    assert a[2].max() == 63  # the last plane is pinned to dmax


BIN = os.path.join(ROOT, "sgm-bench_amd", "bin")


def _run(*args):
    import subprocess
    return subprocess.run(list(args), capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("prog", ["sgm_hip", "sgm_hip_d128", "sgm_hip_p8", "sgm_ms_hip", "gen_scene"])
def test_host_programs_usage(prog):
    """Every host program prints its usage and exits 1 without arguments
    (common/harness.c's exit code for usage and IO errors)."""
    r = _run(os.path.join(BIN, prog))
    assert r.returncode == 1 and "usage" in r.stderr


def test_batch_program_usage():
    """sgm_batch_hip runs its built-in batch without arguments, so an unknown
    argument is what prints its usage (exit 1, before any GPU work)."""
    r = _run(os.path.join(BIN, "sgm_batch_hip"), "--bogus")
    assert r.returncode == 1 and "usage" in r.stderr


def test_harness_rejects_bad_input_before_any_gpu_work(tmp_path):
    """Unreadable files, a corrupt PGM header and mismatched sizes exit 1 on the
    host (common/util.c:23-33, common/harness.c:142-145)."""
    a, b, c = (str(tmp_path / n) for n in ("a.pgm", "b.pgm", "c.pgm"))
    sgm_amd.pgm_write(a, np.zeros((20, 30), np.uint8))
    sgm_amd.pgm_write(b, np.zeros((21, 30), np.uint8))
    with open(c, "wb") as f:
        f.write(b"P2\n30 20\n255\n" + bytes(600))
    exe = os.path.join(BIN, "sgm_hip")
    r = _run(exe, a, b)
    assert r.returncode == 1 and "size mismatch" in r.stderr
    r = _run(exe, a, c)
    assert r.returncode == 1 and "not binary PGM" in r.stderr
    r = _run(exe, a, str(tmp_path / "missing.pgm"))
    assert r.returncode == 1 and "cannot open" in r.stderr
    r = _run(exe, a, a, "--bogus")
    assert r.returncode == 1 and "unknown arg" in r.stderr


def test_gen_scene_writes_the_reference_scene(tmp_path, goldens):
    """gen_scene reference W H DIR seed D writes the reference generator's exact
    PGM files (sha256 of the reference's gen_synthetic output)."""
    import hashlib
    rec = next(r for r in goldens["reference_inputs"] if r["scene"]["W"] == 1920 and r["scene"]["H"] == 800)
    sp = rec["scene"]
    r = _run(os.path.join(BIN, "gen_scene"), "reference", str(sp["W"]), str(sp["H"]), str(tmp_path),
             str(sp["seed"]), str(sp["gen_D"]))
    assert r.returncode == 0, r.stderr
    for n in ("left", "right", "gt"):
        assert hashlib.sha256(open(tmp_path / (n + ".pgm"), "rb").read()).hexdigest() == rec[n + "_sha256"], n


def test_periodic_scene_is_deterministic_and_validated(goldens):
    for r in goldens["ties"]:
        L, R = sgm_amd.make_scene(r["scene"])
        assert sgm_amd.fnv1a64(np.concatenate([L.ravel(), R.ravel()])) == r["input_fnv"], r["name"]
    with pytest.raises(sgm_amd.SgmError):
        sgm_amd.scene_periodic(64, 64, 63, 1, 1)  # period below 2
    with pytest.raises(sgm_amd.SgmError):
        sgm_amd.scene_periodic(64, 64, 63, 65, 1)  # longer than the pattern buffer


def test_host_entry_points_check_shapes_before_the_library():
    """Context.plain / ms / window refuse a right image of another shape and a
    block table of the wrong size before any library call (the library copies
    W*H bytes of each image and reads one block byte per tile)."""
    ctx = sgm_amd.Context.__new__(sgm_amd.Context)  # no device: the checks come first
    L = np.zeros((32, 48), np.uint8)
    with pytest.raises(ValueError, match="differ in shape"):
        ctx.plain(L, np.zeros((32, 47), np.uint8))
    with pytest.raises(ValueError, match="differ in shape"):
        ctx.ms(L, np.zeros((31, 48), np.uint8))
    with pytest.raises(ValueError, match="2-D"):
        ctx.plain(np.zeros(100, np.uint8), np.zeros(100, np.uint8))
    with pytest.raises(ValueError, match="tiles"):
        ctx.window(L, L, np.zeros(5, np.uint8), 64)      # a 48x32 frame has 3x2 tiles
    with pytest.raises(ValueError, match="differ in shape"):
        ctx.window(L, L[:, :40], np.zeros(6, np.uint8), 64)
