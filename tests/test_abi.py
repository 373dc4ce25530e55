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
