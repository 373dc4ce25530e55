"""oracle.py -- ctypes view of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, always as the checker or the timed CPU baseline,
never as a product code path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_P = ctypes.c_void_p
_I = ctypes.c_int
_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_fnv1a64.argtypes = [_P, ctypes.c_size_t]
        L.orc_fnv1a64.restype = ctypes.c_uint64
        L.orc_census9x7.argtypes = [_P, _I, _I, _P]
        L.orc_cost.argtypes = [_P, _P, _I, _I, _I, _I, _P]
        L.orc_step.argtypes = [_P, _P, _P, _I, _I, _I, _I]
        L.orc_path_L.argtypes = [_P, _I, _I, _I, _I, _I, _I, _I, _I, _P]
        L.orc_aggregate_dir.argtypes = [_P, _I, _I, _I, _I, _I, _I, _I, _I, _P]
        L.orc_argmin.argtypes = [_P, ctypes.c_size_t, _I, _P]
        L.orc_sgm_plain.argtypes = [_P, _P, _I, _I, _I, _I, _P]
        L.orc_sgm_plain_omp.argtypes = [_P, _P, _I, _I, _I, _I, _P, _I]
        L.orc_sobelx.argtypes = [_P, _I, _I, _P]
        L.orc_half.argtypes = [_P, _I, _I, _P]
        L.orc_sgm_ms.argtypes = [_P, _P, _I, _I, _I, _P, _P]
        L.orc_sgm_ms_omp.argtypes = [_P, _P, _I, _I, _I, _P, _P, _I]
        L.orc_sgm_window.argtypes = [_P, _P, _I, _I, _I, _I, _P, _P]
        for n in ("orc_sgm_plain", "orc_sgm_plain_omp", "orc_sgm_ms", "orc_sgm_ms_omp",
                  "orc_omp_max_threads", "orc_sgm_window"):
            getattr(L, n).restype = _I
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def _c(a, dt=np.uint8):
    return np.ascontiguousarray(a, dtype=dt)


def fnv1a64(a) -> str:
    a = np.ascontiguousarray(a)
    return "%016x" % lib().orc_fnv1a64(_p(a), a.nbytes)


def census(img) -> np.ndarray:
    img = _c(img)
    H, W = img.shape
    out = np.empty((H, W), np.uint64)
    lib().orc_census9x7(_p(img), W, H, _p(out))
    return out


def cost(cl, cr, D: int, invalid: int = 62) -> np.ndarray:
    cl, cr = _c(cl, np.uint64), _c(cr, np.uint64)
    H, W = cl.shape
    C = np.empty((H, W, D), np.uint8)
    lib().orc_cost(_p(cl), _p(cr), W, H, D, invalid, _p(C))
    return C


def step(prev, c, P1: int, P2: int, sat: bool) -> np.ndarray:
    prev, c = _c(prev), _c(c)
    cur = np.empty_like(prev)
    lib().orc_step(_p(prev), _p(c), _p(cur), prev.size, P1, P2, int(sat))
    return cur


def path_L(C, dx: int, dy: int, P1: int = 20, P2: int = 192, sat: bool = False) -> np.ndarray:
    C = _c(C)
    H, W, D = C.shape
    L = np.empty_like(C)
    lib().orc_path_L(_p(C), W, H, D, dx, dy, P1, P2, int(sat), _p(L))
    return L


def argmin(S) -> np.ndarray:
    S = _c(S, np.uint16)
    H, W, D = S.shape
    out = np.empty((H, W), np.uint8)
    lib().orc_argmin(_p(S), H * W, D, _p(out))
    return out


def sgm_plain(left, right, D: int = 64, paths: int = 4, threads: int = 0) -> np.ndarray:
    """threads == 1: the scalar oracle; otherwise the OpenMP restatement."""
    L, R = _c(left), _c(right)
    H, W = L.shape
    out = np.empty((H, W), np.uint8)
    if threads == 1:
        rc = lib().orc_sgm_plain(_p(L), _p(R), W, H, D, paths, _p(out))
    else:
        rc = lib().orc_sgm_plain_omp(_p(L), _p(R), W, H, D, paths, _p(out), threads)
    if rc:
        raise ValueError("oracle rejected the arguments")
    return out


def path_sums(left, right, D: int = 64, paths: int = 4) -> np.ndarray:
    """The plain path's uint16 S volume [H][W][D], summed from the per-direction
    L volumes of orc_path_L (common/sgm_ref.c:75-99,117-118)."""
    C = cost(census(left), census(right), D)
    dirs = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, 1), (1, -1), (-1, -1)][:paths]
    S = np.zeros(C.shape, np.uint16)
    for dx, dy in dirs:
        S += path_L(C, dx, dy)
    return S


def tie_pixels(S) -> int:
    """Pixels whose path-sum minimum is attained at more than one d: the
    reference's SGM_COUNT_TIES count (common/sgm_ref.c:125-138)."""
    S = np.asarray(S)
    return int(((S == S.min(axis=-1, keepdims=True)).sum(axis=-1) > 1).sum())


def sgm_ms(left, right, D: int = 64, threads: int = 0, want_stage1: bool = False):
    L, R = _c(left), _c(right)
    H, W = L.shape
    out = np.empty((H // 2, W // 2), np.uint8)
    out1 = np.empty((H // 2, W // 2), np.uint8)
    if threads == 1:
        rc = lib().orc_sgm_ms(_p(L), _p(R), W, H, D, _p(out), _p(out1))
    else:
        rc = lib().orc_sgm_ms_omp(_p(L), _p(R), W, H, D, _p(out), _p(out1), threads)
    if rc:
        raise ValueError("oracle rejected the arguments")
    return (out, out1) if want_stage1 else out


def sobelx(img) -> np.ndarray:
    img = _c(img)
    H, W = img.shape
    out = np.empty_like(img)
    lib().orc_sobelx(_p(img), W, H, _p(out))
    return out


def half(img) -> np.ndarray:
    img = _c(img)
    H, W = img.shape
    out = np.empty((H // 2, W // 2), np.uint8)
    lib().orc_half(_p(img), W, H, _p(out))
    return out


def omp_threads() -> int:
    return lib().orc_omp_max_threads()


def sgm_window(left, right, blk, D: int = 128, paths: int = 4) -> np.ndarray:
    """Windowed SGM; blk[ty, tx] = b0 | b1 << 4 (active 16-blocks, inclusive)."""
    L, R = _c(left), _c(right)
    H, W = L.shape
    blk = _c(blk)
    assert blk.shape == ((H + 15) // 16, (W + 15) // 16)
    out = np.empty((H, W), np.uint8)
    if lib().orc_sgm_window(_p(L), _p(R), W, H, D, paths, _p(blk), _p(out)):
        raise ValueError("oracle rejected the arguments")
    return out
