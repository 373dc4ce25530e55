"""sgm_amd -- Python binding of the MI355X SGM engine (libsgm_hip.so) and of its
plain-C host helpers (libsgm_host.so), over ctypes.

The engine itself is the C ABI in include/sgm_hip.h; this module only marshals
numpy arrays (host buffers) or raw device pointers (for example from torch
tensors) into it.  There is no fallback: if the shared libraries are missing,
importing the GPU entry points raises, and every call that fails in the library
raises SgmError with the library's own message.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(os.path.dirname(_HERE), "lib")
REPO = os.path.dirname(os.path.dirname(_HERE))

_P = ctypes.c_void_p
_I = ctypes.c_int


class SgmError(RuntimeError):
    pass


def _load(name: str) -> ctypes.CDLL:
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built; run `make -C sgm-bench_amd` "
                          f"(or __graft_entry__.build()) first")
    return ctypes.CDLL(path)


_host: Optional[ctypes.CDLL] = None
_hip: Optional[ctypes.CDLL] = None


def host_lib() -> ctypes.CDLL:
    global _host
    if _host is None:
        h = _load("libsgm_host.so")
        h.sgm_scene_reference.argtypes = [_I, _I, _I, ctypes.c_uint32, _P, _P, _P]
        h.sgm_scene_reference.restype = _I
        h.sgm_scene_planes.argtypes = [ctypes.POINTER(SceneSpec), _P, _P, _P]
        h.sgm_scene_planes.restype = _I
        h.sgm_scene_periodic.argtypes = [_I, _I, _I, _I, ctypes.c_uint32, _P, _P, _P]
        h.sgm_scene_periodic.restype = _I
        h.sgm_fnv1a64.argtypes = [_P, ctypes.c_size_t]
        h.sgm_fnv1a64.restype = ctypes.c_uint64
        h.sgm_pgm_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(_I), ctypes.POINTER(_I)]
        h.sgm_pgm_read.restype = _P
        h.sgm_pgm_write.argtypes = [ctypes.c_char_p, _P, _I, _I]
        h.sgm_pgm_write.restype = _I
        h.sgm_free.argtypes = [_P]
        h.sgm_now_ms.restype = ctypes.c_double
        _host = h
    return _host


class Times(ctypes.Structure):
    """sgmh_times (include/sgm_hip.h)."""
    _fields_ = [("census_ms", ctypes.c_double), ("cost_ms", ctypes.c_double),
                ("aggregate_ms", ctypes.c_double), ("argmin_ms", ctypes.c_double),
                ("transfer_ms", ctypes.c_double), ("total_ms", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class SceneSpec(ctypes.Structure):
    """sgm_scene_spec (include/sgm_scene.h)."""
    _fields_ = [("W", _I), ("H", _I), ("dmax", _I), ("seed", ctypes.c_uint32),
                ("planes_lo", _I), ("planes_hi", _I), ("occl_permille", _I),
                ("noise", _I), ("slanted", _I)]


class Pair(ctypes.Structure):
    """sgmh_pair (include/sgm_hip.h)."""
    _fields_ = [("d_left", _P), ("d_right", _P), ("W", _I), ("H", _I), ("D", _I),
                ("mode", _I), ("paths", _I), ("d_disp", _P)]


PLAIN, MULTISCALE = 0, 1
BATCH_AUTOTUNE, BATCH_NO_GRAPH, BATCH_CENSUS_32X32, BATCH_XCD_SLICES, BATCH_FUSED = 1, 2, 4, 8, 16
BATCH_PIPELINED = 32
LAYOUT_AUTO, LAYOUT_LINES4, LAYOUT_WAVE, LAYOUT_PLACED = 0, 1, 2, 3  # SGMH_LAYOUT_* (include/sgm_hip.h)


def hip_lib() -> ctypes.CDLL:
    global _hip
    if _hip is None:
        # One HIP runtime per process: when torch is installed, let it load its
        # runtime first so the library binds to the same one (a runtime loaded
        # after another has initialised the device finds no GPU).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        h = _load("libsgm_hip.so")
        h.sgmh_last_error.restype = ctypes.c_char_p
        h.sgmh_ctx_create.argtypes = [ctypes.POINTER(_P), _I]
        h.sgmh_ctx_destroy.argtypes = [_P]
        h.sgmh_ctx_destroy.restype = None
        h.sgmh_set_timing.argtypes = [_P, _I]
        h.sgmh_set_path_layout.argtypes = [_P, _I]
        h.sgmh_get_times.argtypes = [_P, ctypes.POINTER(Times)]
        h.sgmh_plain.argtypes = [_P, _P, _P, _I, _I, _I, _I, _P, ctypes.POINTER(Times)]
        h.sgmh_plain_dev.argtypes = [_P, _P, _P, _I, _I, _I, _I, _P, _P]
        h.sgmh_plain_noalloc_dev.argtypes = [_P, _P, _I, _I, _I, _I, _P, _P]
        h.sgmh_ms.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, ctypes.POINTER(Times)]
        h.sgmh_ms_dev.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, _P]
        h.sgmh_census_dev.argtypes = [_P, _P, _I, _I, _P, _P]
        h.sgmh_sobel_dev.argtypes = [_P, _P, _I, _I, _P, _P]
        h.sgmh_half_dev.argtypes = [_P, _P, _I, _I, _P, _P]
        h.sgmh_path_plain_dev.argtypes = [_P, _P, _P, _I, _I, _I, _I, _I, _P, _P]
        h.sgmh_copy_probe.argtypes = [_P, _P, _P, ctypes.c_size_t, _P]
        h.sgmh_window.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, ctypes.POINTER(Times)]
        h.sgmh_window_dev.argtypes = [_P, _P, _P, _I, _I, _I, _P, _P, _P]
        h.sgmh_cnn_num_params.restype = ctypes.c_size_t
        h.sgmh_cnn_create.argtypes = [_P, ctypes.POINTER(_P)]
        h.sgmh_cnn_destroy.argtypes = [_P]
        h.sgmh_cnn_destroy.restype = None
        h.sgmh_cnn_set_params.argtypes = [_P, _P, ctypes.c_size_t]
        h.sgmh_cnn_get_params.argtypes = [_P, _P, ctypes.c_size_t]
        h.sgmh_cnn_propose_dev.argtypes = [_P, _P, _P, _I, _I, ctypes.c_float, _P, _P, _P]
        h.sgmh_learned_dev.argtypes = [_P, _P, _P, _I, _I, ctypes.c_float, _P, _P, _P]
        h.sgmh_cnn_train_step_dev.argtypes = [_P, _P, _P, _P, _I, _I, ctypes.c_float,
                                              ctypes.POINTER(ctypes.c_float), _P]
        h.sgmh_cnn_get_grads.argtypes = [_P, _P, ctypes.c_size_t]
        h.sgmh_cnn_get_activations.argtypes = [_P, _P, _P]
        h.sgmh_batch_plan.argtypes = [_P, ctypes.POINTER(Pair), _I, _I, ctypes.POINTER(_P)]
        h.sgmh_batch_run.argtypes = [_P, _P]
        h.sgmh_batch_run_loop.argtypes = [_P, _P]
        h.sgmh_batch_describe.argtypes = [_P, ctypes.c_char_p, ctypes.c_size_t]
        h.sgmh_batch_destroy.argtypes = [_P]
        h.sgmh_batch_destroy.restype = None
        for name in ("sgmh_ctx_create", "sgmh_set_timing", "sgmh_set_path_layout", "sgmh_get_times", "sgmh_plain",
                     "sgmh_plain_dev", "sgmh_plain_noalloc_dev", "sgmh_ms", "sgmh_ms_dev", "sgmh_census_dev",
                     "sgmh_sobel_dev", "sgmh_half_dev", "sgmh_path_plain_dev", "sgmh_copy_probe",
                     "sgmh_batch_plan", "sgmh_batch_run", "sgmh_batch_run_loop",
                     "sgmh_batch_describe", "sgmh_window", "sgmh_window_dev",
                     "sgmh_cnn_create", "sgmh_cnn_set_params", "sgmh_cnn_get_params",
                     "sgmh_cnn_propose_dev", "sgmh_learned_dev", "sgmh_cnn_train_step_dev", "sgmh_cnn_get_grads",
                     "sgmh_cnn_get_activations"):
            getattr(h, name).restype = _I
        _hip = h
    return _hip


NOALLOC_SCRATCH_BYTES = 128 << 20  # SGMH_NOALLOC_SCRATCH_BYTES (include/sgm_hip.h)


def plain_noalloc_dev(d_left: int, d_right: int, W: int, H: int, D: int, paths: int,
                      d_disp: int, stream: int = 0) -> None:
    """sgmh_plain_noalloc_dev: the plain path on the current device with no
    device allocation (the library's static scratch array); no context."""
    lib = hip_lib()
    if lib.sgmh_plain_noalloc_dev(d_left, d_right, W, H, D, paths, d_disp, stream or None):
        raise SgmError((lib.sgmh_last_error() or b"").decode())


def noalloc_scratch_bytes(W: int, H: int, D: int, paths: int) -> int:
    """Scratch bytes one sgmh_plain_noalloc_dev call needs (include/sgm_hip.h)."""
    a = ((W * H + 256) * 8 + 255) // 256 * 256
    return 2 * a + paths * W * H * D


def _ptr(a: np.ndarray) -> _P:
    return a.ctypes.data_as(_P)


def _u8(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a


def _pair(left, right):
    """The two images of a pair as C-contiguous uint8 (H, W) arrays of one shape:
    the library copies W*H bytes from each, so a smaller right image would be
    read past its end."""
    L, R = _u8(left), _u8(right)
    if L.ndim != 2:
        raise ValueError(f"images must be 2-D (H, W) arrays, got shape {L.shape}")
    if R.shape != L.shape:
        raise ValueError(f"left {L.shape} and right {R.shape} images differ in shape")
    return L, R


def fnv1a64(a: np.ndarray) -> str:
    """FNV-1a 64 of the bytes of `a`, as the reference prints it (%016llx)."""
    a = np.ascontiguousarray(a)
    return "%016x" % host_lib().sgm_fnv1a64(_ptr(a), a.nbytes)


def scene_reference(W: int, H: int, D: int, seed: int) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """The reference's gen_synthetic scene (common/gen_synthetic.c), as (H, W) arrays."""
    L = np.empty((H, W), np.uint8)
    R = np.empty_like(L)
    G = np.empty_like(L)
    if host_lib().sgm_scene_reference(W, H, D, seed, _ptr(L), _ptr(R), _ptr(G)):
        raise SgmError(f"sgm_scene_reference rejected {W}x{H} D={D}")
    return L, R, G


def scene_planes(W: int, H: int, dmax: int, seed: int, planes: Tuple[int, int] = (8, 12),
                 occl_permille: int = 0, noise: int = 2, slanted: bool = False):
    """Seeded piecewise-planar scene (include/sgm_scene.h), as (H, W) arrays."""
    spec = SceneSpec(W, H, dmax, seed, planes[0], planes[1], occl_permille, noise, int(slanted))
    L = np.empty((H, W), np.uint8)
    R = np.empty_like(L)
    G = np.empty_like(L)
    if host_lib().sgm_scene_planes(ctypes.byref(spec), _ptr(L), _ptr(R), _ptr(G)):
        raise SgmError("sgm_scene_planes rejected the spec")
    return L, R, G


def scene_periodic(W: int, H: int, dmax: int, period: int, seed: int):
    """Tie-rich scene (include/sgm_scene.h: sgm_scene_periodic), as (H, W) arrays."""
    L = np.empty((H, W), np.uint8)
    R = np.empty_like(L)
    G = np.empty_like(L)
    if host_lib().sgm_scene_periodic(W, H, dmax, period, seed, _ptr(L), _ptr(R), _ptr(G)):
        raise SgmError(f"sgm_scene_periodic rejected {W}x{H} dmax={dmax} period={period}")
    return L, R, G


# Named workloads: (W, H, D, generator kwargs).  The plain configurations are
# 4-path; the multiscale ones 8-path two-scale.
WORKLOADS = {
    "P": dict(W=640, H=480, D=64, planes=(8, 12), occl_permille=0, seed=11),
    "A": dict(W=1920, H=1080, D=64, planes=(20, 20), occl_permille=50, seed=21),
    "B": dict(W=1920, H=800, D=64, planes=(20, 20), occl_permille=50, seed=31),
    "C": dict(W=1920, H=1080, D=64, planes=(20, 20), occl_permille=50, seed=41),
}


# The batched workload: 24 pairs of uneven size, D and configuration in one
# submission -- 8 x 640x480 D=64 plain, 8 x 1280x720 D=64 plain,
# 4 x 1920x800 D=128 multiscale 8-path, 4 x 1920x1080 D=128 plain.
def batch_spec():
    out = []
    for i in range(8):
        out.append(dict(name=f"p640_{i}", W=640, H=480, D=64, mode=PLAIN, paths=4,
                        scene=dict(kind="planes", W=640, H=480, dmax=63, seed=200 + i,
                                   planes=[8, 12], occl=0, noise=2, slanted=0)))
    for i in range(8):
        out.append(dict(name=f"p1280_{i}", W=1280, H=720, D=64, mode=PLAIN, paths=4,
                        scene=dict(kind="planes", W=1280, H=720, dmax=63, seed=210 + i,
                                   planes=[8, 12], occl=20, noise=2, slanted=0)))
    for i in range(4):
        out.append(dict(name=f"ms1920x800_{i}", W=1920, H=800, D=128, mode=MULTISCALE, paths=8,
                        scene=dict(kind="planes", W=1920, H=800, dmax=127, seed=220 + i,
                                   planes=[20, 20], occl=50, noise=2, slanted=0)))
    for i in range(4):
        out.append(dict(name=f"p1080_{i}", W=1920, H=1080, D=128, mode=PLAIN, paths=4,
                        scene=dict(kind="planes", W=1920, H=1080, dmax=127, seed=230 + i,
                                   planes=[20, 20], occl=50, noise=2, slanted=0)))
    return out


def make_scene(spec: dict):
    """A scene from its recorded generator call (the fixtures' "scene" dicts)."""
    if spec["kind"] == "reference":
        L, R, _ = scene_reference(spec["W"], spec["H"], spec["gen_D"], spec["seed"])
    elif spec["kind"] == "periodic":
        L, R, _ = scene_periodic(spec["W"], spec["H"], spec["dmax"], spec["period"], spec["seed"])
    else:
        L, R, _ = scene_planes(spec["W"], spec["H"], spec["dmax"], spec["seed"],
                               planes=tuple(spec["planes"]), occl_permille=spec["occl"],
                               noise=spec["noise"], slanted=bool(spec["slanted"]))
    return L, R


def workload_scene(name: str):
    w = WORKLOADS[name]
    return scene_planes(w["W"], w["H"], w["D"] - 1, w["seed"], planes=w["planes"],
                        occl_permille=w["occl_permille"], noise=2)


def pgm_read(path: str) -> np.ndarray:
    W, H = _I(), _I()
    p = host_lib().sgm_pgm_read(path.encode(), ctypes.byref(W), ctypes.byref(H))
    if not p:
        raise SgmError(f"cannot read PGM {path}")
    try:
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_uint8)),
                                     shape=(H.value, W.value)).copy()
    finally:
        host_lib().sgm_free(p)


def pgm_write(path: str, img: np.ndarray) -> None:
    img = _u8(img)
    H, W = img.shape
    if host_lib().sgm_pgm_write(path.encode(), _ptr(img), W, H):
        raise SgmError(f"cannot write PGM {path}")


class Context:
    """One sgmh_ctx: a device, a stream, and the device workspaces."""

    def __init__(self, device: int = 0):
        self._lib = hip_lib()
        h = _P()
        if self._lib.sgmh_ctx_create(ctypes.byref(h), device):
            raise SgmError(self.error())
        self._h = h

    def error(self) -> str:
        return (self._lib.sgmh_last_error() or b"").decode()

    def _ck(self, rc: int) -> None:
        if rc:
            raise SgmError(self.error())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sgmh_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # ---- timing ----
    def set_timing(self, on: bool) -> None:
        self._ck(self._lib.sgmh_set_timing(self._h, int(on)))

    def set_path_layout(self, layout: int) -> None:
        """LAYOUT_AUTO, LAYOUT_LINES4 (four scanlines per wave) or LAYOUT_WAVE
        (one scanline per wave) for this context's plain path kernel."""
        self._ck(self._lib.sgmh_set_path_layout(self._h, int(layout)))

    def times(self) -> dict:
        t = Times()
        self._ck(self._lib.sgmh_get_times(self._h, ctypes.byref(t)))
        return t.as_dict()

    # ---- host-buffer entry points ----
    def plain(self, left: np.ndarray, right: np.ndarray, D: int = 64, paths: int = 4,
              times: Optional[dict] = None) -> np.ndarray:
        L, R = _pair(left, right)
        H, W = L.shape
        out = np.empty((H, W), np.uint8)
        t = Times()
        self._ck(self._lib.sgmh_plain(self._h, _ptr(L), _ptr(R), W, H, D, paths, _ptr(out),
                                      ctypes.byref(t)))
        if times is not None:
            times.update(t.as_dict())
        return out

    def ms(self, left: np.ndarray, right: np.ndarray, D: int = 64, want_stage1: bool = False,
           times: Optional[dict] = None):
        L, R = _pair(left, right)
        H, W = L.shape
        out = np.empty((H // 2, W // 2), np.uint8)
        out1 = np.empty((H // 2, W // 2), np.uint8) if want_stage1 else None
        t = Times()
        self._ck(self._lib.sgmh_ms(self._h, _ptr(L), _ptr(R), W, H, D, _ptr(out),
                                   _ptr(out1) if out1 is not None else None, ctypes.byref(t)))
        if times is not None:
            times.update(t.as_dict())
        return (out, out1) if want_stage1 else out

    # ---- device-pointer entry points (ints are raw device addresses) ----
    def plain_dev(self, d_left: int, d_right: int, W: int, H: int, D: int, paths: int,
                  d_disp: int, stream: int = 0) -> None:
        self._ck(self._lib.sgmh_plain_dev(self._h, d_left, d_right, W, H, D, paths, d_disp,
                                          stream or None))

    def ms_dev(self, d_left: int, d_right: int, W: int, H: int, D: int, d_disp: int,
               d_disp1: int = 0, stream: int = 0) -> None:
        self._ck(self._lib.sgmh_ms_dev(self._h, d_left, d_right, W, H, D, d_disp,
                                       d_disp1 or None, stream or None))

    def census_dev(self, d_img: int, W: int, H: int, d_codes: int, stream: int = 0) -> None:
        self._ck(self._lib.sgmh_census_dev(self._h, d_img, W, H, d_codes, stream or None))

    def sobel_dev(self, d_img: int, W: int, H: int, d_out: int, stream: int = 0) -> None:
        self._ck(self._lib.sgmh_sobel_dev(self._h, d_img, W, H, d_out, stream or None))

    def half_dev(self, d_img: int, W: int, H: int, d_out: int, stream: int = 0) -> None:
        self._ck(self._lib.sgmh_half_dev(self._h, d_img, W, H, d_out, stream or None))

    def window(self, left: np.ndarray, right: np.ndarray, blk: np.ndarray, D: int = 128,
               times: Optional[dict] = None) -> np.ndarray:
        """Windowed SGM; blk[ty, tx] = first block | last block << 4."""
        L, R = _pair(left, right)
        H, W = L.shape
        blk = _u8(blk)
        tiles = ((H + 15) // 16) * ((W + 15) // 16)
        if blk.size != tiles:
            raise ValueError(f"block table has {blk.size} entries; a {W}x{H} frame has {tiles} tiles")
        out = np.empty((H, W), np.uint8)
        t = Times()
        self._ck(self._lib.sgmh_window(self._h, _ptr(L), _ptr(R), W, H, D, _ptr(blk), _ptr(out),
                                       ctypes.byref(t)))
        if times is not None:
            times.update(t.as_dict())
        return out

    def window_dev(self, d_left: int, d_right: int, W: int, H: int, D: int, d_blk: int,
                   d_disp: int, stream: int = 0) -> None:
        self._ck(self._lib.sgmh_window_dev(self._h, d_left, d_right, W, H, D, d_blk, d_disp,
                                           stream or None))

    def batch(self, pairs, flags: int = BATCH_AUTOTUNE) -> "Batch":
        return Batch(self, pairs, flags)

    def copy_probe(self, d_dst: int, d_src: int, nbytes: int, stream: int = 0) -> None:
        self._ck(self._lib.sgmh_copy_probe(self._h, d_dst, d_src, nbytes, stream or None))

    def path_plain_dev(self, d_cl: int, d_cr: int, W: int, H: int, D: int, dx: int, dy: int,
                       d_L: int, stream: int = 0) -> None:
        self._ck(self._lib.sgmh_path_plain_dev(self._h, d_cl, d_cr, W, H, D, dx, dy, d_L,
                                               stream or None))


class Batch:
    """A planned batch (sgmh_batch): pairs = list of dicts with keys d_left,
    d_right, d_disp (device addresses), W, H, D, mode, paths."""

    def __init__(self, ctx: Context, pairs, flags: int = BATCH_AUTOTUNE):
        self._ctx = ctx
        self._lib = ctx._lib
        arr = (Pair * len(pairs))()
        for i, p in enumerate(pairs):
            arr[i] = Pair(p["d_left"], p["d_right"], p["W"], p["H"], p["D"], p["mode"],
                          p["paths"], p["d_disp"])
        h = _P()
        ctx._ck(self._lib.sgmh_batch_plan(ctx.handle, arr, len(pairs), flags, ctypes.byref(h)))
        self._h = h

    def run(self, stream: int = 0) -> None:
        self._ctx._ck(self._lib.sgmh_batch_run(self._h, stream or None))

    def run_loop(self, stream: int = 0) -> None:
        self._ctx._ck(self._lib.sgmh_batch_run_loop(self._h, stream or None))

    def describe(self) -> dict:
        import json
        buf = ctypes.create_string_buffer(65536)
        n = self._lib.sgmh_batch_describe(self._h, buf, len(buf))
        if n >= len(buf):  # truncated: the return value is the full length
            buf = ctypes.create_string_buffer(n + 1)
            n = self._lib.sgmh_batch_describe(self._h, buf, len(buf))
        if n < 0:
            raise SgmError("sgmh_batch_describe failed")
        return json.loads(buf.value.decode())

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sgmh_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Cnn:
    """The window-proposal network (sgmh_cnn).  Device pointers are ints."""

    def __init__(self, ctx: Context, params: Optional[np.ndarray] = None):
        self._ctx = ctx
        self._lib = ctx._lib
        h = _P()
        ctx._ck(self._lib.sgmh_cnn_create(ctx.handle, ctypes.byref(h)))
        self._h = h
        self.nparams = int(self._lib.sgmh_cnn_num_params())
        if params is not None:
            self.set_params(params)

    def set_params(self, theta: np.ndarray) -> None:
        theta = np.ascontiguousarray(theta, np.float32)
        assert theta.size == self.nparams
        self._ctx._ck(self._lib.sgmh_cnn_set_params(self._h, _ptr(theta), theta.size))

    def params(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        self._ctx._ck(self._lib.sgmh_cnn_get_params(self._h, _ptr(out), out.size))
        return out

    def grads(self) -> np.ndarray:
        out = np.empty(self.nparams, np.float32)
        self._ctx._ck(self._lib.sgmh_cnn_get_grads(self._h, _ptr(out), out.size))
        return out

    def activations(self, W: int, H: int):
        f = np.empty((H // 4, W // 4, 32), np.uint16)
        z = np.empty((H // 4, W // 4, 32), np.float32)
        self._ctx._ck(self._lib.sgmh_cnn_get_activations(self._h, _ptr(f), _ptr(z)))
        return (f.astype(np.uint32) << 16).view(np.float32), z

    def propose_dev(self, d_left: int, d_right: int, W: int, H: int, thresh: float, d_blk: int,
                    d_conf: int = 0, stream: int = 0) -> None:
        self._ctx._ck(self._lib.sgmh_cnn_propose_dev(self._h, d_left, d_right, W, H, thresh, d_blk,
                                                      d_conf or None, stream or None))

    def learned_dev(self, d_left: int, d_right: int, W: int, H: int, thresh: float, d_disp: int,
                    d_blk: int = 0, stream: int = 0) -> None:
        """Proposal + windowed SGM at D=128 in one call (census beside the network)."""
        self._ctx._ck(self._lib.sgmh_learned_dev(self._h, d_left, d_right, W, H, thresh, d_blk or None,
                                                  d_disp, stream or None))

    def train_step_dev(self, d_left: int, d_right: int, d_target: int, W: int, H: int, lr: float,
                       stream: int = 0) -> float:
        loss = ctypes.c_float(0.0)
        self._ctx._ck(self._lib.sgmh_cnn_train_step_dev(self._h, d_left, d_right, d_target, W, H,
                                                         lr, ctypes.byref(loss), stream or None))
        return float(loss.value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.sgmh_cnn_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
