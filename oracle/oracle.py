"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg import
this module, and only to check results (or as the timed CPU baseline).  The product
package ``tcstencil_amd`` never imports it.

The arithmetic lives in ``oracle/gold.c`` (a restatement of the reference's
``common/gold_*.cpp``); ``oracle/_ref/`` holds the reference's own gold files
compiled from ``/root/reference`` by ``oracle/Makefile`` (target ``ref``).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

STENCILS = ("star2d1r", "star2d2r", "box2d1r", "box2d2r")
STENCIL_ID = {s: i for i, s in enumerate(STENCILS)}
RADIUS = {"star2d1r": 1, "star2d2r": 2, "box2d1r": 1, "box2d2r": 2}

_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)
_lib = None


def build(quiet: bool = True) -> None:
    """Compile liboracle.so (and the reference gold libraries when the
    reference sources are present)."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.run(["make", "-s", "-C", HERE], check=True, stdout=out)
    if os.path.isdir("/root/reference/common"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True, stdout=out)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name in STENCILS:
            getattr(L, "gold_" + name).argtypes = [_F32P, _F32P, _F32P, ctypes.c_int, ctypes.c_int]
            getattr(L, "gold_" + name + "_f64").argtypes = [_F64P, _F64P, _F64P, ctypes.c_int, ctypes.c_int]
        L.oracle_sweep_f32.argtypes = [ctypes.c_int, _F32P, _F32P, _F32P, ctypes.c_int, ctypes.c_int]
        L.oracle_sweep_f32.restype = ctypes.c_int
        L.oracle_sweep_f64.argtypes = [ctypes.c_int, _F64P, _F64P, _F64P, ctypes.c_int, ctypes.c_int]
        L.oracle_sweep_f64.restype = ctypes.c_int
        L.oracle_sweeps_f64.argtypes = [ctypes.c_int, _F64P, _F64P, _F64P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_sweeps_f64.restype = ctypes.c_int
        L.oracle_error_norms.argtypes = [_F32P, _F64P, ctypes.c_int, ctypes.c_int, ctypes.c_int, _F64P]
        L.oracle_check_error_2d.argtypes = [ctypes.c_int, _F32P, _F32P] + [ctypes.c_int] * 4
        L.oracle_check_error_2d.restype = ctypes.c_double
        L.oracle_check_relative_error_2d.argtypes = [ctypes.c_int, _F32P, _F32P] + [ctypes.c_int] * 4
        L.oracle_check_relative_error_2d.restype = ctypes.c_double
        L.oracle_check_relative_error_3d.argtypes = [ctypes.c_int, _F32P, _F32P] + [ctypes.c_int] * 6
        L.oracle_check_relative_error_3d.restype = ctypes.c_double
        L.oracle_mix64.argtypes = [ctypes.c_uint64]
        L.oracle_mix64.restype = ctypes.c_uint64
        L.oracle_gen_uniform_f32.argtypes = [_F32P, ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                             ctypes.c_uint64, ctypes.c_float, ctypes.c_float]
        _lib = L
    return _lib


def _p32(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F32P)


def _p64(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_F64P)


def gold(stencil: str, A: np.ndarray, B: np.ndarray, C: np.ndarray | None = None) -> np.ndarray:
    """Reference-signature FP32 gold on a square grid.  C defaults to a copy
    of A (the reference caller pre-seeds C from A, common/main_2r.cu:29)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    N = A.shape[0]
    assert A.shape == (N, N)
    C = A.copy() if C is None else C
    getattr(lib(), "gold_" + stencil)(_p32(A), _p32(B), _p32(C), N, B.shape[0])
    return C


def sweep_f32_threads(stencil: str, A: np.ndarray, B: np.ndarray, C: np.ndarray, threads: int) -> None:
    """One FP32 sweep A -> C (C seeded with A's ring by the caller) by the
    oracle's loop on `threads` bands of rows, one Python thread per band (the
    ctypes call releases the interpreter lock).  Each band is the oracle's
    rectangular sweep on rows [lo - r, hi + r) of A, which writes rows
    [lo, hi) of C: the same arithmetic per point as sweep_f32."""
    import threading
    assert A.dtype == C.dtype == np.float32 and A.flags.c_contiguous and C.flags.c_contiguous
    B = np.ascontiguousarray(B, dtype=np.float32)
    rows, cols = A.shape
    r = RADIUS[stencil]
    edges = np.linspace(r, rows - r, threads + 1).astype(int)
    L, sid = lib(), STENCIL_ID[stencil]
    rcs = []

    def band(lo, hi):
        off = (lo - r) * cols * 4
        a = ctypes.cast(A.ctypes.data + off, _F32P)
        c = ctypes.cast(C.ctypes.data + off, _F32P)
        rcs.append(L.oracle_sweep_f32(sid, a, _p32(B), c, hi - lo + 2 * r, cols))

    ts = [threading.Thread(target=band, args=(int(lo), int(hi))) for lo, hi in zip(edges[:-1], edges[1:]) if hi > lo]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert rcs and all(rc == 0 for rc in rcs)


def sweep_f32(stencil: str, A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """One FP32 sweep on a rows x cols grid; the ring keeps A's values."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    C = A.copy()
    rc = lib().oracle_sweep_f32(STENCIL_ID[stencil], _p32(A), _p32(B), _p32(C), A.shape[0], A.shape[1])
    assert rc == 0
    return C


def sweep_f64(stencil: str, A: np.ndarray, B: np.ndarray) -> np.ndarray:
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.ascontiguousarray(B, dtype=np.float64)
    C = A.copy()
    rc = lib().oracle_sweep_f64(STENCIL_ID[stencil], _p64(A), _p64(B), _p64(C), A.shape[0], A.shape[1])
    assert rc == 0
    return C


def sweeps_f64(stencil: str, A: np.ndarray, B: np.ndarray, sweeps: int) -> np.ndarray:
    X = np.array(A, dtype=np.float64, order="C", copy=True)
    B = np.ascontiguousarray(B, dtype=np.float64)
    tmp = np.empty_like(X)
    rc = lib().oracle_sweeps_f64(STENCIL_ID[stencil], _p64(X), _p64(B), _p64(tmp), X.shape[0], X.shape[1], sweeps)
    assert rc == 0
    return X


def error_norms(out: np.ndarray, ref: np.ndarray, r: int) -> dict:
    out = np.ascontiguousarray(out, dtype=np.float32)
    ref = np.ascontiguousarray(ref, dtype=np.float64)
    res = np.zeros(4, dtype=np.float64)
    lib().oracle_error_norms(_p32(out), _p64(ref), out.shape[0], out.shape[1], r, _p64(res))
    return {"max_abs": res[0], "rel_l2": res[1], "rms_abs": res[2], "ref_l2": res[3]}


def check_error_2d(N, output, reference, y_lb, y_ub, x_lb, x_ub) -> float:
    return lib().oracle_check_error_2d(N, _p32(output), _p32(reference), y_lb, y_ub, x_lb, x_ub)


def check_relative_error_2d(N, output, reference, y_lb, y_ub, x_lb, x_ub) -> float:
    return lib().oracle_check_relative_error_2d(N, _p32(output), _p32(reference), y_lb, y_ub, x_lb, x_ub)


def check_relative_error_3d(N, output, reference, z_lb, z_ub, y_lb, y_ub, x_lb, x_ub) -> float:
    """common/common.hpp:94-124 on an N x N x N volume (z slowest)."""
    output = np.ascontiguousarray(output, dtype=np.float32)
    reference = np.ascontiguousarray(reference, dtype=np.float32)
    assert output.size == reference.size == N ** 3
    return lib().oracle_check_relative_error_3d(N, _p32(output), _p32(reference), z_lb, z_ub, y_lb, y_ub,
                                                x_lb, x_ub)


def mix64(z: int) -> int:
    return int(lib().oracle_mix64(z))


def gen_uniform_f32(rows: int, cols: int, seed: int, stream: int, lo: float, hi: float) -> np.ndarray:
    out = np.empty((rows, cols), dtype=np.float32)
    lib().oracle_gen_uniform_f32(_p32(out), rows, cols, seed, stream, lo, hi)
    return out


# ---------------------------------------------------------------------------
# The reference's own gold (compiled from /root/reference by `make ref`).
# ---------------------------------------------------------------------------
_ref_libs: dict = {}


def ref_available(stencil: str, fast: bool = False) -> bool:
    suffix = "_fast" if fast else ""
    return os.path.exists(os.path.join(REF_DIR, f"libgold_{stencil}{suffix}.so"))


def ref_gold(stencil: str, A: np.ndarray, B: np.ndarray, C: np.ndarray | None = None,
             fast: bool = False) -> np.ndarray:
    """Run the reference's `gold` (common/gold_<stencil>.cpp) compiled here."""
    key = (stencil, fast)
    if key not in _ref_libs:
        suffix = "_fast" if fast else ""
        L = ctypes.CDLL(os.path.join(REF_DIR, f"libgold_{stencil}{suffix}.so"))
        L.gold.argtypes = [_F32P, _F32P, _F32P, ctypes.c_int, ctypes.c_int]
        _ref_libs[key] = L
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    N = A.shape[0]
    C = A.copy() if C is None else C
    _ref_libs[key].gold(_p32(A), _p32(B), _p32(C), N, B.shape[0])
    return C


def init_B(stencil: str) -> np.ndarray:
    """The reference driver's default coefficients (common/main_1r.cu:10-16,
    common/main_2r.cu:13-19): B[i] = i with the centre set to -40 (3x3) or
    -290 (5x5)."""
    n = 2 * RADIUS[stencil] + 1
    B = np.arange(n * n, dtype=np.float32)
    B[(n * n) // 2] = -40.0 if n == 3 else -290.0
    return B.reshape(n, n)
