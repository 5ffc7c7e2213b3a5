"""Python front end of the device-pointer API (include/tcstencil.h).

PyTorch supplies device memory and streams only; every computation is a call
into libtcstencil.so.  Device buffers are torch tensors on ``cuda``; streams
default to torch's current stream, so ``torch.cuda.Event`` timing brackets
exactly the kernels launched here.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from ._lib import check, lib


def _torch():
    import torch  # imported lazily: the host drivers need no torch
    return torch


def _stream(stream=None):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t) -> ctypes.c_void_p:
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def _grid_ptr(t, rows: int, cols: int, what: str) -> ctypes.c_void_p:
    """A row-major grid view: unit column stride, any row stride >= cols (the
    C functions take it as the leading dimension)."""
    if not t.is_cuda:
        raise ValueError(f"{what}: expected a device tensor")
    if t.dim() != 2 or t.shape[0] != rows or t.shape[1] != cols:
        raise ValueError(f"{what}: shape {tuple(t.shape)} != grid {rows}x{cols}")
    if t.stride(1) != 1 or t.stride(0) < cols:
        raise ValueError(f"{what}: expected unit column stride and row stride >= {cols}")
    return ctypes.c_void_p(t.data_ptr())


def _fp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def stencil_id(s) -> int:
    return L.STENCIL_ID[s] if isinstance(s, str) else int(s)


def radius(stencil) -> int:
    return lib().tcs_radius(stencil_id(stencil))


class Problem:
    """A tcs_problem descriptor: stencil, method, precision, grid, coefficients."""

    def __init__(self, stencil, rows: int, cols: int, B, method="mfma", precision="fp16",
                 tile=None, wg_per_cu=None, fuse=None, frag: int = 16, coef_lo: bool = False, tuned=None):
        self.stencil = stencil_id(stencil)
        self.method = L.METHOD_ID[method] if isinstance(method, str) else int(method)
        self.precision = L.PREC_ID[precision] if isinstance(precision, str) else int(precision)
        self.B = np.ascontiguousarray(B, dtype=np.float32)
        n = self.B.shape[0]
        self.s = L.ProblemStruct()
        check(lib().tcs_problem_init(ctypes.byref(self.s), self.stencil, self.method,
                                     self.precision, rows, cols, _fp(self.B), n))
        if frag != 16:
            check(lib().tcs_problem_set_frag(ctypes.byref(self.s), frag))
        if coef_lo:
            check(lib().tcs_problem_set_coef_lo(ctypes.byref(self.s), 1))
        if tuned is not None:
            # tuned=(path or "", batched): the cached parameters of that cache file
            self.apply_tuning(*tuned)
        if tile is not None:
            self.set_tile(*tile)
        if wg_per_cu is not None:
            self.s.wg_per_cu = wg_per_cu
        if fuse is not None:
            self.s.fuse = fuse

    # -- descriptor ---------------------------------------------------------
    def set_tile(self, tile_h: int, tile_w: int) -> None:
        check(lib().tcs_problem_set_tile(ctypes.byref(self.s), tile_h, tile_w))

    def apply_tuning(self, path: str = "", batched: bool = False) -> None:
        """Tile, wg_per_cu and sweeps per launch from a tuning cache file ("" or
        None: $TCS_TUNE_CACHE); unchanged when it has no entry for this problem."""
        check(lib().tcs_problem_apply_tuning(ctypes.byref(self.s), (path or "").encode(), int(bool(batched))))

    def set_tile_size(self, tile_size: int) -> None:
        """Apply the reference's TILE_SIZE (bricks per block along x, 1..15)."""
        check(lib().tcs_problem_set_tile_size(ctypes.byref(self.s), int(tile_size)))

    def check_buf(self, buf, what: str = "storage buffer") -> None:
        """A storage buffer of this problem must hold the whole layout (which
        set_tile and the fragment choice re-derive): the kernels address all of it."""
        if buf is None:
            raise ValueError(f"{what}: missing")
        nbytes = buf.numel() * buf.element_size()
        if nbytes < self.buffer_bytes:
            raise ValueError(f"{what}: {nbytes} bytes, this problem's layout needs {self.buffer_bytes}")

    def buf_ptr(self, buf, what: str = "storage buffer") -> ctypes.c_void_p:
        """Device pointer of a checked storage buffer of this problem."""
        self.check_buf(buf, what)
        return _ptr(buf)

    @property
    def ref(self):
        return ctypes.byref(self.s)

    @property
    def rows(self) -> int:
        return self.s.layout.rows

    @property
    def cols(self) -> int:
        return self.s.layout.cols

    @property
    def radius(self) -> int:
        return self.s.radius

    @property
    def wg_per_cu(self) -> int:
        return self.s.wg_per_cu

    @wg_per_cu.setter
    def wg_per_cu(self, n: int) -> None:
        self.s.wg_per_cu = int(n)

    @property
    def fuse(self) -> int:
        return self.s.fuse

    @fuse.setter
    def fuse(self, k: int) -> None:
        self.s.fuse = int(k)

    @property
    def coef_lo(self) -> bool:
        return bool(self.s.coef_lo)

    def set_coef_lo(self, on: bool) -> None:
        """Hi+lo coefficients on or off; re-derives the default tile."""
        check(lib().tcs_problem_set_coef_lo(ctypes.byref(self.s), int(bool(on))))

    @property
    def fused_kernel_name(self) -> str:
        """The kernel of a launch of `fuse` fused sweeps ("" if none): its tile
        and workgroup size may differ from the problem's (large FP16/BF16 4x8
        problems fuse on 8x8 tiles of 8 waves)."""
        n = lib().tcs_fused_kernel_name(self.ref)
        return n.decode() if n else ""

    def edge_kernel_name(self, mode: int = 3) -> str:
        """The kernel run_f32 uses for a sweep with the layout transforms fused
        in (mode 1 first, 2 last, 3 the only sweep), "" when it runs three steps."""
        n = lib().tcs_edge_kernel_name(self.ref, mode)
        return n.decode() if n else ""

    @property
    def tile(self):
        return (self.s.tile_h, self.s.tile_w)

    @property
    def buffer_bytes(self) -> int:
        return int(lib().tcs_buffer_bytes(self.ref))

    @property
    def kernel_name(self) -> str:
        n = lib().tcs_kernel_name(self.ref)
        return n.decode() if n else ""

    @property
    def frag(self) -> int:
        return self.s.frag

    @property
    def mfma_per_brick(self) -> int:
        return int(lib().tcs_mfma_per_brick(self.ref))

    @property
    def mfma_flops_per_brick(self) -> int:
        return int(lib().tcs_mfma_flops_per_brick(self.ref))

    @property
    def interior_points(self) -> int:
        r = self.radius
        return max(self.rows - 2 * r, 0) * max(self.cols - 2 * r, 0)

    # -- device buffers -----------------------------------------------------
    def alloc(self, device="cuda"):
        torch = _torch()
        return torch.empty(self.buffer_bytes, dtype=torch.uint8, device=device)

    def clear(self, buf, stream=None) -> None:
        check(lib().tcs_clear(self.ref, self.buf_ptr(buf), _stream(stream)))

    def pack(self, src, buf, scale_exp: int = 0, stream=None) -> None:
        """Row-major FP32/FP64 device grid -> brick storage."""
        torch = _torch()
        sp = _grid_ptr(src, self.rows, self.cols, "pack source")
        bp = self.buf_ptr(buf, "pack destination")
        fn = {torch.float32: lib().tcs_pack_f32, torch.float64: lib().tcs_pack_f64}.get(src.dtype)
        if fn is None:
            raise TypeError("pack: source must be float32 or float64")
        check(fn(self.ref, sp, src.stride(0), bp, scale_exp, _stream(stream)))

    def unpack(self, buf, out=None, scale_exp: int = 0, ring=None, stream=None):
        """Brick storage -> row-major FP32 device grid (ring copied from `ring`)."""
        torch = _torch()
        bp = self.buf_ptr(buf, "unpack source")
        if out is None:
            out = torch.empty((self.rows, self.cols), dtype=torch.float32, device=buf.device)
        if out.dtype != torch.float32:
            raise TypeError("unpack: output must be float32")
        op = _grid_ptr(out, self.rows, self.cols, "unpack output")
        ringp = None
        if ring is not None:
            ringp = _grid_ptr(ring, self.rows, self.cols, "unpack ring")
            if ring.stride(0) != out.stride(0) or ring.dtype != torch.float32:
                raise ValueError("unpack: ring must be float32 with the output's row stride")
        check(lib().tcs_unpack_f32(self.ref, bp, op, out.stride(0), scale_exp, ringp, _stream(stream)))
        return out

    # -- sweeps -------------------------------------------------------------
    def sweep(self, src, dst, stream=None) -> None:
        self.check_buf(src, "sweep input")
        self.check_buf(dst, "sweep output")
        check(lib().tcs_sweep(self.ref, _ptr(src), _ptr(dst), _stream(stream)))

    def run(self, buf0, buf1, sweeps: int, stream=None):
        """Ping-pong sweeps; returns the buffer holding the result."""
        self.check_buf(buf0, "buf0")
        self.check_buf(buf1, "buf1")
        check(lib().tcs_run(self.ref, _ptr(buf0), _ptr(buf1), sweeps, _stream(stream)))
        return buf1 if sweeps % 2 else buf0

    def run_f32(self, src, sweeps: int, out=None, buf0=None, buf1=None, scale_exp: int = 0, stream=None):
        """Row-major FP32 device grid in -> `sweeps` sweeps -> row-major FP32 out
        (tcs_run_f32: pack and unpack fused into the first and last sweep where
        the problem allows).  buf0/buf1: cleared storage buffers, needed unless
        a fused single sweep runs."""
        torch = _torch()
        if src.dim() != 2 or src.shape[0] != self.rows or src.shape[1] != self.cols or src.dtype != torch.float32:
            raise ValueError(f"run_f32: source must be a {self.rows}x{self.cols} float32 grid")
        if out is None:
            out = torch.empty_like(src)
        if out.stride(0) != src.stride(0) or out.stride(1) != 1 or src.stride(1) != 1:
            raise ValueError("run_f32: source and output need the same row stride and unit column stride")
        for b, name in ((buf0, "buf0"), (buf1, "buf1")):
            if b is not None:
                self.check_buf(b, name)
        b0 = _ptr(buf0) if buf0 is not None else None
        b1 = _ptr(buf1) if buf1 is not None else None
        sp = _grid_ptr(src, self.rows, self.cols, "run_f32 source")
        op = _grid_ptr(out, self.rows, self.cols, "run_f32 output")
        check(lib().tcs_run_f32(self.ref, sp, op, src.stride(0), b0, b1, sweeps, scale_exp,
                                _stream(stream)))
        return out

    def graph(self, buf0, buf1, sweeps: int, stream=None) -> "Graph":
        h = ctypes.c_void_p()
        self.check_buf(buf0, "buf0")
        self.check_buf(buf1, "buf1")
        check(lib().tcs_graph_create(self.ref, _ptr(buf0), _ptr(buf1), sweeps, _stream(stream), ctypes.byref(h)))
        return Graph(h)


class Graph:
    """A captured sweep loop (hipGraphExec)."""

    def __init__(self, handle: ctypes.c_void_p):
        self.h = handle

    def launch(self, stream=None) -> None:
        check(lib().tcs_graph_launch(self.h, _stream(stream)))

    def destroy(self) -> None:
        if self.h:
            check(lib().tcs_graph_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Batch:
    """Batched uneven grids (tcs_batch): problems[i] swept between bufs0[i] and
    bufs1[i] (device tensors from problems[i].alloc())."""

    def __init__(self, problems, bufs0, bufs1):
        n = len(problems)
        if not (n == len(bufs0) == len(bufs1)) or n == 0:
            raise ValueError("batch: need one problem and two buffers per grid")
        arr = (L.ProblemStruct * n)(*[p.s for p in problems])
        for i, (p, b0, b1) in enumerate(zip(problems, bufs0, bufs1)):
            p.check_buf(b0, f"grid {i} buf0")
            p.check_buf(b1, f"grid {i} buf1")
        p0 = (ctypes.c_void_p * n)(*[_ptr(b).value for b in bufs0])
        p1 = (ctypes.c_void_p * n)(*[_ptr(b).value for b in bufs1])
        self.h = ctypes.c_void_p()
        check(lib().tcs_batch_create(ctypes.byref(self.h), n, arr, p0, p1))
        self._keep = (problems, bufs0, bufs1)

    @property
    def interior_points(self) -> int:
        return int(lib().tcs_batch_interior_points(self.h))

    @property
    def launches_per_sweep(self) -> int:
        return int(lib().tcs_batch_launches_per_sweep(self.h))

    def run(self, sweeps: int, stream=None) -> None:
        check(lib().tcs_batch_run(self.h, sweeps, _stream(stream)))

    def graph(self, sweeps: int, stream=None) -> Graph:
        h = ctypes.c_void_p()
        check(lib().tcs_batch_graph_create(self.h, sweeps, _stream(stream), ctypes.byref(h)))
        return Graph(h)

    def destroy(self) -> None:
        if self.h:
            check(lib().tcs_batch_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


# ---------------------------------------------------------------- utilities
def _same_grid(a, b, dtype_a, dtype_b, what: str) -> None:
    """Both tensors are contiguous 2-D grids of one shape with the given dtypes
    (the C functions address rows x cols elements of each)."""
    if a.dim() != 2 or tuple(a.shape) != tuple(b.shape):
        raise ValueError(f"{what}: grids of shapes {tuple(a.shape)} and {tuple(b.shape)}")
    if a.dtype != dtype_a or b.dtype != dtype_b:
        raise TypeError(f"{what}: expected {dtype_a} and {dtype_b}, got {a.dtype} and {b.dtype}")


def oracle_sweep_f64(stencil, src, B, dst, stream=None) -> None:
    """One FP64 sweep (gold's order, no FMA) on row-major device grids."""
    torch = _torch()
    _same_grid(src, dst, torch.float64, torch.float64, "oracle_sweep_f64")
    B = np.ascontiguousarray(B, dtype=np.float64)
    Bp = B.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    check(lib().tcs_oracle_sweep_f64(stencil_id(stencil), _ptr(src), Bp, B.shape[0], _ptr(dst),
                                     src.shape[0], src.shape[1], _stream(stream)))


def oracle_sweeps_f64(stencil, x64, B, sweeps: int, stream=None):
    """`sweeps` FP64 sweeps with ping-pong; returns the result tensor."""
    a = x64.clone()
    b = x64.clone()  # the ring keeps the input values in both buffers
    for _ in range(sweeps):
        oracle_sweep_f64(stencil, a, B, b, stream)
        a, b = b, a
    return a


def error_norms(out32, ref64, r: int, stream=None) -> dict:
    _same_grid(out32, ref64, _torch().float32, _torch().float64, "error_norms")
    res = np.zeros(4, dtype=np.float64)
    check(lib().tcs_error_norms(_ptr(out32), _ptr(ref64), out32.shape[0], out32.shape[1], r,
                                res.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), _stream(stream)))
    return {"max_abs": float(res[0]), "rel_l2": float(res[1]), "rms_abs": float(res[2]),
            "ref_l2": float(res[3])}


def copy(dst, src, stream=None) -> None:
    """Streaming device copy of src's bytes into dst (the bandwidth ceiling kernel)."""
    nbytes = src.numel() * src.element_size()
    if dst.numel() * dst.element_size() < nbytes:
        raise ValueError("copy: destination too small")
    check(lib().tcs_copy(_ptr(dst), _ptr(src), nbytes, _stream(stream)))


def maxabs(x, stream=None) -> float:
    torch = _torch()
    d = torch.zeros(1, dtype=torch.float64, device=x.device)
    fn = lib().tcs_maxabs_f64 if x.dtype == torch.float64 else lib().tcs_maxabs_f32
    check(fn(_ptr(x), x.shape[0], x.shape[1], x.stride(0), _ptr(d), _stream(stream)))
    return float(d.item())


def split_scale_exp(mx: float, problem: "Problem | None" = None, sweeps: int = 1) -> int:
    """SPLIT's per-grid scale exponent for max|x| = mx; with a problem, also
    for the growth its taps allow over `sweeps` sweeps (tcs_split_scale_exp_for)."""
    if problem is None:
        return int(lib().tcs_split_scale_exp(mx))
    return int(lib().tcs_split_scale_exp_for(problem.ref, mx, sweeps))


def scale_exp(problem: "Problem", mx: float, sweeps: int = 1) -> int:
    """The per-grid scale exponent of a run of `sweeps` sweeps of `problem` in
    its precision, given max|x| = mx (tcs_scale_exp_for): FP16 brings
    mx * growth into [2^14, 2^15), SPLIT as split_scale_exp, BF16 0.  Pass it
    to pack / unpack / run_f32."""
    return int(lib().tcs_scale_exp_for(problem.ref, mx, sweeps))


def gen_uniform(rows: int, cols: int, seed: int, stream_id: int = 0, lo=-1.0, hi=1.0,
                dtype=None, device="cuda", stream=None):
    torch = _torch()
    dtype = dtype or torch.float32
    out = torch.empty((rows, cols), dtype=dtype, device=device)
    fn = lib().tcs_gen_uniform_f64 if dtype == torch.float64 else lib().tcs_gen_uniform_f32
    check(fn(_ptr(out), rows, cols, seed, stream_id, lo, hi, _stream(stream)))
    return out


def gen_sinusoid(rows: int, cols: int, seed: int, noise: float = 1e-3, device="cuda", stream=None):
    torch = _torch()
    out = torch.empty((rows, cols), dtype=torch.float64, device=device)
    check(lib().tcs_gen_sinusoid_f64(_ptr(out), rows, cols, seed, noise, _stream(stream)))
    return out


def round_to(x, precision="fp16", stream=None):
    """Round a float32/float64 device grid in place to FP16/BF16 values."""
    torch = _torch()
    p = L.PREC_ID[precision] if isinstance(precision, str) else int(precision)
    fn = lib().tcs_round_f64 if x.dtype == torch.float64 else lib().tcs_round_f32
    check(fn(_ptr(x), x.numel(), p, _stream(stream)))
    return x


def gen_coeffs(stencil, seed: int) -> np.ndarray:
    n = 2 * radius(stencil) + 1
    B = np.zeros((n, n), dtype=np.float32)
    check(lib().tcs_gen_coeffs(stencil_id(stencil), seed, _fp(B), n))
    return B


def params(**overrides) -> L.Params:
    """The param.hpp parameters (defaults, then TCS_* environment, then overrides)."""
    p = L.Params()
    lib().tcs_params_from_env(ctypes.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def host_sweep(stencil, A: np.ndarray, B: np.ndarray, sweeps: int = 1, method="mfma",
               precision="fp16", prm: L.Params | None = None):
    """Status-returning host driver (tcs_host_sweep): host FP32 in, host FP32 out.
    Returns (C, ms_per_sweep)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    C = np.empty_like(A)
    ms = ctypes.c_float(0.0)
    m = L.METHOD_ID[method] if isinstance(method, str) else int(method)
    p = L.PREC_ID[precision] if isinstance(precision, str) else int(precision)
    prm = prm if prm is not None else params(RUN_TIMES=1, SKIP_TIMES=0)
    check(lib().tcs_host_sweep(stencil_id(stencil), m, p, _fp(A), _fp(B), B.shape[0], _fp(C),
                               A.shape[0], A.shape[1], sweeps, ctypes.byref(prm), ctypes.byref(ms)))
    return C, ms.value
