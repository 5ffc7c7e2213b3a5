"""Error model of the product sweeps, for the parity tests (test infrastructure).

Two tools, both built from the oracle's FP64 sweep and a rounding step:

1. ``model_sweeps`` -- the kernels' arithmetic with its roundings made explicit:
   FP64 sweeps with the coefficients the kernel multiplies (``kernel_coeffs``)
   and every sweep's result rounded to the storage format.  The kernel differs
   from it only by the order and precision of the FP32 accumulation, which
   moves an FP16/BF16 result by one unit in the last place where the exact
   value lies within that accumulation error of a rounding boundary.  A
   systematic error of the kernel (a biased coefficient, a dropped or doubled
   term, a 1% scale error) is many times larger than the tolerance this leaves.

2. ``error_bounds`` -- rigorous worst-case bounds on max-abs and L2 error of k
   sweeps against exact arithmetic (the FP64 oracle), from the recursion

      e_j = [rnd(v_j) - v_j] + [v_j - sum c^ x] + sum (c^ - c) x + sum c e_{j-1}

   with v_j the FP32 accumulation of sum c^ x_{j-1}:
      |rnd(v) - v|         <= u |v| + s        (u: unit roundoff, s: half the
                                                smallest subnormal spacing)
      |v - sum c^ x|       <= g sum|c^| |x|    (g = n 2^-24 / (1 - n 2^-24),
                                                n products, any summation order)
      |sum (c^ - c) x|     <= d |x|            (d = sum |c^ - c|)
      |sum c e|            <= sum|c| |e|
   in max-norm, and in L2 with Young's inequality (a stencil of taps w has L2
   operator norm <= sum |w|).  These hold for any input and any order of
   accumulation, hence are loose for random data; the model comparison above is
   the tight check.
"""
from __future__ import annotations

import numpy as np

U = {"fp16": 2.0 ** -11, "bf16": 2.0 ** -8}
# Half the spacing of the smallest subnormals (absolute rounding error there).
SUBNORMAL_HALF = {"fp16": 2.0 ** -25, "bf16": 2.0 ** -134}
N_PRODUCTS = {"star2d1r": 5, "star2d2r": 9, "box2d1r": 9, "box2d2r": 25}


def round_to(x, precision: str) -> np.ndarray:
    """x (FP64) rounded to the nearest FP16 or BF16 value, returned as FP64."""
    x = np.asarray(x, dtype=np.float64)
    if precision == "fp16":
        return x.astype(np.float16).astype(np.float64)
    import torch
    return torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()


def kernel_coeffs(B, precision: str, coef_lo: bool = False) -> np.ndarray:
    """The taps the MFMA kernel multiplies for FP32 coefficients B: rounded to
    the storage format, or with coef_lo the exact sum hi + lo of that rounding
    and of the rounding of the remainder (both FP16/BF16)."""
    B = np.asarray(B, dtype=np.float32).astype(np.float64)
    hi = round_to(B, precision)
    if not coef_lo:
        return hi
    return hi + round_to(B - hi, precision)


def model_sweeps(oracle, stencil: str, A, ceff, sweeps: int, precision: str) -> np.ndarray:
    """FP64 sweeps of A with taps ceff, each result rounded to the storage
    format (the ring keeps A, as in the kernels)."""
    x = np.asarray(A, dtype=np.float64)
    for _ in range(sweeps):
        x = round_to(oracle.sweep_f64(stencil, x, ceff), precision)
    return x


def error_bounds(stencil: str, B, ceff, precision: str, sweeps: int, maxabs0: float, l2_0: float,
                 npoints: int):
    """(max-abs bound, L2 bound) of the error after `sweeps` sweeps against
    exact arithmetic, for inputs exactly representable in the storage format
    with max|x| = maxabs0 and ||x||_2 = l2_0 over the whole grid (npoints
    points).  B: the exact FP32 taps; ceff: the taps the kernel multiplies."""
    B = np.asarray(B, dtype=np.float32).astype(np.float64)
    u, s = U[precision], SUBNORMAL_HALF[precision]
    n = N_PRODUCTS[stencil]
    g = n * 2.0 ** -24 / (1 - n * 2.0 ** -24)
    sc, sch = np.abs(B).sum(), np.abs(ceff).sum()
    d = np.abs(ceff - B).sum()
    grow = max(1.0, sch * (1 + g) * (1 + u))  # the ring keeps its values: norms never below the input's
    e_inf = e_2 = 0.0
    m_inf, m_2 = maxabs0, l2_0
    for _ in range(sweeps):
        step_inf = u * sch * (1 + g) * m_inf + g * sch * m_inf + d * m_inf + s
        step_2 = u * sch * (1 + g) * m_2 + g * sch * m_2 + d * m_2 + s * np.sqrt(npoints)
        e_inf = sc * e_inf + step_inf
        e_2 = sc * e_2 + step_2
        m_inf, m_2 = grow * m_inf + s, grow * m_2 + s * np.sqrt(npoints)
    return e_inf, e_2


def rel_l2(a, b, r: int) -> float:
    """||a - b||_2 / ||b||_2 over the interior [r, rows-r) x [r, cols-r)."""
    a = np.asarray(a, dtype=np.float64)[r:-r, r:-r]
    b = np.asarray(b, dtype=np.float64)[r:-r, r:-r]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# Fraction of points per sweep whose FP16/BF16 result may move by one unit in
# the last place between two FP32 accumulation orders of the same products:
# a CPU emulation summing the taps in reverse order in FP32 moved 0.04-0.06% of
# the points per sweep (FP16, box2d2r/box2d1r/star2d2r, 1-30 sweeps); this is 4x that.
FLIP_RATE = 0.002


def model_tolerance(precision: str, sweeps: int) -> float:
    """Largest rel-L2 between a kernel and model_sweeps from accumulation-order
    flips alone: a fraction FLIP_RATE * sweeps of the points (at most all)
    off by at most one unit in the last place, 2u relative, each."""
    return 2.0 * U[precision] * np.sqrt(min(1.0, FLIP_RATE * sweeps))


def gpu_model_sweeps(T, stencil: str, x64, ceff, precision: str, counts):
    """model_sweeps on the GPU (the library's FP64 oracle sweep with taps ceff,
    then rounding to the storage format) for row-major FP64 device grids;
    returns {count: result} for each count in `counts`."""
    a, b = x64.clone(), x64.clone()
    out, done = {}, 0
    for k in sorted(counts):
        while done < k:
            T.oracle_sweep_f64(stencil, a, ceff, b)
            T.round_to(b, precision)
            a, b = b, a
            done += 1
        out[k] = a.clone()
    return out


def gpu_oracle_sweeps(T, stencil: str, x64, B, counts):
    """The FP64 oracle after each sweep count in `counts` (device grids)."""
    a, b = x64.clone(), x64.clone()
    out, done = {}, 0
    for k in sorted(counts):
        while done < k:
            T.oracle_sweep_f64(stencil, a, np.asarray(B, np.float64), b)
            a, b = b, a
            done += 1
        out[k] = a.clone()
    return out
