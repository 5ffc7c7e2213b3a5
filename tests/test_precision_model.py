"""The error model used by the GPU parity tests, checked on the CPU.

A kernel is emulated here in FP32 (taps summed in reverse order, one FP32
rounding per product-add, then the storage rounding): it must stay within
model_tolerance of model_sweeps and within the rigorous error_bounds of the
FP64 oracle, while planted mistakes (a 1% coefficient bias, a transposed tap
block, coefficients truncated instead of rounded) must fall outside the model
tolerance."""
import numpy as np
import pytest

import precision_model as PM

RADIUS = {"star2d1r": 1, "star2d2r": 2, "box2d1r": 1, "box2d2r": 2}


def coeffs(stencil, rng):
    r = RADIUS[stencil]
    n = 2 * r + 1
    B = rng.uniform(0, 1, (n, n))
    if stencil.startswith("star"):
        m = np.zeros((n, n), bool)
        m[r, :] = m[:, r] = True
        B = B * m
    return (B / B.sum()).astype(np.float32)


def emulate(stencil, A, ceff, sweeps, precision):
    """FP32 accumulation of the taps in reverse row-major order, rounded to the
    storage format after every sweep; ring kept."""
    r = RADIUS[stencil]
    n = 2 * r + 1
    c = np.asarray(ceff, dtype=np.float32)
    x = np.asarray(A, dtype=np.float64)
    rows, cols = x.shape
    taps = [(i, j) for i in range(n) for j in range(n) if c[i, j] != 0][::-1]
    for _ in range(sweeps):
        x32 = x.astype(np.float32)
        acc = np.zeros((rows - 2 * r, cols - 2 * r), np.float32)
        for i, j in taps:
            acc = acc + c[i, j] * x32[i:i + rows - 2 * r, j:j + cols - 2 * r]
        y = x.copy()
        y[r:-r, r:-r] = PM.round_to(acc.astype(np.float64), precision)
        x = y
    return x


@pytest.mark.parametrize("stencil", ["box2d2r", "star2d1r"])
@pytest.mark.parametrize("precision", ["fp16", "bf16"])
@pytest.mark.parametrize("coef_lo", [False, True])
def test_emulated_kernel_within_model_tolerance_and_bounds(oracle, stencil, precision, coef_lo):
    rng = np.random.default_rng(5)
    A = PM.round_to(rng.uniform(-1, 1, (96, 160)), precision)
    B = coeffs(stencil, rng)
    ceff = PM.kernel_coeffs(B, precision, coef_lo)
    r = RADIUS[stencil]
    for k in (1, 6):
        model = PM.model_sweeps(oracle, stencil, A, ceff, k, precision)
        emu = emulate(stencil, A, ceff, k, precision)
        assert PM.rel_l2(emu, model, r) <= PM.model_tolerance(precision, k)
        exact = oracle.sweeps_f64(stencil, A, B.astype(np.float64), k)
        b_inf, b_2 = PM.error_bounds(stencil, B, ceff, precision, k, np.abs(A).max(), np.linalg.norm(A), A.size)
        for got in (emu, model):
            assert np.abs(got - exact).max() <= b_inf
            assert np.linalg.norm((got - exact)[r:-r, r:-r]) <= b_2


@pytest.mark.parametrize("stencil", ["box2d2r", "box2d1r"])
def test_planted_mistakes_exceed_model_tolerance(oracle, stencil):
    rng = np.random.default_rng(6)
    A = PM.round_to(rng.uniform(-1, 1, (96, 160)), "fp16")
    B = coeffs(stencil, rng)
    ceff = PM.kernel_coeffs(B, "fp16")
    r = RADIUS[stencil]
    k = 6
    model = PM.model_sweeps(oracle, stencil, A, ceff, k, "fp16")
    tol = PM.model_tolerance("fp16", k)
    # a 1% scale error of the taps
    assert PM.rel_l2(PM.model_sweeps(oracle, stencil, A, ceff * 1.01, k, "fp16"), model, r) > 10 * tol
    # the tap block transposed (rows and columns of the stencil swapped)
    assert PM.rel_l2(PM.model_sweeps(oracle, stencil, A, ceff.T.copy(), k, "fp16"), model, r) > 10 * tol
    # coefficients truncated toward zero instead of rounded to nearest
    trunc = np.trunc(B.astype(np.float64) * 2.0 ** 14) / 2.0 ** 14
    assert PM.rel_l2(PM.model_sweeps(oracle, stencil, A, trunc, k, "fp16"), model, r) > tol
    # and the hi+lo taps are a different kernel, beyond the tolerance of the plain model
    lo = PM.kernel_coeffs(B, "fp16", coef_lo=True)
    assert PM.rel_l2(PM.model_sweeps(oracle, stencil, A, lo, k, "fp16"), model, r) > tol


def test_kernel_coeffs_hi_lo():
    rng = np.random.default_rng(7)
    B = rng.uniform(0, 1, (5, 5)).astype(np.float32)
    hi = PM.kernel_coeffs(B, "fp16")
    assert np.array_equal(hi, B.astype(np.float16).astype(np.float64))
    both = PM.kernel_coeffs(B, "fp16", coef_lo=True)
    # |c - (hi + lo)| <= half a unit of lo's last place: 2^-22 |c|, or 2^-25 where lo is subnormal
    assert np.all(np.abs(both - B.astype(np.float64)) <= np.maximum(np.abs(B) * 2.0 ** -22, 2.0 ** -25))
    bb = PM.kernel_coeffs(B, "bf16", coef_lo=True)
    assert np.all(np.abs(bb - B.astype(np.float64)) <= np.abs(B) * 2.0 ** -16)
