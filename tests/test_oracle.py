"""The CPU oracle, pinned before it is trusted.

* bitwise against fixtures written by the reference's own gold_*.cpp
  (tests/golden/make_golden.py) and, when built here, against oracle/_ref;
* the FP64 twin against an independent brute force (scipy.signal.correlate2d)
  and closed forms (constant and linear fields);
* planted mistakes (a transposed or flipped coefficient block, a dropped tap)
  must fail those same checks;
* the error norms and the reference-style RMS norms against closed forms;
* the SplitMix64 finaliser against its published test vector.
"""
import glob
import os

import numpy as np
import pytest
from scipy.signal import correlate2d

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
STENCILS = ("star2d1r", "star2d2r", "box2d1r", "box2d2r")
RADIUS = {"star2d1r": 1, "star2d2r": 2, "box2d1r": 1, "box2d2r": 2}


def effective_B(stencil, B):
    """The taps gold actually reads: star stencils use only B's centre row/col."""
    B = np.asarray(B, dtype=np.float64)
    if stencil.startswith("star"):
        r = RADIUS[stencil]
        M = np.zeros_like(B)
        M[r, :] = B[r, :]
        M[:, r] = B[:, r]
        return M
    return B


def brute_force(stencil, A, B):
    """Independent FP64 reference: C = A (cross-correlated with B) on the interior,
    ring copied from A (no kernel flip)."""
    r = RADIUS[stencil]
    C = np.array(A, dtype=np.float64, copy=True)
    C[r:-r, r:-r] = correlate2d(np.asarray(A, np.float64), effective_B(stencil, B), mode="valid")
    return C


def test_fixtures_present():
    assert len(FIXTURES) == 12


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gold_bitwise_vs_reference_fixtures(oracle, path):
    stencil = os.path.basename(path).split("_")[0]
    f = np.load(path)
    C = oracle.gold(stencil, f["A"], f["B"])
    assert C.dtype == np.float32
    assert np.array_equal(C.view(np.uint32), f["C"].view(np.uint32))
    # rectangular generalisation on a square grid is the same loop
    assert np.array_equal(oracle.sweep_f32(stencil, f["A"], f["B"]).view(np.uint32), f["C"].view(np.uint32))


@pytest.mark.parametrize("stencil", STENCILS)
def test_gold_vs_compiled_reference(oracle, stencil):
    if not oracle.ref_available(stencil):
        pytest.skip("oracle/_ref not built (needs /root/reference)")
    rng = np.random.default_rng(7)
    n = 2 * RADIUS[stencil] + 1
    for N in (5, 16, 41):
        A = rng.standard_normal((N, N)).astype(np.float32)
        B = rng.standard_normal((n, n)).astype(np.float32)
        assert np.array_equal(oracle.gold(stencil, A, B), oracle.ref_gold(stencil, A, B))


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("shape", [(9, 9), (23, 40), (64, 17)])
def test_f64_vs_brute_force(oracle, stencil, shape):
    rng = np.random.default_rng(hash((stencil, shape)) % 2**32)
    n = 2 * RADIUS[stencil] + 1
    A = rng.standard_normal(shape)
    B = rng.standard_normal((n, n))
    got = oracle.sweep_f64(stencil, A, B)
    want = brute_force(stencil, A, B)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("stencil", STENCILS)
def test_closed_forms(oracle, stencil):
    r = RADIUS[stencil]
    n = 2 * r + 1
    rng = np.random.default_rng(3)
    B = rng.integers(-8, 9, (n, n)).astype(np.float64)
    Be = effective_B(stencil, B)
    # constant field c -> c * sum(B) in the interior, exactly (small integers)
    A = np.full((20, 30), 3.0)
    C = oracle.sweep_f64(stencil, A, B)
    assert np.all(C[r:-r, r:-r] == 3.0 * Be.sum())
    assert np.all(C[:r] == 3.0) and np.all(C[:, :r] == 3.0)
    # linear ramp a*i + b*j: (a*i + b*j)*sum(B) + a*sum(di*B) + b*sum(dj*B)
    a, b = 2.0, -3.0
    i, j = np.mgrid[0:20, 0:30].astype(np.float64)
    A = a * i + b * j
    d = np.arange(n) - r
    want = (a * i + b * j) * Be.sum() + a * (d[:, None] * Be).sum() + b * (d[None, :] * Be).sum()
    C = oracle.sweep_f64(stencil, A, B)
    assert np.array_equal(C[r:-r, r:-r], want[r:-r, r:-r])


@pytest.mark.parametrize("stencil", STENCILS)
def test_planted_mistakes_are_caught(oracle, stencil):
    """The checks above reject plausible oracle mistakes: a transposed or
    flipped coefficient block, or a dropped tap."""
    rng = np.random.default_rng(11)
    n = 2 * RADIUS[stencil] + 1
    A = rng.standard_normal((15, 15))
    B = rng.standard_normal((n, n))
    want = brute_force(stencil, A, B)
    for wrong in (B.T, B[::-1, ::-1]):
        assert not np.allclose(oracle.sweep_f64(stencil, A, wrong), want, rtol=1e-12, atol=1e-12)
    dropped = effective_B(stencil, B).copy()
    dropped[0, RADIUS[stencil]] = 0.0
    assert not np.allclose(oracle.sweep_f64(stencil, A, dropped), want, rtol=1e-12, atol=1e-12)


def test_multisweep_ping_pong(oracle):
    rng = np.random.default_rng(5)
    A = rng.standard_normal((30, 26))
    B = rng.uniform(0, 1, (5, 5))
    x = A.copy()
    for _ in range(3):
        x = brute_force("box2d2r", x, B)
    np.testing.assert_allclose(oracle.sweeps_f64("box2d2r", A, B, 3), x, rtol=1e-12, atol=1e-12)
    assert np.array_equal(oracle.sweeps_f64("box2d2r", A, B, 0), A)


def test_error_norms_closed_form(oracle):
    rng = np.random.default_rng(9)
    ref = rng.standard_normal((40, 50))
    r = 2
    delta = 2.0 ** -10
    out = (ref + delta).astype(np.float32)
    # build an output whose float32 value differs from ref by a known amount
    out64 = out.astype(np.float64)
    d = out64[r:-r, r:-r] - ref[r:-r, r:-r]
    res = oracle.error_norms(out, ref, r)
    assert res["max_abs"] == np.abs(d).max()
    np.testing.assert_allclose(res["rel_l2"], np.sqrt((d ** 2).sum()) / np.sqrt((ref[r:-r, r:-r] ** 2).sum()), rtol=1e-12)
    np.testing.assert_allclose(res["rms_abs"], np.sqrt((d ** 2).mean()), rtol=1e-12)
    # exact closed form: ref on a dyadic grid, out = ref + delta exactly
    ref2 = np.round(ref * 64) / 64
    out2 = (ref2 + delta).astype(np.float32)
    res2 = oracle.error_norms(out2, ref2, r)
    k = (40 - 2 * r) * (50 - 2 * r)
    assert res2["max_abs"] == delta and res2["rms_abs"] == delta
    np.testing.assert_allclose(res2["rel_l2"], delta * np.sqrt(k) / np.sqrt((ref2[r:-r, r:-r] ** 2).sum()), rtol=1e-12)


def test_reference_style_norms(oracle):
    N = 12
    ref = np.full((N, N), 2.0, dtype=np.float32)
    out = ref + np.float32(0.5)
    # checkError2D: RMS |diff| = 0.5; checkRelativeError2D: RMS diff/ref = 0.25
    assert oracle.check_error_2d(N, out, ref, 1, N - 1, 1, N - 1) == 0.5
    assert oracle.check_relative_error_2d(N, out, ref, 1, N - 1, 1, N - 1) == 0.25
    tiny = np.zeros((N, N), dtype=np.float32)
    out = tiny + np.float32(1e-6)
    np.testing.assert_allclose(oracle.check_relative_error_2d(N, out, tiny, 0, N, 0, N), 1.0, rtol=1e-6)


def test_reference_style_norms_subtract_in_single_precision(oracle):
    """common/common.hpp:47,75,105 subtract two floats in float and widen the
    result.  1.0f - 2^-25 is a tie in float and rounds to 1.0f (even mantissa);
    a difference taken in double would be 1 - 2^-25."""
    N = 3
    out = np.ones((N, N), dtype=np.float32)
    ref = np.full((N, N), 2.0 ** -25, dtype=np.float32)
    assert np.float32(1.0) - np.float32(2.0 ** -25) == np.float32(1.0)
    assert oracle.check_error_2d(N, out, ref, 1, 2, 1, 2) == 1.0
    # |ref| < MIN_NUMBER: the difference is divided by MIN_NUMBER (common.hpp:76)
    assert oracle.check_relative_error_2d(N, out, ref, 1, 2, 1, 2) == 1.0 / 1e-6
    out3 = np.ones((N, N, N), dtype=np.float32)
    ref3 = np.full((N, N, N), 2.0 ** -25, dtype=np.float32)
    assert oracle.check_relative_error_3d(N, out3, ref3, 1, 2, 1, 2, 1, 2) == 1.0 / 1e-6
    # a difference that is exact in float is unchanged: 1.5 - 0.25
    assert oracle.check_error_2d(N, out + np.float32(0.5), out - np.float32(0.75), 0, N, 0, N) == 1.25


def test_reference_style_relative_error_3d(oracle):
    """checkRelativeError3D: closed forms, the MIN_NUMBER floor, and agreement
    with the 2D check on a one-slice box (a transposed index would break it)."""
    N = 10
    ref = np.full((N, N, N), 4.0, dtype=np.float32)
    out = ref + np.float32(1.0)
    assert oracle.check_relative_error_3d(N, out, ref, 1, N - 1, 1, N - 1, 1, N - 1) == 0.25
    tiny = np.zeros((N, N, N), dtype=np.float32)
    np.testing.assert_allclose(oracle.check_relative_error_3d(N, tiny + np.float32(2e-6), tiny, 0, N, 0, N, 0, N),
                               2.0, rtol=1e-6)
    rng = np.random.default_rng(3)
    ref = rng.uniform(0.5, 1.5, (N, N, N)).astype(np.float32)
    out = (ref * rng.uniform(0.9, 1.1, (N, N, N))).astype(np.float32)
    z = 6
    got = oracle.check_relative_error_3d(N, out, ref, z, z + 1, 2, 7, 1, 9)
    want = oracle.check_relative_error_2d(N, out[z], ref[z], 2, 7, 1, 9)
    assert got == want
    # only the chosen box counts: damage outside it changes nothing
    out2 = out.copy()
    out2[0] = 100.0
    assert oracle.check_relative_error_3d(N, out2, ref, z, z + 1, 2, 7, 1, 9) == got
    sub = (out[2:5, 3:8, 1:4] - ref[2:5, 3:8, 1:4]).astype(np.float64) / ref[2:5, 3:8, 1:4]
    np.testing.assert_allclose(oracle.check_relative_error_3d(N, out, ref, 2, 5, 3, 8, 1, 4),
                               np.sqrt((sub ** 2).mean()), rtol=1e-12)


def test_splitmix64_published_vector(oracle):
    # First outputs of SplitMix64 seeded with 1234567 (Vigna's reference C).
    assert oracle.mix64(1234567) == 6457827717110365317
    assert oracle.mix64(1234567 + 0x9E3779B97F4A7C15) == 3203168211198807973


def test_uniform_generator_range(oracle):
    x = oracle.gen_uniform_f32(64, 48, seed=1, stream=0, lo=-1.0, hi=1.0)
    assert x.min() >= -1.0 and x.max() < 1.0
    assert abs(float(x.mean())) < 0.05
    # values are k * 2^-23 - 1 exactly
    k = (x.astype(np.float64) + 1.0) * 2.0 ** 23
    assert np.array_equal(k, np.round(k))
    assert not np.array_equal(x, oracle.gen_uniform_f32(64, 48, seed=2, stream=0, lo=-1.0, hi=1.0))


@pytest.mark.parametrize("stencil", ["star2d1r", "star2d2r", "box2d1r", "box2d2r"])
def test_threaded_bands_equal_one_sweep(oracle, stencil):
    """The row-band threaded sweep (the multi-core CPU baseline) is bitwise the
    single-threaded oracle sweep, for band counts that do and do not divide the
    interior (and more bands than rows)."""
    rng = np.random.default_rng(11)
    A = rng.uniform(-1, 1, (97, 61)).astype(np.float32)
    B = rng.uniform(0, 1, (5, 5) if stencil.endswith("2r") else (3, 3)).astype(np.float32)
    want = oracle.sweep_f32(stencil, A, B)
    for threads in (1, 3, 7, 200):
        C = A.copy()
        oracle.sweep_f32_threads(stencil, A, B, C, threads)
        assert np.array_equal(C, want), threads
