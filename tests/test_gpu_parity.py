"""GPU parity of the HIP kernels against the oracle (run on an MI355X).

Bars (the tolerance is written at each test):
  * dyadic inputs: FP16/BF16/SPLIT MFMA and direct-FMA outputs are bitwise
    equal to the FP64 oracle rounded to the storage format (every product and
    partial sum is exact, so any index, lane-layout or halo error shows);
  * random inputs: max-abs within the derived rounding bound and rel-L2
    within the stated targets (FP16 <= 2e-3 per sweep, SPLIT <= 1e-6);
  * the GPU FP64 oracle is bitwise equal to the CPU FP64 oracle;
  * full size (16384^2): rel-L2 vs the GPU FP64 oracle.
"""
import subprocess
import sys

import numpy as np
import pytest

import precision_model as PM

pytestmark = pytest.mark.gpu

STENCILS = ("star2d1r", "star2d2r", "box2d1r", "box2d2r")
RADIUS = {"star2d1r": 1, "star2d2r": 2, "box2d1r": 1, "box2d2r": 2}
MFMA_TILES = [(2, 4), (4, 4), (2, 8), (4, 8), (8, 8)]


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    assert _t.cuda.is_available(), "GPU tests need a GPU"
    return _t


@pytest.fixture(scope="module")
def T():
    import tcstencil_amd
    return tcstencil_amd


def say(*a):
    print(*a, flush=True)


def dyadic(stencil, rows, cols, seed):
    rng = np.random.default_rng(seed)
    r = RADIUS[stencil]
    n = 2 * r + 1
    A = rng.integers(-64, 65, (rows, cols)) / 64.0
    B = rng.integers(0, 33, (n, n)) / 32.0
    if stencil.startswith("star"):
        m = np.zeros((n, n), bool)
        m[r, :] = m[:, r] = True
        B = B * m
    return A, B.astype(np.float32)


def gpu_sweeps(T, torch, stencil, A, B, method="mfma", precision="fp16", tile=None, sweeps=1,
               scale_exp=0, graph=False, frag=16, fuse=1, coef_lo=False):
    rows, cols = A.shape
    p = T.Problem(stencil, rows, cols, B, method=method, precision=precision, tile=tile, frag=frag, fuse=fuse,
                  coef_lo=coef_lo)
    src = torch.from_numpy(np.ascontiguousarray(A)).cuda()
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(src, b0, scale_exp)
    if graph:
        g = p.graph(b0, b1, sweeps)
        g.launch()
        res = b1 if sweeps % 2 else b0
    else:
        res = p.run(b0, b1, sweeps)
    out = p.unpack(res, scale_exp=scale_exp)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


def round_bf16(torch, x):
    return torch.from_numpy(np.asarray(x, np.float64)).to(torch.bfloat16).to(torch.float64).numpy()


# ------------------------------------------------------------ bitwise pins
@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("tile", MFMA_TILES, ids=lambda t: "t%dx%d" % t)
@pytest.mark.parametrize("method", ["mfma", "direct", "direct_dot2"])
def test_dyadic_exact_fp16(T, torch, oracle, stencil, tile, method):
    for rows, cols in ((70, 150), (33, 17), (130, 260)):
        say("case", stencil, method, tile, rows, cols)
        A, B = dyadic(stencil, rows, cols, seed=rows + cols)
        got = gpu_sweeps(T, torch, stencil, A, B, method=method, tile=tile)
        want = oracle.sweep_f64(stencil, A, B).astype(np.float16).astype(np.float64)
        assert np.array_equal(got, want), f"max diff {np.abs(got - want).max()}"


@pytest.mark.parametrize("stencil", STENCILS)
def test_dyadic_exact_bf16(T, torch, oracle, stencil):
    for tile in MFMA_TILES:
        say("case bf16", stencil, tile)
        A, B = dyadic(stencil, 70, 150, seed=3)
        got = gpu_sweeps(T, torch, stencil, A, B, precision="bf16", tile=tile)
        want = round_bf16(torch, oracle.sweep_f64(stencil, A, B))
        assert np.array_equal(got, want)


@pytest.mark.parametrize("stencil", STENCILS)
def test_dyadic_exact_split(T, torch, oracle, stencil):
    for tile in MFMA_TILES[:4]:
        say("case split", stencil, tile)
        A, B = dyadic(stencil, 70, 150, seed=4)
        got = gpu_sweeps(T, torch, stencil, A, B, precision="split", tile=tile, scale_exp=3)
        assert np.array_equal(got, oracle.sweep_f64(stencil, A, B))


def test_dyadic_exact_multisweep(T, torch, oracle):
    # Exactness survives 3 sweeps while values stay on a coarse dyadic grid:
    # coefficients in {0, 1/2} with one tap keep every value a multiple of 2^-9.
    for stencil in STENCILS:
        say("case multisweep", stencil)
        r = RADIUS[stencil]
        n = 2 * r + 1
        A, _ = dyadic(stencil, 96, 160, seed=8)
        B = np.zeros((n, n), np.float32)
        B[r, r] = 0.5
        B[r, 0] = 0.5
        got = gpu_sweeps(T, torch, stencil, A, B, sweeps=3)
        want = oracle.sweeps_f64(stencil, A, B, 3)
        assert np.array_equal(got, want)


# ----------------------------------------------------------- random inputs
def random_problem(T, torch, stencil, rows, cols, seed):
    x = T.gen_uniform(rows, cols, seed=seed, dtype=torch.float64)
    T.round_to(x, "fp16")
    return x.cpu().numpy(), T.gen_coeffs(stencil, seed).astype(np.float64)


def check_against_models(oracle, stencil, A, B32, got, precision, sweeps, coef_lo=False, method="mfma", what=""):
    """A kernel result after `sweeps` sweeps of A, against (1) the rounding
    model of its arithmetic (FP64 sweeps with the taps it multiplies, rounded
    to the storage format after each sweep) within the accumulation-order
    tolerance, and (2) the FP64 oracle within the rigorous worst-case bounds
    (max-abs and L2; precision_model documents both derivations)."""
    r = RADIUS[stencil]
    ceff = B32.astype(np.float64) if method == "direct" else PM.kernel_coeffs(B32, precision, coef_lo)
    model = PM.model_sweeps(oracle, stencil, A, ceff, sweeps, precision)
    exact = oracle.sweeps_f64(stencil, A, B32.astype(np.float64), sweeps)
    vs_model = PM.rel_l2(got, model, r)
    b_inf, b_2 = PM.error_bounds(stencil, B32, ceff, precision, sweeps, np.abs(A).max(), np.linalg.norm(A), A.size)
    err = (got - exact)[r:-r, r:-r]
    say("  ", what, precision, "k=%d" % sweeps, "vs model %.2e (tol %.2e)" % (vs_model, PM.model_tolerance(precision, sweeps)),
        "max-abs %.2e (bound %.2e)" % (np.abs(err).max(), b_inf),
        "rel-L2 %.2e (bound %.2e)" % (np.linalg.norm(err) / np.linalg.norm(exact[r:-r, r:-r]),
                                      b_2 / np.linalg.norm(exact[r:-r, r:-r])))
    assert vs_model <= PM.model_tolerance(precision, sweeps), (what, vs_model)
    assert np.abs(err).max() <= b_inf, what
    assert np.linalg.norm(err) <= b_2, what


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("method", ["mfma", "direct", "direct_dot2"])
def test_random_fp16_within_bound(T, torch, oracle, stencil, method):
    say("case random fp16", stencil, method)
    A, B = random_problem(T, torch, stencil, 300, 517, seed=1)
    B32 = B.astype(np.float32)
    got = gpu_sweeps(T, torch, stencil, A, B32, method=method)
    ref = oracle.sweep_f64(stencil, A, B)
    res = oracle.error_norms(got.astype(np.float32), ref, RADIUS[stencil])
    say("  ", res)
    # |err| <= u16*(sum|c|*max|x|) [coefficient rounding, MFMA only]
    #        + u16*|out| [output rounding] + FP32 accumulation (< 1e-6); u16 = 2^-11
    bound = (2.0 ** -11) * (1.0 if method == "direct" else 2.0) * np.abs(A).max() + 1e-6
    assert res["max_abs"] <= bound
    assert res["rel_l2"] <= 2e-3
    check_against_models(oracle, stencil, A, B32, got, "fp16", 1, method=method, what=method)
    if method == "mfma":
        lo = gpu_sweeps(T, torch, stencil, A, B32, coef_lo=True)
        check_against_models(oracle, stencil, A, B32, lo, "fp16", 1, coef_lo=True, what="hi+lo")
    ring = np.ones_like(A, bool)
    r = RADIUS[stencil]
    ring[r:-r, r:-r] = False
    assert np.array_equal(got[ring], A[ring])


@pytest.mark.parametrize("stencil", STENCILS)
def test_random_bf16_within_bound(T, torch, oracle, stencil):
    say("case random bf16", stencil)
    x = T.gen_uniform(300, 517, seed=2, dtype=torch.float64)
    T.round_to(x, "bf16")
    A = x.cpu().numpy()
    B = T.gen_coeffs(stencil, 2)
    for coef_lo in (False, True):
        for sweeps in (1, 5):
            got = gpu_sweeps(T, torch, stencil, A, B, precision="bf16", sweeps=sweeps, coef_lo=coef_lo)
            check_against_models(oracle, stencil, A, B, got, "bf16", sweeps, coef_lo=coef_lo,
                                 what="hi+lo" if coef_lo else "plain")


@pytest.mark.parametrize("stencil", STENCILS)
def test_random_split_rel_l2(T, torch, oracle, stencil):
    say("case random split", stencil)
    x = T.gen_uniform(300, 517, seed=3, dtype=torch.float32)
    A = x.cpu().numpy().astype(np.float64)
    B = T.gen_coeffs(stencil, 3)
    e = T.split_scale_exp(float(np.abs(A).max()))
    got = gpu_sweeps(T, torch, stencil, A.astype(np.float32), B, precision="split", scale_exp=e)
    res = oracle.error_norms(got.astype(np.float32), oracle.sweep_f64(stencil, A, B.astype(np.float64)), RADIUS[stencil])
    say("  ", res)
    assert res["rel_l2"] <= 1e-6
    assert res["max_abs"] <= 1e-6


@pytest.mark.parametrize("stencil", STENCILS)
def test_ten_sweeps(T, torch, oracle, stencil):
    """Ten sweeps: FP16 MFMA (plain and hi+lo coefficients) and direct-FMA
    against their rounding models and the rigorous bounds; SPLIT against the
    FP64 oracle with the per-sweep target 1e-6 (asserted alone on one sweep by
    test_random_split_rel_l2) accumulated as independent errors: each sweep
    adds at most 1e-6 ||x_j|| in L2, the stencil does not amplify (L2 operator
    norm <= sum|c| = 1) and errors of different sweeps add in quadrature, so
    ||e_k|| <= 1e-6 sqrt(k) max_j ||x_j|| <= 1e-6 sqrt(k) ||x_0||."""
    say("case ten sweeps", stencil)
    A, B = random_problem(T, torch, stencil, 200, 333, seed=4)
    B32 = B.astype(np.float32)
    r = RADIUS[stencil]
    for method, coef_lo in (("mfma", False), ("mfma", True), ("direct", False)):
        got = gpu_sweeps(T, torch, stencil, A, B32, method=method, sweeps=10, coef_lo=coef_lo)
        check_against_models(oracle, stencil, A, B32, got, "fp16", 10, coef_lo=coef_lo, method=method,
                             what=f"{method} hi+lo={coef_lo}")
    ref = oracle.sweeps_f64(stencil, A, B, 10)
    e = T.split_scale_exp(1.0)
    got = gpu_sweeps(T, torch, stencil, A, B32, precision="split", sweeps=10, scale_exp=e)
    err = np.linalg.norm((got - ref)[r:-r, r:-r])
    say("   split rel-L2 %.2e" % (err / np.linalg.norm(ref[r:-r, r:-r])))
    assert err <= 1e-6 * np.sqrt(10) * np.linalg.norm(A)


# ------------------------------------------------------------------ edges
@pytest.mark.parametrize("stencil", STENCILS)
def test_tiny_and_degenerate_grids(T, torch, oracle, stencil):
    for rows, cols in ((1, 1), (2, 3), (3, 3), (4, 5), (5, 5), (6, 40), (17, 1), (16, 16)):
        say("case tiny", stencil, rows, cols)
        A, B = dyadic(stencil, rows, cols, seed=rows * 31 + cols)
        for method in ("mfma", "direct", "direct_dot2"):
            got = gpu_sweeps(T, torch, stencil, A, B, method=method)
            want = oracle.sweep_f64(stencil, A, B).astype(np.float16).astype(np.float64)
            assert np.array_equal(got, want)


def test_reference_driver_inputs(T, torch, oracle):
    """The reference driver's own case: all-ones input, init_B coefficients
    (common/main_2r.cu:13-28); interior = sum(B) exactly."""
    for stencil in STENCILS:
        say("case init_B", stencil)
        B = oracle.init_B(stencil)
        A = np.ones((96, 96))
        for method in ("mfma", "direct"):
            got = gpu_sweeps(T, torch, stencil, A, B, method=method)
            assert np.array_equal(got, oracle.gold(stencil, A.astype(np.float32), B).astype(np.float64))


def test_nan_propagation(T, torch, oracle):
    """A NaN input poisons at least every output gold poisons (the direct kernel
    exactly those; the MFMA form also the brick rows whose window holds it, the
    packed-half form the outputs whose tap pairs hold it)."""
    A, B = dyadic("box2d2r", 64, 64, seed=5)
    A[30, 33] = np.nan
    ref = oracle.sweep_f64("box2d2r", A, B)
    for method in ("direct", "mfma", "direct_dot2"):
        got = gpu_sweeps(T, torch, "box2d2r", A, B, method=method)
        assert np.all(np.isnan(got[np.isnan(ref)]))
        if method == "direct":
            assert np.array_equal(np.isnan(got), np.isnan(ref))


# ------------------------------------------------------- oracle and utils
def test_gpu_fp64_oracle_bitwise(T, torch, oracle):
    rng = np.random.default_rng(6)
    for stencil in STENCILS:
        n = 2 * RADIUS[stencil] + 1
        A = rng.standard_normal((101, 77))
        B = rng.standard_normal((n, n))
        x = torch.from_numpy(A).cuda()
        y = x.clone()
        T.oracle_sweep_f64(stencil, x, B, y)
        assert np.array_equal(y.cpu().numpy(), oracle.sweep_f64(stencil, A, B))


def test_gpu_error_norms(T, torch, oracle):
    rng = np.random.default_rng(7)
    ref = rng.standard_normal((123, 211))
    out = (ref + rng.standard_normal(ref.shape) * 1e-3).astype(np.float32)
    g = T.error_norms(torch.from_numpy(out).cuda(), torch.from_numpy(ref).cuda(), 2)
    c = oracle.error_norms(out, ref, 2)
    assert g["max_abs"] == c["max_abs"]
    for k in ("rel_l2", "rms_abs", "ref_l2"):
        np.testing.assert_allclose(g[k], c[k], rtol=1e-12)


def test_generators_match_oracle(T, torch, oracle):
    g = T.gen_uniform(77, 131, seed=5, stream_id=0, lo=-1.0, hi=1.0).cpu().numpy()
    assert np.array_equal(g, oracle.gen_uniform_f32(77, 131, 5, 0, -1.0, 1.0))
    x = T.gen_sinusoid(64, 96, seed=1)
    assert float(x.abs().max()) < 1.0 + 1e-3
    assert T.maxabs(x) == float(x.abs().max())


@pytest.mark.parametrize("rows,cols,ld", [(45, 77, 77), (45, 88, 88), (33, 77, 96), (64, 64, 64)])
def test_pack_unpack_roundtrip_and_ring(T, torch, rows, cols, ld):
    """Unaligned leading dimensions take the scalar path; aligned ones the
    16-byte loads and stores, with scalar pieces at ragged right edges."""
    base = T.gen_uniform(rows, ld, seed=9, dtype=torch.float32)
    T.round_to(base, "fp16")
    x = base[:, :cols]
    for precision in ("fp16", "split"):
        p = T.Problem("box2d2r", rows, cols, np.ones((5, 5), np.float32), precision=precision)
        b = p.alloc()
        p.pack(x, b)
        full = torch.full((rows, ld), -3.0, device=x.device)
        assert torch.equal(p.unpack(b, out=full[:, :cols]), x)
        assert bool(full[:, cols:].eq(-3.0).all())
    ring = torch.full_like(x, 7.0)
    y = p.unpack(b, ring=ring)
    assert torch.equal(y[2:-2, 2:-2], x[2:-2, 2:-2])
    edge = torch.ones_like(y, dtype=torch.bool)
    edge[2:-2, 2:-2] = False
    assert bool((y[edge] == 7.0).all())


# ------------------------------------------------ graphs and determinism
def test_graph_replays_match_eager(T, torch):
    A, B = dyadic("box2d2r", 256, 384, seed=10)
    p = T.Problem("box2d2r", 256, 384, B)
    src = torch.from_numpy(A).cuda()
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(src, b0)
    eager = p.unpack(p.run(b0, b1, 5)).clone()
    g = p.graph(b0, b1, 5)
    for rep in range(3):
        p.pack(src, b0)           # eager work between replays
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(p.unpack(b1), eager), f"replay {rep}"
    g.destroy()


def test_runs_are_deterministic(T, torch):
    x = T.gen_uniform(500, 700, seed=11, dtype=torch.float32)
    B = T.gen_coeffs("box2d2r", 11)
    p = T.Problem("box2d2r", 500, 700, B)
    outs = []
    for _ in range(2):
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x, b0)
        outs.append(p.unpack(p.run(b0, b1, 4)))
    assert torch.equal(outs[0], outs[1])


# ------------------------------------------------------------ host drivers
@pytest.mark.parametrize("stencil", STENCILS)
def test_host_sweep_vs_gold(T, torch, oracle, stencil):
    say("case host_sweep", stencil)
    rng = np.random.default_rng(12)
    A = rng.uniform(-1, 1, (250, 190)).astype(np.float32)
    B = T.gen_coeffs(stencil, 12)
    gold = oracle.sweep_f32(stencil, A, B)
    r = RADIUS[stencil]
    for method in ("mfma", "direct", "direct_dot2"):
        C, ms = T.host_sweep(stencil, A, B, method=method)
        # input rounding to FP16 + coefficient rounding (MFMA) + output rounding
        assert np.abs(C - gold)[r:-r, r:-r].max() <= 3 * 2.0 ** -11 + 1e-6
        ring = np.ones_like(A, bool)
        ring[r:-r, r:-r] = False
        assert np.array_equal(C[ring], A[ring]) and ms > 0
    C, _ = T.host_sweep(stencil, A, B, precision="split", sweeps=2)
    want = oracle.sweep_f32(stencil, gold, B)
    assert np.abs(C - want)[r:-r, r:-r].max() <= 1e-5


@pytest.mark.parametrize("amp", [3e5, 1e-6])
def test_host_sweep_fp16_per_grid_scaling(T, oracle, amp):
    """FP16 host driver on inputs beyond binary16's range (|x| up to 3e5 >
    65504, Inf without scaling) and far below its normal range (|x| ~ 1e-6,
    subnormal without scaling: 6% relative steps): the per-grid power-of-two
    scale (tcs_scale_exp_for) keeps the relative error that of normal-range
    FP16.  Bound as for the reference fixtures, in units of max|x|:
    u (2 sum|B| max|A| + max|C|) + g sum|B| max|A|, u = 2^-11, g = 2n 2^-24."""
    say("case host_sweep fp16 scaling", amp)
    rng = np.random.default_rng(31)
    stencil = "box2d2r"
    A = (rng.uniform(-1, 1, (200, 180)) * amp).astype(np.float32)
    B = T.gen_coeffs(stencil, 31)
    gold = oracle.sweep_f32(stencil, A, B)
    r = RADIUS[stencil]
    u, g = 2.0 ** -11, 2 * 25 * 2.0 ** -24
    sb, ma, mc = float(np.abs(B).sum()), float(np.abs(A).max()), float(np.abs(gold).max())
    bound = u * (2 * sb * ma + mc) + g * sb * ma
    for method in ("mfma", "direct"):
        C, _ = T.host_sweep(stencil, A, B, method=method)
        assert np.isfinite(C).all()
        err = float(np.abs(C.astype(np.float64) - gold)[r:-r, r:-r].max())
        say(f"  {method}: max |out - gold| {err:.3e} (bound {bound:.3e})")
        assert err <= bound
        ring = np.ones_like(A, bool)
        ring[r:-r, r:-r] = False
        assert np.array_equal(C[ring], A[ring])


@pytest.mark.parametrize("method", ["mfma", "direct"])
@pytest.mark.parametrize("stencil", STENCILS)
def test_dropin_driver_binary(stencil, method):
    """The main_Xr-style C++ driver linked against the host_code and gold shims."""
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "dropin", "_build", f"main_{method}_{stencil}")
    n = 2 * RADIUS[stencil] + 1
    env = dict(os.environ, TCS_RUN_TIMES="20", TCS_SKIP_TIMES="10")
    out = subprocess.run([exe, "200", str(n)], capture_output=True, text=True, timeout=120, env=env)
    say(out.stdout)
    assert out.returncode == 0, out.stderr
    lines = dict(l.split(":", 1) for l in out.stdout.splitlines() if l.startswith("[Test]"))
    assert "[Time] Time used" in out.stdout
    assert int(lines["[Test] Ring mismatches"]) == 0
    max_err = float(lines["[Test] Max Error"].split()[0])
    max_ref = float(lines["[Test] Max Error"].split()[3].rstrip(")"))
    # |A| <= 0.5; init_B is integral (exact in FP16); FP16 rounding of the
    # input (<= u*sum|B|*0.5) and of the output (<= u*|out|), u = 2^-11.
    sum_abs_b = 72.0 if n == 3 else 578.0
    assert max_err <= 2.0 ** -11 * (0.5 * sum_abs_b + max_ref) * 1.01 + 1e-6


# ---------------------------------------------------------------- full size
@pytest.mark.parametrize("method", ["mfma", "direct"])
def test_full_size_box2d2r_16384(T, torch, method):
    """North-star size: box2d2r 16384^2 FP16, one sweep, vs the GPU FP64 oracle
    (itself bitwise-pinned to the CPU oracle above)."""
    say("case full size", method)
    N = 16384
    B = T.gen_coeffs("box2d2r", 1)
    x64 = T.gen_sinusoid(N, N, seed=1, noise=1e-3)
    T.round_to(x64, "fp16")
    ref = torch.empty_like(x64).copy_(x64)
    T.oracle_sweep_f64("box2d2r", x64, B.astype(np.float64), ref)
    p = T.Problem("box2d2r", N, N, B, method=method)
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(x64, b0)
    del x64
    p.sweep(b0, b1)
    out = p.unpack(b1)
    res = T.error_norms(out, ref, 2)
    say("  ", res)
    assert res["rel_l2"] <= 2e-3
    # coefficient + output rounding, max|x| <= 1 + 1e-3 (sum of amplitudes + noise)
    assert res["max_abs"] <= (2.0 ** -11) * 2.0 * 1.001 + 1e-6


@pytest.mark.parametrize("precision", ["fp16", "split", "bf16lo"])
def test_full_size_fused_bitwise_16384(T, torch, precision):
    """bench.py's fused form at its size: box2d2r 16384^2, 8 sweeps per launch
    on the 8-wave kernels (FP16, and BF16 with hi+lo taps on a 4x8 problem: 8x8
    tiles for the 4x8 problem; SPLIT: 4x8 tiles), 16 sweeps bitwise equal to 16
    single-sweep launches."""
    say("case full size fused", precision)
    N = 16384
    B = T.gen_coeffs("box2d2r", 1)
    x = T.gen_sinusoid(N, N, seed=1, noise=1e-3).float()
    e = T.split_scale_exp(1.0) if precision == "split" else 0
    lo = precision.endswith("lo")
    prec = precision[:-2] if lo else precision
    tile = (4, 8) if lo else None  # hi+lo problems default to 8x8 tiles
    pu = T.Problem("box2d2r", N, N, B, precision=prec, tile=tile, coef_lo=lo)
    pf = T.Problem("box2d2r", N, N, B, precision=prec, tile=tile, coef_lo=lo, fuse=8)
    want = {"fp16": "tcs_mfma_fused8<8, 8, 0, 4, 0>", "split": "tcs_mfma_fused8<4, 8, 0, 4, 2>",
            "bf16lo": "tcs_mfma_fused8<8, 8, 0, 4, 5>"}[precision]
    assert pu.tile == (4, 8) and pf.fused_kernel_name == want
    outs = []
    for p in (pu, pf):
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x, b0, e)
        outs.append(p.unpack(p.run(b0, b1, 16), scale_exp=e))
        del b0, b1
    assert torch.equal(outs[0], outs[1])


@pytest.mark.parametrize("stencil", ["box2d2r", "star2d2r"])
def test_full_size_split_16384(T, torch, stencil):
    """The split hi+lo form at the 16384^2 size, with per-grid max-abs scaling:
    rel-L2 vs the GPU FP64 oracle <= 1e-6 after one sweep and after ten, on the
    sinusoid input (FP64, not rounded: the split pair carries ~22 bits)."""
    say("case full size split", stencil)
    N = 16384
    B = T.gen_coeffs(stencil, 1)
    x64 = T.gen_sinusoid(N, N, seed=1, noise=1e-3)
    e = T.split_scale_exp(T.maxabs(x64))
    p = T.Problem(stencil, N, N, B, precision="split")
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(x64, b0, e)
    for sweeps in (1, 10):
        ref = T.oracle_sweeps_f64(stencil, x64, B.astype(np.float64), sweeps)
        out = p.unpack(p.run(b0, b1, sweeps), scale_exp=e)  # b0 keeps the input: run reads it first
        res = T.error_norms(out, ref, 2)
        say("  ", sweeps, res)
        assert res["rel_l2"] <= 1e-6, (sweeps, res)


# ------------------------------------------------------- temporal blocking
FUSED_TILES = [(2, 4), (4, 4), (2, 8), (4, 8), (8, 8)]


def run_buffers(T, torch, p, A, sweeps, scale_exp=0, graph=False):
    src = torch.from_numpy(np.ascontiguousarray(A)).cuda()
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(src, b0, scale_exp)
    if graph:
        g = p.graph(b0, b1, sweeps)
        g.launch()
        torch.cuda.synchronize()
        g.destroy()
        res = b1 if sweeps % 2 else b0
    else:
        res = p.run(b0, b1, sweeps)
    torch.cuda.synchronize()
    return res.clone()


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("precision", ["fp16", "bf16", "split"])
def test_fused_bitwise_equals_unfused(T, torch, stencil, precision):
    """k sweeps fused in LDS give the same bits as k single-sweep launches."""
    r = RADIUS[stencil]
    B = T.gen_coeffs(stencil, 21)
    x = T.gen_uniform(300, 517, seed=21, dtype=torch.float32).cpu().numpy()
    e = T.split_scale_exp(1.0) if precision == "split" else 0
    for tile in FUSED_TILES:
        ref = {}
        for fuse in (2, 3, 1 + 16 // r):
            sweeps = 2 * fuse + 1  # several launches, uneven split, odd parity
            say("case fused", stencil, precision, tile, fuse, sweeps)
            if sweeps not in ref:
                pu = T.Problem(stencil, 300, 517, B, precision=precision, tile=tile)
                ref[sweeps] = run_buffers(T, torch, pu, x, sweeps, e)
                ref[sweeps] = pu.unpack(ref[sweeps], scale_exp=e)
            pf = T.Problem(stencil, 300, 517, B, precision=precision, tile=tile, fuse=fuse)
            # the 8-wave kernels: FP16/BF16 8x8 tiles; SPLIT 4x8 tiles at 6+ sweeps per launch
            # (these runs launch ceil(sweeps / fuse) rounded to odd: fuse 9 and 17 give 6+)
            if tile == (8, 8) and precision != "split" or tile == (4, 8) and precision == "split" and fuse >= 6:
                assert pf.fused_kernel_name.startswith("tcs_mfma_fused8<"), pf.fused_kernel_name
            got = pf.unpack(run_buffers(T, torch, pf, x, sweeps, e), scale_exp=e)
            assert torch.equal(got, ref[sweeps]), f"fuse {fuse}"


@pytest.mark.parametrize("stencil", ["box2d2r", "star2d1r"])
@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_fused_hi_lo_taps_bitwise_equals_unfused(T, torch, stencil, precision):
    """Fused sweeps with hi+lo taps (kernel precision ids 4 and 5, two MFMAs per
    fragment) give the same bits as the unfused hi+lo sweeps, on the tiles their
    fused kernels run (4-wave workgroups; 8-wave ones on 8x8 tiles, the default of
    hi+lo problems)."""
    r = RADIUS[stencil]
    B = T.gen_coeffs(stencil, 23)
    x = T.gen_uniform(300, 517, seed=23, dtype=torch.float32).cpu().numpy()
    for tile in ((2, 4), (4, 8), (8, 8)):
        for fuse in (2, 1 + 16 // r):
            sweeps = 2 * fuse + 1
            say("case fused hi+lo", stencil, precision, tile, fuse, sweeps)
            pu = T.Problem(stencil, 300, 517, B, precision=precision, tile=tile, coef_lo=True)
            pf = T.Problem(stencil, 300, 517, B, precision=precision, tile=tile, coef_lo=True, fuse=fuse)
            kind = "tcs_mfma_fused8<" if tile == (8, 8) else "tcs_mfma_fused<"
            assert pf.coef_lo and pf.fused_kernel_name.startswith(kind), pf.fused_kernel_name
            want = pu.unpack(run_buffers(T, torch, pu, x, sweeps))
            got = pf.unpack(run_buffers(T, torch, pf, x, sweeps))
            assert torch.equal(got, want), (tile, fuse)


@pytest.mark.parametrize("stencil", ["star2d1r", "box2d2r"])
def test_fused_tiny_grids_bitwise_equals_unfused(T, torch, stencil):
    """Grids of one padded tile or less (1x1 to 17x1) through every fused form:
    4-wave tiles, 8x8 tiles on 8 waves (FP16, and with hi+lo taps), SPLIT 4x8
    on 8 waves at 6+ sweeps per launch; bitwise the single sweeps."""
    r = RADIUS[stencil]
    kmax = 1 + 16 // r
    B = T.gen_coeffs(stencil, 29)
    for rows, cols in ((1, 1), (2, 3), (17, 1), (5, 33), (16, 16)):
        x = T.gen_uniform(rows, cols, seed=rows * 37 + cols, dtype=torch.float32).cpu().numpy()
        for precision, lo in (("fp16", False), ("split", False), ("fp16", True)):
            e = T.split_scale_exp(1.0) if precision == "split" else 0
            for tile in ((2, 4), (4, 8), (8, 8)):
                for fuse in (2, kmax):
                    sweeps = 2 * fuse + 1
                    say("case fused tiny", stencil, rows, cols, precision, lo, tile, fuse)
                    kw = dict(precision=precision, tile=tile, coef_lo=lo)
                    pu = T.Problem(stencil, rows, cols, B, **kw)
                    pf = T.Problem(stencil, rows, cols, B, fuse=fuse, **kw)
                    if tile == (8, 8) and precision != "split" or tile == (4, 8) and precision == "split" and fuse >= 6:
                        assert pf.fused_kernel_name.startswith("tcs_mfma_fused8<"), pf.fused_kernel_name
                    want = pu.unpack(run_buffers(T, torch, pu, x, sweeps, e), scale_exp=e)
                    got = pf.unpack(run_buffers(T, torch, pf, x, sweeps, e), scale_exp=e)
                    assert torch.equal(got, want), (rows, cols, precision, lo, tile, fuse)


@pytest.mark.parametrize("precision", ["fp16", "split"])
def test_fused_nan_propagation(T, torch, oracle, precision):
    """Fused launches with NaN inputs: every point that k FP64 oracle sweeps
    poison is NaN, and wherever the fused and the single-sweep results are both
    finite they are bitwise equal.  (The two forms poison different supersets of
    the oracle's set: a NaN enters every output of the brick rows whose window
    holds it, so it spreads up to a brick per sweep across, and a fused launch
    recomputes only its tile's one-brick ring.)"""
    rows, cols = 200, 400
    B = T.gen_coeffs("box2d2r", 31)
    A = T.gen_uniform(rows, cols, seed=31, dtype=torch.float32).cpu().numpy().astype(np.float64)
    A[70, 131] = np.nan
    A[3, 300] = np.nan  # next to the zero ring
    e = T.split_scale_exp(1.0) if precision == "split" else 0
    for tile, fuse in (((4, 8), 4), ((8, 8), 8), ((2, 4), 9)):
        say("case fused nan", precision, tile, fuse)
        ref = A
        for _ in range(fuse):
            ref = oracle.sweep_f64("box2d2r", ref, B)
        uf = gpu_sweeps(T, torch, "box2d2r", A, B, precision=precision, tile=tile, sweeps=fuse, scale_exp=e)
        fu = gpu_sweeps(T, torch, "box2d2r", A, B, precision=precision, tile=tile, sweeps=fuse, scale_exp=e,
                        fuse=fuse)
        poisoned = np.isnan(ref)
        assert poisoned.any() and np.all(np.isnan(uf[poisoned])) and np.all(np.isnan(fu[poisoned]))
        both = np.isfinite(uf) & np.isfinite(fu)
        assert both.sum() > rows * cols // 4, both.sum()
        assert np.array_equal(uf[both], fu[both]), (tile, fuse)


def test_fused_graph_and_accuracy(T, torch, oracle):
    say("case fused graph")
    A, B = random_problem(T, torch, "box2d2r", 200, 333, seed=22)
    p = T.Problem("box2d2r", 200, 333, B.astype(np.float32), fuse=4)
    eager = run_buffers(T, torch, p, A, 10)
    graphed = run_buffers(T, torch, p, A, 10, graph=True)
    assert torch.equal(eager, graphed)
    out = p.unpack(eager).cpu().numpy().astype(np.float64)
    check_against_models(oracle, "box2d2r", A, B.astype(np.float32), out, "fp16", 10, what="fused 4")


@pytest.mark.parametrize("precision", ["fp16", "split"])
def test_fused8_graph_replays_match_eager(T, torch, precision):
    """The 8-wave fused kernels captured in a hipGraph (FP16: 8x8 tiles, two
    launches of 8 sweeps; SPLIT: 4x8 tiles, two launches of 6): three replays,
    each after eager work on the same buffers, each bitwise the eager run."""
    tile, fuse, sweeps = ((8, 8), 8, 16) if precision == "fp16" else ((4, 8), 6, 12)
    say("case fused8 graph", precision, tile, fuse, sweeps)
    B = T.gen_coeffs("box2d2r", 41)
    x = T.gen_uniform(300, 517, seed=41, dtype=torch.float32)
    e = T.split_scale_exp(1.0) if precision == "split" else 0
    p = T.Problem("box2d2r", 300, 517, B, precision=precision, tile=tile, fuse=fuse)
    assert p.fused_kernel_name.startswith("tcs_mfma_fused8<"), p.fused_kernel_name
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(x, b0, e)
    eager = p.unpack(p.run(b0, b1, sweeps), scale_exp=e).clone()
    g = p.graph(b0, b1, sweeps)
    for rep in range(3):
        p.pack(x, b0, e)  # eager work between replays
        g.launch()
        torch.cuda.synchronize()
        assert torch.equal(p.unpack(b0, scale_exp=e), eager), f"replay {rep}"
    g.destroy()


def test_fused_limits(T, torch):
    B = T.gen_coeffs("box2d2r", 1)
    p = T.Problem("box2d2r", 64, 64, B, fuse=10)  # radius 2 allows at most 9
    b0, b1 = p.alloc(), p.alloc()
    with pytest.raises(T.TcsError, match="fuse"):
        p.run(b0, b1, 10)
    q = T.Problem("box2d2r", 64, 64, B, method="direct", fuse=2)
    with pytest.raises(T.TcsError, match="MFMA"):
        q.run(b0, b1, 4)


# ------------------------------------------------------- batched uneven grids
def test_batch_bitwise_equals_per_grid(T, torch):
    """One launch per sweep (per tile shape) over many uneven grids gives the
    same bits as sweeping each grid on its own; 8x8 problems join the 4x8
    group."""
    rng = np.random.default_rng(31)
    probs, b0s, b1s, refs, xs = [], [], [], [], []
    for i in range(12):
        stencil = ("box2d1r", "star2d2r")[i % 2]
        rows, cols = (int(v) for v in rng.integers(5, 300, 2))
        tile = [(2, 4), (4, 4), (4, 8), (8, 8)][i % 4]
        B = T.gen_coeffs(stencil, 100 + i)
        x = T.gen_uniform(rows, cols, seed=100 + i, dtype=torch.float32)
        p = T.Problem(stencil, rows, cols, B, tile=tile)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x, b0)
        probs.append(p), b0s.append(b0), b1s.append(b1), xs.append(x)
        q0, q1 = p.alloc(), p.alloc()
        q1.zero_()
        p.pack(x, q0)
        refs.append(p.unpack(p.run(q0, q1, 5)))
    say("case batch", [(p.rows, p.cols, p.tile) for p in probs])
    batch = T.Batch(probs, b0s, b1s)
    assert batch.launches_per_sweep == 3
    batch.run(5)
    torch.cuda.synchronize()
    for p, b1, ref in zip(probs, b1s, refs):
        assert torch.equal(p.unpack(b1), ref)
    # the same through a captured graph, replayed from a fresh pack
    for p, b0, b1, x in zip(probs, b0s, b1s, xs):
        p.pack(x, b0)
    g = batch.graph(5)
    g.launch()
    torch.cuda.synchronize()
    for p, b1, ref in zip(probs, b1s, refs):
        assert torch.equal(p.unpack(b1), ref)
    g.destroy()
    batch.destroy()


# ------------------------------------------------------- 32x32 fragments
FRAG32_TILES = [(4, 4), (2, 8), (4, 8), (8, 8)]
FRAG32_FUSED_TILES = [(4, 4), (4, 8), (8, 8)]


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_frag32_dyadic_exact(T, torch, oracle, stencil, precision):
    """v_mfma_f32_32x32x16 kernels: bitwise fp16/bf16(oracle) on exact inputs."""
    for tile in FRAG32_TILES:
        for rows, cols in ((70, 150), (33, 17), (130, 260)):
            say("case frag32", stencil, precision, tile, rows, cols)
            A, B = dyadic(stencil, rows, cols, seed=rows + cols + 5)
            got = gpu_sweeps(T, torch, stencil, A, B, precision=precision, tile=tile, frag=32)
            ref = oracle.sweep_f64(stencil, A, B)
            want = round_bf16(torch, ref) if precision == "bf16" else ref.astype(np.float16).astype(np.float64)
            assert np.array_equal(got, want), f"max diff {np.abs(got - want).max()}"


@pytest.mark.parametrize("stencil", STENCILS)
def test_frag32_random_within_bound(T, torch, oracle, stencil):
    """Random FP16 data: same rounding bound as the 16x16 form, and tile-independent bits."""
    A, B = random_problem(T, torch, stencil, 200, 333, seed=41)
    ref = oracle.sweep_f64(stencil, A, B)
    outs = []
    for tile in FRAG32_TILES:
        say("case frag32 random", stencil, tile)
        outs.append(gpu_sweeps(T, torch, stencil, A, B.astype(np.float32), tile=tile, frag=32))
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    r = RADIUS[stencil]
    err = np.abs(outs[0] - ref)[r:-r, r:-r]
    max_ref = np.abs(ref).max()
    # coefficient rounding (|B| <= 1, sum |B| = 1, |x| <= 1) + output rounding
    assert err.max() <= 2.0 ** -11 * (1.0 + max_ref) * 1.01 + 1e-7
    assert oracle.error_norms(outs[0], ref, r)["rel_l2"] <= 2e-3


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_frag32_fused(T, torch, oracle, stencil, precision):
    """Fused 32x32-fragment sweeps: bitwise equal to the oracle on inputs that
    stay exact through k sweeps, and within rounding of single sweeps on random
    inputs (the ring is computed in blocks offset by one brick, so FP32 sums of
    a point may be grouped differently than in a single-sweep launch)."""
    r = RADIUS[stencil]
    n = 2 * r + 1
    A, _ = dyadic(stencil, 96, 200, seed=9)
    Bx = np.zeros((n, n), np.float32)
    Bx[r, r], Bx[r, 0], Bx[0, r] = 0.5, 0.25, 0.25
    for tile in FRAG32_FUSED_TILES:
        for fuse in (2, 1 + 16 // r):
            say("case frag32 fused exact", stencil, precision, tile, fuse)
            got = gpu_sweeps(T, torch, stencil, A, Bx, precision=precision, tile=tile, frag=32, fuse=fuse,
                             sweeps=fuse + 1)
            want = A
            for _ in range(fuse + 1):
                w = oracle.sweep_f64(stencil, want, Bx.astype(np.float64))
                want = round_bf16(torch, w) if precision == "bf16" else w.astype(np.float16).astype(np.float64)
            assert np.array_equal(got, want)
    Ar, Br = random_problem(T, torch, stencil, 160, 300, seed=43)
    single = gpu_sweeps(T, torch, stencil, Ar, Br.astype(np.float32), precision=precision, frag=32, sweeps=6)
    for tile in FRAG32_FUSED_TILES:
        say("case frag32 fused random", stencil, precision, tile)
        got = gpu_sweeps(T, torch, stencil, Ar, Br.astype(np.float32), precision=precision, tile=tile, frag=32,
                         fuse=3, sweeps=6)
        assert oracle.error_norms(got, single, r)["rel_l2"] <= (2e-3 if precision == "bf16" else 2e-4)


def test_frag32_batch_bitwise_equals_per_grid(T, torch):
    rng = np.random.default_rng(51)
    probs, b0s, b1s, refs = [], [], [], []
    for i in range(8):
        stencil = ("box2d1r", "star2d2r")[i % 2]
        rows, cols = (int(v) for v in rng.integers(5, 300, 2))
        frag, tile = [(32, (4, 4)), (32, (4, 8)), (16, (2, 4)), (16, (4, 8))][i % 4]
        B = T.gen_coeffs(stencil, 200 + i)
        x = T.gen_uniform(rows, cols, seed=200 + i, dtype=torch.float32)
        p = T.Problem(stencil, rows, cols, B, tile=tile, frag=frag)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x, b0)
        probs.append(p), b0s.append(b0), b1s.append(b1)
        q0, q1 = p.alloc(), p.alloc()
        q1.zero_()
        p.pack(x, q0)
        refs.append(p.unpack(p.run(q0, q1, 3)))
    say("case frag32 batch", [(p.rows, p.cols, p.tile, p.frag) for p in probs])
    batch = T.Batch(probs, b0s, b1s)
    assert batch.launches_per_sweep == 4
    batch.run(3)
    torch.cuda.synchronize()
    for p, b1, ref in zip(probs, b1s, refs):
        assert torch.equal(p.unpack(b1), ref)
    batch.destroy()


def test_batch_fused_persistent_and_forked_graph(T, torch):
    """Fused batched launches (per-problem fuse), persistent grids and the
    forked-graph capture give the same bits as each grid run on its own with
    the same settings."""
    rng = np.random.default_rng(61)
    cfgs = [(16, (2, 4), 3, 0), (16, (4, 8), 4, 1), (32, (4, 4), 3, 0), (32, (4, 8), 2, 2), (16, (4, 4), 1, 1)]
    probs, b0s, b1s, refs, xs = [], [], [], [], []
    for i in range(10):
        stencil = ("box2d1r", "star2d2r")[i % 2]
        rows, cols = (int(v) for v in rng.integers(20, 400, 2))
        frag, tile, fuse, wg = cfgs[i % len(cfgs)]
        B = T.gen_coeffs(stencil, 300 + i)
        x = T.gen_uniform(rows, cols, seed=300 + i, dtype=torch.float32)
        p = T.Problem(stencil, rows, cols, B, tile=tile, frag=frag, fuse=fuse, wg_per_cu=wg)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x, b0)
        probs.append(p), b0s.append(b0), b1s.append(b1), xs.append(x)
        # per-grid reference: same tile, fragment and fuse, one tile per workgroup
        pr = T.Problem(stencil, rows, cols, B, tile=tile, frag=frag, fuse=fuse)
        q0, q1 = pr.alloc(), pr.alloc()
        q1.zero_()
        pr.pack(x, q0)
        refs.append(pr.unpack(pr.run(q0, q1, 7)))
    say("case batch fused", [(p.rows, p.cols, p.tile, p.frag, p.fuse, p.wg_per_cu) for p in probs])
    batch = T.Batch(probs, b0s, b1s)
    assert batch.launches_per_sweep == len(cfgs)
    batch.run(7)
    torch.cuda.synchronize()
    for p, b1, ref in zip(probs, b1s, refs):
        assert torch.equal(p.unpack(b1), ref)
    g = batch.graph(7)
    for rep in range(3):
        for p, b0, b1, x in zip(probs, b0s, b1s, xs):
            p.pack(x, b0)
            p.clear(b1)
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        for p, b1, ref in zip(probs, b1s, refs):
            assert torch.equal(p.unpack(b1), ref), f"graph replay {rep}"
    g.destroy()
    batch.destroy()


def test_frag32_full_size_box2d2r_16384(T, torch):
    say("case frag32 full size")
    N = 16384
    B = T.gen_coeffs("box2d2r", 1)
    x64 = T.gen_sinusoid(N, N, seed=1, noise=1e-3)
    T.round_to(x64, "fp16")
    ref = torch.empty_like(x64).copy_(x64)
    T.oracle_sweep_f64("box2d2r", x64, B.astype(np.float64), ref)
    p = T.Problem("box2d2r", N, N, B, frag=32)
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(x64, b0)
    del x64
    p.sweep(b0, b1)
    res = T.error_norms(p.unpack(b1), ref, 2)
    say("  ", res)
    assert res["rel_l2"] <= 2e-3
    assert res["max_abs"] <= (2.0 ** -11) * 2.0 * 1.001 + 1e-6


def test_c1_star2d1r_1024_vs_gold(T, torch, oracle):
    """Config C1 as the reference's driver would check it: star2d1r on 1024^2,
    FP16 input U[-1,1), normalised coefficients, zero ring, one sweep, compared
    with gold_star2d1r (FP32, CPU) on the same input."""
    N = 1024
    x = T.gen_uniform(N, N, seed=1, dtype=torch.float32)
    T.round_to(x, "fp16")
    A = x.cpu().numpy()
    A[0, :] = A[-1, :] = A[:, 0] = A[:, -1] = 0.0
    B = T.gen_coeffs("star2d1r", 1)
    gold = oracle.gold("star2d1r", A, B)
    for method in ("mfma", "direct"):
        C, _ = T.host_sweep("star2d1r", A, B, method=method)
        err = np.abs(C.astype(np.float64) - gold)
        say("case c1", method, err.max())
        assert np.array_equal(C[0], gold[0]) and np.array_equal(C[:, -1], gold[:, -1])
        # coefficient rounding (MFMA) + output rounding, u = 2^-11, |x| <= 1
        assert err.max() <= (2.0 if method == "mfma" else 1.0) * 2.0 ** -11 + 1e-6


def three_steps(T, torch, p, src, sweeps, scale_exp=0):
    """tcs_pack_f32 -> tcs_run -> tcs_unpack_f32 (ring from the source)."""
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b0)
    p.clear(b1)
    p.pack(src, b0, scale_exp)
    return p.unpack(p.run(b0, b1, sweeps), scale_exp=scale_exp, ring=src)


@pytest.mark.parametrize("precision", ["fp16", "bf16", "split", "fp16lo", "bf16lo"])
def test_run_f32_fused_edges_bitwise(T, torch, precision):
    """tcs_run_f32 with the layout transforms fused into the first and last sweep
    (tcs_mfma_edge) equals pack -> run -> unpack bitwise: random inputs, every
    stencil and edge tile, 1-3 sweeps (and 4 with fused middle sweeps), widths
    that do and do not allow 16-byte row accesses, and a source 4 bytes off a
    16-byte boundary.  SPLIT with a per-grid scale exponent; FP16 and BF16 also
    with hi+lo taps (kernel precisions 4 and 5)."""
    lo = precision.endswith("lo")
    precision = precision[:-2] if lo else precision
    e = T.split_scale_exp(1.0) if precision == "split" else 0
    tiles = ((2, 4), (4, 4), (2, 8), (4, 8), (8, 8))
    for stencil in STENCILS:
        B = T.gen_coeffs(stencil, 5)
        for rows, cols in ((70, 150), (97, 517), (64, 128)):
            base = torch.empty(rows * cols + 4, dtype=torch.float32, device="cuda")
            for off in (0, 1):
                src = base[off:off + rows * cols].view(rows, cols)
                src.copy_(T.gen_uniform(rows, cols, seed=rows + cols, dtype=torch.float32))
                for tile in tiles:
                    p = T.Problem(stencil, rows, cols, B, precision=precision, tile=tile, coef_lo=lo)
                    assert p.coef_lo == lo and p.edge_kernel_name(3), p.edge_kernel_name(3)
                    b0, b1 = p.alloc(), p.alloc()
                    p.clear(b0)
                    p.clear(b1)
                    for sweeps, fuse in ((1, 1), (2, 1), (3, 1), (4, 2)):
                        say("case run_f32", stencil, precision, rows, cols, off, tile, sweeps, fuse)
                        p.fuse = fuse
                        got = p.run_f32(src, sweeps, buf0=b0, buf1=b1, scale_exp=e)
                        want = three_steps(T, torch, p, src, sweeps, e)
                        assert torch.equal(got, want), (stencil, rows, cols, off, tile, sweeps, fuse)


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_run_f32_wide_edge_tiles(T, torch, precision):
    """Large FP16/BF16 grids run tcs_run_f32's first, last and only sweeps on
    2x16-brick tiles (rows of 1 KB FP32 segments), the middle sweeps on the
    problem's own tile: bitwise the three steps, with ragged edges (rows and
    columns not multiples of 16), 1-3 sweeps and fused middle sweeps."""
    rows, cols = 8187, 4090
    for stencil in ("box2d2r", "star2d1r"):
        B = T.gen_coeffs(stencil, 9)
        p = T.Problem(stencil, rows, cols, B, precision=precision)
        for mode in (1, 2, 3):
            assert p.edge_kernel_name(mode).startswith("tcs_mfma_edge<2, 16,"), p.edge_kernel_name(mode)
        src = T.gen_uniform(rows, cols, seed=9, dtype=torch.float32)
        b0, b1 = p.alloc(), p.alloc()
        for sweeps, fuse in ((1, 1), (2, 1), (3, 1), (5, 3)):
            say("case run_f32 wide edges", stencil, precision, rows, cols, sweeps, fuse)
            p.fuse = fuse
            got = p.run_f32(src, sweeps, buf0=b0, buf1=b1)
            want = three_steps(T, torch, p, src, sweeps)
            assert torch.equal(got, want), (stencil, sweeps, fuse)
        del b0, b1, src


def test_run_f32_tiny_grids(T, torch, oracle):
    """Grids smaller than one brick or one tile, in one or both directions:
    tcs_run_f32 equals pack -> run -> unpack, and one sweep equals the oracle
    (all ring where the interior is empty)."""
    for stencil in STENCILS:
        B = T.gen_coeffs(stencil, 3)
        for rows, cols in ((1, 1), (5, 3), (3, 40), (17, 1), (33, 17), (6, 6)):
            say("case run_f32 tiny", stencil, rows, cols)
            x = T.gen_uniform(rows, cols, seed=rows * cols, dtype=torch.float32)
            T.round_to(x, "fp16")
            p = T.Problem(stencil, rows, cols, B)
            b0, b1 = p.alloc(), p.alloc()
            p.clear(b0)
            p.clear(b1)
            for sweeps in (1, 2):
                got = p.run_f32(x, sweeps, buf0=b0, buf1=b1)
                assert torch.equal(got, three_steps(T, torch, p, x, sweeps)), (stencil, rows, cols, sweeps)
            got = p.run_f32(x, 1, buf0=b0, buf1=b1).cpu().numpy()
            A = x.cpu().numpy().astype(np.float64)
            want = oracle.sweep_f64(stencil, A, B.astype(np.float64))
            r = 1 if stencil in ("star2d1r", "box2d1r") else 2
            if rows <= 2 * r or cols <= 2 * r:
                assert np.array_equal(got, x.cpu().numpy())
            # the input is FP16-representable: the one-sweep worst-case bound applies
            b_inf, _ = PM.error_bounds(stencil, B, PM.kernel_coeffs(B, "fp16"), "fp16", 1, np.abs(A).max(),
                                       np.linalg.norm(A), A.size)
            assert np.abs(got - want).max() <= b_inf, (stencil, rows, cols)


@pytest.mark.parametrize("precision", ["fp16", "split"])
def test_run_f32_captured_in_a_graph(T, torch, precision):
    """tcs_run_f32 launches only kernels on the caller's stream, so a caller can
    capture it (here in a torch CUDA graph): every replay, with eager work in
    between, equals the eager result."""
    rows, cols = 200, 300
    B = T.gen_coeffs("box2d2r", 2)
    x = T.gen_uniform(rows, cols, seed=12, dtype=torch.float32)
    e = T.split_scale_exp(1.0) if precision == "split" else 0
    p = T.Problem("box2d2r", rows, cols, B, precision=precision)
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b0)
    p.clear(b1)
    eager = p.run_f32(x, 3, buf0=b0, buf1=b1, scale_exp=e).clone()
    out = torch.empty_like(x)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        p.run_f32(x, 3, out=out, buf0=b0, buf1=b1, scale_exp=e)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        p.run_f32(x, 3, out=out, buf0=b0, buf1=b1, scale_exp=e)
    for rep in range(4):
        out.fill_(-1.0)
        p.run_f32(x, 1, buf0=b0, buf1=b1, scale_exp=e)  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), rep
    del g


def test_run_f32_row_strided_views(T, torch):
    """A grid that is a column slice of a wider array (leading dimension 160
    for 150 columns) gives the result of the contiguous grid, fused and in
    three steps, and leaves the columns past the grid untouched."""
    rows, cols, ld = 70, 150, 160
    B = T.gen_coeffs("box2d2r", 5)
    x = T.gen_uniform(rows, cols, seed=4, dtype=torch.float32)
    wide = torch.zeros((rows, ld), dtype=torch.float32, device="cuda")
    wide[:, :cols] = x
    for kw in ({}, {"precision": "split"}):
        p = T.Problem("box2d2r", rows, cols, B, **kw)
        e = T.split_scale_exp(1.0) if kw else 0
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b0)
        p.clear(b1)
        for sweeps in (1, 3):
            say("case run_f32 strided", kw, sweeps)
            want = p.run_f32(x, sweeps, buf0=b0, buf1=b1, scale_exp=e)
            out = torch.full((rows, ld), -5.0, dtype=torch.float32, device="cuda")
            got = p.run_f32(wide[:, :cols], sweeps, out=out[:, :cols], buf0=b0, buf1=b1, scale_exp=e)
            assert torch.equal(got, want), (kw, sweeps)
            assert bool(out[:, cols:].eq(-5.0).all())


@pytest.mark.parametrize("precision", ["fp16", "split"])
def test_run_f32_ignores_prior_buffer_contents(T, torch, precision):
    """tcs_run_f32 gives the three-step result whatever the storage buffers
    held before: both filled with 0xFF bytes (a NaN in every FP16 element,
    ghost ring included), 2 and 3 sweeps, fused edges and the three-step path
    (direct method)."""
    rows, cols = 97, 150
    B = T.gen_coeffs("box2d2r", 7)
    src = T.gen_uniform(rows, cols, seed=13, dtype=torch.float32)
    e = T.split_scale_exp(1.0) if precision == "split" else 0
    kinds = [{"tile": (4, 8)}, {"tile": (2, 4)}] + ([{"method": "direct"}] if precision == "fp16" else [])
    for kw in kinds:
        p = T.Problem("box2d2r", rows, cols, B, precision=precision, **kw)
        for sweeps in (2, 3):
            say("case run_f32 dirty buffers", precision, kw, sweeps)
            want = three_steps(T, torch, p, src, sweeps, e)
            b0, b1 = p.alloc(), p.alloc()
            b0.fill_(255)
            b1.fill_(255)
            got = p.run_f32(src, sweeps, buf0=b0, buf1=b1, scale_exp=e)
            assert torch.equal(got, want), (kw, sweeps)


@pytest.mark.parametrize("tile_size", ["1", "15"])
def test_dropin_driver_reference_tile_size(tile_size):
    """The reference's TILE_SIZE values (default 1, sweep up to 15,
    common/param.hpp:11-13, script/param.py:5) through the drop-in driver:
    mapped to an offered tile, same checks as the default."""
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for method in ("mfma", "direct"):
        exe = os.path.join(root, "tests", "dropin", "_build", f"main_{method}_box2d2r")
        env = dict(os.environ, TCS_RUN_TIMES="5", TCS_SKIP_TIMES="2", TCS_TILE_SIZE=tile_size)
        out = subprocess.run([exe, "200", "5"], capture_output=True, text=True, timeout=120, env=env)
        say(method, tile_size, out.stdout)
        assert out.returncode == 0, out.stdout + out.stderr
        lines = dict(l.split(":", 1) for l in out.stdout.splitlines() if l.startswith("[Test]"))
        assert int(lines["[Test] Ring mismatches"]) == 0
        max_err = float(lines["[Test] Max Error"].split()[0])
        max_ref = float(lines["[Test] Max Error"].split()[3].rstrip(")"))
        assert max_err <= 2.0 ** -11 * (0.5 * 578.0 + max_ref) * 1.01 + 1e-6


def test_run_f32_other_problems_run_three_steps(T, torch):
    """Problems without an edge kernel (direct, 32x32 fragments) take the
    three-step path through tcs_run_f32; persistent-grid problems run their edge
    sweeps one tile per workgroup; a fused single sweep needs no storage
    buffers."""
    rows, cols = 70, 150
    B = T.gen_coeffs("box2d2r", 3)
    src = T.gen_uniform(rows, cols, seed=9, dtype=torch.float32)
    e = T.split_scale_exp(1.0)
    for kw, scale in (({"method": "direct"}, 0), ({"frag": 32}, 0), ({"wg_per_cu": 2}, 0),
                      ({"precision": "split", "wg_per_cu": 2}, e)):
        say("case run_f32 three steps", kw)
        p = T.Problem("box2d2r", rows, cols, B, **kw)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b0)
        p.clear(b1)
        got = p.run_f32(src, 3, buf0=b0, buf1=b1, scale_exp=scale)
        assert torch.equal(got, three_steps(T, torch, p, src, 3, scale))
    p = T.Problem("box2d2r", rows, cols, B)
    assert torch.equal(p.run_f32(src, 1), three_steps(T, torch, p, src, 1))
    with pytest.raises(T.TcsError):
        p.run_f32(src, 2)  # two sweeps need storage buffers
