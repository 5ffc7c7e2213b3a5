"""The benchmark configurations at their full sizes, batched grids and the
reference's golden fixtures, each checked against the oracle (run on an MI355X).

  C2  box2d1r 10240^2 FP16 U[-1,1), 100 sweeps: MFMA (rounded and hi+lo taps)
      and direct-FMA, errors vs the FP64 oracle after 1, 10 and 100 sweeps
  C3  star2d2r and box2d2r 16384^2, 4 sinusoids + noise, 10 sweeps, FP16 and
      BF16 (rounded and hi+lo taps) and split hi+lo
  C4  512 batched uneven grids (box2d1r / star2d2r, 64..2048 per side), 20
      sweeps captured as one hipGraph, every grid checked
  C5  box2d2r 32768x16384 split hi+lo, 256 sweeps, fused 2/4/8 per launch
  batches against the oracle (exact inputs bitwise, random inputs within the
  model), the 12 fixtures written by the reference's gold through the drop-in
  host path.

Large results are compared on the GPU: the library's FP64 oracle sweep is
bitwise the CPU oracle (test_gpu_parity.py::test_gpu_fp64_oracle_bitwise), and
the rounding model is that sweep with the kernel's taps followed by rounding to
the storage format (precision_model.py).  Every result's max-abs and rel-L2
against FP64 is printed.
"""
import glob
import math
import os

import numpy as np
import pytest

import precision_model as PM

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RADIUS = {"star2d1r": 1, "star2d2r": 2, "box2d1r": 1, "box2d2r": 2}
SEED = 1


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


def zero_ring(x, r):
    for v in (x[:r], x[-r:], x[:, :r], x[:, -r:]):
        v.zero_()


def run(T, torch, p, x, sweeps, scale_exp=0, graph=False):
    """Pack x, `sweeps` sweeps (or one hipGraph replay of them), unpack."""
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(x, b0, scale_exp)
    if graph:
        g = p.graph(b0, b1, sweeps)
        g.launch()
        torch.cuda.synchronize()
        g.destroy()
        res = b1 if sweeps % 2 else b0
    else:
        res = p.run(b0, b1, sweeps)
    out = p.unpack(res, scale_exp=scale_exp)
    del b0, b1
    return out


def check_full(T, torch, stencil, x64, B, got, ref, model, precision, sweeps, ceff, l2_0, what):
    """A full-size result against the GPU rounding model (accumulation-order
    tolerance) and the GPU FP64 oracle (rigorous worst-case bounds)."""
    r = RADIUS[stencil]
    e = T.error_norms(got, ref, r)
    vs_model = T.error_norms(got, model, r)["rel_l2"]
    b_inf, b_2 = PM.error_bounds(stencil, B, ceff, precision, sweeps, float(x64.abs().max()), l2_0, x64.numel())
    say(f"  {what} k={sweeps}: rel-L2 {e['rel_l2']:.3e} max-abs {e['max_abs']:.3e} vs FP64; "
        f"vs model {vs_model:.2e} (tol {PM.model_tolerance(precision, sweeps):.2e}); "
        f"bounds max-abs {b_inf:.2e} rel-L2 {b_2 / e['ref_l2']:.2e}")
    assert vs_model <= PM.model_tolerance(precision, sweeps), (what, sweeps, vs_model)
    assert e["max_abs"] <= b_inf, (what, sweeps)
    assert e["rel_l2"] * e["ref_l2"] <= b_2, (what, sweeps)
    return e


# ------------------------------------------------------------------------ C2
def test_c2_box2d1r_10240_100_sweeps(T, torch):
    stencil, N, counts = "box2d1r", 10240, (1, 10, 100)
    say("case C2", stencil, N)
    B = T.gen_coeffs(stencil, SEED)
    x = T.gen_uniform(N, N, seed=SEED, dtype=torch.float64)
    T.round_to(x, "fp16")
    zero_ring(x, 1)
    l2_0 = float(x.norm())
    refs = PM.gpu_oracle_sweeps(T, stencil, x, B, counts)
    errs = {}
    for method, coef_lo in (("mfma", False), ("mfma", True), ("direct", False)):
        ceff = B.astype(np.float64) if method == "direct" else PM.kernel_coeffs(B, "fp16", coef_lo)
        models = PM.gpu_model_sweeps(T, stencil, x, ceff, "fp16", counts)
        p = T.Problem(stencil, N, N, B, method=method, coef_lo=coef_lo)
        what = f"{method}{' hi+lo' if coef_lo else ''}"
        for k in counts:
            got = run(T, torch, p, x, k)
            errs[(what, k)] = check_full(T, torch, stencil, x, B, got, refs[k], models[k], "fp16", k, ceff, l2_0, what)
            del got
        del models
    # hi+lo taps: the accuracy of the FP32-tap direct kernel (within 2x) after 100 sweeps
    assert errs[("mfma hi+lo", 100)]["rel_l2"] <= 2 * errs[("direct", 100)]["rel_l2"]
    # the 100-sweep loop fused 16 per launch and captured as one hipGraph: bitwise the unfused loop
    pf = T.Problem(stencil, N, N, B, fuse=16)
    pu = T.Problem(stencil, N, N, B)
    assert torch.equal(run(T, torch, pf, x, 100, graph=True), run(T, torch, pu, x, 100))


# ------------------------------------------------------------------------ C3
@pytest.mark.parametrize("stencil", ["star2d2r", "box2d2r"])
def test_c3_16384_three_precision_forms(T, torch, stencil):
    N, counts = 16384, (1, 10)
    say("case C3", stencil, N)
    B = T.gen_coeffs(stencil, SEED)
    x64 = T.gen_sinusoid(N, N, seed=SEED, noise=1e-3)
    zero_ring(x64, 2)
    for precision in ("fp16", "bf16"):
        x = x64.clone()
        T.round_to(x, precision)
        l2_0 = float(x.norm())
        refs = PM.gpu_oracle_sweeps(T, stencil, x, B, counts)
        for coef_lo in (False, True):
            ceff = PM.kernel_coeffs(B, precision, coef_lo)
            models = PM.gpu_model_sweeps(T, stencil, x, ceff, precision, counts)
            p = T.Problem(stencil, N, N, B, precision=precision, coef_lo=coef_lo)
            for k in counts:
                got = run(T, torch, p, x, k)
                e = check_full(T, torch, stencil, x, B, got, refs[k], models[k], precision, k, ceff, l2_0,
                               f"{precision}{' hi+lo' if coef_lo else ''}")
                if k == 1 and precision == "fp16":
                    assert e["rel_l2"] <= 2e-3  # the per-sweep FP16 target
                del got
            del models
        del x, refs
    # split hi+lo with per-grid max-abs scaling on the unrounded input: <= 1e-6 per sweep
    e = T.split_scale_exp(T.maxabs(x64))
    p = T.Problem(stencil, N, N, B, precision="split")
    refs = PM.gpu_oracle_sweeps(T, stencil, x64, B, counts)
    for k in counts:
        res = T.error_norms(run(T, torch, p, x64, k, scale_exp=e), refs[k], 2)
        say(f"  split k={k}: rel-L2 {res['rel_l2']:.3e} max-abs {res['max_abs']:.3e}")
        assert res["rel_l2"] <= 1e-6


# ------------------------------------------------------------------------ C4
def c4_specs(T, torch, nprob=512):
    """The C4 workload: stencil per problem by seed, rows and cols log-uniform
    in [64, 2048], FP16 U[-1,1) inputs, per-problem coefficients."""
    rng = np.random.default_rng(SEED)
    kinds = rng.integers(0, 2, nprob)
    rows = np.exp(rng.uniform(math.log(64), math.log(2048), nprob)).astype(int)
    cols = np.exp(rng.uniform(math.log(64), math.log(2048), nprob)).astype(int)
    specs = []
    for i in range(nprob):
        stencil = ("box2d1r", "star2d2r")[kinds[i]]
        x = T.gen_uniform(int(rows[i]), int(cols[i]), seed=SEED + i, dtype=torch.float32)
        T.round_to(x, "fp16")
        specs.append((stencil, int(rows[i]), int(cols[i]), T.gen_coeffs(stencil, SEED + i), x))
    return specs


# The autotuner's choice for C4 per size bucket (smaller side < 16, < 64, >= 64
# bricks): tile, fragment, sweeps per launch, waves per CU; here with 32x32
# fragments and a persistent grid in the middle bucket as well, so that every
# batched kernel kind runs.
C4_CHOICE = {0: ((4, 4), 16, 16, 0), 1: ((4, 8), 32, 8, 2), 2: ((4, 8), 16, 8, 0)}


def test_c4_512_batched_grids_every_grid_vs_oracle(T, torch):
    from tcstencil_amd import autotune as A
    sweeps = 20
    specs = c4_specs(T, torch)
    say("case C4", len(specs), "grids", sum((r - 2) * (c - 2) for _, r, c, _, _ in specs), "points")
    bucket = [A.bucket_of(r, c) for _, r, c, _, _ in specs]
    probs, b0s, b1s, batch = A.build_batch(specs, lambda i: C4_CHOICE[bucket[i]])
    g = batch.graph(sweeps)
    g.launch()
    torch.cuda.synchronize()
    worst = {"rel_l2": 0.0, "max_abs": 0.0, "vs_model": 0.0}
    for i, ((stencil, rows, cols, B, x), p, b0, b1) in enumerate(zip(specs, probs, b0s, b1s)):
        got = p.unpack(b1 if sweeps % 2 else b0)
        x64 = x.double()
        r = RADIUS[stencil]
        ref = PM.gpu_oracle_sweeps(T, stencil, x64, B, [sweeps])[sweeps]
        ceff = PM.kernel_coeffs(B, "fp16")
        model = PM.gpu_model_sweeps(T, stencil, x64, ceff, "fp16", [sweeps])[sweeps]
        e = T.error_norms(got, ref, r)
        vm = T.error_norms(got, model, r)["rel_l2"]
        b_inf, b_2 = PM.error_bounds(stencil, B, ceff, "fp16", sweeps, float(x64.abs().max()), float(x64.norm()),
                                     x64.numel())
        assert vm <= PM.model_tolerance("fp16", sweeps), (i, stencil, rows, cols, p.tile, p.frag, p.fuse, vm)
        assert e["max_abs"] <= b_inf and e["rel_l2"] * e["ref_l2"] <= b_2, (i, stencil, rows, cols)
        worst = {"rel_l2": max(worst["rel_l2"], e["rel_l2"]), "max_abs": max(worst["max_abs"], e["max_abs"]),
                 "vs_model": max(worst["vs_model"], vm)}
    say(f"  C4 after {sweeps} sweeps, worst grid vs FP64: rel-L2 {worst['rel_l2']:.3e} "
        f"max-abs {worst['max_abs']:.3e}; worst vs model {worst['vs_model']:.2e}")
    g.destroy()
    batch.destroy()


# ------------------------------------------------------------------------ C5
def test_c5_32768x16384_split_fused_256_sweeps(T, torch):
    stencil, R, C, sweeps = "box2d2r", 32768, 16384, 256
    say("case C5", stencil, R, C)
    B = T.gen_coeffs(stencil, SEED)
    x = T.gen_uniform(R, C, seed=SEED, dtype=torch.float64)
    zero_ring(x, 2)
    e = T.split_scale_exp(T.maxabs(x))
    # FP64 oracle with the L2 norm of every intermediate sweep
    a, b = x.clone(), x.clone()
    norms2 = []
    for _ in range(sweeps):
        norms2.append(float(a[2:-2, 2:-2].norm()) ** 2)
        T.oracle_sweep_f64(stencil, a, B.astype(np.float64), b)
        a, b = b, a
    ref = a
    del b
    outs = {}
    for fuse in (1, 2, 4, 8):
        p = T.Problem(stencil, R, C, B, precision="split", fuse=fuse)
        outs[fuse] = run(T, torch, p, x, sweeps, scale_exp=e)
        if fuse > 1:
            assert torch.equal(outs[fuse], outs[1]), f"fuse {fuse}"
            del outs[fuse]
    res = T.error_norms(outs[1], ref, 2)
    # each sweep j adds at most 1e-6 ||x_j|| in L2 (the per-sweep split target),
    # the stencil does not amplify (L2 norm <= sum|c| = 1), errors of different
    # sweeps add in quadrature: ||e|| <= 1e-6 sqrt(sum_j ||x_j||^2)
    bound = 1e-6 * math.sqrt(sum(norms2)) / res["ref_l2"]
    say(f"  C5 split after {sweeps} sweeps: rel-L2 {res['rel_l2']:.3e} max-abs {res['max_abs']:.3e} "
        f"(bound {bound:.2e}); fused 2/4/8 bitwise equal to unfused")
    assert res["rel_l2"] <= bound


# ------------------------------------------------------ batches against the oracle
def dyadic_signed(stencil, rows, cols, rng):
    """Exact inputs: A in multiples of 1/64, taps in multiples of 1/32 in
    [-1, 1] (signed, asymmetric); every product and partial sum is exact in
    FP32, so a kernel result is the oracle's value rounded once."""
    r = RADIUS[stencil]
    n = 2 * r + 1
    A = rng.integers(-64, 65, (rows, cols)) / 64.0
    B = rng.integers(-32, 33, (n, n)) / 32.0
    if stencil.startswith("star"):
        m = np.zeros((n, n), bool)
        m[r, :] = m[:, r] = True
        B = B * m
    return A, B.astype(np.float32)


BATCH_CFGS = [(16, (2, 4), 1, 0), (16, (4, 4), 3, 0), (16, (4, 8), 1, 2), (32, (4, 4), 1, 0), (32, (4, 8), 3, 0),
              (16, (4, 8), 3, 1)]


@pytest.mark.parametrize("precision", ["fp16", "bf16"])
def test_batch_exact_inputs_bitwise_vs_oracle(T, torch, oracle, precision):
    """Every batched kernel kind (16x16 and 32x32 fragments, fused and not,
    persistent and not) on exact inputs: one sweep gives the oracle's value
    rounded to the storage format, bitwise, for every grid; three sweeps with
    taps {1/2, 1/2} (FP32 sums stay exact) give the oracle's sweeps with that
    rounding after each, bitwise."""
    rng = np.random.default_rng(81)
    for sweeps in (1, 3):
        probs, b0s, b1s, want = [], [], [], []
        for i in range(12):
            stencil = ("box2d1r", "star2d2r", "box2d2r", "star2d1r")[i % 4]
            rows, cols = (int(v) for v in rng.integers(5, 300, 2))
            frag, tile, fuse, wg = BATCH_CFGS[i % len(BATCH_CFGS)]
            A, B = dyadic_signed(stencil, rows, cols, rng)
            if sweeps == 3:
                r = RADIUS[stencil]
                B = np.zeros_like(B)
                B[r, r] = B[r, 0] = 0.5
            p = T.Problem(stencil, rows, cols, B, precision=precision, tile=tile, frag=frag, fuse=fuse,
                          wg_per_cu=wg)
            b0, b1 = p.alloc(), p.alloc()
            p.clear(b1)
            p.pack(torch.from_numpy(A).cuda(), b0)
            probs.append(p), b0s.append(b0), b1s.append(b1)
            # each sweep's FP32 sum is exact, so the kernel rounds the exact value once per sweep
            want.append(PM.model_sweeps(oracle, stencil, A, B.astype(np.float64), sweeps, precision))
        say("case batch exact", precision, sweeps, [(p.rows, p.cols, p.tile, p.frag, p.fuse) for p in probs])
        batch = T.Batch(probs, b0s, b1s)
        batch.run(sweeps)
        torch.cuda.synchronize()
        for p, b0, b1, w in zip(probs, b0s, b1s, want):
            got = p.unpack(b1 if sweeps % 2 else b0).cpu().numpy().astype(np.float64)
            assert np.array_equal(got, w), (p.rows, p.cols, p.tile, p.frag, p.fuse)
        batch.destroy()


@pytest.mark.parametrize("coef_lo", [False, True])
def test_batch_random_inputs_vs_models(T, torch, oracle, coef_lo):
    """Random FP16 inputs and normalised taps, 8 sweeps in a captured graph:
    every grid within the rounding model's tolerance and the rigorous bounds
    (hi+lo taps: 16x16 fragments)."""
    rng = np.random.default_rng(82)
    cfgs = [c for c in BATCH_CFGS if not (coef_lo and c[0] == 32)]
    sweeps = 8
    probs, b0s, b1s, inputs = [], [], [], []
    for i in range(10):
        stencil = ("box2d1r", "star2d2r", "box2d2r", "star2d1r")[i % 4]
        rows, cols = (int(v) for v in rng.integers(20, 400, 2))
        frag, tile, fuse, wg = cfgs[i % len(cfgs)]
        B = T.gen_coeffs(stencil, 400 + i)
        x = T.gen_uniform(rows, cols, seed=400 + i, dtype=torch.float64)
        T.round_to(x, "fp16")
        p = T.Problem(stencil, rows, cols, B, tile=tile, frag=frag, fuse=fuse, wg_per_cu=wg, coef_lo=coef_lo)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x, b0)
        probs.append(p), b0s.append(b0), b1s.append(b1), inputs.append((stencil, x.cpu().numpy(), B))
    batch = T.Batch(probs, b0s, b1s)
    g = batch.graph(sweeps)
    g.launch()
    torch.cuda.synchronize()
    for p, b0, b1, (stencil, A, B) in zip(probs, b0s, b1s, inputs):
        got = p.unpack(b0).cpu().numpy().astype(np.float64)  # 8 sweeps: result in buf0
        r = RADIUS[stencil]
        ceff = PM.kernel_coeffs(B, "fp16", coef_lo)
        model = PM.model_sweeps(oracle, stencil, A, ceff, sweeps, "fp16")
        exact = oracle.sweeps_f64(stencil, A, B.astype(np.float64), sweeps)
        vm = PM.rel_l2(got, model, r)
        b_inf, b_2 = PM.error_bounds(stencil, B, ceff, "fp16", sweeps, np.abs(A).max(), np.linalg.norm(A), A.size)
        say("  batch grid", stencil, p.rows, p.cols, p.tile, p.frag, p.fuse, "vs model %.2e" % vm,
            "rel-L2 %.2e" % PM.rel_l2(got, exact, r))
        assert vm <= PM.model_tolerance("fp16", sweeps)
        assert np.abs(got - exact).max() <= b_inf
        assert np.linalg.norm((got - exact)[r:-r, r:-r]) <= b_2
    g.destroy()
    batch.destroy()


# --------------------------------------------------- the reference's fixtures
FIXTURES = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_golden_fixtures_through_the_host_driver(T, path):
    """The fixtures hold A, B and the output C of the reference's own gold
    (compiled from common/gold_<stencil>.cpp).  The drop-in host path (FP32 in,
    FP32 out) must reproduce C within the rounding of its precision form:
      FP16 MFMA:   |out - C| <= u(sum|B| max|A| [input rounding]
                                 + sum|B| max|A| [tap rounding]
                                 + max|C| [output rounding]) + g sum|B| max|A|
      direct FMA:  the same without the tap rounding term
      split hi+lo: |out - C| <= (3 * 2^-22 + 2g') sum|B| max|A|, g' = 75 * 2^-24
    with u = 2^-11 and g = 2n * 2^-24 (FP32 sums of n products, here and in
    gold); ring points equal A exactly."""
    d = np.load(path)
    A, B, C = d["A"], d["B"], d["C"]
    stencil = os.path.basename(path).split("_")[0]
    r = RADIUS[stencil]
    n = PM.N_PRODUCTS[stencil]
    u, g = 2.0 ** -11, 2 * n * 2.0 ** -24
    sb, ma, mc = float(np.abs(B).sum()), float(np.abs(A).max()), float(np.abs(C).max())
    ring = np.ones_like(A, bool)
    ring[r:-r, r:-r] = False
    for method, precision in (("mfma", "fp16"), ("direct", "fp16"), ("mfma", "split")):
        out, _ = T.host_sweep(stencil, A, B, method=method, precision=precision)
        err = float(np.abs(out.astype(np.float64) - C)[r:-r, r:-r].max())
        if precision == "split":
            bound = (3 * 2.0 ** -22 + 2 * 75 * 2.0 ** -24) * sb * ma
        else:
            bound = u * (sb * ma * (2 if method == "mfma" else 1) + mc) + g * sb * ma
        say("case fixture", os.path.basename(path), method, precision, "max |out - gold| %.3e (bound %.3e)" % (err, bound))
        assert err <= bound
        assert np.array_equal(out[ring], A[ring])
