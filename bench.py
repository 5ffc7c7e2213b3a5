#!/usr/bin/env python3
"""Benchmark: box2d2r on a 16384 x 16384 FP16 grid, MFMA vs direct-FMA.

One step = one Jacobi sweep of the whole grid (ping-pong between two brick
buffers resident in HBM).  `value` is the MFMA kernel's GStencil/s =
(N-4)^2 * steps / t / 1e9 with t measured by HIP events on the launch stream
around exactly `steps` sweeps.  The same line reports, on the same input and
timed the same way: the MFMA kernel with hi+lo coefficients, the direct-FMA
kernel with FP32 taps (the counterpart of the reference's baseline kernel), its
packed-half form (v_dot2_f32_f16, FP16 taps), and 8 sweeps fused per launch;
the accuracy of each against the FP64 oracle (after 1 and 10 sweeps), the HBM
roofline of the MFMA kernel and the reference's CPU gold timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N = 16384
STENCIL = "box2d2r"
SEED = 1
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E spec
FP16_DENSE_PEAK_TFLOPS = 2500.0
BYTES_PER_UPDATE = 4.0         # FP16 in + FP16 out
USEFUL_FLOPS_PER_POINT = 50    # 25 multiply-adds
METRIC = "GStencil/s + %HBM peak, MFMA vs direct-FMA at stated rel-L2 error vs FP64"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_sweeps(torch, p, b0, b1, steps, warmup):
    """Mean ms per sweep over `steps` sweeps, HIP events on torch's current stream
    (the stream every sweep is launched on)."""
    p.run(b0, b1, warmup + (warmup % 2))
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    p.run(b0, b1, steps)
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def accuracy(torch, T, p, x64, refs):
    """rel-L2 and max-abs vs the FP64 oracle after each sweep count in refs."""
    b0, b1 = p.alloc(), p.alloc()
    out = {}
    for k in sorted(refs):
        p.clear(b1)
        p.pack(x64, b0)
        y = p.unpack(p.run(b0, b1, k))
        out[k] = T.error_norms(y, refs[k], 2)
        del y
    del b0, b1
    return out


TRAFFIC_PROFILE = os.path.join("profiles", "box2d2r_16384_fp16_mfma_pmc.md")


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE passes, calibrated on a copy of known size): the
    bytes, or None when the summary is missing or was taken on another kernel."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    if not os.path.exists(path):
        return None
    text = open(path).read()
    k = re.search(r"traffic_kernel:\s*(.+)", text)
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    if not (k and m) or k.group(1).strip() != kernel:
        return None
    return float(m.group(1))


def cpu_baseline(np):
    """The reference's gold_box2d2r.cpp (compiled -O3 -march=x86-64-v3 from the
    reference sources into oracle/_ref) on two full 16384^2 sweeps (about 11 s),
    1 thread; falls back to the oracle's C restatement."""
    from oracle import oracle as O
    A = np.random.default_rng(SEED).uniform(-1, 1, (N, N)).astype(np.float32)
    B = np.full((5, 5), 1.0 / 25.0, dtype=np.float32)
    kind = "reference" if O.ref_available(STENCIL, fast=True) else "port"
    sweeps = 2
    C = A.copy()  # ping-pong pair; gold leaves the ring, so both keep A's ring
    t0 = time.perf_counter()
    for _ in range(sweeps):
        if kind == "reference":
            O.ref_gold(STENCIL, A, B, C, fast=True)
        else:
            O.gold(STENCIL, A, B, C)
        A, C = C, A
    dt = time.perf_counter() - t0
    return {"value": round(sweeps * (N - 4) ** 2 / dt / 1e9, 4), "unit": "GStencil/s", "cores": 1, "kind": kind,
            "sample": f"{sweeps} box2d2r sweeps of {N}x{N} FP32, single thread, {dt:.1f} s "
                      f"({'oracle/_ref/libgold_box2d2r_fast.so' if kind == 'reference' else 'oracle/_build/liboracle.so'}), "
                      f"host {os.cpu_count()} logical CPUs, {cpu_model()}"}


def cpu_baseline_threads(np, threads):
    """The oracle's restatement of gold_box2d2r.cpp on row bands, one thread
    per band (`threads` host cores), eight ping-pong 16384^2 sweeps: a multi-core
    CPU figure beside the single-thread reference one (single-threaded, the
    restatement and the reference's -O3 build run at the same rate: the loop
    does not vectorise)."""
    from oracle import oracle as O
    A = np.random.default_rng(SEED).uniform(-1, 1, (N, N)).astype(np.float32)
    B = np.full((5, 5), 1.0 / 25.0, dtype=np.float32)
    C = A.copy()
    sweeps = 8
    t0 = time.perf_counter()
    for _ in range(sweeps):
        O.sweep_f32_threads(STENCIL, A, B, C, threads)
        A, C = C, A
    dt = time.perf_counter() - t0
    return {"value": round(sweeps * (N - 4) ** 2 / dt / 1e9, 4), "unit": "GStencil/s", "cores": threads,
            "kind": "port",
            "sample": f"{sweeps} box2d2r sweeps of {N}x{N} FP32 on {threads} row bands, {dt:.1f} s "
                      f"(oracle/_build/liboracle.so), host {os.cpu_count()} logical CPUs, {cpu_model()}"}


def kernel_tile(kernel, tile):
    """The tile the kernel runs: its first two template arguments (a fused launch
    may run another tile than the problem's), else the problem's tile."""
    m = re.match(r"[a-z0-9_]+<(\d+), (\d+),", kernel or "")
    return [int(m.group(1)), int(m.group(2))] if m else list(tile)


def form_record(r, what):
    """One measured form of the sweep for the JSON line."""
    return {"value": round(r["gstencil"], 2), "unit": "GStencil/s", "ms_per_step": round(r["ms"], 5),
            "kernel": r["kernel"], "tile_bricks": kernel_tile(r["kernel"], r["tile"]), "what": what,
            "rel_l2_1sweep": r["acc"][1]["rel_l2"], "max_abs_1sweep": r["acc"][1]["max_abs"],
            "rel_l2_10sweeps": r["acc"][10]["rel_l2"], "max_abs_10sweeps": r["acc"][10]["max_abs"]}


def cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), or 'unknown CPU'."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    # 1000 untimed sweeps (0.18 s of MFMA sweeps): an idle GPU's clocks take about
    # 40 ms of load to settle (tcstencil_amd/timing.py)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "the stencil shards by brick-row slabs but no multi-GPU path is built", file=sys.stderr)
        sys.exit(2)
    if int(os.environ.get("RANK", "0")) != 0:
        sys.exit(0)

    import numpy as np
    import torch

    import tcstencil_amd as T

    torch.cuda.set_device(0)
    B = T.gen_coeffs(STENCIL, SEED)
    log(f"bench: {STENCIL} {N}x{N}, input = 4 sinusoids + U[-1e-3,1e-3) noise (seed {SEED}), FP16")
    x64 = T.gen_sinusoid(N, N, seed=SEED, noise=1e-3)
    T.round_to(x64, "fp16")   # the FP16 input; the FP64 oracle runs on the same values
    for ring in (x64[:2], x64[-2:], x64[:, :2], x64[:, -2:]):
        ring.zero_()          # zero Dirichlet ring of width 2

    # FP64 oracle after 1 and 10 sweeps (accuracy only, outside any timed region)
    refs = {}
    a = x64.clone()
    b = x64.clone()
    for s in range(1, 11):
        T.oracle_sweep_f64(STENCIL, a, B.astype(np.float64), b)
        a, b = b, a
        if s in (1, 10):
            refs[s] = a.clone()
    del a, b
    torch.cuda.synchronize()

    results = {}
    forms = (("mfma", "mfma", {}), ("mfma_hi_lo", "mfma", {"coef_lo": True}), ("direct", "direct", {}),
             ("direct_dot2", "direct_dot2", {}), ("mfma_fused8", "mfma", {"fuse": 8}),
             ("mfma_hi_lo_fused8", "mfma", {"coef_lo": True, "fuse": 8}))
    for name, method, kw in forms:
        p = T.Problem(STENCIL, N, N, B, method=method, precision="fp16", **kw)
        fused = name.endswith("fused8")
        # fused sweeps are bitwise the unfused ones (tests/test_gpu_parity.py)
        acc = results[name[:-len("_fused8")]]["acc"] if fused else accuracy(torch, T, p, x64, refs)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x64, b0)
        steps = args.steps - args.steps % 8 if fused else args.steps
        ms = time_sweeps(torch, p, b0, b1, steps, args.warmup)
        gst = p.interior_points / (ms * 1e-3) / 1e9
        kernel = p.fused_kernel_name if fused else p.kernel_name
        results[name] = {"ms": ms, "gstencil": gst, "kernel": kernel, "tile": p.tile,
                         "mfma_per_brick": p.mfma_per_brick, "acc": acc}
        log(f"bench: {name:17s} {kernel}: {ms * 1e3:.1f} us/sweep, {gst:.1f} GStencil/s, "
            f"rel-L2 1 sweep {acc[1]['rel_l2']:.3e}, 10 sweeps {acc[10]['rel_l2']:.3e}")
        del b0, b1
    del x64, refs

    m = results["mfma"]
    pts = (N - 4) ** 2
    achieved = BYTES_PER_UPDATE * pts / (m["ms"] * 1e-3) / 1e9
    bricks = (N // 16) ** 2
    mfma_tflops = m["mfma_per_brick"] * bricks * (16 * 16 * 32 * 2) / (m["ms"] * 1e-3) / 1e12
    traffic = read_traffic(m["kernel"])

    cpu = cpu_mt = None
    if not args.no_cpu:
        log("bench: timing the CPU gold (two full sweeps, single thread) ...")
        cpu = cpu_baseline(np)
        log(f"bench: CPU {cpu['value']} GStencil/s ({cpu['sample']})")
        cpu_mt = cpu_baseline_threads(np, min(16, os.cpu_count() or 1))
        log(f"bench: CPU, all host cores {cpu_mt['value']} GStencil/s ({cpu_mt['sample']})")

    line = {
        "metric": METRIC,
        "value": round(m["gstencil"], 2),
        "unit": "GStencil/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp16 storage, fp32 accumulate",
        "data": "synthetic: 4 seeded sinusoids + U[-1e-3,1e-3) noise, rounded to FP16; "
                "coefficients U[0,1) normalised to sum 1",
        "config": {"workload": f"{STENCIL} {N}x{N} FP16 Jacobi sweep (zero Dirichlet ring width 2), "
                               "one step = one sweep, MFMA 16x16x32 f16 brick kernel",
                   "stencil": STENCIL, "grid": [N, N], "precision": "fp16", "method": "mfma",
                   "kernel": m["kernel"], "tile_bricks": list(m["tile"])},
        "hbm_frac": round(achieved / HBM_PEAK_GBS, 4),
        "direct_fma": form_record(results["direct"], "FP32 taps, one FMA per tap (baseline/kernel_* counterpart)"),
        "direct_dot2": form_record(results["direct_dot2"], "FP16 taps, v_dot2_f32_f16 per tap pair, FP32 accumulate"),
        "mfma_hi_lo": form_record(results["mfma_hi_lo"], "MFMA with hi+lo FP16 taps (FP32-tap accuracy)"),
        "mfma_fused8": form_record(results["mfma_fused8"], "MFMA, 8 sweeps fused per launch in LDS "
                                                            "(bitwise the unfused sweeps; 0.5 B/update of HBM)"),
        "mfma_hi_lo_fused8": form_record(results["mfma_hi_lo_fused8"], "MFMA with hi+lo FP16 taps, 8 sweeps fused "
                                                                        "per launch (bitwise the unfused hi+lo sweeps)"),
        "mfma_vs_direct": round(m["gstencil"] / results["direct"]["gstencil"], 3),
        # the same FP16 taps, and the fastest direct form measured
        "mfma_vs_direct_dot2": round(m["gstencil"] / results["direct_dot2"]["gstencil"], 3),
        "mfma_vs_best_direct": round(m["gstencil"] / max(results["direct"]["gstencil"],
                                                          results["direct_dot2"]["gstencil"]), 3),
        # equal accuracy class: FP32-exact taps on both sides
        "mfma_hi_lo_vs_direct": round(results["mfma_hi_lo"]["gstencil"] / results["direct"]["gstencil"], 3),
        "accuracy_vs_fp64": {"rel_l2_1sweep": m["acc"][1]["rel_l2"], "max_abs_1sweep": m["acc"][1]["max_abs"],
                             "rel_l2_10sweeps": m["acc"][10]["rel_l2"],
                             "max_abs_10sweeps": m["acc"][10]["max_abs"]},
        "mfma_util": {"issued_tflops": round(mfma_tflops, 1), "peak_tflops": FP16_DENSE_PEAK_TFLOPS,
                      "frac": round(mfma_tflops / FP16_DENSE_PEAK_TFLOPS, 4),
                      "useful_tflops": round(USEFUL_FLOPS_PER_POINT * pts / (m["ms"] * 1e-3) / 1e12, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": TRAFFIC_PROFILE if traffic is not None else None},
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_mt,
        "gpu_clocks": "default power management, not locked (the reference locks clocks, README.md:7)",
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
