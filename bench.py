#!/usr/bin/env python3
"""Benchmark: box2d2r on a 16384 x 16384 FP16 grid, MFMA vs direct-FMA.

One step = one Jacobi sweep of the whole grid (ping-pong between two brick
buffers resident in HBM).  `value` is the MFMA kernel's GStencil/s =
(N-4)^2 * steps / t / 1e9 with t measured by HIP events on the launch stream
around exactly `steps` sweeps.  The same line reports the direct-FMA kernel on
the same input, the accuracy of both against the FP64 oracle (after 1 and 10
sweeps), the HBM roofline of the MFMA kernel and the reference's CPU gold timed
on this host.

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


def read_traffic():
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if any."""
    path = os.path.join(ROOT, "profiles", "box2d2r_16384_fp16_mfma_pmc.md")
    if not os.path.exists(path):
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


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
    ap.add_argument("--warmup", type=int, default=50)
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
    for method in ("mfma", "direct"):
        p = T.Problem(STENCIL, N, N, B, method=method, precision="fp16")
        acc = accuracy(torch, T, p, x64, refs)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x64, b0)
        ms = time_sweeps(torch, p, b0, b1, args.steps, args.warmup)
        gst = p.interior_points / (ms * 1e-3) / 1e9
        results[method] = {"ms": ms, "gstencil": gst, "kernel": p.kernel_name, "tile": p.tile,
                           "mfma_per_brick": p.mfma_per_brick, "acc": acc}
        log(f"bench: {method:6s} {p.kernel_name}: {ms * 1e3:.1f} us/sweep, {gst:.1f} GStencil/s, "
            f"rel-L2 1 sweep {acc[1]['rel_l2']:.3e}, 10 sweeps {acc[10]['rel_l2']:.3e}")
        del b0, b1
    del x64, refs

    m = results["mfma"]
    pts = (N - 4) ** 2
    achieved = BYTES_PER_UPDATE * pts / (m["ms"] * 1e-3) / 1e9
    bricks = (N // 16) ** 2
    mfma_tflops = m["mfma_per_brick"] * bricks * (16 * 16 * 32 * 2) / (m["ms"] * 1e-3) / 1e12
    traffic = read_traffic()

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
        "direct_fma": {"value": round(results["direct"]["gstencil"], 2), "unit": "GStencil/s",
                       "ms_per_step": round(results["direct"]["ms"], 5), "kernel": results["direct"]["kernel"],
                       "rel_l2_1sweep": results["direct"]["acc"][1]["rel_l2"],
                       "max_abs_1sweep": results["direct"]["acc"][1]["max_abs"],
                       "rel_l2_10sweeps": results["direct"]["acc"][10]["rel_l2"]},
        "mfma_vs_direct": round(m["gstencil"] / results["direct"]["gstencil"], 3),
        "accuracy_vs_fp64": {"rel_l2_1sweep": m["acc"][1]["rel_l2"], "max_abs_1sweep": m["acc"][1]["max_abs"],
                             "rel_l2_10sweeps": m["acc"][10]["rel_l2"],
                             "max_abs_10sweeps": m["acc"][10]["max_abs"]},
        "mfma_util": {"issued_tflops": round(mfma_tflops, 1), "peak_tflops": FP16_DENSE_PEAK_TFLOPS,
                      "frac": round(mfma_tflops / FP16_DENSE_PEAK_TFLOPS, 4),
                      "useful_tflops": round(USEFUL_FLOPS_PER_POINT * pts / (m["ms"] * 1e-3) / 1e12, 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic},
        "cpu_baseline": cpu,
        "cpu_baseline_threads": cpu_mt,
        "gpu_clocks": "default power management, not locked (the reference locks clocks, README.md:7)",
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
