"""Profiling harness: one workload per process, optionally under rocprofv3.

Replaces the reference's sweep pipeline -- compile one binary per
(stencil, method, mesh size, tile size) (script/nvcc_compile.py:12-27,
FLAG_DICT script/param.py:1-6), run each under `ncu` and scrape "Duration"
(script/run_file_with_ncu.sh:5), then turn it into CSV
(script/process_data.py:46-57) -- with a single library and runtime parameters:

  python -m tcstencil_amd.harness run --stencil box2d2r --size 16384 --method mfma --sweeps 20
      one workload: pack, warm up (runs of --warmup sweeps, rounded up to whole
      fused launches, until 100 ms have passed: tcstencil_amd/timing.py), time
      `sweeps` back-to-back sweeps with HIP events; --copy-mb adds a streaming copy of known size (the bandwidth
      ceiling, and the calibration kernel for FETCH_SIZE / WRITE_SIZE).
      Prints one JSON line.

  python -m tcstencil_amd.harness edges --stencil box2d2r --size 16384 --sweeps 1
      row-major FP32 in -> sweeps -> row-major FP32 out: tcs_run_f32 (pack and
      unpack fused into the first and last sweep) against the three steps
      pack / run / unpack, HIP-event time per call.  Prints one JSON line.

  python -m tcstencil_amd.harness sweep --sizes 1024 4096 16384 --out results.csv [--rocprof]
      every (stencil, method, size, tile) in the cross product; with --rocprof
      each workload runs in its own `rocprofv3 --kernel-trace --stats` process
      and the kernel's average duration comes from the profiler's stats file,
      otherwise from HIP events.  The CSV keeps the reference's columns
      `type,stencil,mesh size,tile size,time,unit` (process_data.py:14) and adds
      GStencil/s (interior updates, (N-2r)^2 / t), the reference's figure rate
      N^2 / t (draw.py:19) in G points/s, and HBM GB/s.
"""
from __future__ import annotations

import argparse
import csv
import glob
import itertools
import json
import os
import subprocess
import sys
import tempfile

from .timing import warm_chunk, warm_up

HBM_PEAK_GBS = 8000.0
BYTES_PER_UPDATE = {"fp16": 4.0, "bf16": 4.0, "split": 8.0}  # per sweep, unfused; fused: / fuse
TILES = [(2, 4), (4, 4), (2, 8), (4, 8), (8, 8)]


def run_workload(stencil, rows, cols, method="mfma", precision="fp16", tile=None, sweeps=20,
                 warmup=5, copy_mb=0, seed=1, fuse=1, wg_per_cu=0, frag=16, coef_lo=False):
    import numpy as np
    import torch

    from . import api as T

    B = T.gen_coeffs(stencil, seed)
    p = T.Problem(stencil, rows, cols, B, method=method, precision=precision, tile=tile, fuse=fuse,
                  wg_per_cu=wg_per_cu, frag=frag, coef_lo=coef_lo)
    x = T.gen_uniform(rows, cols, seed=seed, dtype=torch.float32)
    e = T.split_scale_exp(1.0) if precision == "split" else 0
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(x, b0, e)
    del x
    chunk = warm_chunk(warmup, fuse)
    warm_up(lambda: p.run(b0, b1, chunk))
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    p.run(b0, b1, sweeps)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / sweeps
    pts = p.interior_points
    out = {"stencil": stencil, "rows": rows, "cols": cols, "method": method, "precision": precision,
           "tile": list(p.tile), "fuse": fuse, "frag": p.frag, "coef_lo": p.coef_lo,
           "kernel": p.fused_kernel_name if fuse > 1 else p.kernel_name, "sweeps": sweeps, "ms_per_sweep": ms,
           "gstencil_s": pts / (ms * 1e-3) / 1e9,
           "hbm_gbs": BYTES_PER_UPDATE[precision] / max(fuse, 1) * pts / (ms * 1e-3) / 1e9}
    out["hbm_frac"] = out["hbm_gbs"] / HBM_PEAK_GBS
    if copy_mb:
        n = int(copy_mb * 2 ** 20) // 16 * 16
        src = torch.empty(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1)
        for _ in range(3):
            T.copy(dst, src)
        torch.cuda.synchronize()
        reps = 10
        ev[0].record()
        for _ in range(reps):
            T.copy(dst, src)
        ev[1].record()
        torch.cuda.synchronize()
        cms = ev[0].elapsed_time(ev[1]) / reps
        out["copy"] = {"bytes_per_launch": 2 * n, "ms": cms, "gbs": 2 * n / (cms * 1e-3) / 1e9}
    del b0, b1
    return out


def run_edges(stencil, rows, cols, precision="fp16", sweeps=1, reps=20, seed=1, tile=None):
    """ms per call of tcs_run_f32 (fused edges) and of pack + run + unpack on
    the same row-major FP32 grid, and whether their outputs are bitwise equal."""
    import torch

    from . import api as T

    B = T.gen_coeffs(stencil, seed)
    p = T.Problem(stencil, rows, cols, B, precision=precision, tile=tuple(tile) if tile else None)
    x = T.gen_uniform(rows, cols, seed=seed, dtype=torch.float32)
    y0, y1 = torch.empty_like(x), torch.empty_like(x)
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b0)
    p.clear(b1)

    def fused():
        p.run_f32(x, sweeps, out=y0, buf0=b0, buf1=b1)

    def three():
        p.pack(x, b0)
        p.unpack(p.run(b0, b1, sweeps), out=y1, ring=x)

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    res = {"stencil": stencil, "rows": rows, "cols": cols, "precision": precision, "sweeps": sweeps,
           "tile": list(p.tile)}
    for name, fn in (("fused_edges", fused), ("three_steps", three)):
        warm_up(fn)
        ev[0].record()
        for _ in range(reps):
            fn()
        ev[1].record()
        torch.cuda.synchronize()
        ms = ev[0].elapsed_time(ev[1]) / reps
        res[name] = {"ms": ms, "gstencil_s": p.interior_points * sweeps / (ms * 1e-3) / 1e9,
                     "fp32_io_gbs": 8 * rows * cols / (ms * 1e-3) / 1e9}
    res["bitwise_equal"] = bool(torch.equal(y0, y1))
    res["speedup"] = res["three_steps"]["ms"] / res["fused_edges"]["ms"]
    return res


def rocprof_kernel_avg_ns(cmd, kernel_prefix, workdir):
    """Run `cmd` under rocprofv3 --kernel-trace --stats; return (average ns, calls)
    of the kernels whose name contains kernel_prefix."""
    full = ["rocprofv3", "--kernel-trace", "--stats", "--output-format", "csv", "-d", workdir, "-o", "run",
            "--"] + cmd
    subprocess.run(full, check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    files = glob.glob(os.path.join(workdir, "**", "*kernel_stats.csv"), recursive=True)
    if not files:
        raise RuntimeError("rocprofv3 wrote no kernel_stats.csv")
    total, calls = 0.0, 0
    with open(files[0]) as f:
        for row in csv.DictReader(f):
            if kernel_prefix in row["Name"]:
                total += float(row["TotalDurationNs"])
                calls += int(row["Calls"])
    if calls == 0:
        raise RuntimeError(f"no kernel matching {kernel_prefix}")
    return total / calls, calls


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m tcstencil_amd.harness")
    sub = ap.add_subparsers(dest="cmd", required=True)
    r = sub.add_parser("run")
    r.add_argument("--stencil", default="box2d2r")
    r.add_argument("--size", type=int, default=16384)
    r.add_argument("--cols", type=int, default=0)
    r.add_argument("--method", default="mfma")
    r.add_argument("--precision", default="fp16")
    r.add_argument("--tile", type=int, nargs=2, default=None)
    r.add_argument("--sweeps", type=int, default=20)
    r.add_argument("--warmup", type=int, default=5, help="sweeps per warm-up run (repeated for 100 ms)")
    r.add_argument("--copy-mb", type=float, default=0)
    r.add_argument("--fuse", type=int, default=1)
    r.add_argument("--wg-per-cu", type=int, default=0)
    r.add_argument("--frag", type=int, default=16, choices=(16, 32))
    r.add_argument("--coef-lo", action="store_true", help="carry every tap as a hi+lo pair (FP16/BF16 MFMA)")
    e = sub.add_parser("edges")
    e.add_argument("--stencil", default="box2d2r")
    e.add_argument("--size", type=int, default=16384)
    e.add_argument("--cols", type=int, default=0)
    e.add_argument("--precision", default="fp16")
    e.add_argument("--sweeps", type=int, default=1)
    e.add_argument("--reps", type=int, default=20)
    e.add_argument("--tile", type=int, nargs=2, default=None)
    s = sub.add_parser("sweep")
    s.add_argument("--stencils", nargs="+", default=["star2d1r", "star2d2r", "box2d1r", "box2d2r"])
    s.add_argument("--methods", nargs="+", default=["mfma", "direct"])
    s.add_argument("--precision", default="fp16")
    s.add_argument("--sizes", type=int, nargs="+", default=[1024, 4096, 16384])
    s.add_argument("--reference-meshes", action="store_true",
                   help="the reference's mesh sweep, 160..7200 step 160 (script/param.py:4)")
    s.add_argument("--tiles", nargs="+", default=["%dx%d" % t for t in TILES])
    s.add_argument("--sweeps", type=int, default=20)
    s.add_argument("--rocprof", action="store_true")
    s.add_argument("--out", default="")
    a = ap.parse_args(argv)

    if a.cmd == "run":
        res = run_workload(a.stencil, a.size, a.cols or a.size, a.method, a.precision,
                           tuple(a.tile) if a.tile else None, a.sweeps, a.warmup, a.copy_mb,
                           fuse=a.fuse, wg_per_cu=a.wg_per_cu, frag=a.frag, coef_lo=a.coef_lo)
        print(json.dumps(res), flush=True)
        return 0

    if a.cmd == "edges":
        print(json.dumps(run_edges(a.stencil, a.size, a.cols or a.size, a.precision, a.sweeps, a.reps,
                                   tile=a.tile)), flush=True)
        return 0

    rows = []
    sizes = [160 * i for i in range(1, 46)] if a.reference_meshes else a.sizes
    for stencil, method, size, tile in itertools.product(a.stencils, a.methods, sizes, a.tiles):
        th, tw = (int(v) for v in tile.split("x"))
        if a.rocprof:
            cmd = [sys.executable, "-m", "tcstencil_amd.harness", "run", "--stencil", stencil,
                   "--size", str(size), "--method", method, "--precision", a.precision,
                   "--tile", str(th), str(tw), "--sweeps", str(a.sweeps)]
            prefix = "tcs_mfma_sweep" if method == "mfma" else "tcs_direct_sweep"
            with tempfile.TemporaryDirectory() as d:
                ns, _ = rocprof_kernel_avg_ns(cmd, prefix, d)
            ms = ns * 1e-6
        else:
            ms = run_workload(stencil, size, size, method, a.precision, (th, tw), a.sweeps)["ms_per_sweep"]
        r = 1 if stencil in ("star2d1r", "box2d1r") else 2
        pts = (size - 2 * r) ** 2
        row = {"type": "tensor" if method == "mfma" else "baseline", "stencil": stencil,
               "mesh size": size, "tile size": f"{th}x{tw}", "time": round(ms * 1e3, 3), "unit": "us",
               "gstencil_s": round(pts / (ms * 1e-3) / 1e9, 1),
               # the reference's figure rate, N^2 / t (script/draw.py:19)
               "gpts_n2_s": round(size * size / (ms * 1e-3) / 1e9, 1),
               "hbm_gbs": round(BYTES_PER_UPDATE[a.precision] * pts / (ms * 1e-3) / 1e9, 1)}
        rows.append(row)
        print(json.dumps(row), flush=True)
    if a.out:
        with open(a.out, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=list(rows[0]))
            w.writeheader()
            w.writerows(rows)
    return 0


if __name__ == "__main__":
    sys.exit(main())
