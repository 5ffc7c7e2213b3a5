"""Measure the five benchmark configurations of the stencil path on one GPU.

  python -m tcstencil_amd.configs [c1 c2 c3 c4 c5]

prints one JSON line per measurement.  Throughput is interior updates per
second, (rows-2r)(cols-2r) * sweeps / t, with t from HIP events around the
sweep loop (inputs resident, after warm-up).  Accuracy is max-abs and
relative-L2 against the library's FP64 oracle sweep on the GPU, run on the
same input values (the FP16/BF16 inputs are FP16/BF16-representable; the SPLIT
input is the FP32 or FP64 grid).  HBM fractions use the algorithmic bytes per
update (4 B FP16/BF16, 8 B SPLIT, divided by the sweeps fused per launch)
against 8 TB/s.

  c1  star2d1r 1024^2, FP16 U[-1,1), one sweep, MFMA and direct
  c2  box2d1r 10240^2, FP16 U[-1,1), 100 sweeps, MFMA vs direct, errors after 1/10/100;
      the loop also fused 8/16 per launch and captured as one hipGraph
  c3  star2d2r and box2d2r 16384^2, 4 sinusoids + noise, 10 sweeps, FP16/BF16/SPLIT MFMA
  c4  512 batched uneven grids (box2d1r/star2d2r, 64..2048 per side), 20 sweeps in a
      hipGraph: tile, MFMA fragment, sweeps per launch and waves per CU autotuned per
      size bucket, vs a fixed tile rule at one sweep per launch and vs one launch per grid
  c5  box2d2r 32768x16384, 256 sweeps, SPLIT, 2/4/8 sweeps fused vs unfused MFMA and
      direct; rel-L2 after 256 sweeps
"""
from __future__ import annotations

import json
import math
import sys

import numpy as np

from .timing import warm_chunk, warm_up

HBM_PEAK_GBS = 8000.0
BYTES = {"fp16": 4.0, "bf16": 4.0, "split": 8.0}
SEED = 1


def emit(**kw):
    print(json.dumps(kw), flush=True)


def _torch():
    import torch
    return torch


def _interior(stencil, rows, cols):
    from . import api as T
    r = T.radius(stencil)
    return (rows - 2 * r) * (cols - 2 * r)


def _zero_ring(x, r):
    for v in (x[:r], x[-r:], x[:, :r], x[:, -r:]):
        v.zero_()


def _fp64_refs(stencil, x64, B, counts):
    """FP64 oracle results after each sweep count in `counts`."""
    from . import api as T
    torch = _torch()
    a, b = x64.clone(), x64.clone()
    out, done = {}, 0
    for k in sorted(counts):
        while done < k:
            T.oracle_sweep_f64(stencil, a, B.astype(np.float64), b)
            a, b = b, a
            done += 1
        out[k] = a.clone()
    del a, b
    torch.cuda.synchronize()
    return out


def _timed_run(p, b0, b1, sweeps, warmup=2):
    torch = _torch()
    chunk = warm_chunk(warmup, p.fuse)
    warm_up(lambda: p.run(b0, b1, chunk))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    res = p.run(b0, b1, sweeps)
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1), res


def _errors(p, x, refs, scale_exp=0):
    """Errors of the problem's result after each count in refs (fresh runs)."""
    from . import api as T
    b0, b1 = p.alloc(), p.alloc()
    out = {}
    for k in sorted(refs):
        p.clear(b1)
        p.pack(x, b0, scale_exp)
        y = p.unpack(p.run(b0, b1, k), scale_exp=scale_exp)
        out[k] = T.error_norms(y, refs[k], p.radius)
    return out


def _timed_graph(p, b0, b1, sweeps):
    """ms of one replay of the sweep loop captured as a hipGraph (after untimed
    replays for the warm-up time)."""
    torch = _torch()
    g = p.graph(b0, b1, sweeps)
    warm_up(g.launch)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    g.launch()
    e1.record()
    torch.cuda.synchronize()
    g.destroy()
    return e0.elapsed_time(e1)


def _measure(stencil, rows, cols, B, x, method, precision, sweeps, fuse=1, scale_exp=0, tile=None, graph=False,
             coef_lo=False):
    from . import api as T
    p = T.Problem(stencil, rows, cols, B, method=method, precision=precision, fuse=fuse, tile=tile,
                  coef_lo=coef_lo)
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(x, b0, scale_exp)
    ms = _timed_graph(p, b0, b1, sweeps) if graph else _timed_run(p, b0, b1, sweeps)[0]
    pts = p.interior_points * sweeps
    gst = pts / (ms * 1e-3) / 1e9
    return p, {"stencil": stencil, "grid": [rows, cols], "method": method, "precision": precision,
               "coef_lo": bool(coef_lo), "fuse": fuse, "tile": list(p.tile), "sweeps": sweeps, "graph": graph,
               "ms": round(ms, 3),
               "gstencil_s": round(gst, 1),
               "hbm_gbs": round(gst * BYTES[precision] / fuse, 1),
               "hbm_frac": round(gst * BYTES[precision] / fuse / HBM_PEAK_GBS, 4),
               "kernel": p.fused_kernel_name if fuse > 1 else p.kernel_name}


def c1():
    from . import api as T
    torch = _torch()
    stencil, N = "star2d1r", 1024
    B = T.gen_coeffs(stencil, SEED)
    x = T.gen_uniform(N, N, seed=SEED, dtype=torch.float64)
    T.round_to(x, "fp16")
    _zero_ring(x, 1)
    refs = _fp64_refs(stencil, x, B, [1])
    for method in ("mfma", "direct"):
        p, rec = _measure(stencil, N, N, B, x, method, "fp16", 200)
        rec["errors_vs_fp64"] = {1: _errors(p, x, refs)[1]}
        emit(config="c1", **rec)
    # A 1024^2 sweep is a few microseconds of work: launch-bound on a stream.
    # The same 200 sweeps as one hipGraph replay, and fused 9 to a launch.
    for method in ("mfma", "direct"):
        emit(config="c1", **_measure(stencil, N, N, B, x, method, "fp16", 200, graph=True)[1])
    emit(config="c1", **_measure(stencil, N, N, B, x, "mfma", "fp16", 198, fuse=9, graph=True)[1])


def c2():
    from . import api as T
    torch = _torch()
    stencil, N = "box2d1r", 10240
    B = T.gen_coeffs(stencil, SEED)
    x = T.gen_uniform(N, N, seed=SEED, dtype=torch.float64)
    T.round_to(x, "fp16")
    _zero_ring(x, 1)
    refs = _fp64_refs(stencil, x, B, [1, 10, 100])
    res = {}
    for method, lo in (("mfma", False), ("mfma", True), ("direct", False), ("direct_dot2", False)):
        p, rec = _measure(stencil, N, N, B, x, method, "fp16", 100, coef_lo=lo)
        rec["errors_vs_fp64"] = _errors(p, x, refs)
        res[method + ("_hi_lo" if lo else "")] = rec["gstencil_s"]
        emit(config="c2", **rec)
    for fuse in (8, 16):
        p, rec = _measure(stencil, N, N, B, x, "mfma", "fp16", 100, fuse=fuse)
        emit(config="c2", **rec)
    # the 100-sweep loop captured once as a hipGraph: no launch gaps between sweeps
    for method, fuse in (("mfma", 1), ("direct", 1), ("mfma", 16)):
        p, rec = _measure(stencil, N, N, B, x, method, "fp16", 100, fuse=fuse, graph=True)
        emit(config="c2", **rec)
    emit(config="c2", mfma_vs_direct=round(res["mfma"] / res["direct"], 3),
         mfma_vs_direct_dot2=round(res["mfma"] / res["direct_dot2"], 3),
         mfma_hi_lo_vs_direct=round(res["mfma_hi_lo"] / res["direct"], 3))


def c3():
    from . import api as T
    torch = _torch()
    N = 16384
    for stencil in ("star2d2r", "box2d2r"):
        B = T.gen_coeffs(stencil, SEED)
        x64 = T.gen_sinusoid(N, N, seed=SEED, noise=1e-3)
        _zero_ring(x64, 2)
        for precision, lo in (("fp16", False), ("fp16", True), ("bf16", False), ("bf16", True), ("split", False)):
            x = x64.clone()
            e = 0
            if precision == "split":
                e = T.split_scale_exp(T.maxabs(x))
            else:
                T.round_to(x, precision)
            refs = _fp64_refs(stencil, x, B, [1, 10])
            p, rec = _measure(stencil, N, N, B, x, "mfma", precision, 10, scale_exp=e, coef_lo=lo)
            rec["errors_vs_fp64"] = _errors(p, x, refs, e)
            rec["input"] = "4 sinusoids + U[-1e-3,1e-3)" + ("" if precision == "split" else f", rounded to {precision}")
            emit(config="c3", **rec)
            del x, refs
        for method in ("direct", "direct_dot2"):
            p, rec = _measure(stencil, N, N, B, x64, method, "fp16", 10)
            emit(config="c3", **rec)
        del x64


def c4(nprob=512, sweeps=20):
    from . import api as T
    from . import autotune as A
    torch = _torch()
    rng = np.random.default_rng(SEED)
    kinds = rng.integers(0, 2, nprob)
    rows = np.exp(rng.uniform(math.log(64), math.log(2048), nprob)).astype(int)
    cols = np.exp(rng.uniform(math.log(64), math.log(2048), nprob)).astype(int)
    specs = []
    for i in range(nprob):
        stencil = ("box2d1r", "star2d2r")[kinds[i]]
        x = T.gen_uniform(int(rows[i]), int(cols[i]), seed=SEED + i, dtype=torch.float32)
        specs.append((stencil, int(rows[i]), int(cols[i]), T.gen_coeffs(stencil, SEED + i), x))
    bucket = [A.bucket_of(r, c) for _, r, c, _, _ in specs]

    def timed(choice):
        probs, b0s, b1s, batch = A.build_batch(specs, choice)
        ms = A.time_batch(batch, sweeps, reps=5)
        rec = (batch.interior_points * sweeps / (ms * 1e-3) / 1e9, ms, batch.launches_per_sweep)
        batch.destroy()
        return rec, probs, b0s, b1s

    # the autotuner's choice per size bucket (tile, fragment, sweeps per launch, waves per CU)
    best, log = A.tune_batch(specs, sweeps)
    (gst_t, ms_t, nl_t), _, _, _ = timed(lambda i: best[bucket[i]])
    # fixed rule, one sweep per launch
    fixed = {0: ((2, 4), 16, 1, 0), 1: ((4, 4), 16, 1, 0), 2: ((4, 8), 16, 1, 0)}
    (gst_f, ms_f, nl_f), probs, b0s, b1s = timed(lambda i: fixed[bucket[i]])
    # one launch per grid per sweep, the same work
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    warm_up(lambda: [p.run(b0, b1, 2) for p, b0, b1 in zip(probs, b0s, b1s)])
    e0.record()
    for p, b0, b1 in zip(probs, b0s, b1s):
        p.run(b0, b1, sweeps)
    e1.record()
    torch.cuda.synchronize()
    ms_each = e0.elapsed_time(e1)
    pts = sum((r - 2 * (1 if s == "box2d1r" else 2)) * (c - 2 * (1 if s == "box2d1r" else 2))
              for s, r, c, _, _ in specs) * sweeps
    gst_e = pts / (ms_each * 1e-3) / 1e9
    emit(config="c4", problems=nprob, sweeps=sweeps, interior_points_per_sweep=pts // sweeps,
         buckets={str(A.BUCKETS[b]): {"grids": bucket.count(b), "tile": list(t), "frag": f, "fuse": k, "wg_per_cu": w}
                  for b, (t, f, k, w) in sorted(best.items())},
         tuned_graph_ms=round(ms_t, 3), tuned_gstencil_s=round(gst_t, 1), tuned_launches_per_sweep=nl_t,
         fixed_graph_ms=round(ms_f, 3), fixed_gstencil_s=round(gst_f, 1), fixed_launches_per_sweep=nl_f,
         fixed_hbm_frac=round(gst_f * 4 / HBM_PEAK_GBS, 4),
         one_launch_per_grid_ms=round(ms_each, 3), one_launch_per_grid_gstencil_s=round(gst_e, 1),
         speedup_tuned=round(ms_each / ms_t, 2), speedup_fixed=round(ms_each / ms_f, 2),
         trials={str(A.BUCKETS[b]): [["%dx%d" % c[0], c[1], c[2], c[3], round(ms, 3)] for c, ms in tr]
                 for b, tr in log.items()})


def c5(sweeps=256):
    from . import api as T
    torch = _torch()
    stencil, R, C = "box2d2r", 32768, 16384
    B = T.gen_coeffs(stencil, SEED)
    x = T.gen_uniform(R, C, seed=SEED, dtype=torch.float64)
    _zero_ring(x, 2)
    e = T.split_scale_exp(T.maxabs(x))
    ref = _fp64_refs(stencil, x, B, [sweeps])[sweeps]
    recs = []
    for fuse in (1, 2, 4, 8):
        best = None
        for tile in ((4, 4), (2, 8), (4, 8)):  # the tile is tuned per sweeps-per-launch count
            p, rec = _measure(stencil, R, C, B, x, "mfma", "split", sweeps, fuse=fuse, scale_exp=e, tile=tile)
            if best is None or rec["ms"] < best[1]["ms"]:
                best = (p, rec)
        p, rec = best
        if fuse in (1, 8):
            rec["errors_vs_fp64"] = {sweeps: _errors(p, x, {sweeps: ref}, e)[sweeps]}
        recs.append(rec)
        emit(config="c5", **rec)
    x16 = x.clone()
    T.round_to(x16, "fp16")
    for method in ("mfma", "direct", "direct_dot2"):
        p, rec = _measure(stencil, R, C, B, x16, method, "fp16", 32)
        emit(config="c5", reference_point=True, **rec)


def main(argv=None):
    names = (argv if argv is not None else sys.argv[1:]) or ["c1", "c2", "c3", "c4", "c5"]
    torch = _torch()
    torch.cuda.set_device(0)
    for n in names:
        globals()[n]()
        torch.cuda.empty_cache()
    return 0


if __name__ == "__main__":
    sys.exit(main())
