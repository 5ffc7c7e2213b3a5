"""Autotuner: time the launch parameters of a problem (MFMA fragment, workgroup
tile, persistent grid, sweeps per launch) and keep the fastest in a cache file.

The reference fixes its tile at compile time (TILE_SIZE, common/param.hpp:11-13)
and sweeps it offline (script/param.py:5, 1..15 bricks per block).  Here these
are runtime fields of the problem descriptor, so the tuner times each candidate
in-process with HIP events on the same data and writes the winners to a tuning
cache (the text format of include/tcstencil.h, tcs_tune_lookup) that problem
set-up reads: tcs_problem_init applies the single-grid entry of
$TCS_TUNE_CACHE, and batched grids apply theirs with
tcs_problem_apply_tuning(..., batched=1) before their buffers are allocated.

  python -m tcstencil_amd.autotune --stencils box2d2r --sizes 16384 --methods mfma direct
prints one JSON line per (stencil, method, precision, size) with the per-tile
times and the winner;
  ... --write-cache tune.txt
also writes (or extends) the cache with one entry per size bucket.
`best_config()` and `tune_cache_entries()` are the library calls.
"""
from __future__ import annotations

import argparse
import itertools
import json
import sys

from .timing import warm_chunk, warm_up

TILES = [(2, 4), (4, 4), (2, 8), (4, 8), (8, 8)]


def tune(stencil, rows, cols, method="mfma", precision="fp16", sweeps=20, warmup=4, seed=1,
         tiles=None, wgs=(0,), fuse=1, coef_lo=False):
    """[((tile, wg_per_cu), ms_per_sweep)] sorted fastest first.  wg_per_cu is the
    number of resident workgroups per CU of the persistent grid (0 = as many as
    fit)."""
    import torch

    from . import api as T

    B = T.gen_coeffs(stencil, seed)
    x = T.gen_uniform(rows, cols, seed=seed, dtype=torch.float32)
    e = T.split_scale_exp(1.0) if precision == "split" else 0
    res = []
    for tile, wg in itertools.product(tiles or TILES, wgs):
        if fuse > 1 and wg:
            continue  # fused launches: one tile per workgroup
        p = T.Problem(stencil, rows, cols, B, method=method, precision=precision, tile=tile, wg_per_cu=wg,
                      fuse=fuse, coef_lo=coef_lo)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x, b0, e)
        chunk = warm_chunk(warmup, fuse)
        warm_up(lambda: p.run(b0, b1, chunk))
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        p.run(b0, b1, sweeps)
        ev[1].record()
        torch.cuda.synchronize()
        res.append(((tile, wg), ev[0].elapsed_time(ev[1]) / sweeps))
        del b0, b1
    return sorted(res, key=lambda t: t[1])


def best_config(stencil, rows, cols, method="mfma", precision="fp16", **kw):
    """(tile, wg_per_cu) of the fastest configuration."""
    return tune(stencil, rows, cols, method, precision, **kw)[0][0]


# ------------------------------------------------------------ tuning cache
# Size buckets (tcs_tune_bucket): the smaller side of a grid in bricks.
BUCKETS = ((0, 16), (16, 64), (64, 256), (256, 1 << 30))


def bucket_of(rows, cols):
    from ._lib import lib
    return int(lib().tcs_tune_bucket(rows, cols))


def device_record():
    """'device <arch> <CUs>' of cuda:0, the first record of a cache file."""
    import torch
    prop = torch.cuda.get_device_properties(0)
    return f"device {prop.gcnArchName.split(':')[0]} {prop.multi_processor_count}"


def entry_line(stencil, method, precision, coef_lo, batched, bucket, frag, tile, wg, fuse):
    return (f"entry {stencil} {method} {precision} {int(bool(coef_lo))} {int(bool(batched))} {bucket} "
            f"{frag} {tile[0]} {tile[1]} {wg} {fuse}")


def write_cache(path, lines, device=None):
    """Write a cache: the device record, then the entry lines (earlier entries
    of the same key win at lookup, so new lines go first and an existing file's
    entries for the same device follow)."""
    import os
    dev = device or device_record()
    old = []
    if os.path.exists(path):
        with open(path) as f:
            recs = [l.strip() for l in f if l.strip() and not l.startswith("#")]
        if recs and recs[0] == dev:
            old = recs[1:]
    with open(path, "w") as f:
        f.write("# tcstencil tuning cache (python -m tcstencil_amd.autotune --write-cache)\n")
        f.write(dev + "\n")
        for l in list(lines) + old:
            f.write(l + "\n")


def tune_cache_entries(stencil, size, method="mfma", precision="fp16", coef_lo=False, sweeps=20,
                       tiles=None, wgs=(0, 2), fuses=(1, 2, 4, 8)):
    """The cache entry for one single-grid problem of `size`^2: the fastest tile
    and persistent-grid setting at one sweep per launch, then (MFMA) the fastest
    sweeps per launch at that tile.  Returns (entry line, log)."""
    ranking = tune(stencil, size, size, method, precision, sweeps=sweeps, tiles=tiles, wgs=wgs,
                   coef_lo=coef_lo)
    (tile, wg), _ = ranking[0]
    fuse = 1
    log = {"tiles": [["%dx%d" % t, w, round(ms * 1e3, 1)] for (t, w), ms in ranking]}
    if method == "mfma":
        r = 1 if stencil in ("star2d1r", "box2d1r") else 2
        cands = [k for k in fuses if r * (k - 1) <= 16]
        fr = [(k, tune(stencil, size, size, method, precision, sweeps=sweeps * 2, tiles=[tile], wgs=(0,),
                       fuse=k, coef_lo=coef_lo)[0][1]) for k in cands]
        fuse = min(fr, key=lambda t: t[1])[0]
        if fuse > 1:
            wg = 0  # fused launches run one tile per workgroup
        log["fuse"] = [[k, round(ms * 1e3, 1)] for k, ms in fr]
    return entry_line(stencil, method, precision, coef_lo, 0, bucket_of(size, size), 16, tile, wg, fuse), log


# ------------------------------------------------------------ batched grids

def max_fuse(stencil):
    r = 1 if stencil in ("star2d1r", "box2d1r") else 2
    return 1 + 16 // r


def build_batch(specs, choice, precision="fp16", cache=None):
    """Problems, buffers and tcs_batch for specs = [(stencil, rows, cols, B, x)],
    choice(i) -> (tile, frag, fuse, wg_per_cu), or with choice None the batched
    entries of the tuning cache file `cache` ("" : $TCS_TUNE_CACHE; defaults
    where it has none); x (device, row-major) is packed into each grid's first
    buffer."""
    from . import api as T
    probs, b0s, b1s = [], [], []
    for i, (stencil, rows, cols, B, x) in enumerate(specs):
        if choice is None:
            p = T.Problem(stencil, rows, cols, B, precision=precision, tuned=(cache or "", True))
            if p.tile not in ((2, 4), (4, 4), (4, 8)):
                p.set_tile(4, 4)  # no entry: a batched tile of the default's height
        else:
            tile, frag, fuse, wg = choice(i)
            p = T.Problem(stencil, rows, cols, B, precision=precision, tile=tile, frag=frag,
                          fuse=min(fuse, max_fuse(stencil)), wg_per_cu=wg)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x, b0)
        probs.append(p), b0s.append(b0), b1s.append(b1)
    return probs, b0s, b1s, T.Batch(probs, b0s, b1s)


def time_batch(batch, sweeps, reps=3):
    """ms for `sweeps` sweeps of the batch, replayed as one hipGraph."""
    import torch
    g = batch.graph(sweeps)
    warm_up(g.launch)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        g.launch()
    ev[1].record()
    torch.cuda.synchronize()
    g.destroy()
    return ev[0].elapsed_time(ev[1]) / reps


def tune_batch(specs, sweeps=20, tiles=((2, 4), (4, 4), (4, 8)), frags=(16, 32), wgs=(0, 2),
               fuses=(1, 4, 8, 16)):
    """Per size bucket (BUCKETS), the fastest (tile, fragment, wg_per_cu, sweeps per
    launch) for a batch of grids: each bucket's grids alone are timed as a batch
    for every tile/fragment/waves-per-CU candidate at one sweep per launch, then
    for every sweeps-per-launch count and tile with the winning fragment and
    waves per CU.  Returns
    ({bucket: (tile, frag, fuse, wg)}, {bucket: [(config, ms)]})."""
    best, log = {}, {}
    for b in range(len(BUCKETS)):
        sub = [s for s in specs if bucket_of(s[1], s[2]) == b]
        if not sub:
            continue
        trials = []

        def trial(tile, frag, fuse, wg):
            if frag == 32 and tile == (2, 4):
                return
            objs = build_batch(sub, lambda i: (tile, frag, fuse, wg))
            trials.append(((tile, frag, fuse, wg), time_batch(objs[3], sweeps)))
            objs[3].destroy()

        for tile, frag, wg in itertools.product(tiles, frags, wgs):
            trial(tile, frag, 1, wg)
        _, frag, _, wg = min(trials, key=lambda t: t[1])[0]
        for fuse, tile in itertools.product(fuses, tiles):
            if fuse > 1:
                trial(tile, frag, fuse, wg)
        best[b] = min(trials, key=lambda t: t[1])[0]
        log[b] = trials
    return best, log


def batch_cache_entries(best, stencils=("box2d1r", "star2d2r"), precision="fp16"):
    """Cache lines for tune_batch's per-bucket winners, for every stencil of the
    batch (sweeps per launch capped by each stencil's radius)."""
    lines = []
    for b, (tile, frag, fuse, wg) in sorted(best.items()):
        for st in stencils:
            lines.append(entry_line(st, "mfma", precision, 0, 1, b, frag, tile, wg, min(fuse, max_fuse(st))))
    return lines


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m tcstencil_amd.autotune")
    ap.add_argument("--stencils", nargs="+", default=["star2d1r", "star2d2r", "box2d1r", "box2d2r"])
    ap.add_argument("--methods", nargs="+", default=["mfma", "direct"])
    ap.add_argument("--precisions", nargs="+", default=["fp16"])
    ap.add_argument("--sizes", type=int, nargs="+", default=[16384])
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--wgs", type=int, nargs="+", default=[0])
    ap.add_argument("--fuse", type=int, nargs="+", default=[1])
    ap.add_argument("--write-cache", default="",
                    help="tune tile, persistent grid and sweeps per launch per problem and write the winners "
                         "to this tuning cache file (one entry per stencil, method, precision and size bucket)")
    a = ap.parse_args(argv)
    if a.write_cache:
        lines = []
        for stencil, method, precision, size in itertools.product(a.stencils, a.methods, a.precisions, a.sizes):
            if method == "direct" and precision != "fp16":
                continue
            line, log = tune_cache_entries(stencil, size, method, precision, sweeps=a.sweeps,
                                           wgs=tuple(a.wgs) if a.wgs != [0] else (0, 2))
            lines.append(line)
            print(json.dumps({"entry": line, "size": size, **log}), flush=True)
        write_cache(a.write_cache, lines)
        return 0
    for stencil, method, precision, size, fuse in itertools.product(a.stencils, a.methods, a.precisions,
                                                                    a.sizes, a.fuse):
        if method == "direct" and (precision != "fp16" or fuse > 1):
            continue
        r = 1 if stencil in ("star2d1r", "box2d1r") else 2
        pts = (size - 2 * r) ** 2
        ranking = tune(stencil, size, size, method, precision, sweeps=a.sweeps, wgs=a.wgs, fuse=fuse)
        print(json.dumps({"stencil": stencil, "method": method, "precision": precision, "size": size,
                          "fuse": fuse,
                          "best": {"tile": list(ranking[0][0][0]), "wg_per_cu": ranking[0][0][1]},
                          "us": {"%dx%d/wg%d" % (t + (w,)): round(ms * 1e3, 1) for (t, w), ms in ranking},
                          "best_gstencil_s": round(pts / (ranking[0][1] * 1e-3) / 1e9, 1)}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
