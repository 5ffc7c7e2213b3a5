"""Tile autotuner: time every offered workgroup tile for a problem and keep the best.

The reference fixes its tile at compile time (TILE_SIZE, common/param.hpp:11-13)
and sweeps it offline (script/param.py:5, 1..15 bricks per block).  Here the
tile is a runtime field of the problem descriptor, so the tuner times each
candidate in-process with HIP events on the same data and returns the ranking:

  python -m tcstencil_amd.autotune --stencils box2d2r --sizes 16384 --methods mfma direct
prints one JSON line per (stencil, method, precision, size) with the per-tile
times and the winner.  `best_config()` is the library call.
"""
from __future__ import annotations

import argparse
import itertools
import json
import sys

TILES = [(2, 4), (4, 4), (2, 8), (4, 8), (8, 8)]


def tune(stencil, rows, cols, method="mfma", precision="fp16", sweeps=20, warmup=4, seed=1,
         tiles=None, wgs=(0,), fuse=1):
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
                      fuse=fuse)
        b0, b1 = p.alloc(), p.alloc()
        p.clear(b1)
        p.pack(x, b0, e)
        p.run(b0, b1, warmup)
        torch.cuda.synchronize()
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


# ------------------------------------------------------------ batched grids
# Size buckets of a batch, by the smaller brick dimension of a grid.
BUCKETS = ((0, 16), (16, 64), (64, 1 << 30))


def bucket_of(rows, cols):
    m = min((rows + 15) // 16, (cols + 15) // 16)
    return next(i for i, (lo, hi) in enumerate(BUCKETS) if lo <= m < hi)


def max_fuse(stencil):
    r = 1 if stencil in ("star2d1r", "box2d1r") else 2
    return 1 + 16 // r


def build_batch(specs, choice, precision="fp16"):
    """Problems, buffers and tcs_batch for specs = [(stencil, rows, cols, B, x)],
    choice(i) -> (tile, frag, fuse, wg_per_cu); x (device, row-major) is packed
    into each grid's first buffer."""
    from . import api as T
    probs, b0s, b1s = [], [], []
    for i, (stencil, rows, cols, B, x) in enumerate(specs):
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
    g.launch()
    torch.cuda.synchronize()
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


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m tcstencil_amd.autotune")
    ap.add_argument("--stencils", nargs="+", default=["star2d1r", "star2d2r", "box2d1r", "box2d2r"])
    ap.add_argument("--methods", nargs="+", default=["mfma", "direct"])
    ap.add_argument("--precisions", nargs="+", default=["fp16"])
    ap.add_argument("--sizes", type=int, nargs="+", default=[16384])
    ap.add_argument("--sweeps", type=int, default=20)
    ap.add_argument("--wgs", type=int, nargs="+", default=[0])
    ap.add_argument("--fuse", type=int, nargs="+", default=[1])
    a = ap.parse_args(argv)
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
