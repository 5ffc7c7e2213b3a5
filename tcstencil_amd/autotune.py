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
        if fuse == 1 and precision == "split" and tile == (8, 8):
            continue  # no unfused SPLIT 8x8 tile
        if fuse > 1 and (tile == (2, 8) or wg):
            continue  # fused launches: tiles 2x4, 4x4, 4x8, 8x8, one tile per workgroup
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
