"""Post-processing of mesh/tile sweep results (the reference's script/process_data.py
and script/draw.py, for this library's sweep CSV).

The sweep CSV is what ``python -m tcstencil_amd.harness sweep --out results.csv``
writes: the reference's columns ``type,stencil,mesh size,tile size,time,unit``
(script/process_data.py:14; ``type`` is ``tensor`` for the MFMA kernel and
``baseline`` for the direct-FMA kernel, as in the reference) plus ``gstencil_s``
and ``hbm_gbs``.  Tile sizes are brick shapes ("4x8") instead of the reference's
single tileX count.

  python -m tcstencil_amd.report results.csv --out-dir report/ [--tile 4x8] [--plots]

writes, next to the reference's three files (script/process_data.py:13-44),
  <stem>_full.csv   every row, reference columns first
  <stem>_tile.csv   one tile shape, without the tile column (pick_tile_csv)
  <stem>_best.csv   fastest tile per (type, stencil, mesh size) (pick_best_csv)
each with both update rates as columns (updates_per_s = (N-2r)^2 / t, this
library's figure; ref_updates_per_s = N^2 / t, script/draw.py:19), and with
--plots the two figures of script/draw.py: update points/s against mesh size for
the chosen tile, MFMA vs direct (draw_tile_1), and against tile shape for four
mesh sizes (draw_mult_tile), one panel per stencil, at the reference's rate by
default (--rate interior for the other).
"""
from __future__ import annotations

import argparse
import csv
import os
import sys

REF_COLUMNS = ["type", "stencil", "mesh size", "tile size", "time", "unit"]
STENCILS = ["box2d2r", "star2d2r", "box2d1r", "star2d1r"]  # script/param.py:2 order


def _radius(stencil: str) -> int:
    return 1 if stencil in ("star2d1r", "box2d1r") else 2


def load_sweep_csv(path: str) -> list[dict]:
    """Rows of a sweep CSV with numeric mesh size and time (microseconds)."""
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            r = dict(r)
            r["mesh size"] = int(r["mesh size"])
            r["time"] = float(r["time"])
            rows.append(r)
    return rows


def time_us(row: dict) -> float:
    unit = row.get("unit", "us")
    scale = {"us": 1.0, "ms": 1e3, "ns": 1e-3, "s": 1e6}[unit]
    return row["time"] * scale


def updates_per_s(row: dict) -> float:
    """Interior updates per second, (N - 2r)^2 / t: the figure this library
    reports everywhere."""
    n, r = row["mesh size"], _radius(row["stencil"])
    return (n - 2 * r) ** 2 / (time_us(row) * 1e-6)


def reference_updates_per_s(row: dict) -> float:
    """The reference's figure rate, N^2 / t (script/draw.py:19: mesh_size *
    mesh_size / time * 1e6 with the time in microseconds)."""
    n = row["mesh size"]
    return n * n / (time_us(row) * 1e-6)


RATES = {"interior": updates_per_s, "reference": reference_updates_per_s}


def with_rates(rows: list[dict]) -> list[dict]:
    """Rows with both update rates as columns (points per second)."""
    return [dict(r, updates_per_s=round(updates_per_s(r)), ref_updates_per_s=round(reference_updates_per_s(r)))
            for r in rows]


def pick_tile(rows: list[dict], tile: str) -> list[dict]:
    """Rows of one tile shape (script/process_data.py:20-27)."""
    return [r for r in rows if r["tile size"] == tile]


def pick_best(rows: list[dict]) -> list[dict]:
    """Fastest row per (type, stencil, mesh size) (script/process_data.py:30-44),
    in first-seen key order."""
    best: dict[tuple, dict] = {}
    for r in rows:
        key = (r["type"], r["stencil"], r["mesh size"])
        if key not in best or time_us(r) < time_us(best[key]):
            best[key] = r
    return list(best.values())


def write_csv(rows: list[dict], path: str, drop_tile: bool = False) -> None:
    cols = [c for c in REF_COLUMNS if not (drop_tile and c == "tile size")]
    extra = [c for c in (rows[0].keys() if rows else []) if c not in REF_COLUMNS]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=cols + extra, extrasaction="ignore")
        w.writeheader()
        w.writerows(rows)


def plot_mesh(rows: list[dict], tile: str, path: str, rate: str = "reference") -> None:
    """Update points/s against mesh size for one tile, MFMA ('tensor') and direct
    ('baseline'), one panel per stencil (script/draw.py:24-68); rate
    'reference' plots N^2 / t as draw.py does, 'interior' (N - 2r)^2 / t."""
    rate_of = RATES[rate]
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    sel = pick_tile(rows, tile)
    fig, axes = plt.subplots(2, 2, figsize=(10, 7))
    for ax, stencil in zip(axes.flat, STENCILS):
        for kind, style, label in (("baseline", ":", "direct FMA"), ("tensor", "-", "MFMA")):
            pts = sorted((r["mesh size"], rate_of(r)) for r in sel
                         if r["stencil"] == stencil and r["type"] == kind)
            if pts:
                ax.plot([p[0] for p in pts], [p[1] for p in pts], style, label=label)
        ax.set_title(f"{stencil} (tile {tile})")
        ax.set_xlabel("mesh size N")
        ax.set_ylabel("update points/s")
        if ax.lines:
            ax.legend()
    fig.tight_layout()
    fig.savefig(path)
    plt.close(fig)


def plot_tiles(rows: list[dict], meshes: list[int], path: str, rate: str = "reference") -> None:
    """MFMA update points/s against tile shape for several mesh sizes, one panel
    per stencil (script/draw.py:94-145)."""
    rate_of = RATES[rate]
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    tiles = sorted({r["tile size"] for r in rows}, key=lambda t: [int(v) for v in t.split("x")])
    fig, axes = plt.subplots(2, 2, figsize=(10, 7))
    for ax, stencil in zip(axes.flat, STENCILS):
        for n in meshes:
            by_tile = {r["tile size"]: rate_of(r) for r in rows
                       if r["stencil"] == stencil and r["type"] == "tensor" and r["mesh size"] == n}
            xs = [t for t in tiles if t in by_tile]
            if xs:
                ax.plot(range(len(xs)), [by_tile[t] for t in xs], "-o", label=f"N={n}")
                ax.set_xticks(range(len(xs)), xs)
        ax.set_title(stencil)
        ax.set_xlabel("tile (bricks)")
        ax.set_ylabel("update points/s")
        if ax.lines:
            ax.legend(fontsize=8)
    fig.tight_layout()
    fig.savefig(path)
    plt.close(fig)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m tcstencil_amd.report")
    ap.add_argument("csv")
    ap.add_argument("--out-dir", default=".")
    ap.add_argument("--tile", default="4x8")
    ap.add_argument("--plots", action="store_true")
    ap.add_argument("--rate", choices=sorted(RATES), default="reference",
                    help="figure rate: the reference's N^2/t (draw.py:19) or interior (N-2r)^2/t")
    ap.add_argument("--meshes", type=int, nargs="+", default=None,
                    help="mesh sizes of the per-tile figure (default: four spread over the data)")
    a = ap.parse_args(argv)
    rows = with_rates(load_sweep_csv(a.csv))
    os.makedirs(a.out_dir, exist_ok=True)
    stem = os.path.join(a.out_dir, os.path.splitext(os.path.basename(a.csv))[0])
    write_csv(rows, stem + "_full.csv")
    write_csv(pick_tile(rows, a.tile), stem + "_tile.csv", drop_tile=True)
    write_csv(pick_best(rows), stem + "_best.csv")
    if a.plots:
        plot_mesh(rows, a.tile, stem + "_mesh.png", a.rate)
        sizes = sorted({r["mesh size"] for r in rows})
        meshes = a.meshes or [sizes[int(q * (len(sizes) - 1))] for q in (0.4, 0.55, 0.8, 1.0)]
        plot_tiles(rows, meshes, stem + "_tiles.png", a.rate)
    print(f"report: {len(rows)} rows -> {stem}_{{full,tile,best}}.csv" + (" + figures" if a.plots else ""))
    return 0


if __name__ == "__main__":
    sys.exit(main())
