"""Sweep post-processing (tcstencil_amd.report): the reference's pick_tile_csv /
pick_best_csv semantics (script/process_data.py:20-44) on a synthetic sweep."""
import csv
import os

import pytest

from tcstencil_amd import report as R


def _write_sweep(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(R.REF_COLUMNS + ["gstencil_s"])
        for r in rows:
            w.writerow(r)


ROWS = [
    # type, stencil, mesh, tile, time, unit, gstencil_s
    ["tensor", "box2d2r", 1024, "4x8", 2.0, "us", 0],
    ["tensor", "box2d2r", 1024, "8x8", 1.5, "us", 0],
    ["tensor", "box2d2r", 2048, "4x8", 5.0, "us", 0],
    ["tensor", "box2d2r", 2048, "8x8", 6.0, "us", 0],
    ["baseline", "box2d2r", 1024, "4x8", 3.0, "us", 0],
    ["tensor", "star2d1r", 1024, "4x8", 1.0, "ms", 0],
]


def test_pick_tile_and_best(tmp_path):
    p = tmp_path / "sweep.csv"
    _write_sweep(p, ROWS)
    rows = R.load_sweep_csv(str(p))
    t = R.pick_tile(rows, "4x8")
    assert [(r["type"], r["stencil"], r["mesh size"]) for r in t] == [
        ("tensor", "box2d2r", 1024), ("tensor", "box2d2r", 2048), ("baseline", "box2d2r", 1024),
        ("tensor", "star2d1r", 1024)]
    best = {(r["type"], r["stencil"], r["mesh size"]): r["tile size"] for r in R.pick_best(rows)}
    assert best == {("tensor", "box2d2r", 1024): "8x8", ("tensor", "box2d2r", 2048): "4x8",
                    ("baseline", "box2d2r", 1024): "4x8", ("tensor", "star2d1r", 1024): "4x8"}


def test_update_rate_uses_interior_points_and_units(tmp_path):
    p = tmp_path / "sweep.csv"
    _write_sweep(p, ROWS)
    rows = R.load_sweep_csv(str(p))
    assert R.updates_per_s(rows[0]) == pytest.approx((1024 - 4) ** 2 / 2e-6)
    assert R.updates_per_s(rows[5]) == pytest.approx((1024 - 2) ** 2 / 1e-3)  # "ms" unit


def test_reference_rate_is_n_squared_over_t(tmp_path):
    """script/draw.py:19 plots mesh_size^2 / time * 1e6 (time in us); both rates
    are columns of every output CSV."""
    p = tmp_path / "sweep.csv"
    _write_sweep(p, ROWS)
    rows = R.load_sweep_csv(str(p))
    assert R.reference_updates_per_s(rows[0]) == pytest.approx(1024 * 1024 / 2.0 * 1e6)
    assert R.reference_updates_per_s(rows[5]) == pytest.approx(1024 * 1024 / 1e-3)
    out = tmp_path / "out"
    assert R.main([str(p), "--out-dir", str(out)]) == 0
    with open(out / "sweep_best.csv") as f:
        rec = next(csv.DictReader(f))
    assert int(rec["ref_updates_per_s"]) == round(1024 * 1024 / 1.5e-6)
    assert int(rec["updates_per_s"]) == round(1020 * 1020 / 1.5e-6)


def test_cli_writes_reference_files_and_figures(tmp_path):
    p = tmp_path / "sweep.csv"
    _write_sweep(p, ROWS)
    out = tmp_path / "out"
    assert R.main([str(p), "--out-dir", str(out), "--plots", "--meshes", "1024", "2048"]) == 0
    for suffix in ("_full.csv", "_tile.csv", "_best.csv", "_mesh.png", "_tiles.png"):
        assert os.path.getsize(out / ("sweep" + suffix)) > 0
    with open(out / "sweep_tile.csv") as f:
        assert f.readline().strip().split(",")[:5] == ["type", "stencil", "mesh size", "time", "unit"]
