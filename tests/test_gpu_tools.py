"""The autotuner's cache and the sweep harness on the GPU.

  * the tuner times candidates and writes a cache file; problems set up from it
    (explicitly and through $TCS_TUNE_CACHE in tcs_problem_init) carry the
    winners and give the same bits as the default launch parameters (tile,
    persistent grid and sweeps per launch never change a result);
  * batched grids set up from the cache's batched entries give, grid by grid,
    the bits of each grid swept alone with the same parameters;
  * `harness sweep` (the replacement of the reference's compile-one-binary-
    per-workload sweep) writes the reference's CSV columns
    (script/process_data.py:14) for every workload, with HIP-event timing and,
    for one workload, rocprofv3's kernel statistics.
"""
import csv
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


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


def sweeps_of(T, torch, p, x, sweeps):
    b0, b1 = p.alloc(), p.alloc()
    p.clear(b1)
    p.pack(x, b0)
    return p.unpack(p.run(b0, b1, sweeps)).clone()


def test_tuner_cache_round_trip(T, torch, tmp_path, monkeypatch):
    from tcstencil_amd import autotune as A
    path = str(tmp_path / "tune.txt")
    lines = []
    for stencil, size in (("box2d2r", 1024), ("star2d1r", 4096)):
        line, log = A.tune_cache_entries(stencil, size, sweeps=6, tiles=[(2, 4), (4, 8), (8, 8)], wgs=(0, 2),
                                         fuses=(1, 4))
        say("case tuner", stencil, size, line, log)
        lines.append(line)
    A.write_cache(path, lines)
    with open(path) as f:
        text = f.read()
    assert text.splitlines()[1] == A.device_record() and len(text.splitlines()) == 4
    for line in lines:
        _, stencil, method, precision, lo, batched, bucket, frag, th, tw, wg, fuse = line.split()
        size = 1024 if stencil == "box2d2r" else 4096
        B = T.gen_coeffs(stencil, 3)
        x = T.gen_uniform(size, size, seed=3, dtype=torch.float32)
        p = T.Problem(stencil, size, size, B, tuned=(path, False))
        assert (p.tile, p.wg_per_cu, p.fuse) == ((int(th), int(tw)), int(wg), int(fuse)), line
        monkeypatch.setenv("TCS_TUNE_CACHE", path)
        q = T.Problem(stencil, size, size, B)  # tcs_problem_init applies the cache itself
        monkeypatch.delenv("TCS_TUNE_CACHE")
        assert (q.tile, q.wg_per_cu, q.fuse) == (p.tile, p.wg_per_cu, p.fuse)
        default = T.Problem(stencil, size, size, B)
        assert torch.equal(sweeps_of(T, torch, p, x, 7), sweeps_of(T, torch, default, x, 7)), line


def test_batch_from_tuned_cache(T, torch, tmp_path):
    """tune_batch's winners per size bucket written as batched cache entries,
    then a batch set up from them: every grid equals the same grid swept alone
    with its tuned parameters, bitwise."""
    import math
    from tcstencil_amd import autotune as A
    rng = np.random.default_rng(9)
    specs = []
    for i in range(24):
        stencil = ("box2d1r", "star2d2r")[i % 2]
        rows, cols = (int(v) for v in np.exp(rng.uniform(math.log(64), math.log(1500), 2)))
        x = T.gen_uniform(rows, cols, seed=50 + i, dtype=torch.float32)
        specs.append((stencil, rows, cols, T.gen_coeffs(stencil, 50 + i), x))
    best, _ = A.tune_batch(specs, sweeps=6, tiles=((2, 4), (4, 8)), frags=(16, 32), wgs=(0,), fuses=(1, 4))
    say("case batch tuner", best)
    path = str(tmp_path / "tune.txt")
    A.write_cache(path, A.batch_cache_entries(best))
    probs, b0s, b1s, batch = A.build_batch(specs, None, cache=path)
    for p, (stencil, rows, cols, _, _) in zip(probs, specs):
        tile, frag, fuse, wg = best[A.bucket_of(rows, cols)]
        assert (p.tile, p.frag, p.fuse, p.wg_per_cu) == (tile, frag, min(fuse, A.max_fuse(stencil)), wg)
    batch.run(6)
    torch.cuda.synchronize()
    for p, b0, (stencil, rows, cols, B, x) in zip(probs, b0s, specs):
        alone = T.Problem(stencil, rows, cols, B, tile=p.tile, frag=p.frag, fuse=p.fuse)
        assert torch.equal(p.unpack(b0), sweeps_of(T, torch, alone, x, 6)), (rows, cols, p.tile, p.frag)
    batch.destroy()


REF_COLUMNS = ["type", "stencil", "mesh size", "tile size", "time", "unit"]


def test_harness_sweep_writes_the_reference_csv(T, tmp_path):
    from tcstencil_amd import harness, report
    out = str(tmp_path / "sweep.csv")
    args = ["sweep", "--sizes", "160", "320", "--stencils", "box2d2r", "star2d1r", "--tiles", "2x4", "4x8",
            "--sweeps", "3", "--out", out]
    assert harness.main(args) == 0
    with open(out) as f:
        rows = list(csv.DictReader(f))
        f.seek(0)
        header = f.readline().strip().split(",")
    assert header[:6] == REF_COLUMNS
    assert len(rows) == 2 * 2 * 2 * 2  # stencils x methods x sizes x tiles
    assert {r["type"] for r in rows} == {"tensor", "baseline"}
    for r in rows:
        t = float(r["time"])
        n = int(r["mesh size"])
        assert r["unit"] == "us" and t > 0
        assert float(r["gpts_n2_s"]) == pytest.approx(n * n / t / 1e3, rel=2e-2, abs=0.1)
    assert report.main([out, "--out-dir", str(tmp_path / "rep")]) == 0
    assert os.path.getsize(tmp_path / "rep" / "sweep_best.csv") > 0


@pytest.mark.skipif(shutil.which("rocprofv3") is None, reason="rocprofv3 not on PATH")
def test_harness_sweep_under_rocprofv3(T, tmp_path):
    """One workload timed from rocprofv3's kernel statistics (--rocprof)."""
    out = str(tmp_path / "sweep.csv")
    cmd = [sys.executable, "-m", "tcstencil_amd.harness", "sweep", "--sizes", "320", "--stencils", "box2d2r",
           "--methods", "mfma", "--tiles", "4x8", "--sweeps", "3", "--rocprof", "--out", out]
    res = subprocess.run(cmd, cwd=ROOT, capture_output=True, text=True, timeout=240)
    say(res.stdout[-2000:], res.stderr[-2000:])
    assert res.returncode == 0
    with open(out) as f:
        rows = list(csv.DictReader(f))
    assert len(rows) == 1 and float(rows[0]["time"]) > 0
