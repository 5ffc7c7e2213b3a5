"""Write the golden fixtures for the CPU oracle from the REFERENCE's own gold.

Run in the container that holds /root/reference, after `make -C oracle ref`
has compiled the reference's common/gold_<stencil>.cpp into
oracle/_ref/libgold_<stencil>.so (-O2 -ffp-contract=off):

    python tests/golden/make_golden.py

Each fixture tests/golden/<stencil>_<case>.npz holds the inputs A (N x N FP32),
B (n x n FP32) and the reference gold's output C (C pre-seeded from A, as the
reference driver does at common/main_2r.cu:29).  Cases:
  main  -- the reference driver's own inputs: all-ones A (common/common.hpp:21)
           and init_B coefficients (common/main_1r.cu:10-16, main_2r.cu:13-19)
  rand  -- A ~ U[-1,1), B ~ U[0,1) normalised to sum 1 (numpy PCG64, seed 1)
  odd   -- N = 37, A ~ N(0,1), B ~ N(0,1) (asymmetric, signed; seed 2)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import oracle as O  # noqa: E402


def cases(stencil):
    n = 2 * O.RADIUS[stencil] + 1
    yield "main", np.ones((32, 32), np.float32), O.init_B(stencil)
    rng = np.random.default_rng(1)
    A = rng.uniform(-1.0, 1.0, (24, 24)).astype(np.float32)
    B = rng.uniform(0.0, 1.0, (n, n))
    yield "rand", A, (B / B.sum()).astype(np.float32)
    rng = np.random.default_rng(2)
    yield "odd", rng.standard_normal((37, 37)).astype(np.float32), rng.standard_normal((n, n)).astype(np.float32)


def main():
    for stencil in O.STENCILS:
        if not O.ref_available(stencil):
            raise SystemExit(f"oracle/_ref/libgold_{stencil}.so missing: run `make -C oracle ref`")
        for name, A, B in cases(stencil):
            C = O.ref_gold(stencil, A, B)
            path = os.path.join(HERE, f"{stencil}_{name}.npz")
            np.savez_compressed(path, A=A, B=B, C=C)
            print("wrote", os.path.relpath(path), A.shape)


if __name__ == "__main__":
    main()
