"""MI355X (gfx950) matrix-core 2D stencils: star2d1r, star2d2r, box2d1r,
box2d2r as MFMA products over a 16x16 brick layout, next to a direct-FMA kernel.

The compute lives in libtcstencil.so (hand-written HIP, C ABI in
include/tcstencil.h); this package is its Python front end.
"""
from ._lib import (BOX2D1R, BOX2D2R, METHOD_DIRECT, METHOD_MFMA, PREC_BF16, PREC_FP16,
                   PREC_SPLIT, STAR2D1R, STAR2D2R, STENCIL_NAMES, TcsError, lib)
from .api import (Batch, Graph, Problem, copy, error_norms, gen_coeffs, gen_sinusoid, gen_uniform,
                  host_sweep, maxabs, oracle_sweep_f64, oracle_sweeps_f64, params, radius,
                  round_to, scale_exp, split_scale_exp)

__all__ = [
    "BOX2D1R", "BOX2D2R", "STAR2D1R", "STAR2D2R", "METHOD_MFMA", "METHOD_DIRECT",
    "PREC_FP16", "PREC_BF16", "PREC_SPLIT", "STENCIL_NAMES", "TcsError", "lib",
    "Batch", "Graph", "Problem", "copy", "error_norms", "gen_coeffs", "gen_sinusoid", "gen_uniform",
    "host_sweep", "maxabs", "oracle_sweep_f64", "oracle_sweeps_f64", "params", "radius",
    "round_to", "scale_exp", "split_scale_exp",
]
