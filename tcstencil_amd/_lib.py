"""ctypes binding of libtcstencil.so (declarations mirror include/tcstencil.h).

The shared library is built in-tree by ``__graft_entry__.build()`` (or
``make -C tcstencil_amd/csrc``).  Loading fails loudly when it is missing:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libtcstencil.so")

# Constants of include/tcstencil.h
STAR2D1R, STAR2D2R, BOX2D1R, BOX2D2R = 0, 1, 2, 3
METHOD_MFMA, METHOD_DIRECT, METHOD_DIRECT_DOT2 = 0, 1, 2
PREC_FP16, PREC_BF16, PREC_SPLIT = 0, 1, 2
OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_HIP, ERR_ALLOC = 0, 1, 2, 3, 4

STENCIL_NAMES = ("star2d1r", "star2d2r", "box2d1r", "box2d2r")
STENCIL_ID = {name: i for i, name in enumerate(STENCIL_NAMES)}
METHOD_ID = {"mfma": METHOD_MFMA, "direct": METHOD_DIRECT, "direct_dot2": METHOD_DIRECT_DOT2}
PREC_ID = {"fp16": PREC_FP16, "bf16": PREC_BF16, "split": PREC_SPLIT}


class Layout(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int32), ("cols", ctypes.c_int32),
                ("nbr", ctypes.c_int32), ("nbc", ctypes.c_int32),
                ("sbr", ctypes.c_int32), ("sbc", ctypes.c_int32),
                ("align", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("plane_elems", ctypes.c_int64)]


class ProblemStruct(ctypes.Structure):
    _fields_ = [("stencil", ctypes.c_int32), ("method", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("radius", ctypes.c_int32),
                ("tile_h", ctypes.c_int32), ("tile_w", ctypes.c_int32),
                ("layout", Layout),
                ("coeff", ctypes.c_float * 25),
                ("coeff_hi", ctypes.c_uint16 * 25),
                ("coeff_lo", ctypes.c_uint16 * 25),
                ("coeff_exp", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32),
                ("fuse", ctypes.c_int32), ("frag", ctypes.c_int32),
                ("coef_lo", ctypes.c_int32)]


class Tuning(ctypes.Structure):
    _fields_ = [("tile_h", ctypes.c_int32), ("tile_w", ctypes.c_int32), ("wg_per_cu", ctypes.c_int32),
                ("fuse", ctypes.c_int32), ("frag", ctypes.c_int32)]


class Params(ctypes.Structure):
    """common/param.hpp:1-13 as a runtime struct (same names and defaults)."""
    _fields_ = [("TOLERANCE", ctypes.c_double), ("MIN_NUMBER", ctypes.c_double),
                ("RUN_TIMES", ctypes.c_int32), ("STEP_TIMES", ctypes.c_int32),
                ("SKIP_TIMES", ctypes.c_int32), ("TILE_SIZE", ctypes.c_int32)]


_P = ctypes.POINTER
_vp = ctypes.c_void_p
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_fp = _P(ctypes.c_float)
_dp = _P(ctypes.c_double)
_prob = _P(ProblemStruct)

# name -> (restype, argtypes)
SIGNATURES = {
    "tcs_status_string": (ctypes.c_char_p, [_i]),
    "tcs_last_error": (ctypes.c_char_p, []),
    "tcs_params_default": (None, [_P(Params)]),
    "tcs_params_from_env": (None, [_P(Params)]),
    "host_star2d1r": (None, [_fp, _fp, _fp, _i, _i]),
    "host_star2d2r": (None, [_fp, _fp, _fp, _i, _i]),
    "host_box2d1r": (None, [_fp, _fp, _fp, _i, _i]),
    "host_box2d2r": (None, [_fp, _fp, _fp, _i, _i]),
    "host_direct_star2d1r": (None, [_fp, _fp, _fp, _i, _i]),
    "host_direct_star2d2r": (None, [_fp, _fp, _fp, _i, _i]),
    "host_direct_box2d1r": (None, [_fp, _fp, _fp, _i, _i]),
    "host_direct_box2d2r": (None, [_fp, _fp, _fp, _i, _i]),
    "tcs_host_sweep": (_i, [_i, _i, _i, _fp, _fp, _i, _fp, _i, _i, _i, _P(Params), _fp]),
    "tcs_layout_init": (_i, [_P(Layout), _i, _i, _i]),
    "tcs_problem_init": (_i, [_prob, _i, _i, _i, _i, _i, _fp, _i]),
    "tcs_radius": (_i, [_i]),
    "tcs_problem_set_tile": (_i, [_prob, _i, _i]),
    "tcs_problem_set_tile_size": (_i, [_prob, _i]),
    "tcs_problem_set_frag": (_i, [_prob, _i]),
    "tcs_problem_set_coef_lo": (_i, [_prob, _i]),
    "tcs_buffer_bytes": (_i64, [_prob]),
    "tcs_tune_bucket": (_i, [_i, _i]),
    "tcs_tune_lookup": (_i, [ctypes.c_char_p, _prob, _i, _P(Tuning)]),
    "tcs_problem_apply_tuning": (_i, [_prob, ctypes.c_char_p, _i]),
    "tcs_pack_f32": (_i, [_prob, _vp, _i64, _vp, _i, _vp]),
    "tcs_pack_f64": (_i, [_prob, _vp, _i64, _vp, _i, _vp]),
    "tcs_unpack_f32": (_i, [_prob, _vp, _vp, _i64, _i, _vp, _vp]),
    "tcs_clear": (_i, [_prob, _vp, _vp]),
    "tcs_copy": (_i, [_vp, _vp, _i64, _vp]),
    "tcs_maxabs_f32": (_i, [_vp, _i, _i, _i64, _vp, _vp]),
    "tcs_maxabs_f64": (_i, [_vp, _i, _i, _i64, _vp, _vp]),
    "tcs_split_scale_exp": (_i, [ctypes.c_double]),
    "tcs_split_scale_exp_for": (_i, [_prob, ctypes.c_double, _i]),
    "tcs_scale_exp_for": (_i, [_prob, ctypes.c_double, _i]),
    "tcs_sweep": (_i, [_prob, _vp, _vp, _vp]),
    "tcs_run": (_i, [_prob, _vp, _vp, _i, _vp]),
    "tcs_run_f32": (_i, [_prob, _vp, _vp, _i64, _vp, _vp, _i, _i, _vp]),
    "tcs_graph_create": (_i, [_prob, _vp, _vp, _i, _vp, _P(_vp)]),
    "tcs_graph_launch": (_i, [_vp, _vp]),
    "tcs_graph_destroy": (_i, [_vp]),
    "tcs_batch_create": (_i, [_P(_vp), _i, _prob, _P(_vp), _P(_vp)]),
    "tcs_batch_run": (_i, [_vp, _i, _vp]),
    "tcs_batch_graph_create": (_i, [_vp, _i, _vp, _P(_vp)]),
    "tcs_batch_destroy": (_i, [_vp]),
    "tcs_batch_interior_points": (_i64, [_vp]),
    "tcs_batch_launches_per_sweep": (_i, [_vp]),
    "tcs_oracle_sweep_f64": (_i, [_i, _vp, _dp, _i, _vp, _i, _i, _vp]),
    "tcs_error_norms": (_i, [_vp, _vp, _i, _i, _i, _dp, _vp]),
    "tcs_error_norms_noalloc": (_i, [_vp, _vp, _i, _i, _i, _vp, _vp]),
    "tcs_gen_uniform_f32": (_i, [_vp, _i, _i, _u64, _u64, ctypes.c_float, ctypes.c_float, _vp]),
    "tcs_gen_uniform_f64": (_i, [_vp, _i, _i, _u64, _u64, ctypes.c_double, ctypes.c_double, _vp]),
    "tcs_gen_sinusoid_f64": (_i, [_vp, _i, _i, _u64, ctypes.c_double, _vp]),
    "tcs_round_f64": (_i, [_vp, _i64, _i, _vp]),
    "tcs_round_f32": (_i, [_vp, _i64, _i, _vp]),
    "tcs_gen_coeffs": (_i, [_i, _u64, _fp, _i]),
    "tcs_kernel_name": (ctypes.c_char_p, [_prob]),
    "tcs_fused_kernel_name": (ctypes.c_char_p, [_prob]),
    "tcs_edge_kernel_name": (ctypes.c_char_p, [_prob, _i]),
    "tcs_mfma_per_brick": (_i, [_prob]),
    "tcs_mfma_flops_per_brick": (_i64, [_prob]),
}

_lib = None


class TcsError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"tcstencil error {status}: {message}")
        self.status = status


def lib() -> ctypes.CDLL:
    """Load libtcstencil.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(or `make -C tcstencil_amd/csrc`); there is no fallback path")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int) -> None:
    if status != OK:
        raise TcsError(status, lib().tcs_last_error().decode())
