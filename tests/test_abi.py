"""The C ABI library loads here (no GPU) and exports every function that
include/tcstencil.h declares; host-side validation and set-up need no GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tcstencil.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][\w \*]*?\b(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "while")))


def test_header_declares_the_dropins():
    names = declared_functions()
    for s in ("star2d1r", "star2d2r", "box2d1r", "box2d2r"):
        assert f"host_{s}" in names and f"host_direct_{s}" in names
    assert len(names) >= 40


def test_library_exports_every_declared_symbol():
    import tcstencil_amd as T
    L = T.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    from tcstencil_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_functions()


@pytest.mark.parametrize("stencil", ["star2d1r", "star2d2r", "box2d1r", "box2d2r"])
def test_host_code_shims_export_host_code(stencil):
    for kind in ("", "direct_"):
        path = os.path.join(ROOT, "tcstencil_amd", f"libtcs_host_code_{kind}{stencil}.so")
        lib = ctypes.CDLL(path)
        assert hasattr(lib, "host_code")


def test_struct_sizes_match_c():
    from tcstencil_amd import _lib
    assert ctypes.sizeof(_lib.Layout) == 40
    assert ctypes.sizeof(_lib.ProblemStruct) == 288
    assert ctypes.sizeof(_lib.Params) == 32


def test_params_mirror_param_hpp(monkeypatch):
    import tcstencil_amd as T
    for k in ("TCS_RUN_TIMES", "TCS_SKIP_TIMES", "TCS_TILE_SIZE"):
        monkeypatch.delenv(k, raising=False)
    p = T.params()
    assert (p.TOLERANCE, p.MIN_NUMBER, p.RUN_TIMES, p.STEP_TIMES, p.SKIP_TIMES) == (0.01, 1e-6, 500, 1, 400)
    monkeypatch.setenv("TCS_RUN_TIMES", "7")
    assert T.params().RUN_TIMES == 7


def test_layout_and_problem_setup():
    import tcstencil_amd as T
    B = np.arange(25, dtype=np.float32).reshape(5, 5)
    p = T.Problem("box2d2r", 16384, 16384, B)
    L = p.s.layout
    assert (L.nbr, L.nbc, L.sbr, L.sbc, L.align) == (1024, 1024, 1026, 1026, 8)
    assert p.buffer_bytes == 1026 * 1026 * 256 * 2
    assert p.tile == (4, 8) and p.mfma_per_brick == 4
    assert p.kernel_name == "tcs_mfma_sweep<4, 8, 0, 4, 0, false>"
    # fused FP16/BF16 launches: 8x8 tiles on 8-wave workgroups for 8x8 problems
    # and for 4x8 problems with at least 2048 8x8 tiles; SPLIT 4x8 problems
    # fuse 6 or more sweeps per launch on 8-wave workgroups
    p.fuse = 8
    assert p.fused_kernel_name == "tcs_mfma_fused8<8, 8, 0, 4, 0>"
    assert T.Problem("box2d2r", 8192, 8192, B, fuse=4).fused_kernel_name == "tcs_mfma_fused8<8, 8, 0, 4, 0>"
    small = T.Problem("box2d2r", 4096, 4096, B, tile=(4, 8), fuse=4)  # 1024 8x8 tiles
    assert small.fused_kernel_name == "tcs_mfma_fused<4, 8, 0, 4, 0>"
    p.set_tile(8, 8)
    assert p.kernel_name == "tcs_mfma_sweep<8, 8, 0, 4, 0, false>"
    assert p.fused_kernel_name == "tcs_mfma_fused8<8, 8, 0, 4, 0>"
    bf = T.Problem("box2d1r", 300, 300, np.ones((3, 3), np.float32), precision="bf16", tile=(8, 8), fuse=3)
    assert bf.fused_kernel_name == "tcs_mfma_fused8<8, 8, 1, 3, 1>"
    sp = T.Problem("box2d2r", 16384, 16384, B, precision="split", tile=(8, 8), fuse=8)
    assert sp.fused_kernel_name == "tcs_mfma_fused<8, 8, 0, 4, 2>"
    sp = T.Problem("box2d2r", 16384, 16384, B, precision="split", fuse=6)
    assert sp.tile == (4, 8) and sp.fused_kernel_name == "tcs_mfma_fused8<4, 8, 0, 4, 2>"
    sp.fuse = 5
    assert sp.fused_kernel_name == "tcs_mfma_fused<4, 8, 0, 4, 2>"
    assert T.Problem("box2d2r", 64, 64, B, precision="split").mfma_per_brick == 12  # hi*hi, hi*lo, lo*hi
    assert list(p.s.coeff) == list(range(25))
    p = T.Problem("star2d1r", 100, 37, np.arange(9, dtype=np.float32).reshape(3, 3))
    c = np.array(p.s.coeff).reshape(5, 5)
    want = np.zeros((5, 5))
    want[1, 2], want[2, 1], want[2, 2], want[2, 3], want[3, 2] = 1, 3, 4, 5, 7
    assert np.array_equal(c, want)
    p.set_tile(2, 8)
    assert p.tile == (2, 8) and p.s.layout.align == 8


def test_split_coefficients_are_exact_hi_lo():
    import tcstencil_amd as T
    rng = np.random.default_rng(2)
    B = rng.uniform(0, 1, (5, 5)).astype(np.float32)
    p = T.Problem("box2d2r", 64, 64, B, precision="split")
    e = p.s.coeff_exp
    c = np.array(p.s.coeff, dtype=np.float32) * np.float32(2.0 ** e)
    hi = np.array(p.s.coeff_hi, dtype=np.uint16).view(np.float16)
    lo = np.array(p.s.coeff_lo, dtype=np.uint16).view(np.float16)
    assert np.array_equal(hi, c.astype(np.float16))  # RNE like numpy
    assert np.array_equal(lo, (c - hi.astype(np.float32)).astype(np.float16))
    assert 1.0 <= np.abs(c).max() < 2.0
    err = np.abs(hi.astype(np.float64) + lo.astype(np.float64) - c.astype(np.float64))
    assert np.all(err <= np.abs(c) * 2.0 ** -21)


def test_invalid_arguments_return_status():
    import tcstencil_amd as T
    with pytest.raises(T.TcsError, match="n = 3"):
        T.Problem("box2d2r", 10, 10, np.ones((3, 3), np.float32))
    with pytest.raises(T.TcsError, match="empty"):
        T.Problem("box2d1r", 0, 10, np.ones((3, 3), np.float32))
    with pytest.raises(T.TcsError, match="direct-FMA"):
        T.Problem("box2d1r", 10, 10, np.ones((3, 3), np.float32), method="direct", precision="bf16")
    p = T.Problem("box2d1r", 64, 64, np.ones((3, 3), np.float32), precision="split")
    p.set_tile(8, 8)
    assert p.kernel_name.startswith("tcs_mfma_sweep<8, 8, 1, 3, 2") and p.fused_kernel_name.startswith("tcs_mfma_fused<8, 8")
    p.set_tile(2, 8)
    assert p.fused_kernel_name.startswith("tcs_mfma_fused<2, 8") and p.kernel_name.startswith("tcs_mfma_sweep<2, 8")
    with pytest.raises(T.TcsError, match="not offered"):
        p.set_tile(3, 8)
    with pytest.raises(T.TcsError, match="non-finite"):
        T.Problem("box2d1r", 64, 64, np.full((3, 3), np.nan, np.float32))
    L = T.lib()
    assert L.tcs_sweep(None, None, None, None) == T._lib.ERR_INVALID
    assert L.tcs_status_string(4) == b"allocation failed"


def test_split_scale_exponent():
    import tcstencil_amd as T
    for mx in (1.0, 0.75, 3e-3, 1e3, 65000.0):
        e = T.split_scale_exp(mx)
        assert 2 ** 9 <= mx * 2.0 ** e < 2 ** 10
    assert T.split_scale_exp(0.0) == 0


def test_split_scale_exponent_covers_tap_growth(oracle):
    """Taps with absolute sum S > 1 (the reference's init_B: S = 386 for
    star2d2r) grow max|x| by up to S per sweep; the problem-aware exponent keeps
    every stored value of the run, scaled, below binary16's 65504, while
    normalised taps (S = 1) keep the plain exponent."""
    import tcstencil_amd as T
    for stencil in ("star2d2r", "box2d1r", "box2d2r"):
        B = oracle.init_B(stencil)
        n = B.shape[0]
        p = T.Problem(stencil, 64, 64, B, precision="split")
        S = float(np.abs(np.array(p.s.coeff, dtype=np.float64)).sum())
        for sweeps in (1, 3):
            e = T.split_scale_exp(1.0, p, sweeps)
            assert 1.0 * S ** sweeps * 2.0 ** e < 2 ** 10, (stencil, sweeps, e)
            assert 1.0 * S ** sweeps * 2.0 ** e >= 2 ** 9
        assert n == 2 * T.radius(stencil) + 1
    q = T.Problem("box2d2r", 64, 64, T.gen_coeffs("box2d2r", 1), precision="split")
    assert T.split_scale_exp(0.7, q, 10) == T.split_scale_exp(0.7)


def test_scale_exponent_per_precision(oracle):
    """tcs_scale_exp_for: FP16 brings max|x| times the run's tap growth into
    [2^14, 2^15), so the largest stored value stays below binary16's 65504
    and values far below its normal range (2^-14) are lifted into it; SPLIT is
    split_scale_exp; BF16, and inputs that are zero or not finite, get 0."""
    import tcstencil_amd as T
    B = T.gen_coeffs("box2d2r", 3)
    p16 = T.Problem("box2d2r", 64, 64, B, precision="fp16")
    for mx in (1.0, 0.75, 3e-7, 1e-3, 7e4, 1e6, 3.4e30):
        e = T.scale_exp(p16, mx)
        assert 2 ** 14 <= mx * 2.0 ** e < 2 ** 15, (mx, e)
        assert mx * 2.0 ** e < 65504.0
    assert T.scale_exp(p16, 0.0) == 0 and T.scale_exp(p16, float("nan")) == 0 and T.scale_exp(p16, float("inf")) == 0
    # taps of absolute sum S > 1 (the reference's init_B) grow the run's values
    Bi = oracle.init_B("star2d2r")
    q16 = T.Problem("star2d2r", 64, 64, Bi, precision="fp16")
    S = float(np.abs(np.array(q16.s.coeff, dtype=np.float64)).sum())
    for sweeps in (1, 2):
        e = T.scale_exp(q16, 1.0, sweeps)
        assert 2 ** 14 <= S ** sweeps * 2.0 ** e < 2 ** 15
    ps = T.Problem("box2d2r", 64, 64, B, precision="split")
    for mx in (1.0, 1e-3, 1e5):
        assert T.scale_exp(ps, mx, 3) == T.split_scale_exp(mx, ps, 3)
    pb = T.Problem("box2d2r", 64, 64, B, precision="bf16")
    assert T.scale_exp(pb, 1e6) == 0 and T.scale_exp(pb, 1e-9) == 0


def test_edge_kernel_choice():
    """tcs_run_f32's fused edge sweeps: FP16/BF16 on 2x16-brick tiles when the
    layout divides into them and gives 4096 tiles or more, SPLIT on 4x4, small
    grids on the problem's tile; modes outside 1..3 and the direct method have
    no edge kernel (the three steps run)."""
    import tcstencil_amd as T
    B = T.gen_coeffs("box2d2r", 1)
    big = T.Problem("box2d2r", 8192, 4096, B)
    assert [big.edge_kernel_name(m) for m in (1, 2, 3)] == [
        "tcs_mfma_edge<2, 16, 0, 4, 0, %d>" % m for m in (1, 2, 3)]
    assert T.Problem("box2d2r", 8192, 4096, B, precision="bf16").edge_kernel_name(3) == "tcs_mfma_edge<2, 16, 0, 4, 1, 3>"
    assert T.Problem("box2d2r", 8192, 4096, B, precision="split").edge_kernel_name(3) == "tcs_mfma_edge<4, 4, 0, 4, 2, 3>"
    small = T.Problem("box2d2r", 1024, 1024, B)
    assert small.edge_kernel_name(3) == "tcs_mfma_edge<%d, %d, 0, 4, 0, 3>" % small.tile
    assert big.edge_kernel_name(0) == "" and big.edge_kernel_name(4) == ""
    assert T.Problem("box2d2r", 8192, 4096, B, method="direct").edge_kernel_name(3) == ""


def test_gen_coeffs_normalised():
    import tcstencil_amd as T
    for s in T.STENCIL_NAMES:
        B = T.gen_coeffs(s, seed=1)
        assert abs(float(B.astype(np.float64).sum()) - 1.0) < 1e-6
        assert B.min() >= 0.0
        if s.startswith("star"):
            r = T.radius(s)
            off = B.copy()
            off[r, :] = 0
            off[:, r] = 0
            assert not off.any()
    assert not np.array_equal(T.gen_coeffs("box2d2r", 1), T.gen_coeffs("box2d2r", 2))


def test_batch_validation_needs_no_gpu():
    import tcstencil_amd as T
    L = T.lib()
    B = np.ones((3, 3), np.float32)
    ok = T.Problem("box2d1r", 40, 40, B, tile=(2, 4))
    bad_tile = T.Problem("box2d1r", 40, 40, B, tile=(2, 8))
    direct = T.Problem("box2d1r", 40, 40, B, method="direct", tile=(2, 4))
    bad32 = T.Problem("box2d1r", 40, 40, B, tile=(2, 8), frag=32)
    too_fused = T.Problem("box2d2r", 40, 40, np.ones((5, 5), np.float32), tile=(2, 4), fuse=10)
    fake = (ctypes.c_void_p * 1)(0x1000)
    fake2 = (ctypes.c_void_p * 1)(0x2000)
    h = ctypes.c_void_p()
    for p, want in ((bad_tile, T._lib.ERR_UNSUPPORTED), (direct, T._lib.ERR_INVALID),
                    (bad32, T._lib.ERR_UNSUPPORTED), (too_fused, T._lib.ERR_INVALID)):
        arr = (T._lib.ProblemStruct * 1)(p.s)
        assert L.tcs_batch_create(ctypes.byref(h), 1, arr, fake, fake2) == want
    arr = (T._lib.ProblemStruct * 1)(ok.s)
    assert L.tcs_batch_create(ctypes.byref(h), 1, arr, fake, fake) == T._lib.ERR_INVALID  # same buffer twice
    assert L.tcs_batch_run(None, 1, None) == T._lib.ERR_INVALID


def test_frag32_setup_and_validation():
    import tcstencil_amd as T
    B = np.ones((5, 5), np.float32) / 25
    p = T.Problem("box2d2r", 16384, 16384, B, frag=32)
    assert p.frag == 32 and p.tile == (8, 8)
    assert p.kernel_name == "tcs_mfma32_sweep<8, 8, 0, 4, 0>"
    assert p.mfma_per_brick == 12 and p.mfma_flops_per_brick == 12 * 32 * 32 * 16 * 2 // 4
    p.fuse = 4
    p.set_tile(4, 8)
    assert p.fused_kernel_name == "tcs_mfma32_fused<4, 8, 0, 4, 0>"
    p.set_tile(2, 4)  # no 2x2-brick split over 4 waves: no kernel
    assert p.kernel_name == "" and p.fused_kernel_name == ""
    q = T.Problem("box2d1r", 100, 100, np.ones((3, 3), np.float32), frag=32, precision="bf16")
    assert q.tile in ((4, 4), (2, 8), (4, 8), (8, 8)) and q.mfma_per_brick == 7
    assert T.Problem("box2d1r", 64, 64, np.ones((3, 3), np.float32)).mfma_flops_per_brick == 3 * 16384
    with pytest.raises(T.TcsError, match="32x32"):
        T.Problem("box2d2r", 64, 64, B, precision="split", frag=32)
    with pytest.raises(T.TcsError, match="32x32"):
        T.Problem("box2d2r", 64, 64, B, method="direct", frag=32)
    with pytest.raises(T.TcsError, match="fragment"):
        T.Problem("box2d2r", 64, 64, B, frag=24)


def test_error_messages_are_per_thread():
    """tcs_last_error is thread-local: an error raised on one thread does not
    replace the message another thread is about to read (the C ABI keeps no
    shared mutable state)."""
    import threading

    import tcstencil_amd as T
    L = T.lib()
    first, second = threading.Barrier(2), threading.Barrier(2)
    seen = {}

    def worker(name, bad_align):
        lay = T._lib.Layout()
        if name == "a":
            assert L.tcs_layout_init(ctypes.byref(lay), 64, 64, bad_align) == T._lib.ERR_INVALID
            first.wait()   # b now raises its own error
            second.wait()
        else:
            first.wait()
            assert L.tcs_layout_init(ctypes.byref(lay), 64, 64, bad_align) == T._lib.ERR_INVALID
            second.wait()
        seen[name] = L.tcs_last_error().decode()

    threads = [threading.Thread(target=worker, args=("a", 3)), threading.Thread(target=worker, args=("b", 5))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert "align 3" in seen["a"] and "align 5" in seen["b"], seen


@pytest.mark.parametrize("frag", [16, 32])
def test_reference_tile_size_maps_to_an_offered_tile(frag):
    """TILE_SIZE of common/param.hpp:11-13 over the reference's sweep 1..15
    (script/param.py:5), its default 1 and values beyond: every value selects
    an offered tile with a kernel (no TCS_ERR_UNSUPPORTED, so the host_<stencil>
    drivers never exit for it); widths 4 for 1..5, 8 from 6 on."""
    import tcstencil_amd as T
    B = T.gen_coeffs("box2d2r", 1)
    for ts in list(range(1, 16)) + [16, 64]:
        p = T.Problem("box2d2r", 4096, 4096, B, frag=frag)
        default = p.tile
        p.set_tile_size(ts)
        th, tw = p.tile
        assert tw == (4 if ts < 6 else 8), (ts, p.tile)
        assert p.kernel_name, (ts, p.tile)
        assert p.fused_kernel_name or frag == 32 and (th, tw) == (2, 8), (ts, p.tile)
        if frag == 16:
            assert th == min(default[0], 4 if tw == 4 else 8)
    p = T.Problem("box2d1r", 300, 300, T.gen_coeffs("box2d1r", 1))
    t0 = p.tile
    p.set_tile_size(0)
    assert p.tile == t0


def test_small_frag32_problems_have_fused_and_batch_tiles():
    """32x32-fragment defaults stop at 4x4, a tile with single-sweep, fused and
    batched kernels, also on small grids."""
    import tcstencil_amd as T
    for n in (64, 300, 1024, 2048):
        p = T.Problem("box2d2r", n, n, T.gen_coeffs("box2d2r", 1), frag=32)
        assert p.tile in ((4, 4), (4, 8), (8, 8)), (n, p.tile)
        assert p.kernel_name and p.fused_kernel_name, (n, p.tile)


class _FakeDeviceTensor:
    """Duck-typed stand-in for a device tensor: enough for the argument checks,
    which must reject it before any library call."""

    is_cuda = True

    def __init__(self, nbytes):
        self.nbytes = nbytes

    def numel(self):
        return self.nbytes

    def element_size(self):
        return 1

    def is_contiguous(self):
        return True

    def data_ptr(self):
        raise AssertionError("the checks must reject the buffer before using its address")


def test_storage_buffers_too_small_are_rejected_before_any_launch():
    import tcstencil_amd as T
    B = T.gen_coeffs("box2d2r", 1)
    p = T.Problem("box2d2r", 100, 200, B)
    small = _FakeDeviceTensor(p.buffer_bytes - 1)
    with pytest.raises(ValueError, match="needs"):
        p.clear(small)
    with pytest.raises(ValueError, match="needs"):
        p.sweep(small, small)
    with pytest.raises(ValueError, match="needs"):
        p.run(small, small, 3)
    with pytest.raises(ValueError, match="needs"):
        p.graph(small, small, 3)
    with pytest.raises(ValueError, match="needs"):
        p.unpack(small)
    # a buffer sized for a smaller tile is too small once the layout is re-aligned
    q = T.Problem("box2d2r", 100, 300, B, tile=(2, 4))
    old = _FakeDeviceTensor(q.buffer_bytes)
    q.set_tile(8, 8)
    assert q.buffer_bytes > old.nbytes
    with pytest.raises(ValueError, match="needs"):
        q.clear(old)
    with pytest.raises(ValueError, match="grid 0 buf1"):
        T.Batch([p, p], [_FakeDeviceTensor(p.buffer_bytes)] * 2,
                [small, _FakeDeviceTensor(p.buffer_bytes)])


def test_hi_lo_coefficient_option():
    """coef_lo selects the hi+lo coefficient kernels (kernel precision ids 4 and
    5: two MFMAs per fragment) for MFMA FP16/BF16 with 16x16 fragments, prefers
    8x8 tiles (fused on 8-wave workgroups), and is ignored by SPLIT and the direct method; 32x32 fragments
    have no such kernel (UNSUPPORTED at launch: no kernel name)."""
    import tcstencil_amd as T
    B = T.gen_coeffs("box2d2r", 1)
    for prec, kp in (("fp16", 4), ("bf16", 5)):
        p = T.Problem("box2d2r", 16384, 16384, B, precision=prec, coef_lo=True)
        assert p.coef_lo and p.tile == (8, 8) and p.mfma_per_brick == 8
        assert p.kernel_name == f"tcs_mfma_sweep<8, 8, 0, 4, {kp}, false>"
        assert p.fused_kernel_name == f"tcs_mfma_fused8<8, 8, 0, 4, {kp}>"
        # 4x8 problems fuse on 8x8 tiles of 8-wave workgroups, as rounded taps do
        p.set_tile(4, 8)
        assert p.kernel_name == f"tcs_mfma_sweep<4, 8, 0, 4, {kp}, false>"
        assert p.fused_kernel_name == f"tcs_mfma_fused8<8, 8, 0, 4, {kp}>"
        q = T.Problem("box2d1r", 16384, 16384, T.gen_coeffs("box2d1r", 1), precision=prec, coef_lo=True)
        assert q.mfma_per_brick == 6
    s = T.Problem("box2d2r", 1024, 1024, B, precision="split", coef_lo=True)
    assert s.mfma_per_brick == 12 and ", 2, false>" in s.kernel_name
    d = T.Problem("box2d2r", 1024, 1024, B, method="direct", coef_lo=True)
    assert d.kernel_name.startswith("tcs_direct_sweep")
    f = T.Problem("box2d2r", 1024, 1024, B, frag=32, coef_lo=True)
    assert f.kernel_name == "" and f.fused_kernel_name == ""
    p = T.Problem("box2d2r", 1024, 1024, B)
    p.set_coef_lo(True)
    assert p.coef_lo and ", 4, false>" in p.kernel_name


def test_tuning_cache_lookup_and_apply(tmp_path, monkeypatch):
    """The tuning cache file (no GPU needed with the 'device any' record): a
    matching entry sets tile, wg_per_cu, fuse and fragment; tcs_problem_init
    applies $TCS_TUNE_CACHE's single-grid entries; batched entries apply only
    when asked for; an explicit fragment choice is kept; malformed files are
    errors; a file for another device changes nothing."""
    import tcstencil_amd as T
    from tcstencil_amd import autotune as A
    B = T.gen_coeffs("box2d2r", 1)
    path = tmp_path / "tune.txt"
    lines = [A.entry_line("box2d2r", "mfma", "fp16", 0, 0, 3, 16, (8, 8), 0, 8),
             A.entry_line("box2d2r", "mfma", "fp16", 0, 0, 1, 16, (2, 4), 2, 1),
             A.entry_line("box2d2r", "mfma", "fp16", 0, 1, 1, 32, (4, 4), 0, 4),
             A.entry_line("box2d2r", "direct", "fp16", 0, 0, 3, 16, (2, 8), 0, 1),
             A.entry_line("box2d2r", "mfma", "fp16", 1, 0, 3, 16, (4, 8), 0, 2)]
    A.write_cache(str(path), lines, device="device any 0")
    monkeypatch.delenv("TCS_TUNE_CACHE", raising=False)
    assert T.Problem("box2d2r", 16384, 16384, B).tile == (4, 8)  # no cache named: defaults
    p = T.Problem("box2d2r", 16384, 16384, B, tuned=(str(path), False))
    assert (p.tile, p.wg_per_cu, p.fuse, p.frag) == ((8, 8), 0, 8, 16)
    monkeypatch.setenv("TCS_TUNE_CACHE", str(path))
    p = T.Problem("box2d2r", 500, 700, B)  # bucket 1 (smaller side 32 bricks)
    assert (p.tile, p.wg_per_cu, p.fuse) == ((2, 4), 2, 1)
    q = T.Problem("box2d2r", 500, 700, B, tuned=("", True))  # the batched entry
    assert (q.tile, q.fuse, q.frag) == ((4, 4), 4, 32)
    d = T.Problem("box2d2r", 16384, 16384, B, method="direct")
    assert d.tile == (2, 8) and d.fuse == 0
    lo = T.Problem("box2d2r", 16384, 16384, B, coef_lo=True)
    assert (lo.tile, lo.fuse) == ((4, 8), 2)
    f = T.Problem("box2d2r", 16384, 16384, B, frag=32)  # entry is for 16x16 fragments: not applied
    assert f.frag == 32 and f.tile == (8, 8) and f.fuse == 0
    t = T._lib.Tuning()
    assert T.lib().tcs_tune_lookup(str(path).encode(), T.Problem("box2d2r", 64, 64, B).ref, 0,
                                   ctypes.byref(t)) == T._lib.ERR_UNSUPPORTED  # bucket 0: no entry
    monkeypatch.delenv("TCS_TUNE_CACHE")
    other = tmp_path / "other.txt"
    A.write_cache(str(other), lines, device="device gfx000 1")
    assert T.Problem("box2d2r", 16384, 16384, B, tuned=(str(other), False)).tile == (4, 8)
    bad = tmp_path / "bad.txt"
    bad.write_text("device any 0\nentry box2d2r mfma fp16 0 0 3 16 8\n")
    with pytest.raises(T.TcsError, match="bad entry"):
        T.Problem("box2d2r", 16384, 16384, B, tuned=(str(bad), False))
    bad.write_text("entry box2d2r mfma fp16 0 0 3 16 8 8 0 1\n")
    with pytest.raises(T.TcsError, match="before the device"):
        T.Problem("box2d2r", 16384, 16384, B, tuned=(str(bad), False))
    assert [T.lib().tcs_tune_bucket(n, n) for n in (64, 240, 256, 1008, 1024, 4080, 4096)] == [0, 0, 1, 1, 2, 2, 3]
