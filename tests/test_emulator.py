"""CPU pin of the kernels' index arithmetic (tests/kernel_emulator.py replays
the HIP kernels thread by thread with bounds and alignment checks).

Dyadic inputs (multiples of 2^-6 in [-1, 1]) and coefficients (multiples of
2^-5) make every product and partial sum exact in FP32 and FP64, so the
emulated kernels must reproduce fp16(FP64 oracle) bit for bit; any wrong
index, transposed operand or dropped halo shows up as a mismatch."""
import numpy as np
import pytest

import kernel_emulator as E

STENCILS = ("star2d1r", "star2d2r", "box2d1r", "box2d2r")
RADIUS = {"star2d1r": 1, "star2d2r": 2, "box2d1r": 1, "box2d2r": 2}
CASES = [(70, 150, 2, 4), (70, 150, 4, 8), (33, 17, 2, 4), (130, 140, 8, 8), (5, 3, 4, 4), (64, 128, 2, 8)]


def dyadic_problem(stencil, rows, cols, seed):
    rng = np.random.default_rng(seed)
    r = RADIUS[stencil]
    n = 2 * r + 1
    A = rng.integers(-64, 65, (rows, cols)) / 64.0
    B = rng.integers(0, 33, (n, n)) / 32.0
    if stencil.startswith("star"):
        mask = np.zeros((n, n), bool)
        mask[r, :] = mask[:, r] = True
        B = B * mask
    c5 = np.zeros((5, 5))
    c5[2 - r:3 + r, 2 - r:3 + r] = B
    return A, B, c5


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("case", CASES, ids=lambda c: "%dx%d_t%dx%d" % c)
@pytest.mark.parametrize("method", ["mfma", "direct"])
def test_emulated_kernel_exact(oracle, stencil, case, method):
    rows, cols, TH, TW = case
    r = RADIUS[stencil]
    A, B, c5 = dyadic_problem(stencil, rows, cols, seed=rows * 1000 + cols)
    L = E.Layout(rows, cols, max(TH, TW))
    st = E.pack(A, L)
    per_xcd = (rows + cols) % 3  # 0 (one tile per workgroup), 1 or 2 persistent workgroups per XCD
    if method == "mfma":
        out, written = E.mfma_sweep(st, L, TH, TW, c5, r, per_xcd=per_xcd)
    else:
        out, written = E.direct_sweep(st, L, TH, TW, c5, stencil, r, per_xcd=per_xcd)
    # every data brick is written, the ghost ring never
    data = np.zeros(L.plane, dtype=bool)
    for bi in range(L.nbr):
        base = ((bi + 1) * L.sbc + 1) * 256
        data[base:base + L.nbc * 256] = True
    assert np.array_equal(written, data)
    got = E.unpack(out, L)
    want = E.f16(oracle.sweep_f64(stencil, A, B))
    assert np.array_equal(got, want)


def test_emulated_unpack_ring_and_roundtrip():
    rng = np.random.default_rng(1)
    A = rng.integers(-64, 65, (37, 45)) / 64.0
    L = E.Layout(37, 45, 4)
    st = E.pack(A, L)
    assert np.array_equal(E.unpack(st, L), A)
    ring = rng.standard_normal((37, 45))
    got = E.unpack(st, L, ring=ring, rs=2)
    assert np.array_equal(got[2:-2, 2:-2], A[2:-2, 2:-2])
    assert np.array_equal(got[:2], ring[:2]) and np.array_equal(got[:, -2:], ring[:, -2:])


def test_storage_index_is_the_reference_brick_map():
    """With rows = cols = N (N % 16 == 0) and align 1 the storage is the
    reference's padded brick layout: pad by one 16-wide ring, then
    out[(i*nb + j)*256 + x*16 + y] = in[(16i + x)*R + 16j + y]
    (layout16/host_box2d2r.cu:61-86)."""
    N = 48
    L = E.Layout(N, N, 1)
    R = N + 32
    nb = R // 16
    i, j, x, y = np.meshgrid(np.arange(nb), np.arange(nb), np.arange(16), np.arange(16), indexing="ij")
    ref_out = (i * nb + j) * 256 + x * 16 + y
    padded_r, padded_c = 16 * i + x, 16 * j + y  # coordinates in the padded grid
    mine = E.stor_index(padded_r - 16, padded_c - 16, L.sbc)
    assert L.sbc == nb and L.plane == R * R
    assert np.array_equal(mine, ref_out)


@pytest.mark.parametrize("ntile_shape", [(9, 4, 2, 4), (64, 64, 2, 4), (3, 200, 1, 8), (4, 1030, 2, 4)])
def test_persistent_schedule_covers_every_tile(ntile_shape):
    rows_b, cols_b, TH, TW = ntile_shape
    L = E.Layout(16 * rows_b, 16 * cols_b, max(TH, TW))
    for per_xcd in (0, 1, 2, 3, 7, 1000):
        origins = list(E._tiles(L, TH, TW, per_xcd))
        assert len(origins) == len(set(origins)) == (L.nbr // TH) * (L.nbc // TW)
    band = list(E._tiles_band(L, TH, TW))
    assert len(band) == len(set(band)) == (L.nbr // TH) * (L.nbc // TW)


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("case", [(70, 150, 2, 4, 3, 4), (40, 70, 4, 4, 2, 4), (50, 140, 2, 8, 5, 4),
                                  (33, 140, 4, 8, 4, 4), (40, 70, 4, 4, 9, 4), (35, 140, 2, 8, 9, 4),
                                  (140, 150, 8, 8, 6, 4), (70, 140, 4, 8, 9, 8), (33, 150, 4, 8, 3, 8),
                                  (140, 150, 8, 8, 5, 8)],
                         ids=lambda c: "%dx%d_t%dx%d_k%d_w%d" % c)
def test_emulated_fused_equals_repeated_sweeps(oracle, stencil, case):
    """k fused sweeps == k single sweeps, bitwise, and both equal the FP64 oracle
    rounded to FP16 after every sweep (each sweep of dyadic FP16 data with
    taps 1/2, 1/4, 1/4 is exact in FP64); workgroups of 4 waves
    (tcs_mfma_fused) and 8 (tcs_mfma_fused8)."""
    rows, cols, TH, TW, k, nw = case
    r = RADIUS[stencil]
    A, _, _ = dyadic_problem(stencil, rows, cols, seed=rows + cols)
    n = 2 * r + 1
    B = np.zeros((n, n))
    B[r, r] = 0.5
    B[r, 0] = 0.25
    B[0, r] = 0.25
    c5 = np.zeros((5, 5))
    c5[2 - r:3 + r, 2 - r:3 + r] = B
    L = E.Layout(rows, cols, max(TH, TW))
    st = E.pack(A, L)
    fused, written = E.mfma_fused(st, L, TH, TW, c5, r, k, nw=nw)
    data = np.zeros(L.plane, dtype=bool)
    for bi in range(L.nbr):
        base = ((bi + 1) * L.sbc + 1) * 256
        data[base:base + L.nbc * 256] = True
    assert np.array_equal(written, data)
    x = st
    for _ in range(k):
        x, _ = E.mfma_sweep(x, L, TH, TW, c5, r)
        x = np.where(data, x, 0.0)
    assert np.array_equal(fused, x)
    # each single sweep of FP16 data is exact in FP64, then rounded to FP16
    ref = A
    for _ in range(k):
        ref = E.f16(oracle.sweep_f64(stencil, ref, B))
    assert np.array_equal(E.unpack(fused, L), ref)


@pytest.mark.parametrize("stencil", ["star2d1r", "box2d2r"])
def test_emulated_fused_tiny_grids(oracle, stencil):
    """Grids smaller than one tile (1x1 up to 17x1, one padded tile of bricks):
    every fused form's image loads and write-backs stay inside the storage
    plane (the emulator's bounds checks) and k fused sweeps equal k single
    sweeps, on 4-wave and 8-wave workgroups, at 2 sweeps per launch and (1x1,
    17x1) at the most."""
    kmax = 1 + 16 // RADIUS[stencil]
    for rows, cols in ((1, 1), (2, 3), (17, 1), (5, 33), (16, 16)):
        for TH, TW, nw in ((2, 4, 4), (4, 8, 4), (4, 8, 8), (8, 8, 4), (8, 8, 8)):
            for k in (2, kmax) if cols == 1 else (2,):
                test_emulated_fused_equals_repeated_sweeps(oracle, stencil, (rows, cols, TH, TW, k, nw))


@pytest.mark.parametrize("stencil", ["star2d1r", "box2d1r"])
def test_emulated_fused_every_ring_depth_radius1(oracle, stencil):
    """Radius 1 at the most sweeps per launch (17): the ring rows of sweep s are
    packed at every depth 16, 15, ..., 1, odd ones included."""
    test_emulated_fused_equals_repeated_sweeps(oracle, stencil, (36, 70, 2, 4, 17, 4))
    test_emulated_fused_equals_repeated_sweeps(oracle, stencil, (36, 70, 4, 8, 17, 8))


CASES32 = [(70, 150, 4, 4), (70, 150, 4, 8), (33, 17, 4, 4), (130, 140, 8, 8), (5, 3, 2, 8), (64, 128, 2, 8)]


def _data_mask(L):
    data = np.zeros(L.plane, dtype=bool)
    for bi in range(L.nbr):
        base = ((bi + 1) * L.sbc + 1) * 256
        data[base:base + L.nbc * 256] = True
    return data


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("case", CASES32, ids=lambda c: "%dx%d_t%dx%d" % c)
def test_emulated_frag32_exact(oracle, stencil, case):
    """The 32x32-fragment kernel (z folded into K, permuted output columns)
    reproduces fp16(FP64 oracle) exactly on dyadic data, and never reads an
    LDS word that staging did not write (no zero pad in this kernel)."""
    rows, cols, TH, TW = case
    r = RADIUS[stencil]
    A, B, c5 = dyadic_problem(stencil, rows, cols, seed=rows * 1000 + cols + 7)
    L = E.Layout(rows, cols, max(TH, TW))
    st = E.pack(A, L)
    out, written = E.mfma32_sweep(st, L, TH, TW, c5, r)
    assert np.array_equal(written, _data_mask(L))
    assert np.array_equal(E.unpack(out, L), E.f16(oracle.sweep_f64(stencil, A, B)))


def test_frag32_slots_cover_each_shift_and_column_once():
    """Every (shift, haloed column) pair of the 36-column window appears in
    exactly one real K slot; the output column map is a permutation."""
    for rs in (1, 2):
        ZLO, NZ, NMF = E._fold32(rs)
        seen = {}
        for mf in range(NMF):
            for h in range(2):
                for e in range(8):
                    zi, k, real = E._slot32(mf, h, e, NZ)
                    if real:
                        assert (zi, k) not in seen
                        seen[(zi, k)] = (mf, h, e)
        assert set(seen) == {(z, k) for z in range(NZ) for k in range(36)}
        assert NMF == (12 if rs == 2 else 7)
    assert sorted(E.jmap32(m) for m in range(32)) == list(range(32))


@pytest.mark.parametrize("stencil", ["box2d2r", "star2d1r"])
@pytest.mark.parametrize("case", [(70, 150, 4, 4, 3), (33, 140, 4, 8, 4), (40, 40, 8, 8, 2)],
                         ids=lambda c: "%dx%d_t%dx%d_k%d" % c)
def test_emulated_frag32_fused_equals_repeated_sweeps(oracle, stencil, case):
    rows, cols, TH, TW, k = case
    r = RADIUS[stencil]
    A, _, _ = dyadic_problem(stencil, rows, cols, seed=rows + cols + 1)
    n = 2 * r + 1
    B = np.zeros((n, n))
    B[r, r] = 0.5
    B[r, 0] = 0.25
    B[0, r] = 0.25
    c5 = np.zeros((5, 5))
    c5[2 - r:3 + r, 2 - r:3 + r] = B
    L = E.Layout(rows, cols, max(TH, TW))
    st = E.pack(A, L)
    fused, written = E.mfma32_fused(st, L, TH, TW, c5, r, k)
    data = _data_mask(L)
    assert np.array_equal(written, data)
    x = st
    for _ in range(k):
        x, _ = E.mfma32_sweep(x, L, TH, TW, c5, r)
        x = np.where(data, x, 0.0)
    assert np.array_equal(fused, x)
    ref = A
    for _ in range(k):
        ref = E.f16(oracle.sweep_f64(stencil, ref, B))
    assert np.array_equal(E.unpack(fused, L), ref)


EDGE_CASES = [(70, 150, 150, 2, 4), (70, 152, 152, 4, 8), (33, 17, 20, 2, 4), (64, 128, 132, 2, 8), (5, 3, 4, 4, 4),
              (40, 61, 64, 4, 4), (37, 250, 252, 2, 16)]


@pytest.mark.parametrize("stencil", STENCILS)
@pytest.mark.parametrize("case", EDGE_CASES, ids=lambda c: "%dx%d_ld%d_t%dx%d" % c)
def test_emulated_rowmajor_edges(stencil, case):
    """tcs_mfma_edge: the row-major FP32 loader and store reproduce pack ->
    sweep -> unpack exactly; columns past `cols` in each row (the leading-
    dimension gap) hold NaN, so any read there shows."""
    rows, cols, ld, TH, TW = case
    r = RADIUS[stencil]
    A, _, c5 = dyadic_problem(stencil, rows, cols, seed=rows + 7 * cols)
    src = np.full((rows, ld), np.nan)
    src[:, :cols] = A
    src = src.ravel()
    vec = ld % 4 == 0
    L = E.Layout(rows, cols, max(TH, TW))
    st = E.pack(A, L)
    bricks, bw = E.mfma_sweep(st, L, TH, TW, c5, r)
    want = E.unpack(bricks, L, ring=A, rs=r)
    inside = np.zeros((rows, ld), dtype=bool)
    inside[:, :cols] = True
    for mode in (1, 2, 3):
        out, written = E.mfma_edge(mode, src, rows, cols, ld, L, TH, TW, c5, r, vec, inp=st)
        if mode == 1:
            assert np.array_equal(written, bw)
            assert np.array_equal(out[bw], bricks[bw])
        else:
            assert np.array_equal(written.reshape(rows, ld), inside)
            assert np.array_equal(out.reshape(rows, ld)[:, :cols], want)
