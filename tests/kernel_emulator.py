"""CPU emulator of the HIP kernels' index arithmetic -- test infrastructure.

Replays, thread by thread (vectorised over threadIdx), the address and lane
arithmetic of tcstencil_amd/csrc/tcs_transform.hip (pack/unpack) and
tcstencil_amd/csrc/tcs_sweep.hip (stage_tile, tcs_mfma_sweep, tcs_direct_sweep):
the storage index of every global load and store, every LDS read and write, the
MFMA operand lane layouts and the accumulator layout.  Every access is checked
against its buffer's bounds and alignment, and reads of LDS words that no
thread wrote come back as NaN, so a staging hole poisons the output.

MFMA semantics assumed (v_mfma_f32_16x16x32_{f16,bf16}, wave64):
  A (16x32): lane l, element e holds A[l % 16][8*(l//16) + e]
  B (32x16): lane l, element e holds B[8*(l//16) + e][l % 16]
  D (16x16): lane l, element v holds D[4*(l//16) + v][l % 16]
and (v_mfma_f32_32x32x16_{f16,bf16}, wave64):
  A (32x16): lane l, element e holds A[l % 32][8*(l//32) + e]
  B (16x32): lane l, element e holds B[8*(l//32) + e][l % 32]
  D (32x32): lane l, element v holds D[8*(v//4) + 4*(l//32) + v%4][l % 32]
The kernels depend only on the row/column placement and on A and B sharing
one K order; the GPU test tests/test_gpu_parity.py checks the real hardware.

Arithmetic here is float64 (exact for the dyadic inputs the tests use), with
FP16 rounding where the kernels round.
"""
from __future__ import annotations

import numpy as np

WG = 256
XCD = 8


def stor_index(gr, gc, sbc):
    gr = np.asarray(gr, dtype=np.int64)
    gc = np.asarray(gc, dtype=np.int64)
    bi = (gr + 16) >> 4
    bj = (gc + 16) >> 4
    return (bi * sbc + bj) * 256 + ((gr & 15) << 4) + (gc & 15)


class Layout:
    def __init__(self, rows, cols, align):
        self.rows, self.cols, self.align = rows, cols, align
        self.nbr = -(-(-(-rows // 16)) // align) * align
        self.nbc = -(-(-(-cols // 16)) // align) * align
        self.sbr, self.sbc = self.nbr + 2, self.nbc + 2
        self.plane = self.sbr * self.sbc * 256


def _check(idx, size, align=1, what=""):
    idx = np.asarray(idx)
    if idx.size == 0:
        return
    assert idx.min() >= 0 and idx.max() < size, f"{what}: index out of bounds [{idx.min()}, {idx.max()}] size {size}"
    assert np.all(idx % align == 0), f"{what}: misaligned access"


def f16(x):
    return np.asarray(x, dtype=np.float64).astype(np.float16).astype(np.float64)


# ------------------------------------------------------------------ pack
def pack(src: np.ndarray, L: Layout) -> np.ndarray:
    """tcs_pack_kernel, FP16 precision: returns the storage plane (float64 of FP16 values)."""
    out = np.full(L.plane, np.nan)
    cpr = 2 * L.sbc
    t = np.arange(16 * L.sbr * cpr, dtype=np.int64)
    sr, cc = t // cpr, t % cpr
    gr, gc0 = sr - 16, 8 * cc - 16
    s = stor_index(gr, gc0, L.sbc)
    _check(s, L.plane - 7, 8, "pack store")
    for e in range(8):
        gc = gc0 + e
        inside = (gr >= 0) & (gr < L.rows) & (gc >= 0) & (gc < L.cols)
        vals = np.zeros(t.shape)
        vals[inside] = src[gr[inside], gc[inside]]
        out[s + e] = f16(vals)
    assert not np.isnan(out).any(), "pack leaves storage elements unwritten"
    return out


# ---------------------------------------------------------------- unpack
def unpack(st: np.ndarray, L: Layout, ring=None, rs=0) -> np.ndarray:
    q4 = (L.cols + 3) // 4
    t = np.arange(L.rows * q4, dtype=np.int64)
    gr, gc0 = t // q4, 4 * (t % q4)
    s = stor_index(gr, gc0, L.sbc)
    _check(s, L.plane - 3, 4, "unpack load")
    out = np.full((L.rows, L.cols), np.nan)
    for e in range(4):
        gc = gc0 + e
        ok = gc < L.cols
        v = st[s[ok] + e]
        if ring is not None:
            g_r, g_c = gr[ok], gc[ok]
            interior = (g_r >= rs) & (g_r < L.rows - rs) & (g_c >= rs) & (g_c < L.cols - rs)
            v = np.where(interior, v, ring[g_r, g_c])
        out[gr[ok], gc[ok]] = v
    return out


# ------------------------------------------------------------ tile staging
class TileGeom:
    """LDS image of TH x TW bricks plus the 2-point frame: image point (r, c),
    r < LR, c < LP, holds grid point (r0 - 2 + r, c0 - FX + c).  Row-major
    (pitch LP, FX = 2) for the 32x32 and direct kernels; strip=True (16x16 MFMA
    kernels, FX = 4): 8-column strips of NR = 16*TH + 8 rows, stored one after
    another."""

    def __init__(self, TH, TW, strip=False):
        self.TH, self.TW, self.strip = TH, TW, strip
        self.FX = 4 if strip else 2
        self.LR = 16 * TH + 4
        self.LP = 16 * TW + 8
        self.NR = 16 * TH + 8
        self.NCH = TH * TW * 32
        self.CPT = -(-self.NCH // WG)
        self.HD = 8 * TW + 2
        self.NH = 4 * self.HD + 32 * TH
        self.HPT = -(-self.NH // WG)
        self.PLANE = self.LP * (self.NR if strip else self.LR)

    def image(self):
        """LDS offsets of every image point (strip padding rows excluded)."""
        r, c = np.meshgrid(np.arange(self.LR), np.arange(self.LP), indexing="ij")
        return self.at(r, c).ravel()

    def at(self, r, c):
        r, c = np.asarray(r), np.asarray(c)
        assert np.all((r >= 0) & (r < self.LR) & (c >= 0) & (c < self.LP)), "image point out of range"
        if self.strip:
            return (c // 8) * (8 * self.NR) + 8 * r + (c % 8)
        return r * self.LP + c


def _pad_col(G: TileGeom, zc: int) -> int:
    """zero_pad: the image columns holding no grid point, one dword each side
    per row (FX = 4: columns 0,1 and 16TW+6,+7; FX = 2: 16TW+4..+7)."""
    if G.FX == 4:
        return 16 * G.TW + 6 if zc & 1 else 0
    return 16 * G.TW + 4 + 2 * (zc & 1)


def stage_tile(inp: np.ndarray, L: Layout, G: TileGeom, gr0: int, gc0: int, check=False,
               zero_pad=True) -> np.ndarray:
    """load_tile + write_tile (+ zero_pad).  check: frame points outside the
    stored bricks read as zero (the fused kernel's image includes the ghost ring).
    The 32x32-fragment kernels skip zero_pad; LDS words nobody wrote stay NaN."""
    lds = np.full(G.PLANE, np.nan)
    tid = np.arange(WG, dtype=np.int64)
    TW, TH = G.TW, G.TH
    for i in range(G.CPT):
        c = tid + i * WG
        c = c[c < G.NCH]
        b, w = c >> 5, c & 31
        br, bc = b // TW, b % TW
        s = stor_index(gr0 + 16 * br + (w >> 1), gc0 + 16 * bc + 8 * (w & 1), L.sbc)
        _check(s, L.plane - 7, 8, "centre load (16 B)")
        lr = 16 * br + (w >> 1) + 2
        lc = 16 * bc + 8 * (w & 1) + G.FX
        if G.FX == 4:  # two 8-byte writes: the chunk straddles two strips
            for q in range(2):
                e = G.at(lr, lc + 4 * q)
                _check(e, G.PLANE - 3, 4, "centre LDS write (8 B)")
                for t in range(4):
                    lds[e + t] = inp[s + 4 * q + t]
        else:
            for q in range(4):  # four dword writes
                e = G.at(lr, lc + 2 * q)
                _check(e, G.PLANE - 1, 2, "centre LDS write (dword)")
                lds[e] = inp[s + 2 * q]
                lds[e + 1] = inp[s + 2 * q + 1]
    for i in range(G.HPT):
        h = tid + i * WG
        lr = np.empty_like(h)
        ld = np.empty_like(h)
        m1 = h < 2 * G.HD
        lr[m1], ld[m1] = h[m1] // G.HD, h[m1] % G.HD
        m2 = (~m1) & (h < 4 * G.HD)
        q = h[m2] - 2 * G.HD
        lr[m2], ld[m2] = 16 * TH + 2 + q // G.HD, q % G.HD
        m3 = (~m1) & (~m2) & (h < 4 * G.HD + 16 * TH)
        lr[m3], ld[m3] = 2 + (h[m3] - 4 * G.HD), 0
        m4 = ~(m1 | m2 | m3)
        lr[m4], ld[m4] = 2 + (h[m4] - 4 * G.HD - 16 * TH), G.HD - 1
        act = h < G.NH
        gr, gc = gr0 - 2 + lr[act], gc0 - 2 + 2 * ld[act]
        inside = np.ones(gr.shape, dtype=bool)
        if check:
            inside = (((gr + 16) >> 4) >= 0) & (((gr + 16) >> 4) < L.sbr) & (((gc + 16) >> 4) >= 0) & (((gc + 16) >> 4) < L.sbc)
        s = stor_index(gr, gc, L.sbc)
        _check(s[inside], L.plane - 1, 2, "halo load (4 B)")
        d = G.at(lr[act], G.FX - 2 + 2 * ld[act]) >> 1
        _check(2 * d, G.PLANE - 1, 2, "halo LDS write")
        v0, v1 = np.zeros(gr.shape), np.zeros(gr.shape)
        v0[inside], v1[inside] = inp[s[inside]], inp[s[inside] + 1]
        lds[2 * d] = v0
        lds[2 * d + 1] = v1
    for zc in range(2 * G.LR if zero_pad else 0):
        d = int(G.at(zc >> 1, _pad_col(G, zc))) >> 1
        assert 0 <= 2 * d < G.PLANE
        lds[2 * d] = 0.0
        lds[2 * d + 1] = 0.0
    return lds


def _tiles_band(L: Layout, TH: int, TW: int):
    """The row-band schedule of the kernels that read row-major FP32 input:
    tile rows in bands of 8 (padded with empty rows), virtual tile v of a band
    at tile row 8 * band + v % 8, tile column (v % (8 tiles_x)) // 8; one
    workgroup per virtual tile, on XCD v % 8, so tile row ty stays on XCD
    ty % 8."""
    tiles_x, tiles_y = L.nbc // TW, L.nbr // TH
    nvirt = -(-tiles_y // XCD) * XCD * tiles_x
    seen = []
    for v in range(nvirt):
        band, rem = divmod(v, XCD * tiles_x)
        ty, tx = XCD * band + (rem & 7), rem >> 3
        if ty >= tiles_y:
            continue
        assert tx < tiles_x and ty % XCD == v % XCD, "a tile row left its XCD"
        seen.append(ty * tiles_x + tx)
        yield 16 * TH * ty, 16 * TW * tx
    assert sorted(seen) == list(range(tiles_y * tiles_x)), "schedule does not cover every tile exactly once"


def _tiles(L: Layout, TH: int, TW: int, per_xcd: int = 0):
    """The tile schedule of tcs_sweep.hip: tile rows padded to tiles_xp (a
    multiple of 8) virtual columns, virtual tiles with tx >= tiles_x empty;
    grid = nvirt workgroups (per_xcd = 0: one per virtual tile) or 8 * per_xcd
    (persistent); workgroup w walks v = w, w + grid, ... below nvirt, skipping
    empty tiles.  Workgroup w runs on XCD w % 8, so tile column tx stays on XCD
    tx % 8."""
    tiles_x = L.nbc // TW
    tiles_xp = -(-tiles_x // XCD) * XCD
    nvirt = (L.nbr // TH) * tiles_xp
    grid = nvirt if per_xcd <= 0 or per_xcd * XCD >= nvirt else per_xcd * XCD
    assert grid % XCD == 0
    seen = []
    for wg in range(grid):
        v = wg
        while v < nvirt:
            ty, tx = divmod(v, tiles_xp)
            if tx < tiles_x:
                assert tx % XCD == wg % XCD, "a tile column left its XCD"
                seen.append(ty * tiles_x + tx)
                yield 16 * TH * ty, 16 * TW * tx
            v += grid
    assert sorted(seen) == list(range((L.nbr // TH) * tiles_x)), "schedule does not cover every tile exactly once"


def _interior(gr, gc, L: Layout, rs: int):
    return (gr >= rs) & (gr < L.rows - rs) & (gc >= rs) & (gc < L.cols - rs)


# ------------------------------------------------------------- MFMA sweep
# (row shift index, column group) of the K slot groups g = 0..3 of MFMA f:
# fragments 0..2 are the column groups 0, 2, 1 with lane group g on shift g;
# radius 2 adds fragment 3, shift 4 with column groups 0, 2, 1, 1 (the last
# against zero taps).  Radius 1 (shifts 0..2) reads shift 3 in lane group 3
# against zero taps.  Fragment 0 of a brick is fragment 1 of its left neighbour.
_KG = (0, 2, 1)


def _kmap(rs, f, g):
    """KMap of tcs_sweep.hip: (row shift index, column group, real) of K slot
    group g (array of lane groups) of MFMA f."""
    g = np.asarray(g)
    nz = 5 if rs == 2 else 3
    if f < 3:
        return g.copy(), np.full_like(g, _KG[f]), g < nz
    return np.full_like(g, 4), np.array([0, 2, 1, 1])[g], g < 3


def _nfrag(rs):
    return 4 if rs == 2 else 3


def _afrag(coeff5, rs, round_coeff=f16):
    ZLO, ZHI = (0, 4) if rs == 2 else (1, 3)
    lane = np.arange(64)
    m, g = lane & 15, lane >> 4
    afrag = np.zeros((_nfrag(rs), 64, 8))
    for f in range(_nfrag(rs)):
        zi, kg, real = _kmap(rs, f, g)
        for e in range(8):
            d = 8 * kg + e - m - 2  # output j is window column j + 2 (the strip image's FX = 4)
            ok = real & (d >= 0) & (d <= 4)
            afrag[f, ok, e] = round_coeff(coeff5[ZLO + zi[ok], d[ok]])
    return afrag, ZLO, ZHI


def _mfma_brick(lds, G: TileGeom, afrag, ZLO, ZHI, br, bc, edge, gr, gcb, L: Layout, rs, row=None, frags_out=None):
    """Ops::load + mfma_brick: returns o[lane][v] (FP16 values) for LDS brick
    (br, bc).  Every lane reads 16 bytes per fragment (KMap).  row[lane]: the
    lane's output row counted from the brick's first row (default lane % 16;
    the fused kernel's paired ring units pass other rows)."""
    lane = np.arange(64)
    m, g = lane & 15, lane >> 4
    row = m if row is None else row
    D = np.zeros((16, 16))
    frags = []
    for f in range(_nfrag(rs)):
        zi, kg, _ = _kmap(rs, f, g)
        off = G.at(16 * br + row + ZLO + zi, 16 * bc + 8 * kg)
        _check(off, G.PLANE - 7, 8, "B operand LDS read (16 B)")
        bfrag = lds[off[:, None] + np.arange(8)[None, :]]  # [lane][e]
        frags.append(bfrag)
        Am = np.zeros((16, 32))
        Bm = np.zeros((32, 16))
        for l in range(64):
            Am[l % 16, 8 * (l // 16):8 * (l // 16) + 8] = afrag[f, l]
            Bm[8 * (l // 16):8 * (l // 16) + 8, l % 16] = bfrag[l]
        D += Am @ Bm
    # lane l holds D[4g+v][n], n = l % 16: output row n, column 4g+v
    o = np.stack([f16(D[4 * g + v, m]) for v in range(4)], axis=1)  # [lane][v]
    if edge:
        for v in range(4):
            cidx = G.at(16 * br + row + 2, 16 * bc + 4 * g + G.FX + v)
            _check(cidx, G.PLANE - 1, 1, "centre read")
            keep = ~_interior(gr, gcb + v, L, rs)
            o[keep, v] = lds[cidx[keep]]
    if frags_out is not None:
        frags_out[:] = frags
    return o


def _check_reuse(frags, prev):
    """The kernel takes fragment 0 of a brick from its left neighbour's
    fragment 1 instead of reading it: both must hold the same LDS words."""
    assert prev is not None, "reuse without a previous unit"
    a, b = frags[0], prev[1]
    assert np.array_equal(np.isnan(a), np.isnan(b)) and np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)]), \
        "fragment 0 differs from the left neighbour's fragment 1"


def _store(out, written, L, gr, gcb, o):
    s = stor_index(gr, gcb, L.sbc)
    _check(s, L.plane - 3, 4, "output store (8 B)")
    for v in range(4):
        assert not written[s + v].any(), "two lanes store the same element"
        out[s + v] = o[:, v]
        written[s + v] = True


def mfma_sweep(inp, L: Layout, TH, TW, coeff5, rs, round_coeff=f16, per_xcd=0):
    """tcs_mfma_sweep for FP16 (one plane).  coeff5: 5x5 FP32 embedding."""
    assert L.nbr % TH == 0 and L.nbc % TW == 0
    G = TileGeom(TH, TW, strip=True)
    afrag, ZLO, ZHI = _afrag(coeff5, rs, round_coeff)
    out = np.zeros(L.plane)
    written = np.zeros(L.plane, dtype=bool)
    m, g = np.arange(64) & 15, np.arange(64) >> 4
    for gr0, gc0 in _tiles(L, TH, TW, per_xcd):
        lds = stage_tile(inp, L, G, gr0, gc0)
        edge = not (gr0 >= rs and gr0 + 16 * TH <= L.rows - rs and gc0 >= rs and gc0 + 16 * TW <= L.cols - rs)
        for wave in range(4):
            prev = None
            for q in range(TH * TW // 4):
                br, bc = _block_brick(TH, TW, 0, wave, q)
                gr, gcb = gr0 + 16 * br + m, gc0 + 16 * bc + 4 * g
                frags = []
                o = _mfma_brick(lds, G, afrag, ZLO, ZHI, br, bc, edge, gr, gcb, L, rs, frags_out=frags)
                if q % _block_rw(TH, TW) != 0:  # sweep_tile16's reuse(q)
                    _check_reuse(frags, prev)
                prev = frags
                _store(out, written, L, gr, gcb, o)
    return out, written


# ------------------------------------------- row-major FP32 edges (tcs_mfma_edge)
def stage_tile_rm(src, rows, cols, ld, G: TileGeom, gr0: int, gc0: int, vec: bool) -> np.ndarray:
    """load_tile_rm + write_tile + zero_pad: the LDS image read from the
    row-major grid src (flat, rows x cols with leading dimension ld).  Every
    load is checked against [0, rows*ld) and its row; vector loads (two 16-byte
    loads per 8-element chunk, when vec and the chunk lies inside the grid)
    also for 16-byte alignment of the element index."""
    n = rows * ld
    lds = np.full(G.PLANE, np.nan)
    tid = np.arange(WG, dtype=np.int64)
    TW, TH = G.TW, G.TH
    for i in range(G.CPT):
        c = tid + i * WG
        c = c[c < G.NCH]
        lrow, cc = c // (2 * TW), c % (2 * TW)  # row-major chunk order over the tile
        gr, gc = gr0 + lrow, gc0 + 8 * cc
        x = np.zeros((c.size, 8))
        vecok = (gr < rows) & (gc + 8 <= cols) & vec
        idx = gr[vecok] * ld + gc[vecok]
        _check(idx, n - 7, 4, "row-major vector load (2 x 16 B)")
        assert np.all(gc[vecok] + 7 < cols), "vector load leaves its row"
        x[vecok] = src[idx[:, None] + np.arange(8)[None, :]]
        for e in range(8):
            ok = ~vecok & (gr < rows) & (gc + e < cols)
            ie = gr[ok] * ld + gc[ok] + e
            _check(ie, n, 1, "row-major load")
            x[ok, e] = src[ie]
        lr, lc = lrow + 2, 8 * cc + G.FX
        step = 4 if G.FX == 4 else 2  # STRIP: two 8-byte writes per chunk, else four dwords
        for q in range(8 // step):
            e0 = G.at(lr, lc + step * q)
            _check(e0, G.PLANE - step + 1, step, "row-major chunk LDS write")
            for t in range(step):
                assert np.all(np.isnan(lds[e0 + t])), "two chunks write one LDS word"
                lds[e0 + t] = f16(x[:, step * q + t])
    for i in range(G.HPT):
        h = tid + i * WG
        h = h[h < G.NH]
        lr = np.empty_like(h)
        ldw = np.empty_like(h)
        m1 = h < 2 * G.HD
        lr[m1], ldw[m1] = h[m1] // G.HD, h[m1] % G.HD
        m2 = (~m1) & (h < 4 * G.HD)
        q = h[m2] - 2 * G.HD
        lr[m2], ldw[m2] = 16 * TH + 2 + q // G.HD, q % G.HD
        m3 = (~m1) & (~m2) & (h < 4 * G.HD + 16 * TH)
        lr[m3], ldw[m3] = 2 + (h[m3] - 4 * G.HD), 0
        m4 = ~(m1 | m2 | m3)
        lr[m4], ldw[m4] = 2 + (h[m4] - 4 * G.HD - 16 * TH), G.HD - 1
        gr, gc = gr0 - 2 + lr, gc0 - 2 + 2 * ldw
        d = G.at(lr, G.FX - 2 + 2 * ldw)
        for e in range(2):
            ok = (gr >= 0) & (gr < rows) & (gc + e >= 0) & (gc + e < cols)
            ie = gr[ok] * ld + gc[ok] + e
            _check(ie, n, 1, "row-major halo load")
            v = np.zeros(h.size)
            v[ok] = src[ie]
            lds[d + e] = f16(v)
    for zc in range(2 * G.LR):
        d = int(G.at(zc >> 1, _pad_col(G, zc)))
        lds[d] = lds[d + 1] = 0.0
    return lds


def mfma_edge(mode, src, rows, cols, ld, L: Layout, TH, TW, coeff5, rs, vec, inp=None):
    """tcs_mfma_edge (FP16): mode 1 row-major src -> bricks, 2 bricks (inp) ->
    row-major, 3 both.  Returns (bricks, written) or (row-major flat, written);
    every row-major store is bounds- and alignment-checked."""
    assert L.nbr % TH == 0 and L.nbc % TW == 0
    G = TileGeom(TH, TW, strip=True)
    afrag, ZLO, ZHI = _afrag(coeff5, rs)
    n = rows * ld
    out = np.zeros(n if mode & 2 else L.plane)
    written = np.zeros(out.size, dtype=bool)
    m, g = np.arange(64) & 15, np.arange(64) >> 4
    OP = 16 * TW + 8  # pitch of the row-major FP16 output image (over the input image)
    assert 16 * TH * OP <= G.PLANE
    for gr0, gc0 in (_tiles_band(L, TH, TW) if mode & 1 else _tiles(L, TH, TW, 0)):
        lds = stage_tile_rm(src, rows, cols, ld, G, gr0, gc0, vec) if mode & 1 else stage_tile(inp, L, G, gr0, gc0)
        edge = not (gr0 >= rs and gr0 + 16 * TH <= L.rows - rs and gc0 >= rs and gc0 + 16 * TW <= L.cols - rs)
        oimg = np.full(G.PLANE, np.nan)
        for wave in range(4):
            for q in range(TH * TW // 4):
                b = wave + 4 * q
                br, bc = b // TW, b % TW
                gr, gcb = gr0 + 16 * br + m, gc0 + 16 * bc + 4 * g
                o = _mfma_brick(lds, G, afrag, ZLO, ZHI, br, bc, edge, gr, gcb, L, rs)
                if not mode & 2:
                    _store(out, written, L, gr, gcb, o)
                    continue
                e0 = (16 * br + m) * OP + 16 * bc + 4 * g  # one 8-byte LDS write per lane
                _check(e0, G.PLANE - 3, 4, "output image write (8 B)")
                for v in range(4):
                    assert np.all(np.isnan(oimg[e0 + v])), "two lanes write one output element"
                    oimg[e0 + v] = o[:, v]
        if not mode & 2:
            continue
        tid = np.arange(WG)
        NP = 16 * TH * 4 * TW  # store_rm4: pieces of 4 results of one row per thread
        assert NP % WG == 0
        for i in range(NP // WG):
            c = tid + i * WG
            lrow, pc = c // (4 * TW), c % (4 * TW)
            e0 = lrow * OP + 4 * pc
            _check(e0, G.PLANE - 3, 4, "output image read (8 B)")
            vals = oimg[e0[:, None] + np.arange(4)[None, :]]
            assert not np.isnan(vals).any(), "output image read an unwritten element"
            for t in range(c.size):
                gr, gc = gr0 + lrow[t], gc0 + 4 * pc[t]
                if gr >= rows:
                    continue
                v = vals[t].copy()
                for e in range(4):
                    if gc + e < cols and not _interior(gr, gc + e, L, rs):
                        assert edge, "ring point in a tile not marked edge"
                        v[e] = src[gr * ld + gc + e]
                base = gr * ld + gc
                if vec and gc + 4 <= cols:
                    _check(np.array([base]), n - 3, 4, "row-major vector store (16 B)")
                    es = range(4)
                else:
                    es = [e for e in range(4) if gc + e < cols]
                    _check(np.array([base + e for e in es], dtype=np.int64), n, 1, "row-major store")
                for e in es:
                    assert not written[base + e], "two threads store the same element"
                    out[base + e] = v[e]
                    written[base + e] = True
    return out, written


def _block_rw(BH, BW, nw=4):
    """BlockRows<BH, BW, NW>::RW: bricks per row of a wave's block."""
    return BW if BH % nw == 0 else BW // 2


def _block_brick(BH, BW, base, wave, q, nw=4):
    """block_brick of tcs_sweep.hip: with BH a multiple of NW wave w takes BH/NW
    whole rows from row w * BH/NW; else the NW waves split a BH x BW brick
    region into NW/2 row bands by 2 column halves, wave w taking the
    (2BH/NW) x (BW/2) block at (w // 2, w % 2)."""
    if BH % nw == 0:
        return base + wave * (BH // nw) + q // BW, base + q % BW
    assert BH % (nw // 2) == 0 and BW % 2 == 0
    hb, wb = 2 * BH // nw, BW // 2
    return base + (wave >> 1) * hb + q // wb, base + (wave & 1) * wb + q % wb


def mfma_fused(inp, L: Layout, TH, TW, coeff5, rs, nsweeps, round_coeff=f16, nw=4):
    """tcs_mfma_fused (nw = 4 waves per workgroup) and tcs_mfma_fused8 (nw = 8):
    nsweeps sweeps per tile in an LDS image of the tile plus a one-brick ring
    and the 2-point frame, rewritten in place between sweeps."""
    assert L.nbr % TH == 0 and L.nbc % TW == 0 and 1 <= nsweeps and rs * (nsweeps - 1) <= 16
    G = TileGeom(TH + 2, TW + 2, strip=True)
    afrag, ZLO, ZHI = _afrag(coeff5, rs, round_coeff)
    out = np.zeros(L.plane)
    written = np.zeros(L.plane, dtype=bool)
    m, g = np.arange(64) & 15, np.arange(64) >> 4
    BH, BW = TH + 2, TW + 2
    for gr0, gc0 in _tiles(L, TH, TW, 0):
        er0, ec0 = gr0 - 16, gc0 - 16
        lds = stage_tile(inp, L, G, er0, ec0, check=True)
        assert not np.isnan(lds[G.image()]).any(), "fused image has unwritten LDS words"
        edge = not (er0 >= rs and er0 + 16 * BH <= L.rows - rs and ec0 >= rs and ec0 + 16 * BW <= L.cols - rs)
        QB = TH * BW // nw
        mp = m + 8 + np.where(m >= 8, 16 * TH, 0)  # paired ring unit lane rows from brick row 0
        for sw in range(nsweeps - 1):
            paired = rs * (nsweeps - 1 - sw) <= 8
            wpr = nw if paired else nw // 2  # waves per ring row
            rc, rx = BW // wpr, BW % wpr
            qt = QB + rc + (rx > 0)
            res, rows_done = [], {}
            for wave in range(nw):
                prev = None
                wi = wave % wpr
                ring_br = 0 if paired or wave < wpr else TH + 1
                ring_bc = wi * rc + min(wi, rx)
                for q in range(qt):
                    if q < QB:
                        br, bc = _block_brick(TH, BW, 0, wave, q, nw)
                        br, row, pair = br + 1, m, False
                    else:
                        j = q - QB
                        if j >= rc and wi >= rx:
                            continue
                        br, bc, row, pair = ring_br, ring_bc + j, (mp if paired else m), paired
                    # every output row of every computed brick, once
                    for rr in np.unique(16 * br + row):
                        key = (int(rr), bc)
                        assert key not in rows_done, "two units compute one brick row"
                        rows_done[key] = True
                    frags = []
                    o = _mfma_brick(lds, G, afrag, ZLO, ZHI, br, bc, edge, er0 + 16 * br + row,
                                    ec0 + 16 * bc + 4 * g, L, rs, row=row, frags_out=frags)
                    if (q % _block_rw(TH, BW, nw) != 0) if q < QB else (q > QB):  # fused_sweep's reuse(q)
                        _check_reuse(frags, prev)
                    prev = frags
                    res.append((br, bc, row, o))
            # rows the sweep must produce: within r * (nsweeps - 1 - sw) of the core
            h = rs * (nsweeps - 1 - sw)
            for rr in range(16 - h, 16 * (TH + 1) + h):
                for bc in range(BW):
                    assert (rr, bc) in rows_done, f"sweep {sw}: image row {rr} brick column {bc} not computed"
            for br, bc, row, o in res:  # one 8-byte write per lane
                d = G.at(16 * br + row + 2, 16 * bc + 4 * g + G.FX)
                _check(d, G.PLANE - 3, 4, "fused write-back (8 B)")
                for v in range(4):
                    lds[d + v] = o[:, v]
        for wave in range(nw):
            prev = None
            for q in range(TH * TW // nw):
                br, bc = _block_brick(TH, TW, 1, wave, q, nw)
                gr, gcb = er0 + 16 * br + m, ec0 + 16 * bc + 4 * g
                frags = []
                o = _mfma_brick(lds, G, afrag, ZLO, ZHI, br, bc, edge, gr, gcb, L, rs, frags_out=frags)
                if q % _block_rw(TH, TW, nw) != 0:  # fused_tile's reuse(q)
                    _check_reuse(frags, prev)
                prev = frags
                _store(out, written, L, gr, gcb, o)
    return out, written


# ------------------------------------------------- MFMA, 32x32 fragments
def _fold32(rs):
    ZLO, ZHI = (0, 4) if rs == 2 else (1, 3)
    NZ = ZHI - ZLO + 1
    NG = 4 * NZ + (NZ + 1) // 2
    return ZLO, NZ, (NG + 1) // 2


def _slot32(mf, h, e, NZ):
    """K slot (group 2*mf + h, element e) -> (shift index, haloed column, real)."""
    if mf < 2 * NZ:
        return mf >> 1, 8 * (2 * (mf & 1) + h) + e, True
    z = 2 * (2 * (mf - 2 * NZ) + h) + (e >> 2)
    return (z if z < NZ else 0), 32 + (e & 3), z < NZ


def jmap32(mu):
    return 16 * (mu >> 4) + 8 * ((mu >> 2) & 1) + 4 * ((mu >> 3) & 1) + (mu & 3)


def _afrag32(coeff5, rs, round_coeff=f16):
    ZLO, NZ, NMF = _fold32(rs)
    af = np.zeros((NMF, 64, 8))
    for mf in range(NMF):
        for l in range(64):
            j, h = jmap32(l % 32), l // 32
            for e in range(8):
                zi, k, real = _slot32(mf, h, e, NZ)
                d = k - j
                if real and 0 <= d <= 4:
                    af[mf, l, e] = round_coeff(coeff5[ZLO + zi, d])
    return af


def _mfma_block32(lds, G: TileGeom, af, rs, br, bc, edge, gr, gcb, L: Layout):
    """Ops32::load + mma_chunk32 + epilogue32: o[lane][v] for the block whose
    window starts at LDS brick (br, bc); lane (i, h) owns row i, columns
    16*(v>>3) + 8h + (v&7)."""
    ZLO, NZ, NMF = _fold32(rs)
    lane = np.arange(64)
    i, h = lane & 31, lane >> 5
    base = (16 * br + i + ZLO) * G.LP + 16 * bc
    D = np.zeros((32, 32))
    for mf in range(NMF):
        if mf < 2 * NZ:
            off = base + (mf >> 1) * G.LP + 8 * (2 * (mf & 1) + h)
            _check(off, G.PLANE - 7, 8, "B operand LDS read (16 B)")
            bfrag = lds[off[:, None] + np.arange(8)[None, :]]
        else:
            p = 2 * (mf - 2 * NZ) + h
            za, zb = np.minimum(2 * p, NZ - 1), np.minimum(2 * p + 1, NZ - 1)
            oa, ob = base + za * G.LP + 32, base + zb * G.LP + 32
            _check(oa, G.PLANE - 3, 4, "B operand LDS read (8 B)")
            _check(ob, G.PLANE - 3, 4, "B operand LDS read (8 B)")
            bfrag = np.concatenate([lds[oa[:, None] + np.arange(4)[None, :]],
                                    lds[ob[:, None] + np.arange(4)[None, :]]], axis=1)
        Am = np.zeros((32, 16))
        Bm = np.zeros((16, 32))
        for l in range(64):
            Am[l % 32, 8 * (l // 32):8 * (l // 32) + 8] = af[mf, l]
            Bm[8 * (l // 32):8 * (l // 32) + 8, l % 32] = bfrag[l]
        D += Am @ Bm
    o = np.zeros((64, 16))
    for v in range(16):
        mu = 8 * (v // 4) + 4 * h + v % 4
        o[:, v] = f16(D[mu, i])
    if edge:
        row = (16 * br + i + 2) * G.LP + 16 * bc + 2
        for v in range(16):
            cb = 16 * (v >> 3) + 8 * h + (v & 7)
            _check(row + cb, G.PLANE, 1, "centre read")
            keep = ~_interior(gr, gcb + cb, L, rs)
            o[keep, v] = lds[row[keep] + cb[keep]]
    return o


def _store32(out, written, L, gr, gcb, h, o):
    for s in range(2):
        st = stor_index(gr, gcb + 16 * s + 8 * h, L.sbc)
        _check(st, L.plane - 7, 8, "output store (16 B)")
        for t in range(8):
            assert not written[st + t].any(), "two lanes store the same element"
            out[st + t] = o[:, 8 * s + t]
            written[st + t] = True


def _blocks32(wave, nq, nblk, bx, base):
    for q in range(nq):
        b = wave + 4 * q
        if b >= nblk:
            break
        yield base + 2 * (b // bx), base + 2 * (b % bx)


def mfma32_sweep(inp, L: Layout, TH, TW, coeff5, rs, round_coeff=f16):
    """tcs_mfma32_sweep for FP16."""
    assert L.nbr % TH == 0 and L.nbc % TW == 0 and TH % 2 == 0 and TW % 2 == 0
    BX, NBLK = TW // 2, (TH // 2) * (TW // 2)
    assert NBLK % 4 == 0
    G = TileGeom(TH, TW)
    af = _afrag32(coeff5, rs, round_coeff)
    out = np.zeros(L.plane)
    written = np.zeros(L.plane, dtype=bool)
    i, h = np.arange(64) & 31, np.arange(64) >> 5
    for gr0, gc0 in _tiles(L, TH, TW, 0):
        lds = stage_tile(inp, L, G, gr0, gc0, zero_pad=False)
        edge = not (gr0 >= rs and gr0 + 16 * TH <= L.rows - rs and gc0 >= rs and gc0 + 16 * TW <= L.cols - rs)
        for wave in range(4):
            for br, bc in _blocks32(wave, NBLK // 4, NBLK, BX, 0):
                gr, gcb = gr0 + 16 * br + i, gc0 + 16 * bc
                o = _mfma_block32(lds, G, af, rs, br, bc, edge, gr, gcb, L)
                assert not np.isnan(o).any(), "block read unwritten LDS"
                _store32(out, written, L, gr, gcb, h, o)
    return out, written


def mfma32_fused(inp, L: Layout, TH, TW, coeff5, rs, nsweeps, round_coeff=f16):
    """tcs_mfma32_fused: the fused scheme over 2x2-brick blocks."""
    assert L.nbr % TH == 0 and L.nbc % TW == 0 and 1 <= nsweeps and rs * (nsweeps - 1) <= 16
    BH, BW = TH + 2, TW + 2
    BXA, NBA = BW // 2, (BH // 2) * (BW // 2)
    QA = -(-NBA // 4)
    BXC, NBC = TW // 2, (TH // 2) * (TW // 2)
    G = TileGeom(BH, BW)
    af = _afrag32(coeff5, rs, round_coeff)
    out = np.zeros(L.plane)
    written = np.zeros(L.plane, dtype=bool)
    i, h = np.arange(64) & 31, np.arange(64) >> 5
    for gr0, gc0 in _tiles(L, TH, TW, 0):
        er0, ec0 = gr0 - 16, gc0 - 16
        lds = stage_tile(inp, L, G, er0, ec0, check=True, zero_pad=False)
        edge = not (er0 >= rs and er0 + 16 * BH <= L.rows - rs and ec0 >= rs and ec0 + 16 * BW <= L.cols - rs)
        for sw in range(nsweeps - 1):
            res = []
            for wave in range(4):
                for br, bc in _blocks32(wave, QA, NBA, BXA, 0):
                    o = _mfma_block32(lds, G, af, rs, br, bc, edge, er0 + 16 * br + i, ec0 + 16 * bc, L)
                    res.append((br, bc, o))
            covered = np.zeros(G.PLANE, dtype=bool)
            for br, bc, o in res:
                d = ((16 * br + i + 2) * G.LP + 16 * bc + 8 * h + 2) >> 1
                for s_ in range(2):
                    for t in range(4):
                        e0 = 2 * (d + 8 * s_ + t)
                        _check(e0, G.PLANE - 1, 2, "fused write-back")
                        assert not covered[e0].any()
                        covered[e0] = covered[e0 + 1] = True
                        lds[e0] = o[:, 8 * s_ + 2 * t]
                        lds[e0 + 1] = o[:, 8 * s_ + 2 * t + 1]
        for wave in range(4):
            for br, bc in _blocks32(wave, NBC // 4, NBC, BXC, 1):
                gr, gcb = er0 + 16 * br + i, ec0 + 16 * bc
                o = _mfma_block32(lds, G, af, rs, br, bc, edge, gr, gcb, L)
                assert not np.isnan(o).any(), "block read unwritten LDS"
                _store32(out, written, L, gr, gcb, h, o)
    return out, written


# ----------------------------------------------------------- direct sweep
def _taps(stencil):
    t = np.zeros((5, 5), dtype=bool)
    if stencil == "star2d1r":
        t[2, 1:4] = True
        t[1, 2] = t[3, 2] = True
    elif stencil == "star2d2r":
        t[2, :] = True
        t[:, 2] = True
    elif stencil == "box2d1r":
        t[1:4, 1:4] = True
    else:
        t[:, :] = True
    return t


def direct_sweep(inp, L: Layout, TH, TW, coeff5, stencil, rs, per_xcd=0):
    G = TileGeom(TH, TW)
    CG = 4 * TW
    RG = WG // CG
    MR = 16 * TH // RG
    assert RG * CG == WG and MR * RG == 16 * TH
    taps = _taps(stencil)
    out = np.zeros(L.plane)
    written = np.zeros(L.plane, dtype=bool)
    tid = np.arange(WG)
    cg, rg = tid % CG, tid // CG
    r0 = rg * MR
    for gr0, gc0 in _tiles(L, TH, TW, per_xcd):
        lds = stage_tile(inp, L, G, gr0, gc0)
        acc = np.zeros((WG, MR, 4))
        for ir in range(MR + 4):
            base = (r0 + ir) * G.LP + 4 * cg
            _check(base, G.PLANE - 7, 4, "direct LDS read (8 B)")
            x = lds[base[:, None] + np.arange(8)[None, :]]
            for o in range(MR):
                dz = ir - o
                if dz < 0 or dz > 4:
                    continue
                for dj in range(5):
                    if taps[dz, dj]:
                        acc[:, o, :] += x[:, dj:dj + 4] * np.float32(coeff5[dz, dj])
        edge = not (gr0 >= rs and gr0 + 16 * TH <= L.rows - rs and gc0 >= rs and gc0 + 16 * TW <= L.cols - rs)
        # results -> row-major output image over the input image (pitch OP);
        # box2d2r stores 8-byte pieces straight from registers
        staged = stencil != "box2d2r"
        OP = 16 * TW + 8
        assert 16 * TH * OP <= G.PLANE
        img = np.full(G.PLANE, np.nan)
        for o in range(MR):
            tr = r0 + o
            gr = gr0 + tr
            gcb = gc0 + 4 * cg
            h = f16(acc[:, o, :])
            if edge:
                cidx = (tr + 2) * G.LP + 4 * cg + 2
                for v in range(4):
                    keep = ~_interior(gr, gcb + v, L, rs)
                    h[keep, v] = lds[cidx[keep] + v]
            if not staged:
                s = stor_index(gr, gcb, L.sbc)
                _check(s, L.plane - 3, 4, "direct store (8 B)")
                for v in range(4):
                    assert not written[s + v].any()
                    out[s + v] = h[:, v]
                    written[s + v] = True
                continue
            d = tr * OP + 4 * cg
            _check(d, G.PLANE - 3, 4, "direct output image write (8 B)")
            for v in range(4):
                img[d + v] = h[:, v]
        if not staged:
            continue
        # whole bricks out: 16-byte pieces, consecutive threads along a brick
        nch = TH * TW * 32
        assert nch % WG == 0
        for i in range(nch // WG):
            c = tid + i * WG
            b, w = c >> 5, c & 31
            br, bc = b // TW, b % TW
            tr, tc = 16 * br + (w >> 1), 16 * bc + 8 * (w & 1)
            src = tr * OP + tc
            _check(src, G.PLANE - 7, 8, "direct output image read (16 B)")
            s = stor_index(gr0 + tr, gc0 + tc, L.sbc)
            _check(s, L.plane - 7, 8, "direct store (16 B)")
            for e in range(8):
                v = img[src + e]
                assert not np.isnan(v).any(), "output image read before written"
                assert not written[s + e].any()
                out[s + e] = v
                written[s + e] = True
    return out, written
