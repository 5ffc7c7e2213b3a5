// One Jacobi sweep over brick storage: the MFMA kernel (the TCStencil path,
// replacing layout16/kernel_<stencil>.cu::mma_run) and the direct-FMA kernel
// (replacing baseline/kernel_<stencil>.cu::kernel_run).
//
// Work unit: a persistent workgroup of 256 threads (4 wave64) walks a run of
// tiles of TH x TW data bricks (16*TH x 16*TW points).  For each tile it writes
// the tile plus a 2-point halo into LDS from registers, issues the global loads
// of its next tile (16-byte loads for the data bricks, 4-byte loads for the
// halo frame) so they are in flight meanwhile, then
//   MFMA:   each wave computes whole 16x16 bricks.  For a brick, with X the
//           20x24 haloed input window (X[0][0] = brick origin - (2,4)),
//               out^T = sum_z  C_z * S_z,   C_z[j][k] = c[z][k-j-2] (0 <= k-j-2 <= 4),
//                                           S_z[k][i] = X[i+z][k]
//           is one v_mfma_f32_16x16x32_{f16,bf16} per row shift z: M = output
//           column j, N = output row i, K = window column k (slots 2..21 carry
//           taps, the rest multiply zero coefficients).  The banded Toeplitz
//           C_z is the reference's init_coe matrix (layout16/host_box2d2r.cu:29-50)
//           transposed and widened to cover the halo columns, so no scalar
//           halo fix-ups are needed.  C_z lives in registers for the whole
//           kernel; S_z is one ds_read_b128 per lane, and a brick's reads
//           are issued while the previous brick's MFMAs run.  The FP32 accumulator of
//           lane l holds output row l%16, columns 4*(l/16)..+3, which is one
//           8-byte store into the row-major brick.
//   DIRECT: each thread computes a 4-column x MR-row strip with FP32 FMAs on
//           FP32 coefficients, sliding down the rows held in registers.
// Points outside the interior [rs, rows-rs) x [rs, cols-rs) keep their input
// bits (the Dirichlet ring, and zero padding beyond the grid).
#include "tcs_internal.hpp"

namespace tcs {

// The MFMA kernel stores each 16x16 output brick whole (512 contiguous bytes
// per wave instruction) and never reads it again in the same sweep, so its
// stores are nontemporal.  The direct kernel writes 32-byte pieces of a line
// from different instructions; nontemporal stores measured 2x slower there.
template <bool NT>
__device__ __forceinline__ void store8(uint16_t *p, u32x2 w) {
    if constexpr (NT)
        __builtin_nontemporal_store(w, reinterpret_cast<u32x2 *>(p));
    else
        *reinterpret_cast<u32x2 *>(p) = w;
}

__device__ __forceinline__ uint16_t f16_bits(float v) {
    return __builtin_bit_cast(uint16_t, static_cast<_Float16>(v));
}
__device__ __forceinline__ uint32_t pack2(uint16_t lo, uint16_t hi) {
    return (uint32_t)lo | ((uint32_t)hi << 16);
}
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
// Two FP32 values rounded to nearest-even FP16 / BF16 in one instruction
// (v_cvt_pk_f16_f32 / v_cvt_pk_bf16_f32), as bits: the first value in the low half.
__device__ __forceinline__ uint32_t f16x2_bits(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, f16x2));
}
__device__ __forceinline__ uint32_t bf16x2_bits(float a, float b) {
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{a, b}, bf16x2));
}

// ------------------------------------------------------------ tile staging
// LDS image of BH x BW data bricks plus a 2-point frame on every side: element
// (r, c), r < LR, c < LC.  Image row r holds global row r0 - 2 + r; image
// column c holds global column c0 - FX + c, so the frame columns are FX - 2,
// FX - 1 and 16*BW + FX, +1, and the other columns hold zeros.  Two layouts:
//   row-major (32x32 MFMA and direct kernels, FX = 2): pitch LC elements, so a
//     row is 16 (mod 32) bytes and rows step through the banks;
//   STRIP (16x16 MFMA kernels, FX = 4): the image is cut into 8-column strips
//     stored one after another, NR = 16*BH + 8 rows of 16 bytes each.  A 16x16
//     MFMA operand fragment (16 lanes = 16 consecutive rows of one 8-column
//     group) then reads 256 contiguous bytes, and since NR = 8 (mod 16), strips
//     two apart start at the same bank: lane groups (0,1) and (2,3) of a
//     ds_read_b128 reading the same rows of strips s and s+2 (or the same
//     address) complete in the minimum 4 LDS cycles.  Measured on gfx950:
//     ds_read_b128 of the 16x16 operand pattern takes 4 cycles per wave in
//     this layout against 6 (unpacked) to 8 (packed) in the row-major one.
//     With the data 4 columns in, a lane's 4 results (brick columns 4g..4g+3)
//     are 8 aligned bytes of one strip row: the fused sweeps write them back
//     with one ds_write_b64 that touches every bank exactly twice (16 lanes
//     of consecutive rows per group, groups 0, 2 on the upper and 1, 3 on the
//     lower half of the 16-byte rows), where two ds_write_b32 of the 2-column
//     offset hit half the banks twice each.
template <int BH_, int BW_, int NPL_, bool STRIP_ = false, int NT_ = kWG>
struct Geom {
    static constexpr int BH = BH_, BW = BW_, NPL = NPL_;
    static constexpr bool STRIP = STRIP_;
    static constexpr int NT = NT_;              // threads of the workgroup that stages the image
    static constexpr int FX = STRIP ? 4 : 2;    // image column of the first data column
    static constexpr int LR = 16 * BH + 4;      // image rows (2 frame rows each side)
    static constexpr int LC = 16 * BW + 8;      // image columns (row-major pitch)
    static constexpr int NR = 16 * BH + 8;      // rows per strip (STRIP)
    static constexpr int NCH = BH * BW * 32;    // 16-byte chunks of the data bricks
    static constexpr int CPT = (NCH + NT - 1) / NT;
    static constexpr int HD = 8 * BW + 2;       // dwords in one top/bottom frame row
    static constexpr int NH = 4 * HD + 32 * BH; // frame dwords: 2 top + 2 bottom rows, left, right
    static constexpr int HPT = (NH + NT - 1) / NT;
    static constexpr int PLANE = STRIP ? LC * NR : LR * LC;  // LDS elements per plane
    // Element offset of image point (r, c).  Offsets add: at(r1 + r2, c1 + c2)
    // = at(r1, c1) + at(r2, c2) whenever c1 is a multiple of 8.
    static __device__ __forceinline__ constexpr int at(int r, int c) {
        if constexpr (STRIP)
            return (c >> 3) * (8 * NR) + 8 * r + (c & 7);
        else
            return r * LC + c;
    }
};

template <int TH, int TW, int NPL, bool STRIP = false>
using Tile = Geom<TH, TW, NPL, STRIP>;

// One tile's global loads, held in registers between issue and the LDS write
// so that the next tile's loads can be in flight while the current one computes.
template <class G>
struct TileRegs {
    u32x4 cv[G::NPL][G::CPT];
    uint32_t hv[G::NPL][G::HPT];
};

// Frame dword h -> (LDS row, LDS dword column): two rows above, two rows below
// (HD dwords each), then the left and right 2-column strips.
template <class G>
__device__ __forceinline__ void halo_coords(int h, int &lr, int &ld) {
    constexpr int HD = G::HD;
    if (h < 2 * HD) {
        lr = h / HD;
        ld = h - lr * HD;
    } else if (h < 4 * HD) {
        const int q = h - 2 * HD;
        const int rr = q / HD;
        lr = 16 * G::BH + 2 + rr;
        ld = q - rr * HD;
    } else if (h < 4 * HD + 16 * G::BH) {
        lr = 2 + (h - 4 * HD);
        ld = 0;
    } else {
        lr = 2 + (h - 4 * HD - 16 * G::BH);
        ld = HD - 1;
    }
}

// LDS element (lr, lc) of plane p holds global point (r0 - 2 + lr, c0 - FX + lc)
// for lr < LR, FX - 2 <= lc < 16*BW + FX + 2; the other columns hold zeros.
// (r0, c0) is the first element of the first data brick.  With CHECK, frame points outside
// the stored bricks (possible when the data bricks include the ghost ring) read
// as zero.  (Nontemporal data-brick loads measured slower.)
template <class G, bool CHECK>
__device__ __forceinline__ void load_tile(const SweepArgs &a, TileRegs<G> &R, int r0, int c0, int tid) {
#pragma unroll
    for (int i = 0; i < G::CPT; ++i) {
        const int c = tid + i * G::NT;
        if (G::NCH % G::NT == 0 || c < G::NCH) {
            const int b = c >> 5, w = c & 31;
            const int br = b / G::BW, bc = b - br * G::BW;
            const int64_t s = stor_index(r0 + 16 * br + (w >> 1), c0 + 16 * bc + 8 * (w & 1), a.sbc);
#pragma unroll
            for (int p = 0; p < G::NPL; ++p)
                R.cv[p][i] = *reinterpret_cast<const u32x4 *>(a.in + p * a.plane + s);
        }
    }
#pragma unroll
    for (int i = 0; i < G::HPT; ++i) {
        const int h = tid + i * G::NT;
        if (h < G::NH) {
            int lr, ld;
            halo_coords<G>(h, lr, ld);
            const int gr = r0 - 2 + lr, gc = c0 - 2 + 2 * ld;
            const bool inside = !CHECK || ((unsigned)((gr + 16) >> 4) < (unsigned)a.sbr &&
                                           (unsigned)((gc + 16) >> 4) < (unsigned)a.sbc);
            const int64_t s = stor_index(gr, gc, a.sbc);
#pragma unroll
            for (int p = 0; p < G::NPL; ++p)
                R.hv[p][i] = inside ? *reinterpret_cast<const uint32_t *>(a.in + p * a.plane + s) : 0u;
        }
    }
}

// Storage-format bits of a row-major FP32 value (times rm_scale_in), rounded
// to nearest as the pack kernel rounds: FP16 or BF16, or for SPLIT the FP16
// pair hi = fp16(x), lo = fp16(x - hi) (plane 0 and plane 1).
template <int PREC>
__device__ __forceinline__ uint16_t rm_encode(const SweepArgs &a, float x, int plane = 0) {
    x *= a.rm_scale_in;
    if constexpr (base_prec(PREC) == TCS_PREC_BF16) {
        return __builtin_bit_cast(uint16_t, static_cast<__bf16>(x));
    } else {
        const _Float16 h = static_cast<_Float16>(x);
        if (plane == 0) return __builtin_bit_cast(uint16_t, h);
        return __builtin_bit_cast(uint16_t, static_cast<_Float16>(x - static_cast<float>(h)));
    }
}

// load_tile from the row-major FP32 grid a.rm_src instead of bricks (FP16/BF16,
// one plane): what the pack kernel would have stored -- the rounded value
// inside the grid, zero for padding and for frame points outside it.  Data
// chunk c (8 elements of a row) is taken in row-major order over the tile,
// chunk cc of tile row lrow, so consecutive threads read consecutive 32 bytes
// of one row (two 16-byte loads when a.rm_vec allows and the chunk lies inside
// the grid); write_tile<G, true> stores them with the same order.
template <class G, int PREC>
__device__ __forceinline__ void load_tile_rm(const SweepArgs &a, TileRegs<G> &R, int r0, int c0, int tid) {
    static_assert(G::NPL == 1 || base_prec(PREC) == TCS_PREC_SPLIT, "one storage plane, or the SPLIT pair");
#pragma unroll
    for (int i = 0; i < G::CPT; ++i) {
        const int c = tid + i * G::NT;
        if (G::NCH % G::NT == 0 || c < G::NCH) {
            const int lrow = c / (2 * G::BW), cc = c - lrow * (2 * G::BW);
            const int gr = r0 + lrow, gc = c0 + 8 * cc;  // both >= 0
            float x[8];
            if (gr < a.rows && a.rm_vec && gc + 8 <= a.cols) {
                const float *src = a.rm_src + gr * a.rm_ld + gc;
                const f32x4 u = *reinterpret_cast<const f32x4 *>(src);
                const f32x4 v = *reinterpret_cast<const f32x4 *>(src + 4);
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    x[e] = u[e];
                    x[4 + e] = v[e];
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e)
                    x[e] = gr < a.rows && gc + e < a.cols ? a.rm_src[gr * a.rm_ld + gc + e] : 0.0f;
            }
#pragma unroll
            for (int p = 0; p < G::NPL; ++p)
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    R.cv[p][i][q] = pack2(rm_encode<PREC>(a, x[2 * q], p), rm_encode<PREC>(a, x[2 * q + 1], p));
        }
    }
#pragma unroll
    for (int i = 0; i < G::HPT; ++i) {
        const int h = tid + i * G::NT;
        if (h < G::NH) {
            int lr, ld;
            halo_coords<G>(h, lr, ld);
            const int gr = r0 - 2 + lr, gc = c0 - 2 + 2 * ld;  // gc even; either may be negative
            const bool row_in = (unsigned)gr < (unsigned)a.rows;
            float x[2];
#pragma unroll
            for (int e = 0; e < 2; ++e)
                x[e] = row_in && (unsigned)(gc + e) < (unsigned)a.cols ? a.rm_src[gr * a.rm_ld + gc + e] : 0.0f;
#pragma unroll
            for (int p = 0; p < G::NPL; ++p) R.hv[p][i] = pack2(rm_encode<PREC>(a, x[0], p), rm_encode<PREC>(a, x[1], p));
        }
    }
}

// ROWMAJOR: the data chunks are in load_tile_rm's row-major order.
template <class G, bool ROWMAJOR = false>
__device__ __forceinline__ void write_tile(const TileRegs<G> &R, uint16_t *lds, int tid) {
    uint32_t *l32 = reinterpret_cast<uint32_t *>(lds);
#pragma unroll
    for (int i = 0; i < G::CPT; ++i) {
        const int c = tid + i * G::NT;
        if (G::NCH % G::NT == 0 || c < G::NCH) {
            int lr, lc;  // lc = FX (mod 8)
            if constexpr (ROWMAJOR) {
                const int lrow = c / (2 * G::BW);
                lr = lrow + 2;
                lc = 8 * (c - lrow * (2 * G::BW)) + G::FX;
            } else {
                const int b = c >> 5, w = c & 31;
                const int br = b / G::BW, bc = b - br * G::BW;
                lr = 16 * br + (w >> 1) + 2;
                lc = 16 * bc + 8 * (w & 1) + G::FX;
            }
#pragma unroll
            for (int p = 0; p < G::NPL; ++p) {
                if constexpr (G::FX == 4) {
                    // the chunk's halves: elements 4..7 of one strip row and
                    // 0..3 of the next strip's, 8 aligned bytes each
                    u32x2 *l64 = reinterpret_cast<u32x2 *>(lds + p * G::PLANE);
                    l64[G::at(lr, lc) >> 2] = u32x2{R.cv[p][i][0], R.cv[p][i][1]};
                    l64[G::at(lr, lc + 4) >> 2] = u32x2{R.cv[p][i][2], R.cv[p][i][3]};
                } else {
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        l32[p * (G::PLANE / 2) + (G::at(lr, lc + 2 * q) >> 1)] = R.cv[p][i][q];
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < G::HPT; ++i) {
        const int h = tid + i * G::NT;
        if (h < G::NH) {
            int lr, ld;
            halo_coords<G>(h, lr, ld);
            const int d = G::at(lr, G::FX - 2 + 2 * ld) >> 1;
#pragma unroll
            for (int p = 0; p < G::NPL; ++p) l32[p * (G::PLANE / 2) + d] = R.hv[p][i];
        }
    }
}

// The image columns that hold no grid point: [0, FX-2) and [16*BW+FX+2, LC),
// one dword of each side per row for STRIP (FX = 4), columns [16*BW+4,
// 16*BW+8) for the row-major image (FX = 2).  The MFMA operands of the first
// and last brick column read them against zero coefficients, so they must be
// finite; write_tile never writes them.
template <class G>
__device__ __forceinline__ void zero_pad(uint16_t *lds, int tid) {
    static_assert(G::FX == 2 || G::FX == 4, "two pad dwords per image row");
    uint32_t *l32 = reinterpret_cast<uint32_t *>(lds);
    for (int zc = tid; zc < 2 * G::LR; zc += G::NT) {
        const int col = G::FX == 4 ? ((zc & 1) ? 16 * G::BW + 6 : 0) : 16 * G::BW + 4 + 2 * (zc & 1);
        const int d = G::at(zc >> 1, col) >> 1;
#pragma unroll
        for (int p = 0; p < G::NPL; ++p) l32[p * (G::PLANE / 2) + d] = 0u;
    }
}

__device__ __forceinline__ bool tile_is_interior(const SweepArgs &a, int gr0, int gc0, int h, int w) {
    return gr0 >= a.rs && gr0 + h <= a.rows - a.rs && gc0 >= a.rs && gc0 + w <= a.cols - a.rs;
}

__device__ __forceinline__ bool point_is_interior(const SweepArgs &a, int gr, int gc) {
    return gr >= a.rs && gr < a.rows - a.rs && gc >= a.rs && gc < a.cols - a.rs;
}

// Tile schedule.  Tile rows are padded to tiles_xp virtual columns (a multiple
// of 8); virtual tiles with tx >= tiles_x are empty.  Workgroup w takes virtual
// tiles v = w, w + gridDim.x, ... (gridDim.x a multiple of 8).  Workgroups are
// dispatched to the 8 XCDs round robin, so every tile of tile column tx runs on
// XCD tx % 8 and vertically adjacent tiles meet in that XCD's L2 for their halo
// rows, while the tiles in flight at any moment form one contiguous band of
// tile rows, which spreads the traffic over every HBM channel.  Measured on
// box2d2r 16384^2 (8x8-brick tiles): this order beats eight XCD-private
// contiguous tile ranges (the ranges keep horizontal neighbours on one XCD,
// but run 13% slower as a plain tile copy) and runs of up to 16 tiles per XCD.
//
// Row-band order (BAND, the kernels that read row-major FP32 input): tile rows
// are grouped in bands of 8 (padded with empty tiles), and within a band
// virtual tile v takes tile column (v % (8 tiles_x)) / 8 of tile row
// 8 band + v % 8.  Tile row ty then runs on XCD ty % 8, so horizontally
// adjacent tiles share an L2.  There the horizontal halo is the costly one: a
// 2-column strip of an FP32 row costs a 128-byte line per row per side, on a
// 512-byte tile row (4x8 bricks: 1.57x the algorithmic input fetched in tile
// column order), against two extra rows of 64 for the vertical halo.
struct Schedule {
    int tile, end, step;
};

template <bool BAND>
__device__ __forceinline__ bool tile_is_empty(const SweepArgs &a, int v) {
    if constexpr (BAND)
        return 8 * (v / (8 * a.tiles_x)) + (v & 7) >= a.tiles_y;
    else
        return v % a.tiles_xp >= a.tiles_x;
}

// First virtual tile >= v (in steps of `step`) that is not empty.
template <bool BAND = false>
__device__ __forceinline__ int skip_empty(const SweepArgs &a, int v, int step) {
    while (v < a.nvirt && tile_is_empty<BAND>(a, v)) v += step;
    return v;
}

template <bool BAND = false>
__device__ __forceinline__ Schedule schedule(const SweepArgs &a) {
    const int step = (int)gridDim.x;
    return Schedule{skip_empty<BAND>(a, (int)blockIdx.x, step), a.nvirt, step};
}

template <int TH, int TW, bool BAND = false>
__device__ __forceinline__ void tile_origin(const SweepArgs &a, int v, int &gr0, int &gc0) {
    int ty, tx;
    if constexpr (BAND) {
        const int band = v / (8 * a.tiles_x), rem = v - band * 8 * a.tiles_x;
        ty = 8 * band + (rem & 7);
        tx = rem >> 3;
    } else {
        ty = v / a.tiles_xp;
        tx = v - ty * a.tiles_xp;
    }
    gr0 = 16 * TH * ty;
    gc0 = 16 * TW * tx;
}


// ------------------------------------------------------------- MFMA kernel
template <int PREC>
struct Frag;
template <>
struct Frag<TCS_PREC_FP16> {
    typedef half8 type;
    static __device__ __forceinline__ f32x4 mma(type x, type y, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(x, y, c, 0, 0, 0);
    }
};
template <>
struct Frag<TCS_PREC_BF16> {
    typedef bf16x8 type;
    static __device__ __forceinline__ f32x4 mma(type x, type y, f32x4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(x, y, c, 0, 0, 0);
    }
};
template <>
struct Frag<TCS_PREC_SPLIT> : Frag<TCS_PREC_FP16> {};
template <>
struct Frag<TCS_PREC_FP16 | kCoefLo> : Frag<TCS_PREC_FP16> {};
template <>
struct Frag<TCS_PREC_BF16 | kCoefLo> : Frag<TCS_PREC_BF16> {};

// K slots of the 16x16x32 MFMAs of one brick.  K slot group g (k = 8g..8g+7)
// of MFMA f multiplies window row m + ZLO + zi, columns 8kg..8kg+7 (operand
// B, lane (m, g)) by the taps c[ZLO + zi][8kg + e - j - 2] (operand A, lane
// (j, g)).  Fragments 0..2 each cover one 8-column group, kg = 0, 2, 1, with
// lane group g on row shift zi = g; radius 2 adds fragment 3 for the fifth
// shift (zi = 4) with groups kg = 0, 2, 1, 1:
//     f = 0: (0,0) (1,0) (2,0) (3,0)     f = 2: (0,1) (1,1) (2,1) (3,1)
//     f = 1: (0,2) (1,2) (2,2) (3,2)     f = 3: (4,0) (4,2) (4,1) (4,1)*
// (* zero taps).  Radius 1 has fragments 0..2 only, lane group 3 on shift 3
// against zero taps.  4 MFMAs per brick for radius 2 (15 of 16 slot groups
// carry taps), 3 for radius 1.
//   Reuse: window columns 16..23 of a brick are columns 0..7 of the brick to
// its right, so fragment 0 of a brick is fragment 1 of its left neighbour,
// lane for lane: a wave that walks a brick row left to right reads 3
// fragments per brick instead of 4 (radius 1: 2 instead of 3).
//   LDS banks: a fragment 0..2 read is one strip, each lane group 16
// consecutive rows; fragment 3 reads the same rows of strips two apart in
// lane groups 0, 1 (the same banks, which gfx950 serves without conflict)
// and one address twice in groups 2, 3.  Each takes the minimum 4 LDS cycles
// per ds_read_b128 (measured); lane groups 0, 1 on strips one apart take 8.
// Every 16x16 kernel (one sweep, fused, batched) uses this one map, so the
// MFMAs add the same 8-column partial sums into the FP32 accumulator in the
// same order everywhere, and k fused sweeps stay bitwise equal to k single
// sweeps.  (The MFMA's FP32 sum depends on the slot order: the same products
// with two slot groups swapped round differently in about half the outputs.)
template <int ZLO, int ZHI>
struct KMap {
    static constexpr int NZ = ZHI - ZLO + 1;
    static_assert(NZ == 3 || NZ == 5, "radius 1 or 2");
    static constexpr int NF = NZ == 5 ? 4 : 3;  // MFMAs (operand fragments) per brick and plane
    static constexpr int REUSE_FROM = 1;          // fragment 0 = the left neighbour's fragment 1
    static __host__ __device__ constexpr int kg_of(int f) { return f == 0 ? 0 : (f == 1 ? 2 : 1); }
    static __device__ __forceinline__ void group(int f, int g, int &zi, int &kg, bool &real) {
        if (f < 3) {
            zi = g;
            kg = kg_of(f);
            real = g < NZ;
        } else {
            zi = 4;
            kg = kg_of(g < 3 ? g : 2);
            real = g < 3;
        }
    }
};

// Coefficient fragments: lane (j, g) of MFMA f holds c[ZLO + zi][8kg + e - j - 2],
// e = 0..7 (zero outside the taps) -- the window at d0 = 8kg - j - 2 of the
// zero-padded tap row, loaded whole from the host-built window table
// (CoefWindows; window 12 is all zeros).  K slot k = 8kg + e is image column
// 16bc + k of the brick's window, and output column j sits at image column
// 16bc + 4 + j (the STRIP image's data start 4 columns in), so tap d of
// output j is slot j + 2 + d.  lo: the second plane of the table (SPLIT, and
// FP16/BF16 with hi+lo coefficients).
template <int ZLO, int ZHI, int PREC>
struct Coef {
    typedef typename Frag<PREC>::type frag_t;
    using K = KMap<ZLO, ZHI>;
    frag_t hi[K::NF], lo[K::NF];

    __device__ __forceinline__ void load(const CoefWindows &cw, int m, int g) {
#pragma unroll
        for (int f = 0; f < K::NF; ++f) {
            int zi, kg;
            bool real;
            K::group(f, g, zi, kg, real);
            const int d0 = 8 * kg - m - 2;
            const int j = real && d0 >= -7 && d0 <= 4 ? d0 + 7 : 12;
            const uint32_t *wh = cw.w[0][ZLO + zi][j];
            hi[f] = __builtin_bit_cast(frag_t, u32x4{wh[0], wh[1], wh[2], wh[3]});
            if constexpr (base_prec(PREC) == TCS_PREC_SPLIT || has_coef_lo(PREC)) {
                const uint32_t *wl = cw.w[1][ZLO + zi][j];
                lo[f] = __builtin_bit_cast(frag_t, u32x4{wl[0], wl[1], wl[2], wl[3]});
            }
        }
    }
};

// Operands of one 16x16 output brick: brick (br, bc) of the LDS image of
// geometry G, whose 20x24 window starts at LDS row 16*br, column 16*bc.
// Fragment f: lane (m, g) holds haloed row m + ZLO + zi, columns 8kg..8kg+7
// (KMap), one 16-byte LDS read; all reads of a brick issue back to back.
template <class G, int ZLO, int ZHI, int PREC>
struct Ops {
    static_assert(G::STRIP, "the K slot order is chosen for the STRIP LDS layout");
    using K = KMap<ZLO, ZHI>;
    static constexpr int NPL = base_prec(PREC) == TCS_PREC_SPLIT ? 2 : 1;
    u32x4 h[K::NF];
    u32x4 l[NPL == 2 ? K::NF : 1];

    // Per-lane element offsets of the fragments within a brick's window:
    // fragments 1 and 2 are whole strips from fragment 0 (compile-time), so
    // only fragment 0 and, for radius 2, fragment 3 need a register.
    struct Offsets {
        int o0, o3;
        __device__ __forceinline__ Offsets(int m, int g) {
            int zi, kg;
            bool real;
            K::group(0, g, zi, kg, real);
            o0 = G::at(m + ZLO + zi, 8 * kg);
            o3 = o0;
            if constexpr (K::NF == 4) {
                K::group(3, g, zi, kg, real);
                o3 = G::at(m + ZLO + zi, 8 * kg);
            }
        }
        __device__ __forceinline__ int of(int f) const {
            return f == 3 ? o3 : o0 + G::at(0, 8 * K::kg_of(f));
        }
    };

    __device__ __forceinline__ void read(const uint16_t *base, const Offsets &off, int f) {
        h[f] = *reinterpret_cast<const u32x4 *>(base + off.of(f));
        if constexpr (NPL == 2) l[f] = *reinterpret_cast<const u32x4 *>(base + G::PLANE + off.of(f));
    }

    // All fragments of brick (br, bc).
    __device__ __forceinline__ void load(const uint16_t *lds, int br, int bc, const Offsets &off) {
        const uint16_t *base = lds + G::at(16 * br, 16 * bc);
#pragma unroll
        for (int f = 0; f < K::NF; ++f) read(base, off, f);
    }

    // Brick (br, bc) whose left neighbour's operands are `left` (same lane
    // rows): fragment 0 is taken from there, the others are read.  `left`
    // may be this object.
    __device__ __forceinline__ void load_after(const uint16_t *lds, int br, int bc, const Offsets &off,
                                               const Ops &left) {
        const u32x4 h0 = left.h[K::REUSE_FROM];
        const u32x4 l0 = left.l[NPL == 2 ? K::REUSE_FROM : 0];
        const uint16_t *base = lds + G::at(16 * br, 16 * bc);
#pragma unroll
        for (int f = 1; f < K::NF; ++f) read(base, off, f);
        h[0] = h0;
        if constexpr (NPL == 2) l[0] = l0;
    }

    // load_after when `reuse` (a compile-time constant after unrolling), else load.
    __device__ __forceinline__ void load(const uint16_t *lds, int br, int bc, const Offsets &off, const Ops &left,
                                         bool reuse) {
        if (reuse)
            load_after(lds, br, bc, off, left);
        else
            load(lds, br, bc, off);
    }
};

// The brick's MFMAs: lane (m, g) accumulates output row m, columns 4g..4g+3
// in FP32.
template <class G, int ZLO, int ZHI, int PREC>
__device__ __forceinline__ f32x4 mfma_acc(const Coef<ZLO, ZHI, PREC> &C, const Ops<G, ZLO, ZHI, PREC> &op) {
    constexpr int NF = KMap<ZLO, ZHI>::NF;
    using F = Frag<PREC>;
    typedef typename F::type frag_t;
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const frag_t bh = __builtin_bit_cast(frag_t, op.h[f]);
        acc = F::mma(C.hi[f], bh, acc);
        // Hi+lo coefficients: the lo taps (c - round(c), rounded) on the same
        // data, so the accumulator sees each tap to 2^-22 (FP16) or 2^-16
        // (BF16) of its value instead of 2^-11 or 2^-8: no drift from the
        // rounded taps' sum over many sweeps.
        if constexpr (has_coef_lo(PREC)) acc = F::mma(C.lo[f], bh, acc);
        if constexpr (PREC == TCS_PREC_SPLIT) {
            // c*x = (ch + cl)(xh + xl) without the cl*xl term: |cl| <= 2^-11 |c|
            // and |xl| <= 2^-11 |x| (round to nearest), so the dropped part of
            // an output is at most 2^-22 * sum |c||x| -- at most 2.4e-7 of the
            // max-abs of the (scaled) input for taps summing to 1, below the
            // SPLIT target of 1e-6, for a quarter of the MFMAs.
            const frag_t bl = __builtin_bit_cast(frag_t, op.l[f]);
            acc = F::mma(C.hi[f], bl, acc);
            acc = F::mma(C.lo[f], bh, acc);
        }
    }
    return acc;
}

// The brick's epilogue: the accumulator of lane (m, g) (output row m, columns
// 4g..4g+3 of brick (br, bc)) in storage format (w[p]: four 16-bit values of
// plane p).  gr, gcb: global row and first column of those outputs; with
// `edge`, points outside the interior keep their input value (read from the
// LDS image).
template <class G, int PREC>
__device__ __forceinline__ void brick_out(const SweepArgs &a, const uint16_t *lds, const f32x4 &acc, int br, int bc,
                                          int m, int g, bool edge, int gr, int gcb, u32x2 (&w)[G::NPL]) {
    constexpr int NPL = G::NPL;
    // Conversions two values at a time (one packed conversion per pair: box2d2r
    // 16384^2, 8 fused sweeps per launch, SPLIT 276 -> 249 us and FP16 107.5
    // -> 101.5 us per sweep on one box; the rounding is the same).
    uint32_t wv[NPL][2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const float x0 = acc[2 * h], x1 = acc[2 * h + 1];
        if constexpr (base_prec(PREC) == TCS_PREC_SPLIT) {
            const f32x2 sv = {x0 * a.out_scale, x1 * a.out_scale};
            const f16x2 hi = __builtin_convertvector(sv, f16x2);
            const f32x2 hf = __builtin_convertvector(hi, f32x2);
            wv[0][h] = __builtin_bit_cast(uint32_t, hi);
            wv[1][h] = f16x2_bits(sv[0] - hf[0], sv[1] - hf[1]);
        } else if constexpr (base_prec(PREC) == TCS_PREC_BF16) {
            wv[0][h] = bf16x2_bits(x0, x1);
        } else {
            wv[0][h] = f16x2_bits(x0, x1);
        }
    }
    // Edge tiles: the input at the 4 output points (columns 4g+4..4g+7 of the
    // image row, 8 aligned bytes of one strip row, read whatever the points
    // are) replaces the results outside the interior: no LDS read waits inside
    // a branch.  SPLIT selects whole words under a mask of the kept halves
    // (8 fused sweeps, box2d2r 16384^2: 252 us per sweep against 268 with
    // 16-bit values), FP16/BF16 16-bit values (91.8 us against 94.3 with words).
    static_assert(G::FX == 4, "a lane's 4 points are 8 aligned bytes of the STRIP image");
    const int ck = G::at(16 * br, 16 * bc) + G::at(m + 2, 4 * g + G::FX);
    if constexpr (NPL == 2) {
#pragma unroll
        for (int p = 0; p < NPL; ++p) w[p] = u32x2{wv[p][0], wv[p][1]};
        if (edge) {
            const bool row_in = gr >= a.rs && gr < a.rows - a.rs;
            uint32_t keep[2] = {0u, 0u};
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gc = gcb + v;
                const bool interior = row_in && gc >= a.rs && gc < a.cols - a.rs;
                keep[v >> 1] |= interior ? 0u : 0xffffu << (16 * (v & 1));
            }
#pragma unroll
            for (int p = 0; p < NPL; ++p) {
                const u32x2 k = *reinterpret_cast<const u32x2 *>(lds + p * G::PLANE + ck);
#pragma unroll
                for (int h = 0; h < 2; ++h) w[p][h] = (k[h] & keep[h]) | (w[p][h] & ~keep[h]);
            }
        }
    } else {
        uint16_t o[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) o[v] = (uint16_t)(wv[0][v >> 1] >> (16 * (v & 1)));
        if (edge) {
            const bool row_in = gr >= a.rs && gr < a.rows - a.rs;
            const u32x2 k = *reinterpret_cast<const u32x2 *>(lds + ck);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gc = gcb + v;
                const bool interior = row_in && gc >= a.rs && gc < a.cols - a.rs;
                o[v] = interior ? o[v] : (uint16_t)(k[v >> 1] >> (16 * (v & 1)));
            }
        }
        w[0] = u32x2{pack2(o[0], o[1]), pack2(o[2], o[3])};
    }
}

// MFMAs and epilogue of one brick.
template <class G, int ZLO, int ZHI, int PREC>
__device__ __forceinline__ void mfma_brick(const SweepArgs &a, const uint16_t *lds,
                                           const Coef<ZLO, ZHI, PREC> &C,
                                           const Ops<G, ZLO, ZHI, PREC> &op, int br, int bc, int m, int g,
                                           bool edge, int gr, int gcb, u32x2 (&w)[G::NPL]) {
    brick_out<G, PREC>(a, lds, mfma_acc<G, ZLO, ZHI, PREC>(C, op), br, bc, m, g, edge, gr, gcb, w);
}

// Brick q of a wave's share of a BH x BW brick region whose first brick is
// (BASE, BASE), split over the NW waves of a workgroup, q row-major within the
// wave's block: with BH a multiple of NW wave w takes whole rows (BH/NW of
// them, from row w * BH/NW), else the waves split the region into NW/2 row
// bands by 2 column halves, wave w taking the (2BH/NW) x (BW/2) block at
// (w / 2, w % 2) (4 waves: the 2 x 2 split).  A brick's offset from its
// block's first brick is a compile-time constant, so a per-brick LDS address
// is one register plus an immediate, and a wave walks each row of its block
// left to right (BlockRows::RW bricks per row: every brick but a row's first
// reuses its left neighbour's fragment).
template <int BH, int BW, int NW = 4>
struct BlockRows {
    static constexpr bool WHOLE = BH % NW == 0;
    static_assert(WHOLE || (BH % (NW / 2) == 0 && BW % 2 == 0), "brick region must split over the waves");
    static constexpr int RW = WHOLE ? BW : BW / 2;  // bricks per row of a wave's block
    static constexpr int Q = BH * BW / NW;          // bricks per wave
};

template <int BH, int BW, int BASE, int NW = 4>
__device__ __forceinline__ void block_brick(int wave, int q, int &br, int &bc) {
    using B = BlockRows<BH, BW, NW>;
    if constexpr (B::WHOLE) {
        br = BASE + wave * (BH / NW) + q / BW;
        bc = BASE + q % BW;
    } else {
        br = BASE + (wave >> 1) * (2 * BH / NW) + q / (BW / 2);
        bc = BASE + (wave & 1) * (BW / 2) + q % (BW / 2);
    }
}

// Operand look-ahead across bricks costs 4*NZ registers (8*NZ for SPLIT); it
// pays for FP16/BF16 and measured slower for SPLIT, whose three MFMA passes per
// shift already hide the LDS latency.
template <int PREC>
constexpr bool kLookahead = base_prec(PREC) != TCS_PREC_SPLIT;

// Single-sweep FP16/BF16 kernels on tiles below 8x8 bricks (LDS for up to 8
// workgroups per CU) trade the look-ahead for 8 waves per SIMD (at most 64
// VGPRs): other workgroups then hide the LDS latency, and the extra resident
// tiles keep more loads in flight (box2d2r 16384^2, 4x8 tile: 193 -> 180 us).
// 8x8 tiles are held to 4 workgroups per CU by LDS and keep the look-ahead;
// persistent grids hold the next tile in registers and keep it too, and so do
// kernels with hi+lo coefficients (their lo fragments do not fit in 64 VGPRs).
template <int TH, int TW, int PREC, bool PERSIST>
constexpr bool kHighOccupancy = base_prec(PREC) == PREC && base_prec(PREC) != TCS_PREC_SPLIT && TH * TW < 64 &&
                                 !PERSIST;

// Row-major FP32 output of 4 results of row gr, columns gc..gc+3 (FP16/BF16
// bits u; SPLIT: the hi bits u and lo bits ul): what the unpack kernel writes -- the stored value times
// rm_scale_out at interior points, the caller's rm_src value on the ring
// (looked up only in edge tiles); nothing outside the grid.  One nontemporal
// 16-byte store when a.rm_vec allows and the piece lies inside the grid (the
// wave's stores together cover whole rows of the tile).
template <int PREC>
__device__ __forceinline__ void store_rm4(const SweepArgs &a, int gr, int gc, u32x2 u, u32x2 ul, bool edge) {
    if (gr >= a.rows) return;
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const uint16_t bits = (uint16_t)(u[e >> 1] >> (16 * (e & 1)));
        float x = base_prec(PREC) == TCS_PREC_BF16 ? static_cast<float>(__builtin_bit_cast(__bf16, bits))
                                        : static_cast<float>(__builtin_bit_cast(_Float16, bits));
        if constexpr (base_prec(PREC) == TCS_PREC_SPLIT)  // hi + lo, as the unpack kernel decodes the pair
            x += static_cast<float>(__builtin_bit_cast(_Float16, (uint16_t)(ul[e >> 1] >> (16 * (e & 1)))));
        v[e] = x * a.rm_scale_out;
    }
    const int64_t row = gr * a.rm_ld;
    if (edge) {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (gc + e < a.cols && !point_is_interior(a, gr, gc + e)) v[e] = a.rm_src[row + gc + e];
    }
    if (a.rm_vec && gc + 4 <= a.cols) {
        __builtin_nontemporal_store(f32x4{v[0], v[1], v[2], v[3]}, reinterpret_cast<f32x4 *>(a.rm_dst + row + gc));
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
            if (gc + e < a.cols) a.rm_dst[row + gc + e] = v[e];
    }
}

// One sweep of TH x TW-brick tiles (tcs_mfma_sweep), or with MODE bit 1 the
// input read from the row-major FP32 grid a.rm_src and with MODE bit 2 the
// output written to the row-major FP32 grid a.rm_dst (tcs_mfma_edge).
template <int TH, int TW, int ZLO, int ZHI, int PREC, bool PERSIST, int MODE>
__device__ __forceinline__ void sweep_tile16(const SweepArgs &a) {
    constexpr int NPL = base_prec(PREC) == TCS_PREC_SPLIT ? 2 : 1;
    constexpr bool LOOKAHEAD = kLookahead<PREC> && !kHighOccupancy<TH, TW, PREC, PERSIST>;
    static_assert(MODE == 0 || !PERSIST, "row-major edges: one tile per workgroup");
    using T = Tile<TH, TW, NPL, true>;
    __shared__ __attribute__((aligned(16))) uint16_t lds[NPL * T::PLANE];

    constexpr bool BAND = (MODE & 1) != 0;
    Schedule sc = schedule<BAND>(a);
    if (sc.tile >= sc.end) return;
    const int tid = threadIdx.x;
    int gr0, gc0;
    tile_origin<TH, TW, BAND>(a, sc.tile, gr0, gc0);
    TileRegs<T> R;
    if constexpr (MODE & 1)
        load_tile_rm<T, PREC>(a, R, gr0, gc0, tid);
    else
        load_tile<T, false>(a, R, gr0, gc0, tid);
    zero_pad<T>(lds, tid);

    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, g = lane >> 4;
    Coef<ZLO, ZHI, PREC> C;
    C.load(a.cw, m, g);
    const typename Ops<T, ZLO, ZHI, PREC>::Offsets off(m, g);
    constexpr int NB = TH * TW / 4;
    u32x2 res[MODE & 2 ? NB : 1][NPL];  // results held for the row-major output

    for (;;) {
        write_tile<T, (MODE & 1) != 0>(R, lds, tid);
        __syncthreads();
        const int cgr0 = gr0, cgc0 = gc0;  // the tile in LDS
        const int next = skip_empty<BAND>(a, sc.tile + sc.step, sc.step);
        if (PERSIST && next < sc.end) {
            tile_origin<TH, TW, BAND>(a, next, gr0, gc0);
            load_tile<T, false>(a, R, gr0, gc0, tid);  // in flight during the compute below
        }
        const bool edge = !tile_is_interior(a, cgr0, cgc0, 16 * TH, 16 * TW);
        // With LOOKAHEAD, operand reads run one brick ahead of the MFMAs.
        Ops<T, ZLO, ZHI, PREC> cur, nxt;
        // Each wave walks its block of the tile row by row (block_brick), so a
        // brick takes its first operand fragment from its left neighbour.
        auto reuse = [](int q) { return q % BlockRows<TH, TW>::RW != 0; };
        if constexpr (LOOKAHEAD) {
            int br, bc;
            block_brick<TH, TW, 0>(wave, 0, br, bc);
            cur.load(lds, br, bc, off);
        }
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            int br, bc;
            block_brick<TH, TW, 0>(wave, q, br, bc);
            if constexpr (LOOKAHEAD) {
                if (q + 1 < NB) {
                    int nbr, nbc;
                    block_brick<TH, TW, 0>(wave, q + 1, nbr, nbc);
                    nxt.load(lds, nbr, nbc, off, cur, reuse(q + 1));
                }
            } else {
                cur.load(lds, br, bc, off, cur, reuse(q));
            }
            __builtin_amdgcn_sched_barrier(0);
            const int gr = cgr0 + 16 * br + m, gcb = cgc0 + 16 * bc + 4 * g;
            u32x2 w[NPL];
            mfma_brick<T, ZLO, ZHI, PREC>(a, lds, C, cur, br, bc, m, g, edge, gr, gcb, w);
            if constexpr (MODE & 2) {
#pragma unroll
                for (int p = 0; p < NPL; ++p) res[q][p] = w[p];
            } else {
                const int64_t s = stor_index(gr, gcb, a.sbc);
#pragma unroll
                for (int p = 0; p < NPL; ++p) store8<true>(a.out + p * a.plane + s, w[p]);
            }
            if constexpr (LOOKAHEAD) cur = nxt;
        }
        if constexpr (MODE & 2) {
            // Row-major output through LDS: the tile's results go into a
            // row-major FP16 image (pitch OP, over the input image, which every
            // wave has finished reading), then every thread writes pieces of 4
            // results of one row, consecutive threads along the row, so each
            // store instruction of a wave writes whole tile rows (16TW * 4
            // contiguous bytes) instead of 16-byte pieces of 16 rows per brick.
            constexpr int OP = 16 * TW + 8;
            static_assert(16 * TH * OP <= T::PLANE, "output image must fit the input image");
            __syncthreads();
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                int br, bc;
                block_brick<TH, TW, 0>(wave, q, br, bc);
#pragma unroll
                for (int p = 0; p < NPL; ++p)  // plane p's image over input plane p
                    *reinterpret_cast<u32x2 *>(lds + p * T::PLANE + (16 * br + m) * OP + 16 * bc + 4 * g) =
                        res[q][p];
            }
            __syncthreads();
            constexpr int NP = 16 * TH * 4 * TW;  // 4-result pieces of the tile
            static_assert(NP % kWG == 0, "whole passes of the workgroup");
#pragma unroll
            for (int i = 0; i < NP / kWG; ++i) {
                const int c = tid + i * kWG;
                const int lrow = c / (4 * TW), pc = c - lrow * (4 * TW);
                const u32x2 u = *reinterpret_cast<const u32x2 *>(lds + lrow * OP + 4 * pc);
                u32x2 ul = {0u, 0u};
                if constexpr (NPL == 2) ul = *reinterpret_cast<const u32x2 *>(lds + T::PLANE + lrow * OP + 4 * pc);
                store_rm4<PREC>(a, cgr0 + lrow, cgc0 + 4 * pc, u, ul, edge);
            }
        }
        // One-tile-per-workgroup launches cover every tile with the grid itself.
        if (!PERSIST || next >= sc.end) break;
        sc.tile = next;
        __syncthreads();  // every wave is done reading LDS before it is overwritten
    }
}

template <int TH, int TW, int ZLO, int ZHI, int PREC, bool PERSIST>
__global__ __launch_bounds__(kWG)
__attribute__((amdgpu_waves_per_eu(kHighOccupancy<TH, TW, PREC, PERSIST> ? 8 : 1, 8)))
void tcs_mfma_sweep(SweepArgs a) {
    sweep_tile16<TH, TW, ZLO, ZHI, PREC, PERSIST, 0>(a);
}

// The first and last sweep of tcs_run_f32 with the layout transforms fused in
// (MODE 1: row-major FP32 in, storage out; 2: storage in, row-major FP32 out;
// 3: both, a whole one-sweep run in one kernel).  Results are bitwise those of
// tcs_pack_f32 / tcs_mfma_sweep / tcs_unpack_f32.  The row-major output holds
// the tile's results in registers until the LDS output image is written, and
// is held to 4 waves per SIMD rather than spill at 8 (box2d2r 16384^2, one
// sweep: 0.50 ms unstaged at 8 waves, 0.86 ms staged with spills, 0.45 ms).
template <int TH, int TW, int ZLO, int ZHI, int PREC, int MODE>
__global__ __launch_bounds__(kWG)
__attribute__((amdgpu_waves_per_eu((MODE & 2) ? (base_prec(PREC) == TCS_PREC_SPLIT ? 3 : 4)
                                                : (kHighOccupancy<TH, TW, PREC, false> ? 8 : 1), 8)))
void tcs_mfma_edge(SweepArgs a) {
    sweep_tile16<TH, TW, ZLO, ZHI, PREC, false, MODE>(a);
}

// Temporal blocking: `nsweeps` sweeps (1 <= nsweeps, radius * (nsweeps - 1) <= 16)
// of one TH x TW tile per workgroup, fused in LDS.  The LDS image holds the tile
// plus a one-brick ring and the 2-point frame; every sweep but the last
// recomputes the ring too (each sweep's valid region shrinks by the radius, and
// the ring is 16 points wide), computing into registers, then writing back in
// place after a barrier.  Values are rounded to the storage format after every
// sweep exactly as between unfused launches, so k fused sweeps are bitwise
// equal to k single sweeps.  Values outside the valid region stay finite
// (stencils of loaded data), which the MFMA form needs: a non-finite operand
// anywhere in a brick row's window reaches all 16 outputs of that row.
template <int TH, int TW, int NPL, bool STRIP = false, int NT = kWG>
using FusedGeom = Geom<TH + 2, TW + 2, NPL, STRIP, NT>;

// One sweep of the fused image but the last: every brick is computed into
// registers, then, after a barrier, written back in place.  Work per wave (NW
// waves): its block of the core band (image brick rows 1..TH, every column,
// the left and right ring bricks included; block_brick), then a run of
// consecutive units of one ring row:
//   !PAIRED  whole bricks of the top (image brick row 0) or bottom (TH+1) row:
//            waves 0..NW/2-1 share the top row, the others the bottom row;
//   PAIRED   (the sweep's results are needed only within 8 points of the core)
//            one product per ring column pair, all NW waves sharing the row of
//            pairs: lanes 0..7 compute rows 8..15 of the top brick, lanes 8..15
//            rows 0..7 of the bottom brick (a lane's operand rows are its own
//            address: lane row m + 8, plus 16 TH rows from lane 8 on); the
//            coefficient operand and the K slots are those of a whole brick,
//            so every computed value has the bits of the unfused sweep.  The
//            other halves keep the previous sweep's values, finite and farther
//            than 8 points from the core.
// A ring row of BW units split over W waves gives the first BW % W waves
// ceil(BW / W) consecutive units and the others floor(BW / W).
template <class G, int TH, int ZLO, int ZHI, int PREC, bool PAIRED, bool LOOKAHEAD, int NW>
__device__ __forceinline__ void fused_sweep(const SweepArgs &a, uint16_t *lds, const Coef<ZLO, ZHI, PREC> &C, int mo,
                                            int g, int wave, bool edge, int er0, int ec0) {
    constexpr int NPL = G::NPL;
    using BR = BlockRows<TH, G::BW, NW>;
    constexpr int QB = BR::Q;                    // core-band bricks per wave
    constexpr int WPR = PAIRED ? NW : NW / 2;    // waves per ring row
    constexpr int RC = G::BW / WPR, RX = G::BW % WPR;
    constexpr int QT = QB + RC + (RX > 0);
    using O = Ops<G, ZLO, ZHI, PREC>;
    const int wi = wave % WPR;                                // position among the ring row's waves
    const int ring_br = PAIRED || wave < WPR ? 0 : TH + 1;   // this wave's ring row
    const int ring_bc = wi * RC + (wi < RX ? wi : RX);       // its first ring unit
    const bool ring_extra = wi < RX;                         // whether it has RC + 1 of them
    const int mp = mo + 8 + (mo >= 8 ? 16 * TH : 0);  // a paired unit's lane row, from image brick row 0
    const typename O::Offsets offs(mo, g);
    // LDS offsets are linear in the row (G::at), so a paired unit's operands
    // are the same lane offsets from a lane-dependent base.
    const uint16_t *ldsp = lds + G::at(mp - mo, 0);
    // Unit q of this wave: its image brick, whether it is a ring pair, and
    // whether it exists (the second wave of a column half has one pair fewer
    // when BW/2 is odd).  Wave-uniform.
    auto unit = [&](int q, int &br, int &bc, bool &pair, bool &valid) {
        pair = false;
        valid = true;
        if (q < QB) {
            block_brick<TH, G::BW, 0, NW>(wave, q, br, bc);  // the core band: image brick rows 1..TH
            br += 1;
        } else {
            const int j = q - QB;
            br = ring_br;
            bc = ring_bc + j;
            pair = PAIRED;
            valid = j < RC || ring_extra;
        }
    };
    // Unit q follows its left neighbour (same lane rows) in this wave's order,
    // so its fragment 0 is the previous unit's fragment 1.
    auto reuse = [&](int q) { return q < QB ? q % BR::RW != 0 : q > QB; };
    u32x2 *l64 = reinterpret_cast<u32x2 *>(lds);
    u32x2 res[QT][NPL];
    O cur, nxt;
    if constexpr (LOOKAHEAD) {
        int br, bc;
        bool pair, valid;
        unit(0, br, bc, pair, valid);
        cur.load(lds, br, bc, offs);
    }
    // A unit's epilogue (conversion to the storage format) runs after the next
    // unit's MFMAs are issued, so the wave does not wait on its own MFMA chain.
    f32x4 pacc = {0.0f, 0.0f, 0.0f, 0.0f};
    int pbr = 0, pbc = 0, prow = 0;
    bool pvalid = false;
#pragma unroll
    for (int q = 0; q < QT; ++q) {
        int br, bc;
        bool pair, valid;
        unit(q, br, bc, pair, valid);
        const int row = pair ? mp : mo;
        f32x4 acc = pacc;
        if (valid) {
            if constexpr (LOOKAHEAD) {
                if (q + 1 < QT) {
                    int nbr, nbc;
                    bool npair, nvalid;
                    unit(q + 1, nbr, nbc, npair, nvalid);
                    if (nvalid) nxt.load(npair ? ldsp : lds, nbr, nbc, offs, cur, reuse(q + 1));
                }
            } else {
                cur.load(pair ? ldsp : lds, br, bc, offs, cur, reuse(q));
            }
            __builtin_amdgcn_sched_barrier(0);
            acc = mfma_acc<G, ZLO, ZHI, PREC>(C, cur);
        }
        if (q > 0 && pvalid)
            brick_out<G, PREC>(a, lds, pacc, pbr, pbc, prow, g, edge, er0 + 16 * pbr + prow, ec0 + 16 * pbc + 4 * g,
                               res[q - 1]);
        pacc = acc;
        pbr = br;
        pbc = bc;
        prow = row;
        pvalid = valid;
        if constexpr (LOOKAHEAD) cur = nxt;
    }
    if (pvalid)
        brick_out<G, PREC>(a, lds, pacc, pbr, pbc, prow, g, edge, er0 + 16 * pbr + prow, ec0 + 16 * pbc + 4 * g,
                           res[QT - 1]);
    __syncthreads();  // every wave has read this sweep's input
#pragma unroll
    for (int q = 0; q < QT; ++q) {
        int br, bc;
        bool pair, valid;
        unit(q, br, bc, pair, valid);
        if (!valid) continue;
        const int row = pair ? mp : mo;
        // the lane's 4 results: 8 aligned bytes of one strip row (one ds_write_b64)
        const int e = G::at(16 * br, 16 * bc) + G::at(row + 2, 4 * g + G::FX);
#pragma unroll
        for (int p = 0; p < NPL; ++p) l64[(p * G::PLANE + e) >> 2] = res[q][p];
    }
    __syncthreads();
}

template <int TH, int TW, int ZLO, int ZHI, int PREC, int NW>
__device__ __forceinline__ void fused_tile(const SweepArgs &a, const CoefWindows &cw, uint16_t *lds, int gr0,
                                           int gc0, int nsweeps) {
    constexpr int NPL = base_prec(PREC) == TCS_PREC_SPLIT ? 2 : 1;
    // Operand look-ahead fits the 4-waves-per-SIMD register budget of the
    // FP16/BF16 kernels (box2d2r 16384^2, 4x8 tile, 8 fused sweeps: 118 ->
    // 115 us per sweep; box2d1r 16 fused: 94 -> 90 us); it measured slightly
    // slower on 4x4 tiles, and SPLIT has no registers left for it.
    constexpr bool LOOKAHEAD = base_prec(PREC) != TCS_PREC_SPLIT && TH * TW >= 32;
    using G = FusedGeom<TH, TW, NPL, true, 64 * NW>;
    constexpr int QC = BlockRows<TH, TW, NW>::Q;  // bricks per wave, core only
    const int tid = threadIdx.x;
    const int er0 = gr0 - 16, ec0 = gc0 - 16;  // global point of image brick (0, 0)
    {
        TileRegs<G> R;
        load_tile<G, true>(a, R, er0, ec0, tid);
        zero_pad<G>(lds, tid);
        write_tile<G>(R, lds, tid);
    }
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, g = lane >> 4;
    Coef<ZLO, ZHI, PREC> C;
    C.load(cw, m, g);
    const typename Ops<G, ZLO, ZHI, PREC>::Offsets off(m, g);
    const bool edge = !tile_is_interior(a, er0, ec0, 16 * G::BH, 16 * G::BW);
    __syncthreads();

    for (int sw = 0; sw + 1 < nsweeps; ++sw) {
        // The lane row, made opaque every sweep: all lane-dependent address
        // parts below derive from it, so the compiler cannot hoist per-brick
        // addresses or edge tests out of the sweep loop (registers per brick
        // would cost occupancy or spill).  With compile-time brick offsets,
        // each address is then one register plus an immediate.
        int mo = m, wv = __builtin_amdgcn_readfirstlane(wave);
        asm volatile("" : "+v"(mo));
        asm volatile("" : "+s"(wv));  // likewise the wave's brick bases (scalar registers)
        // This sweep's results are read by later sweeps only within
        // r * (nsweeps - 1 - sw) points of the core; within 8, the top and
        // bottom ring bricks are computed half each, paired in one product.
        constexpr int R = ZLO == 0 ? 2 : 1;
        if (R * (nsweeps - 1 - sw) <= 8)
            fused_sweep<G, TH, ZLO, ZHI, PREC, true, LOOKAHEAD, NW>(a, lds, C, mo, g, wv, edge, er0, ec0);
        else
            fused_sweep<G, TH, ZLO, ZHI, PREC, false, LOOKAHEAD, NW>(a, lds, C, mo, g, wv, edge, er0, ec0);
    }
    Ops<G, ZLO, ZHI, PREC> cur, nxt;
    if constexpr (LOOKAHEAD) {
        int br, bc;
        block_brick<TH, TW, 1, NW>(wave, 0, br, bc);
        cur.load(lds, br, bc, off);
    }
    // block_brick walks the wave's block row by row: brick q follows its left
    // neighbour unless it starts a row.
    auto reuse = [](int q) { return q % BlockRows<TH, TW, NW>::RW != 0; };
#pragma unroll
    for (int q = 0; q < QC; ++q) {
        int br, bc;
        block_brick<TH, TW, 1, NW>(wave, q, br, bc);
        if constexpr (LOOKAHEAD) {
            if (q + 1 < QC) {
                int nbr, nbc;
                block_brick<TH, TW, 1, NW>(wave, q + 1, nbr, nbc);
                nxt.load(lds, nbr, nbc, off, cur, reuse(q + 1));
            }
        } else {
            cur.load(lds, br, bc, off, cur, reuse(q));
        }
        __builtin_amdgcn_sched_barrier(0);
        const int gr = er0 + 16 * br + m, gcb = ec0 + 16 * bc + 4 * g;
        u32x2 w[NPL];
        mfma_brick<G, ZLO, ZHI, PREC>(a, lds, C, cur, br, bc, m, g, edge, gr, gcb, w);
        const int64_t s = stor_index(gr, gcb, a.sbc);
#pragma unroll
        for (int p = 0; p < NPL; ++p) store8<true>(a.out + p * a.plane + s, w[p]);
        if constexpr (LOOKAHEAD) cur = nxt;
    }
}

// Fused sweeps re-read every brick's operands from LDS each sweep (ring
// included); FP16/BF16 kernels are held to 128 VGPRs for 4 waves per SIMD,
// which LDS allows for tiles up to 4x8 (box2d2r 16384^2, 4x8 tile, 8 sweeps
// fused: 142 -> 134 us per sweep against 2 waves per SIMD).  SPLIT kernels
// (two image planes) are held to as many waves per SIMD as LDS allows
// workgroups per CU times the workgroup's waves per SIMD (NW / 4), at most 3
// (168 VGPRs) for 4-wave and 4 (128 VGPRs) for 8-wave workgroups: 4 waves,
// 2x4, 4x4 and 2x8 tiles 3, 4x8 2, 8x8 1; 8 waves, 4x8 4.  Without the bound a
// 2x8 kernel that grew past 168 VGPRs fell to 2 waves per SIMD and ran 18%
// slower (box2d2r 16384^2, 8 fused sweeps: 391 against 331 us per sweep).
template <int TH, int TW, int PREC, int NW>
constexpr int kFusedMinWaves() {
    if constexpr (base_prec(PREC) != TCS_PREC_SPLIT) {
        return 4;
    } else {
        constexpr int wgs = kLdsBytesPerCU / (2 * 2 * FusedGeom<TH, TW, 2, true>::PLANE);
        constexpr int w = wgs * NW / 4, cap = NW == 4 ? 3 : 4;
        return w < 1 ? 1 : (w > cap ? cap : w);
    }
}

template <int TH, int TW, int ZLO, int ZHI, int PREC, int NW>
__device__ __forceinline__ void fused_kernel(const SweepArgs &a, int nsweeps) {
    constexpr int NPL = base_prec(PREC) == TCS_PREC_SPLIT ? 2 : 1;
    __shared__ __attribute__((aligned(16))) uint16_t lds[NPL * FusedGeom<TH, TW, NPL, true>::PLANE];
    const Schedule sc = schedule(a);
    if (sc.tile >= sc.end) return;
    int gr0, gc0;
    tile_origin<TH, TW>(a, sc.tile, gr0, gc0);
    fused_tile<TH, TW, ZLO, ZHI, PREC, NW>(a, a.cw, lds, gr0, gc0, nsweeps);
}

// One tile per workgroup of 4 waves.
template <int TH, int TW, int ZLO, int ZHI, int PREC>
__global__ __launch_bounds__(kWG)
__attribute__((amdgpu_waves_per_eu(kFusedMinWaves<TH, TW, PREC, 4>(), 8)))
void tcs_mfma_fused(SweepArgs a, int nsweeps) {
    fused_kernel<TH, TW, ZLO, ZHI, PREC, 4>(a, nsweeps);
}

// One tile per workgroup of 8 waves: for images whose LDS size allows few
// workgroups per CU (SPLIT 4x8: 2 workgroups, 4 waves per SIMD instead of 2;
// FP16/BF16 8x8, rounded or hi+lo taps: 2 workgroups, 4 waves per SIMD where
// 4-wave workgroups give 2).
template <int TH, int TW, int ZLO, int ZHI, int PREC>
__global__ __launch_bounds__(2 * kWG)
__attribute__((amdgpu_waves_per_eu(kFusedMinWaves<TH, TW, PREC, 8>(), 8)))
void tcs_mfma_fused8(SweepArgs a, int nsweeps) {
    fused_kernel<TH, TW, ZLO, ZHI, PREC, 8>(a, nsweeps);
}

// ----------------------------------------------- MFMA kernels, 32x32 fragments
// A wave computes a block of 2x2 bricks (32x32 points) with
// v_mfma_f32_32x32x16_{f16,bf16}, the row shift z folded into K.  With X the
// 36x36 haloed input of the block (X[0][0] = block origin - (2,2)),
//     out^T[j][i] = sum_(z,k) C[j][(z,k)] * S[(z,k)][i],
//     C[j][(z,k)] = c[z][k-j] for 0 <= k-j <= 4 (else 0),   S[(z,k)][i] = X[i+z][k].
// The (z, k) slots are packed into groups of 8: per shift four full groups
// (k = 0..7, 8..15, 16..23, 24..31; one ds_read_b128 each) and the half group
// k = 32..35, two shifts sharing one group (two ds_read_b64).  Lane half h
// of MFMA mf holds group 2*mf + h.  Radius 2 (shifts 0..4): 23 groups, 12 MFMAs
// per block; radius 1 (shifts 1..3): 14 groups, 7 MFMAs.  Against the 16x16x32
// form (5 MFMAs and 5 KB of operand reads per brick) this issues 1.2x the MFMA
// flops and reads 3 KB per brick.  Slots without a real (z, k) re-read a half
// row of the block's own window against zero coefficients, so every lane reads
// and every operand is a value of the window.
// Output columns are assigned to the M index so that lane (i, h) = (l%32, l/32)
// ends with row i, columns 8h..8h+7 of the block's left brick in accumulator
// registers 0..7 and of its right brick in 8..15: two 16-byte stores per lane,
// each store instruction writing two whole bricks.
template <int PREC>
struct Frag32;
template <>
struct Frag32<TCS_PREC_FP16> {
    typedef half8 type;
    static __device__ __forceinline__ f32x16 mma(type x, type y, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(x, y, c, 0, 0, 0);
    }
};
template <>
struct Frag32<TCS_PREC_BF16> {
    typedef bf16x8 type;
    static __device__ __forceinline__ f32x16 mma(type x, type y, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(x, y, c, 0, 0, 0);
    }
};

template <int ZLO, int ZHI>
struct Fold32 {
    static constexpr int NZ = ZHI - ZLO + 1;
    static constexpr int NFULL = 4 * NZ;              // full groups
    static constexpr int NG = NFULL + (NZ + 1) / 2;   // + half-group pairs
    static constexpr int NMF = (NG + 1) / 2;          // MFMAs per block
    // Half-group pair of MFMA mf >= 2*NZ for lane half h: shift indices of its
    // two halves (clamped to a real shift when the half is empty).
    static __device__ __forceinline__ int pair(int mf, int h) { return 2 * (mf - 2 * NZ) + h; }
    static __device__ __forceinline__ int half_shift(int p, int s) {
        const int z = 2 * p + s;
        return z < NZ ? z : NZ - 1;
    }
};

// Output column (within the block) of M index mu.
__host__ __device__ constexpr int jmap32(int mu) {
    return 16 * (mu >> 4) + 8 * ((mu >> 2) & 1) + 4 * ((mu >> 3) & 1) + (mu & 3);
}

// Coefficient fragments: lane (mrow, h) = (l%32, l/32) holds A[mrow][8h + e] of
// each MFMA, i.e. c[z][k - j] for output column j = jmap32(mrow) and the slot's
// (z, k).  A full group (k = 8q .. 8q+7) is the window at d0 = 8q - j of row z;
// a half-group pair takes the first four taps of the window at d0 = 32 - j of
// two rows (zeros for a row past the last shift).  All come from CoefWindows.
template <int ZLO, int ZHI, int PREC>
struct Coef32 {
    using F = Fold32<ZLO, ZHI>;
    typedef typename Frag32<PREC>::type frag_t;
    frag_t f[F::NMF];

    static __device__ __forceinline__ int window(int d0) { return d0 >= -7 && d0 <= 4 ? d0 + 7 : 12; }

    __device__ __forceinline__ void load(const CoefWindows &cw, int mrow, int h) {
        const int j = jmap32(mrow);
#pragma unroll
        for (int mf = 0; mf < F::NMF; ++mf) {
            if (mf < 2 * F::NZ) {
                const uint32_t *w = cw.w[0][ZLO + (mf >> 1)][window(8 * (2 * (mf & 1) + h) - j)];
                f[mf] = __builtin_bit_cast(frag_t, u32x4{w[0], w[1], w[2], w[3]});
            } else {
                const int z0 = 2 * F::pair(mf, h), jw = window(32 - j);
                const uint32_t *w0 = cw.w[0][ZLO + (z0 < F::NZ ? z0 : 0)][z0 < F::NZ ? jw : 12];
                const uint32_t *w1 = cw.w[0][ZLO + (z0 + 1 < F::NZ ? z0 + 1 : 0)][z0 + 1 < F::NZ ? jw : 12];
                f[mf] = __builtin_bit_cast(frag_t, u32x4{w0[0], w0[1], w1[0], w1[1]});
            }
        }
    }
};

// Operands of the block whose haloed window starts at LDS row 16*br, column
// 16*bc, in chunks of CH MFMAs (chunk c: MFMAs c*CH .. c*CH+CH-1), so that one
// chunk's reads are in flight while the previous chunk's MFMAs run.  Chunk
// indices are compile-time constants once the block loops are unrolled.
template <class G, int ZLO, int ZHI>
struct Ops32 {
    using F = Fold32<ZLO, ZHI>;
    static constexpr int CH = 4;
    static constexpr int NCK = (F::NMF + CH - 1) / CH;
    u32x4 f[CH];

    __device__ __forceinline__ void load(const uint16_t *lds, int c, int br, int bc, int i, int h) {
        const uint16_t *base = lds + (16 * br + i + ZLO) * G::LC + 16 * bc;
#pragma unroll
        for (int t = 0; t < CH; ++t) {
            const int mf = c * CH + t;
            if (mf >= F::NMF) break;
            if (mf < 2 * F::NZ) {
                f[t] = *reinterpret_cast<const u32x4 *>(base + (mf >> 1) * G::LC + 8 * (2 * (mf & 1) + h));
            } else {
                const int p = F::pair(mf, h);
                const u32x2 x = *reinterpret_cast<const u32x2 *>(base + F::half_shift(p, 0) * G::LC + 32);
                const u32x2 y = *reinterpret_cast<const u32x2 *>(base + F::half_shift(p, 1) * G::LC + 32);
                f[t] = u32x4{x[0], x[1], y[0], y[1]};
            }
        }
    }
};

template <class G, int ZLO, int ZHI, int PREC>
__device__ __forceinline__ f32x16 mma_chunk32(const Coef32<ZLO, ZHI, PREC> &K, const Ops32<G, ZLO, ZHI> &op, int c,
                                              f32x16 acc) {
    using F32 = Frag32<PREC>;
    typedef typename F32::type frag_t;
    constexpr int CH = Ops32<G, ZLO, ZHI>::CH;
#pragma unroll
    for (int t = 0; t < CH; ++t) {
        const int mf = c * CH + t;
        if (mf >= Fold32<ZLO, ZHI>::NMF) break;
        acc = F32::mma(K.f[mf], __builtin_bit_cast(frag_t, op.f[t]), acc);
    }
    return acc;
}

// Epilogue of a block: lane (i, h) gets row i, columns 16s + 8h .. +7 (s = 0, 1:
// left and right brick) in storage format as w[s].  gr: global row of the
// lane's outputs, gcb: global column of the block's first column; with `edge`,
// points outside the interior keep their input value (read from the LDS image).
template <class G, int PREC>
__device__ __forceinline__ void epilogue32(const SweepArgs &a, const uint16_t *lds, const f32x16 &acc, int br,
                                           int bc, int i, int h, bool edge, int gr, int gcb, u32x4 (&w)[2]) {
    uint16_t o[16];
#pragma unroll
    for (int v = 0; v < 16; v += 2) {
        const uint32_t pr = PREC == TCS_PREC_BF16 ? bf16x2_bits(acc[v], acc[v + 1]) : f16x2_bits(acc[v], acc[v + 1]);
        o[v] = (uint16_t)pr;
        o[v + 1] = (uint16_t)(pr >> 16);
    }
    if (edge) {
        const uint16_t *row = lds + (16 * br + i + 2) * G::LC + 16 * bc + 2;
#pragma unroll
        for (int v = 0; v < 16; ++v) {
            const int cb = 16 * (v >> 3) + 8 * h + (v & 7);
            if (!point_is_interior(a, gr, gcb + cb)) o[v] = row[cb];
        }
    }
#pragma unroll
    for (int s = 0; s < 2; ++s)
        w[s] = u32x4{pack2(o[8 * s], o[8 * s + 1]), pack2(o[8 * s + 2], o[8 * s + 3]),
                     pack2(o[8 * s + 4], o[8 * s + 5]), pack2(o[8 * s + 6], o[8 * s + 7])};
}

__device__ __forceinline__ void store16_nt(uint16_t *p, u32x4 w) {
    __builtin_nontemporal_store(w, reinterpret_cast<u32x4 *>(p));
}

// Runs the MFMAs of a wave's blocks -- block b = wave + 4q (q < NQ, b < NBLK) at
// LDS brick (BASE + 2*(b / BX), BASE + 2*(b % BX)) -- and body(q, br, bc, acc)
// on each result.  The operand reads run one chunk ahead of the MFMAs, across
// block boundaries, through two chunk buffers.
template <class G, int ZLO, int ZHI, int PREC, int NQ, int NBLK, int BX, int BASE, class Body>
__device__ __forceinline__ void for_blocks32(const uint16_t *lds, const Coef32<ZLO, ZHI, PREC> &K, int wave,
                                             int i, int h, Body body) {
    static_assert(NBLK >= 4 && NQ * 4 >= NBLK && (NQ - 1) * 4 < NBLK, "blocks per wave");
    constexpr int NCK = Ops32<G, ZLO, ZHI>::NCK;
    Ops32<G, ZLO, ZHI> buf[2];
    buf[0].load(lds, 0, BASE + 2 * (wave / BX), BASE + 2 * (wave % BX), i, h);
    f32x16 acc = {};
#pragma unroll
    for (int n = 0; n < NQ * NCK; ++n) {
        const int q = n / NCK, c = n % NCK;
        const int b = wave + 4 * q;
        if (NQ * 4 != NBLK && b >= NBLK) break;  // wave-uniform
        const int n1 = n + 1, b1 = wave + 4 * (n1 / NCK);
        if (n1 < NQ * NCK && (NQ * 4 == NBLK || b1 < NBLK))
            buf[n1 & 1].load(lds, n1 % NCK, BASE + 2 * (b1 / BX), BASE + 2 * (b1 % BX), i, h);
        __builtin_amdgcn_sched_barrier(0);
        if (c == 0) acc = f32x16{};
        acc = mma_chunk32<G, ZLO, ZHI, PREC>(K, buf[n & 1], c, acc);
        __builtin_amdgcn_sched_barrier(0);
        if (c == NCK - 1) body(q, BASE + 2 * (b / BX), BASE + 2 * (b % BX), acc);
    }
}

template <int TH, int TW, int ZLO, int ZHI, int PREC>
__device__ __forceinline__ void sweep_tile32(const SweepArgs &a, const CoefWindows &cw, uint16_t *lds, int gr0,
                                             int gc0) {
    using T = Tile<TH, TW, 1>;
    constexpr int BX = TW / 2, NBLK = (TH / 2) * BX;
    static_assert(TH % 2 == 0 && TW % 2 == 0 && NBLK % 4 == 0, "tile must split into 2x2-brick blocks over 4 waves");
    const int tid = threadIdx.x;
    {
        TileRegs<T> R;
        load_tile<T, false>(a, R, gr0, gc0, tid);
        write_tile<T>(R, lds, tid);  // columns past 16*TW+4 are never read here
    }
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 31, h = lane >> 5;
    Coef32<ZLO, ZHI, PREC> C;
    C.load(cw, i, h);
    __syncthreads();
    const bool edge = !tile_is_interior(a, gr0, gc0, 16 * TH, 16 * TW);
    for_blocks32<T, ZLO, ZHI, PREC, NBLK / 4, NBLK, BX, 0>(lds, C, wave, i, h, [&](int, int br, int bc, const f32x16 &acc) {
        const int gr = gr0 + 16 * br + i, gcb = gc0 + 16 * bc;
        u32x4 w[2];
        epilogue32<T, PREC>(a, lds, acc, br, bc, i, h, edge, gr, gcb, w);
        store16_nt(a.out + stor_index(gr, gcb + 8 * h, a.sbc), w[0]);
        store16_nt(a.out + stor_index(gr, gcb + 16 + 8 * h, a.sbc), w[1]);
    });
}

template <int TH, int TW, int ZLO, int ZHI, int PREC>
__global__ __launch_bounds__(kWG) void tcs_mfma32_sweep(SweepArgs a) {
    __shared__ __attribute__((aligned(16))) uint16_t lds[Tile<TH, TW, 1>::PLANE];
    const Schedule sc = schedule(a);
    if (sc.tile >= sc.end) return;
    int gr0, gc0;
    tile_origin<TH, TW>(a, sc.tile, gr0, gc0);
    sweep_tile32<TH, TW, ZLO, ZHI, PREC>(a, a.cw, lds, gr0, gc0);
}

// Temporal blocking with 32x32 fragments: the scheme of tcs_mfma_fused (one-
// brick ring recomputed by every sweep but the last, per-sweep rounding, bitwise
// equal to single sweeps) over 2x2-brick blocks.  The (TH+2) x (TW+2) image
// splits into blocks from brick (0, 0); the core of the last sweep into blocks
// from brick (1, 1).
template <int TH, int TW, int ZLO, int ZHI, int PREC>
__device__ __forceinline__ void fused_tile32(const SweepArgs &a, const CoefWindows &cw, uint16_t *lds, int gr0,
                                             int gc0, int nsweeps) {
    using G = FusedGeom<TH, TW, 1>;
    constexpr int BXA = G::BW / 2, NBA = (G::BH / 2) * BXA, QA = (NBA + 3) / 4;
    constexpr int BXC = TW / 2, NBC = (TH / 2) * BXC;
    static_assert(TH % 2 == 0 && TW % 2 == 0 && NBC % 4 == 0, "tile must split into 2x2-brick blocks over 4 waves");
    const int tid = threadIdx.x;
    const int er0 = gr0 - 16, ec0 = gc0 - 16;  // global point of image brick (0, 0)
    {
        TileRegs<G> R;
        load_tile<G, true>(a, R, er0, ec0, tid);
        write_tile<G>(R, lds, tid);
    }
    const int lane = tid & 63, wave = tid >> 6;
    const int i = lane & 31, h = lane >> 5;
    Coef32<ZLO, ZHI, PREC> C;
    C.load(cw, i, h);
    const bool edge = !tile_is_interior(a, er0, ec0, 16 * G::BH, 16 * G::BW);
    __syncthreads();

    uint32_t *l32 = reinterpret_cast<uint32_t *>(lds);
    for (int sw = 0; sw + 1 < nsweeps; ++sw) {
        u32x4 res[QA][2];
        for_blocks32<G, ZLO, ZHI, PREC, QA, NBA, BXA, 0>(lds, C, wave, i, h, [&](int q, int br, int bc, const f32x16 &acc) {
            epilogue32<G, PREC>(a, lds, acc, br, bc, i, h, edge, er0 + 16 * br + i, ec0 + 16 * bc, res[q]);
        });
        __syncthreads();  // every wave has read this sweep's input
#pragma unroll
        for (int qq = 0; qq < QA; ++qq) {
            const int b = wave + 4 * qq;
            if (QA * 4 != NBA && b >= NBA) break;
            const int br = 2 * (b / BXA), bc = 2 * (b % BXA);
            const int d = ((16 * br + i + 2) * G::LC + 16 * bc + 8 * h + 2) >> 1;
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int t = 0; t < 4; ++t) l32[d + 8 * s + t] = res[qq][s][t];
        }
        __syncthreads();
    }
    for_blocks32<G, ZLO, ZHI, PREC, NBC / 4, NBC, BXC, 1>(lds, C, wave, i, h, [&](int, int br, int bc, const f32x16 &acc) {
        const int gr = er0 + 16 * br + i, gcb = ec0 + 16 * bc;
        u32x4 w[2];
        epilogue32<G, PREC>(a, lds, acc, br, bc, i, h, edge, gr, gcb, w);
        store16_nt(a.out + stor_index(gr, gcb + 8 * h, a.sbc), w[0]);
        store16_nt(a.out + stor_index(gr, gcb + 16 + 8 * h, a.sbc), w[1]);
    });
}

template <int TH, int TW, int ZLO, int ZHI, int PREC>
__global__ __launch_bounds__(kWG) void tcs_mfma32_fused(SweepArgs a, int nsweeps) {
    __shared__ __attribute__((aligned(16))) uint16_t lds[FusedGeom<TH, TW, 1>::PLANE];
    const Schedule sc = schedule(a);
    if (sc.tile >= sc.end) return;
    int gr0, gc0;
    tile_origin<TH, TW>(a, sc.tile, gr0, gc0);
    fused_tile32<TH, TW, ZLO, ZHI, PREC>(a, a.cw, lds, gr0, gc0, nsweeps);
}

// Batched uneven grids: one launch sweeps every tile of a list drawn from many
// problems (each with its own grid, buffers, radius and coefficients).  The
// tile list is problem-major, and tiles are handed to XCDs in contiguous runs,
// so neighbouring tiles of one problem share an L2.  The radius selects one of
// two instantiations of the same brick code with a workgroup-uniform branch.
template <int TH, int TW, int ZLO, int ZHI, int PREC>
__device__ __forceinline__ void batch_tile(const SweepArgs &a, const CoefWindows &cw, uint16_t *lds, int gr0,
                                           int gc0) {
    constexpr int NPL = base_prec(PREC) == TCS_PREC_SPLIT ? 2 : 1;
    constexpr bool LOOKAHEAD = kLookahead<PREC> && !kHighOccupancy<TH, TW, PREC, false>;
    using T = Tile<TH, TW, NPL, true>;
    const int tid = threadIdx.x;
    {
        TileRegs<T> R;
        load_tile<T, false>(a, R, gr0, gc0, tid);
        zero_pad<T>(lds, tid);
        write_tile<T>(R, lds, tid);
    }
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, g = lane >> 4;
    Coef<ZLO, ZHI, PREC> C;
    C.load(cw, m, g);
    const typename Ops<T, ZLO, ZHI, PREC>::Offsets off(m, g);
    __syncthreads();
    const bool edge = !tile_is_interior(a, gr0, gc0, 16 * TH, 16 * TW);
    constexpr int NB = TH * TW / 4;
    Ops<T, ZLO, ZHI, PREC> cur, nxt;
    auto reuse = [](int q) { return q % BlockRows<TH, TW>::RW != 0; };
    if constexpr (LOOKAHEAD) {
        int br, bc;
        block_brick<TH, TW, 0>(wave, 0, br, bc);
        cur.load(lds, br, bc, off);
    }
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        int br, bc;
        block_brick<TH, TW, 0>(wave, q, br, bc);
        if constexpr (LOOKAHEAD) {
            if (q + 1 < NB) {
                int nbr, nbc;
                block_brick<TH, TW, 0>(wave, q + 1, nbr, nbc);
                nxt.load(lds, nbr, nbc, off, cur, reuse(q + 1));
            }
        } else {
            cur.load(lds, br, bc, off, cur, reuse(q));
        }
        __builtin_amdgcn_sched_barrier(0);
        const int gr = gr0 + 16 * br + m, gcb = gc0 + 16 * bc + 4 * g;
        u32x2 w[NPL];
        mfma_brick<T, ZLO, ZHI, PREC>(a, lds, C, cur, br, bc, m, g, edge, gr, gcb, w);
        const int64_t s = stor_index(gr, gcb, a.sbc);
#pragma unroll
        for (int p = 0; p < NPL; ++p) store8<true>(a.out + p * a.plane + s, w[p]);
        if constexpr (LOOKAHEAD) cur = nxt;
    }
}

// One tile of a batch launch: tile idx of the list, FUSED: nsweeps sweeps in
// LDS (temporal blocking), else one sweep.  Reads buf[parity] and writes
// buf[parity ^ 1] of its problem; the radius selects one of two inlined
// instantiations with a workgroup-uniform branch.
template <int TH, int TW, int PREC, int FRAG, bool FUSED>
__device__ __forceinline__ void batch_body(const BatchProblem *__restrict__ probs,
                                           const BatchTile *__restrict__ tiles, int idx, int parity, int nsweeps,
                                           uint16_t *lds) {
    const BatchTile bt = tiles[idx];
    const BatchProblem &P = probs[bt.prob];
    SweepArgs a;
    a.in = P.buf[parity];
    a.out = P.buf[parity ^ 1];
    a.plane = P.plane;
    a.rows = P.rows;
    a.cols = P.cols;
    a.sbr = P.sbr;
    a.sbc = P.sbc;
    a.tiles_x = P.tiles_x;
    a.tiles_xp = P.tiles_x;  // the tile list holds real tile indices
    a.nvirt = 0;
    a.tiles_y = 0;
    a.rs = P.rs;
    a.out_scale = P.out_scale;
    int gr0, gc0;
    tile_origin<TH, TW>(a, bt.tile, gr0, gc0);
    if constexpr (FUSED && FRAG == 32) {
        if (a.rs == 2)
            fused_tile32<TH, TW, 0, 4, PREC>(a, P.cw, lds, gr0, gc0, nsweeps);
        else
            fused_tile32<TH, TW, 1, 3, PREC>(a, P.cw, lds, gr0, gc0, nsweeps);
    } else if constexpr (FUSED) {
        if (a.rs == 2)
            fused_tile<TH, TW, 0, 4, PREC, 4>(a, P.cw, lds, gr0, gc0, nsweeps);
        else
            fused_tile<TH, TW, 1, 3, PREC, 4>(a, P.cw, lds, gr0, gc0, nsweeps);
    } else if constexpr (FRAG == 32) {
        if (a.rs == 2)
            sweep_tile32<TH, TW, 0, 4, PREC>(a, P.cw, lds, gr0, gc0);
        else
            sweep_tile32<TH, TW, 1, 3, PREC>(a, P.cw, lds, gr0, gc0);
    } else {
        if (a.rs == 2)
            batch_tile<TH, TW, 0, 4, PREC>(a, P.cw, lds, gr0, gc0);
        else
            batch_tile<TH, TW, 1, 3, PREC>(a, P.cw, lds, gr0, gc0);
    }
}

// One launch covers `ntiles` tiles of the list.  Tiles are handed to XCDs in
// contiguous runs of tiles_per_xcd (the list is problem-major, so a problem's
// neighbouring tiles share an L2).  PERSIST: workgroup w on XCD x = w % 8
// walks the XCD's run with stride gridDim.x / 8; else one tile per workgroup,
// which needs no loop state and lets small tiles run 8 waves per SIMD.
template <int TH, int TW, int PREC, int FRAG, bool FUSED, bool PERSIST>
__global__ __launch_bounds__(kWG)
__attribute__((amdgpu_waves_per_eu(FRAG == 16 && !PERSIST && base_prec(PREC) != TCS_PREC_SPLIT
                                       ? (FUSED ? 4 : (kHighOccupancy<TH, TW, PREC, false> ? 8 : 1))
                                       : 1,
                                   8)))
void tcs_mfma_batch(const BatchProblem *__restrict__ probs, const BatchTile *__restrict__ tiles, int ntiles,
                    int tiles_per_xcd, int parity, int nsweeps) {
    constexpr int NPL = base_prec(PREC) == TCS_PREC_SPLIT ? 2 : 1;
    constexpr int PLANE = FUSED ? FusedGeom<TH, TW, NPL, FRAG == 16>::PLANE : Tile<TH, TW, NPL, FRAG == 16>::PLANE;
    __shared__ __attribute__((aligned(16))) uint16_t lds[NPL * PLANE];
    const int x = blockIdx.x % kXCD, step = gridDim.x / kXCD;
    const int begin = x * tiles_per_xcd;
    const int end = begin + tiles_per_xcd < ntiles ? begin + tiles_per_xcd : ntiles;
    const int first = begin + (int)blockIdx.x / kXCD;
    if constexpr (PERSIST) {
        for (int idx = first; idx < end; idx += step) {
            if (idx != first) __syncthreads();  // the previous tile is done with LDS
            batch_body<TH, TW, PREC, FRAG, FUSED>(probs, tiles, idx, parity, nsweeps, lds);
        }
    } else {
        if (first < end) batch_body<TH, TW, PREC, FRAG, FUSED>(probs, tiles, first, parity, nsweeps, lds);
    }
}

// ----------------------------------------------------------- direct kernel
template <int STENCIL>
__host__ __device__ constexpr bool tap(int dz, int dj) {
    if (STENCIL == TCS_STAR2D1R) return (dz == 2 && dj >= 1 && dj <= 3) || (dj == 2 && (dz == 1 || dz == 3));
    if (STENCIL == TCS_STAR2D2R) return dz == 2 || dj == 2;
    if (STENCIL == TCS_BOX2D1R) return dz >= 1 && dz <= 3 && dj >= 1 && dj <= 3;
    return true;
}

// Contiguous tap run [first, last] of stencil row dz (every row of the four
// stencils is one run): star rows off the centre hold the middle tap only.
template <int STENCIL>
__host__ __device__ constexpr int tap_first(int dz) {
    for (int dj = 0; dj < 5; ++dj)
        if (tap<STENCIL>(dz, dj)) return dj;
    return 5;
}
template <int STENCIL>
__host__ __device__ constexpr int tap_last(int dz) {
    for (int dj = 4; dj >= 0; --dj)
        if (tap<STENCIL>(dz, dj)) return dj;
    return -1;
}

typedef _Float16 half2_t __attribute__((ext_vector_type(2)));

// Direct kernels store their results through an LDS image of the tile as whole
// bricks (16-byte pieces, two bricks per wave instruction) rather than as
// 8-byte pieces of 8 bricks per instruction, except box2d2r, which is bound by
// its arithmetic and runs faster without the two extra barriers.  Measured on
// 16384^2, best tile each, staged vs not (us): star2d1r 179 vs 228, box2d1r
// 198 vs 237, star2d2r 198 vs 237, box2d2r 295 vs 284.
template <int STENCIL>
constexpr bool kDirectStaged = STENCIL != TCS_BOX2D2R;

__device__ __forceinline__ float dot2(uint32_t x, uint32_t c, float acc) {
    return __builtin_amdgcn_fdot2(__builtin_bit_cast(half2_t, x), __builtin_bit_cast(half2_t, c), acc, false);
}

// The direct kernel (FP32 taps, one FMA per tap and point) and, with DOT2, its
// packed-half form: taps rounded to FP16 (as the MFMA kernel's), each pair of
// adjacent taps of a row one v_dot2_f32_f16 (two FP16 products added into the
// FP32 accumulator), the odd tap of a run paired with a zero half.  Per point of
// box2d2r that is 15 dot products instead of 25 FMAs and 2 conversions.  A
// non-finite input reaches the outputs whose pairs hold it (times a zero tap),
// a superset of gold's set as for the MFMA kernel.
template <int TH, int TW, int STENCIL, bool PERSIST, bool DOT2 = false>
__global__ __launch_bounds__(kWG) void tcs_direct_sweep(SweepArgs a) {
    using T = Tile<TH, TW, 1>;
    constexpr int CG = 4 * TW;       // 4-column groups across the tile
    constexpr int RG = kWG / CG;     // row groups
    constexpr int MR = 16 * TH / RG; // output rows per thread
    static_assert(RG * CG == kWG && MR * RG == 16 * TH, "tile does not split into strips");
    __shared__ __attribute__((aligned(16))) uint16_t lds[T::PLANE];

    Schedule sc = schedule(a);
    if (sc.tile >= sc.end) return;
    const int tid = threadIdx.x;
    int gr0, gc0;
    tile_origin<TH, TW>(a, sc.tile, gr0, gc0);
    TileRegs<T> R;
    load_tile<T, false>(a, R, gr0, gc0, tid);
    zero_pad<T>(lds, tid);

    const int cg = tid % CG, rg = tid / CG;
    const int r0 = rg * MR;  // first output row of this strip (tile coordinates)

    // DOT2: FP16 tap pairs (c[dz][j], c[dz][j+1]) and single taps (c[dz][j], 0),
    // rounded to nearest like the MFMA kernel's taps.
    uint32_t cpair[DOT2 ? 25 : 1], csing[DOT2 ? 25 : 1];
    if constexpr (DOT2) {
#pragma unroll
        for (int k = 0; k < 25; ++k) {
            const uint16_t c0 = f16_bits(a.cf[k]);
            const uint16_t c1 = (k % 5) < 4 ? f16_bits(a.cf[k + 1]) : (uint16_t)0;
            // uniform across the wave: held in scalar registers
            cpair[k] = __builtin_amdgcn_readfirstlane(pack2(c0, c1));
            csing[k] = __builtin_amdgcn_readfirstlane((uint32_t)c0);
        }
    }

    for (;;) {
    write_tile<T>(R, lds, tid);
    __syncthreads();
    const int cgr0 = gr0, cgc0 = gc0;  // the tile in LDS
    const int next = skip_empty(a, sc.tile + sc.step, sc.step);
    if (PERSIST && next < sc.end) {
        tile_origin<TH, TW>(a, next, gr0, gc0);
        load_tile<T, false>(a, R, gr0, gc0, tid);  // in flight during the compute below
    }
    float acc[MR][4];
#pragma unroll
    for (int o = 0; o < MR; ++o)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[o][v] = 0.0f;

    // Input row r0 + ir (LDS row, haloed) feeds output row o with dz = ir - o.
    // Terms are added in (dz, dj) row-major order, the order of gold's k-loop.
#pragma unroll
    for (int ir = 0; ir < MR + 4; ++ir) {
        bool used = false;
#pragma unroll
        for (int o = 0; o < MR; ++o)
#pragma unroll
            for (int dj = 0; dj < 5; ++dj)
                if (ir - o >= 0 && ir - o <= 4 && tap<STENCIL>(ir - o, dj)) used = true;
        if (!used) continue;
        const uint16_t *row = lds + (r0 + ir) * T::LC + 4 * cg;
        const u32x2 lo = *reinterpret_cast<const u32x2 *>(row);
        const u32x2 hi = *reinterpret_cast<const u32x2 *>(row + 4);
        const uint32_t w4[4] = {lo[0], lo[1], hi[0], hi[1]};
        float x[8];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            x[2 * t] = static_cast<float>(__builtin_bit_cast(_Float16, (uint16_t)(w4[t] & 0xffffu)));
            x[2 * t + 1] = static_cast<float>(__builtin_bit_cast(_Float16, (uint16_t)(w4[t] >> 16)));
        }
#pragma unroll
        for (int o = 0; o < MR; ++o) {
            const int dz = ir - o;
            if (dz < 0 || dz > 4) continue;
            if constexpr (DOT2) {
                // Output column v takes x[v + dj], dj in the row's run: pairs
                // (x[k], x[k+1]) are whole dwords for even k and an alignbit of
                // two dwords for odd k; a last single x[k] is a dword with its
                // other half zeroed.
                const int ta = tap_first<STENCIL>(dz), tb = tap_last<STENCIL>(dz);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float sacc = acc[o][v];
#pragma unroll
                    for (int k = v + ta; k <= v + tb; k += 2) {
                        if (k + 1 <= v + tb) {
                            const uint32_t xp = (k & 1) ? __builtin_amdgcn_alignbit(w4[(k + 1) >> 1], w4[k >> 1], 16)
                                                        : w4[k >> 1];
                            sacc = dot2(xp, cpair[dz * 5 + k - v], sacc);
                        } else {
                            const uint32_t xs = (k & 1) ? (w4[k >> 1] >> 16) : (w4[k >> 1] & 0xffffu);
                            sacc = dot2(xs, csing[dz * 5 + k - v], sacc);
                        }
                    }
                    acc[o][v] = sacc;
                }
            } else {
#pragma unroll
                for (int dj = 0; dj < 5; ++dj) {
                    if (!tap<STENCIL>(dz, dj)) continue;
                    const float c = a.cf[dz * 5 + dj];
#pragma unroll
                    for (int v = 0; v < 4; ++v) acc[o][v] = __builtin_fmaf(x[v + dj], c, acc[o][v]);
                }
            }
        }
    }

    const bool edge = !tile_is_interior(a, cgr0, cgc0, 16 * TH, 16 * TW);
    u32x2 res[MR];
#pragma unroll
    for (int o = 0; o < MR; ++o) {
        const int tr = r0 + o;                  // tile row
        const int gr = cgr0 + tr;
        const int gcb = cgc0 + 4 * cg;
        uint16_t h[4];
#pragma unroll
        for (int v = 0; v < 4; v += 2) {
            const uint32_t pr = f16x2_bits(acc[o][v], acc[o][v + 1]);
            h[v] = (uint16_t)pr;
            h[v + 1] = (uint16_t)(pr >> 16);
        }
        if (edge) {
            // the input points, two dwords read whatever the points are, then selected
            const int cidx = (tr + 2) * T::LC + 4 * cg + 2;
            const uint32_t *l32 = reinterpret_cast<const uint32_t *>(lds);
            const uint32_t k01 = l32[cidx >> 1], k23 = l32[(cidx >> 1) + 1];
            const bool row_in = gr >= a.rs && gr < a.rows - a.rs;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gc = gcb + v;
                const bool interior = row_in && gc >= a.rs && gc < a.cols - a.rs;
                h[v] = interior ? h[v] : (uint16_t)((v < 2 ? k01 : k23) >> (16 * (v & 1)));
            }
        }
        res[o] = u32x2{pack2(h[0], h[1]), pack2(h[2], h[3])};
        if constexpr (!kDirectStaged<STENCIL>) store8<false>(a.out + stor_index(gr, gcb, a.sbc), res[o]);
    }
    if constexpr (kDirectStaged<STENCIL>) {
    // The results go out as whole bricks, as the MFMA kernel's do: a row-major
    // image of the tile in LDS (over the input image, which every thread has
    // finished reading), then 16-byte pieces with consecutive lanes along a
    // brick, so each store instruction of a wave writes two whole bricks.
    constexpr int OP = 16 * TW + 8;  // output image pitch (elements)
    static_assert(16 * TH * OP <= T::PLANE, "output image must fit the input image");
    __syncthreads();
#pragma unroll
    for (int o = 0; o < MR; ++o) *reinterpret_cast<u32x2 *>(lds + (r0 + o) * OP + 4 * cg) = res[o];
    __syncthreads();
    constexpr int NCH = TH * TW * 32;  // 16-byte pieces of the tile's bricks
    static_assert(NCH % kWG == 0, "whole passes of the workgroup");
#pragma unroll
    for (int i = 0; i < NCH / kWG; ++i) {
        const int c = tid + i * kWG;
        const int b = c >> 5, w = c & 31;
        const int br = b / TW, bc = b - br * TW;
        const int tr = 16 * br + (w >> 1), tc = 16 * bc + 8 * (w & 1);
        const u32x4 v = *reinterpret_cast<const u32x4 *>(lds + tr * OP + tc);
        store16_nt(a.out + stor_index(cgr0 + tr, cgc0 + tc, a.sbc), v);
    }
    }
    // One-tile-per-workgroup launches cover every tile with the grid itself.
    if (!PERSIST || next >= sc.end) break;
    sc.tile = next;
    __syncthreads();  // every wave is done reading LDS before it is overwritten
    }
}

// ---------------------------------------------------------------- dispatch
// This file is compiled once per part (-DTCS_PART=n): parts 1..6 each
// instantiate one group of kernels and export its tables, part 0 holds the
// selection and launch code over all of them.  The parts compile in parallel.
//   1 FP16, 2 BF16, 3 SPLIT (16x16 fragments: single sweep, fused, edge, batch)
//   4 FP16 and 5 BF16 with hi+lo coefficients (the same kernel kinds)
//   6 32x32 fragments (single sweep and fused) and the direct-FMA kernel
//   7 the packed-half direct kernel
#ifndef TCS_PART
#error "compile with -DTCS_PART=0..7"
#endif

typedef void (*sweep_fn)(SweepArgs);
typedef void (*fused_fn)(SweepArgs, int);
typedef void (*batch_fn)(const BatchProblem *, const BatchTile *, int, int, int, int);

struct KernelEntry {
    int method, prec, radius, stencil, th, tw;  // prec: kernel precision id; stencil = -1: any of that radius
    bool persist;
    int frag;  // MFMA fragment edge (16 or 32); 0 for the direct kernel
    sweep_fn fn;
    const char *name;  // as rocprofv3 reports it (demangled, without the namespace)
};

// Fused (temporally blocked) MFMA kernels; nt threads per workgroup; the
// entry serves launches of at least min_sweeps sweeps.
struct FusedEntry {
    int prec, radius, th, tw, frag, nt, min_sweeps;
    fused_fn fn;
    const char *name;
};

// Batched kernels (16x16 and 32x32 fragments, one sweep or several fused).
struct BatchEntry {
    int prec, th, tw, frag;
    bool fused, persist;
    batch_fn fn;
    const char *name;
};

// Edge kernels (tcs_run_f32): 16x16 fragments, tiles 2x4, 4x4, 2x8, 4x8, one
// tile per workgroup.
struct EdgeEntry {
    int prec, radius, th, tw, mode;
    sweep_fn fn;
    const char *name;
};

struct PartTables {
    const KernelEntry *sweep;
    int nsweep;
    const FusedEntry *fused;
    int nfused;
    const BatchEntry *batch;
    int nbatch;
    const EdgeEntry *edge;
    int nedge;
};

#define TCS_N(a) (int)(sizeof(a) / sizeof((a)[0]))

#define TCS_MFMA(TH, TW, ZLO, ZHI, P, R, PS) \
    {TCS_METHOD_MFMA, P, R, -1, TH, TW, PS, 16, tcs_mfma_sweep<TH, TW, ZLO, ZHI, P, PS>, \
     "tcs_mfma_sweep<" #TH ", " #TW ", " #ZLO ", " #ZHI ", " #P ", " #PS ">"}
#define TCS_MFMA_TILES(ZLO, ZHI, P, R, PS) \
    TCS_MFMA(2, 4, ZLO, ZHI, P, R, PS), TCS_MFMA(4, 4, ZLO, ZHI, P, R, PS), TCS_MFMA(2, 8, ZLO, ZHI, P, R, PS), \
    TCS_MFMA(4, 8, ZLO, ZHI, P, R, PS), TCS_MFMA(8, 8, ZLO, ZHI, P, R, PS)
// Every 16x16 single-sweep tile of precision P, both radii, one tile per
// workgroup and persistent.
#define TCS_MFMA_PREC(P)                                                        \
    TCS_MFMA_TILES(0, 4, P, 2, false), TCS_MFMA_TILES(1, 3, P, 1, false), \
        TCS_MFMA_TILES(0, 4, P, 2, true), TCS_MFMA_TILES(1, 3, P, 1, true)
#define TCS_DIRECT(TH, TW, S, PS) \
    {TCS_METHOD_DIRECT, TCS_PREC_FP16, -1, S, TH, TW, PS, 0, tcs_direct_sweep<TH, TW, S, PS>, \
     "tcs_direct_sweep<" #TH ", " #TW ", " #S ", " #PS ">"}
#define TCS_DIRECT_TILES(S, PS) \
    TCS_DIRECT(2, 4, S, PS), TCS_DIRECT(4, 4, S, PS), TCS_DIRECT(2, 8, S, PS), TCS_DIRECT(4, 8, S, PS), \
    TCS_DIRECT(8, 8, S, PS)
#define TCS_DOT2(TH, TW, S, PS) \
    {TCS_METHOD_DIRECT_DOT2, TCS_PREC_FP16, -1, S, TH, TW, PS, 0, tcs_direct_sweep<TH, TW, S, PS, true>, \
     "tcs_direct_sweep<" #TH ", " #TW ", " #S ", " #PS ", true>"}
#define TCS_DOT2_TILES(S, PS) \
    TCS_DOT2(2, 4, S, PS), TCS_DOT2(4, 4, S, PS), TCS_DOT2(2, 8, S, PS), TCS_DOT2(4, 8, S, PS), TCS_DOT2(8, 8, S, PS)
#define TCS_MFMA32(TH, TW, ZLO, ZHI, P, R) \
    {TCS_METHOD_MFMA, P, R, -1, TH, TW, false, 32, tcs_mfma32_sweep<TH, TW, ZLO, ZHI, P>, \
     "tcs_mfma32_sweep<" #TH ", " #TW ", " #ZLO ", " #ZHI ", " #P ">"}
#define TCS_MFMA32_TILES(ZLO, ZHI, P, R) \
    TCS_MFMA32(4, 4, ZLO, ZHI, P, R), TCS_MFMA32(2, 8, ZLO, ZHI, P, R), TCS_MFMA32(4, 8, ZLO, ZHI, P, R), \
    TCS_MFMA32(8, 8, ZLO, ZHI, P, R)

#define TCS_FUSED(TH, TW, ZLO, ZHI, P, R) \
    {P, R, TH, TW, 16, kWG, 1, tcs_mfma_fused<TH, TW, ZLO, ZHI, P>, \
     "tcs_mfma_fused<" #TH ", " #TW ", " #ZLO ", " #ZHI ", " #P ">"}
#define TCS_FUSED8(TH, TW, ZLO, ZHI, P, R, MINS) \
    {P, R, TH, TW, 16, 2 * kWG, MINS, tcs_mfma_fused8<TH, TW, ZLO, ZHI, P>, \
     "tcs_mfma_fused8<" #TH ", " #TW ", " #ZLO ", " #ZHI ", " #P ">"}
#define TCS_FUSED_TILES(ZLO, ZHI, P, R)                                                           \
    TCS_FUSED(2, 4, ZLO, ZHI, P, R), TCS_FUSED(4, 4, ZLO, ZHI, P, R), TCS_FUSED(2, 8, ZLO, ZHI, P, R), \
        TCS_FUSED(4, 8, ZLO, ZHI, P, R)
#define TCS_FUSED_TILES88(ZLO, ZHI, P, R) TCS_FUSED_TILES(ZLO, ZHI, P, R), TCS_FUSED(8, 8, ZLO, ZHI, P, R)
#define TCS_FUSED_PREC88(P) TCS_FUSED_TILES88(0, 4, P, 2), TCS_FUSED_TILES88(1, 3, P, 1)
// FP16/BF16, rounded or hi+lo taps: 8x8 tiles on 8-wave workgroups
// (fused_problem maps large 4x8 problems to them), smaller tiles on 4 waves.
// With hi+lo taps, box2d2r 16384^2, 8 sweeps per launch, us per sweep: 8x8 on
// 8 waves 128.9, 4x8 on 4 waves 135.7, 8x8 on 4 waves 140.1 (radius 1, 16
// sweeps: 98.7, 102.3, 115.7).
#define TCS_FUSED_PREC(P) \
    TCS_FUSED8(8, 8, 0, 4, P, 2, 1), TCS_FUSED8(8, 8, 1, 3, P, 1, 1), TCS_FUSED_TILES(0, 4, P, 2), \
        TCS_FUSED_TILES(1, 3, P, 1)
#define TCS_FUSED32(TH, TW, ZLO, ZHI, P, R) \
    {P, R, TH, TW, 32, kWG, 1, tcs_mfma32_fused<TH, TW, ZLO, ZHI, P>, \
     "tcs_mfma32_fused<" #TH ", " #TW ", " #ZLO ", " #ZHI ", " #P ">"}
#define TCS_FUSED32_TILES(ZLO, ZHI, P, R) \
    TCS_FUSED32(4, 4, ZLO, ZHI, P, R), TCS_FUSED32(4, 8, ZLO, ZHI, P, R), TCS_FUSED32(8, 8, ZLO, ZHI, P, R)

#define TCS_BATCH(TH, TW, P, F, FU, PS)                                          \
    {P, TH, TW, F, FU, PS, tcs_mfma_batch<TH, TW, P, F, FU, PS>,               \
     "tcs_mfma_batch<" #TH ", " #TW ", " #P ", " #F ", " #FU ", " #PS ">"}
#define TCS_BATCH16(P, FU, PS) TCS_BATCH(2, 4, P, 16, FU, PS), TCS_BATCH(4, 4, P, 16, FU, PS), TCS_BATCH(4, 8, P, 16, FU, PS)
#define TCS_BATCH32(P, FU, PS) TCS_BATCH(4, 4, P, 32, FU, PS), TCS_BATCH(4, 8, P, 32, FU, PS)
#define TCS_BATCH16_PREC(P) \
    TCS_BATCH16(P, false, false), TCS_BATCH16(P, true, false), TCS_BATCH16(P, false, true), TCS_BATCH16(P, true, true)
#define TCS_BATCH32_PREC(P) \
    TCS_BATCH32(P, false, false), TCS_BATCH32(P, true, false), TCS_BATCH32(P, false, true), TCS_BATCH32(P, true, true)

#define TCS_EDGE(TH, TW, ZLO, ZHI, P, R, M)                          \
    {P, R, TH, TW, M, tcs_mfma_edge<TH, TW, ZLO, ZHI, P, M>,          \
     "tcs_mfma_edge<" #TH ", " #TW ", " #ZLO ", " #ZHI ", " #P ", " #M ">"}
#define TCS_EDGE_MODES(TH, TW, ZLO, ZHI, P, R) \
    TCS_EDGE(TH, TW, ZLO, ZHI, P, R, 1), TCS_EDGE(TH, TW, ZLO, ZHI, P, R, 2), TCS_EDGE(TH, TW, ZLO, ZHI, P, R, 3)
#define TCS_EDGE_TILES(ZLO, ZHI, P, R)                                                                   \
    TCS_EDGE_MODES(2, 4, ZLO, ZHI, P, R), TCS_EDGE_MODES(4, 4, ZLO, ZHI, P, R), TCS_EDGE_MODES(2, 8, ZLO, ZHI, P, R), \
        TCS_EDGE_MODES(4, 8, ZLO, ZHI, P, R), TCS_EDGE_MODES(2, 16, ZLO, ZHI, P, R)
#define TCS_EDGE_PREC(P) TCS_EDGE_TILES(0, 4, P, 2), TCS_EDGE_TILES(1, 3, P, 1)

// The tables below spell kernel precision ids as literals (0 FP16, 1 BF16,
// 2 SPLIT, 4 and 5 FP16 and BF16 with hi+lo coefficients) so that the kernel
// names carry the numbers rocprofv3 prints.
static_assert(TCS_PREC_FP16 == 0 && TCS_PREC_BF16 == 1 && TCS_PREC_SPLIT == 2 && kCoefLo == 4, "precision ids");
static_assert(TCS_STAR2D1R == 0 && TCS_STAR2D2R == 1 && TCS_BOX2D1R == 2 && TCS_BOX2D2R == 3, "stencil ids");

// The 16x16-fragment kernels of one kernel precision: single sweep, fused,
// edge, and (FP16/BF16, with or without hi+lo coefficients) batched.
#define TCS_PART_16(NAME, P, FUSED, BATCH...)                                                 \
    static const KernelEntry kSweep##NAME[] = {TCS_MFMA_PREC(P)};                             \
    static const FusedEntry kFused##NAME[] = {FUSED};                                         \
    static const EdgeEntry kEdge##NAME[] = {TCS_EDGE_PREC(P)};                                \
    static const BatchEntry kBatch##NAME[] = {BATCH};                                         \
    extern const PartTables kPart##NAME = {kSweep##NAME, TCS_N(kSweep##NAME), kFused##NAME,   \
                                           TCS_N(kFused##NAME), kBatch##NAME, TCS_N(kBatch##NAME), \
                                           kEdge##NAME, TCS_N(kEdge##NAME)};

#if TCS_PART == 1
TCS_PART_16(Fp16, 0, TCS_FUSED_PREC(0), TCS_BATCH16_PREC(0), TCS_BATCH32_PREC(0))
#elif TCS_PART == 2
TCS_PART_16(Bf16, 1, TCS_FUSED_PREC(1), TCS_BATCH16_PREC(1), TCS_BATCH32_PREC(1))
#elif TCS_PART == 3
// SPLIT (8x8: two workgroups per CU at 72 KB of LDS); no batched kernels.
static const KernelEntry kSweepSplit[] = {TCS_MFMA_PREC(2)};
// 4x8 tiles on 8-wave workgroups for launches of 6 or more sweeps, first
// (select_fused takes the first match): box2d2r 16384^2, us per sweep, 8 waves
// against 4: k = 6 277 / 280, 7 275 / 281, 8 274 / 281, 9 273 / 283; k = 5
// ties (278 / 278), k = 4 296 / 287 and k = 2 374 / 341 favour 4 waves.
static const FusedEntry kFusedSplit[] = {TCS_FUSED8(4, 8, 0, 4, 2, 2, 6), TCS_FUSED8(4, 8, 1, 3, 2, 1, 6),
                                         TCS_FUSED_PREC88(2)};
static const EdgeEntry kEdgeSplit[] = {TCS_EDGE_PREC(2)};
extern const PartTables kPartSplit = {kSweepSplit, TCS_N(kSweepSplit), kFusedSplit, TCS_N(kFusedSplit),
                                      nullptr, 0, kEdgeSplit, TCS_N(kEdgeSplit)};
#elif TCS_PART == 4
TCS_PART_16(Fp16Lo, 4, TCS_FUSED_PREC(4), TCS_BATCH16_PREC(4))
#elif TCS_PART == 5
TCS_PART_16(Bf16Lo, 5, TCS_FUSED_PREC(5), TCS_BATCH16_PREC(5))
#elif TCS_PART == 6
// 32x32 fragments (FP16 and BF16, tiles whose 2x2-brick blocks split evenly over
// the 4 waves, one tile per workgroup) and the direct-FMA kernel.
static const KernelEntry kSweepOther[] = {
    TCS_MFMA32_TILES(0, 4, 0, 2), TCS_MFMA32_TILES(1, 3, 0, 1), TCS_MFMA32_TILES(0, 4, 1, 2), TCS_MFMA32_TILES(1, 3, 1, 1),
    TCS_DIRECT_TILES(0, false), TCS_DIRECT_TILES(1, false), TCS_DIRECT_TILES(2, false), TCS_DIRECT_TILES(3, false),
    TCS_DIRECT_TILES(0, true), TCS_DIRECT_TILES(1, true), TCS_DIRECT_TILES(2, true), TCS_DIRECT_TILES(3, true),
};
static const FusedEntry kFusedOther[] = {
    TCS_FUSED32_TILES(0, 4, 0, 2), TCS_FUSED32_TILES(1, 3, 0, 1), TCS_FUSED32_TILES(0, 4, 1, 2), TCS_FUSED32_TILES(1, 3, 1, 1),
};
extern const PartTables kPartOther = {kSweepOther, TCS_N(kSweepOther), kFusedOther, TCS_N(kFusedOther),
                                      nullptr, 0, nullptr, 0};
#elif TCS_PART == 7
// The packed-half direct kernels.
static const KernelEntry kSweepDot2[] = {
    TCS_DOT2_TILES(0, false), TCS_DOT2_TILES(1, false), TCS_DOT2_TILES(2, false), TCS_DOT2_TILES(3, false),
    TCS_DOT2_TILES(0, true), TCS_DOT2_TILES(1, true), TCS_DOT2_TILES(2, true), TCS_DOT2_TILES(3, true),
};
extern const PartTables kPartDot2 = {kSweepDot2, TCS_N(kSweepDot2), nullptr, 0, nullptr, 0, nullptr, 0};
#elif TCS_PART == 0
extern const PartTables kPartFp16, kPartBf16, kPartSplit, kPartFp16Lo, kPartBf16Lo, kPartOther, kPartDot2;
static const PartTables *const kParts[] = {&kPartFp16, &kPartBf16, &kPartSplit, &kPartFp16Lo, &kPartBf16Lo,
                                           &kPartOther, &kPartDot2};

static int frag_of(const tcs_problem *p) { return p->frag == 32 ? 32 : 16; }

static const KernelEntry *select_kernel(const tcs_problem *p) {
    const int kp = kernel_prec(p);
    for (const PartTables *t : kParts)
        for (int i = 0; i < t->nsweep; ++i) {
            const KernelEntry &k = t->sweep[i];
            if (k.method != p->method || k.th != p->tile_h || k.tw != p->tile_w || k.persist != (p->wg_per_cu > 0))
                continue;
            if (k.method == TCS_METHOD_MFMA ? (k.prec != kp || k.frag != frag_of(p) || k.radius != p->radius)
                                            : k.stencil != p->stencil)
                continue;
            return &k;
        }
    return nullptr;
}

// The kernel of a launch of `nsweeps` fused sweeps of problem p.
static const FusedEntry *select_fused(const tcs_problem *p, int nsweeps) {
    if (p->method != TCS_METHOD_MFMA) return nullptr;
    const int kp = kernel_prec(p);
    for (const PartTables *t : kParts)
        for (int i = 0; i < t->nfused; ++i) {
            const FusedEntry &k = t->fused[i];
            if (k.prec == kp && k.radius == p->radius && k.th == p->tile_h && k.tw == p->tile_w &&
                k.frag == frag_of(p) && nsweeps >= k.min_sweeps)
                return &k;
        }
    return nullptr;
}

// The problem a fused launch runs: the caller's, except that FP16/BF16 (rounded
// or hi+lo taps) with 16x16 fragments fuses 4x8 problems on 8x8 tiles (8-wave
// workgroups, tcs_mfma_fused8) where the layout divides into 8x8 tiles and
// makes at least 2048 of them (brick results depend on the tile in no way).
// An 8x8 tile recomputes 12% fewer ring bricks than two 4x8 tiles over 8
// sweeps, and its 56 KB image holds 2 workgroups of 8 waves per CU, 4 waves
// per SIMD as 4x8 tiles on 4-wave workgroups do (box2d2r 16384^2, us per
// sweep, 8x8 on 8 waves against 4x8 on 4: k = 4 107.5 / 113.5, 6 100.5 / 103.5,
// 8 96 / 100.5, 9 95.5 / 99.5).
static tcs_problem fused_problem(const tcs_problem *p) {
    tcs_problem q = *p;
    const int kp = kernel_prec(p);
    const tcs_layout &L = p->layout;
    if (p->method == TCS_METHOD_MFMA && frag_of(p) == 16 &&
        (base_prec(kp) == TCS_PREC_FP16 || base_prec(kp) == TCS_PREC_BF16) &&
        p->tile_h == 4 && p->tile_w == 8 && L.nbr % 8 == 0 && L.nbc % 8 == 0 &&
        (int64_t)(L.nbr / 8) * (L.nbc / 8) >= 2048)
        q.tile_h = 8;
    return q;
}

static const BatchEntry *select_batch(int prec, int th, int tw, int frag, bool fused, bool persist) {
    for (const PartTables *t : kParts)
        for (int i = 0; i < t->nbatch; ++i) {
            const BatchEntry &k = t->batch[i];
            if (k.prec == prec && k.th == th && k.tw == tw && k.frag == frag && k.fused == fused &&
                k.persist == persist)
                return &k;
        }
    return nullptr;
}

const char *batch_kernel_name(int precision, int tile_h, int tile_w, int frag, int nsweeps, bool persist) {
    const BatchEntry *k = select_batch(precision, tile_h, tile_w, frag, nsweeps > 1, persist);
    return k ? k->name : nullptr;
}

int launch_batch(int precision, int tile_h, int tile_w, int frag, const BatchProblem *d_probs,
                 const BatchTile *d_tiles, int ntiles, int parity, int nsweeps, int grid_per_xcd,
                 hipStream_t stream) {
    const BatchEntry *k = select_batch(precision, tile_h, tile_w, frag, nsweeps > 1, grid_per_xcd > 0);
    if (!k) return set_error(TCS_ERR_UNSUPPORTED, "no batched kernel for precision %d tile %dx%d fragment %d",
                             precision, tile_h, tile_w, frag);
    if (ntiles <= 0) return TCS_OK;
    const int tpx = (ntiles + kXCD - 1) / kXCD;
    const int per_xcd = grid_per_xcd > 0 && grid_per_xcd < tpx ? grid_per_xcd : tpx;
    hipLaunchKernelGGL(k->fn, dim3((unsigned)(per_xcd * kXCD)), dim3(kWG), 0, stream, d_probs, d_tiles, ntiles, tpx,
                       parity, nsweeps);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

const char *fused_kernel_name(const tcs_problem *p) {
    const tcs_problem q = fused_problem(p);
    const FusedEntry *k = select_fused(&q, p->fuse > 1 ? p->fuse : 1);
    return k ? k->name : nullptr;
}

const char *sweep_kernel_name(const tcs_problem *p) {
    const KernelEntry *k = select_kernel(p);
    return k ? k->name : nullptr;
}

int mfma_per_brick(const tcs_problem *p) {
    if (p->method != TCS_METHOD_MFMA) return 0;
    if (frag_of(p) == 32) return p->radius == 2 ? Fold32<0, 4>::NMF : Fold32<1, 3>::NMF;  // per 2x2-brick block
    // Radius 2 packs its K slots into 4 MFMAs per brick (3 for radius 1);
    // SPLIT issues 3 per fragment (hi*hi, hi*lo, lo*hi), hi+lo coefficients 2.
    const int nf = p->radius == 2 ? KMap<0, 4>::NF : KMap<1, 3>::NF;
    const int kp = kernel_prec(p);
    return base_prec(kp) == TCS_PREC_SPLIT ? 3 * nf : (has_coef_lo(kp) ? 2 * nf : nf);
}

int64_t mfma_flops_per_brick(const tcs_problem *p) {
    if (p->method != TCS_METHOD_MFMA) return 0;
    if (frag_of(p) == 32) return (int64_t)mfma_per_brick(p) * (32 * 32 * 16 * 2) / 4;
    return (int64_t)mfma_per_brick(p) * (16 * 16 * 32 * 2);
}

static int make_args(const tcs_problem *p, const void *d_in, void *d_out, SweepArgs &a) {
    const tcs_layout &L = p->layout;
    if (L.nbr % p->tile_h != 0 || L.nbc % p->tile_w != 0)
        return set_error(TCS_ERR_INVALID, "layout %dx%d bricks is not a multiple of tile %dx%d", L.nbr,
                         L.nbc, p->tile_h, p->tile_w);
    a.in = static_cast<const uint16_t *>(d_in);
    a.out = static_cast<uint16_t *>(d_out);
    a.plane = L.plane_elems;
    a.rows = L.rows;
    a.cols = L.cols;
    a.sbr = L.sbr;
    a.sbc = L.sbc;
    a.tiles_x = L.nbc / p->tile_w;
    a.tiles_xp = (a.tiles_x + kXCD - 1) / kXCD * kXCD;
    a.tiles_y = L.nbr / p->tile_h;
    a.nvirt = a.tiles_y * a.tiles_xp;
    a.rs = p->radius;
    a.out_scale = ldexpf(1.0f, -p->coeff_exp);
    for (int i = 0; i < 25; ++i) a.cf[i] = p->coeff[i];
    build_coef_windows(p, &a.cw);
    a.rm_src = nullptr;
    a.rm_dst = nullptr;
    a.rm_ld = 0;
    a.rm_scale_in = a.rm_scale_out = 1.0f;
    a.rm_vec = 0;
    return TCS_OK;
}

int launch_fused(const tcs_problem *p_in, const void *d_in, void *d_out, int nsweeps, hipStream_t stream) {
    const tcs_problem q = fused_problem(p_in);
    const tcs_problem *p = &q;
    const FusedEntry *k = select_fused(p, nsweeps);
    if (!k)
        return set_error(TCS_ERR_UNSUPPORTED, "no fused kernel for method %d precision %d tile %dx%d", p->method,
                         kernel_prec(p), p->tile_h, p->tile_w);
    if (nsweeps < 1 || p->radius * (nsweeps - 1) > 16)
        return set_error(TCS_ERR_INVALID, "fused launch of %d sweeps: radius %d allows 1..%d", nsweeps, p->radius,
                         1 + 16 / p->radius);
    SweepArgs a;
    int st = make_args(p, d_in, d_out, a);
    if (st) return st;
    hipLaunchKernelGGL(k->fn, dim3((unsigned)a.nvirt), dim3((unsigned)k->nt), 0, stream, a, nsweeps);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

static const EdgeEntry *select_edge(const tcs_problem *p, int mode) {
    if (p->method != TCS_METHOD_MFMA || frag_of(p) != 16 || p->wg_per_cu > 0) return nullptr;
    const int kp = kernel_prec(p);
    for (const PartTables *t : kParts)
        for (int i = 0; i < t->nedge; ++i) {
            const EdgeEntry &k = t->edge[i];
            if (k.prec == kp && k.radius == p->radius && k.th == p->tile_h && k.tw == p->tile_w && k.mode == mode)
                return &k;
        }
    return nullptr;
}

// The problem the edge kernels run: the caller's on one tile per workgroup,
// except that larger tiles run as the measured-fastest edge tile, which
// divides the layout (brick results depend on neither).  SPLIT: 4x4 for tiles of 32 or more
// bricks -- holding both planes' results for the staged output costs
// registers (box2d2r 16384^2, one sweep: 4x8 0.62 ms, 2x8 0.61, 4x4 0.57, 2x4
// 0.61, 2x16 0.62, 1x16 0.62).  FP16/BF16: 2x16 where it fits (one sweep 0.38 ms
// against 0.40 for 4x8, 0.40 for 1x16 and 0.44 for 8x8): the row-major FP32 side
// streams 1 KB row segments instead of 512 bytes (the same access pattern as a
// plain copy: 387 us for 4x8-brick tiles, 362 for 2x16, 322 linear).
static tcs_problem edge_problem(const tcs_problem *p) {
    tcs_problem q = *p;
    q.wg_per_cu = 0;  // one tile per workgroup: persistence does not change results
    const bool split = p->precision == TCS_PREC_SPLIT;
    const int tw = split ? 4 : 8;
    if (p->tile_h * p->tile_w >= (split ? 32 : 64) && p->layout.nbr % 4 == 0 && p->layout.nbc % tw == 0) {
        q.tile_h = 4;
        q.tile_w = tw;
    }
    // FP16/BF16 edges on 2x16 tiles (32 rows of 1 KB FP32 row segments) where the layout
    // divides into them and they still make 4096 tiles or more.
    const int64_t wide = (int64_t)(p->layout.nbr / 2) * (p->layout.nbc / 16);
    if (!split && p->layout.nbr % 2 == 0 && p->layout.nbc % 16 == 0 && wide >= 4096) {
        q.tile_h = 2;
        q.tile_w = 16;
    }
    return q;
}

bool edge_supported(const tcs_problem *p) {
    const tcs_problem q = edge_problem(p);
    return select_edge(&q, 3) != nullptr;
}

const char *edge_kernel_name(const tcs_problem *p, int mode) {
    const tcs_problem q = edge_problem(p);
    const EdgeEntry *k = select_edge(&q, mode);
    return k ? k->name : nullptr;
}

int launch_edge(const tcs_problem *pp, int mode, const void *d_in, void *d_out, const float *src, float *dst,
                int64_t ld, int scale_exp, hipStream_t stream) {
    const tcs_problem q = edge_problem(pp);
    const tcs_problem *p = &q;
    const EdgeEntry *k = select_edge(p, mode);
    if (!k) return set_error(TCS_ERR_UNSUPPORTED, "no row-major edge kernel for this problem");
    SweepArgs a;
    int st = make_args(p, d_in, d_out, a);
    if (st) return st;
    a.rm_src = src;
    a.rm_dst = dst;
    a.rm_ld = ld;
    a.rm_scale_in = ldexpf(1.0f, scale_exp);
    a.rm_scale_out = ldexpf(1.0f, -scale_exp);
    a.rm_vec = ld % 4 == 0 && (reinterpret_cast<uintptr_t>(src) & 15) == 0 &&
               (reinterpret_cast<uintptr_t>(dst) & 15) == 0;
    if (mode & 1) a.nvirt = (a.tiles_y + kXCD - 1) / kXCD * kXCD * a.tiles_x;  // row-band order
    hipLaunchKernelGGL(k->fn, dim3((unsigned)a.nvirt), dim3(kWG), 0, stream, a);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

int launch_sweep(const tcs_problem *p, const void *d_in, void *d_out, hipStream_t stream) {
    const KernelEntry *k = select_kernel(p);
    if (!k)
        return set_error(TCS_ERR_UNSUPPORTED, "no kernel for stencil %d method %d precision %d tile %dx%d",
                         p->stencil, p->method, kernel_prec(p), p->tile_h, p->tile_w);
    SweepArgs a;
    int st = make_args(p, d_in, d_out, a);
    if (st) return st;
    // One workgroup per virtual tile, or (wg_per_cu > 0) a persistent grid of
    // wg_per_cu workgroups per CU; either way a multiple of 8.
    unsigned grid = (unsigned)a.nvirt;
    if (k->persist) {
        int dev = 0, cus = 0;
        TCS_HIP_CHECK(hipGetDevice(&dev));
        TCS_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        int per_xcd = cus * p->wg_per_cu / kXCD;
        if (per_xcd < 1) per_xcd = 1;
        if (per_xcd * kXCD < a.nvirt) grid = (unsigned)(per_xcd * kXCD);
    }
    hipLaunchKernelGGL(k->fn, dim3(grid), dim3(kWG), 0, stream, a);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}
#endif  // TCS_PART

}  // namespace tcs
