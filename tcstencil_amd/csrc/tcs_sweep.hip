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
//           20x20 haloed input (X[0][0] = brick origin - (2,2)),
//               out^T = sum_z  C_z * S_z,   C_z[j][k] = c[z][k-j] (0 <= k-j <= 4),
//                                           S_z[k][i] = X[i+z][k]
//           is one v_mfma_f32_16x16x32_{f16,bf16} per row shift z: M = output
//           column j, N = output row i, K = haloed column k (20 of 32 slots
//           used, the rest multiply zero coefficients; slots 24..31 read no LDS).  The banded Toeplitz
//           C_z is the reference's init_coe matrix (layout16/host_box2d2r.cu:29-50)
//           transposed and widened to cover the halo columns, so no scalar
//           halo fix-ups are needed.  C_z lives in registers for the whole
//           kernel; S_z is one ds_read_b128 per lane.  The FP32 accumulator of
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

// ------------------------------------------------------------ tile staging
template <int TH, int TW, int NPL>
struct Tile {
    static constexpr int LR = 16 * TH + 4;      // LDS rows (2 halo rows each side)
    static constexpr int LP = 16 * TW + 8;      // LDS pitch in elements; 2*LP = 16 (mod 32) bytes
    static constexpr int NCH = TH * TW * 32;    // 16-byte chunks of the data bricks
    static constexpr int CPT = NCH / kWG;       // chunks per thread
    static constexpr int HD = 8 * TW + 2;       // dwords in one top/bottom halo row
    static constexpr int NH = 4 * HD + 32 * TH; // halo dwords: 2 top + 2 bottom rows, left, right
    static constexpr int HPT = (NH + kWG - 1) / kWG;
    static constexpr int PLANE = LR * LP;       // LDS elements per plane
    static_assert(NCH % kWG == 0, "tile must hold a multiple of 8 bricks");
};

// One tile's global loads, held in registers between issue and the LDS write
// so that the next tile's loads are in flight while the current one computes.
template <int TH, int TW, int NPL>
struct TileRegs {
    using T = Tile<TH, TW, NPL>;
    u32x4 cv[NPL][T::CPT];
    uint32_t hv[NPL][T::HPT];
};

// Halo frame dword h -> (LDS row, LDS dword column): two rows above, two rows
// below (HD dwords each), then the left and right 2-column strips.
template <int TH, int TW>
__device__ __forceinline__ void halo_coords(int h, int &lr, int &ld) {
    constexpr int HD = 8 * TW + 2;
    if (h < 2 * HD) {
        lr = h / HD;
        ld = h - lr * HD;
    } else if (h < 4 * HD) {
        const int q = h - 2 * HD;
        const int rr = q / HD;
        lr = 16 * TH + 2 + rr;
        ld = q - rr * HD;
    } else if (h < 4 * HD + 16 * TH) {
        lr = 2 + (h - 4 * HD);
        ld = 0;
    } else {
        lr = 2 + (h - 4 * HD - 16 * TH);
        ld = HD - 1;
    }
}

// LDS element (lr, lc) of plane p holds global point (gr0 - 2 + lr, gc0 - 2 + lc)
// for lr < LR, lc < 16*TW + 4; columns [16*TW+4, LP) hold zeros.
template <int TH, int TW, int NPL>
__device__ __forceinline__ void load_tile(const SweepArgs &a, TileRegs<TH, TW, NPL> &R, int gr0, int gc0,
                                          int tid) {
    using T = Tile<TH, TW, NPL>;
#pragma unroll
    for (int i = 0; i < T::CPT; ++i) {
        const int c = tid + i * kWG;
        const int b = c >> 5, w = c & 31;
        const int br = b / TW, bc = b - br * TW;
        const int64_t s = stor_index(gr0 + 16 * br + (w >> 1), gc0 + 16 * bc + 8 * (w & 1), a.sbc);
#pragma unroll
        for (int p = 0; p < NPL; ++p)
            R.cv[p][i] = *reinterpret_cast<const u32x4 *>(a.in + p * a.plane + s);
    }
#pragma unroll
    for (int i = 0; i < T::HPT; ++i) {
        const int h = tid + i * kWG;
        if (h < T::NH) {
            int lr, ld;
            halo_coords<TH, TW>(h, lr, ld);
            const int64_t s = stor_index(gr0 - 2 + lr, gc0 - 2 + 2 * ld, a.sbc);
#pragma unroll
            for (int p = 0; p < NPL; ++p)
                R.hv[p][i] = *reinterpret_cast<const uint32_t *>(a.in + p * a.plane + s);
        }
    }
}

template <int TH, int TW, int NPL>
__device__ __forceinline__ void write_tile(const TileRegs<TH, TW, NPL> &R, uint16_t *lds, int tid) {
    using T = Tile<TH, TW, NPL>;
    uint32_t *l32 = reinterpret_cast<uint32_t *>(lds);
#pragma unroll
    for (int i = 0; i < T::CPT; ++i) {
        const int c = tid + i * kWG;
        const int b = c >> 5, w = c & 31;
        const int br = b / TW, bc = b - br * TW;
        const int lr = 16 * br + (w >> 1) + 2;
        const int lc = 16 * bc + 8 * (w & 1) + 2;
        const int d = (lr * T::LP + lc) >> 1;  // lc even: dword aligned, not 16-byte aligned
#pragma unroll
        for (int p = 0; p < NPL; ++p) {
#pragma unroll
            for (int q = 0; q < 4; ++q) l32[p * (T::PLANE / 2) + d + q] = R.cv[p][i][q];
        }
    }
#pragma unroll
    for (int i = 0; i < T::HPT; ++i) {
        const int h = tid + i * kWG;
        if (h < T::NH) {
            int lr, ld;
            halo_coords<TH, TW>(h, lr, ld);
            const int d = ((lr * T::LP) >> 1) + ld;
#pragma unroll
            for (int p = 0; p < NPL; ++p) l32[p * (T::PLANE / 2) + d] = R.hv[p][i];
        }
    }
}

// Columns [16*TW+4, 16*TW+8) of every LDS row: read by the MFMA operand of the
// last brick column (against zero coefficients), never written by write_tile.
template <int TH, int TW, int NPL>
__device__ __forceinline__ void zero_pad(uint16_t *lds, int tid) {
    using T = Tile<TH, TW, NPL>;
    uint32_t *l32 = reinterpret_cast<uint32_t *>(lds);
    for (int zc = tid; zc < 2 * T::LR; zc += kWG) {
        const int d = (((zc >> 1) * T::LP) >> 1) + 8 * TW + 2 + (zc & 1);
#pragma unroll
        for (int p = 0; p < NPL; ++p) l32[p * (T::PLANE / 2) + d] = 0u;
    }
}

__device__ __forceinline__ bool tile_is_interior(const SweepArgs &a, int gr0, int gc0, int h, int w) {
    return gr0 >= a.rs && gr0 + h <= a.rows - a.rs && gc0 >= a.rs && gc0 + w <= a.cols - a.rs;
}

__device__ __forceinline__ bool point_is_interior(const SweepArgs &a, int gr, int gc) {
    return gr >= a.rs && gr < a.rows - a.rs && gc >= a.rs && gc < a.cols - a.rs;
}

// Persistent tile schedule.  Workgroup w runs on XCD w % 8; XCD x owns the
// contiguous tile range [x*tpx, (x+1)*tpx), and its gridDim.x/8 workgroups step
// through that range together, so the tiles in flight on one XCD are
// neighbours and their halos meet in that XCD's L2.
struct Schedule {
    int tile, end, step;
};

__device__ __forceinline__ Schedule schedule(const SweepArgs &a) {
    const int x = blockIdx.x % kXCD, k = blockIdx.x / kXCD;
    const int begin = x * a.tiles_per_xcd;
    const int end = begin + a.tiles_per_xcd < a.ntiles ? begin + a.tiles_per_xcd : a.ntiles;
    return Schedule{begin + k, end, (int)gridDim.x / kXCD};
}

template <int TH, int TW>
__device__ __forceinline__ void tile_origin(const SweepArgs &a, int tile, int &gr0, int &gc0) {
    const int ty = tile / a.tiles_x, tx = tile - ty * a.tiles_x;
    gr0 = 16 * TH * ty;
    gc0 = 16 * TW * tx;
}

__device__ __forceinline__ uint16_t f16_bits(float v) {
    return __builtin_bit_cast(uint16_t, static_cast<_Float16>(v));
}
__device__ __forceinline__ uint16_t bf16_bits(float v) {
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(v));
}
__device__ __forceinline__ uint32_t pack2(uint16_t lo, uint16_t hi) {
    return (uint32_t)lo | ((uint32_t)hi << 16);
}

// ------------------------------------------------------------- MFMA kernel
// Operand reuse across row shifts through DPP cuts LDS reads ~4x; measured
// time-neutral on 16384^2 sweeps, so it is a build option, off by default.
#ifndef TCS_DPP_SHIFT
#define TCS_DPP_SHIFT 0
#endif

// Lane n of each 16-lane row takes lane n+1's value (DPP row_shl:1); lane 15
// of each row receives 0.
__device__ __forceinline__ u32x4 rows_up(u32x4 v) {
    u32x4 r;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        r[q] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[q], 0x101, 0xf, 0xf, true);
    return r;
}

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

template <int TH, int TW, int ZLO, int ZHI, int PREC, bool PERSIST>
__global__ __launch_bounds__(kWG) void tcs_mfma_sweep(SweepArgs a) {
    constexpr int NPL = PREC == TCS_PREC_SPLIT ? 2 : 1;
    constexpr int NZ = ZHI - ZLO + 1;
    using T = Tile<TH, TW, NPL>;
    using F = Frag<PREC>;
    typedef typename F::type frag_t;
    __shared__ __attribute__((aligned(16))) uint16_t lds[NPL * T::PLANE];

    Schedule sc = schedule(a);
    if (sc.tile >= sc.end) return;
    const int tid = threadIdx.x;
    int gr0, gc0;
    tile_origin<TH, TW>(a, sc.tile, gr0, gc0);
    TileRegs<TH, TW, NPL> R;
    load_tile<TH, TW, NPL>(a, R, gr0, gc0, tid);
    zero_pad<TH, TW, NPL>(lds, tid);

    // Coefficient fragments.  Lane l holds A[m][k] for m = l%16 and
    // k = 8*(l/16) + e, e = 0..7: the coefficient at row offset z and column
    // offset k - m when 0 <= k - m <= 4, else 0.
    const int lane = tid & 63, wave = tid >> 6;
    const int m = lane & 15, g = lane >> 4;
    frag_t ahi[NZ], alo[NZ];
#pragma unroll
    for (int zi = 0; zi < NZ; ++zi) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int d = 8 * g + e - m;
            if constexpr (PREC == TCS_PREC_SPLIT) {
                uint16_t hb = 0, lb = 0;
#pragma unroll
                for (int dd = 0; dd < 5; ++dd)
                    if (d == dd) {
                        hb = a.ch[(ZLO + zi) * 5 + dd];
                        lb = a.cl[(ZLO + zi) * 5 + dd];
                    }
                ahi[zi][e] = __builtin_bit_cast(_Float16, hb);
                alo[zi][e] = __builtin_bit_cast(_Float16, lb);
            } else {
                float v = 0.0f;
#pragma unroll
                for (int dd = 0; dd < 5; ++dd)
                    if (d == dd) v = a.cf[(ZLO + zi) * 5 + dd];
                if constexpr (PREC == TCS_PREC_BF16)
                    ahi[zi][e] = static_cast<__bf16>(v);
                else
                    ahi[zi][e] = static_cast<_Float16>(v);
            }
        }
    }
    // Slot group 3 (k = 24..31) only meets zero coefficients: those lanes skip
    // the LDS read and feed zeros.
    const bool reads = g < 3;
    constexpr int NB = TH * TW / 4;

    for (;;) {
    write_tile<TH, TW, NPL>(R, lds, tid);
    __syncthreads();
    const int cgr0 = gr0, cgc0 = gc0;  // the tile in LDS
    const int next = sc.tile + sc.step;
    if (PERSIST && next < sc.end) {
        tile_origin<TH, TW>(a, next, gr0, gc0);
        load_tile<TH, TW, NPL>(a, R, gr0, gc0, tid);  // in flight during the compute below
    }
    const bool edge = !tile_is_interior(a, cgr0, cgc0, 16 * TH, 16 * TW);
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int b = wave + 4 * q;
        const int br = b / TW, bc = b - br * TW;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        // Operand of row shift z: lane (m, g) holds haloed row m + z, columns
        // 8g..8g+7.  With TCS_DPP_SHIFT, shift z+1 is shift z moved up one lane
        // within each 16-lane row (DPP row_shl:1) and only lane 15 of each row
        // reads a new LDS row.
        u32x4 rh = {0u, 0u, 0u, 0u}, rl = {0u, 0u, 0u, 0u};
#pragma unroll
        for (int zi = 0; zi < NZ; ++zi) {
            const int off = (16 * br + m + ZLO + zi) * T::LP + 16 * bc + 8 * g;
            const bool fetch = reads && (TCS_DPP_SHIFT == 0 || zi == 0 || m == 15);
            if (TCS_DPP_SHIFT && zi > 0) {
                rh = rows_up(rh);
                if constexpr (PREC == TCS_PREC_SPLIT) rl = rows_up(rl);
            }
            if (fetch) {
                rh = *reinterpret_cast<const u32x4 *>(lds + off);
                if constexpr (PREC == TCS_PREC_SPLIT) rl = *reinterpret_cast<const u32x4 *>(lds + T::PLANE + off);
            }
            const frag_t bh = __builtin_bit_cast(frag_t, rh);
            acc = F::mma(ahi[zi], bh, acc);
            if constexpr (PREC == TCS_PREC_SPLIT) {
                const frag_t bl = __builtin_bit_cast(frag_t, rl);
                acc = F::mma(ahi[zi], bl, acc);
                acc = F::mma(alo[zi], bh, acc);
                acc = F::mma(alo[zi], bl, acc);
            }
        }
        // Lane holds output row m, columns 4g..4g+3 of brick (br, bc).
        const int gr = cgr0 + 16 * br + m;
        const int gcb = cgc0 + 16 * bc + 4 * g;
        uint16_t o[NPL][4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            if constexpr (PREC == TCS_PREC_SPLIT) {
                const float s = acc[v] * a.out_scale;
                const _Float16 hi = static_cast<_Float16>(s);
                o[0][v] = __builtin_bit_cast(uint16_t, hi);
                o[1][v] = f16_bits(s - static_cast<float>(hi));
            } else if constexpr (PREC == TCS_PREC_BF16) {
                o[0][v] = bf16_bits(acc[v]);
            } else {
                o[0][v] = f16_bits(acc[v]);
            }
        }
        if (edge) {
            const int cidx = (16 * br + m + 2) * T::LP + 16 * bc + 4 * g + 2;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (!point_is_interior(a, gr, gcb + v)) {
#pragma unroll
                    for (int p = 0; p < NPL; ++p) o[p][v] = lds[p * T::PLANE + cidx + v];
                }
        }
        const int64_t s = stor_index(gr, gcb, a.sbc);
#pragma unroll
        for (int p = 0; p < NPL; ++p) {
            u32x2 w = {pack2(o[p][0], o[p][1]), pack2(o[p][2], o[p][3])};
            store8<true>(a.out + p * a.plane + s, w);
        }
    }
    // One-tile-per-workgroup launches cover every tile with the grid itself.
    if (!PERSIST || next >= sc.end) break;
    sc.tile = next;
    __syncthreads();  // every wave is done reading LDS before it is overwritten
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

template <int TH, int TW, int STENCIL, bool PERSIST>
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
    TileRegs<TH, TW, 1> R;
    load_tile<TH, TW, 1>(a, R, gr0, gc0, tid);
    zero_pad<TH, TW, 1>(lds, tid);

    const int cg = tid % CG, rg = tid / CG;
    const int r0 = rg * MR;  // first output row of this strip (tile coordinates)

    for (;;) {
    write_tile<TH, TW, 1>(R, lds, tid);
    __syncthreads();
    const int cgr0 = gr0, cgc0 = gc0;  // the tile in LDS
    const int next = sc.tile + sc.step;
    if (PERSIST && next < sc.end) {
        tile_origin<TH, TW>(a, next, gr0, gc0);
        load_tile<TH, TW, 1>(a, R, gr0, gc0, tid);  // in flight during the compute below
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
        const uint16_t *row = lds + (r0 + ir) * T::LP + 4 * cg;
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
#pragma unroll
            for (int dj = 0; dj < 5; ++dj) {
                if (!tap<STENCIL>(dz, dj)) continue;
                const float c = a.cf[dz * 5 + dj];
#pragma unroll
                for (int v = 0; v < 4; ++v) acc[o][v] = __builtin_fmaf(x[v + dj], c, acc[o][v]);
            }
        }
    }

    const bool edge = !tile_is_interior(a, cgr0, cgc0, 16 * TH, 16 * TW);
#pragma unroll
    for (int o = 0; o < MR; ++o) {
        const int tr = r0 + o;                  // tile row
        const int gr = cgr0 + tr;
        const int gcb = cgc0 + 4 * cg;
        uint16_t h[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) h[v] = f16_bits(acc[o][v]);
        if (edge) {
            const int cidx = (tr + 2) * T::LP + 4 * cg + 2;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (!point_is_interior(a, gr, gcb + v)) h[v] = lds[cidx + v];
        }
        u32x2 w = {pack2(h[0], h[1]), pack2(h[2], h[3])};
        store8<false>(a.out + stor_index(gr, gcb, a.sbc), w);
    }
    // One-tile-per-workgroup launches cover every tile with the grid itself.
    if (!PERSIST || next >= sc.end) break;
    sc.tile = next;
    __syncthreads();  // every wave is done reading LDS before it is overwritten
    }
}

// ---------------------------------------------------------------- dispatch
typedef void (*sweep_fn)(SweepArgs);

struct KernelEntry {
    int method, prec, radius, stencil, th, tw;  // stencil = -1: any stencil of that radius
    bool persist;
    sweep_fn fn;
    const char *name;  // as rocprofv3 reports it (demangled, without the namespace)
};

#define TCS_MFMA(TH, TW, ZLO, ZHI, P, R, PS) \
    {TCS_METHOD_MFMA, P, R, -1, TH, TW, PS, tcs_mfma_sweep<TH, TW, ZLO, ZHI, P, PS>, \
     "tcs_mfma_sweep<" #TH ", " #TW ", " #ZLO ", " #ZHI ", " #P ", " #PS ">"}
#define TCS_MFMA_TILES(ZLO, ZHI, P, R, PS) \
    TCS_MFMA(2, 4, ZLO, ZHI, P, R, PS), TCS_MFMA(4, 4, ZLO, ZHI, P, R, PS), TCS_MFMA(2, 8, ZLO, ZHI, P, R, PS), \
    TCS_MFMA(4, 8, ZLO, ZHI, P, R, PS)
#define TCS_MFMA_ALL(PS)                                                                  \
    TCS_MFMA_TILES(0, 4, 0, 2, PS), TCS_MFMA(8, 8, 0, 4, 0, 2, PS), TCS_MFMA_TILES(1, 3, 0, 1, PS), \
    TCS_MFMA(8, 8, 1, 3, 0, 1, PS), TCS_MFMA_TILES(0, 4, 1, 2, PS), TCS_MFMA(8, 8, 0, 4, 1, 2, PS), \
    TCS_MFMA_TILES(1, 3, 1, 1, PS), TCS_MFMA(8, 8, 1, 3, 1, 1, PS), TCS_MFMA_TILES(0, 4, 2, 2, PS), \
    TCS_MFMA_TILES(1, 3, 2, 1, PS)
#define TCS_DIRECT(TH, TW, S, PS) \
    {TCS_METHOD_DIRECT, TCS_PREC_FP16, -1, S, TH, TW, PS, tcs_direct_sweep<TH, TW, S, PS>, \
     "tcs_direct_sweep<" #TH ", " #TW ", " #S ", " #PS ">"}
#define TCS_DIRECT_TILES(S, PS) \
    TCS_DIRECT(2, 4, S, PS), TCS_DIRECT(4, 4, S, PS), TCS_DIRECT(2, 8, S, PS), TCS_DIRECT(4, 8, S, PS), \
    TCS_DIRECT(8, 8, S, PS)

// Tile shapes offered (bricks, TH x TW), each one-tile-per-workgroup (false)
// and persistent (true).  SPLIT has no 8x8 tile: two planes of 132 x 136
// elements would need 71.8 KB of LDS.
static const KernelEntry kKernels[] = {
    TCS_MFMA_ALL(false), TCS_MFMA_ALL(true),
    TCS_DIRECT_TILES(0, false), TCS_DIRECT_TILES(1, false), TCS_DIRECT_TILES(2, false), TCS_DIRECT_TILES(3, false),
    TCS_DIRECT_TILES(0, true), TCS_DIRECT_TILES(1, true), TCS_DIRECT_TILES(2, true), TCS_DIRECT_TILES(3, true),
};
static_assert(TCS_PREC_FP16 == 0 && TCS_PREC_BF16 == 1 && TCS_PREC_SPLIT == 2, "precision ids");
static_assert(TCS_STAR2D1R == 0 && TCS_STAR2D2R == 1 && TCS_BOX2D1R == 2 && TCS_BOX2D2R == 3, "stencil ids");

static const KernelEntry *select_kernel(const tcs_problem *p) {
    for (const KernelEntry &k : kKernels) {
        if (k.method != p->method || k.prec != p->precision || k.th != p->tile_h || k.tw != p->tile_w ||
            k.persist != (p->wg_per_cu > 0))
            continue;
        if (k.method == TCS_METHOD_MFMA ? k.radius == p->radius : k.stencil == p->stencil) return &k;
    }
    return nullptr;
}

const char *sweep_kernel_name(const tcs_problem *p) {
    const KernelEntry *k = select_kernel(p);
    return k ? k->name : nullptr;
}

int mfma_per_brick(const tcs_problem *p) {
    if (p->method != TCS_METHOD_MFMA) return 0;
    const int nz = 2 * p->radius + 1;
    return p->precision == TCS_PREC_SPLIT ? 4 * nz : nz;
}

int launch_sweep(const tcs_problem *p, const void *d_in, void *d_out, hipStream_t stream) {
    const KernelEntry *k = select_kernel(p);
    if (!k)
        return set_error(TCS_ERR_UNSUPPORTED, "no kernel for stencil %d method %d precision %d tile %dx%d",
                         p->stencil, p->method, p->precision, p->tile_h, p->tile_w);
    const tcs_layout &L = p->layout;
    if (L.nbr % p->tile_h != 0 || L.nbc % p->tile_w != 0)
        return set_error(TCS_ERR_INVALID, "layout %dx%d bricks is not a multiple of tile %dx%d", L.nbr,
                         L.nbc, p->tile_h, p->tile_w);
    SweepArgs a;
    a.in = static_cast<const uint16_t *>(d_in);
    a.out = static_cast<uint16_t *>(d_out);
    a.plane = L.plane_elems;
    a.rows = L.rows;
    a.cols = L.cols;
    a.sbc = L.sbc;
    a.tiles_x = L.nbc / p->tile_w;
    a.ntiles = (L.nbr / p->tile_h) * a.tiles_x;
    a.tiles_per_xcd = (a.ntiles + kXCD - 1) / kXCD;
    a.rs = p->radius;
    a.out_scale = ldexpf(1.0f, -p->coeff_exp);
    for (int i = 0; i < 25; ++i) {
        a.cf[i] = p->coeff[i];
        a.ch[i] = p->coeff_hi[i];
        a.cl[i] = p->coeff_lo[i];
    }
    // One workgroup per tile, or (wg_per_cu > 0) a persistent grid of wg_per_cu
    // workgroups per CU; either way a multiple of 8 so every XCD gets the same.
    int per_xcd = a.tiles_per_xcd;
    if (k->persist) {
        int dev = 0, cus = 0;
        TCS_HIP_CHECK(hipGetDevice(&dev));
        TCS_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        per_xcd = cus * p->wg_per_cu / kXCD;
        if (per_xcd < 1) per_xcd = 1;
        if (per_xcd > a.tiles_per_xcd) per_xcd = a.tiles_per_xcd;
    }
    const unsigned grid = (unsigned)(per_xcd * kXCD);
    hipLaunchKernelGGL(k->fn, dim3(grid), dim3(kWG), 0, stream, a);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

}  // namespace tcs
