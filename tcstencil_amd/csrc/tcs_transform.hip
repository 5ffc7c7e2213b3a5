// Brick-layout transform and data utilities on the GPU:
//   pack    row-major FP32/FP64 -> brick storage (FP16, BF16 or hi+lo FP16 pair),
//           ghost ring and padding zeroed (replaces the host-side array_float2half,
//           padding_2D and brick_layout_2D of layout16/host_box2d2r.cu:61-86, :131-141)
//   unpack  brick storage -> row-major FP32 (replaces reverse_layout_2D and the
//           interior copy, layout16/host_box2d2r.cu:88-99, :196-204)
//   clear, max-abs, rounding to FP16/BF16 values, and the seeded generators.
#include <math.h>

#include "tcs_internal.hpp"

namespace tcs {

// --------------------------------------------------------------- conversion
struct Enc2 {
    uint16_t hi, lo;
};

// The FP16 value of x is fp16(fp32(x)) for FP64 sources (two roundings,
// defined that way on the CPU and the GPU alike); lo = fp16(x - hi) keeps the
// remainder for the SPLIT pair (the subtraction is exact in the source type).
template <typename T>
__device__ __forceinline__ Enc2 encode(T x, int prec) {
    Enc2 e;
    if (prec == TCS_PREC_BF16) {
        e.hi = __builtin_bit_cast(uint16_t, static_cast<__bf16>(static_cast<float>(x)));
        e.lo = 0;
    } else {
        const _Float16 h = static_cast<_Float16>(static_cast<float>(x));
        e.hi = __builtin_bit_cast(uint16_t, h);
        e.lo = __builtin_bit_cast(uint16_t, static_cast<_Float16>(static_cast<float>(x - static_cast<T>(h))));
    }
    return e;
}

__device__ __forceinline__ float decode(uint16_t hi, uint16_t lo, int prec) {
    if (prec == TCS_PREC_BF16) return static_cast<float>(__builtin_bit_cast(__bf16, hi));
    const float h = static_cast<float>(__builtin_bit_cast(_Float16, hi));
    if (prec == TCS_PREC_SPLIT) return h + static_cast<float>(__builtin_bit_cast(_Float16, lo));
    return h;
}

// ------------------------------------------------------- pack and unpack
// Both transforms give one wave a span of 16 rows x 64 columns: four bricks
// side by side in one brick row.  Lane l handles the 4-column piece
// c4 = l % 16 (brick 4*span + c4/4 of the row, brick columns 4*(c4%4)..+3)
// in the rows x = 4k + l/16, k = 0..3.  For each k the 64 lanes cover four
// complete rows of 256 bytes (FP32) and, in each of the four bricks, four
// complete brick rows (128 contiguous bytes of FP16), so every memory
// instruction of the wave moves whole 128-byte lines on both sides.
constexpr int kSpanBricks = 4;

// Pack: spans over the stored bricks (ghost ring included; points outside the
// grid are stored as zero).  FP32 pieces inside the grid are read with one
// 16-byte load when vec (16-byte aligned source and leading dimension).
template <typename T>
__global__ void tcs_pack_kernel(const T *__restrict__ src, int64_t ld, uint16_t *__restrict__ dst,
                                int64_t plane, int rows, int cols, int sbr, int sbc, int prec, T scale, int vec) {
    const int spans_per_row = (sbc + kSpanBricks - 1) / kSpanBricks;
    const int64_t total = 64 * (int64_t)sbr * spans_per_row;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t span = t >> 6;
        const int l = (int)(t & 63), c4 = l & 15;
        const int bi = (int)(span / spans_per_row);
        const int bj = kSpanBricks * (int)(span - (int64_t)bi * spans_per_row) + (c4 >> 2);
        if (bj >= sbc) continue;
        const int cb = 4 * (c4 & 3), gc = 16 * bj - 16 + cb;
        const bool col_vec = vec && gc >= 0 && gc + 4 <= cols;  // FP32 sources only (the host sets vec)
        const int64_t brick = ((int64_t)bi * sbc + bj) * 256 + cb;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = 4 * k + (l >> 4), gr = 16 * bi - 16 + x;
            const bool row_in = (unsigned)gr < (unsigned)rows;
            T v[4];
            bool loaded = false;
            if constexpr (sizeof(T) == 4) {
                if (row_in && col_vec) {
                    const f32x4 w = *reinterpret_cast<const f32x4 *>(src + gr * ld + gc);
#pragma unroll
                    for (int e = 0; e < 4; ++e) v[e] = w[e];
                    loaded = true;
                }
            }
            if (!loaded) {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    v[e] = row_in && (unsigned)(gc + e) < (unsigned)cols ? src[gr * ld + gc + e] : (T)0;
            }
            Enc2 en[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) en[e] = encode<T>(v[e] * scale, prec);
            const int64_t d = brick + 16 * x;
            __builtin_nontemporal_store(
                u32x2{(uint32_t)en[0].hi | ((uint32_t)en[1].hi << 16), (uint32_t)en[2].hi | ((uint32_t)en[3].hi << 16)},
                reinterpret_cast<u32x2 *>(dst + d));
            if (prec == TCS_PREC_SPLIT)
                __builtin_nontemporal_store(u32x2{(uint32_t)en[0].lo | ((uint32_t)en[1].lo << 16),
                                                  (uint32_t)en[2].lo | ((uint32_t)en[3].lo << 16)},
                                            reinterpret_cast<u32x2 *>(dst + plane + d));
        }
    }
}

// Unpack: spans over the data bricks; writes the points inside the grid, the
// ring points (outside [rs, rows-rs) x [rs, cols-rs)) from `ring` when given.
// Pieces inside the grid are stored with one 16-byte store when vec.
__global__ void tcs_unpack_kernel(const uint16_t *__restrict__ src, int64_t plane, float *__restrict__ dst,
                                  int64_t ld, int rows, int cols, int nbr, int nbc, int sbc, int prec, float scale,
                                  const float *__restrict__ ring, int rs, int vec) {
    const int spans_per_row = (nbc + kSpanBricks - 1) / kSpanBricks;
    const int64_t total = 64 * (int64_t)nbr * spans_per_row;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t span = t >> 6;
        const int l = (int)(t & 63), c4 = l & 15;
        const int i = (int)(span / spans_per_row);
        const int j = kSpanBricks * (int)(span - (int64_t)i * spans_per_row) + (c4 >> 2);
        const int cb = 4 * (c4 & 3), gc = 16 * j + cb;
        if (j >= nbc || gc >= cols) continue;
        const int64_t brick = ((int64_t)(i + 1) * sbc + (j + 1)) * 256 + cb;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int x = 4 * k + (l >> 4), gr = 16 * i + x;
            if (gr >= rows) continue;
            const int64_t s = brick + 16 * x;
            const u32x2 h = *reinterpret_cast<const u32x2 *>(src + s);
            u32x2 lw = {0u, 0u};
            if (prec == TCS_PREC_SPLIT) lw = *reinterpret_cast<const u32x2 *>(src + plane + s);
            float v[4];
#pragma unroll
            for (int e = 0; e < 4; ++e)
                v[e] = decode((uint16_t)(h[e >> 1] >> (16 * (e & 1))), (uint16_t)(lw[e >> 1] >> (16 * (e & 1))), prec) *
                       scale;
            const int64_t o = gr * ld + gc;
            if (ring && !(gr >= rs && gr < rows - rs && gc >= rs && gc + 4 <= cols - rs)) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int c = gc + e;
                    if (c < cols && !(gr >= rs && gr < rows - rs && c >= rs && c < cols - rs)) v[e] = ring[o + e];
                }
            }
            if (vec && gc + 4 <= cols) {
                __builtin_nontemporal_store(f32x4{v[0], v[1], v[2], v[3]}, reinterpret_cast<f32x4 *>(dst + o));
            } else {
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (gc + e < cols) dst[o + e] = v[e];
            }
        }
    }
}

// ---------------------------------------------------------------- utilities
__global__ void tcs_fill_kernel(u32x4 *__restrict__ p, int64_t n16) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n16;
         t += (int64_t)gridDim.x * blockDim.x)
        p[t] = u32x4{0u, 0u, 0u, 0u};
}

// Zero the one-brick ghost ring of a storage buffer (every plane): ring brick
// b < 2*sbc is in stored brick row 0 or sbr-1, the rest in stored brick column
// 0 or sbc-1 of rows 1..sbr-2.  One thread per 16-byte chunk (32 per brick).
__global__ void tcs_clear_ring_kernel(uint16_t *__restrict__ buf, int64_t plane, int nplanes, int sbr, int sbc) {
    const int64_t nring = 2 * (int64_t)sbc + 2 * (int64_t)(sbr - 2);
    const int64_t total = nring * 32;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = t >> 5;
        int64_t bi, bj;
        if (b < 2 * (int64_t)sbc) {
            bi = b < sbc ? 0 : sbr - 1;
            bj = b < sbc ? b : b - sbc;
        } else {
            const int64_t q = b - 2 * (int64_t)sbc;
            bi = 1 + (q < sbr - 2 ? q : q - (sbr - 2));
            bj = q < sbr - 2 ? 0 : sbc - 1;
        }
        const int64_t e = (bi * sbc + bj) * 256 + 8 * (t & 31);
        for (int p = 0; p < nplanes; ++p) *reinterpret_cast<u32x4 *>(buf + p * plane + e) = u32x4{0u, 0u, 0u, 0u};
    }
}

// Streaming copy: the measured HBM ceiling next to the 8 TB/s spec figure, and
// the known-size kernel that calibrates the FETCH_SIZE / WRITE_SIZE counters.
// One pass, four 16-byte chunks per thread (all loads before the stores),
// nontemporal loads and stores.
__global__ void tcs_copy_kernel(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, int64_t n16) {
    const int64_t base = (int64_t)blockIdx.x * blockDim.x * 4 + threadIdx.x;
    u32x4 v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t t = base + (int64_t)i * blockDim.x;
        if (t < n16) v[i] = __builtin_nontemporal_load(src + t);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int64_t t = base + (int64_t)i * blockDim.x;
        if (t < n16) __builtin_nontemporal_store(v[i], dst + t);
    }
}

__global__ void tcs_zero_u64_kernel(unsigned long long *p) { *p = 0ull; }

// max |x|: the bit pattern of a non-negative double orders like its value
// (NaN above +inf), so an integer atomicMax on the bits is exact and
// independent of the launch order.
template <typename T>
__global__ void tcs_maxabs_kernel(const T *__restrict__ src, int rows, int cols, int64_t ld,
                                  unsigned long long *out) {
    __shared__ unsigned long long red[256];
    unsigned long long best = 0ull;
    const int64_t total = (int64_t)rows * cols;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = t / cols, c = t - r * cols;
        const double v = fabs((double)src[r * ld + c]);
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        best = b > best ? b : best;
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, red[0]);
}

template <typename T>
__global__ void tcs_round_kernel(T *__restrict__ p, int64_t n, int prec) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x) {
        const float f = static_cast<float>(p[t]);
        float r;
        if (prec == TCS_PREC_BF16)
            r = static_cast<float>(static_cast<__bf16>(f));
        else
            r = static_cast<float>(static_cast<_Float16>(f));
        p[t] = static_cast<T>(r);
    }
}

// --------------------------------------------------------------- generators
__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t stream_key(uint64_t seed, uint64_t stream) {
    return mix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
}
__host__ __device__ inline double u24(uint64_t key, uint64_t idx) {
    return (double)(mix64(key + idx * 0x9E3779B97F4A7C15ull) >> 40) * (1.0 / 16777216.0);
}

template <typename T>
__global__ void tcs_gen_uniform_kernel(T *__restrict__ dst, int rows, int cols, uint64_t key, double lo,
                                       double span) {
    const int64_t total = (int64_t)rows * cols;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        // span * u is exact in double (24-bit u, span at most 53 bits wide), so a
        // fused or unfused evaluation rounds identically.
        dst[t] = static_cast<T>(lo + span * u24(key, (uint64_t)t));
    }
}

struct SinParams {
    double kx[4], ky[4], phase[4], amp[4];
};

__global__ void tcs_gen_sinusoid_kernel(double *__restrict__ dst, int rows, int cols, SinParams sp,
                                        uint64_t noise_key, double noise) {
    const int64_t total = (int64_t)rows * cols;
    const double two_pi = 6.283185307179586476925286766559;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = t / cols, j = t - i * cols;
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            v += sp.amp[q] * sin(two_pi * (sp.kx[q] * (double)i / rows + sp.ky[q] * (double)j / cols) + sp.phase[q]);
        v += -noise + 2.0 * noise * u24(noise_key, (uint64_t)t);
        dst[t] = v;
    }
}

static unsigned grid_for(int64_t n, int block = 256) {
    int64_t g = (n + block - 1) / block;
    if (g > 65536) g = 65536;
    if (g < 1) g = 1;
    return (unsigned)g;
}

}  // namespace tcs

using namespace tcs;

static int check_problem(const tcs_problem *p) {
    if (!p) return set_error(TCS_ERR_INVALID, "null problem");
    if (p->layout.rows < 1 || p->layout.cols < 1 || p->layout.plane_elems <= 0)
        return set_error(TCS_ERR_INVALID, "problem has no layout");
    return TCS_OK;
}

template <typename T>
static int pack_impl(const tcs_problem *p, const T *d_src, int64_t ld, void *d_dst, int scale_exp,
                     void *stream) {
    int st = check_problem(p);
    if (st) return st;
    if (!d_src || !d_dst) return set_error(TCS_ERR_INVALID, "pack: null pointer");
    if (ld < p->layout.cols) return set_error(TCS_ERR_INVALID, "pack: ld %lld < cols %d", (long long)ld, p->layout.cols);
    const tcs_layout &L = p->layout;
    const int64_t total = 64 * (int64_t)L.sbr * ((L.sbc + kSpanBricks - 1) / kSpanBricks);
    const T scale = (T)ldexp(1.0, scale_exp);
    const int vec = sizeof(T) == 4 && ld % 4 == 0 && (reinterpret_cast<uintptr_t>(d_src) & 15) == 0;
    hipLaunchKernelGGL(tcs_pack_kernel<T>, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream, d_src, ld,
                       (uint16_t *)d_dst, L.plane_elems, L.rows, L.cols, L.sbr, L.sbc, p->precision, scale, vec);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

extern "C" int tcs_pack_f32(const tcs_problem *p, const float *d_src, int64_t ld, void *d_dst, int scale_exp,
                            void *stream) {
    return pack_impl<float>(p, d_src, ld, d_dst, scale_exp, stream);
}

extern "C" int tcs_pack_f64(const tcs_problem *p, const double *d_src, int64_t ld, void *d_dst, int scale_exp,
                            void *stream) {
    return pack_impl<double>(p, d_src, ld, d_dst, scale_exp, stream);
}

extern "C" int tcs_unpack_f32(const tcs_problem *p, const void *d_src, float *d_dst, int64_t ld, int scale_exp,
                              const float *d_ring, void *stream) {
    int st = check_problem(p);
    if (st) return st;
    if (!d_src || !d_dst) return set_error(TCS_ERR_INVALID, "unpack: null pointer");
    const tcs_layout &L = p->layout;
    if (ld < L.cols) return set_error(TCS_ERR_INVALID, "unpack: ld %lld < cols %d", (long long)ld, L.cols);
    const int64_t total = 64 * (int64_t)L.nbr * ((L.nbc + kSpanBricks - 1) / kSpanBricks);
    const int vec = ld % 4 == 0 && (reinterpret_cast<uintptr_t>(d_dst) & 15) == 0;
    hipLaunchKernelGGL(tcs_unpack_kernel, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream,
                       (const uint16_t *)d_src, L.plane_elems, d_dst, ld, L.rows, L.cols, L.nbr, L.nbc, L.sbc,
                       p->precision, ldexpf(1.0f, -scale_exp), d_ring, p->radius, vec);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

extern "C" int tcs_copy(void *d_dst, const void *d_src, int64_t bytes, void *stream) {
    if (!d_dst || !d_src || bytes < 0 || bytes % 16 != 0)
        return set_error(TCS_ERR_INVALID, "copy: null pointer or size not a multiple of 16");
    const int64_t n16 = bytes / 16;
    if (n16 == 0) return TCS_OK;
    const int64_t g = (n16 + 4 * 256 - 1) / (4 * 256);
    if (g > 0x7fffffff) return set_error(TCS_ERR_INVALID, "copy: too large");
    hipLaunchKernelGGL(tcs_copy_kernel, dim3((unsigned)g), dim3(256), 0, (hipStream_t)stream, (const u32x4 *)d_src,
                       (u32x4 *)d_dst, n16);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

extern "C" int tcs_clear(const tcs_problem *p, void *d_buf, void *stream) {
    int st = check_problem(p);
    if (st) return st;
    if (!d_buf) return set_error(TCS_ERR_INVALID, "clear: null pointer");
    const int64_t n16 = tcs_buffer_bytes(p) / 16;
    hipLaunchKernelGGL(tcs_fill_kernel, dim3(grid_for(n16)), dim3(256), 0, (hipStream_t)stream, (u32x4 *)d_buf, n16);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

int tcs::clear_ring(const tcs_problem *p, void *d_buf, hipStream_t stream) {
    const tcs_layout &L = p->layout;
    const int64_t total = (2 * (int64_t)L.sbc + 2 * (int64_t)(L.sbr - 2)) * 32;
    hipLaunchKernelGGL(tcs_clear_ring_kernel, dim3(grid_for(total)), dim3(256), 0, stream, (uint16_t *)d_buf,
                       L.plane_elems, p->precision == TCS_PREC_SPLIT ? 2 : 1, L.sbr, L.sbc);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

template <typename T>
static int maxabs_impl(const T *d_src, int rows, int cols, int64_t ld, double *d_out, void *stream) {
    if (!d_src || !d_out) return set_error(TCS_ERR_INVALID, "maxabs: null pointer");
    if (rows < 1 || cols < 1 || ld < cols) return set_error(TCS_ERR_INVALID, "maxabs: bad shape");
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(tcs_zero_u64_kernel, dim3(1), dim3(1), 0, s, (unsigned long long *)d_out);
    hipLaunchKernelGGL(tcs_maxabs_kernel<T>, dim3(grid_for((int64_t)rows * cols)), dim3(256), 0, s, d_src, rows,
                       cols, ld, (unsigned long long *)d_out);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

extern "C" int tcs_maxabs_f32(const float *d_src, int rows, int cols, int64_t ld, double *d_out, void *stream) {
    return maxabs_impl<float>(d_src, rows, cols, ld, d_out, stream);
}
extern "C" int tcs_maxabs_f64(const double *d_src, int rows, int cols, int64_t ld, double *d_out, void *stream) {
    return maxabs_impl<double>(d_src, rows, cols, ld, d_out, stream);
}

template <typename T>
static int round_impl(T *d_buf, int64_t count, int precision, void *stream) {
    if (!d_buf || count < 0) return set_error(TCS_ERR_INVALID, "round: bad buffer");
    if (precision != TCS_PREC_FP16 && precision != TCS_PREC_BF16)
        return set_error(TCS_ERR_INVALID, "round: precision must be FP16 or BF16");
    if (count == 0) return TCS_OK;
    hipLaunchKernelGGL(tcs_round_kernel<T>, dim3(grid_for(count)), dim3(256), 0, (hipStream_t)stream, d_buf, count,
                       precision);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

extern "C" int tcs_round_f64(double *d_buf, int64_t count, int precision, void *stream) {
    return round_impl<double>(d_buf, count, precision, stream);
}
extern "C" int tcs_round_f32(float *d_buf, int64_t count, int precision, void *stream) {
    return round_impl<float>(d_buf, count, precision, stream);
}

template <typename T>
static int gen_uniform_impl(T *d_dst, int rows, int cols, uint64_t seed, uint64_t stream_id, double lo, double hi,
                            void *stream) {
    if (!d_dst || rows < 1 || cols < 1) return set_error(TCS_ERR_INVALID, "gen_uniform: bad arguments");
    const int64_t total = (int64_t)rows * cols;
    hipLaunchKernelGGL(tcs_gen_uniform_kernel<T>, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream, d_dst,
                       rows, cols, stream_key(seed, stream_id), lo, hi - lo);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

extern "C" int tcs_gen_uniform_f32(float *d_dst, int rows, int cols, uint64_t seed, uint64_t stream_id, float lo,
                                   float hi, void *stream) {
    return gen_uniform_impl<float>(d_dst, rows, cols, seed, stream_id, (double)lo, (double)hi, stream);
}
extern "C" int tcs_gen_uniform_f64(double *d_dst, int rows, int cols, uint64_t seed, uint64_t stream_id, double lo,
                                   double hi, void *stream) {
    return gen_uniform_impl<double>(d_dst, rows, cols, seed, stream_id, lo, hi, stream);
}

// Four sinusoids: wavenumbers kx, ky uniform integers in [1, 64], phase
// U[0, 2pi), amplitude U[0.5, 1)/4, all drawn from stream 1 of `seed`; the
// noise uses stream 2.
extern "C" int tcs_gen_sinusoid_f64(double *d_dst, int rows, int cols, uint64_t seed, double noise, void *stream) {
    if (!d_dst || rows < 1 || cols < 1) return set_error(TCS_ERR_INVALID, "gen_sinusoid: bad arguments");
    SinParams sp;
    const uint64_t key = stream_key(seed, 1);
    for (int q = 0; q < 4; ++q) {
        sp.kx[q] = 1.0 + floor(64.0 * u24(key, 4 * q + 0));
        sp.ky[q] = 1.0 + floor(64.0 * u24(key, 4 * q + 1));
        sp.phase[q] = 6.283185307179586476925286766559 * u24(key, 4 * q + 2);
        sp.amp[q] = 0.25 * (0.5 + 0.5 * u24(key, 4 * q + 3));
    }
    const int64_t total = (int64_t)rows * cols;
    hipLaunchKernelGGL(tcs_gen_sinusoid_kernel, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream, d_dst,
                       rows, cols, sp, stream_key(seed, 2), noise);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

// Coefficients U[0,1) from stream 3 of `seed`, normalised in FP64 to sum 1 and
// rounded to FP32; star stencils draw only the 2n-1 taps of the centre row and
// column (in row-major order of B).
extern "C" int tcs_gen_coeffs(int stencil, uint64_t seed, float *B, int n) {
    const int r = tcs_radius(stencil);
    if (r < 0) return set_error(TCS_ERR_INVALID, "gen_coeffs: bad stencil %d", stencil);
    if (!B || n != 2 * r + 1) return set_error(TCS_ERR_INVALID, "gen_coeffs: n must be %d", 2 * r + 1);
    const bool star = stencil == TCS_STAR2D1R || stencil == TCS_STAR2D2R;
    const uint64_t key = stream_key(seed, 3);
    double v[25], sum = 0.0;
    for (int k = 0; k < n * n; ++k) {
        const int i = k / n, j = k % n;
        v[k] = (!star || i == r || j == r) ? u24(key, (uint64_t)k) : 0.0;
        sum += v[k];
    }
    for (int k = 0; k < n * n; ++k) B[k] = (float)(sum > 0.0 ? v[k] / sum : 0.0);
    return TCS_OK;
}
