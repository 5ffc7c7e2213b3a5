// Internal definitions shared by the HIP translation units of libtcstencil.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/tcstencil.h"

namespace tcs {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

constexpr int kWG = 256;  // threads per workgroup of every tiled kernel (4 wave64)

// Kernel precision ids: the storage precision (TCS_PREC_FP16, _BF16, _SPLIT)
// in the low two bits, plus kCoefLo for FP16/BF16 MFMA kernels that carry each
// coefficient as a hi+lo pair of that format (a second MFMA pass per fragment
// with the lo taps, on the same data operand, into the same accumulator).
constexpr int kCoefLo = 4;
__host__ __device__ constexpr int base_prec(int kp) { return kp & 3; }
__host__ __device__ constexpr bool has_coef_lo(int kp) { return (kp & kCoefLo) != 0; }
constexpr int kXCD = 8;   // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr int kLdsBytesPerCU = 160 * 1024;  // gfx950 LDS per CU

// Storage element of global grid point (gr, gc), gr, gc >= -16: bricks are
// 16x16 row-major, brick-row-major, behind a one-brick ghost ring, so stored
// brick (bi+1, bj+1) holds data brick (bi, bj).  (-2 & 15) == 14, so negative
// coordinates inside the ghost ring map to the right brick row/column.
__host__ __device__ inline int64_t stor_index(int gr, int gc, int sbc) {
    const int bi = (gr + 16) >> 4, bj = (gc + 16) >> 4;
    return ((int64_t)bi * sbc + bj) * 256 + ((gr & 15) << 4) + (gc & 15);
}

// Coefficient operand windows of the 16x16x32 MFMA, built on the host
// (build_coef_windows).  Lane (m, g) of the A operand holds c[z][8g + e - m],
// e = 0..7 (zero outside the taps 0..4): the window at d0 = 8g - m of the
// zero-padded tap row.  Only d0 in [-7, 4] gives a non-zero window, so
// w[p][z][j] (j < 12) holds the window at d0 = j - 7 and w[p][z][12] zeros, in
// storage bits (plane p: FP16 or BF16 taps; SPLIT: hi, then lo).  A lane loads
// its fragments straight from here (16 bytes each), with no per-element selects.
struct CoefWindows {
    uint32_t w[2][5][13][4];
};

// Arguments of one sweep launch, passed by value.
struct SweepArgs {
    const uint16_t *in;
    uint16_t *out;
    int64_t plane;         // elements per storage plane (SPLIT: lo plane offset)
    int32_t rows, cols;    // logical grid
    int32_t sbr, sbc;      // stored brick rows, stored bricks per brick row
    int32_t tiles_x;       // tiles per tile row
    int32_t tiles_xp;      // tiles_x rounded up to a multiple of 8 (virtual tile row)
    int32_t nvirt;         // virtual tiles: tile rows * tiles_xp (row-band order: see tcs_sweep.hip)
    int32_t tiles_y;       // tile rows
    int32_t rs;            // stencil radius: points closer than rs to the edge are copied
    float out_scale;       // SPLIT: 2^-coeff_exp
    float cf[25];          // 5x5 FP32 coefficient embedding (direct kernel)
    CoefWindows cw;        // MFMA operand windows (FP16/BF16 taps, or SPLIT hi and lo)
    // Row-major FP32 edges (tcs_mfma_edge): the first sweep reads rm_src instead
    // of bricks, the last writes rm_dst (ring points copied from rm_src).
    const float *rm_src;
    float *rm_dst;
    int64_t rm_ld;         // leading dimension of rm_src and rm_dst, elements
    float rm_scale_in;     // 2^scale_exp, applied before rounding to storage format
    float rm_scale_out;    // 2^-scale_exp, applied to stored values on output
    int32_t rm_vec;        // 1: rm_src, rm_dst and rm_ld allow 16-byte accesses
};

// One problem of a batch, device-resident (tcs_batch.cpp builds them).
struct BatchProblem {
    uint16_t *buf[2];      // ping-pong storage buffers
    int64_t plane;
    int32_t rows, cols, sbr, sbc;
    int32_t rs;            // radius
    int32_t tiles_x;       // tiles per tile row of this problem
    float out_scale;
    CoefWindows cw;        // MFMA operand windows of the problem's coefficients
};

// One tile of a batch launch: problem index and tile index within it.
struct BatchTile {
    int32_t prob, tile;
};

// Kernel launchers (tcs_sweep.hip).
// Batched launch: nsweeps > 1 fuses that many sweeps per tile; grid_per_xcd > 0
// caps the workgroups per XCD (persistent grid), else one workgroup per tile.
int launch_batch(int precision, int tile_h, int tile_w, int frag, const BatchProblem *d_probs,
                 const BatchTile *d_tiles, int ntiles, int parity, int nsweeps, int grid_per_xcd,
                 hipStream_t stream);
const char *batch_kernel_name(int precision, int tile_h, int tile_w, int frag, int nsweeps, bool persist);
int launch_sweep(const tcs_problem *p, const void *d_in, void *d_out, hipStream_t stream);
// Edge sweeps for tcs_run_f32: mode 1 reads row-major FP32 src and writes
// storage d_out, mode 2 reads storage d_in and writes row-major FP32 dst (ring
// from src), mode 3 both.  edge_supported: a kernel exists for the problem.
bool edge_supported(const tcs_problem *p);
const char *edge_kernel_name(const tcs_problem *p, int mode);
int launch_edge(const tcs_problem *p, int mode, const void *d_in, void *d_out, const float *src, float *dst,
                int64_t ld, int scale_exp, hipStream_t stream);
int launch_fused(const tcs_problem *p, const void *d_in, void *d_out, int nsweeps, hipStream_t stream);
// Zero the ghost ring of a storage buffer of this problem (tcs_transform.hip).
int clear_ring(const tcs_problem *p, void *d_buf, hipStream_t stream);
const char *sweep_kernel_name(const tcs_problem *p);
const char *fused_kernel_name(const tcs_problem *p);
int mfma_per_brick(const tcs_problem *p);
int64_t mfma_flops_per_brick(const tcs_problem *p);

// Error reporting (tcs_host.cpp).
int set_error(int status, const char *fmt, ...);
int hip_error(hipError_t e, const char *what);

// Host FP16/BF16 rounding helpers (round to nearest even).
uint16_t f32_to_f16_bits(float f);
uint16_t f32_to_bf16_bits(float f);
float f16_bits_to_f32(uint16_t h);
float bf16_bits_to_f32(uint16_t h);

// Kernel precision id of a problem (tcs_host.cpp): its precision, with
// kCoefLo for MFMA FP16/BF16 problems with coef_lo set.
int kernel_prec(const tcs_problem *p);

// Tuning cache entry applied to p (tcs_tune.cpp); want_frag != 0 keeps an
// explicitly chosen fragment (entries for the other one are skipped).
int apply_tuning(tcs_problem *p, const char *path, int batched, int want_frag);

// MFMA operand windows of a problem's coefficients (tcs_host.cpp).
void build_coef_windows(const tcs_problem *p, CoefWindows *cw);

}  // namespace tcs

// Return a non-zero status from the enclosing function.
#define TCS_TRY(expr)          \
    do {                       \
        int _st = (expr);      \
        if (_st) return _st;   \
    } while (0)

#define TCS_HIP_CHECK(expr)                                          \
    do {                                                             \
        hipError_t _e = (expr);                                      \
        if (_e != hipSuccess) return ::tcs::hip_error(_e, #expr);    \
    } while (0)
