// Host side of the C ABI: status reporting, parameters, problem set-up, sweep
// loops, hipGraph capture and the drop-in host_<stencil> drivers that replace
// host_code of layout16/host_<stencil>.cu and baseline/host_<stencil>.cu.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <thread>
#include <vector>

#include "tcs_internal.hpp"

namespace tcs {

static thread_local char g_err[512];

int set_error(int status, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return status;
}

int hip_error(hipError_t e, const char *what) {
    return set_error(e == hipErrorOutOfMemory ? TCS_ERR_ALLOC : TCS_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

// float -> IEEE binary16 bits, round to nearest even (overflow to inf, NaN kept quiet).
uint16_t f32_to_f16_bits(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return sign | 0x7e00u;
    if (ax >= 0x477ff000u) return sign | 0x7c00u;  // |f| >= 65520 rounds to inf
    if (ax >= 0x38800000u) {                       // normal binary16 range
        const uint32_t m = ax - 0x38000000u;       // rebias the exponent by 127 - 15
        return sign | (uint16_t)((m + 0x0fffu + ((m >> 13) & 1u)) >> 13);
    }
    if (ax < 0x33000000u) return sign;  // below 2^-25 (2^-25 itself ties to even = 0)
    const uint32_t e = ax >> 23;
    const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e;  // value / 2^-24 = mant >> shift
    uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return sign | (uint16_t)q;
}

// float -> bfloat16 bits, round to nearest even (NaN kept quiet).
uint16_t f32_to_bf16_bits(float f) {
    uint32_t x;
    memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    return (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

void build_coef_windows(const tcs_problem *p, CoefWindows *cw) {
    memset(cw, 0, sizeof *cw);
    for (int pl = 0; pl < 2; ++pl)
        for (int z = 0; z < 5; ++z)
            for (int j = 0; j < 12; ++j)
                for (int e = 0; e < 8; ++e) {
                    const int h = j - 7 + e;  // tap of window element e
                    if (h < 0 || h > 4) continue;
                    const int k = z * 5 + h;
                    uint16_t bits;
                    if (p->precision == TCS_PREC_SPLIT) {
                        bits = pl == 0 ? p->coeff_hi[k] : p->coeff_lo[k];
                    } else {
                        // hi = round(c); lo = round(c - hi), c - hi exact in FP32
                        const bool bf = p->precision == TCS_PREC_BF16;
                        const float c = p->coeff[k];
                        const uint16_t hb = bf ? f32_to_bf16_bits(c) : f32_to_f16_bits(c);
                        if (pl == 0) {
                            bits = hb;
                        } else {
                            if (!has_coef_lo(kernel_prec(p))) continue;
                            const float rest = c - (bf ? bf16_bits_to_f32(hb) : f16_bits_to_f32(hb));
                            bits = bf ? f32_to_bf16_bits(rest) : f32_to_f16_bits(rest);
                        }
                    }
                    cw->w[pl][z][j][e >> 1] |= (uint32_t)bits << (16 * (e & 1));
                }
}

float bf16_bits_to_f32(uint16_t h) {
    const uint32_t x = (uint32_t)h << 16;
    float f;
    memcpy(&f, &x, 4);
    return f;
}

int kernel_prec(const tcs_problem *p) {
    const bool lo = p->coef_lo && p->method == TCS_METHOD_MFMA &&
                    (p->precision == TCS_PREC_FP16 || p->precision == TCS_PREC_BF16);
    return p->precision | (lo ? kCoefLo : 0);
}

float f16_bits_to_f32(uint16_t h) {
    const uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    uint32_t x;
    if (e == 0) {
        const float v = ldexpf((float)m, -24);
        memcpy(&x, &v, 4);
        x |= sign;
    } else if (e == 31) {
        x = sign | 0x7f800000u | (m << 13);
    } else {
        x = sign | ((e + 112u) << 23) | (m << 13);
    }
    float f;
    memcpy(&f, &x, 4);
    return f;
}

}  // namespace tcs

using namespace tcs;

// ------------------------------------------------------------------ status
extern "C" const char *tcs_status_string(int status) {
    switch (status) {
    case TCS_OK: return "ok";
    case TCS_ERR_INVALID: return "invalid argument";
    case TCS_ERR_UNSUPPORTED: return "unsupported";
    case TCS_ERR_HIP: return "HIP runtime error";
    case TCS_ERR_ALLOC: return "allocation failed";
    }
    return "unknown status";
}

extern "C" const char *tcs_last_error(void) { return g_err; }

// ------------------------------------------------------------------ params
extern "C" void tcs_params_default(tcs_params *p) {
    if (!p) return;
    p->TOLERANCE = 0.01;
    p->MIN_NUMBER = 1e-6;
    p->RUN_TIMES = 500;
    p->STEP_TIMES = 1;
    p->SKIP_TIMES = 400;
    p->TILE_SIZE = 0;
}

extern "C" void tcs_params_from_env(tcs_params *p) {
    if (!p) return;
    tcs_params_default(p);
    const char *v;
    if ((v = getenv("TCS_TOLERANCE"))) p->TOLERANCE = atof(v);
    if ((v = getenv("TCS_MIN_NUMBER"))) p->MIN_NUMBER = atof(v);
    if ((v = getenv("TCS_RUN_TIMES"))) p->RUN_TIMES = atoi(v);
    if ((v = getenv("TCS_STEP_TIMES"))) p->STEP_TIMES = atoi(v);
    if ((v = getenv("TCS_SKIP_TIMES"))) p->SKIP_TIMES = atoi(v);
    if ((v = getenv("TCS_TILE_SIZE"))) p->TILE_SIZE = atoi(v);
}

// ------------------------------------------------------------ problem set-up
extern "C" int tcs_radius(int stencil) {
    switch (stencil) {
    case TCS_STAR2D1R: case TCS_BOX2D1R: return 1;
    case TCS_STAR2D2R: case TCS_BOX2D2R: return 2;
    }
    return -1;
}

extern "C" int tcs_layout_init(tcs_layout *L, int rows, int cols, int align) {
    if (!L) return set_error(TCS_ERR_INVALID, "layout: null pointer");
    if (rows < 1 || cols < 1) return set_error(TCS_ERR_INVALID, "layout: grid %dx%d is empty", rows, cols);
    if (align != 1 && align != 2 && align != 4 && align != 8)
        return set_error(TCS_ERR_INVALID, "layout: align %d not in {1,2,4,8}", align);
    if (rows > (1 << 28) || cols > (1 << 28)) return set_error(TCS_ERR_INVALID, "layout: grid too large");
    memset(L, 0, sizeof *L);
    L->rows = rows;
    L->cols = cols;
    L->align = align;
    L->nbr = ((rows + 15) / 16 + align - 1) / align * align;
    L->nbc = ((cols + 15) / 16 + align - 1) / align * align;
    L->sbr = L->nbr + 2;
    L->sbc = L->nbc + 2;
    L->plane_elems = (int64_t)L->sbr * L->sbc * 256;
    return TCS_OK;
}

static bool tile_offered(int th, int tw) {
    return (th == 2 && tw == 4) || (th == 4 && tw == 4) || (th == 2 && tw == 8) || (th == 4 && tw == 8) ||
           (th == 8 && tw == 8);
}

static bool frag32(const tcs_problem *p) { return p->frag == 32; }

// Default tile: the fastest shape measured on 16384^2 grids for this method,
// precision and radius, stepped down until the grid has at least 4096 tiles
// (16 per CU), so that smaller grids still fill the chip.  Over the mesh sweep
// 160..7200 x all stencils this picks within 2.1% (MFMA) and 2.5% (direct) of
// the best tile on average; a 1024-tile threshold was 4.0% and 6.5% off.
static void default_tile(const tcs_problem *p, int *th, int *tw) {
    static const int kPref[5][2] = {{8, 8}, {4, 8}, {4, 4}, {2, 8}, {2, 4}};
    // 32x32 fragments need 2x2-brick blocks, 4 per tile; 4x4 is the smallest
    // tile with single-sweep, fused and batched 32x32 kernels.
    const int last = frag32(p) ? 2 : 4;
    int first;
    if (p->method == TCS_METHOD_MFMA)
        // 8x8 for 32x32 fragments and for hi+lo coefficients (4 waves per SIMD:
        // box2d2r 16384^2 FP16 193.8 us against 222.7 us on 4x8), else 4x8
        // (warm GPU, box2d2r 16384^2: FP16 177.4 us against 180.0 on 8x8, BF16
        // 177.5 against 180.0, SPLIT 358.6 against 383.5)
        first = frag32(p) || has_coef_lo(kernel_prec(p)) ? 0 : 1;
    else
        first = 0;  // direct kernels: 8x8 (16384^2, all stencils, both forms; 4x8 within 1-2%)
    const long br = (p->layout.rows + 15) / 16, bc = (p->layout.cols + 15) / 16;
    int pick = last;
    for (int i = first; i <= last; ++i) {
        const long tiles = ((br + kPref[i][0] - 1) / kPref[i][0]) * ((bc + kPref[i][1] - 1) / kPref[i][1]);
        if (tiles >= 4096) {
            pick = i;
            break;
        }
    }
    *th = kPref[pick][0];
    *tw = kPref[pick][1];
}

extern "C" int tcs_problem_set_tile(tcs_problem *p, int tile_h, int tile_w);

// The default tile of the problem's method, precision and fragment, then the
// single-grid entry of $TCS_TUNE_CACHE when that names a file (keeping the
// fragment when want_frag is set).
static int derive_tile(tcs_problem *p, int want_frag = 0) {
    int th, tw;
    default_tile(p, &th, &tw);
    TCS_TRY(tcs_problem_set_tile(p, th, tw));
    const char *cache = getenv("TCS_TUNE_CACHE");
    if (!cache || !*cache) return TCS_OK;
    p->wg_per_cu = 0;  // the launch parameters are re-derived with the tile
    p->fuse = 0;
    return apply_tuning(p, cache, 0, want_frag);
}

extern "C" int tcs_problem_set_tile(tcs_problem *p, int tile_h, int tile_w) {
    if (!p) return set_error(TCS_ERR_INVALID, "set_tile: null problem");
    if (!tile_offered(tile_h, tile_w))
        return set_error(TCS_ERR_UNSUPPORTED, "tile %dx%d is not offered", tile_h, tile_w);
    const int align = tile_h > tile_w ? tile_h : tile_w;
    int st = tcs_layout_init(&p->layout, p->layout.rows, p->layout.cols, align);
    if (st) return st;
    p->tile_h = tile_h;
    p->tile_w = tile_w;
    return TCS_OK;
}

// The reference's TILE_SIZE (common/param.hpp:11-13) is the number of bricks a
// thread block walks along x, on blocks one brick tall (layout16/host_box2d2r.cu:
// 166-168); its sweep covers 1..15 (script/param.py:5).  Here it selects the
// offered tile whose width is nearest: 1..5 -> 4 bricks wide, 6 and above -> 8,
// keeping the current tile height where that shape is offered (else the
// nearest offered height).  Brick results do not depend on the tile.
extern "C" int tcs_problem_set_tile_size(tcs_problem *p, int tile_size) {
    if (!p) return set_error(TCS_ERR_INVALID, "set_tile_size: null problem");
    if (tile_size <= 0) return TCS_OK;  // the tuned default stays
    const int tw = tile_size < 6 ? 4 : 8;
    static const int kHeights[3] = {2, 4, 8};
    int th = p->tile_h > 0 ? p->tile_h : 4, best = -1;
    for (int h : kHeights) {
        const bool ok = tile_offered(h, tw) && (!frag32(p) || (h % 2 == 0 && (h / 2) * (tw / 2) % 4 == 0));
        if (ok && (best < 0 || abs(h - th) < abs(best - th))) best = h;
    }
    return tcs_problem_set_tile(p, best, tw);
}

extern "C" int tcs_problem_init(tcs_problem *p, int stencil, int method, int precision, int rows, int cols,
                                const float *B, int n) {
    if (!p) return set_error(TCS_ERR_INVALID, "problem: null pointer");
    const int r = tcs_radius(stencil);
    if (r < 0) return set_error(TCS_ERR_INVALID, "problem: unknown stencil %d", stencil);
    if (method != TCS_METHOD_MFMA && method != TCS_METHOD_DIRECT && method != TCS_METHOD_DIRECT_DOT2)
        return set_error(TCS_ERR_INVALID, "problem: unknown method %d", method);
    if (precision != TCS_PREC_FP16 && precision != TCS_PREC_BF16 && precision != TCS_PREC_SPLIT)
        return set_error(TCS_ERR_INVALID, "problem: unknown precision %d", precision);
    if (method != TCS_METHOD_MFMA && precision != TCS_PREC_FP16)
        return set_error(TCS_ERR_UNSUPPORTED, "problem: the direct-FMA kernels run FP16 storage only");
    if (!B) return set_error(TCS_ERR_INVALID, "problem: null coefficients");
    if (n != 2 * r + 1) return set_error(TCS_ERR_INVALID, "problem: n = %d, stencil needs n = %d", n, 2 * r + 1);
    if (rows < 1 || cols < 1) return set_error(TCS_ERR_INVALID, "problem: grid %dx%d is empty", rows, cols);
    memset(p, 0, sizeof *p);
    p->frag = 16;
    p->stencil = stencil;
    p->method = method;
    p->precision = precision;
    p->radius = r;
    p->layout.rows = rows;
    p->layout.cols = cols;
    const bool star = stencil == TCS_STAR2D1R || stencil == TCS_STAR2D2R;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            if (!star || i == r || j == r) p->coeff[(i + 2 - r) * 5 + (j + 2 - r)] = B[i * n + j];
    float maxc = 0.0f;
    for (int k = 0; k < 25; ++k) {
        if (!isfinite(p->coeff[k])) return set_error(TCS_ERR_INVALID, "problem: non-finite coefficient");
        maxc = fmaxf(maxc, fabsf(p->coeff[k]));
    }
    if (precision == TCS_PREC_SPLIT) {
        // Scale so that the largest |coefficient| lies in [1, 2): the lo parts
        // then stay in binary16's normal range for all but tiny taps.
        p->coeff_exp = maxc > 0.0f ? -ilogbf(maxc) : 0;
        for (int k = 0; k < 25; ++k) {
            const float c = ldexpf(p->coeff[k], p->coeff_exp);
            const uint16_t hb = f32_to_f16_bits(c);
            p->coeff_hi[k] = hb;
            p->coeff_lo[k] = f32_to_f16_bits(c - f16_bits_to_f32(hb));
        }
    }
    return derive_tile(p);
}

extern "C" int tcs_problem_set_frag(tcs_problem *p, int frag) {
    if (!p) return set_error(TCS_ERR_INVALID, "set_frag: null problem");
    if (frag != 16 && frag != 32) return set_error(TCS_ERR_INVALID, "set_frag: fragment %d not in {16, 32}", frag);
    if (frag == 32 && (p->method != TCS_METHOD_MFMA || p->precision == TCS_PREC_SPLIT))
        return set_error(TCS_ERR_UNSUPPORTED, "set_frag: 32x32 fragments need method MFMA with FP16 or BF16");
    p->frag = frag;
    return derive_tile(p, frag);
}

extern "C" int tcs_problem_set_coef_lo(tcs_problem *p, int on) {
    if (!p) return set_error(TCS_ERR_INVALID, "set_coef_lo: null problem");
    p->coef_lo = on ? 1 : 0;
    return derive_tile(p);
}

extern "C" int64_t tcs_buffer_bytes(const tcs_problem *p) {
    if (!p) return 0;
    return p->layout.plane_elems * 2 * (p->precision == TCS_PREC_SPLIT ? 2 : 1);
}

// SPLIT storage holds x * 2^s as an FP16 pair, so every value a run produces
// must stay below binary16's 65504 after scaling: with taps of absolute sum
// S > 1 a sweep can grow max|x| by S, so the exponent is taken for
// max|x| * max(1, S)^sweeps (growth capped at 2^60).
// The factor by which `sweeps` sweeps of p can grow max|x|: S^sweeps for taps
// of absolute sum S > 1 (capped at 2^60), else 1.
static double run_growth(const tcs_problem *p, int sweeps) {
    double s = 0.0;
    for (int k = 0; k < 25; ++k) s += fabs((double)p->coeff[k]);
    double growth = 1.0;
    for (int i = 0; i < sweeps && s > 1.0 && growth < 0x1p60; ++i) growth *= s;
    return growth > 0x1p60 ? 0x1p60 : growth;
}

extern "C" int tcs_split_scale_exp_for(const tcs_problem *p, double maxabs, int sweeps) {
    if (!p) return tcs_split_scale_exp(maxabs);
    return tcs_split_scale_exp(maxabs * run_growth(p, sweeps));
}

extern "C" int tcs_scale_exp_for(const tcs_problem *p, double maxabs, int sweeps) {
    if (!p) return 0;
    if (p->precision == TCS_PREC_SPLIT) return tcs_split_scale_exp_for(p, maxabs, sweeps);
    if (p->precision != TCS_PREC_FP16) return 0;
    const double grown = maxabs * run_growth(p, sweeps);
    if (!(grown > 0.0) || !isfinite(grown)) return 0;
    return 14 - ilogb(grown);  // grown * 2^e in [2^14, 2^15)
}

extern "C" int tcs_split_scale_exp(double maxabs) {
    // Bring max|x| to [2^9, 2^10): the lo parts keep full precision and a
    // stencil with |sum of coefficients| up to 32 cannot overflow binary16.
    if (!(maxabs > 0.0) || !isfinite(maxabs)) return 0;
    return 9 - ilogb(maxabs);
}

extern "C" const char *tcs_kernel_name(const tcs_problem *p) { return p ? sweep_kernel_name(p) : nullptr; }
extern "C" const char *tcs_fused_kernel_name(const tcs_problem *p) { return p ? fused_kernel_name(p) : nullptr; }
extern "C" const char *tcs_edge_kernel_name(const tcs_problem *p, int mode) {
    return p && mode >= 1 && mode <= 3 ? edge_kernel_name(p, mode) : nullptr;
}
extern "C" int tcs_mfma_per_brick(const tcs_problem *p) { return p ? mfma_per_brick(p) : 0; }
extern "C" int64_t tcs_mfma_flops_per_brick(const tcs_problem *p) { return p ? mfma_flops_per_brick(p) : 0; }

// ---------------------------------------------------------------- sweeping
extern "C" int tcs_sweep(const tcs_problem *p, const void *d_in, void *d_out, void *stream) {
    if (!p) return set_error(TCS_ERR_INVALID, "sweep: null problem");
    if (!d_in || !d_out) return set_error(TCS_ERR_INVALID, "sweep: null buffer");
    if (d_in == d_out) return set_error(TCS_ERR_INVALID, "sweep: in and out must differ (Jacobi)");
    return launch_sweep(p, d_in, d_out, (hipStream_t)stream);
}

extern "C" int tcs_run(const tcs_problem *p, void *d_buf0, void *d_buf1, int sweeps, void *stream) {
    if (!p) return set_error(TCS_ERR_INVALID, "run: null problem");
    if (sweeps < 0) return set_error(TCS_ERR_INVALID, "run: negative sweep count");
    void *buf[2] = {d_buf0, d_buf1};
    const int fuse = p->fuse > 1 ? p->fuse : 1;
    if (fuse == 1) {
        for (int s = 0; s < sweeps; ++s) {
            int st = tcs_sweep(p, buf[s & 1], buf[(s + 1) & 1], stream);
            if (st) return st;
        }
        return TCS_OK;
    }
    if (p->method != TCS_METHOD_MFMA) return set_error(TCS_ERR_UNSUPPORTED, "run: fused sweeps need the MFMA method");
    if (p->radius * (fuse - 1) > 16)
        return set_error(TCS_ERR_INVALID, "run: fuse %d exceeds 1 + 16/radius = %d", fuse, 1 + 16 / p->radius);
    if (!d_buf0 || !d_buf1 || d_buf0 == d_buf1) return set_error(TCS_ERR_INVALID, "run: need two distinct buffers");
    if (sweeps == 0) return TCS_OK;
    // Launch count with the parity of `sweeps` (result in buf[sweeps % 2]),
    // sweeps spread evenly: each launch gets at most `fuse`.
    int nl = (sweeps + fuse - 1) / fuse;
    if ((nl & 1) != (sweeps & 1)) ++nl;
    const int base = sweeps / nl, extra = sweeps % nl;
    for (int i = 0; i < nl; ++i) {
        int st = launch_fused(p, buf[i & 1], buf[(i + 1) & 1], base + (i < extra ? 1 : 0), (hipStream_t)stream);
        if (st) return st;
    }
    return TCS_OK;
}

// Row-major FP32 in and out.  Fused edges (pack into the first sweep, unpack
// into the last) where an edge kernel exists; otherwise the three steps.
extern "C" int tcs_run_f32(const tcs_problem *p, const float *d_src, float *d_dst, int64_t ld, void *d_buf0,
                           void *d_buf1, int sweeps, int scale_exp, void *stream) {
    if (!p) return set_error(TCS_ERR_INVALID, "run_f32: null problem");
    if (!d_src || !d_dst) return set_error(TCS_ERR_INVALID, "run_f32: null source or destination");
    if (d_src == d_dst) return set_error(TCS_ERR_INVALID, "run_f32: source and destination must differ");
    if (sweeps < 0) return set_error(TCS_ERR_INVALID, "run_f32: negative sweep count");
    if (ld < p->layout.cols)
        return set_error(TCS_ERR_INVALID, "run_f32: ld %lld < cols %d", (long long)ld, p->layout.cols);
    const hipStream_t s = (hipStream_t)stream;
    // The storage buffers' prior contents do not matter: every data brick is
    // written before it is read, and the ghost rings (read as the frame of the
    // edge tiles, where the MFMA operands multiply them, if only by zero
    // coefficients) are cleared here.
    if (sweeps >= 1 && edge_supported(p)) {
        if (sweeps == 1) return launch_edge(p, 3, nullptr, nullptr, d_src, d_dst, ld, scale_exp, s);
        if (!d_buf0 || !d_buf1 || d_buf0 == d_buf1)
            return set_error(TCS_ERR_INVALID, "run_f32: %d sweeps need two distinct storage buffers", sweeps);
        TCS_TRY(clear_ring(p, d_buf0, s));
        if (sweeps > 2) TCS_TRY(clear_ring(p, d_buf1, s));
        TCS_TRY(launch_edge(p, 1, nullptr, d_buf0, d_src, nullptr, ld, scale_exp, s));
        // Sweeps 2 .. sweeps-1 ping-pong from d_buf0; the last reads the result.
        TCS_TRY(tcs_run(p, d_buf0, d_buf1, sweeps - 2, stream));
        void *last = (sweeps - 2) % 2 ? d_buf1 : d_buf0;
        return launch_edge(p, 2, last, nullptr, d_src, d_dst, ld, scale_exp, s);
    }
    if (!d_buf0 || !d_buf1 || d_buf0 == d_buf1)
        return set_error(TCS_ERR_INVALID, "run_f32: this problem needs two distinct storage buffers");
    TCS_TRY(tcs_pack_f32(p, d_src, ld, d_buf0, scale_exp, stream));
    TCS_TRY(clear_ring(p, d_buf1, s));
    TCS_TRY(tcs_run(p, d_buf0, d_buf1, sweeps, stream));
    return tcs_unpack_f32(p, sweeps % 2 ? d_buf1 : d_buf0, d_dst, ld, scale_exp, d_src, stream);
}

extern "C" int tcs_graph_create(const tcs_problem *p, void *d_buf0, void *d_buf1, int sweeps, void *stream,
                                void **graph_exec) {
    if (!graph_exec) return set_error(TCS_ERR_INVALID, "graph: null output");
    *graph_exec = nullptr;
    hipStream_t cs = (hipStream_t)stream;
    bool own = false;
    if (!cs) {
        TCS_HIP_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        own = true;
    }
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        if (own) (void)hipStreamDestroy(cs);
        return hip_error(e, "hipStreamBeginCapture");
    }
    const int st = tcs_run(p, d_buf0, d_buf1, sweeps, cs);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(cs, &g);
    if (own) (void)hipStreamDestroy(cs);
    if (st) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return hip_error(e, "hipGraphInstantiate");
    *graph_exec = ge;
    return TCS_OK;
}

extern "C" int tcs_graph_launch(void *graph_exec, void *stream) {
    if (!graph_exec) return set_error(TCS_ERR_INVALID, "graph: null exec");
    TCS_HIP_CHECK(hipGraphLaunch((hipGraphExec_t)graph_exec, (hipStream_t)stream));
    return TCS_OK;
}

extern "C" int tcs_graph_destroy(void *graph_exec) {
    if (!graph_exec) return TCS_OK;
    TCS_HIP_CHECK(hipGraphExecDestroy((hipGraphExec_t)graph_exec));
    return TCS_OK;
}

// ----------------------------------------------------------- host drivers
namespace {

// Device allocations of one host call, released in every exit path (after
// the call's stream has drained, so that no copy or kernel an error path left
// in flight still uses them).
struct DeviceScope {
    void *ptr[8] = {};
    int n = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~DeviceScope() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (int i = 0; i < n; ++i) (void)hipFree(ptr[i]);
        if (ev[0]) (void)hipEventDestroy(ev[0]);
        if (ev[1]) (void)hipEventDestroy(ev[1]);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int alloc(void **out, size_t bytes) {
        hipError_t e = hipMalloc(out, bytes);
        if (e != hipSuccess) return set_error(TCS_ERR_ALLOC, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        ptr[n++] = *out;
        return TCS_OK;
    }
};

// Host <-> device copies of pageable host arrays through two pinned staging
// chunks: the CPU copy of one chunk (split over several threads) overlaps the
// DMA of the other.  Pageable copies measured 28-30 GB/s host to device and
// 10-17 GB/s back on MI355X; pinned DMA runs at ~57 GB/s.
struct PinnedStage {
    static constexpr size_t kChunk = 32u << 20;
    void *buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~PinnedStage() {
        for (int i = 0; i < 2; ++i) {
            if (buf[i]) (void)hipHostFree(buf[i]);
            if (ev[i]) (void)hipEventDestroy(ev[i]);
        }
    }
    int init(size_t bytes) {
        const size_t len = bytes < kChunk ? bytes : kChunk;
        for (int i = 0; i < 2; ++i) {
            hipError_t e = hipHostMalloc(&buf[i], len, hipHostMallocDefault);
            if (e != hipSuccess) return hip_error(e, "hipHostMalloc (staging)");
            e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
            if (e != hipSuccess) return hip_error(e, "hipEventCreate (staging)");
        }
        return TCS_OK;
    }
};

// memcpy split over up to 8 threads (one CPU thread reaches ~10-20 GB/s).
static void parallel_copy(void *dst, const void *src, size_t bytes) {
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt < 1 ? 1 : (nt > 8 ? 8 : nt);
    if (bytes < (4u << 20) || nt == 1) {
        memcpy(dst, src, bytes);
        return;
    }
    const size_t part = (bytes / nt + 63) & ~(size_t)63;
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) {
        const size_t off = part * t;
        if (off >= bytes) break;
        const size_t len = off + part > bytes ? bytes - off : part;
        th.emplace_back([=] { memcpy(static_cast<char *>(dst) + off, static_cast<const char *>(src) + off, len); });
    }
    memcpy(dst, src, part < bytes ? part : bytes);
    for (auto &t : th) t.join();
}

static int staged_h2d(void *d_dst, const void *h_src, size_t bytes, hipStream_t s, PinnedStage &st) {
    for (size_t off = 0, i = 0; off < bytes; off += PinnedStage::kChunk, ++i) {
        const size_t len = bytes - off < PinnedStage::kChunk ? bytes - off : PinnedStage::kChunk;
        const int k = (int)(i & 1);
        if (i >= 2) TCS_HIP_CHECK(hipEventSynchronize(st.ev[k]));  // chunk i-2's DMA has drained buf[k]
        parallel_copy(st.buf[k], static_cast<const char *>(h_src) + off, len);
        TCS_HIP_CHECK(hipMemcpyAsync(static_cast<char *>(d_dst) + off, st.buf[k], len, hipMemcpyHostToDevice, s));
        TCS_HIP_CHECK(hipEventRecord(st.ev[k], s));
    }
    return TCS_OK;
}

static int staged_d2h(void *h_dst, const void *d_src, size_t bytes, hipStream_t s, PinnedStage &st) {
    const size_t n = (bytes + PinnedStage::kChunk - 1) / PinnedStage::kChunk;
    auto issue = [&](size_t i) -> int {
        const size_t off = i * PinnedStage::kChunk;
        const size_t len = bytes - off < PinnedStage::kChunk ? bytes - off : PinnedStage::kChunk;
        TCS_HIP_CHECK(hipMemcpyAsync(st.buf[i & 1], static_cast<const char *>(d_src) + off, len,
                                     hipMemcpyDeviceToHost, s));
        TCS_HIP_CHECK(hipEventRecord(st.ev[i & 1], s));
        return TCS_OK;
    };
    if (n == 0) return TCS_OK;
    int st0 = issue(0);
    if (st0) return st0;
    for (size_t i = 0; i < n; ++i) {
        if (i + 1 < n) {
            const int st1 = issue(i + 1);  // buf[(i+1)&1] was emptied by the CPU copy of chunk i-1
            if (st1) return st1;
        }
        const size_t off = i * PinnedStage::kChunk;
        const size_t len = bytes - off < PinnedStage::kChunk ? bytes - off : PinnedStage::kChunk;
        TCS_HIP_CHECK(hipEventSynchronize(st.ev[i & 1]));
        parallel_copy(static_cast<char *>(h_dst) + off, st.buf[i & 1], len);
    }
    return TCS_OK;
}

}  // namespace

static int host_sweep_impl(int stencil, int method, int precision, const float *h_A, const float *h_B, int n,
                           float *h_C, int rows, int cols, int sweeps, const tcs_params *params, float *ms_rep,
                           float *ms_per_sweep) {
    if (!h_A || !h_B || !h_C) return set_error(TCS_ERR_INVALID, "host: null host array");
    if (sweeps < 1) return set_error(TCS_ERR_INVALID, "host: sweeps must be >= 1");
    tcs_params prm;
    if (params)
        prm = *params;
    else
        tcs_params_from_env(&prm);
    tcs_problem p;
    TCS_TRY(tcs_problem_init(&p, stencil, method, precision, rows, cols, h_B, n));
    TCS_TRY(tcs_problem_set_tile_size(&p, prm.TILE_SIZE));

    PinnedStage stage;  // declared first: released after ds has drained the stream
    DeviceScope ds;
    TCS_HIP_CHECK(hipStreamCreateWithFlags(&ds.stream, hipStreamNonBlocking));
    TCS_HIP_CHECK(hipEventCreate(&ds.ev[0]));
    TCS_HIP_CHECK(hipEventCreate(&ds.ev[1]));
    const size_t grid_bytes = (size_t)rows * cols * sizeof(float);
    const size_t buf_bytes = (size_t)tcs_buffer_bytes(&p);
    void *d_src, *d_dst, *buf0, *buf1, *d_max;
    TCS_TRY(ds.alloc(&d_src, grid_bytes));
    TCS_TRY(ds.alloc(&d_dst, grid_bytes));
    TCS_TRY(ds.alloc(&buf0, buf_bytes));
    TCS_TRY(ds.alloc(&buf1, buf_bytes));
    TCS_TRY(ds.alloc(&d_max, sizeof(double)));
    hipStream_t s = ds.stream;
    TCS_TRY(stage.init(grid_bytes));
    TCS_TRY(staged_h2d(d_src, h_A, grid_bytes, s, stage));

    // Per-grid scaling (FP16 and SPLIT storage): a power of two that keeps the
    // run inside binary16's range.
    int scale_exp = 0;
    if (precision == TCS_PREC_SPLIT || precision == TCS_PREC_FP16) {
        TCS_TRY(tcs_maxabs_f32((const float *)d_src, rows, cols, cols, (double *)d_max, s));
        double mx = 0.0;
        TCS_HIP_CHECK(hipMemcpyAsync(&mx, d_max, sizeof mx, hipMemcpyDeviceToHost, s));
        TCS_HIP_CHECK(hipStreamSynchronize(s));
        scale_exp = tcs_scale_exp_for(&p, mx, sweeps);
    }
    TCS_TRY(tcs_pack_f32(&p, (const float *)d_src, cols, buf0, scale_exp, s));
    TCS_TRY(tcs_clear(&p, buf1, s));

    // Timing as the reference does (layout16/host_box2d2r.cu:171-193): RUN_TIMES
    // repetitions, the first SKIP_TIMES untimed; every repetition starts from
    // the packed input.
    const int reps = prm.RUN_TIMES > 0 ? prm.RUN_TIMES : 1;
    const int skip = prm.SKIP_TIMES < reps ? (prm.SKIP_TIMES > 0 ? prm.SKIP_TIMES : 0) : 0;
    double sum_ms = 0.0;
    for (int t = 0; t < reps; ++t) {
        if (t > 0 && sweeps > 1) TCS_TRY(tcs_pack_f32(&p, (const float *)d_src, cols, buf0, scale_exp, s));
        TCS_HIP_CHECK(hipEventRecord(ds.ev[0], s));
        TCS_TRY(tcs_run(&p, buf0, buf1, sweeps, s));
        TCS_HIP_CHECK(hipEventRecord(ds.ev[1], s));
        TCS_HIP_CHECK(hipEventSynchronize(ds.ev[1]));
        float ms = 0.0f;
        TCS_HIP_CHECK(hipEventElapsedTime(&ms, ds.ev[0], ds.ev[1]));
        if (t >= skip) sum_ms += ms;
    }
    const double per_rep = sum_ms / (reps - skip);
    if (ms_rep) *ms_rep = (float)(per_rep / (prm.STEP_TIMES > 0 ? prm.STEP_TIMES : 1));
    if (ms_per_sweep) *ms_per_sweep = (float)(per_rep / sweeps);

    void *result = (sweeps & 1) ? buf1 : buf0;
    TCS_TRY(tcs_unpack_f32(&p, result, (float *)d_dst, cols, scale_exp, (const float *)d_src, s));
    TCS_TRY(staged_d2h(h_C, d_dst, grid_bytes, s, stage));
    TCS_HIP_CHECK(hipStreamSynchronize(s));
    return TCS_OK;
}

extern "C" int tcs_host_sweep(int stencil, int method, int precision, const float *h_A, const float *h_B, int n,
                              float *h_C, int rows, int cols, int sweeps, const tcs_params *params,
                              float *ms_per_sweep) {
    return host_sweep_impl(stencil, method, precision, h_A, h_B, n, h_C, rows, cols, sweeps, params, nullptr,
                           ms_per_sweep);
}

// The reference's driver contract: one sweep, timing printed, exit(-1) on error.
static void dropin(int stencil, int method, float *h_A, float *h_B, float *h_C, int N, int n) {
    float ms = 0.0f;
    const int st =
        host_sweep_impl(stencil, method, TCS_PREC_FP16, h_A, h_B, n, h_C, N, N, 1, nullptr, &ms, nullptr);
    if (st != TCS_OK) {
        printf("HIP error : %s, %s\n", tcs_status_string(st), tcs_last_error());
        exit(-1);
    }
    printf("[Time] Time used: %f ms\n", ms);
}

extern "C" void host_star2d1r(float *h_A, float *h_B, float *h_C, int N, int n) { dropin(TCS_STAR2D1R, TCS_METHOD_MFMA, h_A, h_B, h_C, N, n); }
extern "C" void host_star2d2r(float *h_A, float *h_B, float *h_C, int N, int n) { dropin(TCS_STAR2D2R, TCS_METHOD_MFMA, h_A, h_B, h_C, N, n); }
extern "C" void host_box2d1r(float *h_A, float *h_B, float *h_C, int N, int n) { dropin(TCS_BOX2D1R, TCS_METHOD_MFMA, h_A, h_B, h_C, N, n); }
extern "C" void host_box2d2r(float *h_A, float *h_B, float *h_C, int N, int n) { dropin(TCS_BOX2D2R, TCS_METHOD_MFMA, h_A, h_B, h_C, N, n); }
extern "C" void host_direct_star2d1r(float *h_A, float *h_B, float *h_C, int N, int n) { dropin(TCS_STAR2D1R, TCS_METHOD_DIRECT, h_A, h_B, h_C, N, n); }
extern "C" void host_direct_star2d2r(float *h_A, float *h_B, float *h_C, int N, int n) { dropin(TCS_STAR2D2R, TCS_METHOD_DIRECT, h_A, h_B, h_C, N, n); }
extern "C" void host_direct_box2d1r(float *h_A, float *h_B, float *h_C, int N, int n) { dropin(TCS_BOX2D1R, TCS_METHOD_DIRECT, h_A, h_B, h_C, N, n); }
extern "C" void host_direct_box2d2r(float *h_A, float *h_B, float *h_C, int N, int n) { dropin(TCS_BOX2D2R, TCS_METHOD_DIRECT, h_A, h_B, h_C, N, n); }
