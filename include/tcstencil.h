/*
 * tcstencil.h -- C ABI of the MI355X (gfx950) matrix-core 2D stencil library.
 *
 * Four stencils (star2d1r, star2d2r, box2d1r, box2d2r) run as small dense MFMA
 * products over a 16x16 brick layout, next to a plain direct-FMA kernel.  All
 * signatures use plain C types: pointers, sizes and ints.  `stream` is a
 * hipStream_t passed as void* (NULL = the default stream).
 *
 * Two layers:
 *   1. Drop-in drivers with the reference's host_code signature
 *      (layout16/host_<stencil>.cu, baseline/host_<stencil>.cu).
 *   2. A device-pointer API returning status codes: a problem descriptor, the
 *      brick pack/unpack transform, sweeps, hipGraph capture, FP64 oracle sweep
 *      and error norms, seeded generators.  It keeps no global mutable state;
 *      every call takes the caller's stream.
 */
#ifndef TCSTENCIL_H
#define TCSTENCIL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- constants */
enum { TCS_STAR2D1R = 0, TCS_STAR2D2R = 1, TCS_BOX2D1R = 2, TCS_BOX2D2R = 3 };
/* MFMA: the matrix-core kernel.  DIRECT: one FP32 FMA per tap and point with
 * FP32 taps (the drop-in for baseline/kernel_<stencil>.cu).  DIRECT_DOT2: the
 * direct kernel on packed halves, v_dot2_f32_f16 per pair of adjacent taps
 * (taps rounded to FP16, FP32 accumulation): the fastest plain-arithmetic form,
 * measured beside the MFMA kernel. */
enum { TCS_METHOD_MFMA = 0, TCS_METHOD_DIRECT = 1, TCS_METHOD_DIRECT_DOT2 = 2 };
/* FP16: FP16 storage, FP16 coefficients, MFMA f16 with FP32 accumulation.
 * BF16: the same in BF16.
 * SPLIT: storage is a hi+lo FP16 pair (two planes) with a per-grid power-of-two
 *        scale; coefficients are split hi+lo too; three MFMA passes
 *        (hi*hi, hi*lo, lo*hi; lo*lo is below 2^-22 of the result) accumulate
 *        into one FP32 accumulator. */
enum { TCS_PREC_FP16 = 0, TCS_PREC_BF16 = 1, TCS_PREC_SPLIT = 2 };

enum {
    TCS_OK = 0,
    TCS_ERR_INVALID = 1,     /* bad argument (null pointer, size, stencil, n...) */
    TCS_ERR_UNSUPPORTED = 2, /* valid combination that this build does not provide */
    TCS_ERR_HIP = 3,         /* a HIP runtime call failed */
    TCS_ERR_ALLOC = 4        /* device or host allocation failed */
};

/* Problem parameters of the reference (common/param.hpp:1-13), same names and
 * defaults; tcs_params_from_env() overrides each from TCS_<NAME>.  TILE_SIZE is
 * the reference's bricks per block along x (1..15 in its sweep); the host
 * drivers apply it with tcs_problem_set_tile_size; 0 means "use the tuned
 * default". */
typedef struct tcs_params {
    double TOLERANCE;  /* 0.01 */
    double MIN_NUMBER; /* 1e-6 */
    int32_t RUN_TIMES; /* 500 */
    int32_t STEP_TIMES;/* 1 */
    int32_t SKIP_TIMES;/* 400 */
    int32_t TILE_SIZE; /* 0 (tuned) -- the reference's compile-time default is 1 */
} tcs_params;

/* Brick layout of one rows x cols grid.  Bricks are 16x16, stored brick-row-major,
 * each brick row-major (256 elements, 512 B in FP16): the reference's
 * brick_layout_2D map (layout16/host_box2d2r.cu:75-86).  Around the nbr x nbc
 * data bricks sits a one-brick ghost ring (all zero), like the reference's 16-wide
 * padding ring (layout16/host_box2d2r.cu:61-73, :136-141).  nbr/nbc are rounded up
 * to a multiple of `align` so that tiled kernels need no bounds checks. */
typedef struct tcs_layout {
    int32_t rows, cols;  /* logical grid, including the Dirichlet ring of width r */
    int32_t nbr, nbc;    /* data bricks (padded to a multiple of align) */
    int32_t sbr, sbc;    /* stored bricks = nbr + 2, nbc + 2 */
    int32_t align;       /* 1, 2, 4 or 8 */
    int32_t reserved;
    int64_t plane_elems; /* sbr * sbc * 256: elements of one storage plane */
} tcs_layout;

/* Everything a sweep needs, by value: no device allocations behind it. */
typedef struct tcs_problem {
    int32_t stencil, method, precision;
    int32_t radius;          /* 1 or 2 */
    int32_t tile_h, tile_w;  /* bricks per workgroup tile (0 = tuned default) */
    tcs_layout layout;
    float coeff[25];         /* 5x5 embedding, coeff[(di+2)*5 + (dj+2)], FP32 */
    uint16_t coeff_hi[25];   /* SPLIT: FP16 bits of hi(coeff * 2^coeff_exp) */
    uint16_t coeff_lo[25];   /* SPLIT: FP16 bits of lo(coeff * 2^coeff_exp) */
    int32_t coeff_exp;       /* SPLIT: coefficient scale exponent */
    int32_t wg_per_cu;       /* 0: one workgroup per tile; n > 0: a persistent grid
                                of n workgroups per CU that prefetch their next tile */
    int32_t fuse;            /* sweeps fused per launch by tcs_run (temporal blocking,
                                MFMA only): 1..1+16/radius; 0 or 1 = unfused */
    int32_t frag;            /* MFMA fragment edge: 16 (v_mfma_f32_16x16x32, one brick per
                                product; the default) or 32 (v_mfma_f32_32x32x16 over
                                2x2-brick blocks, FP16/BF16); set with tcs_problem_set_frag */
    int32_t coef_lo;         /* MFMA FP16/BF16 with 16x16 fragments: 1 = each FP32 tap c is
                                carried as hi = round(c) plus lo = round(c - hi) in the
                                storage format, a second MFMA pass per fragment on the same
                                data (taps exact to 2^-22 FP16 / 2^-16 BF16 instead of
                                2^-11 / 2^-8); 0 = rounded taps only.  Ignored by SPLIT
                                (which splits its taps anyway) and the direct method. */
} tcs_problem;

/* ---------------------------------------------------------- status, params */
const char *tcs_status_string(int status);
/* Message of the last failing call on this thread ("" if none). */
const char *tcs_last_error(void);
void tcs_params_default(tcs_params *p);
void tcs_params_from_env(tcs_params *p);

/* ------------------------------------------------------------- drop-in layer */
/* Replace `extern "C" void host_code(float *h_A, float *h_B, float *h_C, int N, int n)`
 * of layout16/host_box2d2r.cu:129 (and siblings layout16/host_box2d1r.cu,
 * layout16/host_star2d1r.cu, layout16/host_star2d2r.cu), declared by the caller at
 * common/main_2r.cu:10 / common/main_1r.cu:7.  One Jacobi sweep of the N x N FP32
 * row-major grid h_A with the n x n coefficients h_B (n = 2r+1, row stride n),
 * on the MFMA FP16 path; h_C receives the interior [r,N-r)^2 and, on the ring,
 * h_A's values (gold's convention, common/main_2r.cu:29).  Any N >= 1 (the
 * reference needs N % 16 == 0).  The kernel is launched RUN_TIMES times on the
 * same input and "[Time] Time used: %f ms" is printed, as the reference does
 * (layout16/host_box2d2r.cu:177-193).  On error it prints the message and
 * exits with status -1, like the reference's check_error (common/common.hpp:10-16). */
void host_star2d1r(float *h_A, float *h_B, float *h_C, int N, int n);
void host_star2d2r(float *h_A, float *h_B, float *h_C, int N, int n);
void host_box2d1r(float *h_A, float *h_B, float *h_C, int N, int n);
void host_box2d2r(float *h_A, float *h_B, float *h_C, int N, int n);

/* The same on the direct-FMA kernel: replaces host_code of baseline/host_box2d2r.cu:91
 * (and siblings baseline/host_{box2d1r,star2d1r,star2d2r}.cu). */
void host_direct_star2d1r(float *h_A, float *h_B, float *h_C, int N, int n);
void host_direct_star2d2r(float *h_A, float *h_B, float *h_C, int N, int n);
void host_direct_box2d1r(float *h_A, float *h_B, float *h_C, int N, int n);
void host_direct_box2d2r(float *h_A, float *h_B, float *h_C, int N, int n);

/* Status-returning generalisation of the drivers: rows x cols, any method and
 * precision, `sweeps` ping-pong sweeps, no printing, no exit.  The per-sweep
 * kernel time (mean over RUN_TIMES-SKIP_TIMES timed repetitions of the whole
 * sweep loop, divided by sweeps) goes to *ms_per_sweep when non-NULL. */
int tcs_host_sweep(int stencil, int method, int precision, const float *h_A, const float *h_B,
                   int n, float *h_C, int rows, int cols, int sweeps, const tcs_params *params,
                   float *ms_per_sweep);

/* ------------------------------------------------------------- problem setup */
int tcs_layout_init(tcs_layout *L, int rows, int cols, int align);
/* Builds the descriptor from the n x n coefficient block B (row stride n,
 * n = 2r+1 of the stencil).  Star stencils use only B's centre row and column,
 * like gold. */
int tcs_problem_init(tcs_problem *p, int stencil, int method, int precision, int rows, int cols,
                     const float *B, int n);
int tcs_radius(int stencil);
/* Choose the workgroup tile (bricks, rows x cols): 2x4, 4x4, 2x8, 4x8 or 8x8.
 * With 16x16 fragments single-sweep and fused kernels exist for all five
 * (FP16/BF16, with rounded or hi+lo taps, fuse large 4x8 problems -- at least 2048 8x8
 * tiles -- on 8x8 tiles of 8-wave workgroups, which measure faster at every
 * fused depth; SPLIT fuses 6 or more sweeps per launch of 4x8 problems on
 * 8-wave workgroups); a missing combination (32x32 fragments: see tcs_problem_set_frag) fails at
 * launch with TCS_ERR_UNSUPPORTED (tcs_kernel_name / tcs_fused_kernel_name
 * return NULL for it).  Re-derives the layout's alignment, so buffers packed
 * for another tile must be packed again. */
int tcs_problem_set_tile(tcs_problem *p, int tile_h, int tile_w);
/* Apply the reference's TILE_SIZE (common/param.hpp:11-13: bricks one thread
 * block walks along x on blocks one brick tall, layout16/host_box2d2r.cu:166-168;
 * swept over 1..15 by script/param.py:5).  Every value maps to an offered tile:
 * 1..5 -> 4 bricks wide, 6 and above -> 8 bricks wide, at the current tile
 * height where that shape is offered (else the nearest offered height).
 * tile_size <= 0 keeps the tuned default.  Brick results do not depend on the
 * tile, so any TILE_SIZE gives the same output bits. */
int tcs_problem_set_tile_size(tcs_problem *p, int tile_size);
/* Choose the MFMA fragment (16 or 32, see tcs_problem.frag) and re-derive the
 * default tile for it (32: 4x4, 2x8, 4x8 or 8x8 single-sweep tiles; 4x4, 4x8, 8x8
 * fused).  32 needs method MFMA and FP16 or BF16. */
int tcs_problem_set_frag(tcs_problem *p, int frag);
/* Set coef_lo (hi+lo coefficients, see tcs_problem.coef_lo) and re-derive the
 * default tile for it (8x8 where the grid allows: the hi+lo kernels run 4 waves
 * per SIMD).  Buffers packed for another tile must be packed again. */
int tcs_problem_set_coef_lo(tcs_problem *p, int on);
/* ------------------------------------------------------------ tuning cache */
/* Parameters the autotuner (python -m tcstencil_amd.autotune --write-cache)
 * measured fastest, kept in a text file per device: "device <arch> <CUs>", then
 * "entry <stencil> <method> <precision> <coef_lo> <batched> <bucket> <frag>
 * <tile_h> <tile_w> <wg_per_cu> <fuse>" lines.  This replaces the reference's
 * compile-time TILE_SIZE (common/param.hpp:11-13) swept offline
 * (script/param.py:5).  Tile, wg_per_cu and fuse never change a result; the
 * fragment (MFMA only) changes the grouping of the FP32 sums. */
typedef struct tcs_tuning {
    int32_t tile_h, tile_w, wg_per_cu, fuse, frag;
} tcs_tuning;
/* Size bucket of a rows x cols grid: the smaller side in bricks in [0,16),
 * [16,64), [64,256) or [256,inf) -> 0..3. */
int tcs_tune_bucket(int rows, int cols);
/* The cached parameters for problem p (single grid: batched = 0; a grid of a
 * tcs_batch: 1) from the file at `path` (NULL or "": $TCS_TUNE_CACHE).  Entries
 * apply only on the device their file names.  TCS_ERR_UNSUPPORTED: no file or
 * no matching entry; TCS_ERR_INVALID: unreadable or malformed file. */
int tcs_tune_lookup(const char *path, const tcs_problem *p, int batched, tcs_tuning *out);
/* Set p's tile, wg_per_cu and (MFMA) fragment and fuse from its cache entry; unchanged when
 * there is none.  Re-derives the layout, so do it before allocating buffers.
 * tcs_problem_init, tcs_problem_set_frag and tcs_problem_set_coef_lo apply the
 * single-grid entry of $TCS_TUNE_CACHE themselves when it is set (re-deriving
 * tile, wg_per_cu and fuse together; an explicit fragment is kept); batched grids
 * call this with batched = 1 before allocating their buffers (tcs_batch_create
 * takes each grid's descriptor, tile included, as given: the tile fixes the
 * grid's storage layout). */
int tcs_problem_apply_tuning(tcs_problem *p, const char *path, int batched);

/* Bytes of one storage buffer for this problem (2 B/elem; SPLIT: two planes). */
int64_t tcs_buffer_bytes(const tcs_problem *p);

/* --------------------------------------------------------- device transforms */
/* Row-major source (FP32 or FP64, leading dimension ld elements) -> brick storage
 * in the problem's precision, ghost ring and padding zeroed.  scale_exp: the
 * source is multiplied by 2^scale_exp (SPLIT's per-grid scaling; 0 otherwise). */
int tcs_pack_f32(const tcs_problem *p, const float *d_src, int64_t ld, void *d_dst, int scale_exp,
                 void *stream);
int tcs_pack_f64(const tcs_problem *p, const double *d_src, int64_t ld, void *d_dst, int scale_exp,
                 void *stream);
/* Brick storage -> row-major FP32 (times 2^-scale_exp).  When d_ring is non-NULL
 * (row-major FP32 with the same ld), points outside the interior are copied from
 * it instead, so the ring keeps the caller's exact input values. */
int tcs_unpack_f32(const tcs_problem *p, const void *d_src, float *d_dst, int64_t ld,
                   int scale_exp, const float *d_ring, void *stream);
/* Zero a whole storage buffer (a fill kernel, safe under graph capture). */
int tcs_clear(const tcs_problem *p, void *d_buf, void *stream);
/* Device-to-device streaming copy (bytes % 16 == 0): the measured bandwidth
 * ceiling reported next to the HBM spec figure. */
int tcs_copy(void *d_dst, const void *d_src, int64_t bytes, void *stream);
/* max |x| over a row-major FP32/FP64 grid into *d_out (one double, device). */
int tcs_maxabs_f32(const float *d_src, int rows, int cols, int64_t ld, double *d_out, void *stream);
int tcs_maxabs_f64(const double *d_src, int rows, int cols, int64_t ld, double *d_out, void *stream);
/* Per-grid scale exponent for SPLIT storage given max|x| (host helper). */
int tcs_split_scale_exp(double maxabs);
/* The same for a run of `sweeps` sweeps of problem p: taps whose absolute sum S
 * exceeds 1 can grow max|x| by S per sweep, so the exponent is that of
 * max|x| * S^sweeps (growth capped at 2^60), keeping every stored value below
 * binary16's largest finite value.  With S <= 1 it equals tcs_split_scale_exp. */
int tcs_split_scale_exp_for(const tcs_problem *p, double maxabs, int sweeps);
/* Per-grid scale exponent e for a run of `sweeps` sweeps of problem p in its
 * precision, given max|x| of the input (the pack and unpack calls multiply by
 * 2^e and 2^-e, exact powers of two):
 *   SPLIT  tcs_split_scale_exp_for;
 *   FP16   max|x| * S^sweeps brought to [2^14, 2^15) (S = the taps' absolute sum
 *          when above 1, growth capped at 2^60): values beyond binary16's
 *          65504 stay finite and values below its normal range (6.1e-5) keep
 *          full precision; results for values already in the normal range are
 *          those of e = 0 (the rounding is scale-free there);
 *   BF16   0 (the FP32 exponent range).
 * The host drivers (tcs_host_sweep, host_<stencil>) apply it; device callers
 * may pass it to tcs_pack_f32 / tcs_unpack_f32 / tcs_run_f32.  The reference
 * converts to binary16 unscaled (common/common.hpp:146-160, array_float2half),
 * so |x| > 65504 becomes Inf there.  No counterpart otherwise. */
int tcs_scale_exp_for(const tcs_problem *p, double maxabs, int sweeps);

/* ----------------------------------------------------------------- sweeping */
/* One sweep d_in -> d_out (distinct storage buffers of this problem; d_out's
 * ghost ring must be zero, e.g. from tcs_clear or a previous pack). */
int tcs_sweep(const tcs_problem *p, const void *d_in, void *d_out, void *stream);
/* `sweeps` ping-pong sweeps between d_buf0 and d_buf1; the result is in
 * buf[sweeps % 2].  Unfused, sweep s reads buf[s % 2] and writes buf[(s+1) % 2].
 * With p->fuse = k > 1 the sweeps run as temporally blocked launches of at most
 * k sweeps each (LDS-resident, bitwise equal to unfused sweeps), balanced so
 * that the launch count has the parity of `sweeps`. */
int tcs_run(const tcs_problem *p, void *d_buf0, void *d_buf1, int sweeps, void *stream);
/* Row-major FP32 grid in (d_src) -> `sweeps` sweeps -> row-major FP32 grid out
 * (d_dst; both rows x cols with leading dimension ld, distinct): bitwise the
 * result of tcs_pack_f32(d_src -> d_buf0, scale_exp), tcs_run(d_buf0, d_buf1,
 * sweeps) and tcs_unpack_f32(result -> d_dst, scale_exp, d_ring = d_src).  For
 * MFMA problems with 16x16 fragments (any of the five tiles; the two edge
 * sweeps run one tile per workgroup, maybe on a smaller tile that divides the
 * layout), the pack is fused into the first sweep and the unpack
 * into the last, so one sweep is one kernel reading and writing FP32 and
 * touching no storage buffer (d_buf0/d_buf1 may then be NULL); otherwise the
 * three steps run.  The storage buffers' prior contents do not matter: their
 * ghost rings are cleared here (a fill kernel, safe under graph capture) and
 * every data brick is written before it is read.
 * This is the GPU-resident edge of the reference's host driver: its host-side
 * conversion, padding and brick layout (layout16/host_box2d2r.cu:131-141) and
 * the reverse (:196-204), run on the device and fused into the sweeps. */
int tcs_run_f32(const tcs_problem *p, const float *d_src, float *d_dst, int64_t ld, void *d_buf0, void *d_buf1,
                int sweeps, int scale_exp, void *stream);
/* Capture tcs_run into a hipGraphExec (returned as void*) and replay it. */
int tcs_graph_create(const tcs_problem *p, void *d_buf0, void *d_buf1, int sweeps, void *stream,
                     void **graph_exec);
int tcs_graph_launch(void *graph_exec, void *stream);
int tcs_graph_destroy(void *graph_exec);

/* ---------------------------------------------------- batched uneven grids */
/* Many independent grids (own size, stencil, coefficients and buffers) swept
 * together.  Problems that share a tile shape, fragment (frag), sweeps per launch
 * (fuse) and wg_per_cu form a group; a group advances all its problems with one
 * launch per sweep, or per `fuse` sweeps fused in LDS (balanced like tcs_run),
 * on a grid of one workgroup per tile or, with wg_per_cu > 0, a persistent grid
 * of wg_per_cu workgroups per CU.  A sweep of the whole batch is a handful of
 * launches instead of one per grid.  All problems: MFMA, the same precision
 * (FP16 or BF16); tiles 2x4, 4x4 or 4x8 (16x16 fragments), 4x4 or 4x8 (32x32);
 * 8x8 problems run on 4x8 tiles (same storage layout, same results).
 * d_buf0[i] / d_buf1[i] are problem i's two storage buffers (d_buf1 ghost ring
 * zero); after tcs_batch_run the results are in buf[sweeps % 2].  The batch owns
 * device copies of the descriptors (allocated here, freed by tcs_batch_destroy).
 * tcs_batch_graph_create captures each group's launches as an independent
 * chain, so groups run concurrently on replay; tcs_batch_run issues them in
 * order on one stream.  tcs_batch_launches_per_sweep returns the group count. */
typedef struct tcs_batch tcs_batch;
int tcs_batch_create(tcs_batch **b, int nprob, const tcs_problem *probs, void *const *d_buf0,
                     void *const *d_buf1);
int tcs_batch_run(tcs_batch *b, int sweeps, void *stream);
int tcs_batch_graph_create(tcs_batch *b, int sweeps, void *stream, void **graph_exec);
int tcs_batch_destroy(tcs_batch *b);
int64_t tcs_batch_interior_points(const tcs_batch *b);
int tcs_batch_launches_per_sweep(const tcs_batch *b);

/* ------------------------------------------------ FP64 oracle and error norms */
/* One FP64 sweep on row-major grids (ld = cols), gold's loop and accumulation
 * order with no fused multiply-add (common/gold_*.cpp); the ring of d_out is not
 * written.  Used only as the accuracy oracle, never as a product sweep. */
int tcs_oracle_sweep_f64(int stencil, const double *d_in, const double *B, int n, double *d_out,
                         int rows, int cols, void *stream);
/* Norms over the interior [r,rows-r) x [r,cols-r) of row-major FP32 output vs
 * FP64 reference, reduced in FP64 (deterministic two-pass):
 * out[0]=max|o-ref|, out[1]=||o-ref||_2/||ref||_2, out[2]=RMS abs error
 * (checkError2D, common/common.hpp:38-64), out[3]=||ref||_2.  Host result. */
int tcs_error_norms(const float *d_out, const double *d_ref, int rows, int cols, int r,
                    double *h_out4, void *stream);

/* No counterpart in the reference: tcs_error_norms for callers that cannot
 * allocate (code under hipGraph capture, serving loops, harnesses that forbid
 * hipMalloc).  The same two-pass FP64 reduction and the same four results, but
 * the partials go to a static 48 KiB device scratch array of the library (one
 * per GPU, held until the process ends) and the results to d_out4 (4 doubles,
 * device memory); nothing is allocated, copied to the host or synchronised.
 * Any grid size (the partial pass uses at most 2048 blocks).  Calls that
 * overlap on one GPU share that GPU's scratch and race: serialise them. */
int tcs_error_norms_noalloc(const float *d_out, const double *d_ref, int rows, int cols, int r,
                            double *d_out4, void *stream);

/* ------------------------------------------------------------- generators */
/* Counter-based (SplitMix64) fills, identical on every launch shape.
 * uniform: element (i,j) = (T)(lo + (hi-lo) * u), u = top 24 bits of
 *          mix64(key + (i*cols+j)*gamma) / 2^24, key = mix64(seed ^ stream*C).
 * sinusoid: sum of 4 sinusoids (integer wavenumbers 1..64 per axis, random
 *          phase and amplitude, drawn from seed) + uniform noise in
 *          [-noise, noise). */
int tcs_gen_uniform_f32(float *d_dst, int rows, int cols, uint64_t seed, uint64_t stream_id,
                        float lo, float hi, void *stream);
int tcs_gen_uniform_f64(double *d_dst, int rows, int cols, uint64_t seed, uint64_t stream_id,
                        double lo, double hi, void *stream);
int tcs_gen_sinusoid_f64(double *d_dst, int rows, int cols, uint64_t seed, double noise,
                         void *stream);
/* Round a row-major FP64/FP32 grid in place to FP16 (or BF16) values, kept in
 * the wider type, so that the FP64 oracle sees exactly the FP16 input. */
int tcs_round_f64(double *d_buf, int64_t count, int precision, void *stream);
int tcs_round_f32(float *d_buf, int64_t count, int precision, void *stream);
/* Host: coefficients U[0,1) normalised to sum 1 (in FP64, then rounded to FP32)
 * into the n x n block B (star stencils: only centre row and column non-zero). */
int tcs_gen_coeffs(int stencil, uint64_t seed, float *B, int n);

/* ------------------------------------------------------------ introspection */
/* Kernel name (as rocprofv3 reports it) used for this problem, or NULL. */
const char *tcs_kernel_name(const tcs_problem *p);
/* Kernel name of the fused (temporally blocked) launch of p->fuse sweeps for
 * this problem, or NULL. */
const char *tcs_fused_kernel_name(const tcs_problem *p);
/* Kernel name of tcs_run_f32's sweep with the layout transforms fused in (mode 1:
 * row-major FP32 in, storage out; 2: storage in, row-major FP32 out; 3: both, a
 * one-sweep run), or NULL when tcs_run_f32 runs the three steps instead.  Its
 * tile may differ from the problem's (results do not depend on the tile). */
const char *tcs_edge_kernel_name(const tcs_problem *p, int mode);
/* MFMA instructions issued per 16x16 brick, or per 2x2-brick block with 32x32
 * fragments (0 for the direct method). */
int tcs_mfma_per_brick(const tcs_problem *p);
/* MFMA flops issued per 16x16 output brick by one sweep (0 for the direct method). */
int64_t tcs_mfma_flops_per_brick(const tcs_problem *p);

#ifdef __cplusplus
}
#endif
#endif /* TCSTENCIL_H */
