/*
 * CPU oracle for the 2D stencil path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only to check results (or as the timed CPU baseline); the product
 * library (tcstencil_amd/) never links, loads or calls it.
 *
 * Every function restates the reference's CPU gold loop with the same loop order
 * and the same left-to-right accumulation order (file:line cited per function).
 * Build with -ffp-contract=off so that no multiply-add is fused: the FP32 results
 * are then bitwise identical to the reference's gold_*.cpp compiled the same way
 * (pinned by tests/golden fixtures made from the reference sources).
 *
 * Parity pins (see tests/test_oracle.py):
 *   gold_* (FP32)       bitwise vs fixtures from the reference gold_*.cpp
 *   oracle_sweep_f64    vs scipy.signal.correlate2d (independent brute force),
 *                       constant and linear-ramp closed forms
 *   oracle_error_norms  closed forms on synthetic differences
 *   oracle_check_*      closed forms (reference common/common.hpp:38-92)
 *   oracle_mix64 / gen  bitwise vs the product's own generator on the GPU,
 *                       and the published SplitMix64 test vector
 */
#ifndef TCS_ORACLE_GOLD_H
#define TCS_ORACLE_GOLD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Stencil ids (same numbering as the product's include/tcstencil.h). */
enum { ORACLE_STAR2D1R = 0, ORACLE_STAR2D2R = 1, ORACLE_BOX2D1R = 2, ORACLE_BOX2D2R = 3 };

/* Reference-signature gold loops: N x N row-major FP32, B is (2r+1)x(2r+1) with
 * row stride 2r+1 (the reference ignores n).  Only C[r..N-r)^2 is written. */
void gold_star2d1r(float *h_A, float *h_B, float *h_C, int N, int n);
void gold_star2d2r(float *h_A, float *h_B, float *h_C, int N, int n);
void gold_box2d1r(float *h_A, float *h_B, float *h_C, int N, int n);
void gold_box2d2r(float *h_A, float *h_B, float *h_C, int N, int n);

/* FP64 twins (same loop and accumulation order). */
void gold_star2d1r_f64(double *h_A, double *h_B, double *h_C, int N, int n);
void gold_star2d2r_f64(double *h_A, double *h_B, double *h_C, int N, int n);
void gold_box2d1r_f64(double *h_A, double *h_B, double *h_C, int N, int n);
void gold_box2d2r_f64(double *h_A, double *h_B, double *h_C, int N, int n);

/* Rectangular generalisation: rows x cols row-major, same orders as above.
 * Returns 0 on success, -1 on a bad stencil id. */
int oracle_sweep_f32(int stencil, const float *A, const float *B, float *C, int rows, int cols);
int oracle_sweep_f64(int stencil, const double *A, const double *B, double *C, int rows, int cols);

/* `sweeps` Jacobi sweeps with ping-pong; A is overwritten with the result
 * (tmp must hold rows*cols doubles).  The ring of width r keeps A's values. */
int oracle_sweeps_f64(int stencil, double *A, const double *B, double *tmp, int rows, int cols, int sweeps);

int oracle_radius(int stencil);

/* North-star error norms over the interior [r,rows-r) x [r,cols-r):
 * out[0]=max|o-ref|, out[1]=||o-ref||_2/||ref||_2, out[2]=RMS abs, out[3]=||ref||_2. */
void oracle_error_norms(const float *o, const double *ref, int rows, int cols, int r, double *out);

/* Restatements of common/common.hpp:38-64 (checkError2D) and :66-92
 * (checkRelativeError2D), without the printing. */
double oracle_check_error_2d(int N, const float *output, const float *reference,
                             int y_lb, int y_ub, int x_lb, int x_ub);
double oracle_check_relative_error_2d(int N, const float *output, const float *reference,
                                      int y_lb, int y_ub, int x_lb, int x_ub);
/* Restatement of common/common.hpp:94-124 (checkRelativeError3D), without the printing. */
double oracle_check_relative_error_3d(int N, const float *output, const float *reference, int z_lb,
                                      int z_ub, int y_lb, int y_ub, int x_lb, int x_ub);

/* SplitMix64 finaliser and the counter-based generators the product uses. */
uint64_t oracle_mix64(uint64_t z);
void oracle_gen_uniform_f32(float *dst, int rows, int cols, uint64_t seed, uint64_t stream,
                            float lo, float hi);

#ifdef __cplusplus
}
#endif
#endif
