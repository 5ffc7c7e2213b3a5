/* Test infrastructure: exports the reference's single `gold` symbol
 * (common/gold_<stencil>.cpp:3-4) for one stencil, forwarding to the oracle's
 * gold_<stencil>, so a main_Xr-style driver links against it unchanged.
 * Build with -DORACLE_SHIM_TARGET=gold_<stencil>. */
#include "gold.h"

#ifndef ORACLE_SHIM_TARGET
#error "define ORACLE_SHIM_TARGET"
#endif

void gold(float *h_A, float *h_B, float *h_C, int N, int n) { ORACLE_SHIM_TARGET(h_A, h_B, h_C, N, n); }
