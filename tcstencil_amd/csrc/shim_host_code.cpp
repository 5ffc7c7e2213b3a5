// Link-time drop-in for the reference's single `host_code` symbol: the
// reference picks one stencil and method per binary at link time
// (script/param.py:14-33), so each shim library exports `host_code` and
// forwards to one driver of libtcstencil.  Build with -DTCS_SHIM_TARGET=<driver>.
#include "../../include/tcstencil.h"

#ifndef TCS_SHIM_TARGET
#error "define TCS_SHIM_TARGET, e.g. -DTCS_SHIM_TARGET=host_box2d2r"
#endif

extern "C" void host_code(float *h_A, float *h_B, float *h_C, int N, int n) {
    TCS_SHIM_TARGET(h_A, h_B, h_C, N, n);
}
