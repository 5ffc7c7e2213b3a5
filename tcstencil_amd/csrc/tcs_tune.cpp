// Tuning cache: the parameters the autotuner measured fastest (MFMA fragment,
// tile, persistent grid, sweeps per launch) per device, stencil, method,
// precision, coefficient form, single or batched grid and size bucket, kept in
// a text file that problem set-up reads.
// The reference fixes its tile at compile time (common/param.hpp:11-13) and
// sweeps it offline (script/param.py:5, script/nvcc_compile.py:55-67); here the
// sweep's result is a file the library consults at run time.
//
// File format (one record per line, '#' starts a comment):
//   device <gfx arch> <compute units>        (first record; "device any 0" matches every device)
//   entry <stencil> <method> <precision> <coef_lo> <batched> <bucket> <frag> <tile_h> <tile_w> <wg_per_cu> <fuse>
// with the buckets of tcs_tune_bucket below.  Tile, persistent grid and sweeps
// per launch never change a result (brick results do not depend on them, fused
// sweeps are bitwise the unfused ones), so a tuned problem computes exactly
// what the default one does with the same fragment; the 16x16 and 32x32
// fragments sum a point's products in different groupings (FP32 rounding
// level).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "tcs_internal.hpp"

using namespace tcs;

namespace {

const char *const kStencilNames[] = {"star2d1r", "star2d2r", "box2d1r", "box2d2r"};
const char *const kMethodNames[] = {"mfma", "direct", "direct_dot2"};
const char *const kPrecNames[] = {"fp16", "bf16", "split"};

int index_of(const char *s, const char *const *names, int n) {
    for (int i = 0; i < n; ++i)
        if (!strcmp(s, names[i])) return i;
    return -1;
}

// Architecture name (up to the first ':') and CU count of the current device;
// false when no device is usable.
bool current_device(char *arch, size_t len, int *cus) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    snprintf(arch, len, "%s", prop.gcnArchName);
    char *c = strchr(arch, ':');
    if (c) *c = '\0';
    *cus = prop.multiProcessorCount;
    return true;
}

}  // namespace

// Size bucket of a grid: the smaller side in bricks, [0,16), [16,64), [64,256),
// [256, inf) -> 0..3.
extern "C" int tcs_tune_bucket(int rows, int cols) {
    const int m = ((rows < cols ? rows : cols) + 15) / 16;
    return m < 16 ? 0 : m < 64 ? 1 : m < 256 ? 2 : 3;
}

extern "C" int tcs_tune_lookup(const char *path, const tcs_problem *p, int batched, tcs_tuning *out) {
    if (!p || !out) return set_error(TCS_ERR_INVALID, "tune_lookup: null argument");
    if (!path || !*path) path = getenv("TCS_TUNE_CACHE");
    if (!path || !*path) return set_error(TCS_ERR_UNSUPPORTED, "tune_lookup: no cache file");
    FILE *f = fopen(path, "r");
    if (!f) return set_error(TCS_ERR_INVALID, "tune_lookup: cannot open %s", path);
    const int bucket = tcs_tune_bucket(p->layout.rows, p->layout.cols);
    const int lo = has_coef_lo(kernel_prec(p)) ? 1 : 0;
    char line[512];
    bool device_ok = false, seen_device = false, found = false;
    int lineno = 0;
    while (fgets(line, sizeof line, f)) {
        ++lineno;
        char *hash = strchr(line, '#');
        if (hash) *hash = '\0';
        char kind[16] = "";
        if (sscanf(line, "%15s", kind) != 1) continue;
        if (!strcmp(kind, "device")) {
            char arch[64] = "";
            int cus = 0;
            if (sscanf(line, "%*s %63s %d", arch, &cus) != 2) {
                fclose(f);
                return set_error(TCS_ERR_INVALID, "tune cache %s:%d: bad device record", path, lineno);
            }
            seen_device = true;
            if (!strcmp(arch, "any")) {
                device_ok = true;
            } else {
                char cur[64];
                int cur_cus = 0;
                device_ok = current_device(cur, sizeof cur, &cur_cus) && !strcmp(cur, arch) && cur_cus == cus;
            }
            continue;
        }
        if (strcmp(kind, "entry")) {
            fclose(f);
            return set_error(TCS_ERR_INVALID, "tune cache %s:%d: unknown record '%s'", path, lineno, kind);
        }
        char st[16], me[16], pr[16];
        int clo, ba, bu;
        tcs_tuning t;
        if (sscanf(line, "%*s %15s %15s %15s %d %d %d %d %d %d %d %d", st, me, pr, &clo, &ba, &bu, &t.frag, &t.tile_h,
                   &t.tile_w, &t.wg_per_cu, &t.fuse) != 11 ||
            (t.frag != 16 && t.frag != 32)) {
            fclose(f);
            return set_error(TCS_ERR_INVALID, "tune cache %s:%d: bad entry", path, lineno);
        }
        if (!seen_device) {
            fclose(f);
            return set_error(TCS_ERR_INVALID, "tune cache %s:%d: entry before the device record", path, lineno);
        }
        if (!device_ok || found) continue;
        if (index_of(st, kStencilNames, 4) == p->stencil && index_of(me, kMethodNames, 3) == p->method &&
            index_of(pr, kPrecNames, 3) == p->precision && clo == lo && ba == (batched ? 1 : 0) && bu == bucket) {
            *out = t;
            found = true;  // the first matching entry wins; keep checking the rest of the file's syntax
        }
    }
    fclose(f);
    if (!found) return set_error(TCS_ERR_UNSUPPORTED, "tune cache %s: no entry for this problem and device", path);
    return TCS_OK;
}

// The entry for p, applied unless want_frag is non-zero and differs from the
// entry's fragment (an explicit fragment choice is kept).
int tcs::apply_tuning(tcs_problem *p, const char *path, int batched, int want_frag) {
    if (!p) return set_error(TCS_ERR_INVALID, "apply_tuning: null problem");
    tcs_tuning t;
    const int st = tcs_tune_lookup(path, p, batched, &t);
    if (st == TCS_ERR_UNSUPPORTED) return TCS_OK;  // no cache or no entry: the defaults stay
    if (st) return st;
    if (want_frag && p->method == TCS_METHOD_MFMA && t.frag != want_frag) return TCS_OK;
    if (t.wg_per_cu < 0 || t.fuse < 0) return set_error(TCS_ERR_INVALID, "apply_tuning: bad cached values");
    tcs_problem q = *p;
    if (p->method == TCS_METHOD_MFMA) q.frag = t.frag;
    TCS_TRY(tcs_problem_set_tile(&q, t.tile_h, t.tile_w));
    q.wg_per_cu = t.wg_per_cu;
    if (p->method == TCS_METHOD_MFMA) q.fuse = t.fuse;
    *p = q;
    return TCS_OK;
}

extern "C" int tcs_problem_apply_tuning(tcs_problem *p, const char *path, int batched) {
    return apply_tuning(p, path, batched, 0);
}
