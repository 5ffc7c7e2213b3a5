// FP64 accuracy oracle on the GPU and the FP64 error-norm reduction.  These
// are the only places the library computes in FP64; no product sweep does.
//
// The sweep follows gold's loop and accumulation order (common/gold_star2d1r.cpp:10-12,
// common/gold_star2d2r.cpp:10-14, common/gold_box2d1r.cpp:11-13,
// common/gold_box2d2r.cpp:11-13) with contraction off, so at any size it is
// bitwise equal to the CPU gold loop run in FP64 with -ffp-contract=off.
#include <math.h>

#include <mutex>

#include "tcs_internal.hpp"

#pragma clang fp contract(off)

namespace tcs {

struct Coef64 {
    double b[25];  // n x n, row stride n
};

template <int STENCIL>
__global__ void tcs_oracle_sweep_f64_kernel(const double *__restrict__ A, Coef64 cb, double *__restrict__ C,
                                            int rows, int cols) {
    constexpr int r = (STENCIL == TCS_STAR2D1R || STENCIL == TCS_BOX2D1R) ? 1 : 2;
    constexpr int n = 2 * r + 1;
    const int64_t iw = cols - 2 * r;
    const int64_t total = (int64_t)(rows - 2 * r) * iw;
    const double *B = cb.b;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = r + t / iw, j = r + t % iw;
        const int64_t ld = cols;
        double v;
        if constexpr (STENCIL == TCS_STAR2D1R) {
            v = A[(i - 1) * ld + j] * B[0 * 3 + 1] + A[i * ld + j] * B[1 * 3 + 1] + A[(i + 1) * ld + j] * B[2 * 3 + 1] +
                A[i * ld + j - 1] * B[1 * 3 + 0] + A[i * ld + j + 1] * B[1 * 3 + 2];
        } else if constexpr (STENCIL == TCS_STAR2D2R) {
            v = A[(i - 2) * ld + j] * B[0 * 5 + 2] + A[(i - 1) * ld + j] * B[1 * 5 + 2] + A[i * ld + j] * B[2 * 5 + 2] +
                A[(i + 1) * ld + j] * B[3 * 5 + 2] + A[(i + 2) * ld + j] * B[4 * 5 + 2] +
                A[i * ld + j - 2] * B[2 * 5 + 0] + A[i * ld + j - 1] * B[2 * 5 + 1] +
                A[i * ld + j + 1] * B[2 * 5 + 3] + A[i * ld + j + 2] * B[2 * 5 + 4];
        } else {
            v = 0.0;
#pragma unroll
            for (int k = 0; k < n * n; ++k) v += A[(i - r + k / n) * ld + (j - r + k % n)] * B[k];
        }
        C[i * ld + j] = v;
    }
}

// Pass 1: per-block partials of (max|d|, sum d^2, sum ref^2) in a fixed order.
__global__ void tcs_norm_partial_kernel(const float *__restrict__ o, const double *__restrict__ ref, int rows,
                                        int cols, int r, double *__restrict__ part) {
    __shared__ double smax[256], sse[256], ssr[256];
    const int64_t iw = cols - 2 * r;
    const int64_t total = (int64_t)(rows - 2 * r) * iw;
    double mx = 0.0, se = 0.0, sr = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = r + t / iw, j = r + t % iw;
        const double rv = ref[i * cols + j];
        const double d = (double)o[i * cols + j] - rv;
        const double ad = fabs(d);
        if (ad > mx || ad != ad) mx = ad;
        se += d * d;
        sr += rv * rv;
    }
    smax[threadIdx.x] = mx;
    sse[threadIdx.x] = se;
    ssr[threadIdx.x] = sr;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double a = smax[threadIdx.x], b = smax[threadIdx.x + s];
            smax[threadIdx.x] = (b > a || b != b) ? b : a;
            sse[threadIdx.x] += sse[threadIdx.x + s];
            ssr[threadIdx.x] += ssr[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = smax[0];
        part[3 * blockIdx.x + 1] = sse[0];
        part[3 * blockIdx.x + 2] = ssr[0];
    }
}

// Pass 2: one thread folds the partials in block order.
__global__ void tcs_norm_final_kernel(const double *__restrict__ part, int nblocks, int64_t count,
                                      double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double mx = 0.0, se = 0.0, sr = 0.0;
    for (int b = 0; b < nblocks; ++b) {
        const double m = part[3 * b];
        if (m > mx || m != m) mx = m;
        se += part[3 * b + 1];
        sr += part[3 * b + 2];
    }
    out[0] = mx;
    out[1] = sr > 0.0 ? sqrt(se) / sqrt(sr) : sqrt(se);
    out[2] = count > 0 ? sqrt(se / (double)count) : 0.0;
    out[3] = sqrt(sr);
}

// Partial blocks of the norm reduction: at most this many (the grid of the
// partial kernel is capped here whatever the grid size).
constexpr int kNormMaxBlocks = 2048;

// Scratch of tcs_error_norms_noalloc: the per-block partials, one copy per GPU
// for the life of the process (48 KiB).  Host code never uses this name as an
// address: it takes the device address from hipGetSymbolAddress (norm_scratch).
static __device__ __align__(256) double scratch[3 * kNormMaxBlocks];

}  // namespace tcs

using namespace tcs;

namespace {

// Device address of `scratch` on the current device, looked up once per device
// (hipGetSymbolAddress with that device current) and kept.
int norm_scratch(double **out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static double *addr[kMaxDevices] = {};
    int dev = 0;
    TCS_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(TCS_ERR_UNSUPPORTED, "error_norms_noalloc: device %d", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
        addr[dev] = static_cast<double *>(p);
    }
    *out = addr[dev];
    return TCS_OK;
}

// Blocks of the partial pass for an interior of `total` points.
int norm_blocks(int64_t total) {
    const int64_t nb = (total + 255) / 256;
    return nb > kNormMaxBlocks ? kNormMaxBlocks : (int)nb;
}

}  // namespace

extern "C" int tcs_oracle_sweep_f64(int stencil, const double *d_in, const double *B, int n, double *d_out,
                                    int rows, int cols, void *stream) {
    const int r = tcs_radius(stencil);
    if (r < 0) return set_error(TCS_ERR_INVALID, "oracle: bad stencil %d", stencil);
    if (!d_in || !d_out || !B) return set_error(TCS_ERR_INVALID, "oracle: null pointer");
    if (n != 2 * r + 1) return set_error(TCS_ERR_INVALID, "oracle: n must be %d", 2 * r + 1);
    if (rows < 1 || cols < 1) return set_error(TCS_ERR_INVALID, "oracle: bad shape");
    if (rows <= 2 * r || cols <= 2 * r) return TCS_OK;  // no interior
    Coef64 cb;
    for (int k = 0; k < 25; ++k) cb.b[k] = k < n * n ? B[k] : 0.0;
    const int64_t total = (int64_t)(rows - 2 * r) * (cols - 2 * r);
    int64_t g = (total + 255) / 256;
    if (g > 65536) g = 65536;
    hipStream_t s = (hipStream_t)stream;
    switch (stencil) {
    case TCS_STAR2D1R: hipLaunchKernelGGL(tcs_oracle_sweep_f64_kernel<TCS_STAR2D1R>, dim3((unsigned)g), dim3(256), 0, s, d_in, cb, d_out, rows, cols); break;
    case TCS_STAR2D2R: hipLaunchKernelGGL(tcs_oracle_sweep_f64_kernel<TCS_STAR2D2R>, dim3((unsigned)g), dim3(256), 0, s, d_in, cb, d_out, rows, cols); break;
    case TCS_BOX2D1R: hipLaunchKernelGGL(tcs_oracle_sweep_f64_kernel<TCS_BOX2D1R>, dim3((unsigned)g), dim3(256), 0, s, d_in, cb, d_out, rows, cols); break;
    default: hipLaunchKernelGGL(tcs_oracle_sweep_f64_kernel<TCS_BOX2D2R>, dim3((unsigned)g), dim3(256), 0, s, d_in, cb, d_out, rows, cols); break;
    }
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}

extern "C" int tcs_error_norms(const float *d_out, const double *d_ref, int rows, int cols, int r, double *h_out4,
                               void *stream) {
    if (!d_out || !d_ref || !h_out4) return set_error(TCS_ERR_INVALID, "error_norms: null pointer");
    if (r < 0 || rows <= 2 * r || cols <= 2 * r) return set_error(TCS_ERR_INVALID, "error_norms: no interior");
    const int64_t total = (int64_t)(rows - 2 * r) * (cols - 2 * r);
    const int nb = norm_blocks(total);
    hipStream_t s = (hipStream_t)stream;
    double *buf = nullptr;
    if (hipMalloc(&buf, sizeof(double) * (3 * (size_t)nb + 4)) != hipSuccess)
        return set_error(TCS_ERR_ALLOC, "error_norms: hipMalloc failed");
    hipLaunchKernelGGL(tcs_norm_partial_kernel, dim3(nb), dim3(256), 0, s, d_out, d_ref, rows, cols, r, buf);
    hipLaunchKernelGGL(tcs_norm_final_kernel, dim3(1), dim3(64), 0, s, buf, nb, total, buf + 3 * nb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(h_out4, buf + 3 * nb, 4 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(buf);
    if (e != hipSuccess) return hip_error(e, "error_norms");
    return TCS_OK;
}

// The same two passes with the partials in the library's static scratch and
// the four results written to the caller's device memory: no allocation, no
// copy to the host and no synchronisation, so the call can be captured in a
// graph or issued from a loop that must not allocate.  Every call on a GPU
// uses that GPU's one scratch array: calls that overlap on one GPU (other
// streams, other threads) race and must be serialised by the caller.  The
// partial pass writes every partial the final pass reads, so nothing needs
// clearing between calls.
extern "C" int tcs_error_norms_noalloc(const float *d_out, const double *d_ref, int rows, int cols, int r,
                                       double *d_out4, void *stream) {
    if (!d_out || !d_ref || !d_out4) return set_error(TCS_ERR_INVALID, "error_norms_noalloc: null pointer");
    if (r < 0 || rows <= 2 * r || cols <= 2 * r)
        return set_error(TCS_ERR_INVALID, "error_norms_noalloc: no interior");
    const int64_t total = (int64_t)(rows - 2 * r) * (cols - 2 * r);
    const int nb = norm_blocks(total);
    if (nb > kNormMaxBlocks)
        return set_error(TCS_ERR_INVALID, "error_norms_noalloc: %d partial blocks exceed the scratch (%d)", nb,
                         kNormMaxBlocks);
    double *part = nullptr;
    TCS_TRY(norm_scratch(&part));
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(tcs_norm_partial_kernel, dim3(nb), dim3(256), 0, s, d_out, d_ref, rows, cols, r, part);
    hipLaunchKernelGGL(tcs_norm_final_kernel, dim3(1), dim3(64), 0, s, part, nb, total, d_out4);
    TCS_HIP_CHECK(hipGetLastError());
    return TCS_OK;
}
