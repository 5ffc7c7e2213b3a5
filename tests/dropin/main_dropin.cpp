// A driver in the shape of the reference's common/main_2r.cu / main_1r.cu,
// linked (like the reference's) against exactly one `host_code` and one `gold`:
// here the host_code shim of libtcstencil and the oracle's gold shim.  Unlike
// the reference it fills h_A with the reference's own random formula
// (common/common.hpp:20, commented out there) and runs the error checks that
// the reference leaves commented out (common/main_2r.cu:32-41).
//   usage: main_<method>_<stencil> N n
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

extern "C" void host_code(float *h_A, float *h_B, float *h_C, int N, int n);
extern "C" void gold(float *h_A, float *h_B, float *h_C, int N, int n);
extern "C" double oracle_check_error_2d(int N, const float *output, const float *reference, int y_lb, int y_ub,
                                        int x_lb, int x_ub);
extern "C" double oracle_check_relative_error_2d(int N, const float *output, const float *reference, int y_lb,
                                                 int y_ub, int x_lb, int x_ub);

// common/main_1r.cu:10-16 and common/main_2r.cu:13-19.
static void init_B(float *B, int n) {
    for (int i = 0; i < n * n; i++) B[i] = (float)i;
    B[(n * n) / 2] = n == 3 ? -40.0f : -290.0f;
}

int main(int argc, char **argv) {
    const int N = argc > 1 ? atoi(argv[1]) : 256;
    const int n = argc > 2 ? atoi(argv[2]) : 5;
    const int r = n / 2;
    srand(1);
    float *h_A = new float[(size_t)N * N];
    for (size_t i = 0; i < (size_t)N * N; i++) h_A[i] = ((float)rand()) / (float)((RAND_MAX - 1) / 1) - 0.5f;
    float *h_B = new float[n * n];
    init_B(h_B, n);
    float *h_C = new float[(size_t)N * N];
    float *C = new float[(size_t)N * N];
    memcpy(C, h_A, sizeof(float) * N * N);

    host_code(h_A, h_B, h_C, N, n);
    gold(h_A, h_B, C, N, n);

    double max_err = 0.0, max_ref = 0.0;
    size_t ring_mismatch = 0;
    for (int i = 0; i < N; i++)
        for (int j = 0; j < N; j++) {
            const size_t q = (size_t)i * N + j;
            if (i >= r && i < N - r && j >= r && j < N - r) {
                max_err = fmax(max_err, fabs((double)h_C[q] - (double)C[q]));
                max_ref = fmax(max_ref, fabs((double)C[q]));
            } else if (memcmp(&h_C[q], &h_A[q], sizeof(float)) != 0) {
                ring_mismatch++;
            }
        }
    printf("[Test] RMS Error: %e\n", oracle_check_error_2d(N, h_C, C, r, N - r, r, N - r));
    printf("[Test] Relative RMS Error: %e\n", oracle_check_relative_error_2d(N, h_C, C, r, N - r, r, N - r));
    printf("[Test] Max Error: %e (max |gold| %e)\n", max_err, max_ref);
    printf("[Test] Ring mismatches: %zu\n", ring_mismatch);
    delete[] h_A;
    delete[] h_B;
    delete[] h_C;
    delete[] C;
    return 0;
}
