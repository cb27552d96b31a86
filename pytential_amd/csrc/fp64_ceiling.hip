// FP64 throughput ceilings of the GPU it runs on, for the roofline figures of
// bench.py: v_mfma_f64_16x16x4f64 with eight independent accumulators per
// wave, and v_fma_f64 with eight independent chains per lane, no memory
// traffic in the timed loops (one store per lane at the end), 2048 blocks of
// 256 threads.  Prints one JSON line {"mfma_f64_tflops": ..,
// "valu_fma_f64_tflops": ..}; the vendor peaks are 78.6 TFLOP/s for both.
#include <hip/hip_runtime.h>

#include <cstdio>

typedef double ceil_d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_ceil_mfma(int iters, double *out)
{
    ceil_d4 acc[8];
    for (int i = 0; i < 8; ++i) acc[i] = ceil_d4{0, 0, 0, 0};
    const double a = 1.0 + threadIdx.x * 1e-9, b = 1.0 - threadIdx.x * 1e-9;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
            acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
    }
    double s = 0;
    for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) k_ceil_fma(int iters, double *out)
{
    double x[8];
    for (int i = 0; i < 8; ++i) x[i] = threadIdx.x * 1e-9 + i;
    const double a = 1.0 - 1e-12, b = 1e-13;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int i = 0; i < 8; ++i) x[i] = __builtin_fma(x[i], a, b);
    }
    double s = 0;
    for (int i = 0; i < 8; ++i) s += x[i];
    out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

#define CK(x)                                                               \
    do {                                                                    \
        if ((x) != hipSuccess) {                                            \
            fprintf(stderr, "fp64_ceiling: %s failed\n", #x);               \
            return 1;                                                       \
        }                                                                   \
    } while (0)

int main()
{
    const int blocks = 2048;
    double *out = nullptr;
    CK(hipMalloc(&out, sizeof(double) * blocks * 256));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    float ms_mfma = 0.f, ms_fma = 0.f;
    const int it_mfma = 2000, it_fma = 20000;
    hipLaunchKernelGGL(k_ceil_mfma, dim3(blocks), dim3(256), 0, 0, 10, out);
    CK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_ceil_mfma, dim3(blocks), dim3(256), 0, 0, it_mfma, out);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms_mfma, e0, e1));
    hipLaunchKernelGGL(k_ceil_fma, dim3(blocks), dim3(256), 0, 0, 10, out);
    CK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(k_ceil_fma, dim3(blocks), dim3(256), 0, 0, it_fma, out);
    CK(hipEventRecord(e1, 0));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms_fma, e0, e1));
    CK(hipGetLastError());
    // per wave one MFMA is 16 x 16 x 4 multiply-adds; per lane one FMA
    const double f_mfma = 2048.0 * 8 * it_mfma * (blocks * 4.0);
    const double f_fma = 2.0 * 8 * it_fma * (blocks * 256.0);
    printf("{\"mfma_f64_tflops\": %.2f, \"valu_fma_f64_tflops\": %.2f}\n",
           f_mfma / (ms_mfma * 1e-3) / 1e12, f_fma / (ms_fma * 1e-3) / 1e12);
    CK(hipFree(out));
    return 0;
}
