// Plain device-to-device copy of a known size, used to calibrate the
// FETCH_SIZE / WRITE_SIZE hardware counters against bytes actually moved:
// profiling this kernel and dividing the counter reading by the bytes copied
// gives the factor applied to the same counters of the engine's kernels.
// Not part of the engine's C ABI; `bench.py --copy-calibration` drives it.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

__global__ void __launch_bounds__(256) hbm_copy_calibration(const float4* __restrict__ src,
                                                            float4* __restrict__ dst, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// Copies `bytes` (rounded down to 16) `iters` times; returns the mean kernel
// time in ms, or a negative value on a HIP error.
extern "C" double hbm_copy_run(size_t bytes, int iters)
{
    const size_t n = bytes / sizeof(float4);
    if (n == 0 || iters < 1) return -1.0;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess) return -2.0;
    if (hipMalloc(&b, n * sizeof(float4)) != hipSuccess) {
        (void)hipFree(a);
        return -2.0;
    }
    double ms_out = -3.0;
    hipEvent_t e0, e1;
    if (hipMemset(a, 0x3c, n * sizeof(float4)) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
        const int grid = 256 * 8 * 4;  // 4 waves of 256-thread groups per CU on 256 CUs
        hbm_copy_calibration<<<grid, 256>>>(a, b, n);
        (void)hipEventRecord(e0, nullptr);
        for (int i = 0; i < iters; i++) hbm_copy_calibration<<<grid, 256>>>(a, b, n);
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) == hipSuccess && hipGetLastError() == hipSuccess) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            ms_out = ms / iters;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return ms_out;
}
