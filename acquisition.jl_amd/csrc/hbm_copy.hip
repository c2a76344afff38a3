// Plain streaming kernels of known size, outside the engine's C ABI:
//  - a device-to-device copy used to calibrate the FETCH_SIZE / WRITE_SIZE
//    hardware counters against bytes actually moved (profiling it and
//    dividing the counter reading by the bytes copied gives the factor
//    applied to the same counters of the engine's kernels);
//  - read-only and write-only streams, which with the copy give the HBM
//    bandwidth this device reaches in practice.
// `bench.py --copy-calibration` and `bench.py --hbm-stream` drive them.
#include <hip/hip_runtime.h>
#include <cstdint>

// Each workgroup moves one contiguous chunk of 4 x 256 float4; every thread
// keeps 4 loads in flight.
__global__ void __launch_bounds__(256) hbm_copy_calibration(const float4* __restrict__ src,
                                                            float4* __restrict__ dst, size_t n)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (base + u * 256 < n) v[u] = src[base + u * 256];
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (base + u * 256 < n) dst[base + u * 256] = v[u];
}

__global__ void __launch_bounds__(256) hbm_read_stream(const float4* __restrict__ src, float* __restrict__ sink,
                                                       size_t n)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    float s = 0.f;
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (base + u * 256 < n) {
            const float4 v = src[base + u * 256];
            s += v.x + v.y + v.z + v.w;
        }
    if (s == 1.2345678e30f) sink[0] = s;  // keeps the loads; the inputs are zero
}

__global__ void __launch_bounds__(256) hbm_write_stream(float4* __restrict__ dst, size_t n)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (base + u * 256 < n) dst[base + u * 256] = make_float4(1.f, 2.f, 3.f, 4.f);
}

// kind 0: copy, 1: read, 2: write, over `bytes` (rounded down to 16), `iters`
// timed launches after one untimed. Returns the mean kernel time in ms, or a
// negative value on a HIP error.
extern "C" double hbm_stream_run(size_t bytes, int iters, int kind)
{
    const size_t n = bytes / sizeof(float4);
    if (n == 0 || iters < 1 || kind < 0 || kind > 2) return -1.0;
    float4 *a = nullptr, *b = nullptr;
    float* sink = nullptr;
    if (hipMalloc(&a, n * sizeof(float4)) != hipSuccess) return -2.0;
    if (hipMalloc(&b, n * sizeof(float4)) != hipSuccess || hipMalloc(&sink, sizeof(float)) != hipSuccess) {
        (void)hipFree(a);
        (void)hipFree(b);
        return -2.0;
    }
    double ms_out = -3.0;
    hipEvent_t e0, e1;
    const unsigned grid = (unsigned)((n + 1023) / 1024);
    auto launch = [&]() {
        if (kind == 0) hbm_copy_calibration<<<grid, 256>>>(a, b, n);
        else if (kind == 1) hbm_read_stream<<<grid, 256>>>(a, sink, n);
        else hbm_write_stream<<<grid, 256>>>(b, n);
    };
    if (hipMemset(a, 0, n * sizeof(float4)) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
        hipEventCreate(&e1) == hipSuccess) {
        launch();
        (void)hipEventRecord(e0, nullptr);
        for (int i = 0; i < iters; i++) launch();
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
    (void)hipFree(sink);
    return ms_out;
}

// The copy alone (counter calibration).
extern "C" double hbm_copy_run(size_t bytes, int iters) { return hbm_stream_run(bytes, iters, 0); }
