// copy_calib -- a plain device copy of a known byte count, used to calibrate
// rocprofv3's FETCH_SIZE / WRITE_SIZE counters against HBM bytes.
//   smip_copy_calib <bytes> <launches>
// Each launch reads <bytes> and writes <bytes> once (16-byte vector accesses,
// grid-stride), so per dispatch FETCH_SIZE and WRITE_SIZE should read
// <bytes>/1024 KiB each when the counters measure HBM traffic one to one.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

__global__ void copy16(const int4* __restrict__ src, int4* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int main(int argc, char** argv)
{
    const size_t bytes = argc > 1 ? strtoull(argv[1], nullptr, 10) : (size_t)256 << 20;
    const int launches = argc > 2 ? atoi(argv[2]) : 3;
    const size_t n = bytes / sizeof(int4);
    int4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * sizeof(int4)) != hipSuccess || hipMalloc(&b, n * sizeof(int4)) != hipSuccess) {
        std::fprintf(stderr, "allocation failed\n");
        return 1;
    }
    if (hipMemset(a, 1, n * sizeof(int4)) != hipSuccess)
        return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0);
    for (int k = 0; k < launches; ++k)
        hipLaunchKernelGGL(copy16, dim3(4096), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1);
    if (hipEventSynchronize(e1) != hipSuccess)
        return 1;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    std::printf("copy %zu bytes x %d launches: %.3f ms/launch, %.1f GB/s (read+write)\n", n * sizeof(int4), launches,
        ms / launches, 2.0 * n * sizeof(int4) * launches / (ms * 1e-3) / 1e9);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
