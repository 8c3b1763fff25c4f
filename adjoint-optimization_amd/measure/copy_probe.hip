// copy_probe.hip -- a plain device-to-device copy kernel of known size (1 GiB read + 1 GiB written,
// larger than the 256 MB Infinity Cache).  Profiled with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE it
// gives the factors that turn those counters into HBM bytes for the fused kernels' traffic figures.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copyk(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

int main()
{
    const size_t bytes = 1024ull << 20;
    const size_t n = bytes / 16;
    float4 *a, *b;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        printf("alloc failed\n");
        return 1;
    }
    if (hipMemset(a, 1, bytes) != hipSuccess || hipMemset(b, 0, bytes) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    for (int r = 0; r < 5; r++) {
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(copyk, dim3(8192), dim3(256), 0, 0, a, b, n);
        (void)hipEventRecord(e1);
        if (hipEventSynchronize(e1) != hipSuccess) return 1;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        printf("copy %zu MiB: %.3f ms, %.1f GB/s (read+write)\n", bytes >> 20, ms, 2.0 * bytes / ms / 1e6);
    }
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
