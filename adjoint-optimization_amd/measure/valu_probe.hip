// valu_probe.hip -- FP32 vector-ALU throughput of this GPU: 8 independent non-packed FMA chains per
// thread over 4096 workgroups of 256 threads, and chains of exp2 with a multiply.  The kernels'
// instruction rates are compared with these measured peaks.
#include <hip/hip_runtime.h>
#include <cstdio>
template <int CH>
__global__ void fma_chains(float *out, float x, float y, int iters)
{
    float a[CH];
#pragma unroll
    for (int i = 0; i < CH; i++) a[i] = threadIdx.x * 1e-3f + i;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < CH; i++) a[i] = fmaf(a[i], x, y);
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < CH; i++) s += a[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}
template <int CH>
__global__ void exp_chains(float *out, float x, int iters)
{
    float a[CH];
#pragma unroll
    for (int i = 0; i < CH; i++) a[i] = threadIdx.x * 1e-6f + i * 1e-3f;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < CH; i++) a[i] = __builtin_amdgcn_exp2f(a[i] * x);
    }
    float s = 0.f;
#pragma unroll
    for (int i = 0; i < CH; i++) s += a[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}
int main()
{
    const int blocks = 256 * 16, threads = 256, iters = 4096;
    float *out;
    if (hipMalloc(&out, sizeof(float) * blocks * threads) != hipSuccess) return 1;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0); (void)hipEventCreate(&e1);
    for (int rep = 0; rep < 3; rep++) {
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(fma_chains<8>, dim3(blocks), dim3(threads), 0, 0, out, 0.999f, 1e-3f, iters);
        (void)hipEventRecord(e1);
        (void)hipEventSynchronize(e1);
        float ms; (void)hipEventElapsedTime(&ms, e0, e1);
        double lane_ops = (double)blocks * threads * iters * 8;
        printf("fma: %.3f ms, %.1f G lane-FMA/s = %.1f TFLOP/s, %.0f G wave-instr/s\n", ms, lane_ops / ms / 1e6,
               2 * lane_ops / ms / 1e9, lane_ops / 64 / ms / 1e6);
        (void)hipEventRecord(e0);
        hipLaunchKernelGGL(exp_chains<8>, dim3(blocks), dim3(threads), 0, 0, out, 0.5f, iters / 4);
        (void)hipEventRecord(e1);
        (void)hipEventSynchronize(e1);
        (void)hipEventElapsedTime(&ms, e0, e1);
        double ex = (double)blocks * threads * (iters / 4) * 8;
        printf("exp2+mul: %.3f ms, %.0f G wave-pairs/s\n", ms, ex / 64 / ms / 1e6);
    }
    return 0;
}
