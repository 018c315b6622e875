// Issue-rate probe for the roofline of the FP32 series kernels: independent
// FMA chains (16 per lane, 2048 steps) in FP32, compiled without packed
// FP32 instructions as the library is, and in FP64; prints TFLOP/s.  Not
// part of the library: `make fma_probe`, then ./fma_probe on the GPU.
#include <hip/hip_runtime.h>
#include <cstdio>
constexpr int kChains = 16;
constexpr int kIter = 2048;
template <class T>
__global__ __launch_bounds__(256) void fma_kernel(const T* __restrict__ in, T* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    T a = in[0], b = in[1];
    T acc[kChains];
#pragma unroll
    for (int c = 0; c < kChains; ++c) acc[c] = in[2 + c] + (T)t;
    for (int i = 0; i < kIter; ++i) {
#pragma unroll
        for (int c = 0; c < kChains; ++c) acc[c] = fma(acc[c], a, b);
    }
    T s = 0;
#pragma unroll
    for (int c = 0; c < kChains; ++c) s += acc[c];
    out[t] = s;
}
template <class T>
double run(int blocks) {
    T h[2 + kChains];
    h[0] = (T)0.999;
    h[1] = (T)0.001;
    for (int c = 0; c < kChains; ++c) h[2 + c] = (T)c;
    T *in = nullptr, *out = nullptr;
    if (hipMalloc(&in, sizeof(h)) != hipSuccess ||
        hipMalloc(&out, sizeof(T) * (size_t)blocks * 256) != hipSuccess ||
        hipMemcpy(in, h, sizeof(h), hipMemcpyHostToDevice) != hipSuccess) {
        fprintf(stderr, "fma_probe: device set-up failed\n");
        return -1.0;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    hipLaunchKernelGGL(fma_kernel<T>, dim3(blocks), dim3(256), 0, 0, in, out);
    (void)hipEventRecord(e0, 0);
    for (int r = 0; r < 5; ++r)
        hipLaunchKernelGGL(fma_kernel<T>, dim3(blocks), dim3(256), 0, 0, in, out);
    (void)hipEventRecord(e1, 0);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipFree(in);
    (void)hipFree(out);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const double flops = 2.0 * kChains * kIter * 256.0 * blocks * 5;
    return flops / (ms * 1e-3) / 1e12;
}
int main() {
    for (int blocks : {2048, 8192}) {
        const double f32 = run<float>(blocks), f64 = run<double>(blocks);
        printf("blocks %d: fp32 %.1f TFLOP/s, fp64 %.1f TFLOP/s\n", blocks, f32, f64);
        fflush(stdout);
    }
    return 0;
}
