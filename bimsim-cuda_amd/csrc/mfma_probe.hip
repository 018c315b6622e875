// Probe of v_mfma_f32_4x4x1_16b_f32 as the harmonic kernels use it (A
// operand broadcast from one block of four lanes, CBSZ 4): its lane layout,
// its rounding against fmaf, its issue rate with the accumulators tied in
// AGPRs, the issue rates of plain VALU forms, and MFMA and VALU work
// interleaved in every wave of a SIMD.  Not part of the library: `make mfma_probe`,
// then ./mfma_probe on the GPU.  Every kernel reads and writes only the
// buffers allocated in main, at the indices its grid covers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

typedef float f32x4 __attribute__((ext_vector_type(4)));

#define CHECK(x)                                                                       \
    do {                                                                               \
        hipError_t e_ = (x);                                                           \
        if (e_ != hipSuccess) {                                                        \
            std::fprintf(stderr, "mfma_probe: %s at line %d\n", hipGetErrorString(e_), \
                         __LINE__);                                                    \
            return 1;                                                                  \
        }                                                                              \
    } while (0)

// one wave: d = mfma(a, b, c) with the A block broadcast from block kAbid
template <int kAbid>
__global__ void layout_kernel(const float* a, const float* b, const float* c, float* d) {
    const int l = threadIdx.x;
    f32x4 acc = {c[4 * l], c[4 * l + 1], c[4 * l + 2], c[4 * l + 3]};
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(a[l], b[l], acc, 4, kAbid, 0);
    for (int i = 0; i < 4; ++i) d[4 * l + i] = acc[i];
}

// kMode 0: 8 independent MFMA chains per wave (accumulators tied in AGPRs,
// the eight products in one asm statement with one leading s_nop as in the
// harmonic kernels); 1: 16 independent v_fma_f32 chains; 2: 16 v_mul_f32;
// 3: 16 v_fma_f64; 4: both 0 and 1 in every iteration of every wave (8
// waves per SIMD, so that each can issue one kind while another's other
// kind runs, if the SIMD allows it)
__device__ __forceinline__ void mfma8(f32x4* acc, float a, float b) {
    asm volatile(
        "s_nop 1\n\t"
        "v_mfma_f32_4x4x1_16b_f32 %0, %8, %9, %0 cbsz:4 abid:0\n\t"
        "v_mfma_f32_4x4x1_16b_f32 %1, %8, %9, %1 cbsz:4 abid:1\n\t"
        "v_mfma_f32_4x4x1_16b_f32 %2, %8, %9, %2 cbsz:4 abid:2\n\t"
        "v_mfma_f32_4x4x1_16b_f32 %3, %8, %9, %3 cbsz:4 abid:3\n\t"
        "v_mfma_f32_4x4x1_16b_f32 %4, %8, %9, %4 cbsz:4 abid:4\n\t"
        "v_mfma_f32_4x4x1_16b_f32 %5, %8, %9, %5 cbsz:4 abid:5\n\t"
        "v_mfma_f32_4x4x1_16b_f32 %6, %8, %9, %6 cbsz:4 abid:6\n\t"
        "v_mfma_f32_4x4x1_16b_f32 %7, %8, %9, %7 cbsz:4 abid:7\n\t"
        "s_nop 0"
        : "+a"(acc[0]), "+a"(acc[1]), "+a"(acc[2]), "+a"(acc[3]), "+a"(acc[4]), "+a"(acc[5]),
          "+a"(acc[6]), "+a"(acc[7])
        : "v"(a), "v"(b));
}

template <int kMode>
__global__ __launch_bounds__(256) void rate_kernel(const float* in, float* out, int n) {
    const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
    const float a = in[l], b = in[64 + l], y = in[128 + l], z = in[129];
    float s = 0.f;
    (void)w;
    if constexpr (kMode == 4) {
        f32x4 acc[8];
        for (int k = 0; k < 8; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, (float)k};
        float x[16];
        for (int k = 0; k < 16; ++k) x[k] = 1.f + 1e-3f * l + k;
        for (int it = 0; it < n; ++it) {
            mfma8(acc, a, b);
#pragma unroll
            for (int k = 0; k < 16; ++k) x[k] = fmaf(x[k], y, z);
        }
        asm volatile("s_nop 7\n\ts_nop 7"
                     : "+a"(acc[0]), "+a"(acc[1]), "+a"(acc[2]), "+a"(acc[3]), "+a"(acc[4]),
                       "+a"(acc[5]), "+a"(acc[6]), "+a"(acc[7]));
        for (int k = 0; k < 8; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
        for (int k = 0; k < 16; ++k) s += x[k];
    }
    if constexpr (kMode == 0) {
        f32x4 acc[8];
        for (int k = 0; k < 8; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, (float)k};
        for (int it = 0; it < n; ++it) mfma8(acc, a, b);
        asm volatile("s_nop 7\n\ts_nop 7"
                     : "+a"(acc[0]), "+a"(acc[1]), "+a"(acc[2]), "+a"(acc[3]), "+a"(acc[4]),
                       "+a"(acc[5]), "+a"(acc[6]), "+a"(acc[7]));
        for (int k = 0; k < 8; ++k) s += acc[k][0] + acc[k][1] + acc[k][2] + acc[k][3];
    }
    if constexpr (kMode >= 1 && kMode <= 3) {
        {
            if constexpr (kMode == 3) {
                double x[16];
                for (int k = 0; k < 16; ++k) x[k] = 1.0 + 1e-3 * l + k;
                for (int it = 0; it < n; ++it) {
#pragma unroll
                    for (int k = 0; k < 16; ++k) x[k] = fma(x[k], (double)y, (double)z);
                }
                for (int k = 0; k < 16; ++k) s += (float)x[k];
            } else {
                float x[16];
                for (int k = 0; k < 16; ++k) x[k] = 1.f + 1e-3f * l + k;
                for (int it = 0; it < n; ++it) {
#pragma unroll
                    for (int k = 0; k < 16; ++k) x[k] = kMode == 2 ? x[k] * y : fmaf(x[k], y, z);
                }
                for (int k = 0; k < 16; ++k) s += x[k];
            }
        }
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

static unsigned bits(float f) {
    unsigned u;
    std::memcpy(&u, &f, 4);
    return u;
}

int main() {
    std::vector<float> ha(256), hb(64), hc(256), hd(256);
    float *da, *db, *dc, *dd, *dout;
    CHECK(hipMalloc(&da, 256 * sizeof(float)));
    CHECK(hipMalloc(&db, 64 * sizeof(float)));
    CHECK(hipMalloc(&dc, 256 * sizeof(float)));
    CHECK(hipMalloc(&dd, 256 * sizeof(float)));
    // lane layout: integer operands, every product exact
    for (int l = 0; l < 64; ++l) {
        ha[l] = 1.f + l;
        hb[l] = 1000.f + l;
    }
    for (int i = 0; i < 256; ++i) hc[i] = 0.f;
    CHECK(hipMemcpy(da, ha.data(), 64 * sizeof(float), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(db, hb.data(), 64 * sizeof(float), hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dc, hc.data(), 256 * sizeof(float), hipMemcpyHostToDevice));
    hipLaunchKernelGGL((layout_kernel<5>), dim3(1), dim3(64), 0, 0, da, db, dc, dd);
    CHECK(hipMemcpy(hd.data(), dd, 256 * sizeof(float), hipMemcpyDeviceToHost));
    int bad = 0;
    for (int l = 0; l < 64; ++l)
        for (int i = 0; i < 4; ++i)
            if (hd[4 * l + i] != ha[4 * 5 + i] * hb[l]) ++bad;
    std::printf("layout (cbsz 4, abid 5): lane l gets d[i] = A[lane 20 + i] * B[lane l]: %s\n",
                bad ? "MISMATCH" : "yes");
    // rounding against fmaf over random operands of spread exponents
    int diff_fma = 0, diff_two = 0, total = 0;
    std::srand(7);
    auto rnd = []() {
        return (float)((std::rand() / (double)RAND_MAX - 0.5) * std::ldexp(1.0, std::rand() % 40 - 20));
    };
    for (int rep = 0; rep < 200; ++rep) {
        for (int l = 0; l < 64; ++l) {
            ha[l] = rnd();
            hb[l] = rnd();
        }
        for (int i = 0; i < 256; ++i) hc[i] = rnd();
        CHECK(hipMemcpy(da, ha.data(), 64 * sizeof(float), hipMemcpyHostToDevice));
        CHECK(hipMemcpy(db, hb.data(), 64 * sizeof(float), hipMemcpyHostToDevice));
        CHECK(hipMemcpy(dc, hc.data(), 256 * sizeof(float), hipMemcpyHostToDevice));
        hipLaunchKernelGGL((layout_kernel<2>), dim3(1), dim3(64), 0, 0, da, db, dc, dd);
        CHECK(hipMemcpy(hd.data(), dd, 256 * sizeof(float), hipMemcpyDeviceToHost));
        for (int l = 0; l < 64; ++l)
            for (int i = 0; i < 4; ++i) {
                const float av = ha[8 + i], bv = hb[l], cv = hc[4 * l + i];
                ++total;
                if (bits(hd[4 * l + i]) != bits(std::fmaf(av, bv, cv))) ++diff_fma;
                if (bits(hd[4 * l + i]) != bits(av * bv + cv)) ++diff_two;
            }
    }
    std::printf("rounding: %d results, %d differ from fmaf, %d from product-then-sum\n", total,
                diff_fma, diff_two);
    std::fflush(stdout);
    // issue rates, 8 waves per SIMD
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount, grid = cus * 8, n = 4000;
    const double clk = prop.clockRate * 1e3;
    for (int i = 0; i < 256; ++i) ha[i] = i < 128 ? 1.f / (1 + i) : 0.999999f;
    ha[129] = 1e-7f;
    CHECK(hipMemcpy(da, ha.data(), 256 * sizeof(float), hipMemcpyHostToDevice));
    CHECK(hipMalloc(&dout, (size_t)grid * 256 * sizeof(float)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const char* names[] = {"MFMA 4x4x1 f32 (8 chains)", "v_fma_f32 (16 chains)",
                           "v_mul_f32 (16 chains)", "v_fma_f64 (16 chains)",
                           "MFMA and v_fma_f32 in every wave"};
    float ms[5];
    for (int k = 0; k < 5; ++k) {
        ms[k] = 1e30f;
        for (int rep = 0; rep < 3; ++rep) {
            CHECK(hipEventRecord(e0));
            switch (k) {
                case 0: hipLaunchKernelGGL(rate_kernel<0>, dim3(grid), dim3(256), 0, 0, da, dout, n); break;
                case 1: hipLaunchKernelGGL(rate_kernel<1>, dim3(grid), dim3(256), 0, 0, da, dout, n); break;
                case 2: hipLaunchKernelGGL(rate_kernel<2>, dim3(grid), dim3(256), 0, 0, da, dout, n); break;
                case 3: hipLaunchKernelGGL(rate_kernel<3>, dim3(grid), dim3(256), 0, 0, da, dout, n); break;
                default: hipLaunchKernelGGL(rate_kernel<4>, dim3(grid), dim3(256), 0, 0, da, dout, n); break;
            }
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float t = 0.f;
            CHECK(hipEventElapsedTime(&t, e0, e1));
            if (t < ms[k]) ms[k] = t;
        }
        // instructions per SIMD: 8 waves x n x (8 MFMA or 16 VALU)
        const double per = k == 4 ? 0.0 : 8.0 * n * (k == 0 ? 8 : 16);
        if (k < 4)
            std::printf("%-38s %.3f ms  %.2f cycles per instruction per SIMD\n", names[k], ms[k],
                        ms[k] * 1e-3 * clk / per);
        else
            std::printf("%-38s %.3f ms  (the two alone: %.3f ms added, %.3f ms the longer)\n",
                        names[k], ms[k], ms[0] + ms[1], ms[0] > ms[1] ? ms[0] : ms[1]);
        std::fflush(stdout);
    }
    return 0;
}
