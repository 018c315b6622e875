// Detector-side images of the interpolating extended-source drivers
// (reference microscope.cu:37-240):
//   rotate and add      InterpolateAndIntegrateUd / Ufd (microscope.cu:89-132)
//   ring sums           simRing_Ud / simRing_Ufd (microscope.cu:138-240)
//   radial line image   InterpolateAndIntegrateUdFromLine / Ufd (:37-86)
// The reference samples float textures (bilinear, border addressing, +0.5
// texel offsets); here the image itself is sampled with exact bilinear
// weights at the same coordinates, in the slice's precision.  Coordinates
// the reference keeps in float are rounded to float.  Images are small
// (the detector grid); the kernels are one thread per output sample, and
// one workgroup per ring for the sums.  Built with -ffp-contract=off.
#include "internal.h"

namespace bimsim {
namespace {

constexpr int kThreads = 256;
constexpr int kRingMax = 1024;   // threads of a ring workgroup

// bilinear at data coordinates (x, y) with zero outside the image
template <class T>
__device__ inline T fetch_border(const T* __restrict__ S, int W, int H, float xf, float yf) {
    const double x = xf, y = yf;
    const double fx = floor(x), fy = floor(y);
    const T wx = (T)(x - fx), wy = (T)(y - fy);
    const int x0 = (int)fx, y0 = (int)fy;
    auto at = [&](int xx, int yy) -> T {
        if (xx < 0 || yy < 0 || xx >= W || yy >= H) return (T)0;
        return S[(size_t)yy * W + xx];
    };
    const T a = ((T)1 - wx) * at(x0, y0) + wx * at(x0 + 1, y0);
    const T b = ((T)1 - wx) * at(x0, y0 + 1) + wx * at(x0 + 1, y0 + 1);
    return ((T)1 - wy) * a + wy * b;
}

template <class T>
__global__ __launch_bounds__(kThreads) void rotate_add_kernel(const T* __restrict__ src,
                                                              T* __restrict__ D, int uR, int vR,
                                                              double deg, T scale) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= uR * vR) return;
    const int iv = i / uR, iu = i - iv * uR;
    const int uc = (int)(uR / 2.0), vc = (int)(vR / 2.0);
    const double cs = cos(deg), sn = sin(deg);
    const float nx = (float)((double)uc + (double)(iu - uc) * cs - (double)(iv - vc) * sn);
    const float ny = (float)((double)vc + (double)(iu - uc) * sn + (double)(iv - vc) * cs);
    D[i] = D[i] + scale * fetch_border(src, uR, vR, nx, ny);
}

// One workgroup per ring; the sum follows the reference's tree exactly.
template <class T>
__global__ __launch_bounds__(kRingMax) void ring_sums_kernel(const T* __restrict__ src,
                                                             T* __restrict__ line, int uR, int vR,
                                                             int ns, int nrings) {
    __shared__ T s[kRingMax];
    const int r = blockIdx.x;
    const int t = threadIdx.x;
    if (r >= nrings) return;   // uniform over the workgroup
    if (t < ns) {
        double th = (2.0 * (double)kPIf / ns) * (double)(t + 1);
        if (sizeof(T) == sizeof(float)) th = (double)(float)th;   // a float angle
        const int uc = (int)(uR / 2.0), vc = (int)(vR / 2.0);
        const float nx = (float)((double)uc + (double)r * cos(th));
        const float ny = (float)((double)vc - (double)r * sin(th));
        s[t] = fetch_border(src, uR, vR, nx, ny);
    }
    __syncthreads();
    for (int step = ns / 2; step >= 1; step >>= 1) {
        if (t < step) s[t] = s[t] + s[t + step];
        __syncthreads();
    }
    if (t == 0) line[r] = line[r] + s[0];
}

template <class T>
__global__ __launch_bounds__(kThreads) void from_line_kernel(const T* __restrict__ line, int len,
                                                             T* __restrict__ D, int uR, int vR,
                                                             T scale) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= uR * vR) return;
    const int iv = i / uR, iu = i - iv * uR;
    const int uc = (int)(uR / 2.0), vc = (int)(vR / 2.0);
    const double du = iu - uc, dv = iv - vc;
    // float arithmetic in the reference: the square sum is exact and the
    // double square root rounds to the correctly rounded float one
    const float rf = (float)sqrt(du * du + dv * dv);
    const double x = rf;
    const double fx = floor(x);
    const T w = (T)(x - fx);
    const int x0 = (int)fx;
    const T a = (x0 >= 0 && x0 < len) ? line[x0] : (T)0;
    const T b = (x0 + 1 >= 0 && x0 + 1 < len) ? line[x0 + 1] : (T)0;
    D[i] = scale * (((T)1 - w) * a + w * b);
}

}  // namespace
}  // namespace bimsim

using namespace bimsim;

extern "C" {

int bimsim_detector_rotate_add(bimsim_ctx* ctx, int prec, double deg, int uR, int vR,
                               double scale, const void* src, void* D) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, src && D, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 30), "bad sizes");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const dim3 grid((unsigned)(((long long)uR * vR + kThreads - 1) / kThreads));
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(rotate_add_kernel<float>, grid, dim3(kThreads), 0, ctx->stream,
                           (const float*)src, (float*)D, uR, vR, deg, (float)scale);
    else
        hipLaunchKernelGGL(rotate_add_kernel<double>, grid, dim3(kThreads), 0, ctx->stream,
                           (const double*)src, (double*)D, uR, vR, deg, scale);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int bimsim_detector_ring_sums(bimsim_ctx* ctx, int prec, int uR, int vR, int n_samples,
                              int n_rings, const void* src, void* line) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, src && line, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 30), "bad sizes");
    BIMSIM_CHECK_ARG(ctx, n_samples >= 1 && n_samples < kRingMax && n_rings >= 0,
                     "ring sums need 1 <= samples <= 1023 (the reference's workgroup limit)");
    if (n_rings == 0) return BIMSIM_OK;
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int threads = (n_samples + 63) / 64 * 64;
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(ring_sums_kernel<float>, dim3(n_rings), dim3(threads), 0, ctx->stream,
                           (const float*)src, (float*)line, uR, vR, n_samples, n_rings);
    else
        hipLaunchKernelGGL(ring_sums_kernel<double>, dim3(n_rings), dim3(threads), 0, ctx->stream,
                           (const double*)src, (double*)line, uR, vR, n_samples, n_rings);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int bimsim_detector_from_line(bimsim_ctx* ctx, int prec, int uR, int vR, double scale,
                              const void* line, int len, void* D) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, line && D, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 30) && len >= 1,
                     "bad sizes");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const dim3 grid((unsigned)(((long long)uR * vR + kThreads - 1) / kThreads));
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(from_line_kernel<float>, grid, dim3(kThreads), 0, ctx->stream,
                           (const float*)line, len, (float*)D, uR, vR, (float)scale);
    else
        hipLaunchKernelGGL(from_line_kernel<double>, grid, dim3(kThreads), 0, ctx->stream,
                           (const double*)line, len, (double*)D, uR, vR, scale);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // extern "C"
