// Element-wise operations on field slices (complex, row-major, u fastest) and
// scalar slices (real): the fieldslice / scalarslice / microscopeStruct
// building blocks that the host API exposes one by one.
//   field -> scalar       fieldslice::Mag / Intensity / Real / Imag
//                         (fieldslice.cu:6-27, 110-138, 156-230)
//   scale                 fieldslice::ScaleField (fieldslice.cu:141-153, 243-254)
//   crop                  fieldslice::crop (fieldslice.cu:319-362)
//   detector resample     fieldslice::IntegrateAndResample / ResampleAndSave
//                         (fieldslice.cu:30-108, 205-229)
//   absorbance etc.       microscopeStruct::getAbsorbance / getTransmittance
//                         (microscope.cu:722-791)
//   objective band-pass   bandpass kernel (microscope.cu:495-526)
//   back-propagation      backpropagate kernel (microscope.cu:423-489, 592-605)
// Every kernel is a grid-stride-free one-thread-per-element map; these are
// HBM-bound and small next to the near-field series.  Built with
// -ffp-contract=off so that the FP64 forms round like the reference's ops.
#include "batch.h"
#include "internal.h"

namespace bimsim {
namespace {

constexpr int kThreads = 256;

template <class R> struct Cplx;
template <> struct Cplx<float> { using T = float2; };
template <> struct Cplx<double> { using T = double2; };

inline dim3 grid1(size_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

// rts::complex::abs: sqrt(r*r + i*i)
template <class R>
__device__ inline R cabs2(typename Cplx<R>::T z) {
    return z.x * z.x + z.y * z.y;
}

// field_intensity (fieldslice.cu:6-27): |x|^2, or |x|^2 + |y|^2 + |z|^2 for a
// vector field (Y and Z both given); |.| = sqrt(r*r + i*i) as rts::complex::abs
template <class R>
__device__ inline R intensity3(const typename Cplx<R>::T* X, const typename Cplx<R>::T* Y,
                               const typename Cplx<R>::T* Z, size_t i) {
    const R xm = sqrt(cabs2<R>(X[i]));
    if (Y && Z) {
        const R ym = sqrt(cabs2<R>(Y[i]));
        const R zm = sqrt(cabs2<R>(Z[i]));
        return xm * xm + ym * ym + zm * zm;
    }
    return xm * xm;
}

template <class R>
__global__ void field_to_scalar_kernel(const typename Cplx<R>::T* __restrict__ X,
                                       const typename Cplx<R>::T* __restrict__ Y,
                                       const typename Cplx<R>::T* __restrict__ Z, size_t n, int op,
                                       R* __restrict__ S) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    R v;
    switch (op) {
        case BIMSIM_SLICE_INTENSITY: v = intensity3<R>(X, Y, Z, i); break;
        // Mag: field_intensity then field_sqrt (fieldslice.cu:156-171)
        case BIMSIM_SLICE_MAGNITUDE: v = sqrt(intensity3<R>(X, Y, Z, i)); break;
        case BIMSIM_SLICE_REAL: v = X[i].x; break;
        default: v = X[i].y; break;
    }
    S[i] = v;
}

template <class R>
__global__ void field_scale_kernel(typename Cplx<R>::T* __restrict__ F, size_t n, R s) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    typename Cplx<R>::T z = F[i];
    z.x *= s;
    z.y *= s;
    F[i] = z;
}

template <class C>
__global__ void crop_kernel(const C* __restrict__ src, int uR, int u, int v, int su, int sv,
                            C* __restrict__ dst) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= (size_t)su * sv) return;
    const int iv = (int)(i / su), iu = (int)(i - (size_t)iv * su);
    dst[i] = src[(size_t)(v + iv) * uR + (u + iu)];
}

// resample_intensity (accumulate = 1: D += scale*I/(ss*ss)) and
// resample_intensity_without_integrating (accumulate = 0: D = I/(ss*ss)).
// The field row stride is Du*ss (the detector-resolution field).
// resample_intensity / resample_intensity_without_integrating
// (fieldslice.cu:30-108): each of Y, Z adds its |.|^2 when given
template <class R>
__global__ void resample_kernel(const typename Cplx<R>::T* __restrict__ X,
                                const typename Cplx<R>::T* __restrict__ Y,
                                const typename Cplx<R>::T* __restrict__ Z, int Du, int Dv, int ss,
                                R scale, int accumulate, R* __restrict__ D) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= (size_t)Du * Dv) return;
    const int iv = (int)(i / Du), iu = (int)(i - (size_t)iv * Du);
    R I = 0;
    for (int ix = 0; ix < ss; ++ix)
        for (int iy = 0; iy < ss; ++iy) {
            const size_t fi = (size_t)(iv * ss + iy) * ss * Du + (size_t)iu * ss + ix;
            const R xm = sqrt(cabs2<R>(X[fi]));
            const R ym = Y ? sqrt(cabs2<R>(Y[fi])) : R(0);
            const R zm = Z ? sqrt(cabs2<R>(Z[fi])) : R(0);
            I += Y || Z ? xm * xm + ym * ym + zm * zm : xm * xm;
        }
    const R den = (R)(ss * ss);
    if (accumulate) D[i] += scale * I / den;
    else D[i] = I / den;
}

template <class R>
__global__ void ratio_kernel(const R* __restrict__ D, const R* __restrict__ Di, size_t n, int op,
                             R* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const R t = D[i] / Di[i];
    out[i] = op == BIMSIM_RATIO_ABSORBANCE ? -log10(t) : t;
}

// bandpass (microscope.cu:495-526): zero the samples whose spatial frequency
// lies outside [NAin/lambda, NAout/lambda]; frequencies in double.
template <class R>
__global__ void bandpass_kernel(typename Cplx<R>::T* __restrict__ F, int uR, int vR, double du,
                                double dv, double fmin, double fmax, int quirks) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= (size_t)uR * vR) return;
    const int iv = (int)(i / uR), iu = (int)(i - (size_t)iv * uR);
    double u, v;
    if (iu <= uR / 2) u = (double)iu * du;
    else u = -(double)(quirks ? uR - 1 - iu : uR - iu) * du;
    if (iv <= vR / 2) v = (double)iv * dv;
    else v = -(double)(quirks ? vR - 1 - iv : vR - iv) * dv;
    const double fmag = sqrt(u * u + v * v);
    if (fmag < fmin || fmag > fmax) {
        typename Cplx<R>::T z;
        z.x = 0;
        z.y = 0;
        F[i] = z;
    }
}

// backpropagate (microscope.cu:423-489), on the angular spectrum: samples
// with kx^2 + ky^2 < k^2 times -exp(i kz d), kz = sqrt(k^2 - kx^2 - ky^2),
// the others 0.  The reference keeps Kx, kx, k and kz in float whatever
// its precision, so they are float here too; the phase factor and the
// product are in the field's precision.  Frequencies as in bandpass (the
// negative ones one sample off with quirks).
template <class R>
__global__ void backprop_kernel(typename Cplx<R>::T* __restrict__ F, int uR, int vR, double du,
                                double dv, double lambda, float d, int quirks) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= (size_t)uR * vR) return;
    const int iv = (int)(i / uR), iu = (int)(i - (size_t)iv * uR);
    const float twopi = 2 * kPIf;
    // float Kx = 2*PI * uR * du: the last product in the field's precision
    const float Kx = (float)((R)(twopi * (float)uR) * (R)du);
    const float Ky = (float)((R)(twopi * (float)vR) * (R)dv);
    const float dkx = Kx / (float)uR, dky = Ky / (float)vR;
    const int nu = iu <= uR / 2 ? iu : -(quirks ? uR - 1 - iu : uR - iu);
    const int nv = iv <= vR / 2 ? iv : -(quirks ? vR - 1 - iv : vR - iv);
    const float kx = (float)((R)nu * (R)dkx), ky = (float)((R)nv * (R)dky);
    const float k = (float)((R)twopi / (R)lambda);
    typename Cplx<R>::T z = F[i];
    if (kx * kx + ky * ky < k * k) {
        const float kz = sqrtf(k * k - kx * kx - ky * ky);
        const R ph = (R)(kz * d);
        const R sr = -cos(ph), si = -sin(ph);
        const R zr = z.x * sr - z.y * si, zi = z.x * si + z.y * sr;
        z.x = zr;
        z.y = zi;
    } else {
        z.x = 0;
        z.y = 0;
    }
    F[i] = z;
}

bool prec_ok(int prec) { return prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST; }

__global__ void zero_words(unsigned* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kThreads)
        p[i] = 0u;
}

}  // namespace

int zero_launch(bimsim_ctx* ctx, void* p, size_t bytes) {
    const size_t n = bytes / 4;
    if (n == 0) return BIMSIM_OK;
    const size_t blocks = std::min<size_t>((n + kThreads - 1) / kThreads, 65535);
    hipLaunchKernelGGL(zero_words, dim3((unsigned)blocks), dim3(kThreads), 0, ctx->stream,
                       (unsigned*)p, n);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // namespace bimsim

using namespace bimsim;

extern "C" {

int bimsim_field3_to_scalar(bimsim_ctx* ctx, int prec, int op, const void* X, const void* Y,
                            const void* Z, long long n, void* S) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, X && S && n > 0 && prec_ok(prec), "bad field-to-scalar arguments");
    BIMSIM_CHECK_ARG(ctx, (Y == nullptr) == (Z == nullptr), "Y and Z must both be given or both NULL");
    BIMSIM_CHECK_ARG(ctx, op >= BIMSIM_SLICE_INTENSITY && op <= BIMSIM_SLICE_IMAG, "unknown op");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(field_to_scalar_kernel<float>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (const float2*)X, (const float2*)Y, (const float2*)Z, (size_t)n, op,
                           (float*)S);
    else
        hipLaunchKernelGGL(field_to_scalar_kernel<double>, grid1(n), dim3(kThreads), 0,
                           ctx->stream, (const double2*)X, (const double2*)Y, (const double2*)Z,
                           (size_t)n, op, (double*)S);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int bimsim_field_to_scalar(bimsim_ctx* ctx, int prec, int op, const void* F, long long n, void* S) {
    return bimsim_field3_to_scalar(ctx, prec, op, F, nullptr, nullptr, n, S);
}

int bimsim_field_scale(bimsim_ctx* ctx, int prec, void* F, long long n, double v) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, F && n > 0 && prec_ok(prec), "bad field-scale arguments");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(field_scale_kernel<float>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (float2*)F, (size_t)n, (float)v);
    else
        hipLaunchKernelGGL(field_scale_kernel<double>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (double2*)F, (size_t)n, v);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int bimsim_field_crop(bimsim_ctx* ctx, int prec, const void* src, int uR, int vR, int u, int v,
                      int su, int sv, void* dst) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, src && dst && prec_ok(prec), "bad crop arguments");
    BIMSIM_CHECK_ARG(ctx, su > 0 && sv > 0 && u >= 0 && v >= 0 && u + su <= uR && v + sv <= vR,
                     "crop window outside the field");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)su * sv;
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(crop_kernel<float2>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (const float2*)src, uR, u, v, su, sv, (float2*)dst);
    else
        hipLaunchKernelGGL(crop_kernel<double2>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (const double2*)src, uR, u, v, su, sv, (double2*)dst);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int bimsim_detector_resample3(bimsim_ctx* ctx, int prec, const void* X, const void* Y,
                              const void* Z, int Du, int Dv, int ss, double scale, int accumulate,
                              void* D) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, X && D && Du > 0 && Dv > 0 && ss > 0 && prec_ok(prec),
                     "bad detector-resample arguments");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)Du * Dv;
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(resample_kernel<float>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (const float2*)X, (const float2*)Y, (const float2*)Z, Du, Dv, ss,
                           (float)scale, accumulate, (float*)D);
    else
        hipLaunchKernelGGL(resample_kernel<double>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (const double2*)X, (const double2*)Y, (const double2*)Z, Du, Dv, ss,
                           scale, accumulate, (double*)D);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int bimsim_detector_resample(bimsim_ctx* ctx, int prec, const void* F, int Du, int Dv, int ss,
                             double scale, int accumulate, void* D) {
    return bimsim_detector_resample3(ctx, prec, F, nullptr, nullptr, Du, Dv, ss, scale, accumulate, D);
}

int bimsim_slice_ratio(bimsim_ctx* ctx, int prec, int op, const void* D, const void* Di,
                       long long n, void* out) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, D && Di && out && n > 0 && prec_ok(prec), "bad ratio arguments");
    BIMSIM_CHECK_ARG(ctx, op == BIMSIM_RATIO_ABSORBANCE || op == BIMSIM_RATIO_TRANSMITTANCE,
                     "unknown op");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(ratio_kernel<float>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (const float*)D, (const float*)Di, (size_t)n, op, (float*)out);
    else
        hipLaunchKernelGGL(ratio_kernel<double>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (const double*)D, (const double*)Di, (size_t)n, op, (double*)out);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int bimsim_field_bandpass(bimsim_ctx* ctx, int prec, int quirks, void* F, int uR, int vR,
                          double du, double dv, double NAin, double NAout, double lambda) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, F && uR > 0 && vR > 0 && prec_ok(prec), "bad band-pass arguments");
    BIMSIM_CHECK_ARG(ctx, lambda > 0.0, "lambda must be positive");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)uR * vR;
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(bandpass_kernel<float>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (float2*)F, uR, vR, du, dv, NAin / lambda, NAout / lambda, quirks);
    else
        hipLaunchKernelGGL(bandpass_kernel<double>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (double2*)F, uR, vR, du, dv, NAin / lambda, NAout / lambda, quirks);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int bimsim_field_backpropagate(bimsim_ctx* ctx, int prec, int quirks, void* F, int uR, int vR,
                               double du, double dv, double lambda, double d) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, F && uR > 0 && vR > 0 && prec_ok(prec), "bad back-propagation arguments");
    BIMSIM_CHECK_ARG(ctx, lambda > 0.0, "lambda must be positive");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const size_t n = (size_t)uR * vR;
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(backprop_kernel<float>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (float2*)F, uR, vR, du, dv, lambda, (float)d, quirks);
    else
        hipLaunchKernelGGL(backprop_kernel<double>, grid1(n), dim3(kThreads), 0, ctx->stream,
                           (double2*)F, uR, vR, du, dv, lambda, (float)d, quirks);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // extern "C"
