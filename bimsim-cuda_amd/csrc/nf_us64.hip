// Scattered/internal field of spheres under Monte-Carlo plane waves
// (nearfieldStruct::scalarUs, reference nfScalarUs.cu:494-658).
//
// FP64 kernels of this file are built with -ffp-contract=off:
//   us_parity_kernel   the reference structure (precompute + optPc_gpuScalarUsp,
//                      nfScalarUs.cu:17-48,449-491): per pixel, per wave, the
//                      full calc_Us / calc_Ui recurrences in double.
//   us_coef64_kernel   per (pixel, sphere) wave-independent products
//                      c_l = B_l h_l(kd) or A_l j_l(knd) (the only part of
//                      calc_Us/calc_Ui that does not depend on the wave).
//   us_series64_kernel the hoisted FP64 form: same per-term operations as
//                      the parity kernel, bit for bit.
//   wave tables        per (sphere, wave) phase weights in FP64.
// The FP32 form lives in us_series.hip.
#include "nf_us.h"
#include "us_series.h"
#include "refmath.h"

namespace bimsim {
namespace {

constexpr int kThreads = 256;

__device__ inline dc ld2(const double2* p, int i) {
    const double2 v = p[i];
    return dc{v.x, v.y};
}

// calc_Us, reference nfScalarUs.cu:50-95
__device__ dc calc_Us(double kd, double ct, int Nl, const double2* B) {
    double j[2], y[2];
    ref::init_j(kd, j);
    ref::init_y(kd, y);
    double P0 = 1.0, P1 = ct;
    dc Us = {0.0, 0.0};
    for (int l = 0; l <= Nl; ++l) {
        if (l == 0) {
            Us = cadd(Us, cscl(cmul(ld2(B, 0), dc{j[0], y[0]}), P0));
        } else {
            if (l > 1) {
                ref::shift_legendre(l, ct, P0, P1);
                ref::shift_j(l, kd, j);
                ref::shift_y(l, kd, y);
            }
            Us = cadd(Us, cscl(cmul(ld2(B, l), dc{j[1], y[1]}), P1));
        }
    }
    return Us;
}

// calc_Ui, reference nfScalarUs.cu:201-245
__device__ dc calc_Ui(dc knd, double ct, int Nl, const double2* A) {
    if (knd.r < (double)kEpsFloat) return ld2(A, 0);
    dc j[2];
    ref::init_jc(knd, j);
    double P0 = 1.0, P1 = ct;
    dc Ui = {0.0, 0.0};
    for (int l = 0; l <= Nl; ++l) {
        if (l == 0) {
            Ui = cadd(Ui, cscl(cmul(ld2(A, 0), j[0]), P0));
        } else {
            if (l > 1) {
                ref::shift_legendre(l, ct, P0, P1);
                ref::shift_jc(l, knd, j);
            }
            Ui = cadd(Ui, cscl(cmul(ld2(A, l), j[1]), P1));
        }
    }
    return Ui;
}

__device__ inline void pixel_of(int pix, int uR, int& iu, int& iv) {
    iv = pix / uR;
    iu = pix - iv * uR;
}

__global__ __launch_bounds__(kThreads) void us_parity_kernel(
    QuadArg q, int uR, int vR, d3 f, double kmag, UsSphereBatch batch, int ns,
    const double2* __restrict__ A, const double2* __restrict__ B,
    const double* __restrict__ waves, int nk, double sA, double2* U, int accumulate) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= uR * vR) return;
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const double u = (double)iu / (double)uR, v = (double)iv / (double)vR;
    const d3 p = quad_at(q, u, v);
    dc acc = {0.0, 0.0};
    if (accumulate) acc = ld2(U, pix);
    for (int s = 0; s < ns; ++s) {
        const SphereDev& S = batch.s[s];
        const d3 ps = mk(S.p);
        d3 r = sub(p, ps);
        const double d = len(r);
        const double kd = kmag * d;
        const dc knd = cscl(dc{S.n[0], S.n[1]}, kmag * d);
        r = norm(r);
        const d3 c = sub(ps, f);
        const double2* As = A + S.coeff_off;
        const double2* Bs = B + S.coeff_off;
        dc sum = {0.0, 0.0};
        const bool inside = d <= S.a;
        for (int w = 0; w < nk; ++w) {
            const d3 k = mk(waves + 3 * w);
            const double ct = dot(k, r);
            const dc phase = cexp_i(kmag * dot(k, c));
            const dc term = inside ? calc_Ui(knd, ct, S.Nl, As) : calc_Us(kd, ct, S.Nl, Bs);
            sum = cadd(sum, cmul(cscl(phase, sA), term));
        }
        acc = cadd(acc, sum);
    }
    U[pix] = make_double2(acc.r, acc.i);
}

// Wave-independent products of calc_Us / calc_Ui for one pixel and sphere.
// kind: 0 = outside (Hankel), 1 = inside (Bessel), 2 = inside near the
// centre (calc_Ui returns A_0 for every wave).
template <class Sink>
__device__ inline int pixel_coefficients(double d, double kmag, const SphereDev& S,
                                         const double2* As, const double2* Bs, Sink&& sink) {
    const int Nl = S.Nl;
    if (d <= S.a) {
        const dc knd = cscl(dc{S.n[0], S.n[1]}, kmag * d);
        if (knd.r < (double)kEpsFloat) {
            sink(0, ld2(As, 0));
            for (int l = 1; l <= Nl; ++l) sink(l, dc{0.0, 0.0});
            return 2;
        }
        dc j[2];
        ref::init_jc(knd, j);
        sink(0, cmul(ld2(As, 0), j[0]));
        if (Nl >= 1) sink(1, cmul(ld2(As, 1), j[1]));
        for (int l = 2; l <= Nl; ++l) {
            ref::shift_jc(l, knd, j);
            sink(l, cmul(ld2(As, l), j[1]));
        }
        return 1;
    }
    const double kd = kmag * d;
    double j[2], y[2];
    ref::init_j(kd, j);
    ref::init_y(kd, y);
    sink(0, cmul(ld2(Bs, 0), dc{j[0], y[0]}));
    if (Nl >= 1) sink(1, cmul(ld2(Bs, 1), dc{j[1], y[1]}));
    for (int l = 2; l <= Nl; ++l) {
        ref::shift_j(l, kd, j);
        ref::shift_y(l, kd, y);
        sink(l, cmul(ld2(Bs, l), dc{j[1], y[1]}));
    }
    return 0;
}

// Hoisted FP64 form, one sphere: c_l in coef[l*npix + pix], r-hat in
// rh[c*npix + pix], kind in kind[pix].
__global__ __launch_bounds__(kThreads) void us_coef64_kernel(
    QuadArg q, int uR, int vR, double kmag, SphereDev S, const double2* __restrict__ A,
    const double2* __restrict__ B, double2* __restrict__ coef, double* __restrict__ rh,
    int* __restrict__ kind) {
    const int npix = uR * vR;
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= npix) return;
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const d3 p = quad_at(q, (double)iu / (double)uR, (double)iv / (double)vR);
    d3 r = sub(p, mk(S.p));
    const double d = len(r);
    r = norm(r);
    rh[pix] = r.x;
    rh[npix + pix] = r.y;
    rh[2 * npix + pix] = r.z;
    kind[pix] = pixel_coefficients(d, kmag, S, A + S.coeff_off, B + S.coeff_off,
                                   [&](int l, dc c) { coef[(size_t)l * npix + pix] = make_double2(c.r, c.i); });
}

// Per-wave weights (scale*subA) * exp(i kmag k.c) for one sphere (FP64).
__global__ void phase64_kernel(d3 c, double kmag, double sA, const double* __restrict__ waves,
                               int nk, double2* __restrict__ ph) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nk) return;
    const dc e = cscl(cexp_i(kmag * dot(mk(waves + 3 * w), c)), sA);
    ph[w] = make_double2(e.r, e.i);
}

__global__ __launch_bounds__(kThreads) void us_series64_kernel(
    int npix, int Nl, const double2* __restrict__ coef, const double* __restrict__ rh,
    const int* __restrict__ kind, const double* __restrict__ waves, const double2* __restrict__ ph,
    int nk, double2* U, int accumulate) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= npix) return;
    const d3 r = {rh[pix], rh[npix + pix], rh[2 * npix + pix]};
    const int kd = kind[pix];
    dc sum = {0.0, 0.0};
    for (int w = 0; w < nk; ++w) {
        dc term;
        if (kd == 2) {
            term = ld2(coef, pix);
        } else {
            const double ct = dot(mk(waves + 3 * w), r);
            double P0 = 1.0, P1 = ct;
            term = dc{0.0, 0.0};
            for (int l = 0; l <= Nl; ++l) {
                const dc c = ld2(coef, l * npix + pix);
                if (l == 0) {
                    term = cadd(term, cscl(c, P0));
                } else {
                    if (l > 1) ref::shift_legendre(l, ct, P0, P1);
                    term = cadd(term, cscl(c, P1));
                }
            }
        }
        sum = cadd(sum, cmul(ld2(ph, w), term));
    }
    dc acc = {0.0, 0.0};
    if (accumulate) acc = ld2(U, pix);
    acc = cadd(acc, sum);
    U[pix] = make_double2(acc.r, acc.i);
}

__global__ void clear_c64(double2* U, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) U[i] = make_double2(0.0, 0.0);
}
__global__ void clear_c32(float2* U, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) U[i] = make_float2(0.f, 0.f);
}
// U (+)= the complex128 field W rounded to complex64
__global__ void narrow_add_kernel(const double2* __restrict__ W, float2* __restrict__ U, int n,
                                  int accumulate) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double2 w = W[i];
    float2 v = make_float2(0.f, 0.f);
    if (accumulate) v = U[i];
    U[i] = make_float2(v.x + (float)w.x, v.y + (float)w.y);
}

SphereDev to_dev(const bimsim_sphere& S) {
    SphereDev d;
    for (int c = 0; c < 3; ++c) d.p[c] = S.p[c];
    d.a = S.a;
    d.n[0] = S.n[0];
    d.n[1] = S.n[1];
    d.Nl = S.Nl;
    d.coeff_off = S.coeff_off;
    return d;
}

}  // namespace
}  // namespace bimsim

namespace bimsim {

int us_wide_fp32(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR, const double focus[3],
                 double kmag, const bimsim_sphere* spheres, int ns, const void* A, const void* B,
                 const double* waves, int nk, double subA, float2* U, int accumulate) {
    const int npix = uR * vR;
    void* wide = nullptr;
    BIMSIM_TRY(scratch_get(ctx, kSlotWide, sizeof(double2) * (size_t)npix, &wide));
    BIMSIM_TRY(bimsim_nf_us(ctx, pos, uR, vR, focus, kmag, spheres, ns, A, B, waves, nk, subA,
                            BIMSIM_FP64, BIMSIM_QUIRKS_REF, wide, 0));
    hipLaunchKernelGGL(narrow_add_kernel, dim3((npix + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       ctx->stream, (const double2*)wide, U, npix, accumulate ? 1 : 0);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // namespace bimsim

using namespace bimsim;

extern "C" int bimsim_nf_us(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                            const double focus[3], double kmag, const bimsim_sphere* spheres,
                            int ns, const void* A, const void* B, const double* waves, int nk,
                            double subA, int prec, int quirks, void* U, int accumulate) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, pos && focus && U, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 28),
                     "slice resolution must be positive and at most 2^28 pixels");
    BIMSIM_CHECK_ARG(ctx, ns >= 0 && nk > 0, "ns must be >= 0 and nk > 0");
    BIMSIM_CHECK_ARG(ctx, prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST, "unknown precision");
    BIMSIM_CHECK_ARG(ctx, kmag > 0.0 && std::isfinite(kmag), "kmag must be positive");
    if (!quirks)
        return set_error(ctx, BIMSIM_E_UNSUPPORTED,
                         "scattered-field series without the reference recurrences is not built");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int npix = uR * vR;
    const int pix_blocks = (npix + kThreads - 1) / kThreads;
    if (ns == 0) {
        if (!accumulate) {
            if (prec == BIMSIM_FP32_FAST)
                hipLaunchKernelGGL(clear_c32, dim3(pix_blocks), dim3(kThreads), 0, ctx->stream, (float2*)U, npix);
            else
                hipLaunchKernelGGL(clear_c64, dim3(pix_blocks), dim3(kThreads), 0, ctx->stream, (double2*)U, npix);
            BIMSIM_HIP_CHECK(ctx, hipGetLastError());
        }
        return BIMSIM_OK;
    }
    BIMSIM_CHECK_ARG(ctx, spheres && A && B && waves, "null sphere/coefficient/wave pointer");
    int lmax = 0;
    for (int s = 0; s < ns; ++s) {
        BIMSIM_CHECK_ARG(ctx, spheres[s].Nl >= 1 && spheres[s].a > 0.0 && spheres[s].coeff_off >= 0,
                         "sphere needs Nl >= 1, a > 0, coeff_off >= 0");
        lmax = std::max(lmax, spheres[s].Nl);
    }
    const QuadArg q = to_quad(*pos);
    const d3 f = mk(focus);
    // ptype scale = (1.0f/nk) (nfScalarUs.cu:461): float arithmetic
    const double sA = (double)(1.0f / (float)nk) * subA;
    const int tmr = timer_begin(ctx, kStageUs);
    if (prec == BIMSIM_FP64_PARITY) {
        for (int s0 = 0; s0 < ns; s0 += kUsBatch) {
            const int cnt = std::min(kUsBatch, ns - s0);
            UsSphereBatch b{};
            for (int i = 0; i < cnt; ++i) b.s[i] = to_dev(spheres[s0 + i]);
            hipLaunchKernelGGL(us_parity_kernel, dim3(pix_blocks), dim3(kThreads), 0, ctx->stream,
                               q, uR, vR, f, kmag, b, cnt, (const double2*)A, (const double2*)B,
                               waves, nk, sA, (double2*)U, (s0 > 0 || accumulate) ? 1 : 0);
            BIMSIM_HIP_CHECK(ctx, hipGetLastError());
        }
    } else if (prec == BIMSIM_FP64) {
        void *coef = nullptr, *rh = nullptr, *misc = nullptr, *ph = nullptr;
        BIMSIM_TRY(scratch_get(ctx, kSlotCoef, sizeof(double2) * (size_t)(lmax + 1) * npix, &coef));
        BIMSIM_TRY(scratch_get(ctx, kSlotRhat, sizeof(double) * 3 * (size_t)npix, &rh));
        BIMSIM_TRY(scratch_get(ctx, kSlotMisc, sizeof(int) * (size_t)npix, &misc));
        BIMSIM_TRY(scratch_get(ctx, kSlotWaves, sizeof(double2) * (size_t)nk, &ph));
        for (int s = 0; s < ns; ++s) {
            const SphereDev S = to_dev(spheres[s]);
            const d3 c = sub(mk(S.p), f);
            hipLaunchKernelGGL(phase64_kernel, dim3((nk + 255) / 256), dim3(256), 0, ctx->stream,
                               c, kmag, sA, waves, nk, (double2*)ph);
            hipLaunchKernelGGL(us_coef64_kernel, dim3(pix_blocks), dim3(kThreads), 0, ctx->stream,
                               q, uR, vR, kmag, S, (const double2*)A, (const double2*)B,
                               (double2*)coef, (double*)rh, (int*)misc);
            hipLaunchKernelGGL(us_series64_kernel, dim3(pix_blocks), dim3(kThreads), 0, ctx->stream,
                               npix, S.Nl, (const double2*)coef, (const double*)rh,
                               (const int*)misc, waves, (const double2*)ph, nk, (double2*)U,
                               (s > 0 || accumulate) ? 1 : 0);
            BIMSIM_HIP_CHECK(ctx, hipGetLastError());
        }
    } else if (series_capacity(lmax + 1) < 0) {
        // orders beyond the FP32 series kernels: the FP64 form, rounded
        timer_end(ctx, tmr);
        return us_wide_fp32(ctx, pos, uR, vR, focus, kmag, spheres, ns, A, B, waves, nk, subA,
                            (float2*)U, accumulate);
    } else {
        // FP32 series: a batch of one instance
        SeriesBatch batch;
        SeriesInst I{};
        I.q = q;
        I.f = f;
        I.kmag = kmag;
        I.sA = sA;
        I.uR = uR;
        I.vR = vR;
        I.s_begin = 0;
        I.ns = ns;
        I.nk = nk;
        I.accumulate = accumulate ? 1 : 0;
        batch.inst.push_back(I);
        for (int s = 0; s < ns; ++s) {
            SeriesSph S{};
            S.s = to_dev(spheres[s]);
            S.inst = 0;
            batch.sph.push_back(S);
        }
        const SeriesPlan* plan = nullptr;
        BIMSIM_TRY(series_cached_plan(ctx, batch, waves, (const double2*)A, (const double2*)B, &plan));
        BIMSIM_TRY(series_plan_launch(ctx, *plan, waves, (const double2*)A, (const double2*)B,
                                     (float2*)U));
    }
    timer_end(ctx, tmr);
    return BIMSIM_OK;
}
