// FP64 Mie coefficients on the GPU: sphere::calcCoeff (reference
// sphere.cpp:21-97) with the complex spherical Bessel functions of
// zj_device.h.  Built with -ffp-contract=off.
//
// The work is a few long chains of dependent FP64 steps per sphere, so the
// kernels spread it over lanes and keep it out of global memory: a block
// takes up to 32 spheres and computes, into LDS, J of ka, Y of ka and J of
// kna (the Y of kna is not used), one wave per part and one lane per sphere
// in it, so that the three parts' different code paths run side by side on
// different waves rather than one after another in one wave; the
// recurrences reread their arrays, so LDS latency instead of a global round
// trip per step.  Then every lane takes (sphere, order) pairs of the
// coefficient formula.
#include "batch.h"
#include "internal.h"
#include "mie_device.h"

namespace bimsim {
namespace {

using namespace mie;

struct SphereBatch {
    SphereDev s[kSphPerLaunch];
};

// Spheres [b spb, b spb + spb) of the batch passed by value, at one kmag.
// kGlobal: the Bessel arrays (8 wstride values per sphere) are in work,
// for orders too high for LDS; otherwise in LDS (a separate instantiation,
// so that its accesses compile to LDS instructions).
template <bool kGlobal>
__global__ __launch_bounds__(kMieThreads) void mie_coeff_kernel(
    SphereBatch batch, int ns, int spb, double kmag, int quirks, zc* work, int wstride,
    double2* A, double2* B) {
    extern __shared__ zc lds[];
    __shared__ SphereDev sh[kSphPerLaunch];
    __shared__ double kh[kSphPerLaunch];
    const int s0 = (int)blockIdx.x * spb;
    const int cnt = min(spb, ns - s0);
    const int t = (int)threadIdx.x;
    if (t < cnt) {
        sh[t] = batch.s[s0 + t];
        kh[t] = kmag;
    }
    __syncthreads();
    if (kGlobal)
        mie_block(sh, kh, cnt, quirks, work + (size_t)8 * s0 * wstride, wstride, A, B);
    else
        mie_block(sh, kh, cnt, quirks, lds, wstride, A, B);
}

// Spheres [b spb, b spb + spb) of a batch, sphere t at wavenumber kmag[t].
template <bool kGlobal>
__global__ __launch_bounds__(kMieThreads) void mie_batch_kernel(
    const SphereDev* __restrict__ sph, const double* __restrict__ kmag, int n, int spb,
    int quirks, zc* work, int wstride, double2* A, double2* B) {
    extern __shared__ zc lds[];
    __shared__ SphereDev sh[kSphPerLaunch];
    __shared__ double kh[kSphPerLaunch];
    const int s0 = (int)blockIdx.x * spb;
    const int cnt = min(spb, n - s0);
    const int t = (int)threadIdx.x;
    if (t < cnt) {
        sh[t] = sph[s0 + t];
        kh[t] = kmag[s0 + t];
    }
    __syncthreads();
    if (kGlobal)
        mie_block(sh, kh, cnt, quirks, work + (size_t)8 * s0 * wstride, wstride, A, B);
    else
        mie_block(sh, kh, cnt, quirks, lds, wstride, A, B);
}

}  // namespace

int mie_batch_launch(bimsim_ctx* ctx, const SphereDev* d_sph, const double* d_kmag, int n, int lmax,
                     int quirks, void* work, double2* A, double2* B) {
    if (n <= 0) return BIMSIM_OK;
    const int wstride = lmax + 2;
    const int spb = mie_spheres_per_block(wstride);
    size_t lds = sizeof(zc) * 8 * (size_t)wstride * spb;
    zc* w = nullptr;
    if (lds > std::min(kMieLds, ctx->lds_max)) {   // very high orders: global arrays
        lds = 0;
        w = (zc*)work;
    }
    if (w)
        hipLaunchKernelGGL(mie_batch_kernel<true>, dim3((n + spb - 1) / spb), dim3(kMieThreads),
                           0, ctx->stream, d_sph, d_kmag, n, spb, quirks, w, wstride, A, B);
    else
        hipLaunchKernelGGL(mie_batch_kernel<false>, dim3((n + spb - 1) / spb), dim3(kMieThreads),
                           lds, ctx->stream, d_sph, d_kmag, n, spb, quirks, w, wstride, A, B);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // namespace bimsim

using namespace bimsim;

extern "C" int bimsim_mie_coeffs(bimsim_ctx* ctx, const bimsim_sphere* spheres, int ns,
                                 double kmag, int quirks, void* A_out, void* B_out) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, ns >= 0, "ns must be >= 0");
    if (ns == 0) return BIMSIM_OK;
    BIMSIM_CHECK_ARG(ctx, spheres && A_out && B_out, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, kmag > 0.0 && std::isfinite(kmag), "kmag must be positive");
    int lmax = 0;
    for (int s = 0; s < ns; ++s) {
        BIMSIM_CHECK_ARG(ctx, spheres[s].Nl >= 2 && spheres[s].Nl <= 4096,
                         "sphere order Nl must be in [2, 4096]");
        BIMSIM_CHECK_ARG(ctx, spheres[s].a > 0.0 && spheres[s].coeff_off >= 0,
                         "sphere radius must be positive and coeff_off >= 0");
        lmax = std::max(lmax, spheres[s].Nl);
    }
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int wstride = lmax + 2;
    const int spb = mie_spheres_per_block(wstride);
    size_t lds = sizeof(zj::zc) * 8 * (size_t)wstride * spb;
    void* work = nullptr;
    if (lds > std::min(kMieLds, ctx->lds_max)) {   // very high orders: global arrays
        lds = 0;
        BIMSIM_TRY(scratch_get(ctx, kSlotMisc,
                               sizeof(zj::zc) * 8 * (size_t)wstride * kSphPerLaunch, &work));
    }
    for (int s0 = 0; s0 < ns; s0 += kSphPerLaunch) {
        const int cnt = std::min(kSphPerLaunch, ns - s0);
        SphereBatch b{};
        for (int i = 0; i < cnt; ++i) {
            const bimsim_sphere& S = spheres[s0 + i];
            SphereDev& d = b.s[i];
            for (int c = 0; c < 3; ++c) d.p[c] = S.p[c];
            d.a = S.a;
            d.n[0] = S.n[0];
            d.n[1] = S.n[1];
            d.Nl = S.Nl;
            d.coeff_off = S.coeff_off;
        }
        if (work)
            hipLaunchKernelGGL(mie_coeff_kernel<true>, dim3((cnt + spb - 1) / spb),
                               dim3(kMieThreads), 0, ctx->stream, b, cnt, spb, kmag, quirks,
                               (zj::zc*)work, wstride, (double2*)A_out, (double2*)B_out);
        else
            hipLaunchKernelGGL(mie_coeff_kernel<false>, dim3((cnt + spb - 1) / spb),
                               dim3(kMieThreads), lds, ctx->stream, b, cnt, spb, kmag, quirks,
                               nullptr, wstride, (double2*)A_out, (double2*)B_out);
        BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    }
    return BIMSIM_OK;
}
