// FP64 Mie coefficients on the GPU: sphere::calcCoeff (reference
// sphere.cpp:21-97) with the complex spherical Bessel functions of
// zj_device.h.  One thread per sphere.  Built with -ffp-contract=off.
#include "batch.h"
#include "internal.h"
#include "zj_device.h"

namespace bimsim {
namespace {

constexpr int kSphPerLaunch = 32;

struct SphereBatch {
    SphereDev s[kSphPerLaunch];
};

using zj::zc;

// Coefficients of one sphere at wavenumber kmag (work: 8 arrays of wstride).
__device__ void mie_one(const SphereDev& S, double kmag, int quirks, zc* w, int wstride,
                        double2* A, double2* B) {
    const int nl = S.Nl;
    zc *j1 = w, *y1 = w + wstride, *jp1 = w + 2 * wstride, *yp1 = w + 3 * wstride;
    zc *j2 = w + 4 * wstride, *y2 = w + 5 * wstride, *jp2 = w + 6 * wstride, *yp2 = w + 7 * wstride;
    const zc nc = {S.n[0], S.n[1]};
    const zc ka = {kmag * S.a, 0.0};
    const zc kna = zj::zsc(S.a, zj::zsc(kmag, nc));
    zj::sph_jy_cplx(nl, ka, quirks != 0, j1, y1, jp1, yp1);
    zj::sph_jy_cplx(nl, kna, quirks != 0, j2, y2, jp2, yp2);
    const zc ci = {0.0, 1.0};
    // i^l by exact repeated squaring, as std::pow(complex, int)
    const zc ipow[4] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {-0.0, -1.0}};
    for (int l = 0; l <= nl; ++l) {
        const zc h = zj::zadd(j1[l], zj::zmul(ci, y1[l]));
        const zc hp = zj::zadd(jp1[l], zj::zmul(ci, yp1[l]));
        const zc num_a = zj::zsub(zj::zmul(j1[l], hp), zj::zmul(jp1[l], h));
        const zc den = zj::zsub(zj::zmul(j2[l], hp), zj::zmul(zj::zmul(h, jp2[l]), nc));
        const zc f = zj::zsc(2.0 * l + 1.0, ipow[l & 3]);
        const zc An = zj::zmul(f, zj::zdiv(num_a, den));
        const zc num_b = zj::zsub(zj::zmul(zj::zmul(j1[l], jp2[l]), nc), zj::zmul(j2[l], jp1[l]));
        const zc Bn = zj::zmul(f, zj::zdiv(num_b, den));
        A[S.coeff_off + l] = make_double2(An.r, An.i);
        B[S.coeff_off + l] = make_double2(Bn.r, Bn.i);
    }
}

__global__ void mie_coeff_kernel(SphereBatch batch, int ns, double kmag, int quirks, zc* work,
                                 int wstride, double2* A, double2* B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ns) return;
    mie_one(batch.s[t], kmag, quirks, work + (size_t)t * 8 * wstride, wstride, A, B);
}

// One thread per sphere of a batch, sphere t at wavenumber kmag[t].
__global__ void mie_batch_kernel(const SphereDev* __restrict__ sph, const double* __restrict__ kmag,
                                 int n, int quirks, zc* work, int wstride, double2* A, double2* B) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    mie_one(sph[t], kmag[t], quirks, work + (size_t)t * 8 * wstride, wstride, A, B);
}

}  // namespace

int mie_batch_launch(bimsim_ctx* ctx, const SphereDev* d_sph, const double* d_kmag, int n, int lmax,
                     int quirks, void* work, double2* A, double2* B) {
    if (n <= 0) return BIMSIM_OK;
    hipLaunchKernelGGL(mie_batch_kernel, dim3((n + 63) / 64), dim3(64), 0, ctx->stream, d_sph,
                       d_kmag, n, quirks, (zc*)work, lmax + 2, A, B);
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
    void* work = nullptr;
    BIMSIM_TRY(scratch_get(ctx, kSlotMisc,
                           sizeof(zj::zc) * 8 * (size_t)wstride * kSphPerLaunch, &work));
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
        hipLaunchKernelGGL(mie_coeff_kernel, dim3(1), dim3(64), 0, ctx->stream, b, cnt, kmag,
                           quirks, (zj::zc*)work, wstride, (double2*)A_out, (double2*)B_out);
        BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    }
    return BIMSIM_OK;
}
