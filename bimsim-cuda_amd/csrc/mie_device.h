// Device side of the FP64 Mie coefficients (sphere::calcCoeff, reference
// sphere.cpp:21-97): the three Bessel parts of a block of spheres, one wave
// each, and the coefficient formula per (sphere, order).  Shared by the
// coefficient kernels and the batch's fused pre-pass launch; every file
// that includes it is built with -ffp-contract=off.
#pragma once

#include <algorithm>

#include "internal.h"
#include "zj_device.h"

namespace bimsim {
namespace mie {

using zj::zc;

constexpr int kSphPerLaunch = 32;   // spheres per block at most
constexpr size_t kMieLds = 61440;   // LDS budget of a block's Bessel arrays (plus 2 KB static)

constexpr int kLanes = 3;            // parts per sphere: one wave each
constexpr int kWave = 64;
constexpr int kMieThreads = kLanes * kWave;
static_assert(kSphPerLaunch <= kWave, "one lane per sphere in each part's wave");

// Spheres per block so that 2 sets x 4 arrays x wstride complex values
// per sphere fit the LDS budget.
inline int mie_spheres_per_block(int wstride) {
    const size_t per = 2 * 4 * sizeof(zc) * (size_t)wstride;
    return (int)std::max<size_t>(1, std::min<size_t>(kSphPerLaunch, kMieLds / per));
}

// One part of the Bessel sets of sphere S into its work w (j, y, j', y' of
// ka, then of kna, wstride apart): role 0 J of ka, 1 Y of ka, 2 J of kna.
__device__ inline void mie_bessel(const SphereDev& S, double kmag, int role, int quirks, zc* w,
                           int wstride) {
    const zc nc = {S.n[0], S.n[1]};
    const zc z = role < 2 ? zc{kmag * S.a, 0.0} : zj::zsc(S.a, zj::zsc(kmag, nc));
    zc* o = role < 2 ? w : w + 4 * wstride;
    zj::sph_jy_cplx(S.Nl, z, quirks != 0, o, o + wstride, o + 2 * wstride, o + 3 * wstride,
                    role == 1, role != 1);
}

// A_l and B_l of sphere S from its two Bessel sets (w: ka set, then kna).
__device__ inline void mie_order(const SphereDev& S, int l, const zc* w, int wstride, double2* A,
                          double2* B) {
    const zc *j1 = w, *y1 = w + wstride, *jp1 = w + 2 * wstride, *yp1 = w + 3 * wstride;
    const zc *j2 = w + 4 * wstride, *jp2 = w + 6 * wstride;
    const zc nc = {S.n[0], S.n[1]};
    const zc ci = {0.0, 1.0};
    // i^l by exact repeated squaring, as std::pow(complex, int)
    const zc ipow[4] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {-0.0, -1.0}};
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

// Coefficients of the block's cnt spheres (sh, wavenumbers kh, in LDS);
// lds holds 8 wstride complex values per sphere.  Wave r < 3 computes part
// r of sphere (lane); every thread of the block (at least kMieThreads)
// then takes (sphere, order) pairs.
__device__ __forceinline__ void mie_block(const SphereDev* sh, const double* kh, int cnt,
                                          int quirks, zc* lds, int wstride, double2* A,
                                          double2* B) {
    const int t = (int)threadIdx.x;
    const int s = t % kWave;
    if (s < cnt && t < kMieThreads)
        mie_bessel(sh[s], kh[s], t / kWave, quirks, lds + (size_t)8 * s * wstride, wstride);
    __syncthreads();
    int lm = 0;
    for (int i = 0; i < cnt; ++i) lm = max(lm, sh[i].Nl);
    const int per = lm + 1;
    for (int idx = t; idx < cnt * per; idx += (int)blockDim.x) {
        const int i = idx / per, l = idx - i * per;
        if (l <= sh[i].Nl) mie_order(sh[i], l, lds + (size_t)8 * i * wstride, wstride, A, B);
    }
}


}  // namespace mie
}  // namespace bimsim
