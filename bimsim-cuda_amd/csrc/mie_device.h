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
constexpr int kMieThreads = (kLanes + 1) * kWave;   // and a wave for the start orders
static_assert(kSphPerLaunch <= kWave, "one lane per sphere in each part's wave");

// Spheres per block so that 2 sets x 4 arrays x wstride complex values
// per sphere fit the LDS budget.
inline int mie_spheres_per_block(int wstride) {
    const size_t per = 2 * 4 * sizeof(zc) * (size_t)wstride;
    return (int)std::max<size_t>(1, std::min<size_t>(kSphPerLaunch, kMieLds / per));
}

// The argument of part `role` of sphere S: ka for J and Y of ka (roles 0,
// 1), kna for J of kna (role 2).
__device__ inline zc mie_arg(const SphereDev& S, double kmag, int role) {
    const zc nc = {S.n[0], S.n[1]};
    return role < 2 ? zc{kmag * S.a, 0.0} : zj::zsc(S.a, zj::zsc(kmag, nc));
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
// lds holds 8 wstride complex values per sphere (j, y, j', y' of ka, then
// of kna).  Wave r < 3 computes part r of sphere (lane): J of ka, Y of ka,
// J of kna, as sph_jy_cplx (cbessjyva_sph) with the array clearing left out
// (every order up to Nl is written on every path).  The start orders of the
// Miller recurrences depend only on |z| and Nl, so wave 3 computes them, one
// secant search per lane, while waves 0-2 evaluate the start values; the
// parts then finish with them.  Every thread of the block (at least
// kMieThreads) then takes (sphere, order) pairs of the coefficient formula.
__device__ __forceinline__ void mie_block(const SphereDev* sh, const double* kh, int cnt,
                                          int quirks, zc* lds, int wstride, double2* A,
                                          double2* B) {
    __shared__ zj::CjyvaOrders ord[kSphPerLaunch][2];   // (sphere, ka / kna)
    const int t = (int)threadIdx.x;
    const int s = t % kWave, role = t / kWave;
    const bool part = s < cnt && role < kLanes;
    zc z = {0.0, 0.0};
    zc* o = lds;
    bool want_y = false, want_j = false;
    zj::CjyvaStart st = {};
    if (part) {
        z = mie_arg(sh[s], kh[s], role);
        o = lds + (size_t)8 * s * wstride + (role < 2 ? 0 : 4 * wstride);
        want_y = role == 1;
        want_j = role != 1 || z.i != 0.0;
        st = zj::cjyva_start(sh[s].Nl + 0.5, z, o, o + wstride, o + 2 * wstride, o + 3 * wstride,
                             quirks != 0, want_y, want_j);
    } else if (role == kLanes) {
        // one search per lane: (sphere, argument, m1 or m2)
        for (int q = s; q < 4 * cnt; q += kWave) {
            const int i = q >> 2, arg = (q >> 1) & 1, which = q & 1;
            const double v = sh[i].Nl + 0.5;
            const int m = zj::cjyva_order(zj::zabs(mie_arg(sh[i], kh[i], 2 * arg)), (int)v, which);
            if (which)
                ord[i][arg].m2 = m;
            else
                ord[i][arg].m1 = m;
        }
    }
    __syncthreads();
    if (part) {
        if (!st.done)
            zj::cjyva_rest(st, z, &ord[s][role < 2 ? 0 : 1], o, o + wstride, o + 2 * wstride,
                           o + 3 * wstride, want_y, want_j);
        zj::sph_jy_scale(sh[s].Nl, z, quirks != 0, o, o + wstride, o + 2 * wstride,
                         o + 3 * wstride, want_y, want_j);
    }
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
