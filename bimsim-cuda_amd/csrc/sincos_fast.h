// FP64 helpers of the FP32 forms' per-pixel set-up: sin and cos for the
// moderate arguments of their start values, without the general sincos's
// large-argument path; a distance and its reciprocal, and a reciprocal,
// from the hardware estimates and Newton steps.
#pragma once

#include <hip/hip_runtime.h>

namespace bimsim {

namespace sincos_detail {
// x = n pi/2 + r: 2/pi and pi/2 split into 31, 32 and 53 bits
constexpr double kTwoOverPi = 0.6366197723675814;
constexpr double kP1 = 1.5707963267341256;
constexpr double kP2 = 6.077100506303966e-11;
constexpr double kP3 = 2.0222662487959506e-21;
// sin r = r + r^3 (S0 + z S1 + ... + z^5 S5), z = r^2, and
// cos r = 1 - z/2 + z^2 (C0 + z C1 + ... + z^4 C4) on |r| <= pi/4: least
// squares fits of the relative error on 400 Chebyshev nodes in z, made by
// sincos_fit.py; the polynomials are within 4e-18 (sin) and 1e-16 (cos) of
// the functions, and the evaluation below within an ulp of them.
constexpr double kS[6] = {-0.1666666666666663,     0.0083333333333219066,
                          -0.00019841269829452244, 2.7557313580959792e-06,
                          -2.505074222357712e-08,  1.5895943103357368e-10};
constexpr double kC[5] = {0.041666666666593299, -0.0013888888877274068, 2.4801580582864965e-05,
                          -2.755550362671201e-07, 2.0644015741436523e-09};

// A constant held in a scalar register pair: the FMAs then take it as their
// scalar operand, where a folded literal costs two VALU moves (VALU
// instructions take no 64-bit literals).
__host__ __device__ __forceinline__ double scalar_const(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("" : "+s"(c));
#endif
    return c;
}
}  // namespace sincos_detail

// sin and cos of 0 <= x < 2^20 (the kd of radial start values): x =
// n pi/2 + r by Cody and Waite's reduction (n P1 and n P2 exact for
// n < 2^21), the polynomials above on |r| <= pi/4, and the quadrant
// n mod 4: within about two ulps of sin and cos over the range, in some 25
// VALU instructions against some 160 of the general sincos, whose
// large-argument reduction this range never needs.
__host__ __device__ __forceinline__ void sincos_reduced(double x, double* sn, double* cs) {
    using namespace sincos_detail;
    const double n = rint(x * scalar_const(kTwoOverPi));
    double r = fma(-n, scalar_const(kP1), x);
    r = fma(-n, scalar_const(kP2), r);
    r = fma(-n, scalar_const(kP3), r);
    const double z = r * r;
    double ps = scalar_const(kS[5]);
#pragma unroll
    for (int k = 4; k >= 0; --k) ps = fma(ps, z, scalar_const(kS[k]));
    const double sr = fma(ps * z, r, r);
    double pc = scalar_const(kC[4]);
#pragma unroll
    for (int k = 3; k >= 0; --k) pc = fma(pc, z, scalar_const(kC[k]));
    pc = fma(pc, z, -0.5);
    const double cr = fma(pc, z, 1.0);
    const int q = (int)n & 3;
    const double sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
    *sn = (q & 2) ? -sv : sv;
    *cs = ((q + 1) & 2) ? -cv : cv;
}

// The distance d = sqrt(s) of a pixel from a centre and 1/d, from s =
// |r|^2 >= 0: v_rsq_f64's estimate y, one step of Goldschmidt's coupled
// iteration from g = s y -> sqrt(s), h = y/2 -> 1/(2 sqrt(s)), two
// corrections g += (s - g^2) h, and a Newton step for 1/g from 2h.  d comes
// out correctly rounded and 1/d within two ulps (for estimates y within
// 2^-22), in some 15 VALU instructions against some 30 of the square root
// and the division.  s is never denormal here; s = 0 (a pixel at the centre
// itself) gives d = 0 and 1/d = inf, as the division does.
__device__ __forceinline__ void dist_rsq(double s, double& d, double& inv_d) {
    const double y = __builtin_amdgcn_rsq(s);
    double g = s * y, h = 0.5 * y;
    const double e = fma(-g, h, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    g = fma(fma(-g, g, s), h, g);
    g = fma(fma(-g, g, s), h, g);
    double v = h + h;
    v = fma(v, fma(-g, v, 1.0), v);
    const bool zero = s == 0.0;
    d = zero ? 0.0 : g;
    inv_d = zero ? __builtin_inf() : v;
}

// 1/x for a finite x > 0 that is not denormal: v_rcp_f64's estimate and two
// Newton steps, within an ulp or two (some 5 VALU instructions against some
// 11 of the division).
__device__ __forceinline__ double rcp_newton(double x) {
    double y = __builtin_amdgcn_rcp(x);
    y = fma(y, fma(-x, y, 1.0), y);
    y = fma(y, fma(-x, y, 1.0), y);
    return y;
}

}  // namespace bimsim
