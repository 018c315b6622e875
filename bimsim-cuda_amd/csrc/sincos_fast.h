// FP64 sin and cos for the moderate arguments of the FP32 forms' radial
// start values, without the general sincos's large-argument path.
#pragma once

#include <hip/hip_runtime.h>

namespace bimsim {

// sin and cos of 0 <= x < 2^20 (the kd of radial start values): x =
// n pi/2 + r by Cody and Waite's reduction with pi/2 split into 31, 32 and
// 53 bits (n P1 and n P2 exact for n < 2^21), the Taylor series of sin r
// and cos r on |r| <= pi/4 up to r^17 and r^18 (the next terms are below
// 1e-19), and the quadrant n mod 4: about 60 VALU instructions against some
// 160 of the general sincos, whose large-argument reduction this range
// never needs.
__constant__ static const double kSinCos[21] = {
    0.6366197723675814,       // 2/pi
    1.5707963267341256,       // P1
    6.077100506303966e-11,    // P2
    2.0222662487959506e-21,   // P3
    // (-1)^k / (2k+1)!, k = 1 .. 8
    -0.16666666666666666, 0.008333333333333333, -0.0001984126984126984,
    2.7557319223985893e-06, -2.505210838544172e-08, 1.6059043836821613e-10,
    -7.647163731819816e-13, 2.8114572543455206e-15,
    // (-1)^k / (2k)!, k = 1 .. 9
    -0.5, 0.041666666666666664, -0.001388888888888889, 2.48015873015873e-05,
    -2.755731922398589e-07, 2.08767569878681e-09, -1.1470745597729725e-11,
    4.779477332387385e-14, -1.5619206968586225e-16};

__device__ __forceinline__ void sincos_reduced(double x, double* sn, double* cs) {
    const double n = rint(x * kSinCos[0]);
    double r = fma(-n, kSinCos[1], x);
    r = fma(-n, kSinCos[2], r);
    r = fma(-n, kSinCos[3], r);
    const double r2 = r * r;
    double ps = kSinCos[11];
#pragma unroll
    for (int k = 10; k >= 4; --k) ps = fma(ps, r2, kSinCos[k]);
    const double sr = fma(ps * r2, r, r);
    double pc = kSinCos[20];
#pragma unroll
    for (int k = 19; k >= 12; --k) pc = fma(pc, r2, kSinCos[k]);
    const double cr = fma(pc, r2, 1.0);
    const int q = (int)n & 3;
    const double sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
    *sn = (q & 2) ? -sv : sv;
    *cs = ((q + 1) & 2) ? -cv : cv;
}

}  // namespace bimsim
