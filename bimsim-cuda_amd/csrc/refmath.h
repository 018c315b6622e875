// In-kernel recurrences of the reference in FP64, operation for operation:
// rts/math/legendre.h:9-25 and rts/math/spherical_bessel.h:85-141 as driven
// by calc_Us / calc_Ui / gpuScalarUf (nfScalarUs.cu:50-95,201-245;
// nfScalarUf.cu:46-175).  Translation units including this header are built
// with -ffp-contract=off so each product and sum rounds separately.
// Threshold compares run in double (the working precision).
#pragma once

#include "common.h"

namespace bimsim {
namespace ref {

__device__ inline void shift_legendre(int n, double x, double& P0, double& P1) {
    const double Pn = ((double)(2 * n - 1) * x * P1 - (double)(n - 1) * P0) / (double)n;
    P0 = P1;
    P1 = Pn;
}

__device__ inline void init_j(double x, double* j) {
    j[0] = sin(x) / x;
    j[1] = j[0] / x - cos(x) / x;
}
__device__ inline void init_y(double x, double* y) {
    y[0] = -cos(x) / x;
    y[1] = y[0] / x - sin(x) / x;
}
__device__ inline void shift_j(int n, double x, double* b) {
    double bn = ((double)(2 * n - 1) * b[1]) / x - b[0];
    if ((double)n > x)
        if (bn < kJMin || bn > kJMax) bn = 0.0;
    b[0] = b[1];
    b[1] = bn;
}
__device__ inline void shift_y(int n, double x, double* b) {
    double bn = ((double)(2 * n - 1) * b[1]) / x - b[0];
    if (bn < kYMin || ((double)n > x && bn > 0.0)) {
        bn = 0.0;
        b[1] = 0.0;
    }
    b[0] = b[1];
    b[1] = bn;
}
__device__ inline void init_jc(dc x, dc* j) {
    j[0] = cdiv_naive(csin_r(x), x);
    j[1] = csub(cdiv_naive(j[0], x), cdiv_naive(ccos_r(x), x));
}
__device__ inline void shift_jc(int n, dc x, dc* b) {
    dc bn = csub(cdiv_naive(cscl(b[1], (double)(2 * n - 1)), x), b[0]);
    if ((double)n > x.r)
        if (bn.r < kJMin || bn.r > kJMax) bn = dc{0.0, 0.0};
    b[0] = b[1];
    b[1] = bn;
}

}  // namespace ref
}  // namespace bimsim
