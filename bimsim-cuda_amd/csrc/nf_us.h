// Scattered-field (nfScalarUs) data layouts shared by the FP64 set-up kernels
// and the FP32 series kernel.
#pragma once

#include "internal.h"

namespace bimsim {

// Spheres per launch of the per-sphere kernels (records travel as kernel
// arguments).
constexpr int kUsBatch = 32;

struct UsSphereBatch {
    SphereDev s[kUsBatch];
};

// One Monte-Carlo wave as seen by the FP32 series kernel: 2*k (so that the
// dot product with r-hat is the Legendre argument 2x of Q_1) and the wave's
// complex weight (1/nk) * subA * exp(i kmag k.c) for the sphere.
struct WaveF {
    float kx2, ky2, kz2, pr, pi, pad0, pad1, pad2;
};

// The scattered field of an FP32 call whose orders exceed the FP32 series
// kernels (Nl + 1 > 112): the hoisted FP64 form into context scratch, then
// U (+)= that field rounded to complex64.
int us_wide_fp32(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR, const double focus[3],
                 double kmag, const bimsim_sphere* spheres, int ns, const void* A, const void* B,
                 const double* waves, int nk, double subA, float2* U, int accumulate);

}  // namespace bimsim
