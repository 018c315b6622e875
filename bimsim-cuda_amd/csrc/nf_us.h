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

}  // namespace bimsim
