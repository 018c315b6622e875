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

// Per-sphere offsets into the FP32 series buffers.
struct SeriesMeta {
    int nb[kUsBatch];         // blocks of 8 orders (ceil((Nl+1)/8))
    int nl1[kUsBatch];        // orders Nl + 1
    long long coff[kUsBatch]; // element offset of c'[l][pix] for the sphere
    long long roff[kUsBatch]; // element offset of r-hat[pix]
    long long woff[kUsBatch]; // element offset of the sphere's wave table
};

// Order capacities of the register-resident FP32 series kernel.
constexpr int kSeriesBuckets[] = {16, 24, 32, 40, 48, 64, 80, 96, 112};
constexpr int kSeriesMaxL = 112;

// Launch the FP32 series kernel for `ns` spheres whose coefficients, unit
// vectors and wave tables are already in place.  `L` is the order capacity
// (a multiple of 8 >= max Nl+1).  nk_pad is a multiple of 2.  Results are
// added to (accumulate) or stored into U (float2); `partial` receives split
// partial sums when splits > 1.
int launch_us_series_f32(bimsim_ctx* ctx, int L, const float2* cp, const float4* rh,
                         const WaveF* wt, const SeriesMeta& meta, int ns, int nk_pad,
                         int npix, float2* U, int accumulate, int first_batch);

}  // namespace bimsim
