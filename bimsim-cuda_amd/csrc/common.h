// Shared device/host helpers of the gfx950 BIMSim library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/bimsim_hip.h"

namespace bimsim {

// Constants of the reference (dataTypes.h:17, nearfield.h:11,
// rts/math/spherical_bessel.h:8-10).
constexpr float kPIf = 3.14159f;
constexpr float kEpsFloat = 0.000001f;
constexpr double kJMin = 0.0145;
constexpr double kJMax = 0.4;
constexpr double kYMin = -1e33;

// ------------------------------------------------------------- geometry
struct d3 {
    double x, y, z;
};
__host__ __device__ inline d3 mk(const double* p) { return d3{p[0], p[1], p[2]}; }
__host__ __device__ inline d3 add(d3 a, d3 b) { return d3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline d3 sub(d3 a, d3 b) { return d3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline d3 scale(d3 a, double s) { return d3{a.x * s, a.y * s, a.z * s}; }
// rts::vector::dot / len accumulate from zero in index order
__host__ __device__ inline double dot(d3 a, d3 b) {
    double s = 0.0;
    s += a.x * b.x;
    s += a.y * b.y;
    s += a.z * b.z;
    return s;
}
__host__ __device__ inline double len(d3 a) { return sqrt(dot(a, a)); }
__host__ __device__ inline d3 norm(d3 a) {
    double l = len(a);
    return d3{a.x / l, a.y / l, a.z / l};
}

// Slice geometry passed by value to kernels.
struct QuadArg {
    d3 A, X, Y;
};
__host__ __device__ inline QuadArg to_quad(const bimsim_quad& q) {
    return QuadArg{mk(q.A), mk(q.X), mk(q.Y)};
}
// rts::quad::p(a, b) = (A + X*a) + Y*b, quad.h:129-140
__host__ __device__ inline d3 quad_at(const QuadArg& q, double u, double v) {
    return add(add(q.A, scale(q.X, u)), scale(q.Y, v));
}

// ------------------------------------------------------ rts::complex (FP64)
// Arithmetic exactly as rts::complex (reference rts/math/complex.h:19-396):
// naive product and quotient, scaling by a real, exp/sin/cos via real parts.
struct dc {
    double r, i;
};
__host__ __device__ inline dc cadd(dc a, dc b) { return dc{a.r + b.r, a.i + b.i}; }
__host__ __device__ inline dc csub(dc a, dc b) { return dc{a.r - b.r, a.i - b.i}; }
__host__ __device__ inline dc cmul(dc a, dc b) {
    return dc{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__host__ __device__ inline dc cdiv_naive(dc a, dc b) {
    double den = b.r * b.r + b.i * b.i;
    return dc{(a.r * b.r + a.i * b.i) / den, (-a.r * b.i + a.i * b.r) / den};
}
__host__ __device__ inline dc cscl(dc a, double s) { return dc{a.r * s, a.i * s}; }
__device__ inline dc cexp_i(double phi) { return dc{1.0 * cos(phi), 1.0 * sin(phi)}; }

// Kahan-compensated complex float accumulation: acc += y with the running
// compensation cmp (needs FMA contraction off, as every file here has).
__device__ __forceinline__ void kahan(float& acc_r, float& acc_i, float& cmp_r, float& cmp_i,
                                      float yr0, float yi0) {
    const float yr = yr0 - cmp_r, yi = yi0 - cmp_i;
    const float tr = acc_r + yr, ti = acc_i + yi;
    cmp_r = (tr - acc_r) - yr;
    cmp_i = (ti - acc_i) - yi;
    acc_r = tr;
    acc_i = ti;
}
__device__ inline dc csin_r(dc a) { return dc{sin(a.r) * cosh(a.i), cos(a.r) * sinh(a.i)}; }
__device__ inline dc ccos_r(dc a) { return dc{cos(a.r) * cosh(a.i), -(sin(a.r) * sinh(a.i))}; }

// Legendre scaling s_l = (2l)! / (4^l (l!)^2) linking P_l to the rescaled
// polynomials Q_l = P_l / s_l, which obey Q_l = 2x Q_{l-1} - g_l Q_{l-2}
// with g_l = 4 (l-1)^2 / ((2l-1)(2l-3)) (Q_0 = 1, Q_1 = 2x).  Q_l grows like
// sqrt(pi l) instead of decaying, so the float recurrence stays in range.
__host__ __device__ constexpr double legq_g(int l) {
    return 4.0 * (double)(l - 1) * (double)(l - 1) / ((2.0 * l - 1.0) * (2.0 * l - 3.0));
}

// Per-sphere record used by the device kernels (built on the host from
// bimsim_sphere).
struct SphereDev {
    double p[3];
    double a;
    double n[2];
    int Nl;
    int coeff_off;
};

}  // namespace bimsim
