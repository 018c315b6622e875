// Device side of the incident field (reference nfScalarUf.cu:9-217): the
// field at one pixel by the series in reference order, by the complex64
// form's tabled series, or as a plane wave.  Shared by the incident-field
// kernels and the batch's fused pre-pass launch; every file that includes
// it is built with -ffp-contract=off.
#pragma once

#include <type_traits>

#include "batch.h"
#include "geometry.h"
#include "internal.h"
#include "refmath.h"
#include "sincos_fast.h"

namespace bimsim {
namespace ufd {

__device__ inline void pixel_of(int pix, int uR, int& iu, int& iv) {
    iv = pix / uR;
    iu = pix - iv * uR;
}

template <class Out>
__device__ inline void store(Out* U, int i, dc v);
template <>
__device__ inline void store<double2>(double2* U, int i, dc v) { U[i] = make_double2(v.r, v.i); }
template <>
__device__ inline void store<float2>(float2* U, int i, dc v) {
    U[i] = make_float2((float)v.r, (float)v.i);
}

// gpuScalarUf, nfScalarUf.cu:46-175: the field at pixel pix
__device__ inline dc uf_series_value(const QuadArg& q, int uR, int vR, int pix, d3 f, d3 k, double kmag,
                              double amp, double ca, double cb, int m) {
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const d3 p = quad_at(q, (double)iu / (double)uR, (double)iv / (double)vR);
    d3 r = sub(p, f);
    const double d = len(r);
    if (d < (double)kEpsFloat) return dc{amp * 2.0 * (double)kPIf * (ca - cb), 0.0};
    r = norm(r);
    const dc im = {0.0, 1.0};
    dc il = {1.0, 0.0};
    double j[2];
    const double kd = kmag * d;
    ref::init_j(kd, j);
    const double ct = dot(k, r);
    double P0 = 1.0, P1 = ct;
    double Pa[3] = {1.0, ca, 1.0}, Pb[3] = {1.0, cb, 1.0};
    dc sum = {0.0, 0.0};
    for (int l = 0; l <= m; ++l) {
        double jl, Pl;
        if (l == 0) {
            jl = j[0];
            Pl = P0;
        } else if (l == 1) {
            jl = j[1];
            Pl = P1;
            Pa[2] = Pa[0];
            ref::shift_legendre(l + 1, ca, Pa[0], Pa[1]);
            Pb[2] = Pb[0];
            ref::shift_legendre(l + 1, cb, Pb[0], Pb[1]);
        } else {
            ref::shift_j(l, kd, j);
            ref::shift_legendre(l, ct, P0, P1);
            jl = j[1];
            Pl = P1;
            Pa[2] = Pa[0];
            ref::shift_legendre(l + 1, ca, Pa[0], Pa[1]);
            Pb[2] = Pb[0];
            ref::shift_legendre(l + 1, cb, Pb[0], Pb[1]);
        }
        const double D = ((Pa[1] - Pa[2]) - Pb[1]) + Pb[2];
        sum = cadd(sum, cscl(cscl(cscl(il, jl), Pl), D));
        il = cmul(il, im);
    }
    return cscl(cscl(cscl(sum, 2.0), (double)kPIf), amp);
}

// The series of uf_series_value for complex64 fields, with the per-order
// constants from uf_fast_table (T[4l ..]: (i^l sign) D_l, (2l-1)/l,
// (l-1)/l, 2l-1): j_l and P_l by their upward recurrences multiplying by
// 1/(kd) and by the tabled ratios with FMAs; the j_l threshold rule of
// shift_j (spherical_bessel.h:109-122) unchanged (l > kd as an integer
// compare).  Differs from the reference order by rounding only.
__device__ inline dc uf_fast_value(const QuadArg& q, int uR, int vR, int pix, d3 f, d3 k, double kmag,
                            double amp, double ca, double cb, int m, const double* __restrict__ T,
                            bool flat) {
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const d3 p = quad_at(q, (double)iu / (double)uR, (double)iv / (double)vR);
    d3 r = sub(p, f);
    // the distance and sin, cos of kd by the FP32 forms' set-up helpers
    // (sincos_fast.h), their reciprocals by Newton steps
    double d, inv_d;
    dist_rsq(dot(r, r), d, inv_d);
    if (d < (double)kEpsFloat) return dc{amp * 2.0 * (double)kPIf * (ca - cb), 0.0};
    r = scale(r, inv_d);
    const double kd = kmag * d;
    const double inv = rcp_newton(kd);
    // l > kd  <=>  l >= floor(kd) + 1 for kd >= 0
    const int lk = (int)fmin(floor(kd), 1e9) + 1;
    double sn, cs;
    if (__all(kd < 1048576.0)) sincos_reduced(kd, &sn, &cs);
    else sincos(kd, &sn, &cs);
    double j0 = sn * inv;
    double j1 = fma(j0, inv, -(cs * inv));
    // one order: j_l from the recurrence (threshold rule where l > kd), then
    // P_l and the term (t: the order's table entry)
    // (chk: apply the rule; a block of orders skips it when no lane of the
    // wave has l > kd there)
    auto jstep = [&](int l, const double* t, bool chk) {
        double jn = fma(t[3] * j1, inv, -j0);
        if (chk && l >= lk && (jn < kJMin || jn > kJMax)) jn = 0.0;
        j0 = j1;
        j1 = jn;
        return jn;
    };
    // the table is read four orders at a time, so that one wait on the
    // scalar loads covers four orders
    constexpr int kB = 4;
    if (flat) {
        // k.r = 0 at every pixel: P_l(0) = -((l-1)/l) P_{l-2}(0) and 0 for odd
        // l, so the series is real; the flat table holds D_l P_l(0) for the
        // even orders (uf_fast_table), one FMA per term
        double ar = j0 * T[0];
        auto fstep = [&](int l, const double* t, bool chk) {
            const double jn = jstep(l, t, chk);
            if (!(l & 1)) ar = fma(jn, t[0], ar);
        };
        int l = 2;
        for (; l + kB - 1 <= m; l += kB) {
            // two orders zeroed by the threshold rule zero every later one
            // (the recurrence then stays at 0): the rest of the series adds
            // nothing once that holds on every lane
            if (__all(j0 == 0.0 && j1 == 0.0)) break;
            double tb[4 * kB];
#pragma unroll
            for (int x = 0; x < 4 * kB; ++x) tb[x] = T[4 * l + x];
            if (__any(l + kB - 1 >= lk)) {
#pragma unroll
                for (int b = 0; b < kB; ++b) fstep(l + b, tb + 4 * b, true);
            } else {
#pragma unroll
                for (int b = 0; b < kB; ++b) fstep(l + b, tb + 4 * b, false);
            }
        }
        if (!(j0 == 0.0 && j1 == 0.0))
            for (; l <= m; ++l) fstep(l, T + 4 * l, true);
        return dc{((ar * 2.0) * (double)kPIf) * amp, 0.0};
    }
    const double ct = dot(k, r);
    double P0 = 1.0, P1 = ct;
    double ar = j0 * T[0];
    double ai = (j1 * ct) * T[4];
    auto step = [&](int l, const double* t, double acc, bool chk) {
        const double jn = jstep(l, t, chk);
        const double Pn = fma(t[1] * ct, P1, -(t[2] * P0));
        P0 = P1;
        P1 = Pn;
        return fma(jn * Pn, t[0], acc);
    };
    int l = 2;
    for (; l + kB - 1 <= m; l += kB) {   // l even: ar, ai, ar, ai
        if (__all(j0 == 0.0 && j1 == 0.0)) break;   // (as in the flat case)
        double tb[4 * kB];
#pragma unroll
        for (int x = 0; x < 4 * kB; ++x) tb[x] = T[4 * l + x];
        if (__any(l + kB - 1 >= lk)) {
            ar = step(l, tb, ar, true);
            ai = step(l + 1, tb + 4, ai, true);
            ar = step(l + 2, tb + 8, ar, true);
            ai = step(l + 3, tb + 12, ai, true);
        } else {
            ar = step(l, tb, ar, false);
            ai = step(l + 1, tb + 4, ai, false);
            ar = step(l + 2, tb + 8, ar, false);
            ai = step(l + 3, tb + 12, ai, false);
        }
    }
    if (!(j0 == 0.0 && j1 == 0.0))
        for (; l <= m; ++l) {
            if (l & 1) ai = step(l, T + 4 * l, ai, true);
            else ar = step(l, T + 4 * l, ar, true);
        }
    return dc{((ar * 2.0) * (double)kPIf) * amp, ((ai * 2.0) * (double)kPIf) * amp};
}

// gpuScalarUfp, nfScalarUf.cu:9-43: the plane wave at pixel pix
__device__ inline dc uf_plane_value(const QuadArg& q, int uR, int vR, int pix, d3 f, d3 k,
                                    double kmag, double amp) {
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const d3 p = quad_at(q, (double)iu / (double)uR, (double)iv / (double)vR);
    const d3 r = sub(p, f);
    const double kr = kmag * dot(k, r);
    return cscl(cexp_i(kr), amp);
}

// The Monte-Carlo incident field at pixel pix: the plane waves of
// gpuScalarUfp (nfScalarUf.cu:9-43) along the nk directions W, each with
// the weight sA = (1.0f/nk) subA that the scattered field gives the same
// wave (nfScalarUs.cu:461,572), summed in wave order in FP64.
__device__ inline dc uf_mc_value(const QuadArg& q, int uR, int vR, int pix, d3 f, double kmag,
                                 const double* __restrict__ W, int nk, double sA) {
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const d3 p = quad_at(q, (double)iu / (double)uR, (double)iv / (double)vR);
    const d3 r = sub(p, f);
    dc sum = {0.0, 0.0};
    for (int w = 0; w < nk; ++w) {
        const d3 K = {W[3 * w], W[3 * w + 1], W[3 * w + 2]};
        sum = cadd(sum, cscl(cexp_i(kmag * dot(K, r)), sA));
    }
    return sum;
}

// Pixel pix of batch instance I into its field (complex64 fields use the
// instance's uf_fast_table row at tabs + I.tab; their Monte-Carlo instances
// are left to uf_mc_batch_launch).
template <class Out>
__device__ inline void uf_batch_pixel(const UfInst& I, int pix, const double* __restrict__ tabs,
                                      const double* __restrict__ waves, Out* __restrict__ Uf) {
    dc v;
    if (I.mode == BIMSIM_UF_PLANE) {
        v = uf_plane_value(I.q, I.uR, I.vR, pix, I.f, I.k, I.kmag, I.amp);
    } else if (I.mode == BIMSIM_UF_MC) {
        if constexpr (std::is_same<Out, float2>::value) return;
        v = uf_mc_value(I.q, I.uR, I.vR, pix, I.f, I.kmag, waves + 3 * I.wave_off, I.nk, I.sA);
    } else if constexpr (std::is_same<Out, float2>::value) {   // tables for every instance
        v = uf_fast_value(I.q, I.uR, I.vR, pix, I.f, I.k, I.kmag, I.amp, I.ca, I.cb, I.m,
                          tabs + I.tab, I.flat != 0);
    } else {
        v = uf_series_value(I.q, I.uR, I.vR, pix, I.f, I.k, I.kmag, I.amp, I.ca, I.cb, I.m);
    }
    store(Uf + I.off, pix, v);
}

// Work items of batch instance I's incident field: its pixels, or with
// mirror symmetry (uf_mirror) its canonical pixels, or with the diagonal
// symmetry as well the canonical pixels on and below the diagonal.
__device__ inline int uf_batch_work(const UfInst& I) { return uf_work_items(I); }

// Slice column (row) of canonical column (row) index c.
__device__ inline int uf_canonical(int c, int nl, int h) { return c < nl ? c : h + (c - nl); }

// Stores v at pixel (iu, iv) and at its mirror images in the slice (column
// mx - iu, row my - iv, both), those inside the slice and distinct.
template <class Out>
__device__ inline void uf_store_mirrors(const UfInst& I, Out* F, int iu, int iv, dc v) {
    const int pu = I.mx - iu, pv = I.my - iv;
    const bool mu = pu >= 0 && pu < I.uR && pu != iu, mv = pv >= 0 && pv < I.vR && pv != iv;
    store(F, iv * I.uR + iu, v);
    if (mu) store(F, iv * I.uR + pu, v);
    if (mv) store(F, pv * I.uR + iu, v);
    if (mu && mv) store(F, pv * I.uR + pu, v);
}

// Work item w of batch instance I: the pixel w, or the canonical pixel w
// stored at each of its mirror images in the slice (its own column and
// mx - that column, its row and my - that row, those inside the slice),
// and with the diagonal symmetry (uf_mirror: column i and row i at the same
// offsets from the focus) at the transposed pixel and its mirror images
// too.
template <class Out>
__device__ inline void uf_batch_item(const UfInst& I, int w, const double* __restrict__ tabs,
                                     const double* __restrict__ waves, Out* __restrict__ Uf) {
    if (I.ncx <= 0) {
        uf_batch_pixel(I, w, tabs, waves, Uf);
        return;
    }
    int cx, cy;
    if (I.diag) {
        // tile t = ty (ty + 1) / 2 + tx, tx <= ty, of 8 x 8 canonical pixels
        const int t = w >> 6, lane = w & 63;
        int ty = (int)((sqrtf(8.f * (float)t + 1.f) - 1.f) * 0.5f);
        while (ty > 0 && ty * (ty + 1) / 2 > t) --ty;
        while ((ty + 1) * (ty + 2) / 2 <= t) ++ty;
        const int tx = t - ty * (ty + 1) / 2;
        cx = 8 * tx + (lane & 7);
        cy = 8 * ty + (lane >> 3);
        if (cx > cy || cy >= I.ncx) return;
    } else {
        cy = w / I.ncx;
        cx = w - cy * I.ncx;
    }
    const int iu = uf_canonical(cx, I.nlx, I.hx), iv = uf_canonical(cy, I.nly, I.hy);
    dc v;
    if constexpr (std::is_same<Out, float2>::value)
        v = uf_fast_value(I.q, I.uR, I.vR, iv * I.uR + iu, I.f, I.k, I.kmag, I.amp, I.ca, I.cb,
                          I.m, tabs + I.tab, I.flat != 0);
    else
        v = uf_series_value(I.q, I.uR, I.vR, iv * I.uR + iu, I.f, I.k, I.kmag, I.amp, I.ca, I.cb,
                            I.m);
    Out* F = Uf + I.off;
    uf_store_mirrors(I, F, iu, iv, v);
    if (I.diag && cx != cy) uf_store_mirrors(I, F, iv, iu, v);
}

}  // namespace ufd
}  // namespace bimsim
