// FP64 fractional-order Bessel functions on the device, following the
// published Zhang & Jin routines (S. Zhang, J. Jin, "Computation of Special
// Functions", Wiley 1996: GAMMA2, MSTA1, MSTA2, JYV, CJYVA) in the same steps
// as the reference's C++ transcription (gamma.cpp:14-82, bessjy.cpp:268-319
// and 536-728, cbessjy.cpp:456-737), so that Mie coefficients and radial
// LUTs agree with the reference's host code to rounding.
//
// The complex helpers follow the C99/glibc formulas that the reference's
// std::complex<double> calls reach on Linux: Smith's quotient (libgcc
// __divdc3), csqrt, clog via hypot/atan2, ccos/csin via cosh/sinh, and
// pow(z, v) = polar(exp(v log|z|), v arg z) (libstdc++).  The order
// recurrences and the asymptotic series of CJYVA multiply by 1/z and 1/z^2
// (one Smith quotient each per call) where the reference divides at every
// step: a rounding-level difference that takes long chains of dependent
// divisions off a device thread.
#pragma once

#include <hip/hip_runtime.h>

namespace bimsim {
namespace zj {

struct zc {
    double r, i;
};

__host__ __device__ inline zc z_(double r, double i) { return zc{r, i}; }
__host__ __device__ inline zc zadd(zc a, zc b) { return zc{a.r + b.r, a.i + b.i}; }
__host__ __device__ inline zc zsub(zc a, zc b) { return zc{a.r - b.r, a.i - b.i}; }
__host__ __device__ inline zc zneg(zc a) { return zc{-a.r, -a.i}; }
__host__ __device__ inline zc zmul(zc a, zc b) {
    return zc{a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
__host__ __device__ inline zc zsc(double s, zc a) { return zc{a.r * s, a.i * s}; }
__host__ __device__ inline zc zdivr(zc a, double s) { return zc{a.r / s, a.i / s}; }
// Smith's algorithm
__host__ __device__ inline zc zdiv(zc x, zc y) {
    const double a = x.r, b = x.i, c = y.r, d = y.i;
    if (fabs(c) < fabs(d)) {
        const double ratio = c / d;
        const double denom = (c * ratio) + d;
        return zc{((a * ratio) + b) / denom, ((b * ratio) - a) / denom};
    }
    const double ratio = d / c;
    const double denom = (d * ratio) + c;
    return zc{((b * ratio) + a) / denom, (b - (a * ratio)) / denom};
}
__host__ __device__ inline zc rdiv(double x, zc y) { return zdiv(zc{x, 0.0}, y); }
__host__ __device__ inline double zabs(zc a) { return hypot(a.r, a.i); }

__device__ inline zc zsqrt(zc z) {
    const double x = z.r, y = z.i;
    if (y == 0.0) {
        if (x < 0.0) return zc{0.0, copysign(sqrt(-x), y)};
        return zc{fabs(sqrt(x)), copysign(0.0, y)};
    }
    if (x == 0.0) {
        const double r = sqrt(0.5 * fabs(y));
        return zc{r, copysign(r, y)};
    }
    const double d = hypot(x, y);
    double r, s;
    if (x > 0.0) {
        r = sqrt(0.5 * (d + x));
        s = 0.5 * (y / r);
    } else {
        s = sqrt(0.5 * (d - x));
        r = fabs(0.5 * (y / s));
    }
    return zc{r, copysign(s, y)};
}
__device__ inline zc zlog(zc z) { return zc{log(hypot(z.r, z.i)), atan2(z.i, z.r)}; }
__device__ inline zc zpow(zc z, double v) {
    if (z.i == 0.0 && z.r > 0.0) return zc{pow(z.r, v), 0.0};
    const zc t = zlog(z);
    const double rho = exp(v * t.r), th = v * t.i;
    return zc{rho * cos(th), rho * sin(th)};
}
__device__ inline zc zcos(zc z) { return zc{cosh(z.i) * cos(z.r), -(sinh(z.i) * sin(z.r))}; }
__device__ inline zc zsin(zc z) { return zc{sin(z.r) * cosh(z.i), cos(z.r) * sinh(z.i)}; }
__device__ inline zc zexp(zc z) {
    const double e = exp(z.r);
    return zc{e * cos(z.i), e * sin(z.i)};
}

// 1/Gamma(z) = sum c_k z^k, Abramowitz & Stegun 6.1.34 (25 terms, as the
// reference gamma.cpp:20-45).
__device__ inline double gamma_fn(double x) {
    const double c[25] = {1.0,
                          0.5772156649015329,
                          -0.6558780715202538,
                          -0.420026350340952e-1,
                          0.1665386113822915,
                          -0.421977345555443e-1,
                          -0.9621971527877e-2,
                          0.7218943246663e-2,
                          -0.11651675918591e-2,
                          -0.2152416741149e-3,
                          0.1280502823882e-3,
                          -0.201348547807e-4,
                          -0.12504934821e-5,
                          0.1133027232e-5,
                          -0.2056338417e-6,
                          0.6116095e-8,
                          0.50020075e-8,
                          -0.11812746e-8,
                          0.1043427e-9,
                          0.77823e-11,
                          -0.36968e-11,
                          0.51e-12,
                          -0.206e-13,
                          -0.54e-14,
                          0.14e-14};
    if (x > 171.0) return 1e308;
    if (x == (double)(int)x) {
        if (x > 0.0) {
            double f = 1.0;
            for (int k = 2; k < x; ++k) f *= k;
            return f;
        }
        return 1e308;
    }
    const bool big = fabs(x) > 1.0;
    double z = x, prod = 1.0;
    if (big) {
        z = fabs(x);
        const int m = (int)z;
        for (int k = 1; k <= m; ++k) prod *= (z - k);
        z -= m;
    }
    double acc = c[24];
    for (int k = 23; k >= 0; --k) acc = acc * z + c[k];
    double g = 1.0 / (acc * z);
    if (big) {
        g *= prod;
        if (x < 0.0) g = -M_PI / (x * g * sin(M_PI * x));
    }
    return g;
}

__device__ inline double envelope(int n, double x) {
    return 0.5 * log10(6.28 * n) - n * log10(1.36 * x / n);
}
__device__ inline int secant(int n0, double x, double target) {
    double f0 = envelope(n0, x) - target;
    int n1 = n0 + 5;
    double f1 = envelope(n1, x) - target;
    int nn = n1;
    for (int it = 0; it < 20; ++it) {
        nn = (int)(n1 - (n1 - n0) / (1.0 - f0 / f1));
        const double f = envelope(nn, x) - target;
        if (abs(nn - n1) < 1) break;
        n0 = n1;
        f0 = f1;
        n1 = nn;
        f1 = f;
    }
    return nn;
}
// Miller start order for a relative magnitude 10^-mp (MSTA1)
__device__ inline int msta1(double x, int mp) {
    const double a0 = fabs(x);
    return secant((int)(1.1 * a0) + 1, a0, (double)mp);
}
// Miller start order for mp significant digits at order n (MSTA2)
__device__ inline int msta2(double x, int n, int mp) {
    const double a0 = fabs(x);
    const double half = 0.5 * mp;
    const double ejn = envelope(n, a0);
    double target;
    int n0;
    if (ejn <= half) {
        target = mp;
        n0 = (int)(1.1 * a0);
        if (n0 < 1) n0 = 1;
    } else {
        target = half + ejn;
        n0 = n;
    }
    return secant(n0, a0, target) + 10;
}

__device__ inline double sq(double v) { return v * v; }

// Start orders of CJYVA's Miller recurrence for |z| = a0 and top order n
// (MSTA1 for 10^-200, MSTA2 for 15 digits at order n).  They depend on
// nothing else, so a caller may compute them beside cjyva_start.
struct CjyvaOrders {
    int m1, m2;
};
__device__ inline CjyvaOrders cjyva_orders(double a0, int n) {
    return CjyvaOrders{msta1(a0, 200), msta2(a0, n, 15)};
}
// One of the two (which = 0: m1, 1: m2) with the same secant code for both,
// so that lanes computing different ones stay together: msta1 and msta2
// differ only in the start and target of the search.
__device__ inline int cjyva_order(double a0, int n, int which) {
    int n0;
    double target;
    if (which == 0) {
        n0 = (int)(1.1 * a0) + 1;
        target = 200.0;
    } else {
        const double half = 0.5 * 15;
        const double ejn = envelope(n, a0);
        if (ejn <= half) {
            target = 15;
            n0 = (int)(1.1 * a0);
            if (n0 < 1) n0 = 1;
        } else {
            target = half + ejn;
            n0 = n;
        }
    }
    return secant(n0, a0, target) + (which ? 10 : 0);
}

// CJYVA up to its recurrences: J, Y of orders v0 and v0 + 1 (done: |z| is
// below 1e-100 and every array is written).
struct CjyvaStart {
    zc j0, j1, y0, y1;
    double a0, v0;
    int n;
    bool done;
};

// CJYVA's first part (see cjyva): the power series or asymptotic
// expansions for orders v0 and v0 + 1; writes the arrays only when done.
__device__ inline CjyvaStart cjyva_start(double v, zc z, zc* cjv, zc* cyv, zc* cjvp, zc* cyvp,
                                         bool typo, bool want_y, bool want_j) {
    const double eps = 1e-15, euler = 0.5772156649015329;
    const zc one = {1.0, 0.0}, zero = {0.0, 0.0};
    const double a0 = zabs(z);
    zc z1 = z;
    const zc z2 = zmul(z, z);
    int n = (int)v;
    const double v0 = v - n;
    const double pv0 = M_PI * v0, pv1 = M_PI * (1.0 + v0);
    if (a0 < 1.0e-100) {
        for (int k = 0; k <= n; ++k) {
            cjv[k] = zero;
            cyv[k] = zc{-1e308, 0.0};
            cjvp[k] = zero;
            cyvp[k] = zc{1e308, 0.0};
        }
        if (v0 == 0.0) {
            cjv[0] = one;
            cjvp[1] = zc{0.5, 0.0};
        } else {
            cjvp[0] = zc{1e308, 0.0};
        }
        return CjyvaStart{zero, zero, zero, zero, a0, v0, n, true};
    }
    if (z1.r < 0.0) z1 = zneg(z);
    zc j0, j1, y0 = zero, y1 = zero;
    if (a0 <= 12.0) {
        for (int l = 0; l < 2; ++l) {
            const double vl = v0 + l;
            zc s = one, t = one;
            for (int k = 1; k <= 40; ++k) {
                t = zmul(t, zdivr(zsc(-0.25, z2), k * (k + vl)));
                s = zadd(s, t);
                if (zabs(t) < zabs(s) * eps) break;
            }
            const double g = gamma_fn(1.0 + vl);
            const zc sc = zdivr(zpow(zsc(0.5, z1), vl), g);
            if (l == 0) j0 = zmul(s, sc);
            else j1 = zmul(s, sc);
        }
    } else {
        const int kz = a0 >= 50.0 ? 8 : (a0 >= 35.0 ? 10 : 11);
        const zc z2r = rdiv(1.0, z2);   // the terms divide by k(2k -+ 1) z^2
        // the four series (P and Q of orders v0 and v0 + 1) are independent:
        // one loop steps all four, each with the operations it has alone
        double mu[2];
        zc pp[2], tp[2], qq[2], tq[2];
#pragma unroll
        for (int l = 0; l < 2; ++l) {
            mu[l] = 4.0 * (l + v0) * (l + v0);
            pp[l] = tp[l] = qq[l] = tq[l] = one;
        }
        for (int k = 1; k <= kz; ++k) {
#pragma unroll
            for (int l = 0; l < 2; ++l) {
                const zc nump = zsc(mu[l] - sq(4.0 * k - 1.0),
                                    zsc(mu[l] - sq(4.0 * k - 3.0), zsc(-0.78125e-2, tp[l])));
                tp[l] = zmul(zsc(1.0 / (k * (2.0 * k - 1.0)), nump), z2r);
                pp[l] = zadd(pp[l], tp[l]);
                const zc numq = zsc(mu[l] - sq(4.0 * k + 1.0),
                                    zsc(mu[l] - sq(4.0 * k - 1.0), zsc(-0.78125e-2, tq[l])));
                tq[l] = zmul(zsc(1.0 / (k * (2.0 * k + 1.0)), numq), z2r);
                qq[l] = zadd(qq[l], tq[l]);
            }
        }
        const zc amp = zsqrt(rdiv(M_2_PI, z1));   // the same for both orders
#pragma unroll
        for (int l = 0; l < 2; ++l) {
            const zc p = pp[l];
            const zc q = zmul(qq[l], rdiv(0.125 * (mu[l] - 1.0), z1));
            const zc chi = zsub(z1, zc{(0.5 * (l + v0) + 0.25) * M_PI, 0.0});
            const zc c = zcos(chi), s = zsin(chi);
            const zc jj = zmul(amp, zsub(zmul(p, c), zmul(q, s)));
            if (l == 0) {
                j0 = jj;
                y0 = typo ? zmul(amp, zadd(zadd(zmul(p, s), q), c))
                          : zmul(amp, zadd(zmul(p, s), zmul(q, c)));
            } else {
                j1 = jj;
                y1 = zmul(amp, zadd(zmul(p, s), zmul(q, c)));
            }
        }
    }
    if (a0 <= 12.0 && want_y) {
        if (v0 != 0.0) {
            zc u[2];
            for (int l = 0; l < 2; ++l) {
                const double vl = v0 + l;
                zc s = one, t = one;
                for (int k = 1; k <= 40; ++k) {
                    t = zmul(t, zdivr(zsc(-0.25, z2), k * (k - vl)));
                    s = zadd(s, t);
                    if (zabs(t) < zabs(s) * eps) break;
                }
                const double g = gamma_fn(1.0 - vl);
                u[l] = zmul(s, zdivr(zpow(rdiv(2.0, z1), vl), g));
            }
            y0 = zdivr(zsub(zsc(cos(pv0), j0), u[0]), sin(pv0));
            y1 = zdivr(zsub(zsc(cos(pv1), j1), u[1]), sin(pv1));
        } else {
            const zc ec = zadd(zlog(zsc(0.5, z1)), zc{euler, 0.0});
            zc cs0 = zero, r0 = one;
            double w0 = 0.0;
            for (int k = 1; k <= 30; ++k) {
                w0 += 1.0 / k;
                r0 = zmul(r0, zdivr(zsc(-0.25, z2), (double)(k * k)));
                cs0 = zadd(cs0, zsc(w0, r0));
            }
            y0 = zsc(M_2_PI, zsub(zmul(ec, j0), cs0));
            zc cs1 = one, r1 = one;
            double w1 = 0.0;
            for (int k = 1; k <= 30; ++k) {
                w1 += 1.0 / k;
                r1 = zmul(r1, zdivr(zsc(-0.25, z2), k * (k + 1.0)));
                cs1 = zadd(cs1, zsc(2.0 * w1 + 1.0 / (k + 1.0), r1));
            }
            y1 = zsc(M_2_PI, zsub(zsub(zmul(ec, j1), rdiv(1.0, z1)), zmul(zsc(0.25, z1), cs1)));
        }
    }
    if (z.r < 0.0) {
        const zc f0 = zexp(zc{0.0, pv0}), f1 = zexp(zc{0.0, pv1});
        const zc ci = {0.0, 1.0};
        if (z.i < 0.0) {
            y0 = zsub(zmul(f0, y0), zmul(zsc(cos(pv0), zsc(2.0, ci)), j0));
            y1 = zsub(zmul(f1, y1), zmul(zsc(cos(pv1), zsc(2.0, ci)), j1));
            j0 = zdiv(j0, f0);
            j1 = zdiv(j1, f1);
        } else if (z.i > 0.0) {
            y0 = zadd(zdiv(y0, f0), zmul(zsc(cos(pv0), zsc(2.0, ci)), j0));
            y1 = zadd(zdiv(y1, f1), zmul(zsc(cos(pv1), zsc(2.0, ci)), j1));
            j0 = zmul(j0, f0);
            j1 = zmul(j1, f1);
        }
    }
    return CjyvaStart{j0, j1, y0, y1, a0, v0, n, false};
}

// The rest of CJYVA from cjyva_start's values: the recurrences for orders
// v0 + 2 .. v0 + n and the derivatives.  ord: the start orders of
// cjyva_orders(st.a0, st.n), or null to compute them here where needed.
__device__ inline void cjyva_rest(const CjyvaStart& st, zc z, const CjyvaOrders* ord, zc* cjv,
                                  zc* cyv, zc* cjvp, zc* cyvp, bool want_y, bool want_j) {
    const zc one = {1.0, 0.0}, zero = {0.0, 0.0};
    const zc j0 = st.j0, j1 = st.j1, y0 = st.y0, y1 = st.y1;
    const double a0 = st.a0, v0 = st.v0;
    int n = st.n;
    if (want_j) {
        cjv[0] = j0;
        cjv[1] = j1;
    }
    const zc zr = rdiv(1.0, z);
    if (n >= 2 && n <= (int)(0.25 * a0)) {
        zc f0 = j0, f1 = j1;
        for (int k = 2; k <= n && want_j; ++k) {
            const zc f = zsub(zmul(zsc(2.0 * (k + v0 - 1.0), f1), zr), f0);
            cjv[k] = f;
            f0 = f1;
            f1 = f;
        }
    } else if (n >= 2) {
        int m = ord ? ord->m1 : msta1(a0, 200);
        if (m < n) {
            for (int k = m + 1; k <= n; ++k) {
                if (want_j) cjv[k] = cjvp[k] = zero;
                if (want_y) cyv[k] = cyvp[k] = zero;
            }
            n = m;
        } else if (want_j) {
            m = ord ? ord->m2 : msta2(a0, n, 15);
        }
        if (want_j) {
            // each step's factor is formed one step ahead, beside the
            // previous step's product; orders above n go to the spare entry
            // n + 1 (m >= n here), so the step has no branch
            zc f2 = zero, f1 = {1.0e-100, 0.0}, f = zero;
            double c = 2.0 * (v0 + m + 1.0);
            for (int k = m; k >= 0; --k) {
                const double cn = 2.0 * (v0 + (k - 1) + 1.0);
                f = zsub(zmul(zsc(c, f1), zr), f2);
                cjv[min(k, n + 1)] = f;
                f2 = f1;
                f1 = f;
                c = cn;
            }
            const zc cs = zabs(j0) > zabs(j1) ? zdiv(j0, f) : zdiv(j1, f2);
            #pragma unroll 4
            for (int k = 0; k <= n; ++k) cjv[k] = zmul(cjv[k], cs);
        }
    }
    if (want_j) {
        cjvp[0] = zsub(zmul(zsc(v0, cjv[0]), zr), cjv[1]);
        #pragma unroll 4
        for (int k = 1; k <= n; ++k) cjvp[k] = zadd(zmul(zsc(-(k + v0), cjv[k]), zr), cjv[k - 1]);
    }
    if (!want_y) return;
    cyv[0] = y0;
    cyv[1] = y1;
    int lb = 0;
    if (z.i != 0.0) {
        // |y_{k-2}| and |y_{k-1}| carried along (one magnitude per order)
        const double ya0 = zabs(y0);
        zc g0 = y0, g1 = y1;
        double a_g0 = ya0, a_g1 = zabs(y1);
        for (int k = 2; k <= n; ++k) {
            const zc g = zsub(zmul(zsc(2.0 * (v0 + k - 1.0), g1), zr), g0);
            const double yak = zabs(g), ya1 = a_g0;
            if (yak < ya0 && yak < ya1) lb = k;
            cyv[k] = g;
            g0 = g1;
            g1 = g;
            a_g0 = a_g1;
            a_g1 = yak;
        }
    } else {
        // the same recurrence (each factor formed one step ahead); lb is
        // read only by the complex-z correction
        zc g0 = y0, g1 = y1;
        double c = 2.0 * (v0 + 2 - 1.0);
        for (int k = 2; k <= n; ++k) {
            const double cn = 2.0 * (v0 + (k + 1) - 1.0);
            const zc g = zsub(zmul(zsc(c, g1), zr), g0);
            cyv[k] = g;
            g0 = g1;
            g1 = g;
            c = cn;
        }
    }
    int lb0 = 0;
    if (lb > 4 && z.i != 0.0) {
        const zc piz = zsc(M_PI, z);
        // the reference iterates until lb settles; a device thread must end,
        // so the pass is capped (it settles in one or two rounds in practice)
        for (int round = 0; lb != lb0 && round < 64; ++round) {
            zc h0 = zero, h1 = zero, h2 = one;
            lb0 = lb;
            for (int k = lb; k >= 1; --k) {
                h0 = zsub(zmul(zsc(2.0 * (k + v0), h1), zr), h2);
                h2 = h1;
                h1 = h0;
            }
            const zc p12 = h0, p22 = h2;
            h2 = zero;
            h1 = one;
            for (int k = lb; k >= 1; --k) {
                h0 = zsub(zmul(zsc(2.0 * (k + v0), h1), zr), h2);
                h2 = h1;
                h1 = h0;
            }
            const zc p11 = h0, p21 = h2;
            if (lb == n)
                cjv[lb + 1] = zsub(zmul(zsc(2.0 * (lb + v0), cjv[lb]), zr), cjv[lb - 1]);
            if (zabs(cjv[0]) > zabs(cjv[1])) {
                cyv[lb + 1] = zdiv(zsub(zmul(cjv[lb + 1], y0), zdiv(zsc(2.0, p11), piz)), cjv[0]);
                cyv[lb] = zdiv(zadd(zmul(cjv[lb], y0), zdiv(zsc(2.0, p12), piz)), cjv[0]);
            } else {
                cyv[lb + 1] = zdiv(zsub(zmul(cjv[lb + 1], y1), zdiv(zsc(2.0, p21), piz)), cjv[1]);
                cyv[lb] = zdiv(zadd(zmul(cjv[lb], y1), zdiv(zsc(2.0, p22), piz)), cjv[1]);
            }
            zc yl2 = cyv[lb + 1], yl1 = cyv[lb];
            for (int k = lb - 1; k >= 0; --k) {
                const zc yk = zsub(zmul(zsc(2.0 * (k + v0 + 1.0), yl1), zr), yl2);
                cyv[k] = yk;
                yl2 = yl1;
                yl1 = yk;
            }
            yl1 = cyv[lb];
            yl2 = cyv[lb + 1];
            for (int k = lb + 1; k < n; ++k) {
                const zc yk = zsub(zmul(zsc(2.0 * (k + v0), yl2), zr), yl1);
                cyv[k + 1] = yk;
                yl1 = yl2;
                yl2 = yk;
            }
            double a_prev = zabs(cyv[1]);
            for (int k = 2; k <= n; ++k) {
                const double a_k = zabs(cyv[k]);
                if (a_k < a_prev) lb = k;
                a_prev = a_k;
            }
        }
    }
    cyvp[0] = zsub(zmul(zsc(v0, cyv[0]), zr), cyv[1]);
    #pragma unroll 4
    for (int k = 1; k <= n; ++k) cyvp[k] = zsub(cyv[k - 1], zmul(zsc(k + v0, cyv[k]), zr));
}

// Complex-argument J_{v0+k}, Y_{v0+k} and derivatives, k = 0..floor(v)
// (CJYVA).  Arrays hold floor(v)+2 entries.  typo reproduces the reference's
// y_{v0} asymptotic defect (cbessjy.cpp:541).
//
// The two halves can run apart (on different lanes): want_y = false stops
// after J and J' (their values do not depend on the Y part) and leaves cyv,
// cyvp untouched; want_j = false, for real z only (the complex-z Y
// correction reads J), skips the Miller recurrence and leaves cjv, cjvp
// untouched.  Each call writes only the arrays of the halves it computes.
__device__ inline void cjyva(double v, zc z, zc* cjv, zc* cyv, zc* cjvp, zc* cyvp, bool typo,
                             bool want_y = true, bool want_j = true) {
    const CjyvaStart st = cjyva_start(v, z, cjv, cyv, cjvp, cyvp, typo, want_y, want_j);
    if (!st.done) cjyva_rest(st, z, nullptr, cjv, cyv, cjvp, cyvp, want_y, want_j);
}

// Real-argument J_{v0+k}, Y_{v0+k} and derivatives, k = 0..floor(v) (JYV,
// reference bessjy.cpp:536-707).  Arrays hold floor(v)+2 entries; orders the
// Miller start truncates stay 0.
__device__ inline void jyv(double v, double x, double* jv, double* yv, double* djv, double* dyv) {
    const double eps = 1e-15, euler = 0.5772156649015329;
    const double x2 = x * x;
    int n = (int)v;
    const double v0 = v - n;
    if (x < 1e-15) {
        for (int k = 0; k <= n; ++k) {
            jv[k] = 0.0;
            yv[k] = -1e308;
            djv[k] = 0.0;
            dyv[k] = 1e308;
        }
        if (v0 == 0.0) {
            jv[0] = 1.0;
            djv[1] = 0.5;
        } else {
            djv[0] = 1e308;
        }
        return;
    }
    double bj0 = 0.0, bj1 = 0.0, by0 = 0.0, by1 = 0.0;
    if (x <= 12.0) {
        for (int l = 0; l < 2; ++l) {
            const double vl = v0 + l;
            double s = 1.0, t = 1.0;
            for (int k = 1; k <= 40; ++k) {
                t *= -0.25 * x2 / (k * (k + vl));
                s += t;
                if (fabs(t) < fabs(s) * eps) break;
            }
            const double val = s * (pow(0.5 * x, vl) / gamma_fn(1.0 + vl));
            if (l == 0) bj0 = val;
            else bj1 = val;
        }
    } else {
        const int kz = x >= 50.0 ? 8 : (x >= 35.0 ? 10 : 11);
        for (int l = 0; l < 2; ++l) {
            const double mu = 4.0 * (l + v0) * (l + v0);
            double p = 1.0, tp = 1.0;
            for (int k = 1; k <= kz; ++k) {
                tp *= (-0.78125e-2) * (mu - sq(4.0 * k - 3.0)) * (mu - sq(4.0 * k - 1.0)) /
                      (k * (2.0 * k - 1.0) * x2);
                p += tp;
            }
            double q = 1.0, tq = 1.0;
            for (int k = 1; k <= kz; ++k) {
                tq *= (-0.78125e-2) * (mu - sq(4.0 * k - 1.0)) * (mu - sq(4.0 * k + 1.0)) /
                      (k * (2.0 * k + 1.0) * x2);
                q += tq;
            }
            q *= 0.125 * (mu - 1.0) / x;
            const double chi = x - (0.5 * (l + v0) + 0.25) * M_PI;
            const double amp = sqrt(M_2_PI / x);
            const double c = cos(chi), sn = sin(chi);
            if (l == 0) {
                bj0 = amp * (p * c - q * sn);
                by0 = amp * (p * sn + q * c);
            } else {
                bj1 = amp * (p * c - q * sn);
                by1 = amp * (p * sn + q * c);
            }
        }
    }
    jv[0] = bj0;
    jv[1] = bj1;
    djv[0] = v0 * jv[0] / x - jv[1];
    djv[1] = -(1.0 + v0) * jv[1] / x + jv[0];
    if (n >= 2 && n <= (int)(0.9 * x)) {
        double f0 = bj0, f1 = bj1;
        for (int k = 2; k <= n; ++k) {
            const double f = 2.0 * (k + v0 - 1.0) * f1 / x - f0;
            jv[k] = f;
            f0 = f1;
            f1 = f;
        }
    } else if (n >= 2) {
        int m = msta1(x, 200);
        if (m < n) {
            for (int k = m + 1; k <= n; ++k) jv[k] = yv[k] = djv[k] = dyv[k] = 0.0;
            n = m;
        } else {
            m = msta2(x, n, 15);
        }
        double f2 = 0.0, f1 = 1.0e-100, f = 0.0;
        for (int k = m; k >= 0; --k) {
            f = 2.0 * (v0 + k + 1.0) * f1 / x - f2;
            if (k <= n) jv[k] = f;
            f2 = f1;
            f1 = f;
        }
        const double cs = fabs(bj0) > fabs(bj1) ? bj0 / f : bj1 / f2;
        for (int k = 0; k <= n; ++k) jv[k] *= cs;
    }
    for (int k = 2; k <= n; ++k) djv[k] = -(k + v0) * jv[k] / x + jv[k - 1];
    if (x <= 12.0) {
        if (v0 != 0.0) {
            double bu[2];
            for (int l = 0; l < 2; ++l) {
                const double vl = v0 + l;
                double s = 1.0, t = 1.0;
                for (int k = 1; k <= 40; ++k) {
                    t *= -0.25 * x2 / (k * (k - vl));
                    s += t;
                    if (fabs(t) < fabs(s) * eps) break;
                }
                bu[l] = s * (pow(2.0 / x, vl) / gamma_fn(1.0 - vl));
            }
            const double pv0 = M_PI * v0, pv1 = M_PI * (1.0 + v0);
            by0 = (bj0 * cos(pv0) - bu[0]) / sin(pv0);
            by1 = (bj1 * cos(pv1) - bu[1]) / sin(pv1);
        } else {
            const double ec = log(0.5 * x) + euler;
            double cs0 = 0.0, w0 = 0.0, r0 = 1.0;
            for (int k = 1; k <= 30; ++k) {
                w0 += 1.0 / k;
                r0 *= -0.25 * x2 / (k * k);
                cs0 += r0 * w0;
            }
            by0 = M_2_PI * (ec * bj0 - cs0);
            double cs1 = 1.0, w1 = 0.0, r1 = 1.0;
            for (int k = 1; k <= 30; ++k) {
                w1 += 1.0 / k;
                r1 *= -0.25 * x2 / (k * (k + 1));
                cs1 += r1 * (2.0 * w1 + 1.0 / (k + 1.0));
            }
            by1 = M_2_PI * (ec * bj1 - 1.0 / x - 0.25 * x * cs1);
        }
    }
    yv[0] = by0;
    yv[1] = by1;
    double g0 = by0, g1 = by1;
    for (int k = 2; k <= n; ++k) {
        const double g = 2.0 * (v0 + k - 1.0) * g1 / x - g0;
        yv[k] = g;
        g0 = g1;
        g1 = g;
    }
    dyv[0] = v0 * yv[0] / x - yv[1];
    for (int k = 1; k <= n; ++k) dyv[k] = -(k + v0) * yv[k] / x + yv[k - 1];
}

// Spherical j_l, y_l (real argument) for l = 0..nl (bessjyv_sph,
// bessjy.cpp:709-728); arrays hold nl+2 entries.
__device__ inline void sph_jy_real(int nl, double x, bool quirks, double* j, double* y, double* jp,
                                   double* yp) {
    for (int k = 0; k <= nl + 1; ++k) j[k] = y[k] = jp[k] = yp[k] = 0.0;
    jyv(nl + 0.5, x, j, y, jp, yp);
    const double pi = quirks ? 3.14159 : M_PI;
    for (int k = 0; k <= nl; ++k) {
        const double s = sqrt(pi / (x * 2.0));
        j[k] = j[k] * s;
        y[k] = y[k] * s;
        jp[k] = -1.0 / (x * 2.0) * j[k] + jp[k] * s;
        yp[k] = -1.0 / (x * 2.0) * y[k] + yp[k] * s;
    }
}

// The cylindrical-to-spherical step of sph_jy_cplx on orders 0..nl:
// j_l = sqrt(pi / (2z)) J_{l+1/2}, j'_l = -j_l / (2z) + sqrt(pi / (2z)) J'
// (pi = 3.14159 with quirks, as the reference).
__device__ inline void sph_jy_scale(int nl, zc z, bool quirks, zc* j, zc* y, zc* jp, zc* yp,
                                    bool want_y, bool want_j) {
    const double pi = quirks ? 3.14159 : M_PI;
    const zc z2 = zsc(2.0, z);
    // the reference evaluates these two per order; they do not depend on it
    const zc s = zsqrt(rdiv(pi, z2));
    const zc inv = rdiv(-1.0, z2);
#pragma unroll 4
    for (int k = 0; k <= nl; ++k) {
        if (want_j) {
            const zc jk = zmul(j[k], s), jpk = jp[k];
            j[k] = jk;
            jp[k] = zadd(zmul(inv, jk), zmul(jpk, s));
        }
        if (want_y) {
            const zc yk = zmul(y[k], s), ypk = yp[k];
            y[k] = yk;
            yp[k] = zadd(zmul(inv, yk), zmul(ypk, s));
        }
    }
}

// Spherical j_l, y_l, j'_l, y'_l for l = 0..nl (cbessjyva_sph,
// cbessjy.cpp:718-737): scale by sqrt(pi/(2z)); the reference uses
// pi = 3.14159 there.
// want_j / want_y select the halves as in cjyva (want_j is forced for
// complex z when Y is wanted); only the selected arrays are written.
__device__ inline void sph_jy_cplx(int nl, zc z, bool quirks, zc* j, zc* y, zc* jp, zc* yp,
                                   bool want_y = true, bool want_j = true) {
    if (want_y && z.i != 0.0) want_j = true;
    for (int k = 0; k <= nl + 1; ++k) {
        if (want_j) j[k] = jp[k] = zc{0.0, 0.0};
        if (want_y) y[k] = yp[k] = zc{0.0, 0.0};
    }
    cjyva(nl + 0.5, z, j, y, jp, yp, quirks, want_y, want_j);
    sph_jy_scale(nl, z, quirks, j, y, jp, yp, want_y, want_j);
}

}  // namespace zj
}  // namespace bimsim
