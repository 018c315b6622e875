// ORACLE -- test infrastructure only.  Nothing in the product links this file;
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use it.
//
// Restatement of the Zhang & Jin fractional-order Bessel routines used by the
// reference (S. Zhang and J. Jin, "Computation of Special Functions", Wiley
// 1996: routines GAMMA2, MSTA1, MSTA2, JYV, CJYVA), step for step as the
// reference's C++ transcription runs them, so that the FP64 Mie coefficients
// and radial LUTs of the reference can be reproduced to rounding:
//   gamma        reference gamma.cpp:14-82 (1/Gamma series, A&S 6.1.34)
//   msta1/msta2  reference bessjy.cpp:268-319
//   jyv          reference bessjy.cpp:536-707  (real argument)
//   cjyva        reference cbessjy.cpp:456-716 (complex argument)
//   spherical    reference bessjy.cpp:709-728, cbessjy.cpp:718-737
// With quirks on, the reference's defects are kept: the y_{v0} asymptotic typo
// (cbessjy.cpp:541) and the sqrt(3.14159/(2z)) scale (bessjy.cpp:18).
// Orders the Miller start truncates (msta1 < n) are returned as 0; the
// reference leaves them uninitialised (bessjy.cpp:630-631, cbessjy.cpp:620-621).
#include "oracle_internal.h"

#include <cmath>
#include <cstdlib>
#include <vector>

namespace orc {

namespace {

const double kPiRef = 3.14159;             // reference bessjy.cpp:18
const double kEuler = 0.5772156649015329;   // reference bessel.h:9
const double kSeriesEps = 1e-15;            // reference bessel.h:8

// Coefficients c_1..c_25 of 1/Gamma(z) = sum c_k z^k (Abramowitz & Stegun
// 6.1.34); the same 25 terms as reference gamma.cpp:20-45.
const double kInvGamma[25] = {
    1.0,
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

// log10 of the Bessel envelope used to pick the Miller start order.
double envelope(int n, double x) {
    return 0.5 * std::log10(6.28 * n) - n * std::log10(1.36 * x / n);
}

// Secant search on the envelope shared by msta1/msta2.
int secant_order(int n0, double x, double target) {
    double f0 = envelope(n0, x) - target;
    int n1 = n0 + 5;
    double f1 = envelope(n1, x) - target;
    int nn = n1;
    for (int it = 0; it < 20; ++it) {
        nn = (int)(n1 - (n1 - n0) / (1.0 - f0 / f1));
        double f = envelope(nn, x) - target;
        if (std::abs(nn - n1) < 1) break;
        n0 = n1;
        f0 = f1;
        n1 = nn;
        f1 = f;
    }
    return nn;
}

double sq(double v) { return v * v; }

}  // namespace

double zj_gamma(double x) {
    if (x > 171.0) return 1e308;
    if (x == (int)x) {
        if (x > 0.0) {
            double f = 1.0;
            for (int i = 2; i < x; ++i) f *= i;
            return f;
        }
        return 1e308;
    }
    double z = x, prod = 1.0;
    const bool big = std::fabs(x) > 1.0;
    if (big) {
        z = std::fabs(x);
        int m = (int)z;
        for (int k = 1; k <= m; ++k) prod *= (z - k);
        z -= m;
    }
    double acc = kInvGamma[24];
    for (int k = 23; k >= 0; --k) acc = acc * z + kInvGamma[k];
    double g = 1.0 / (acc * z);
    if (big) {
        g *= prod;
        if (x < 0.0) g = -M_PI / (x * g * std::sin(M_PI * x));
    }
    return g;
}

int zj_msta1(double x, int mp) {
    double a0 = std::fabs(x);
    return secant_order((int)(1.1 * a0) + 1, a0, (double)mp);
}

int zj_msta2(double x, int n, int mp) {
    double a0 = std::fabs(x);
    double half = 0.5 * mp;
    double ejn = envelope(n, a0);
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
    return secant_order(n0, a0, target) + 10;
}

// ---------------------------------------------------------------- real JYV
// jv, yv, djv, dyv must hold floor(v)+2 entries.  Returns the top order
// actually filled (the Miller start may truncate it).
int zj_jyv(double v, double x, double* jv, double* yv, double* djv, double* dyv) {
    const double x2 = x * x;
    int n = (int)v;
    const double v0 = v - n;
    if (x < 0.0 || v < 0.0) return -1;
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
        return n;
    }
    double bj[2], by[2] = {0.0, 0.0};
    if (x <= 12.0) {
        // Ascending series for J_{v0}, J_{v0+1}.
        for (int l = 0; l < 2; ++l) {
            const double vl = v0 + l;
            double s = 1.0, t = 1.0;
            for (int k = 1; k <= 40; ++k) {
                t *= -0.25 * x2 / (k * (k + vl));
                s += t;
                if (std::fabs(t) < std::fabs(s) * kSeriesEps) break;
            }
            bj[l] = s * (std::pow(0.5 * x, vl) / zj_gamma(1.0 + vl));
        }
    } else {
        // Hankel asymptotic expansion, P and Q with kz terms.
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
            const double amp = std::sqrt(M_2_PI / x);
            const double c = std::cos(chi), s = std::sin(chi);
            bj[l] = amp * (p * c - q * s);
            by[l] = amp * (p * s + q * c);
        }
    }
    jv[0] = bj[0];
    jv[1] = bj[1];
    djv[0] = v0 * jv[0] / x - jv[1];
    djv[1] = -(1.0 + v0) * jv[1] / x + jv[0];
    if (n >= 2 && n <= (int)(0.9 * x)) {
        double f0 = bj[0], f1 = bj[1];
        for (int k = 2; k <= n; ++k) {
            double f = 2.0 * (k + v0 - 1.0) * f1 / x - f0;
            jv[k] = f;
            f0 = f1;
            f1 = f;
        }
    } else if (n >= 2) {
        int m = zj_msta1(x, 200);
        if (m < n) {
            for (int k = m + 1; k <= n; ++k) jv[k] = yv[k] = djv[k] = dyv[k] = 0.0;
            n = m;
        } else {
            m = zj_msta2(x, n, 15);
        }
        double f2 = 0.0, f1 = 1.0e-100, f = 0.0;
        for (int k = m; k >= 0; --k) {
            f = 2.0 * (v0 + k + 1.0) * f1 / x - f2;
            if (k <= n) jv[k] = f;
            f2 = f1;
            f1 = f;
        }
        const double cs = std::fabs(bj[0]) > std::fabs(bj[1]) ? bj[0] / f : bj[1] / f2;
        for (int k = 0; k <= n; ++k) jv[k] *= cs;
    }
    for (int k = 2; k <= n; ++k) djv[k] = -(k + v0) * jv[k] / x + jv[k - 1];
    if (x <= 12.0) {
        if (v0 != 0.0) {
            // Y from J_{-v}: Y_v = (J_v cos(v pi) - J_{-v}) / sin(v pi).
            double bu[2];
            for (int l = 0; l < 2; ++l) {
                const double vl = v0 + l;
                double s = 1.0, t = 1.0;
                for (int k = 1; k <= 40; ++k) {
                    t *= -0.25 * x2 / (k * (k - vl));
                    s += t;
                    if (std::fabs(t) < std::fabs(s) * kSeriesEps) break;
                }
                bu[l] = s * (std::pow(2.0 / x, vl) / zj_gamma(1.0 - vl));
            }
            const double pv0 = M_PI * v0, pv1 = M_PI * (1.0 + v0);
            by[0] = (bj[0] * std::cos(pv0) - bu[0]) / std::sin(pv0);
            by[1] = (bj[1] * std::cos(pv1) - bu[1]) / std::sin(pv1);
        } else {
            const double ec = std::log(0.5 * x) + kEuler;
            double cs0 = 0.0, w0 = 0.0, r0 = 1.0;
            for (int k = 1; k <= 30; ++k) {
                w0 += 1.0 / k;
                r0 *= -0.25 * x2 / (k * k);
                cs0 += r0 * w0;
            }
            by[0] = M_2_PI * (ec * bj[0] - cs0);
            double cs1 = 1.0, w1 = 0.0, r1 = 1.0;
            for (int k = 1; k <= 30; ++k) {
                w1 += 1.0 / k;
                r1 *= -0.25 * x2 / (k * (k + 1));
                cs1 += r1 * (2.0 * w1 + 1.0 / (k + 1.0));
            }
            by[1] = M_2_PI * (ec * bj[1] - 1.0 / x - 0.25 * x * cs1);
        }
    }
    yv[0] = by[0];
    yv[1] = by[1];
    double g0 = by[0], g1 = by[1];
    for (int k = 2; k <= n; ++k) {
        double g = 2.0 * (v0 + k - 1.0) * g1 / x - g0;
        yv[k] = g;
        g0 = g1;
        g1 = g;
    }
    dyv[0] = v0 * yv[0] / x - yv[1];
    for (int k = 1; k <= n; ++k) dyv[k] = -(k + v0) * yv[k] / x + yv[k - 1];
    return n;
}

// ------------------------------------------------------------- complex CJYVA
int zj_cjyva(double v, cd z, cd* cjv, cd* cyv, cd* cjvp, cd* cyvp, bool typo) {
    const cd one(1.0, 0.0), zero(0.0, 0.0), ci(0.0, 1.0);
    const double a0 = std::abs(z);
    cd z1 = z;
    const cd z2 = z * z;
    int n = (int)v;
    const double v0 = v - n;
    const double pv0 = M_PI * v0, pv1 = M_PI * (1.0 + v0);
    if (a0 < 1.0e-100) {
        for (int k = 0; k <= n; ++k) {
            cjv[k] = zero;
            cyv[k] = cd(-1e308, 0);
            cjvp[k] = zero;
            cyvp[k] = cd(1e308, 0);
        }
        if (v0 == 0.0) {
            cjv[0] = one;
            cjvp[1] = cd(0.5, 0.0);
        } else {
            cjvp[0] = cd(1e308, 0);
        }
        return n;
    }
    if (z1.real() < 0.0) z1 = -z;
    cd j0, j1, y0, y1;
    if (a0 <= 12.0) {
        for (int l = 0; l < 2; ++l) {
            const double vl = v0 + l;
            cd s = one, t = one;
            for (int k = 1; k <= 40; ++k) {
                t *= -0.25 * z2 / (k * (k + vl));
                s += t;
                if (std::abs(t) < std::abs(s) * kSeriesEps) break;
            }
            const double g = zj_gamma(1.0 + vl);
            const cd scale = std::pow(0.5 * z1, vl) / g;
            (l == 0 ? j0 : j1) = s * scale;
        }
    } else {
        const int kz = a0 >= 50.0 ? 8 : (a0 >= 35.0 ? 10 : 11);
        for (int l = 0; l < 2; ++l) {
            const double mu = 4.0 * (l + v0) * (l + v0);
            cd p = one, tp = one;
            for (int k = 1; k <= kz; ++k) {
                tp = -0.78125e-2 * tp * (mu - sq(4.0 * k - 3.0)) * (mu - sq(4.0 * k - 1.0)) /
                     (k * (2.0 * k - 1.0) * z2);
                p += tp;
            }
            cd q = one, tq = one;
            for (int k = 1; k <= kz; ++k) {
                tq = -0.78125e-2 * tq * (mu - sq(4.0 * k - 1.0)) * (mu - sq(4.0 * k + 1.0)) /
                     (k * (2.0 * k + 1.0) * z2);
                q += tq;
            }
            q *= 0.125 * (mu - 1.0) / z1;
            const cd chi = z1 - (0.5 * (l + v0) + 0.25) * M_PI;
            const cd amp = std::sqrt(M_2_PI / z1);
            const cd c = std::cos(chi), s = std::sin(chi);
            if (l == 0) {
                j0 = amp * (p * c - q * s);
                // Reference cbessjy.cpp:541 adds q and cos(chi) instead of
                // multiplying them.
                y0 = typo ? amp * (p * s + q + c) : amp * (p * s + q * c);
            } else {
                j1 = amp * (p * c - q * s);
                y1 = amp * (p * s + q * c);
            }
        }
    }
    if (a0 <= 12.0) {
        if (v0 != 0.0) {
            cd u[2];
            for (int l = 0; l < 2; ++l) {
                const double vl = v0 + l;
                cd s = one, t = one;
                for (int k = 1; k <= 40; ++k) {
                    t *= -0.25 * z2 / (k * (k - vl));
                    s += t;
                    if (std::abs(t) < std::abs(s) * kSeriesEps) break;
                }
                const double g = zj_gamma(1.0 - vl);
                const cd scale = std::pow(2.0 / z1, vl) / g;
                u[l] = s * scale;
            }
            y0 = (j0 * std::cos(pv0) - u[0]) / std::sin(pv0);
            y1 = (j1 * std::cos(pv1) - u[1]) / std::sin(pv1);
        } else {
            const cd ec = std::log(0.5 * z1) + kEuler;
            cd cs0 = zero, r0 = one;
            double w0 = 0.0;
            for (int k = 1; k <= 30; ++k) {
                w0 += 1.0 / k;
                r0 *= -0.25 * z2 / (double)(k * k);
                cs0 += r0 * w0;
            }
            y0 = M_2_PI * (ec * j0 - cs0);
            cd cs1 = one, r1 = one;
            double w1 = 0.0;
            for (int k = 1; k <= 30; ++k) {
                w1 += 1.0 / k;
                r1 *= -0.25 * z2 / (k * (k + 1.0));
                cs1 += r1 * (2.0 * w1 + 1.0 / (k + 1.0));
            }
            y1 = M_2_PI * (ec * j1 - 1.0 / z1 - 0.25 * z1 * cs1);
        }
    }
    if (z.real() < 0.0) {
        const cd f0 = std::exp(pv0 * ci), f1 = std::exp(pv1 * ci);
        if (z.imag() < 0.0) {
            y0 = f0 * y0 - 2.0 * ci * std::cos(pv0) * j0;
            y1 = f1 * y1 - 2.0 * ci * std::cos(pv1) * j1;
            j0 /= f0;
            j1 /= f1;
        } else if (z.imag() > 0.0) {
            y0 = y0 / f0 + 2.0 * ci * std::cos(pv0) * j0;
            y1 = y1 / f1 + 2.0 * ci * std::cos(pv1) * j1;
            j0 *= f0;
            j1 *= f1;
        }
    }
    cjv[0] = j0;
    cjv[1] = j1;
    if (n >= 2 && n <= (int)(0.25 * a0)) {
        cd f0 = j0, f1 = j1;
        for (int k = 2; k <= n; ++k) {
            cd f = 2.0 * (k + v0 - 1.0) * f1 / z - f0;
            cjv[k] = f;
            f0 = f1;
            f1 = f;
        }
    } else if (n >= 2) {
        int m = zj_msta1(a0, 200);
        if (m < n) {
            for (int k = m + 1; k <= n; ++k) cjv[k] = cyv[k] = cjvp[k] = cyvp[k] = zero;
            n = m;
        } else {
            m = zj_msta2(a0, n, 15);
        }
        cd f2 = zero, f1(1.0e-100, 0.0), f = zero;
        for (int k = m; k >= 0; --k) {
            f = 2.0 * (v0 + k + 1.0) * f1 / z - f2;
            if (k <= n) cjv[k] = f;
            f2 = f1;
            f1 = f;
        }
        const cd cs = std::abs(j0) > std::abs(j1) ? j0 / f : j1 / f2;
        for (int k = 0; k <= n; ++k) cjv[k] *= cs;
    }
    cjvp[0] = v0 * cjv[0] / z - cjv[1];
    for (int k = 1; k <= n; ++k) cjvp[k] = -(k + v0) * cjv[k] / z + cjv[k - 1];
    cyv[0] = y0;
    cyv[1] = y1;
    // Forward Y recurrence; lb marks the last order where |Y| dips.
    const double ya0 = std::abs(y0);
    int lb = 0;
    {
        cd g0 = y0, g1 = y1;
        for (int k = 2; k <= n; ++k) {
            cd g = 2.0 * (v0 + k - 1.0) * g1 / z - g0;
            const double yak = std::abs(g), ya1 = std::abs(g0);
            if (yak < ya0 && yak < ya1) lb = k;
            cyv[k] = g;
            g0 = g1;
            g1 = g;
        }
    }
    // For complex z, rebuild Y around the dip from the Wronskian with J
    // (backward recurrence of the two Neumann polynomial sequences).
    int lb0 = 0;
    if (lb > 4 && z.imag() != 0.0) {
        while (lb != lb0) {
            cd h0, h1, h2;
            h2 = one;
            h1 = zero;
            lb0 = lb;
            for (int k = lb; k >= 1; --k) {
                h0 = 2.0 * (k + v0) * h1 / z - h2;
                h2 = h1;
                h1 = h0;
            }
            const cd p12 = h0, p22 = h2;
            h2 = zero;
            h1 = one;
            for (int k = lb; k >= 1; --k) {
                h0 = 2.0 * (k + v0) * h1 / z - h2;
                h2 = h1;
                h1 = h0;
            }
            const cd p11 = h0, p21 = h2;
            if (lb == n) cjv[lb + 1] = 2.0 * (lb + v0) * cjv[lb] / z - cjv[lb - 1];
            if (std::abs(cjv[0]) > std::abs(cjv[1])) {
                cyv[lb + 1] = (cjv[lb + 1] * y0 - 2.0 * p11 / (M_PI * z)) / cjv[0];
                cyv[lb] = (cjv[lb] * y0 + 2.0 * p12 / (M_PI * z)) / cjv[0];
            } else {
                cyv[lb + 1] = (cjv[lb + 1] * y1 - 2.0 * p21 / (M_PI * z)) / cjv[1];
                cyv[lb] = (cjv[lb] * y1 + 2.0 * p22 / (M_PI * z)) / cjv[1];
            }
            cd yl2 = cyv[lb + 1], yl1 = cyv[lb];
            for (int k = lb - 1; k >= 0; --k) {
                cd yk = 2.0 * (k + v0 + 1.0) * yl1 / z - yl2;
                cyv[k] = yk;
                yl2 = yl1;
                yl1 = yk;
            }
            yl1 = cyv[lb];
            yl2 = cyv[lb + 1];
            for (int k = lb + 1; k < n; ++k) {
                cd yk = 2.0 * (k + v0) * yl2 / z - yl1;
                cyv[k + 1] = yk;
                yl1 = yl2;
                yl2 = yk;
            }
            for (int k = 2; k <= n; ++k)
                if (std::abs(cyv[k]) < std::abs(cyv[k - 1])) lb = k;
        }
    }
    cyvp[0] = v0 * cyv[0] / z - cyv[1];
    for (int k = 1; k <= n; ++k) cyvp[k] = cyv[k - 1] - (k + v0) * cyv[k] / z;
    return n;
}

// ------------------------------------------------------- spherical wrappers
void sph_jy_real(int nl, double x, bool quirks, double* j, double* y, double* jp, double* yp) {
    for (int k = 0; k <= nl + 1; ++k) j[k] = y[k] = jp[k] = yp[k] = 0.0;
    zj_jyv(nl + 0.5, x, j, y, jp, yp);
    const double pi = quirks ? kPiRef : M_PI;
    for (int k = 0; k <= nl; ++k) {
        j[k] = j[k] * std::sqrt(pi / (x * 2.0));
        y[k] = y[k] * std::sqrt(pi / (x * 2.0));
        jp[k] = -1.0 / (x * 2.0) * j[k] + jp[k] * std::sqrt(pi / (x * 2.0));
        yp[k] = -1.0 / (x * 2.0) * y[k] + yp[k] * std::sqrt(pi / (x * 2.0));
    }
}

void sph_jy_cplx(int nl, cd z, bool quirks, cd* j, cd* y, cd* jp, cd* yp) {
    for (int k = 0; k <= nl + 1; ++k) j[k] = y[k] = jp[k] = yp[k] = cd(0.0, 0.0);
    zj_cjyva(nl + 0.5, z, j, y, jp, yp, quirks);
    const double pi = quirks ? kPiRef : M_PI;
    for (int k = 0; k <= nl; ++k) {
        j[k] = j[k] * std::sqrt(pi / (z * 2.0));
        y[k] = y[k] * std::sqrt(pi / (z * 2.0));
        jp[k] = -1.0 / (z * 2.0) * j[k] + jp[k] * std::sqrt(pi / (z * 2.0));
        yp[k] = -1.0 / (z * 2.0) * y[k] + yp[k] * std::sqrt(pi / (z * 2.0));
    }
}

}  // namespace orc
