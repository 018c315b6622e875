// ORACLE -- test infrastructure only (see oracle.h for the contract).
//
// CPU restatement of the reference near-field and imaging path.  Each routine
// names the reference file:line it follows.  Templates run with T = double
// (the reference with ptype = double, the plain FP64 form) or T = float (the
// reference's own single-precision program).  rc<T> reproduces the
// arithmetic of rts::complex (reference rts/math/complex.h): naive products
// and quotients, exp/sin/cos through real cos/sin/cosh/sinh.
#include "oracle.h"
#include "oracle_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {
namespace {

const float kPIf = 3.14159f;       // reference dataTypes.h:17
const float kEpsFloat = 0.000001f; // reference nearfield.h:11
const double kJMin = 0.0145;       // rts/math/spherical_bessel.h:8
const double kJMax = 0.4;          // rts/math/spherical_bessel.h:9
const double kYMin = -1e33;        // rts/math/spherical_bessel.h:10

// ------------------------------------------------------------ rts::complex
template <class T>
struct rc {
    T r, i;
    rc() : r(0), i(0) {}
    rc(T a, T b) : r(a), i(b) {}
};
template <class T> rc<T> operator+(rc<T> a, rc<T> b) { return rc<T>(a.r + b.r, a.i + b.i); }
template <class T> rc<T> operator-(rc<T> a, rc<T> b) { return rc<T>(a.r - b.r, a.i - b.i); }
template <class T> rc<T> operator*(rc<T> a, rc<T> b) {
    return rc<T>(a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r);
}
template <class T> rc<T> operator/(rc<T> a, rc<T> b) {
    T den = b.r * b.r + b.i * b.i;
    return rc<T>((a.r * b.r + a.i * b.i) / den, (-a.r * b.i + a.i * b.r) / den);
}
// complex<T>::operator*(T) and operator*(double, complex<T>) both scale by a T.
template <class T> rc<T> scl(rc<T> a, T s) { return rc<T>(a.r * s, a.i * s); }
template <class T> rc<T> cexp_(rc<T> a) {
    T e = std::exp(a.r);
    return rc<T>(e * std::cos(a.i), e * std::sin(a.i));
}
template <class T> rc<T> csin_(rc<T> a) {
    return rc<T>(std::sin(a.r) * std::cosh(a.i), std::cos(a.r) * std::sinh(a.i));
}
template <class T> rc<T> ccos_(rc<T> a) {
    return rc<T>(std::cos(a.r) * std::cosh(a.i), -(std::sin(a.r) * std::sinh(a.i)));
}

// ------------------------------------------------------ rts::vector / quad
template <class T>
struct v3 {
    T v[3];
    v3() { v[0] = v[1] = v[2] = 0; }
    v3(T a, T b, T c) { v[0] = a; v[1] = b; v[2] = c; }
    T len() const {
        T s = 0;
        for (int i = 0; i < 3; ++i) s += v[i] * v[i];
        return std::sqrt(s);
    }
    T dot(const v3& o) const {
        T s = 0;
        for (int i = 0; i < 3; ++i) s += v[i] * o.v[i];
        return s;
    }
    v3 norm() const {
        T l = len();
        return v3(v[0] / l, v[1] / l, v[2] / l);
    }
};
template <class T> v3<T> operator+(v3<T> a, v3<T> b) { return v3<T>(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
template <class T> v3<T> operator-(v3<T> a, v3<T> b) { return v3<T>(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
template <class T> v3<T> operator*(v3<T> a, T s) { return v3<T>(a.v[0] * s, a.v[1] * s, a.v[2] * s); }
template <class T> v3<T> mk3(const double* p) { return v3<T>((T)p[0], (T)p[1], (T)p[2]); }

template <class T>
struct quad {
    v3<T> A, X, Y;
    explicit quad(const bimsim_quad& q) : A(mk3<T>(q.A)), X(mk3<T>(q.X)), Y(mk3<T>(q.Y)) {}
    v3<T> at(T a, T b) const { return (A + X * a) + Y * b; }   // quad.h:129-140
};

// Nearest point of triangle (A,B,C) to p; rts/math/triangle.h:62-184
// (Eberly, "Distance between point and triangle in 3D").
template <class T>
T tri_dist(v3<T> A, v3<T> B, v3<T> C, v3<T> p) {
    v3<T> E0 = B - A, E1 = C - A, D = A - p;
    T a = E0.dot(E0), b = E0.dot(E1), c = E1.dot(E1), d = E0.dot(D), e = E1.dot(D);
    T det = a * c - b * b, s = b * e - c * d, t = b * d - a * e;
    if (s + t <= det) {
        if (s < 0) {
            if (t < 0) {
                s = 0;
                t = 0;
            } else {
                s = 0;
                t = (e >= 0 ? 0 : (-e >= c ? 1 : -e / c));
            }
        } else if (t < 0) {
            s = (d >= 0 ? 0 : (-d >= a ? 1 : -d / a));
            t = 0;
        } else {
            T inv = (T)1.0 / det;
            s *= inv;
            t *= inv;
        }
    } else {
        if (s < 0) {
            s = 0;
            t = 1;
        } else if (t < 0) {
            s = 1;
            t = 0;
        } else {
            T numer = c + e - b - d;
            if (numer <= 0) {
                s = 0;
            } else {
                T denom = a - 2 * b + c;
                s = (numer >= denom ? 1 : numer / denom);
            }
            t = 1 - s;
        }
    }
    v3<T> n = (A + E0 * s) + E1 * t;   // triangle::_p: A + s*E0 + t*E1
    return (p - n).len();
}

template <class T>
void quad_dist(const quad<T>& q, v3<T> p, T* dmin, T* dmax) {
    // quad.h:179-193: two triangles (A, A+X, A+Y) and (A+X+Y, A+X, A+Y)
    T d0 = tri_dist(q.A, q.A + q.X, q.A + q.Y, p);
    T d1 = tri_dist((q.A + q.X) + q.Y, q.A + q.X, q.A + q.Y, p);
    *dmin = d0 < d1 ? d0 : d1;
    T da = (q.A - p).len(), db = ((q.A + q.X) - p).len(), dc = ((q.A + q.Y) - p).len();
    T dd = (((q.A + q.X) + q.Y) - p).len();
    *dmax = std::max(da, std::max(db, std::max(dc, dd)));
}

// ----------------------------------------------------------- recurrences
// rts/math/legendre.h:18-25
template <class T>
inline void shift_legendre(int n, T x, T& P0, T& P1) {
    T Pn = ((T)(2 * n - 1) * x * P1 - (T)(n - 1) * P0) / (T)n;
    P0 = P1;
    P1 = Pn;
}
// rts/math/spherical_bessel.h:85-141 (real argument).  Thresholds are
// compared in the working precision.
template <class T>
inline void init_j(T x, T* j) {
    j[0] = std::sin(x) / x;
    j[1] = j[0] / x - std::cos(x) / x;
}
template <class T>
inline void init_y(T x, T* y) {
    y[0] = -std::cos(x) / x;
    y[1] = y[0] / x - std::sin(x) / x;
}
template <class T>
inline void shift_j(int n, T x, T* b) {
    T bn = ((T)(2 * n - 1) * b[1]) / x - b[0];
    if ((T)n > x)
        if ((double)bn < kJMin || (double)bn > kJMax) bn = 0;
    b[0] = b[1];
    b[1] = bn;
}
template <class T>
inline void shift_y(int n, T x, T* b) {
    T bn = ((T)(2 * n - 1) * b[1]) / x - b[0];
    if ((double)bn < kYMin || ((T)n > x && bn > 0)) {
        bn = 0;
        b[1] = 0;
    }
    b[0] = b[1];
    b[1] = bn;
}
// complex argument (calc_Ui): spherical_bessel.h:85-92,102-119 with T = complex
template <class T>
inline void init_jc(rc<T> x, rc<T>* j) {
    j[0] = csin_(x) / x;
    j[1] = j[0] / x - ccos_(x) / x;
}
template <class T>
inline void shift_jc(int n, rc<T> x, rc<T>* b) {
    rc<T> bn = scl(b[1], (T)(2 * n - 1)) / x - b[0];
    if ((T)n > x.r)
        if ((double)bn.r < kJMin || (double)bn.r > kJMax) bn = rc<T>(0, 0);
    b[0] = b[1];
    b[1] = bn;
}

// ------------------------------------------------------ per-pixel series
// calc_Us, reference nfScalarUs.cu:50-95
template <class T>
rc<T> calc_Us(T kd, T ct, int Nl, const rc<T>* B) {
    T j[2], y[2], P0 = 1, P1 = ct;
    init_j(kd, j);
    init_y(kd, y);
    rc<T> Us;
    for (int l = 0; l <= Nl; ++l) {
        if (l == 0) {
            Us = Us + scl(B[0] * rc<T>(j[0], y[0]), P0);
        } else {
            if (l > 1) {
                shift_legendre(l, ct, P0, P1);
                shift_j(l, kd, j);
                shift_y(l, kd, y);
            }
            Us = Us + scl(B[l] * rc<T>(j[1], y[1]), P1);
        }
    }
    return Us;
}

// calc_Ui, reference nfScalarUs.cu:201-245
template <class T>
rc<T> calc_Ui(rc<T> knd, T ct, int Nl, const rc<T>* A) {
    if (knd.r < (T)kEpsFloat) return A[0];
    rc<T> j[2];
    init_jc(knd, j);
    T P0 = 1, P1 = ct;
    rc<T> Ui;
    for (int l = 0; l <= Nl; ++l) {
        if (l == 0) {
            Ui = Ui + scl(A[0] * j[0], P0);
        } else {
            if (l > 1) {
                shift_legendre(l, ct, P0, P1);
                shift_jc(l, knd, j);
            }
            Ui = Ui + scl(A[l] * j[1], P1);
        }
    }
    return Ui;
}

// Corrected-Bessel forms (quirks off): exact j_l/y_l from the Zhang & Jin
// routines, same Legendre recurrence and operation order.
template <class T>
rc<T> calc_Us_exact(T kd, T ct, int Nl, const rc<T>* B) {
    std::vector<double> j(Nl + 2), y(Nl + 2), jp(Nl + 2), yp(Nl + 2);
    sph_jy_real(Nl, (double)kd, false, j.data(), y.data(), jp.data(), yp.data());
    T P0 = 1, P1 = ct;
    rc<T> Us;
    for (int l = 0; l <= Nl; ++l) {
        if (l > 1) shift_legendre(l, ct, P0, P1);
        T P = l == 0 ? P0 : P1;
        Us = Us + scl(B[l] * rc<T>((T)j[l], (T)y[l]), P);
    }
    return Us;
}
template <class T>
rc<T> calc_Ui_exact(rc<T> knd, T ct, int Nl, const rc<T>* A) {
    if (knd.r < (T)kEpsFloat) return A[0];
    std::vector<cd> j(Nl + 2), y(Nl + 2), jp(Nl + 2), yp(Nl + 2);
    sph_jy_cplx(Nl, cd((double)knd.r, (double)knd.i), false, j.data(), y.data(), jp.data(),
                yp.data());
    T P0 = 1, P1 = ct;
    rc<T> Ui;
    for (int l = 0; l <= Nl; ++l) {
        if (l > 1) shift_legendre(l, ct, P0, P1);
        T P = l == 0 ? P0 : P1;
        Ui = Ui + scl(A[l] * rc<T>((T)j[l].real(), (T)j[l].imag()), P);
    }
    return Ui;
}

template <class T>
std::vector<rc<T>> load_c(const double* p, int n) {
    std::vector<rc<T>> v(n);
    for (int i = 0; i < n; ++i) v[i] = rc<T>((T)p[2 * i], (T)p[2 * i + 1]);
    return v;
}

// precompute + optPc_gpuScalarUsp (nfScalarUs.cu:17-48,449-491); with nk = 1
// this is also gpuScalarUsp (nfScalarUs.cu:335-391).
template <class T>
void nf_us_sphere(const quad<T>& q, int uR, int vR, v3<T> f, T kmag, const bimsim_sphere& s,
                  const double* Ad, const double* Bd, const double* waves, int nk, T subA,
                  bool quirks, rc<T>* U, int nthreads) {
    const int Nl = s.Nl;
    std::vector<rc<T>> A = load_c<T>(Ad + 2 * s.coeff_off, Nl + 1);
    std::vector<rc<T>> B = load_c<T>(Bd + 2 * s.coeff_off, Nl + 1);
    std::vector<v3<T>> K(nk);
    for (int w = 0; w < nk; ++w) K[w] = mk3<T>(waves + 3 * w);
    const v3<T> ps = mk3<T>(s.p);
    const v3<T> c = ps - f;                       // nfScalarUs.cu:590
    const rc<T> n((T)s.n[0], (T)s.n[1]);
    const T a = (T)s.a;
    const T scale = (T)(1.0f / nk);               // nfScalarUs.cu:461
    std::vector<rc<T>> phase(nk);
    for (int w = 0; w < nk; ++w) phase[w] = cexp_(rc<T>(0, kmag * K[w].dot(c)));
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int iv = 0; iv < vR; ++iv) {
        for (int iu = 0; iu < uR; ++iu) {
            const int i = iv * uR + iu;
            T u = (T)iu / (T)uR, v = (T)iv / (T)vR;
            v3<T> r = q.at(u, v) - ps;
            T d = r.len();
            T kd = kmag * d;
            rc<T> knd = scl(n, kmag * d);
            r = r.norm();
            rc<T> sum;
            for (int w = 0; w < nk; ++w) {
                T ct = K[w].dot(r);
                rc<T> term;
                if (d <= a)
                    term = quirks ? calc_Ui(knd, ct, Nl, A.data()) : calc_Ui_exact(knd, ct, Nl, A.data());
                else
                    term = quirks ? calc_Us(kd, ct, Nl, B.data()) : calc_Us_exact(kd, ct, Nl, B.data());
                sum = sum + scl(phase[w], scale * subA) * term;
            }
            U[i] = U[i] + sum;
        }
    }
}

// gpuScalarUf (nfScalarUf.cu:46-175) and gpuScalarUfLut (nfScalarUfLut.cu:11-123)
template <class T>
rc<T> uf_pixel(v3<T> p, v3<T> f, v3<T> k, T kmag, T amp, T ca, T cb, int m, bool quirks,
               const T* lut, int dR, T dmin, T dmax) {
    v3<T> r = p - f;
    T d = r.len();
    if (lut ? d == 0 : d < (T)kEpsFloat)
        return rc<T>(amp * (T)2 * (T)kPIf * (ca - cb), 0);
    r = r.norm();
    const rc<T> im(0, 1);
    rc<T> il(1, 0);
    T kd = kmag * d;
    T j[2] = {0, 0};
    std::vector<double> jx, yx, jpx, ypx;
    if (!lut) {
        if (quirks) {
            init_j(kd, j);
        } else {
            jx.resize(m + 2); yx.resize(m + 2); jpx.resize(m + 2); ypx.resize(m + 2);
            sph_jy_real(m, (double)kd, false, jx.data(), yx.data(), jpx.data(), ypx.data());
        }
    }
    T ct = k.dot(r);
    T P0 = 1, P1 = ct;
    T Pa[3] = {1, ca, 1}, Pb[3] = {1, cb, 1};
    T di = 0;
    if (lut) di = ((d - dmin) / (dmax - dmin)) * (T)(dR - 1);
    rc<T> sum;
    for (int l = 0; l <= m; ++l) {
        T jl, Pl;
        if (l == 0) {
            jl = j[0];
            Pl = P0;
        } else if (l == 1) {
            jl = j[1];
            Pl = P1;
            Pa[2] = Pa[0];
            shift_legendre(l + 1, ca, Pa[0], Pa[1]);
            Pb[2] = Pb[0];
            shift_legendre(l + 1, cb, Pb[0], Pb[1]);
        } else {
            if (!lut && quirks) shift_j(l, kd, j);
            shift_legendre(l, ct, P0, P1);
            jl = j[1];
            Pl = P1;
            Pa[2] = Pa[0];
            shift_legendre(l + 1, ca, Pa[0], Pa[1]);
            Pb[2] = Pb[0];
            shift_legendre(l + 1, cb, Pb[0], Pb[1]);
        }
        if (lut) {
            // tex2D(texJ, l + 0.5, di + 0.5): column l, linear in di, clamped
            T y = di;
            T fy = std::floor(y);
            T wy = y - fy;
            int i0 = (int)fy, i1 = i0 + 1;
            i0 = std::min(std::max(i0, 0), dR - 1);
            i1 = std::min(std::max(i1, 0), dR - 1);
            jl = ((T)1 - wy) * lut[i0 * (m + 1) + l] + wy * lut[i1 * (m + 1) + l];
        } else if (!quirks) {
            jl = (T)jx[l];
        }
        T D = ((Pa[1] - Pa[2]) - Pb[1]) + Pb[2];
        sum = sum + scl(scl(scl(il, jl), Pl), D);
        il = il * im;
    }
    return scl(scl(scl(sum, (T)2), (T)kPIf), amp);
}

template <class T>
void nf_uf_impl(const bimsim_quad& qd, int uR, int vR, const double* fd, const double* kd_,
                T kmag, T amp, T ca, T cb, int m, int mode, bool quirks, const double* lutd,
                int dR, T dmin, T dmax, double* Uf, int nthreads) {
    quad<T> q(qd);
    v3<T> f = mk3<T>(fd), k = mk3<T>(kd_);
    std::vector<T> lut;
    if (mode == BIMSIM_UF_LUT) {
        lut.resize((size_t)dR * (m + 1));
        for (size_t i = 0; i < lut.size(); ++i) lut[i] = (T)lutd[i];
    }
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int iv = 0; iv < vR; ++iv) {
        for (int iu = 0; iu < uR; ++iu) {
            T u = (T)iu / (T)uR, v = (T)iv / (T)vR;
            v3<T> p = q.at(u, v);
            rc<T> val;
            if (mode == BIMSIM_UF_PLANE) {
                // gpuScalarUfp, nfScalarUf.cu:9-43
                v3<T> r = p - f;
                T kr = kmag * k.dot(r);
                val = scl(cexp_(rc<T>(0, kr)), amp);
            } else {
                val = uf_pixel(p, f, k, kmag, amp, ca, cb, m, quirks,
                               mode == BIMSIM_UF_LUT ? lut.data() : nullptr, dR, dmin, dmax);
            }
            const int i = iv * uR + iu;
            Uf[2 * i] = val.r;
            Uf[2 * i + 1] = val.i;
        }
    }
}

template <class T>
void to_rc(const double* d, rc<T>* c, size_t n) {
    for (size_t i = 0; i < n; ++i) c[i] = rc<T>((T)d[2 * i], (T)d[2 * i + 1]);
}
template <class T>
void from_rc(const rc<T>* c, double* d, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        d[2 * i] = (double)c[i].r;
        d[2 * i + 1] = (double)c[i].i;
    }
}

// ------------------------------------------------------------------- LUTs
// sphere::calcUp dimensions, sphere.cpp:222-241
template <class T>
void lut_dims(const bimsim_quad& qd, const bimsim_sphere& s, int R, int* aR, int* dR, T* dmin,
              T* dmax) {
    quad<T> q(qd);
    v3<T> p = mk3<T>(s.p);
    quad_dist(q, p, dmin, dmax);
    T a = (T)s.a;
    T a_inter = 0;
    if (*dmin < a) a_inter = std::sqrt(a - *dmin);   // sphere.cpp:233 (sic)
    v3<T> diag = q.at(0, 0) - q.at(1, 1);
    *aR = (int)((T)1 + (T)(2u * (unsigned)R) * a_inter / diag.len());
    *dR = 2 * R;
}

// calcHankelLut, sphere.cpp:154-203: h[ir][l] = (j_l, y_l)(k r)
template <class T>
void lut_hankel(int nl, T k, T a, T dmin, T dmax, int dR, double* h) {
    std::vector<double> j(nl + 2), y(nl + 2), jp(nl + 2), yp(nl + 2);
    T lo = std::max(a, dmin);
    T dr = (dmax - lo) / (T)(dR - 1);
    for (int ir = 0; ir < dR; ++ir) {
        T r = (T)ir * dr + lo;
        double kr = (double)(k * r);
        sph_jy_real(nl, kr, true, j.data(), y.data(), jp.data(), yp.data());
        for (int l = 0; l <= nl; ++l) {
            h[2 * ((size_t)ir * (nl + 1) + l)] = (double)(T)j[l];
            h[2 * ((size_t)ir * (nl + 1) + l) + 1] = (double)(T)y[l];
        }
    }
}

// calcBesselLut, sphere.cpp:99-152: j[ir][l] = j_l(k n r), r = ir*a/(aR-1)
template <class T>
void lut_bessel(int nl, T k, T a, rc<T> n, int aR, double* jo) {
    std::vector<cd> j(nl + 2), y(nl + 2), jp(nl + 2), yp(nl + 2);
    T dr = a / (T)(aR - 1);
    for (int ir = 0; ir < aR; ++ir) {
        T r = (T)ir * dr;
        rc<T> knr = scl(scl(n, k), r);
        if (ir != 0)
            sph_jy_cplx(nl, cd((double)knr.r, (double)knr.i), true, j.data(), y.data(), jp.data(),
                        yp.data());
        for (int l = 0; l <= nl; ++l) {
            double re, imv;
            if (ir == 0) {
                re = l == 0 ? 1.0 : 0.0;
                imv = 0.0;
            } else {
                re = (double)(T)j[l].real();
                imv = (double)(T)j[l].imag();
            }
            jo[2 * ((size_t)ir * (nl + 1) + l)] = re;
            jo[2 * ((size_t)ir * (nl + 1) + l) + 1] = imv;
        }
    }
}

// gpuScalarUsp / gpuScalarUip of sphere.cu:4-99: L[itheta][ir]
template <class T>
void lut_eval(int nl, const double* coef, const double* radial, int nr, int thetaR,
              double* out) {
    std::vector<rc<T>> C = load_c<T>(coef, nl + 1);
    // (PI) / (thetaR - 1) is float arithmetic in the reference for any ptype
    const T dtheta = (T)(kPIf / (float)(thetaR - 1));
    for (int it = 0; it < thetaR; ++it) {
        T theta = dtheta * (T)it;
        T ct = std::cos(theta);
        for (int ir = 0; ir < nr; ++ir) {
            T P0 = 1, P1 = ct;
            rc<T> acc;
            for (int l = 0; l <= nl; ++l) {
                const double* hp = radial + 2 * ((size_t)ir * (nl + 1) + l);
                rc<T> h((T)hp[0], (T)hp[1]);
                if (l == 0) {
                    acc = acc + scl(C[l] * h, P0);
                } else {
                    if (l > 1) shift_legendre(l, ct, P0, P1);
                    acc = acc + scl(C[l] * h, P1);
                }
            }
            size_t o = (size_t)it * nr + ir;
            out[2 * o] = (double)acc.r;
            out[2 * o + 1] = (double)acc.i;
        }
    }
}

// Bilinear fetch of a texture with unnormalised coordinates (x+0.5, y+0.5)
// and clamp addressing: exact weights (the reference hardware quantises
// them to 1/256).
template <class T>
rc<T> bilinear_clamp(const std::vector<rc<T>>& L, int W, int H, T x, T y) {
    T fx = std::floor(x), fy = std::floor(y);
    T wx = x - fx, wy = y - fy;
    int x0 = (int)fx, y0 = (int)fy;
    int xs[2] = {std::min(std::max(x0, 0), W - 1), std::min(std::max(x0 + 1, 0), W - 1)};
    int ys[2] = {std::min(std::max(y0, 0), H - 1), std::min(std::max(y0 + 1, 0), H - 1)};
    rc<T> a = scl(L[(size_t)ys[0] * W + xs[0]], (T)1 - wx) + scl(L[(size_t)ys[0] * W + xs[1]], wx);
    rc<T> b = scl(L[(size_t)ys[1] * W + xs[0]], (T)1 - wx) + scl(L[(size_t)ys[1] * W + xs[1]], wx);
    return scl(a, (T)1 - wy) + scl(b, wy);
}

// gpuScalarUpLut, nfScalarUpLut.cu:11-97
template <class T>
void nf_us_lut(const bimsim_quad& qd, int uR, int vR, const double* fd, T kmag,
               const bimsim_sphere& s, const double* Uspd, int dR, const double* Uipd, int aR,
               int thetaR, T dmin, T dmax, const double* waves, int nk, T subA, double* Ud,
               int nthreads) {
    quad<T> q(qd);
    v3<T> f = mk3<T>(fd), ps = mk3<T>(s.p);
    const T a = (T)s.a;
    std::vector<rc<T>> Usp((size_t)dR * thetaR), Uip((size_t)std::max(aR, 1) * thetaR);
    to_rc(Uspd, Usp.data(), Usp.size());
    if (Uipd) to_rc(Uipd, Uip.data(), (size_t)aR * thetaR);
    std::vector<v3<T>> K(nk);
    for (int w = 0; w < nk; ++w) K[w] = mk3<T>(waves + 3 * w);
    const v3<T> c = ps - f;
    const T lo = std::max(a, dmin);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
    for (int iv = 0; iv < vR; ++iv) {
        for (int iu = 0; iu < uR; ++iu) {
            T u = (T)iu / (T)uR, v = (T)iv / (T)vR;
            v3<T> r = q.at(u, v) - ps;
            T d = r.len();
            // the reference declares these texture coordinates float
            const float di = (float)(((d - lo) / (dmax - lo)) * (T)(dR - 1));
            const float ai = (float)(((d - dmin) / (a - dmin)) * (T)(aR - 1));
            rc<T> sum;
            for (int w = 0; w < nk; ++w) {
                r = r.norm();
                T ct = K[w].dot(r);
                if (ct < -1) ct = -1;
                if (ct > 1) ct = 1;
                // at the sphere centre r-hat is 0/0: the reference's texture
                // coordinate is NaN there; every row of the Uip table's first
                // column is A_0, so any angle gives the defined value
                if (ct != ct) ct = 1;
                const float ti = (float)((std::acos(ct) / (T)kPIf) * (T)(thetaR - 1));
                rc<T> phase = cexp_(rc<T>(0, kmag * K[w].dot(c)));
                rc<T> val = d < a ? bilinear_clamp(Uip, aR, thetaR, (T)ai, (T)ti)
                                  : bilinear_clamp(Usp, dR, thetaR, (T)di, (T)ti);
                sum = sum + scl(phase, (T)(1.0f / nk) * subA) * val;
            }
            const int i = iv * uR + iu;
            Ud[2 * i] = (double)((T)Ud[2 * i] + sum.r);
            Ud[2 * i + 1] = (double)((T)Ud[2 * i + 1] + sum.i);
        }
    }
}

// ------------------------------------------------------------------- FFT
// Mixed-radix recursive DFT (decimation in time) in double.  sign = -1 is
// the forward transform e^{-2 pi i jk/N} of cufftExecC2C(CUFFT_FORWARD).
void fft1(cd* x, int n, int stride, int sign, cd* tmp) {
    if (n == 1) return;
    int p = 2;
    while (n % p) ++p;
    const int m = n / p;
    // gather the p decimated sub-sequences into tmp, transform each
    for (int r = 0; r < p; ++r)
        for (int j = 0; j < m; ++j) tmp[r * m + j] = x[(j * p + r) * stride];
    std::vector<cd> sub(m > 1 ? m : 1);
    for (int r = 0; r < p; ++r) fft1(tmp + r * m, m, 1, sign, sub.data());
    for (int k = 0; k < n; ++k) {
        cd acc(0, 0);
        const int km = k % m;
        for (int r = 0; r < p; ++r) {
            // twiddle e^{sign 2 pi i r k / n}, angle reduced exactly
            long long e = ((long long)r * k) % n;
            double ang = sign * 2.0 * M_PI * (double)e / (double)n;
            acc += tmp[r * m + km] * cd(std::cos(ang), std::sin(ang));
        }
        x[k * stride] = acc;
    }
}

void fft2_impl(cd* a, int uR, int vR, int sign) {
    std::vector<cd> tmp(std::max(uR, vR));
    for (int iv = 0; iv < vR; ++iv) fft1(a + (size_t)iv * uR, uR, 1, sign, tmp.data());
    for (int iu = 0; iu < uR; ++iu) fft1(a + iu, vR, uR, sign, tmp.data());
}

// bandpass, microscope.cu:495-526
template <class T>
void bandpass(rc<T>* U, int uR, int vR, T du, T dv, T NAin, T NAout, T lambda, bool quirks) {
    for (int iv = 0; iv < vR; ++iv)
        for (int iu = 0; iu < uR; ++iu) {
            T u, v;
            if (iu <= uR / 2) u = (T)iu * du;
            else u = -(T)(quirks ? uR - 1 - iu : uR - iu) * du;
            if (iv <= vR / 2) v = (T)iv * dv;
            else v = -(T)(quirks ? vR - 1 - iv : vR - iv) * dv;
            T fmag = std::sqrt(u * u + v * v);
            if (fmag < NAin / lambda || fmag > NAout / lambda) U[(size_t)iv * uR + iu] = rc<T>(0, 0);
        }
}

template <class T>
void detector_impl(bool quirks, double* Ud, double* Ufd, int uR, int vR, T du, T dv, T NAin,
                   T NAout, T lambda, int padding, int Nd, int ss, double* D, double* Di,
                   double* Abs) {
    const size_t N = (size_t)uR * vR;
    double* fields[2] = {Ud, Ufd};
    double* dets[2] = {D, Di};
    const int S = Nd * ss;   // Ud.R[0]
    for (int fi = 0; fi < 2; ++fi) {
        std::vector<cd> a(N);
        for (size_t i = 0; i < N; ++i) a[i] = cd(fields[fi][2 * i], fields[fi][2 * i + 1]);
        fft2_impl(a.data(), uR, vR, -1);
        std::vector<rc<T>> F(N);
        for (size_t i = 0; i < N; ++i) F[i] = rc<T>((T)a[i].real(), (T)a[i].imag());
        bandpass(F.data(), uR, vR, du, dv, NAin, NAout, lambda, quirks);
        for (size_t i = 0; i < N; ++i) a[i] = cd((double)F[i].r, (double)F[i].i);
        fft2_impl(a.data(), uR, vR, +1);
        // ScaleField(1.0 / (R0*R1)), fieldslice.cpp:90
        const T sc = (T)(1.0 / ((unsigned)uR * (unsigned)vR));
        for (size_t i = 0; i < N; ++i) {
            F[i] = scl(rc<T>((T)a[i].real(), (T)a[i].imag()), sc);
            fields[fi][2 * i] = (double)F[i].r;
            fields[fi][2 * i + 1] = (double)F[i].i;
        }
        // crop (fieldslice.cu:319-362) + resample_intensity (fieldslice.cu:71-108)
        const int off = padding * S;
        for (int iv = 0; iv < Nd; ++iv)
            for (int iu = 0; iu < Nd; ++iu) {
                T I = 0;
                for (int ix = 0; ix < ss; ++ix)
                    for (int iy = 0; iy < ss; ++iy) {
                        int cu = iu * ss + ix, cv = iv * ss + iy;
                        rc<T> val = F[(size_t)(off + cv) * uR + (off + cu)];
                        T xm = std::sqrt(val.r * val.r + val.i * val.i);
                        I += xm * xm;
                    }
                size_t o = (size_t)iv * Nd + iu;
                dets[fi][o] = (double)((T)dets[fi][o] + (T)(1.0f * I) / (T)(ss * ss));
            }
    }
    if (Abs)
        for (size_t o = 0; o < (size_t)Nd * Nd; ++o)
            Abs[o] = (double)(-std::log10((T)D[o] / (T)Di[o]));
}

// rotation resample, nfInterpolate.cu:13-66; border addressing
template <class T>
void interpolate_impl(const double* in, double* out, int uR, int vR, T deg, bool cfloat) {
    T uc, vc;
    if (cfloat) {
        uc = (T)(uR / 2.0);
        vc = (T)(vR / 2.0);
    } else {
        uc = (T)(int)(uR / 2.0);
        vc = (T)(int)(vR / 2.0);
    }
    const T cs = std::cos(deg), sn = std::sin(deg);
    auto at = [&](int x, int y, int c) -> T {
        if (x < 0 || y < 0 || x >= uR || y >= vR) return 0;
        return (T)in[2 * ((size_t)y * uR + x) + c];
    };
    for (int iv = 0; iv < vR; ++iv)
        for (int iu = 0; iu < uR; ++iu) {
            float nx, ny;
            if (cfloat) {
                nx = (float)(uc + ((T)iu - uc) * cs - ((T)iv - vc) * sn);
                ny = (float)(vc + ((T)iu - uc) * sn + ((T)iv - vc) * cs);
            } else {
                int iuc = (int)uc, ivc = (int)vc;
                nx = (float)(uc + (T)(iu - iuc) * cs - (T)(iv - ivc) * sn);
                ny = (float)(vc + (T)(iu - iuc) * sn + (T)(iv - ivc) * cs);
            }
            T x = (T)nx, y = (T)ny;
            T fx = std::floor(x), fy = std::floor(y);
            T wx = x - fx, wy = y - fy;
            int x0 = (int)fx, y0 = (int)fy;
            for (int c = 0; c < 2; ++c) {
                T a = ((T)1 - wx) * at(x0, y0, c) + wx * at(x0 + 1, y0, c);
                T b = ((T)1 - wx) * at(x0, y0 + 1, c) + wx * at(x0 + 1, y0 + 1, c);
                out[2 * ((size_t)iv * uR + iu) + c] = (double)(((T)1 - wy) * a + wy * b);
            }
        }
}

// ---------------------------------------------------------- back-propagation
// backpropagate, microscope.cu:423-489: float Kx, dkx, kx, ky, k, kz (the
// kernel declares them float); the phase factor and the product in T.
template <class T>
void backprop_impl(double* U, int uR, int vR, T du, T dv, T lambda, float d, bool quirks) {
    const float twopi = 2 * kPIf;
    const float Kx = (float)((T)(twopi * (float)uR) * du);
    const float Ky = (float)((T)(twopi * (float)vR) * dv);
    const float dkx = Kx / (float)uR, dky = Ky / (float)vR;
    const float k = (float)((T)twopi / lambda);
    for (int iv = 0; iv < vR; ++iv)
        for (int iu = 0; iu < uR; ++iu) {
            const int nu = iu <= uR / 2 ? iu : -(quirks ? uR - 1 - iu : uR - iu);
            const int nv = iv <= vR / 2 ? iv : -(quirks ? vR - 1 - iv : vR - iv);
            const float kx = (float)((T)nu * (T)dkx), ky = (float)((T)nv * (T)dky);
            double* z = U + 2 * ((size_t)iv * uR + iu);
            if (kx * kx + ky * ky < k * k) {
                const float kz = std::sqrt(k * k - kx * kx - ky * ky);
                const T ph = (T)(kz * d);
                const T sr = -std::cos(ph), si = -std::sin(ph);
                const T zr = (T)z[0], zi = (T)z[1];
                z[0] = (double)(zr * sr - zi * si);
                z[1] = (double)(zr * si + zi * sr);
            } else {
                z[0] = z[1] = 0.0;
            }
        }
}

// ---------------------------------------- extended-source detector images
// Real images sampled like the reference's float textures at the same
// coordinates: bilinear (exact weights), zero outside.
template <class T>
T fetch_border_real(const double* S, int W, int H, float xf, float yf) {
    const T x = (T)xf, y = (T)yf;
    const T fx = std::floor(x), fy = std::floor(y);
    const T wx = x - fx, wy = y - fy;
    const int x0 = (int)fx, y0 = (int)fy;
    auto at = [&](int xx, int yy) -> T {
        if (xx < 0 || yy < 0 || xx >= W || yy >= H) return 0;
        return (T)S[(size_t)yy * W + xx];
    };
    const T a = ((T)1 - wx) * at(x0, y0) + wx * at(x0 + 1, y0);
    const T b = ((T)1 - wx) * at(x0, y0 + 1) + wx * at(x0 + 1, y0 + 1);
    return ((T)1 - wy) * a + wy * b;
}

// InterpolateAndIntegrateUd / Ufd, microscope.cu:89-132
template <class T>
void det_rotate_add_impl(const double* src, double* D, int uR, int vR, T deg, T scale) {
    const int uc = (int)(uR / 2.0), vc = (int)(vR / 2.0);
    const T cs = std::cos(deg), sn = std::sin(deg);
    for (int iv = 0; iv < vR; ++iv)
        for (int iu = 0; iu < uR; ++iu) {
            const float nx = (float)((T)uc + (T)(iu - uc) * cs - (T)(iv - vc) * sn);
            const float ny = (float)((T)vc + (T)(iu - uc) * sn + (T)(iv - vc) * cs);
            const size_t i = (size_t)iv * uR + iu;
            D[i] = (double)((T)D[i] + scale * fetch_border_real<T>(src, uR, vR, nx, ny));
        }
}

// simRing_Ud / simRing_Ufd, microscope.cu:138-240: one ring per radius, the
// samples summed by the workgroup's tree (stride n/2 halving to 1)
template <class T>
void det_ring_sums_impl(const double* src, double* line, int uR, int vR, int ns, int nrings) {
    const int uc = (int)(uR / 2.0), vc = (int)(vR / 2.0);
    std::vector<T> s(ns);
    for (int r = 0; r < nrings; ++r) {
        for (int t = 0; t < ns; ++t) {
            const T th = (T)((2.0 * (double)kPIf / ns) * (double)(t + 1));
            const float nx = (float)((T)uc + (T)r * std::cos(th));
            const float ny = (float)((T)vc - (T)r * std::sin(th));
            s[t] = fetch_border_real<T>(src, uR, vR, nx, ny);
        }
        for (int step = ns / 2; step >= 1; step >>= 1)
            for (int t = 0; t < step; ++t) s[t] = s[t] + s[t + step];
        line[r] = (double)((T)line[r] + s[0]);
    }
}

// InterpolateAndIntegrateUdFromLine / UfdFromLine, microscope.cu:37-86
template <class T>
void det_from_line_impl(const double* line, int len, double* D, int uR, int vR, T scale) {
    const int uc = (int)(uR / 2.0), vc = (int)(vR / 2.0);
    for (int iv = 0; iv < vR; ++iv)
        for (int iu = 0; iu < uR; ++iu) {
            const float du = (float)(iu - uc), dv = (float)(iv - vc);
            const float r = std::sqrt(du * du + dv * dv);
            const T x = (T)r;
            const T fx = std::floor(x);
            const T w = x - fx;
            const int x0 = (int)fx;
            const T a = (x0 >= 0 && x0 < len) ? (T)line[x0] : (T)0;
            const T b = (x0 + 1 >= 0 && x0 + 1 < len) ? (T)line[x0 + 1] : (T)0;
            D[(size_t)iv * uR + iu] = (double)(scale * (((T)1 - w) * a + w * b));
        }
}

// ----------------------------------------------------------------- random
// glibc random(3) TYPE_3 additive generator, as rand() after srand(seed).
struct GlibcRand {
    std::vector<int32_t> r;
    size_t k;
    explicit GlibcRand(unsigned seed) {
        r.resize(344);
        r[0] = (int32_t)(seed == 0 ? 1 : seed);
        for (int i = 1; i < 31; ++i) {
            int64_t w = (16807LL * r[i - 1]) % 2147483647LL;
            if (w < 0) w += 2147483647LL;
            r[i] = (int32_t)w;
        }
        for (int i = 31; i < 34; ++i) r[i] = r[i - 31];
        for (int i = 34; i < 344; ++i) r[i] = (int32_t)((uint32_t)r[i - 31] + (uint32_t)r[i - 3]);
        k = 344;
    }
    int next() {
        int32_t v = (int32_t)((uint32_t)r[k - 31] + (uint32_t)r[k - 3]);
        r.push_back(v);
        ++k;
        return (int)((uint32_t)v >> 1);
    }
};

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// rotation (0,0,1) -> k of montecarlo.cpp:23-40 (quaternion toMatrix3)
template <class T>
void rotation_to(const T k[3], T M[3][3]) {
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) M[a][b] = a == b ? 1 : 0;
    // k.dot((0,0,1)) accumulated from zero like rts::vector::dot
    T ca = (((T)0 + k[0] * (T)0) + k[1] * (T)0) + k[2] * (T)1;
    if (ca == (T)1.0) return;
    // axis = z x k, normalised
    T ax[3] = {(T)0 * k[2] - (T)1 * k[1], (T)1 * k[0] - (T)0 * k[2], (T)0 * k[1] - (T)0 * k[0]};
    T l = std::sqrt(((T)0 + ax[0] * ax[0]) + ax[1] * ax[1] + ax[2] * ax[2]);
    for (int i = 0; i < 3; ++i) ax[i] = ax[i] / l;
    T angle = std::acos(ca);
    T w = (T)std::cos(angle / 2.0);
    T x = ax[0] * (T)std::sin(angle / 2.0);
    T y = ax[1] * (T)std::sin(angle / 2.0);
    T z = ax[2] * (T)std::sin(angle / 2.0);
    T x2 = x + x, y2 = y + y, z2 = z + z;
    T xx = x * x2, xy = x * y2, xz = x * z2, yy = y * y2, yz = y * z2, zz = z * z2;
    T wx = w * x2, wy = w * y2, wz = w * z2;
    M[0][0] = (T)1.0 - (yy + zz);
    M[0][1] = xy - wz;
    M[0][2] = xz + wy;
    M[1][0] = xy + wz;
    M[1][1] = (T)1.0 - (xx + zz);
    M[1][2] = yz - wx;
    M[2][0] = xz - wy;
    M[2][1] = yz + wx;
    M[2][2] = (T)1.0 - (xx + yy);
}

}  // namespace
}  // namespace orc

using namespace orc;

// ======================================================================= C API
extern "C" {

double orc_gamma(double x) { return zj_gamma(x); }
int orc_msta1(double x, int mp) { return zj_msta1(x, mp); }
int orc_msta2(double x, int n, int mp) { return zj_msta2(x, n, mp); }

int orc_sph_jy(int nl, double x, int quirks, double* j, double* y, double* jp, double* yp) {
    if (nl < 0) return -1;
    sph_jy_real(nl, x, quirks != 0, j, y, jp, yp);
    return 0;
}

int orc_sph_jy_c(int nl, double zr, double zi, int quirks, double* j, double* y, double* jp,
                 double* yp) {
    if (nl < 0) return -1;
    std::vector<cd> J(nl + 2), Y(nl + 2), JP(nl + 2), YP(nl + 2);
    sph_jy_cplx(nl, cd(zr, zi), quirks != 0, J.data(), Y.data(), JP.data(), YP.data());
    for (int k = 0; k <= nl + 1; ++k) {
        j[2 * k] = J[k].real(); j[2 * k + 1] = J[k].imag();
        y[2 * k] = Y[k].real(); y[2 * k + 1] = Y[k].imag();
        jp[2 * k] = JP[k].real(); jp[2 * k + 1] = JP[k].imag();
        yp[2 * k] = YP[k].real(); yp[2 * k + 1] = YP[k].imag();
    }
    return 0;
}

int orc_sphere_order(double a, double lambda) {
    // sphere.h:71-74, ptype = float
    float x = (2 * kPIf * (float)a) / (float)lambda;
    return (int)std::ceil(x + 4 * std::pow(x, (float)(1.0 / 3.0)) + 2);
}

// sphere::calcCoeff, sphere.cpp:21-97 (kmag passed in: 2*PI/lambda)
int orc_mie_coeffs(int nl, double kmag, double a, double nr, double ni, int quirks, double* A,
                   double* B) {
    if (nl < 0) return -1;
    const cd nc(nr, ni);
    const cd kna = nc * kmag * a;
    const cd ka(kmag * a, 0.0);
    std::vector<cd> j1(nl + 2), y1(nl + 2), jp1(nl + 2), yp1(nl + 2);
    std::vector<cd> j2(nl + 2), y2(nl + 2), jp2(nl + 2), yp2(nl + 2);
    sph_jy_cplx(nl, ka, quirks != 0, j1.data(), y1.data(), jp1.data(), yp1.data());
    sph_jy_cplx(nl, kna, quirks != 0, j2.data(), y2.data(), jp2.data(), yp2.data());
    const cd ci(0, 1);
    for (int l = 0; l <= nl; ++l) {
        cd h = j1[l] + ci * y1[l];
        cd hp = jp1[l] + ci * yp1[l];
        cd num_a = j1[l] * hp - jp1[l] * h;
        cd den = j2[l] * hp - h * jp2[l] * nc;
        cd An = (2.0 * l + 1.0) * std::pow(ci, l) * (num_a / den);
        cd num_b = j1[l] * jp2[l] * nc - j2[l] * jp1[l];
        cd Bn = (2.0 * l + 1.0) * std::pow(ci, l) * (num_b / den);
        A[2 * l] = An.real(); A[2 * l + 1] = An.imag();
        B[2 * l] = Bn.real(); B[2 * l + 1] = Bn.imag();
    }
    return 0;
}

int orc_sbessel_up(int nl, double x, int fp32, double* j, double* y) {
    auto run = [&](auto xt) {
        using T = decltype(xt);
        T jj[2], yy[2];
        init_j(xt, jj);
        init_y(xt, yy);
        j[0] = jj[0]; y[0] = yy[0];
        if (nl >= 1) { j[1] = jj[1]; y[1] = yy[1]; }
        for (int l = 2; l <= nl; ++l) {
            shift_j(l, xt, jj);
            shift_y(l, xt, yy);
            j[l] = jj[1];
            y[l] = yy[1];
        }
    };
    if (fp32) run((float)x); else run(x);
    return 0;
}

int orc_sbessel_up_c(int nl, double xr, double xi, int fp32, double* j) {
    auto run = [&](auto xt) {
        using T = decltype(xt);
        rc<T> x((T)xr, (T)xi), jj[2];
        init_jc(x, jj);
        j[0] = jj[0].r; j[1] = jj[0].i;
        if (nl >= 1) { j[2] = jj[1].r; j[3] = jj[1].i; }
        for (int l = 2; l <= nl; ++l) {
            shift_jc(l, x, jj);
            j[2 * l] = jj[1].r;
            j[2 * l + 1] = jj[1].i;
        }
    };
    if (fp32) run(0.0f); else run(0.0);
    return 0;
}

int orc_legendre_up(int nl, double x, int fp32, double* P) {
    auto run = [&](auto xt) {
        using T = decltype(xt);
        T P0 = 1, P1 = xt;
        P[0] = 1;
        if (nl >= 1) P[1] = xt;
        for (int l = 2; l <= nl; ++l) {
            shift_legendre(l, xt, P0, P1);
            P[l] = P1;
        }
    };
    if (fp32) run((float)x); else run(x);
    return 0;
}

int orc_quad_dist(const bimsim_quad* q, const double p[3], int fp32, double* dmin, double* dmax) {
    if (fp32) {
        float a, b;
        quad_dist(quad<float>(*q), mk3<float>(p), &a, &b);
        *dmin = a; *dmax = b;
    } else {
        quad_dist(quad<double>(*q), mk3<double>(p), dmin, dmax);
    }
    return 0;
}

// nearfieldStruct::calcBesselLut, nearfield.cpp:222-283
int orc_uf_bessel_lut(int m, double kmag, double dmin, double dmax, int dR, int fp32,
                      double* lut) {
    auto run = [&](auto zero) {
        using T = decltype(zero);
        std::vector<double> j(m + 2), y(m + 2), jp(m + 2), yp(m + 2);
        T k = (T)kmag, lo = (T)dmin;
        T dr = ((T)dmax - lo) / (T)(dR - 1);
        for (int id = 0; id < dR; ++id) {
            T d = (T)id * dr + lo;
            double kd = (double)(k * d);
            sph_jy_real(m, kd, true, j.data(), y.data(), jp.data(), yp.data());
            for (int l = 0; l <= m; ++l) {
                T jv = (T)j[l];
                if (std::isnan(jv) || std::isinf(jv)) jv = (kd == 0 && l == 0) ? 1 : 0;
                lut[(size_t)id * (m + 1) + l] = (double)jv;
            }
        }
    };
    if (fp32) run(0.0f); else run(0.0);
    return 0;
}

int orc_nf_uf(const bimsim_quad* q, int uR, int vR, const double f[3], const double k[3],
              double kmag, double amp, double ca, double cb, int m, int mode, int fp32,
              int quirks, const double* lut, int dR, double dmin, double dmax, double* Uf,
              int nthreads) {
    if (uR <= 0 || vR <= 0 || m < 0) return -1;
    if (mode == BIMSIM_UF_LUT && (!lut || dR < 2)) return -1;
    if (fp32)
        nf_uf_impl<float>(*q, uR, vR, f, k, (float)kmag, (float)amp, (float)ca, (float)cb, m,
                          mode, quirks != 0, lut, dR, (float)dmin, (float)dmax, Uf, nthreads);
    else
        nf_uf_impl<double>(*q, uR, vR, f, k, kmag, amp, ca, cb, m, mode, quirks != 0, lut, dR,
                           dmin, dmax, Uf, nthreads);
    return 0;
}

// Monte-Carlo incident field: per pixel the sum over the waves of the plane
// wave of gpuScalarUfp (nfScalarUf.cu:9-43: p = ABCD(u, v), r = p - f,
// exp(i kmag k.r) A) with A = (1.0f/nk) subA, the weight the scattered
// field gives each wave (nfScalarUs.cu:461 scale, :572 subA), in wave order.
int orc_nf_uf_mc(const bimsim_quad* q, int uR, int vR, const double f[3], double kmag,
                 const double* waves, int nk, double subA, int fp32, double* Uf, int nthreads) {
    if (uR <= 0 || vR <= 0 || nk <= 0) return -1;
    auto run = [&](auto zero) {
        using T = decltype(zero);
        quad<T> qq(*q);
        const v3<T> fv = mk3<T>(f);
        const T w = (T)(1.0f / nk) * (T)subA;
        std::vector<v3<T>> K(nk);
        for (int i = 0; i < nk; ++i) K[i] = mk3<T>(waves + 3 * i);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 1)
        for (int iv = 0; iv < vR; ++iv) {
            for (int iu = 0; iu < uR; ++iu) {
                T u = (T)iu / (T)uR, v = (T)iv / (T)vR;
                const v3<T> r = qq.at(u, v) - fv;
                rc<T> sum;
                for (int i = 0; i < nk; ++i) sum = sum + scl(cexp_(rc<T>(0, (T)kmag * K[i].dot(r))), w);
                const int px = iv * uR + iu;
                Uf[2 * px] = (double)sum.r;
                Uf[2 * px + 1] = (double)sum.i;
            }
        }
    };
    if (fp32) run(0.0f); else run(0.0);
    return 0;
}

int orc_nf_us(const bimsim_quad* q, int uR, int vR, const double f[3], double kmag,
              const bimsim_sphere* s, int ns, const double* A, const double* B,
              const double* waves, int nk, double subA, int fp32, int quirks, double* U,
              int accumulate, int nthreads) {
    if (uR <= 0 || vR <= 0 || nk <= 0 || ns < 0) return -1;
    const size_t N = (size_t)uR * vR;
    auto run = [&](auto zero) {
        using T = decltype(zero);
        std::vector<rc<T>> Uc(N);
        if (accumulate) to_rc(U, Uc.data(), N);
        quad<T> qq(*q);
        for (int is = 0; is < ns; ++is)
            nf_us_sphere(qq, uR, vR, mk3<T>(f), (T)kmag, s[is], A, B, waves, nk, (T)subA,
                         quirks != 0, Uc.data(), nthreads);
        from_rc(Uc.data(), U, N);
    };
    if (fp32) run(0.0f); else run(0.0);
    return 0;
}

// nfSumUf.cu:6-117
int orc_nf_sum_uf(const bimsim_quad* q, int uR, int vR, const bimsim_sphere* s, int ns, int fp32,
                  const double* Uf, double* U) {
    const size_t N = (size_t)uR * vR;
    if (ns == 0) {
        std::memcpy(U, Uf, sizeof(double) * 2 * N);
        return 0;
    }
    auto run = [&](auto zero) {
        using T = decltype(zero);
        quad<T> qq(*q);
        for (int iv = 0; iv < vR; ++iv)
            for (int iu = 0; iu < uR; ++iu) {
                T u = (T)iu / (T)uR, v = (T)iv / (T)vR;
                v3<T> p = qq.at(u, v);
                bool inside = false;
                for (int is = 0; is < ns && !inside; ++is)
                    if ((p - mk3<T>(s[is].p)).len() < (T)s[is].a) inside = true;
                if (inside) continue;
                size_t i = (size_t)iv * uR + iu;
                U[2 * i] = (double)((T)U[2 * i] + (T)Uf[2 * i]);
                U[2 * i + 1] = (double)((T)U[2 * i + 1] + (T)Uf[2 * i + 1]);
            }
    };
    if (fp32) run(0.0f); else run(0.0);
    return 0;
}

int orc_sphere_lut_dims(const bimsim_quad* q, const bimsim_sphere* s, int R, int fp32, int* aR,
                        int* dR, double* dmin, double* dmax) {
    if (fp32) {
        float a, b;
        lut_dims<float>(*q, *s, R, aR, dR, &a, &b);
        *dmin = a; *dmax = b;
    } else {
        lut_dims<double>(*q, *s, R, aR, dR, dmin, dmax);
    }
    return 0;
}

int orc_lut_hankel(int nl, double kmag, double a, double dmin, double dmax, int dR, int fp32,
                   double* h) {
    if (dR < 2) return -1;
    if (fp32) lut_hankel<float>(nl, (float)kmag, (float)a, (float)dmin, (float)dmax, dR, h);
    else lut_hankel<double>(nl, kmag, a, dmin, dmax, dR, h);
    return 0;
}

int orc_lut_bessel(int nl, double kmag, double a, double nr, double ni, int aR, int fp32,
                   double* j) {
    if (aR < 1) return -1;
    if (fp32) lut_bessel<float>(nl, (float)kmag, (float)a, rc<float>((float)nr, (float)ni), aR, j);
    else lut_bessel<double>(nl, kmag, a, rc<double>(nr, ni), aR, j);
    return 0;
}

int orc_lut_eval(int nl, const double* coef, const double* radial, int nr, int thetaR, int fp32,
                 double* lut) {
    if (thetaR < 2 || nr < 1) return -1;
    if (fp32) lut_eval<float>(nl, coef, radial, nr, thetaR, lut);
    else lut_eval<double>(nl, coef, radial, nr, thetaR, lut);
    return 0;
}

int orc_nf_us_lut(const bimsim_quad* q, int uR, int vR, const double f[3], double kmag,
                  const bimsim_sphere* s, const double* Usp, int dR, const double* Uip, int aR,
                  int thetaR, double dmin, double dmax, const double* waves, int nk,
                  double subA, int fp32, double* U, int nthreads) {
    if (fp32)
        nf_us_lut<float>(*q, uR, vR, f, (float)kmag, *s, Usp, dR, Uip, aR, thetaR, (float)dmin,
                         (float)dmax, waves, nk, (float)subA, U, nthreads);
    else
        nf_us_lut<double>(*q, uR, vR, f, kmag, *s, Usp, dR, Uip, aR, thetaR, dmin, dmax, waves,
                          nk, subA, U, nthreads);
    return 0;
}

int orc_nf_interpolate(const double* Uin, double* Uout, int uR, int vR, double deg,
                       int centre_float, int fp32) {
    if (fp32) interpolate_impl<float>(Uin, Uout, uR, vR, (float)deg, centre_float != 0);
    else interpolate_impl<double>(Uin, Uout, uR, vR, deg, centre_float != 0);
    return 0;
}

int orc_backpropagate(double* U, int uR, int vR, double du, double dv, double lambda, double d,
                      int quirks, int fp32) {
    if (fp32)
        backprop_impl<float>(U, uR, vR, (float)du, (float)dv, (float)lambda, (float)d, quirks != 0);
    else
        backprop_impl<double>(U, uR, vR, du, dv, lambda, (float)d, quirks != 0);
    return 0;
}

int orc_det_rotate_add(const double* src, double* D, int uR, int vR, double deg, double scale,
                       int fp32) {
    if (fp32) det_rotate_add_impl<float>(src, D, uR, vR, (float)deg, (float)scale);
    else det_rotate_add_impl<double>(src, D, uR, vR, deg, scale);
    return 0;
}

int orc_det_ring_sums(const double* src, double* line, int uR, int vR, int n_samples,
                      int n_rings, int fp32) {
    if (n_samples < 1) return -1;
    if (fp32) det_ring_sums_impl<float>(src, line, uR, vR, n_samples, n_rings);
    else det_ring_sums_impl<double>(src, line, uR, vR, n_samples, n_rings);
    return 0;
}

int orc_det_from_line(const double* line, int len, double* D, int uR, int vR, double scale,
                      int fp32) {
    if (fp32) det_from_line_impl<float>(line, len, D, uR, vR, (float)scale);
    else det_from_line_impl<double>(line, len, D, uR, vR, scale);
    return 0;
}

int orc_fft2(double* data, int uR, int vR, int inverse) {
    const size_t N = (size_t)uR * vR;
    std::vector<cd> a(N);
    for (size_t i = 0; i < N; ++i) a[i] = cd(data[2 * i], data[2 * i + 1]);
    fft2_impl(a.data(), uR, vR, inverse ? +1 : -1);
    for (size_t i = 0; i < N; ++i) {
        data[2 * i] = a[i].real();
        data[2 * i + 1] = a[i].imag();
    }
    return 0;
}

int orc_detector(int fp32, int quirks, double* U, double* Uf, int uR, int vR, double du,
                 double dv, double NAin, double NAout, double lambda, int padding, int Nd,
                 int ss, double* D, double* Di, double* Abs) {
    if ((2 * padding + 1) * Nd * ss > std::min(uR, vR)) return -1;
    if (fp32)
        detector_impl<float>(quirks != 0, U, Uf, uR, vR, (float)du, (float)dv, (float)NAin,
                             (float)NAout, (float)lambda, padding, Nd, ss, D, Di, Abs);
    else
        detector_impl<double>(quirks != 0, U, Uf, uR, vR, du, dv, NAin, NAout, lambda, padding,
                              Nd, ss, D, Di, Abs);
    return 0;
}

int orc_glibc_rand(unsigned int seed, int n, int* out) {
    GlibcRand g(seed);
    for (int i = 0; i < n; ++i) out[i] = g.next();
    return 0;
}

// mcSampleNA, montecarlo.cpp:9-71
int orc_mc_sample_na(double* out, int N, const double k[3], double NAin, double NAout, int rng,
                     unsigned long long seed) {
    if (N < 0) return -1;
    if (rng == BIMSIM_RNG_GLIBC) {
        float kk[3] = {(float)k[0], (float)k[1], (float)k[2]}, M[3][3];
        rotation_to(kk, M);
        float inPhi = std::asin((float)NAin), outPhi = std::asin((float)NAout);
        float inZ = std::cos(inPhi), outZ = std::cos(outPhi);
        float rangeZ = inZ - outZ;
        GlibcRand g((unsigned)seed);
        for (int i = 0; i < N; ++i) {
            float z = ((float)g.next() / (float)2147483647) * rangeZ + outZ;
            float theta = ((float)g.next() / (float)2147483647) * 2 * (float)kPIf;
            float phi = std::acos(z);
            float c[3] = {1 * std::cos(theta) * std::sin(phi), 1 * std::sin(theta) * std::sin(phi),
                          1 * std::cos(phi)};
            for (int r = 0; r < 3; ++r) {
                float acc = 0;
                for (int cc = 0; cc < 3; ++cc) acc += M[r][cc] * c[cc];
                out[3 * i + r] = acc;
            }
        }
    } else {
        double M[3][3];
        rotation_to(k, M);
        double inZ = std::cos(std::asin(NAin)), outZ = std::cos(std::asin(NAout));
        double rangeZ = inZ - outZ;
        SplitMix64 g(seed);
        for (int i = 0; i < N; ++i) {
            double z = g.uniform() * rangeZ + outZ;
            double theta = g.uniform() * 2 * (double)kPIf;
            double phi = std::acos(z);
            double c[3] = {std::cos(theta) * std::sin(phi), std::sin(theta) * std::sin(phi),
                           std::cos(phi)};
            for (int r = 0; r < 3; ++r) {
                double acc = 0;
                for (int cc = 0; cc < 3; ++cc) acc += M[r][cc] * c[cc];
                out[3 * i + r] = acc;
            }
        }
    }
    return 0;
}

}  // extern "C"
