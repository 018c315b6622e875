// Host types of the BIMSim API (reference dataTypes.h:6-41, rts/math/
// vector.h, point.h, quad.h).  The reference fixes its arithmetic type at
// compile time (PRECISION_SINGLE / PRECISION_DOUBLE); here host parameters
// are double and the device arithmetic is a run-time setting
// (bimsim::settings().precision), so one build serves both forms.
#pragma once

#include <cmath>
#include <complex>
#include <ostream>
#include <sstream>
#include <stdexcept>
#include <string>

#include "../../include/bimsim_hip.h"

typedef double ptype;
typedef std::complex<ptype> bsComplex;

#ifndef PI
#define PI 3.14159f   // the reference's float constant (dataTypes.h:17)
#endif

namespace rts {

// 3-vector with the members the BIMSim host code uses.
template <class T, int N = 3>
struct vector {
    T v[3];
    vector() : v{0, 0, 0} {}
    vector(T x, T y, T z) : v{x, y, z} {}
    T& operator[](int i) { return v[i]; }
    T operator[](int i) const { return v[i]; }
    vector operator+(const vector& b) const { return vector(v[0] + b.v[0], v[1] + b.v[1], v[2] + b.v[2]); }
    vector operator-(const vector& b) const { return vector(v[0] - b.v[0], v[1] - b.v[1], v[2] - b.v[2]); }
    vector operator*(T s) const { return vector(v[0] * s, v[1] * s, v[2] * s); }
    vector operator/(T s) const { return vector(v[0] / s, v[1] / s, v[2] / s); }
    bool operator==(const vector& b) const { return v[0] == b.v[0] && v[1] == b.v[1] && v[2] == b.v[2]; }
    T dot(const vector& b) const {
        T s = 0;
        for (int i = 0; i < 3; ++i) s += v[i] * b.v[i];
        return s;
    }
    T len() const { return std::sqrt(dot(*this)); }
    vector norm() const { return *this / len(); }
    vector cross(const vector& b) const {
        return vector(v[1] * b.v[2] - v[2] * b.v[1], v[2] * b.v[0] - v[0] * b.v[2],
                      v[0] * b.v[1] - v[1] * b.v[0]);
    }
    // (r, theta, phi) -> (x, y, z), theta the azimuth (vector.h:73-84)
    vector sph2cart() const {
        return vector(v[0] * std::cos(v[1]) * std::sin(v[2]), v[0] * std::sin(v[1]) * std::sin(v[2]),
                      v[0] * std::cos(v[2]));
    }
    vector cart2sph() const {
        vector s;
        s[0] = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        s[1] = std::atan2(v[1], v[0]);
        s[2] = std::acos(v[2] / s[0]);
        return s;
    }
    std::string toStr() const {
        std::stringstream ss;
        ss << "[" << v[0] << ", " << v[1] << ", " << v[2] << "]";
        return ss.str();
    }
};

// a position; differences of points are vectors
template <class T, int N = 3>
struct point : vector<T, N> {
    point() {}
    point(T x, T y, T z) : vector<T, N>(x, y, z) {}
    point(const vector<T, N>& v) : vector<T, N>(v) {}
};

// Slice geometry: corner A and edges X, Y (quad.h:16-215).
template <class T, int N = 3>
struct quad {
    point<T, N> A;
    vector<T, N> X, Y;
    quad() {}
    quad(point<T, N> a, point<T, N> b, point<T, N> c) : A(a), X(b - a), Y(c - a) {}
    // from two corners and a normal (quad.h:55-84)
    quad(point<T, N> pMin, point<T, N> pMax, vector<T, N> normal) {
        A = pMin;
        const vector<T, N> v0 = pMax - pMin;
        const vector<T, N> v1 = v0.cross(normal);
        const point<T, N> B = A + v0 * 0.5 + v1 * 0.5;
        const point<T, N> C = A + v0 * 0.5 - v1 * 0.5;
        X = B - A;
        Y = C - A;
    }
    point<T, N> p(T a, T b) const { return A + X * a + Y * b; }
    point<T, N> operator()(T a, T b) const { return p(a, b); }
    // scale about the centre (quad.h:156-172)
    quad operator*(T s) const {
        const point<T, N> c = A + X * 0.5 + Y * 0.5;
        quad r;
        r.X = X * s;
        r.Y = Y * s;
        r.A = c - r.X * 0.5 - r.Y * 0.5;
        return r;
    }
    bimsim_quad abi() const {
        bimsim_quad q;
        for (int i = 0; i < 3; ++i) {
            q.A[i] = A[i];
            q.X[i] = X[i];
            q.Y[i] = Y[i];
        }
        return q;
    }
    std::string toStr() const {
        std::stringstream ss;
        ss << "A = " << A.toStr() << std::endl;
        ss << "B = " << (A + X).toStr() << std::endl;
        ss << "C = " << (A + X + Y).toStr() << std::endl;
        ss << "D = " << (A + Y).toStr() << std::endl;
        return ss.str();
    }
};

}  // namespace rts

typedef rts::vector<ptype, 3> bsVector;
typedef rts::point<ptype, 3> bsPoint;
typedef rts::quad<ptype, 3> bsRect;

template <class T, int N>
std::ostream& operator<<(std::ostream& os, const rts::vector<T, N>& v) {
    return os << v.toStr();
}
template <class T, int N>
std::ostream& operator<<(std::ostream& os, const rts::quad<T, N>& q) {
    return os << q.toStr();
}

namespace bimsim {

// Process-wide settings that replace the reference's compile-time switches.
struct Settings {
    int device = 0;
    int precision = BIMSIM_FP32_FAST;   // device arithmetic of the fields
    int quirks = BIMSIM_QUIRKS_REF;     // reproduce the reference's numerics
    int mc_rng = BIMSIM_RNG_GLIBC;      // Monte-Carlo stream (glibc rand after srand(0))
    unsigned long long mc_seed = 0;
    bool verbose = false;
};
Settings& settings();

// The library context of the current settings (created on first use on
// settings().device, on that device's default stream).
bimsim_ctx* context();
void release_context();

// Throws std::runtime_error with the library's message when rc != OK.
void check(int rc, const char* what);

// Bytes of one field sample / scalar sample at a precision.
inline size_t field_bytes(int prec) { return prec == BIMSIM_FP32_FAST ? 8 : 16; }
inline size_t scalar_bytes(int prec) { return prec == BIMSIM_FP32_FAST ? 4 : 8; }

}  // namespace bimsim
