// Host geometry of the slice (rts::quad::dist / dist_max, reference
// rts/math/quad.h:179-205 with the triangle nearest point of
// rts/math/triangle.h:62-184, after D. Eberly, "Distance Between Point and
// Triangle in 3D").  Double precision.
#pragma once

#include <algorithm>
#include <cmath>

#include "common.h"

namespace bimsim {

inline double tri_dist(d3 A, d3 B, d3 C, d3 p) {
    const d3 E0 = sub(B, A), E1 = sub(C, A), D = sub(A, p);
    const double a = dot(E0, E0), b = dot(E0, E1), c = dot(E1, E1), d = dot(E0, D), e = dot(E1, D);
    const double det = a * c - b * b;
    double s = b * e - c * d, t = b * d - a * e;
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
            const double inv = 1.0 / det;
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
            const double numer = c + e - b - d;
            if (numer <= 0) {
                s = 0;
            } else {
                const double denom = a - 2 * b + c;
                s = (numer >= denom ? 1 : numer / denom);
            }
            t = 1 - s;
        }
    }
    const d3 n = add(add(A, scale(E0, s)), scale(E1, t));
    return len(sub(p, n));
}

// Distance from p to the nearest and the farthest point of the slice.
inline void quad_dist(const QuadArg& q, d3 p, double* dmin, double* dmax) {
    const d3 AX = add(q.A, q.X), AY = add(q.A, q.Y), AXY = add(add(q.A, q.X), q.Y);
    const double d0 = tri_dist(q.A, AX, AY, p);
    const double d1 = tri_dist(AXY, AX, AY, p);
    *dmin = d0 < d1 ? d0 : d1;
    const double da = len(sub(q.A, p)), db = len(sub(AX, p)), dc = len(sub(AY, p));
    const double dd = len(sub(AXY, p));
    *dmax = std::max(da, std::max(db, std::max(dc, dd)));
}

}  // namespace bimsim
