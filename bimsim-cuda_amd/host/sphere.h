// sphere (reference sphere.h:14-92, sphere.cpp): position, radius, material
// index, series order, Mie coefficients and the scattered/internal field
// tables used by the look-up path.
#pragma once

#include <sstream>
#include <string>
#include <vector>

#include "dataTypes.h"
#include "fieldslice.h"

struct sphere {
    bsPoint p;
    ptype a = 0;
    unsigned int iMaterial = 0;

    // Usp(theta, r) / Uip(theta, rho) tables (R[0] = radial samples,
    // R[1] = angular samples) and their radial range
    fieldslice Usp;
    fieldslice Uip;
    ptype d_min = 0;
    ptype d_max = 0;

    int Nl = 0;
    bsComplex n;
    std::vector<bsComplex> B;   // external coefficients B_l, l = 0..Nl
    std::vector<bsComplex> A;   // internal coefficients A_l

    sphere(ptype x = 0.0, ptype y = 0.0, ptype z = 0.0, ptype a = 0.0, int m = 0, int ang = 128);

    // Copies drop the tables, as in the reference (sphere.cpp:274-296).
    sphere(const sphere& rhs);
    sphere& operator=(const sphere& rhs);

    std::string toStr() const {
        std::stringstream ss;
        ss << p << ", " << a << ", " << iMaterial;
        return ss.str();
    }

    // Nl = ceil(x + 4 x^(1/3) + 2), x = 2 PI a / lambda (sphere.h:71-74)
    void calcNl(ptype lambda);

    // A_l, B_l in FP64 on the GPU (sphere.cpp:21-97)
    void calcCoeff(ptype lambda, bsComplex n);

    // Radial tables on the host: j[ia*(Nl+1) + l] = j_l(k n rho) over
    // rho in [0, a] (aR samples) and h[ir*(Nl+1) + l] = h_l(k r) over
    // r in [max(a, d_min), d_max] (rR samples), scaled like the reference
    // (sphere.cpp:99-220).
    void calcLut(bsComplex* j, bsComplex* h, ptype lambda, bsComplex n, int aR, int rR);
    void calcBesselLut(bsComplex* j, ptype k, bsComplex n, int aR);
    void calcHankelLut(bsComplex* h, ptype k, int rR);

    // Table ranges and sizes for a slice and resolution R, then Usp and Uip
    // (sphere.cpp:222-272; thetaR = 1000, defaults.h:83).
    void calcUp(ptype lambda, bsComplex n, bsRect nfPlane, unsigned int R);

    // The sphere as the C ABI sees it (coefficients at coeff_off).
    bimsim_sphere abi(int coeff_off) const;
};
