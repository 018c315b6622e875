// ORACLE -- test infrastructure only (see oracle.h).
#pragma once

#include <complex>

namespace orc {

using cd = std::complex<double>;

double zj_gamma(double x);
int zj_msta1(double x, int mp);
int zj_msta2(double x, int n, int mp);
int zj_jyv(double v, double x, double* jv, double* yv, double* djv, double* dyv);
int zj_cjyva(double v, cd z, cd* cjv, cd* cyv, cd* cjvp, cd* cyvp, bool typo);

// Spherical j_l, y_l (and derivatives) for l = 0..nl; arrays hold nl+2.
void sph_jy_real(int nl, double x, bool quirks, double* j, double* y, double* jp, double* yp);
void sph_jy_cplx(int nl, cd z, bool quirks, cd* j, cd* y, cd* jp, cd* yp);

}  // namespace orc
