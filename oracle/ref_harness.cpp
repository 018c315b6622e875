// ORACLE CHECKER -- test infrastructure only.
//
// C entry points around the reference's own host sources, compiled where they
// lie under the reference tree by `make -C oracle ref` into oracle/_ref/
// (never committed, never shipped as product):
//   bessjy.cpp, cbessjy.cpp, gamma.cpp   FP64 Zhang & Jin Bessel functions
//   montecarlo.cpp                       mcSampleNA
//   rts/math/spherical_bessel.h, legendre.h (CUDA_CALLABLE templates,
//   instantiated here for the host)
// The tests use these to pin the oracle's restatement.  Note on the template
// instantiation with T = double: the unqualified real(x) in
// spherical_bessel.h resolves to ::real(float) in this translation unit (no
// `using namespace std` before the header), so its threshold compares run
// after rounding to float.  The oracle compares in the working precision;
// the two agree unless a value lies within one float ulp of a threshold.
#include <complex>
#include <cstdlib>
#include <vector>

#include "dataTypes.h"
#include "montecarlo.h"
#include "rts/math/legendre.h"
#include "rts/math/spherical_bessel.h"

int bessjyv_sph(int v, double z, double& vm, double* cjv, double* cyv, double* cjvp,
                double* cyvp);
int cbessjyva_sph(int v, std::complex<double> z, double& vm, std::complex<double>* cjv,
                  std::complex<double>* cyv, std::complex<double>* cjvp,
                  std::complex<double>* cyvp);
int msta1(double x, int mp);
int msta2(double x, int n, int mp);
// gamma.cpp defines gamma() with C linkage (libm declares the same name);
// the library is linked with -Bsymbolic so every call inside it binds to the
// reference's definition, not to libm's legacy gamma (= lgamma).
#include <math.h>

extern "C" {

double ref_gamma(double x) { return ::gamma(x); }
int ref_msta1(double x, int mp) { return msta1(x, mp); }
int ref_msta2(double x, int n, int mp) { return msta2(x, n, mp); }

// Arrays hold nl+2 entries, zero-filled first (the reference leaves orders
// above its Miller start unwritten).
int ref_sph_jy(int nl, double x, double* j, double* y, double* jp, double* yp) {
    for (int k = 0; k < nl + 2; ++k) j[k] = y[k] = jp[k] = yp[k] = 0.0;
    double vm = 0;
    return bessjyv_sph(nl, x, vm, j, y, jp, yp);
}

int ref_sph_jy_c(int nl, double zr, double zi, double* j, double* y, double* jp, double* yp) {
    std::vector<std::complex<double> > J(nl + 2), Y(nl + 2), JP(nl + 2), YP(nl + 2);
    double vm = 0;
    int rc = cbessjyva_sph(nl, std::complex<double>(zr, zi), vm, &J[0], &Y[0], &JP[0], &YP[0]);
    for (int k = 0; k < nl + 2; ++k) {
        j[2 * k] = J[k].real(); j[2 * k + 1] = J[k].imag();
        y[2 * k] = Y[k].real(); y[2 * k + 1] = Y[k].imag();
        jp[2 * k] = JP[k].real(); jp[2 * k + 1] = JP[k].imag();
        yp[2 * k] = YP[k].real(); yp[2 * k + 1] = YP[k].imag();
    }
    return rc;
}

// mcSampleNA after srand(seed); out: N*3 floats widened to double.
int ref_mc_sample_na(double* out, int N, double kx, double ky, double kz, double NAin,
                     double NAout, unsigned int seed) {
    std::vector<bsVector> s(N > 0 ? N : 1);
    srand(seed);
    mcSampleNA(&s[0], N, bsVector((ptype)kx, (ptype)ky, (ptype)kz), (ptype)NAin, (ptype)NAout);
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < 3; ++c) out[3 * i + c] = s[i][c];
    return 0;
}

// In-kernel recurrences as calc_Us/gpuScalarUf drive them: j, y for l=0..nl.
int ref_sbessel_up(int nl, double x, int fp32, double* j, double* y) {
    if (fp32) {
        float xf = (float)x, jj[2], yy[2];
        rts::init_sbesselj<float>(xf, jj);
        rts::init_sbessely<float>(xf, yy);
        j[0] = jj[0]; y[0] = yy[0]; j[1] = jj[1]; y[1] = yy[1];
        for (int l = 2; l <= nl; ++l) {
            rts::shift_sbesselj<float>(l, xf, jj);
            rts::shift_sbessely<float>(l, xf, yy);
            j[l] = jj[1]; y[l] = yy[1];
        }
    } else {
        double jj[2], yy[2];
        rts::init_sbesselj<double>(x, jj);
        rts::init_sbessely<double>(x, yy);
        j[0] = jj[0]; y[0] = yy[0]; j[1] = jj[1]; y[1] = yy[1];
        for (int l = 2; l <= nl; ++l) {
            rts::shift_sbesselj<double>(l, x, jj);
            rts::shift_sbessely<double>(l, x, yy);
            j[l] = jj[1]; y[l] = yy[1];
        }
    }
    return 0;
}

// Complex j_l recurrence of calc_Ui (T = rts::complex<float>).
int ref_sbessel_up_c32(int nl, double xr, double xi, double* j) {
    rts::complex<float> x((float)xr, (float)xi), jj[2];
    rts::init_sbesselj<rts::complex<float> >(x, jj);
    j[0] = jj[0].r; j[1] = jj[0].i; j[2] = jj[1].r; j[3] = jj[1].i;
    for (int l = 2; l <= nl; ++l) {
        rts::shift_sbesselj<rts::complex<float> >(l, x, jj);
        j[2 * l] = jj[1].r; j[2 * l + 1] = jj[1].i;
    }
    return 0;
}

int ref_legendre_up(int nl, double x, int fp32, double* P) {
    if (fp32) {
        float P0, P1, xf = (float)x;
        rts::init_legendre<float>(xf, P0, P1);
        P[0] = P0; P[1] = P1;
        for (int l = 2; l <= nl; ++l) { rts::shift_legendre<float>(l, xf, P0, P1); P[l] = P1; }
    } else {
        double P0, P1;
        rts::init_legendre<double>(x, P0, P1);
        P[0] = P0; P[1] = P1;
        for (int l = 2; l <= nl; ++l) { rts::shift_legendre<double>(l, x, P0, P1); P[l] = P1; }
    }
    return 0;
}

}  // extern "C"
