/*
 * ORACLE -- test infrastructure only.
 *
 * CPU restatement of the reference BIMSim near-field and imaging path
 * (seberisha/bimsim-cuda), used as the checker of the HIP library.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Every near-field/detector routine follows the reference operation order.
 * fp32 = 0 evaluates with ptype = double (the reference built with
 * PRECISION_DOUBLE: the plain FP64 form); fp32 = 1 evaluates with
 * ptype = float, modelling the reference's own single-precision program.
 * Complex arrays are interleaved (re, im) doubles; fields use the reference
 * layout element i = iv*uR + iu.  The Mie coefficients and LUT radial Bessel
 * functions are always FP64 (the reference computes them in double).
 *
 * Parity pins (tests/): the Zhang & Jin routines against the reference's own
 * bessjy.cpp/cbessjy.cpp/gamma.cpp compiled into oracle/_ref and against
 * scipy; the sampler against the reference montecarlo.cpp in oracle/_ref and
 * glibc rand(); recurrences against the reference headers compiled into
 * oracle/_ref; the series against closed forms (n = 1 spheres, plane-wave
 * limits, annulus integrals); the FFT against numpy.fft.
 */
#ifndef BIMSIM_ORACLE_H
#define BIMSIM_ORACLE_H

#include "../include/bimsim_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Zhang & Jin special functions (reference gamma.cpp, bessjy.cpp, cbessjy.cpp) */
double orc_gamma(double x);
int orc_msta1(double x, int mp);
int orc_msta2(double x, int n, int mp);
int orc_sph_jy(int nl, double x, int quirks, double* j, double* y, double* jp, double* yp);
int orc_sph_jy_c(int nl, double zr, double zi, int quirks, double* j, double* y, double* jp,
                 double* yp);

/* sphere.h:71-74 and sphere.cpp:21-97 */
int orc_sphere_order(double a, double lambda);
int orc_mie_coeffs(int nl, double kmag, double a, double nr, double ni, int quirks, double* A,
                   double* B);

/* In-kernel recurrences as calc_Us/calc_Ui/gpuScalarUf run them
 * (rts/math/spherical_bessel.h:85-141, rts/math/legendre.h:9-25); l = 0..nl. */
int orc_sbessel_up(int nl, double x, int fp32, double* j, double* y);
int orc_sbessel_up_c(int nl, double xr, double xi, int fp32, double* j);
int orc_legendre_up(int nl, double x, int fp32, double* P);

/* rts::quad::dist / dist_max (rts/math/quad.h:179-205) */
int orc_quad_dist(const bimsim_quad* q, const double p[3], int fp32, double* dmin, double* dmax);

/* Near field (nfScalarUf.cu, nfScalarUfLut.cu, nfScalarUs.cu, nfSumUf.cu) */
int orc_uf_bessel_lut(int m, double kmag, double dmin, double dmax, int dR, int fp32,
                      double* lut);
int orc_nf_uf(const bimsim_quad* q, int uR, int vR, const double f[3], const double k[3],
              double kmag, double amp, double cos_alpha, double cos_beta, int m, int mode,
              int fp32, int quirks, const double* lut, int dR, double dmin, double dmax,
              double* Uf, int nthreads);
int orc_nf_uf_mc(const bimsim_quad* q, int uR, int vR, const double f[3], double kmag,
                 const double* waves, int nk, double subA, int fp32, double* Uf, int nthreads);
int orc_nf_us(const bimsim_quad* q, int uR, int vR, const double f[3], double kmag,
              const bimsim_sphere* s, int ns, const double* A, const double* B,
              const double* waves, int nk, double subA, int fp32, int quirks, double* U,
              int accumulate, int nthreads);
int orc_nf_sum_uf(const bimsim_quad* q, int uR, int vR, const bimsim_sphere* s, int ns,
                  int fp32, const double* Uf, double* U);

/* LUT path (sphere.cpp:99-272, sphere.cu:4-149, nfScalarUpLut.cu:11-97) */
int orc_sphere_lut_dims(const bimsim_quad* q, const bimsim_sphere* s, int R, int fp32,
                        int* aR, int* dR, double* dmin, double* dmax);
int orc_lut_hankel(int nl, double kmag, double a, double dmin, double dmax, int dR, int fp32,
                   double* h);
int orc_lut_bessel(int nl, double kmag, double a, double nr, double ni, int aR, int fp32,
                   double* j);
int orc_lut_eval(int nl, const double* coef, const double* radial, int nr, int thetaR,
                 int fp32, double* lut);
int orc_nf_us_lut(const bimsim_quad* q, int uR, int vR, const double f[3], double kmag,
                  const bimsim_sphere* s, const double* Usp, int dR, const double* Uip, int aR,
                  int thetaR, double dmin, double dmax, const double* waves, int nk,
                  double subA, int fp32, double* U, int nthreads);

/* Rotation resample (nfInterpolate.cu:13-66); centre_float selects the
 * gpuInterpolateUf centre (float) instead of gpuInterpolateU (int). */
int orc_nf_interpolate(const double* Uin, double* Uout, int uR, int vR, double deg,
                       int centre_float, int fp32);

/* Extended-source detector images (microscope.cu:37-240); real uR x vR
 * images, centre ((int)(uR/2.0), (int)(vR/2.0)), zero border:
 * D += scale * src rotated by deg (InterpolateAndIntegrateUd, :89-132);
 * line[r] += the reference's tree sum over a ring of n_samples points
 * (simRing_Ud, :196-240); D = scale * line(|p - centre|) (:37-60). */
/* backpropagate (microscope.cu:423-489) on an angular spectrum U (uR x vR
 * complex, in place); Kx, kx, k and kz in float as the reference. */
int orc_backpropagate(double* U, int uR, int vR, double du, double dv, double lambda, double d,
                      int quirks, int fp32);

int orc_det_rotate_add(const double* src, double* D, int uR, int vR, double deg, double scale,
                       int fp32);
int orc_det_ring_sums(const double* src, double* line, int uR, int vR, int n_samples,
                      int n_rings, int fp32);
int orc_det_from_line(const double* line, int len, double* D, int uR, int vR, double scale,
                      int fp32);

/* Detector (microscope.cu:495-526,608-662,722-791; fieldslice.cpp:25-94;
 * fieldslice.cu:71-153,319-362).  U and Uf are band-passed in place. */
int orc_fft2(double* data, int uR, int vR, int inverse);
int orc_detector(int fp32, int quirks, double* U, double* Uf, int uR, int vR, double du,
                 double dv, double NAin, double NAout, double lambda, int padding, int Nd,
                 int ss, double* D, double* Di, double* Abs);

/* Monte-Carlo directions (montecarlo.cpp:9-71) and the glibc rand() stream */
int orc_mc_sample_na(double* out, int N, const double k[3], double NAin, double NAout, int rng,
                     unsigned long long seed);
int orc_glibc_rand(unsigned int seed, int n, int* out);

#ifdef __cplusplus
}
#endif

#endif
