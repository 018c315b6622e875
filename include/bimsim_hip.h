/*
 * bimsim_hip.h -- C ABI of the MI355X (gfx950) BIMSim near-field and imaging path.
 *
 * Plain C: POD structs, pointers and sizes only.  Every entry point returns a
 * bimsim_status (0 = success, negative = error) and never exits the process;
 * bimsim_last_error() gives the message of the last failure on a context.
 *
 * Buffer conventions (match the reference field containers):
 *   - fields are complex, interleaved (re, im), row-major with u fastest:
 *     element i = iv*uR + iu   (fieldslice::x_hat, reference fieldslice.h:16,
 *     indexing nfScalarUs.cu:346).  FP64 precisions use double pairs,
 *     BIMSIM_FP32_FAST uses float pairs.
 *   - detector images are real, Nd*Nd, same indexing (scalarslice::S,
 *     reference scalarslice.h:10); double for FP64 precisions, float for FP32.
 *   - every device pointer argument is caller-owned device memory; the library
 *     never frees caller memory.  Scratch memory belongs to the context.
 *   - calls are asynchronous on the context stream; bimsim_sync() waits.
 *     One context per host thread; calls on one context are serialised.
 */
#ifndef BIMSIM_HIP_H
#define BIMSIM_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define BIMSIM_ABI_VERSION 2

typedef struct bimsim_ctx bimsim_ctx;

typedef enum {
    BIMSIM_OK = 0,
    BIMSIM_E_INVALID = -1,      /* bad argument (shape, pointer, range)        */
    BIMSIM_E_HIP = -2,          /* HIP runtime error                            */
    BIMSIM_E_NOMEM = -3,        /* device or host allocation failed             */
    BIMSIM_E_UNSUPPORTED = -4,  /* mode/size not supported by this build        */
    BIMSIM_E_NODEVICE = -5      /* no usable gfx950 device                      */
} bimsim_status;

/* Arithmetic form of a call.
 *   FP64_PARITY: the reference algorithm in double, reference operation order
 *                and loop structure (per-wave recurrences), FMA contraction off.
 *   FP64:        same arithmetic, wave-independent Bessel/Hankel terms hoisted
 *                out of the wave loop (identical operation sequence per term).
 *   FP32_FAST:   Bessel terms and Mie products in double, per-wave Legendre
 *                series, plane-wave sum and FFT in float with compensated
 *                accumulation. */
typedef enum {
    BIMSIM_FP64_PARITY = 0,
    BIMSIM_FP64 = 1,
    BIMSIM_FP32_FAST = 2
} bimsim_precision;

/* Reference behaviours that are numerical defects, reproduced when set
 * (default).  With 0 the library uses the corrected forms:
 *   - y_{v0} asymptotic typo for |z|>12 (reference cbessjy.cpp:541)
 *   - upward j_l/y_l recurrences with the convergence thresholds
 *     (rts/math/spherical_bessel.h:8-10,102-141); off = Miller downward j_l
 *   - sqrt(3.14159/2z) scale of the spherical wrappers (bessjy.cpp:18,719-723)
 */
typedef enum {
    BIMSIM_QUIRKS_OFF = 0,
    BIMSIM_QUIRKS_REF = 1
} bimsim_quirks;

/* Incident-field evaluation (nearfieldStruct::scalarUf / scalarUfLut /
 * gpuScalarUfp, reference nfScalarUf.cu:9-217, nfScalarUfLut.cu:11-188).
 * BIMSIM_UF_MC is the Monte-Carlo sum of the plane waves that also drive the
 * scattered field (bimsim_nf_uf_mc); the reference sums only Us that way. */
typedef enum {
    BIMSIM_UF_SERIES = 0,
    BIMSIM_UF_PLANE = 1,
    BIMSIM_UF_LUT = 2,
    BIMSIM_UF_MC = 3
} bimsim_uf_mode;

/* Scattered-field evaluation (scalarUs / scalarUpLut, reference
 * nfScalarUs.cu:494-658, nfScalarUpLut.cu:342-495). */
typedef enum {
    BIMSIM_US_SERIES = 0,
    BIMSIM_US_LUT = 1
} bimsim_us_mode;

/* Monte-Carlo direction generators for bimsim_mc_sample_na. */
typedef enum {
    BIMSIM_RNG_GLIBC = 0,       /* glibc rand() stream after srand(seed)        */
    BIMSIM_RNG_SPLITMIX64 = 1   /* SplitMix64(seed) uniform doubles             */
} bimsim_rng;

/* Scalar views of a field slice (fieldslice::Intensity / Mag / Real / Imag). */
typedef enum {
    BIMSIM_SLICE_INTENSITY = 0,
    BIMSIM_SLICE_MAGNITUDE = 1,
    BIMSIM_SLICE_REAL = 2,
    BIMSIM_SLICE_IMAG = 3
} bimsim_slice_op;

/* Detector ratios (microscopeStruct::getAbsorbance / getTransmittance). */
typedef enum {
    BIMSIM_RATIO_ABSORBANCE = 0,
    BIMSIM_RATIO_TRANSMITTANCE = 1
} bimsim_ratio_op;

/* rts::quad<ptype,3> (reference rts/math/quad.h:30-33): corner A, edges X, Y.
 * Pixel (iu, iv) of a uR x vR slice sits at A + X*(iu/uR) + Y*(iv/vR). */
typedef struct {
    double A[3];
    double X[3];
    double Y[3];
} bimsim_quad;

/* One sphere at one wavelength (reference sphere.h:14-92). */
typedef struct {
    double p[3];       /* centre                                         */
    double a;          /* radius                                         */
    double n[2];       /* complex refractive index (re, im)              */
    int    Nl;         /* series order, sphere::calcNl (sphere.h:71-74)  */
    int    coeff_off;  /* element offset of A_l/B_l in packed arrays     */
} bimsim_sphere;

/* ---------------------------------------------------------------- context */

/* hip_stream: a hipStream_t to run on, or NULL for the device's default
 * stream (where the reference launches everything).  All work of the context
 * is ordered on that stream.  Replaces the implicit device/stream of
 * main.cpp:46 (cudaSetDevice). */
int bimsim_ctx_create(int device, void* hip_stream, bimsim_ctx** out);
int bimsim_ctx_destroy(bimsim_ctx* ctx);
int bimsim_sync(bimsim_ctx* ctx);
const char* bimsim_last_error(const bimsim_ctx* ctx);
int bimsim_abi_version(void);

/* Stage timers (nearfieldStruct::t_Uf / t_Us, reference nearfield.h:58-59,
 * rts/cuda/timer.h:4-21).  With timing enabled every call records hipEvents
 * around its stages on the context stream without waiting.
 * bimsim_get_timings waits for them and returns, per stage, the summed
 * milliseconds and span counts since the previous call, then resets:
 * ms[0] incident field, ms[1] scattered field (whole bimsim_nf_us),
 * ms[2] detector, ms[3] scattered-field series kernel launches only. */
int bimsim_set_timing(bimsim_ctx* ctx, int enable);

/* Launch-shape tuning of the FP32 scattered-field series (on by default):
 * when a call's work queue is short, the first call of each shape times the
 * candidate wave splits and grid sizes on its own inputs and keeps the
 * fastest for the context.  No counterpart in the reference. */
int bimsim_set_autotune(bimsim_ctx* ctx, int enable);

/* Algorithm of the FP32 scattered-field series (BIMSIM_FP32_FAST):
 *   BIMSIM_SERIES_DIRECT    per pixel, wave and order, as the reference
 *                           (rescaled Legendre recurrence per wave);
 *   BIMSIM_SERIES_HARMONIC  the addition theorem P_l(k.r) = 4 pi/(2l+1)
 *                           sum_m Y_lm(k) Y*_lm(r) moves the wave sum onto
 *                           each sphere: (Nl+1)(Nl+2)/2 terms per pixel and
 *                           sphere instead of nk (Nl+1);
 *   BIMSIM_SERIES_AUTO      harmonic when every instance has >= 16 waves.
 * Both sum the same series; they differ by rounding.  No counterpart in the
 * reference. */
typedef enum {
    BIMSIM_SERIES_AUTO = 0,
    BIMSIM_SERIES_DIRECT = 1,
    BIMSIM_SERIES_HARMONIC = 2
} bimsim_series_algo;
int bimsim_set_series_algorithm(bimsim_ctx* ctx, int algo);
int bimsim_get_timings(bimsim_ctx* ctx, float* ms, int* counts);

/* ------------------------------------------------------------ host helpers */

/* sphere::calcNl (reference sphere.h:71-74): ceil(x + 4 x^(1/3) + 2),
 * x = 2*PI*a/lambda evaluated in float with PI = 3.14159f. */
int bimsim_sphere_order(double a, double lambda);

/* mcSampleNA (reference montecarlo.cpp:9-71): N unit directions uniform in
 * cos(phi) over the condenser annulus [NAin, NAout], rotated from +z to k.
 * With BIMSIM_RNG_GLIBC the draws replay glibc rand() after srand(seed) in
 * float arithmetic; nearfieldStruct::calcWaves uses srand(NULL) (seed 0,
 * nearfield.cpp:102).  out: N*3 doubles (x, y, z) on the host. */
int bimsim_mc_sample_na(double* out, int N, const double k[3], double NAin,
                        double NAout, int rng, unsigned long long seed);

/* ------------------------------------------------------ device entry points */

/* Mie coefficients on the GPU in FP64 (sphere::calcCoeff, reference
 * sphere.cpp:21-97 with cbessjyva_sph, cbessjy.cpp:456-737).
 * spheres: host array; A_out/B_out: device complex128 arrays, sphere s writes
 * elements [coeff_off, coeff_off + Nl] of each.  kmag = 2*PI/lambda. */
int bimsim_mie_coeffs(bimsim_ctx* ctx, const bimsim_sphere* spheres, int ns,
                      double kmag, int quirks, void* A_out, void* B_out);

/* Incident field Uf over a uR x vR slice (nearfieldStruct::calcUf,
 * reference nearfield.cpp:168-192; kernels nfScalarUf.cu:9-175 and
 * nfScalarUfLut.cu:11-123).  mode = bimsim_uf_mode.  Uf is overwritten. */
int bimsim_nf_uf(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                 const double focus[3], const double kdir[3], double kmag,
                 double amp, double cos_alpha, double cos_beta, int m,
                 int mode, int prec, int quirks, void* Uf);

/* Incident field as the Monte-Carlo sum of nk plane waves: Uf(p) =
 * sum_w (1.0f/nk) subA exp(i kmag K_w.(p - f)), each term the plane wave of
 * gpuScalarUfp (reference nfScalarUf.cu:9-43) with the weights the
 * scattered field gives the same waves (nfScalarUs.cu:461,572).  waves:
 * device, nk*3 doubles (the list nearfieldStruct::calcWaves draws,
 * nearfield.cpp:96-108).  FP64 forms sum the waves per pixel in order;
 * BIMSIM_FP32_FAST factors each phase into its row and column parts
 * (exp(i k K.(A-f)) exp(i k K.X iu/uR) exp(i k K.Y iv/vR)) and sums the
 * products over LDS-staged tiles.  Uf is overwritten.  The reference has
 * no such mode (its Uf is the Legendre series or one plane wave); it is
 * the focused field that converges to the series as nk grows. */
int bimsim_nf_uf_mc(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                    const double focus[3], double kmag, const double* waves, int nk,
                    double subA, int prec, void* Uf);

/* Scattered/internal field of ns spheres summed over nk plane waves
 * (nearfieldStruct::scalarUs, reference nfScalarUs.cu:494-658: precompute +
 * optPc_gpuScalarUsp, or gpuScalarUsp for a single plane wave).
 * waves: device, nk*3 doubles.  subA: amplitude per the reference
 * (nfScalarUs.cu:572 for Monte-Carlo, A for a plane wave).
 * A, B: device complex128 packed coefficients (bimsim_mie_coeffs).
 * accumulate = 0 clears U first (U.clear_gpu(), nfScalarUs.cu:507).
 * BIMSIM_FP32_FAST serves orders up to Nl = 111 with the FP32 series; a
 * call with a larger Nl computes the field in the FP64 form and rounds it
 * to complex64 (same result, FP64 speed; batches do the same per
 * instance). */
int bimsim_nf_us(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                 const double focus[3], double kmag,
                 const bimsim_sphere* spheres, int ns, const void* A,
                 const void* B, const double* waves, int nk, double subA,
                 int prec, int quirks, void* U, int accumulate);

/* U += Uf wherever the pixel lies outside every sphere; U = Uf when ns == 0
 * (nearfieldStruct::sumUf, reference nfSumUf.cu:6-117). */
int bimsim_nf_sum_uf(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                     const bimsim_sphere* spheres, int ns, int prec,
                     const void* Uf, void* U);

/* Detector (microscopeStruct::SimulateImaging, reference microscope.cu:
 * 608-662 and 722-791; fieldslice.cpp:25-94; fieldslice.cu:71-153,319-362):
 * FFT of U and Uf, objective band-pass NAin/lambda <= |f| <= NAout/lambda
 * with du = 1/|X|, dv = 1/|Y|, inverse FFT / (uR*vR), crop of Nd*ss samples
 * at offset padding*Nd*ss, |.|^2 averaged over ss x ss into D and Di.
 * D and Di accumulate (+=) like resample_intensity; Abs (may be NULL) gets
 * -log10(D/Di).  write_filtered != 0 leaves the band-passed fields in U and
 * Uf like the reference (in-place FFT); otherwise U/Uf are left unchanged.
 * quirks = BIMSIM_QUIRKS_REF keeps the reference's negative-frequency index
 * -(uR-1-iu) of the band-pass (microscope.cu:510-518); 0 uses -(uR-iu). */
int bimsim_detector(bimsim_ctx* ctx, int prec, int quirks, void* U, void* Uf, int uR,
                    int vR, double du, double dv, double NAin, double NAout,
                    double lambda, int padding, int Nd, int ss, void* D,
                    void* Di, void* Abs, int write_filtered);

/* The detector on the near field zero-filled to (fill uR) x (fill vR) for
 * the FFT: U and Uf sit at the corner [0, uR) x [0, vR) of zero arrays, the
 * objective band-pass acts on the finer frequency grid (du/fill, dv/fill),
 * and the crop is the field's own uR x uR samples there (Nd = uR = vR,
 * ss = 1); D, Di accumulate and Abs = -log10(D/Di) as bimsim_detector.  It
 * keeps the band-pass from wrapping the field around its edges.  No
 * counterpart in the reference, whose --padding instead simulates a larger
 * near field (microscope.cu:641-655,709-720); with fill = 1 it is
 * bimsim_detector with padding 0.  U and Uf are left unchanged. */
int bimsim_detector_zero_filled(bimsim_ctx* ctx, int prec, int quirks, const void* U,
                                const void* Uf, int uR, int vR, int fill, double du, double dv,
                                double NAin, double NAout, double lambda, void* D, void* Di,
                                void* Abs);

/* bimsim_detector (write_filtered = 0: U and Uf are left unchanged) without
 * any device allocation.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory, such as code under HIP graph capture,
 * tight serving loops and harnesses that forbid hipMalloc.  Its work memory
 * (twiddles, pass-band columns, compact spectrum, intensity rows) is a
 * fixed 64 MiB device array that each GPU the library runs on holds until
 * the process ends; calls needing more return BIMSIM_E_INVALID.  FFT sizes
 * must have a fixed plan (64..4096 of 2^k and 3 2^k, rows fitting LDS),
 * else BIMSIM_E_UNSUPPORTED.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (other streams, other threads or contexts)
 * race and must be serialised by the caller. */
int bimsim_detector_noalloc(bimsim_ctx* ctx, int prec, int quirks, const void* U, const void* Uf,
                            int uR, int vR, double du, double dv, double NAin, double NAout,
                            double lambda, int padding, int Nd, int ss, void* D, void* Di,
                            void* Abs);

/* ------------------------------------------- slice building blocks
 * The steps of microscopeStruct::SimulateImaging one at a time, for host
 * code that calls them separately.  Fields are complex128 (FP64
 * precisions) or complex64 (BIMSIM_FP32_FAST), scalar slices double or
 * float, all row-major with u fastest. */

/* In-place 2D FFT of a field (fieldslice::toAngularSpectrum, inverse = 0;
 * fromAngularSpectrum incl. the 1/(uR vR) of ScaleField, inverse = 1;
 * reference fieldslice.cpp:25-94).  uR, vR products of 2, 3 and 5. */
int bimsim_field_fft(bimsim_ctx* ctx, int prec, void* F, int uR, int vR, int inverse);

/* Objective band-pass mask on an angular spectrum (bandpass kernel,
 * reference microscope.cu:495-526): zero where |f| < NAin/lambda or
 * |f| > NAout/lambda. */
int bimsim_field_bandpass(bimsim_ctx* ctx, int prec, int quirks, void* F, int uR, int vR,
                          double du, double dv, double NAin, double NAout, double lambda);

/* Propagation of an angular spectrum over a distance d (backpropagate
 * kernel of microscopeStruct::applyBackpropagation, reference
 * microscope.cu:423-489,592-605): samples with kx^2 + ky^2 < k^2 are
 * multiplied by -exp(i kz d), kz = sqrt(k^2 - kx^2 - ky^2), the others
 * set to 0; kx = 2 pi u du over the band-pass frequencies (quirks: the
 * negative ones one sample off), k = 2 pi/lambda, in float as the
 * reference. */
int bimsim_field_backpropagate(bimsim_ctx* ctx, int prec, int quirks, void* F, int uR, int vR,
                               double du, double dv, double lambda, double d);

/* S[i] = op(F[i]) (fieldslice::Intensity / Mag / Real / Imag, reference
 * fieldslice.cu:6-27,110-230); op = bimsim_slice_op. */
int bimsim_field_to_scalar(bimsim_ctx* ctx, int prec, int op, const void* F, long long n,
                           void* S);

/* The same for a vector field of three components (a fieldslice with
 * scalarField = false: x_hat, y_hat, z_hat): INTENSITY and MAGNITUDE use
 * |X|^2 + |Y|^2 + |Z|^2 (field_intensity with y and z, reference
 * fieldslice.cu:6-27,156-171), REAL and IMAG use X alone
 * (fieldslice.cu:110-138,173-203).  Y and Z are both given or both NULL
 * (then as bimsim_field_to_scalar). */
int bimsim_field3_to_scalar(bimsim_ctx* ctx, int prec, int op, const void* X, const void* Y,
                            const void* Z, long long n, void* S);

/* F *= v (fieldslice::ScaleField, reference fieldslice.cu:141-153). */
int bimsim_field_scale(bimsim_ctx* ctx, int prec, void* F, long long n, double v);

/* dst (su x sv) = src[v.., u..] (fieldslice::crop, reference
 * fieldslice.cu:319-362). */
int bimsim_field_crop(bimsim_ctx* ctx, int prec, const void* src, int uR, int vR, int u, int v,
                      int su, int sv, void* dst);

/* Detector pixels from a (Du ss) x (Dv ss) field: accumulate != 0 gives
 * D += scale |F|^2 averaged over ss x ss (resample_intensity); 0 gives
 * D = the average (resample_intensity_without_integrating); reference
 * fieldslice.cu:30-108. */
int bimsim_detector_resample(bimsim_ctx* ctx, int prec, const void* F, int Du, int Dv, int ss,
                             double scale, int accumulate, void* D);

/* The same over a vector field: each pixel sums |X|^2 + |Y|^2 + |Z|^2 per
 * sample (resample_intensity with y_hat and z_hat, reference
 * fieldslice.cu:30-108); Y and Z may be NULL. */
int bimsim_detector_resample3(bimsim_ctx* ctx, int prec, const void* X, const void* Y,
                              const void* Z, int Du, int Dv, int ss, double scale, int accumulate,
                              void* D);

/* out = -log10(D/Di) or D/Di (calc_absorbance / calc_transmittance,
 * reference microscope.cu:722-791); op = bimsim_ratio_op. */
int bimsim_slice_ratio(bimsim_ctx* ctx, int prec, int op, const void* D, const void* Di,
                       long long n, void* out);

/* -------------------------------------------------- look-up-table paths */

/* Radial j_l table of the incident field (nearfieldStruct::calcBesselLut,
 * reference nearfield.cpp:222-283): lut[id*(m+1) + l] = j_l(kmag d) for
 * d = dmin + id (dmax - dmin)/(dR - 1), NaN/inf replaced by 1 (l = 0 at
 * d = 0) or 0.  Computed in FP64 on the GPU; stored as double, or float for
 * BIMSIM_FP32_FAST. */
int bimsim_uf_bessel_lut(bimsim_ctx* ctx, int m, double kmag, double dmin, double dmax,
                         int dR, int prec, void* lut);

/* Incident field from that table (gpuScalarUfLut, reference
 * nfScalarUfLut.cu:11-123).  bimsim_nf_uf with BIMSIM_UF_LUT builds the
 * table itself (dR = 2 max(uR, vR), nfScalarUfLut.cu:130-134). */
int bimsim_nf_uf_lut(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                     const double focus[3], const double kdir[3], double amp,
                     double cos_alpha, double cos_beta, int m, const void* lut, int dR,
                     double dmin, double dmax, int prec, void* Uf);

/* Table sizes of a sphere (sphere::calcUp, reference sphere.cpp:222-241):
 * d_min/d_max from the slice, aR = 1 + 2R sqrt(a - d_min)/|diagonal|,
 * dR = 2R (or dR_override > 0).  R = max(uR, vR).  Host only. */
int bimsim_sphere_lut_dims(const bimsim_quad* pos, const bimsim_sphere* s, int R,
                           int dR_override, int* aR, int* dR, double* dmin, double* dmax);

/* Scattered / internal field tables of one sphere (sphere::calcHankelLut,
 * calcBesselLut and the gpuScalarUsp / gpuScalarUip kernels, reference
 * sphere.cpp:99-220, sphere.cu:4-149): Usp[itheta*dR + ir] and
 * Uip[itheta*aR + ia] with theta = itheta * PI/(thetaR - 1).  Radial Bessel
 * functions in FP64 on the GPU; tables complex128 (FP64 precisions) or
 * complex64 (BIMSIM_FP32_FAST).  Uip may be NULL. */
int bimsim_sphere_lut(bimsim_ctx* ctx, const bimsim_sphere* s, const void* A, const void* B,
                      double kmag, double dmin, double dmax, int dR, int aR, int thetaR,
                      int prec, void* Usp, void* Uip);

/* The radial tables alone (sphere::calcHankelLut / calcBesselLut, reference
 * sphere.cpp:99-203), device complex128, row-major with l fastest:
 * h[ir*(Nl+1) + l] = (j_l + i y_l)(kmag r), r over [max(a, dmin), dmax] in
 * dR samples; j[ia*(Nl+1) + l] = j_l(kmag n rho), rho over [0, a] in aR
 * samples with rho = 0 -> delta_l0.  Either output may be NULL. */
int bimsim_sphere_radial_lut(bimsim_ctx* ctx, const bimsim_sphere* s, double kmag, double dmin,
                             double dmax, int dR, int aR, void* h, void* j);

/* U += scattered field of one sphere from its tables under nk waves
 * (gpuScalarUpLut, reference nfScalarUpLut.cu:11-97; texture fetches as
 * bilinear interpolation with exact weights and clamped edges). */
int bimsim_nf_us_lut(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                     const double focus[3], double kmag, const bimsim_sphere* s,
                     const void* Usp, int dR, const void* Uip, int aR, int thetaR,
                     double dmin, double dmax, const double* waves, int nk, double subA,
                     int prec, void* U);

/* Rotation of the saved fields by deg about the slice centre
 * (nearfieldStruct::Interpolate, reference nfInterpolate.cu:13-115): Uf with
 * the float centre of gpuInterpolateUf, U with the integer centre of
 * gpuInterpolateU; bilinear, zero outside.  Either pair may be NULL. */
int bimsim_nf_interpolate(bimsim_ctx* ctx, int prec, double deg, int uR, int vR,
                          const void* U_src, const void* Uf_src, void* U, void* Uf);

/* ------------------------------- extended sources, detector-side images
 * Images are uR x vR real samples (float for FP32_FAST, double otherwise);
 * the centre is ((int)(uR/2.0), (int)(vR/2.0)). */

/* D += scale * src rotated by deg about the centre, bilinear with a zero
 * border (InterpolateAndIntegrateUd / InterpolateAndIntegrateUfd, reference
 * microscope.cu:89-132). */
int bimsim_detector_rotate_add(bimsim_ctx* ctx, int prec, double deg, int uR, int vR,
                               double scale, const void* src, void* D);

/* line[r] += sum over n_samples points of the ring of radius r about the
 * centre, for r = 0..n_rings-1: point t at angle (2 PI/n_samples)(t+1),
 * (uc + r cos, vc - r sin), bilinear with a zero border; the sum is the
 * reference's tree (step n_samples/2, halving to 1), which leaves out points
 * when n_samples is not a power of two (simRing_Ud / simRing_Ufd, reference
 * microscope.cu:138-240).  1 <= n_samples <= 1023. */
int bimsim_detector_ring_sums(bimsim_ctx* ctx, int prec, int uR, int vR, int n_samples,
                              int n_rings, const void* src, void* line);

/* D = scale * line at the distance |(iu, iv) - centre| (float), linear
 * between samples with a zero border outside [0, len)
 * (InterpolateAndIntegrateUdFromLine / ...UfdFromLine, reference
 * microscope.cu:37-86). */
int bimsim_detector_from_line(bimsim_ctx* ctx, int prec, int uR, int vR, double scale,
                              const void* line, int len, void* D);

/* ---------------------------------------------------------------- batches
 * Many independent absorbance images (spectral sweeps, ragged sets of
 * spheres and grids) in one pass: every instance is a microscopeStruct::
 * Simulate (reference microscope.cu:849-857) with its own slice, spheres,
 * wavelength, optics and waves.  The Mie coefficients of all spheres, the
 * incident fields of all instances and the Uf + Us sums run as one launch
 * each; the FP32 scattered-field series runs as one persistent work-queue
 * launch per order bucket over (instance, pixel block, wave split) items.
 * With use_graph the whole pass is captured once into a HIP graph and
 * bimsim_batch_launch replays it. */

typedef struct {
    bimsim_quad pos;        /* near-field slice (nearfieldStruct::pos)          */
    int    N;               /* samples per side of the near field (uR = vR = N) */
    int    m;               /* incident-field order (--field-order)             */
    double focus[3];
    double kdir[3];         /* incident direction                               */
    double lambda;          /* wavelength; kmag = 2*PI/lambda, PI = 3.14159f    */
    double amp;             /* incident amplitude A                             */
    double condenser[2];    /* NAin, NAout                                      */
    double objective[2];    /* NAin, NAout                                      */
    int    uf_mode;         /* incident field (bimsim_uf_mode): BIMSIM_UF_SERIES
                               (Legendre series), BIMSIM_UF_PLANE (plane wave
                               along kdir for Uf and a single wave for Us) or
                               BIMSIM_UF_MC (Monte-Carlo sum of the waves)   */
    int    sphere_begin;    /* spheres [sphere_begin, sphere_begin + ns)        */
    int    ns;
    int    wave_begin;      /* Monte-Carlo waves [wave_begin, wave_begin + nk)  */
    int    nk;
    int    padding;         /* detector (microscopeStruct::padding, ss, Nd)     */
    int    Nd;
    int    ss;
    long long image_off;    /* element offset of the instance's Nd x Nd images  */
    int    zero_fill;       /* >= 2: the detector of bimsim_detector_zero_filled
                               with this factor (Nd = N, padding 0, ss 1);
                               0 or 1: the reference detector               */
} bimsim_instance;

typedef struct bimsim_batch bimsim_batch;

/* spheres: host array; n is the refractive index at the owning instance's
 * wavelength, Nl <= 0 asks for sphere::calcNl, coeff_off is ignored.
 * waves: device, rows of 3 doubles.  D, Di (and Abs unless NULL): device
 * images, float for BIMSIM_FP32_FAST and double otherwise; they are
 * overwritten.  The device buffers must stay valid while the batch lives.
 * Creation runs the pass once (results in D/Di/Abs) and synchronises. */
int bimsim_batch_create(bimsim_ctx* ctx, const bimsim_instance* inst, int n,
                        const bimsim_sphere* spheres, int n_spheres,
                        const double* waves, int prec, int quirks, void* D,
                        void* Di, void* Abs, int use_graph, bimsim_batch** out);
/* Runs the pass again (asynchronously on the context stream). */
int bimsim_batch_launch(bimsim_batch* batch);
int bimsim_batch_destroy(bimsim_batch* batch);
/* Create, run once and destroy. */
int bimsim_batch_run(bimsim_ctx* ctx, const bimsim_instance* inst, int n,
                     const bimsim_sphere* spheres, int n_spheres, const double* waves,
                     int prec, int quirks, void* D, void* Di, void* Abs);

#ifdef __cplusplus
}
#endif

#endif /* BIMSIM_HIP_H */
