// nearfieldStruct and sphere over the C ABI (reference nearfield.cpp,
// sphere.cpp, nfScalarUf.cu:177-217, nfScalarUfLut.cu:125-188,
// nfScalarUs.cu:494-658, nfScalarUpLut.cu:342-495, nfSumUf.cu:51-117,
// nfInterpolate.cu:68-115).  Host-derived constants are computed as the
// reference does with ptype = double: k = 2*PI/lambda with the float PI,
// cos(asin(NA)) and the condenser amplitude subA.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <sstream>

#include "nearfield.h"

using namespace bimsim;

namespace {

void hip_ok(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("bimsim: ") + what + ": " + hipGetErrorString(e));
}

void grow(void** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes) return;
    if (*p) hip_ok(hipFree(*p), "hipFree");
    *p = nullptr;
    hip_ok(hipMalloc(p, bytes), "hipMalloc");
    *cap = bytes;
}

double kmag_of(ptype lambda) { return 2 * PI / lambda; }

}  // namespace

// --------------------------------------------------------------------- sphere

sphere::sphere(ptype x, ptype y, ptype z, ptype a_, int m, int ang) : p(x, y, z), a(a_), iMaterial(m) {
    (void)ang;
}

sphere::sphere(const sphere& rhs)
    : p(rhs.p), a(rhs.a), iMaterial(rhs.iMaterial), Nl(rhs.Nl), n(rhs.n), B(rhs.B), A(rhs.A) {}

sphere& sphere::operator=(const sphere& rhs) {
    p = rhs.p;
    a = rhs.a;
    iMaterial = rhs.iMaterial;
    Nl = rhs.Nl;
    n = rhs.n;
    B = rhs.B;
    A = rhs.A;
    return *this;
}

bimsim_sphere sphere::abi(int coeff_off) const {
    bimsim_sphere s;
    for (int i = 0; i < 3; ++i) s.p[i] = p[i];
    s.a = a;
    s.n[0] = n.real();
    s.n[1] = n.imag();
    s.Nl = Nl;
    s.coeff_off = coeff_off;
    return s;
}

void sphere::calcNl(ptype lambda) { Nl = bimsim_sphere_order(a, lambda); }

void sphere::calcCoeff(ptype lambda, bsComplex n_) {
    n = n_;
    bimsim_ctx* ctx = context();
    const bimsim_sphere s = abi(0);
    void *dA = nullptr, *dB = nullptr;
    const size_t bytes = 16 * (size_t)(Nl + 1);
    hip_ok(hipMalloc(&dA, bytes), "hipMalloc");
    hip_ok(hipMalloc(&dB, bytes), "hipMalloc");
    const int rc = bimsim_mie_coeffs(ctx, &s, 1, kmag_of(lambda), settings().quirks, dA, dB);
    A.assign(Nl + 1, bsComplex());
    B.assign(Nl + 1, bsComplex());
    if (rc == BIMSIM_OK) {
        hip_ok(hipMemcpy(A.data(), dA, bytes, hipMemcpyDeviceToHost), "download");
        hip_ok(hipMemcpy(B.data(), dB, bytes, hipMemcpyDeviceToHost), "download");
    }
    (void)hipFree(dA);
    (void)hipFree(dB);
    check(rc, "calcCoeff");
}

static void radial_tables(const sphere& sp, double kmag, double dmin, double dmax, int dR, int aR,
                          bsComplex* h, bsComplex* j) {
    bimsim_ctx* ctx = context();
    const bimsim_sphere s = sp.abi(0);
    const size_t row = 16 * (size_t)(sp.Nl + 1);
    void *dh = nullptr, *dj = nullptr;
    if (h) hip_ok(hipMalloc(&dh, row * dR), "hipMalloc");
    if (j) hip_ok(hipMalloc(&dj, row * aR), "hipMalloc");
    const int rc = bimsim_sphere_radial_lut(ctx, &s, kmag, dmin, dmax, dR, aR, dh, dj);
    if (rc == BIMSIM_OK) {
        if (h) hip_ok(hipMemcpy(h, dh, row * dR, hipMemcpyDeviceToHost), "download");
        if (j) hip_ok(hipMemcpy(j, dj, row * aR, hipMemcpyDeviceToHost), "download");
    }
    (void)hipFree(dh);
    (void)hipFree(dj);
    check(rc, "radial tables");
}

void sphere::calcBesselLut(bsComplex* j, ptype k, bsComplex n_, int aR) {
    n = n_;
    radial_tables(*this, k, 0, 0, 0, aR, nullptr, j);
}

void sphere::calcHankelLut(bsComplex* h, ptype k, int rR) {
    radial_tables(*this, k, d_min, d_max, rR, 0, h, nullptr);
}

void sphere::calcLut(bsComplex* j, bsComplex* h, ptype lambda, bsComplex n_, int aR, int rR) {
    const double k = kmag_of(lambda);
    calcBesselLut(j, k, n_, aR);
    calcHankelLut(h, k, rR);
}

void sphere::calcUp(ptype lambda, bsComplex n_, bsRect nfPlane, unsigned int R) {
    n = n_;
    bimsim_ctx* ctx = context();
    const bimsim_quad q = nfPlane.abi();
    bimsim_sphere s = abi(0);
    int aR = 0, dR = 0;
    double dmin = 0, dmax = 0;
    check(bimsim_sphere_lut_dims(&q, &s, (int)R, 0, &aR, &dR, &dmin, &dmax), "sphere_lut_dims");
    d_min = dmin;
    d_max = dmax;
    const int thetaR = 1000;   // DEFAULT_SPHERE_THETA_R
    const int prec = settings().precision;
    Usp.prec = Uip.prec = prec;
    Usp.R[0] = dR;
    Usp.R[1] = thetaR;
    Usp.init_gpu();
    Uip.R[0] = aR;
    Uip.R[1] = thetaR;
    Uip.init_gpu();
    const size_t bytes = 16 * (size_t)(Nl + 1);
    void *dA = nullptr, *dB = nullptr;
    hip_ok(hipMalloc(&dA, bytes), "hipMalloc");
    hip_ok(hipMalloc(&dB, bytes), "hipMalloc");
    hip_ok(hipMemcpy(dA, A.data(), bytes, hipMemcpyHostToDevice), "upload");
    hip_ok(hipMemcpy(dB, B.data(), bytes, hipMemcpyHostToDevice), "upload");
    const int rc = bimsim_sphere_lut(ctx, &s, dA, dB, kmag_of(lambda), dmin, dmax, dR, aR, thetaR,
                                     prec, Usp.x_hat, Uip.x_hat);
    const int rs = bimsim_sync(ctx);
    (void)hipFree(dA);
    (void)hipFree(dB);
    check(rc, "calcUp");
    check(rs, "calcUp");
}

// ------------------------------------------------------------ nearfieldStruct

nearfieldStruct::nearfieldStruct() {}

nearfieldStruct::~nearfieldStruct() { free_device(); }

void nearfieldStruct::free_device() {
    if (d_waves) (void)hipFree(d_waves);
    if (d_A) (void)hipFree(d_A);
    if (d_B) (void)hipFree(d_B);
    d_waves = d_A = d_B = nullptr;
    d_waves_cap = d_coef_cap = 0;
}

void nearfieldStruct::init() {
    U.scalarField = Uf.scalarField = scalarSim;
    U.prec = Uf.prec = settings().precision;
    U.init_gpu();
    Uf.init_gpu();
}

void nearfieldStruct::destroy() {
    U.kill_gpu();
    Uf.kill_gpu();
    savedU.kill_gpu();
    savedUf.kill_gpu();
    free_device();
}

void nearfieldStruct::setPos(bsPoint pMin, bsPoint pMax, bsVector normal) {
    pos = bsRect(pMin, pMax, normal);
}

void nearfieldStruct::setRes(int x_res, int y_res) {
    U.R[0] = Uf.R[0] = x_res;
    U.R[1] = Uf.R[1] = y_res;
}

std::string nearfieldStruct::toStr() {
    std::stringstream ss;
    ss << "---------Timings-------------" << std::endl;
    ss << "Uf = " << t_Uf << "ms" << std::endl;
    ss << "Us = " << t_Us << "ms" << std::endl;
    ss << std::endl;
    ss << "------Field Parameters-------" << std::endl;
    ss << "Wavelength: " << lambda << "um" << std::endl;
    ss << "K Vector (x, y, z): " << k << std::endl;
    ss << "K Vector (r, theta, phi): " << k.cart2sph() << std::endl;
    ss << "Condenser NA: " << condenser[0] << " to " << condenser[1] << std::endl;
    ss << "Number of points and rings: " << extSource[0] << " to " << extSource[1] << std::endl;
    ss << "Focal Point: " << focus[0] << ", " << focus[1] << ", " << focus[2] << std::endl;
    if (lut_us) ss << "LUT Parameters --- min: " << d_min << "   max: " << d_max << std::endl;
    ss << pos << std::endl;
    ss << std::endl << "---------Materials-----------" << std::endl;
    ss << "Number of Materials: " << mVector.size() << std::endl;
    ss << "Refractive Indices at lambda = " << lambda << "um" << std::endl;
    for (size_t i = 0; i < mVector.size(); ++i) ss << " " << i << ": " << mVector[i](lambda) << std::endl;
    ss << "---------Spheres-------------" << std::endl;
    ss << "Number of Spheres: " << sVector.size() << std::endl;
    for (const sphere& s : sVector) ss << s.toStr() << std::endl;
    return ss.str();
}

void nearfieldStruct::calcWaves() {
    // srand(NULL) before mcSampleNA (nearfield.cpp:96-108): the glibc stream
    // of seed 0 unless the settings pick another generator
    inWaves.assign(nWaves, bsVector());
    if (nWaves <= 0) return;
    std::vector<double> w(3 * (size_t)nWaves);
    const double kk[3] = {k[0], k[1], k[2]};
    check(bimsim_mc_sample_na(w.data(), nWaves, kk, condenser[0], condenser[1], settings().mc_rng,
                              settings().mc_seed),
          "calcWaves");
    // the reference stores the directions as float vectors
    for (int i = 0; i < nWaves; ++i)
        inWaves[i] = bsVector((float)w[3 * i], (float)w[3 * i + 1], (float)w[3 * i + 2]);
}

void nearfieldStruct::calcSpheres() {
    for (sphere& s : sVector) {
        s.calcNl(lambda);
        if (s.iMaterial >= mVector.size())
            throw std::runtime_error("BIMSIM Error - A material is not loaded for a sphere");
        const bsComplex n = mVector[s.iMaterial](lambda);
        s.calcCoeff(lambda, n);
        s.n = n;
        if (lut_us) s.calcUp(lambda, n, pos, std::max(U.R[0], U.R[1]));
    }
}

void nearfieldStruct::calcUf() {
    if (!scalarSim) {
        if (lut_uf) vectorUfLut();
        else vectorUf();
        return;
    }
    if (lut_uf) scalarUfLut();
    else scalarUf();
}

// Vector simulation (nearfieldStruct::vectorUf, nfVectorUf.cu:162-191): the
// scalar focused field (or plane wave) in the x component, y and z zero.
void nearfieldStruct::vectorUf() {
    scalarUf();
    Uf.clear_yz();
}

// The reference's vectorUfLut has an empty body (nfVectorUfLut.cu:7-10), which
// leaves Uf zero; this one evaluates the tabulated form into x.
void nearfieldStruct::vectorUfLut() {
    scalarUfLut();
    Uf.clear_yz();
}

static float stage_ms(bimsim_ctx* ctx, int stage) {
    float ms[4] = {0, 0, 0, 0};
    int counts[4] = {0, 0, 0, 0};
    check(bimsim_get_timings(ctx, ms, counts), "timings");
    return ms[stage];
}

void nearfieldStruct::scalarUf() {
    bimsim_ctx* ctx = context();
    check(bimsim_set_timing(ctx, 1), "timing");
    const bimsim_quad q = pos.abi();
    const double f[3] = {focus[0], focus[1], focus[2]};
    const double kd[3] = {k[0], k[1], k[2]};
    const double ca = std::cos(std::asin(condenser[0]));
    const double cb = std::cos(std::asin(condenser[1]));
    check(bimsim_nf_uf(ctx, &q, (int)Uf.R[0], (int)Uf.R[1], f, kd, kmag_of(lambda), A, ca, cb, m,
                       planeWave ? BIMSIM_UF_PLANE : BIMSIM_UF_SERIES, Uf.prec, settings().quirks,
                       Uf.x_hat),
          "scalarUf");
    t_Uf = stage_ms(ctx, 0);
}

void nearfieldStruct::scalarUfLut() {
    bimsim_ctx* ctx = context();
    check(bimsim_set_timing(ctx, 1), "timing");
    const bimsim_quad q = pos.abi();
    const double f[3] = {focus[0], focus[1], focus[2]};
    const double kd[3] = {k[0], k[1], k[2]};
    const double ca = std::cos(std::asin(condenser[0]));
    const double cb = std::cos(std::asin(condenser[1]));
    // the table range (nfScalarUfLut.cu:129-130), kept for toStr
    bimsim_sphere at_focus{};
    for (int i = 0; i < 3; ++i) at_focus.p[i] = focus[i];
    at_focus.a = 1;
    int aR, dR;
    double dmin, dmax;
    check(bimsim_sphere_lut_dims(&q, &at_focus, 2, 0, &aR, &dR, &dmin, &dmax), "scalarUfLut");
    d_min = dmin;
    d_max = dmax;
    check(bimsim_nf_uf(ctx, &q, (int)Uf.R[0], (int)Uf.R[1], f, kd, kmag_of(lambda), A, ca, cb, m,
                       planeWave ? BIMSIM_UF_PLANE : BIMSIM_UF_LUT, Uf.prec, settings().quirks,
                       Uf.x_hat),
          "scalarUfLut");
    t_Uf = stage_ms(ctx, 0);
}

void nearfieldStruct::calcBesselLut(ptype* j, ptype dmin, ptype dmax, int dR) {
    bimsim_ctx* ctx = context();
    void* dj = nullptr;
    const size_t bytes = sizeof(double) * (size_t)dR * (m + 1);
    hip_ok(hipMalloc(&dj, bytes), "hipMalloc");
    const int rc = bimsim_uf_bessel_lut(ctx, m, kmag_of(lambda), dmin, dmax, dR, BIMSIM_FP64, dj);
    if (rc == BIMSIM_OK) hip_ok(hipMemcpy(j, dj, bytes, hipMemcpyDeviceToHost), "download");
    (void)hipFree(dj);
    check(rc, "calcBesselLut");
}

void nearfieldStruct::calcUs() {
    if (!scalarSim) {
        // the reference has no vector scattered field: it prints a note and
        // leaves U as it is (nearfield.cpp:160-165); here U is zero, so the
        // vector image is that of the incident field outside the spheres
        U.clear_gpu();
        return;
    }
    if (lut_us) scalarUpLut();
    else scalarUs();
}

void nearfieldStruct::upload_waves(const double* w, int nk) {
    grow(&d_waves, &d_waves_cap, sizeof(double) * 3 * (size_t)nk);
    hip_ok(hipMemcpy(d_waves, w, sizeof(double) * 3 * (size_t)nk, hipMemcpyHostToDevice), "upload");
}

void nearfieldStruct::upload_coefficients() {
    size_t total = 0;
    for (const sphere& s : sVector) total += (size_t)s.Nl + 1;
    const size_t bytes = 16 * std::max<size_t>(total, 1);
    if (d_coef_cap < bytes) {
        size_t cap = d_coef_cap;
        grow(&d_A, &cap, bytes);
        cap = d_coef_cap;
        grow(&d_B, &cap, bytes);
        d_coef_cap = bytes;
    }
    size_t off = 0;
    for (const sphere& s : sVector) {
        const size_t b = 16 * ((size_t)s.Nl + 1);
        if (s.A.size() != (size_t)s.Nl + 1 || s.B.size() != (size_t)s.Nl + 1)
            throw std::runtime_error("bimsim: sphere coefficients missing (calcSpheres first)");
        hip_ok(hipMemcpy((char*)d_A + 16 * off, s.A.data(), b, hipMemcpyHostToDevice), "upload");
        hip_ok(hipMemcpy((char*)d_B + 16 * off, s.B.data(), b, hipMemcpyHostToDevice), "upload");
        off += (size_t)s.Nl + 1;
    }
}

void nearfieldStruct::scalarUs() {
    if (sVector.empty()) return;
    bimsim_ctx* ctx = context();
    check(bimsim_set_timing(ctx, 1), "timing");
    upload_coefficients();
    std::vector<bimsim_sphere> S;
    int off = 0;
    for (const sphere& s : sVector) {
        S.push_back(s.abi(off));
        off += s.Nl + 1;
    }
    int nk;
    double subA;
    if (planeWave) {
        const double w[3] = {k[0], k[1], k[2]};
        upload_waves(w, 1);
        nk = 1;
        subA = A;
    } else {
        nk = (int)inWaves.size();
        std::vector<double> w(3 * (size_t)nk);
        for (int i = 0; i < nk; ++i)
            for (int c = 0; c < 3; ++c) w[3 * i + c] = inWaves[i][c];
        upload_waves(w.data(), nk);
        subA = 2 * PI * A * ((1 - std::cos(std::asin(condenser[1]))) - (1 - std::cos(std::asin(condenser[0]))));
    }
    if (nk == 0) {
        U.clear_gpu();
        return;
    }
    const bimsim_quad q = pos.abi();
    const double f[3] = {focus[0], focus[1], focus[2]};
    check(bimsim_nf_us(ctx, &q, (int)U.R[0], (int)U.R[1], f, kmag_of(lambda), S.data(), (int)S.size(),
                       d_A, d_B, (const double*)d_waves, nk, subA, U.prec, settings().quirks, U.x_hat, 0),
          "scalarUs");
    t_Us = stage_ms(ctx, 1);
}

void nearfieldStruct::scalarUpLut() {
    if (sVector.empty()) return;
    bimsim_ctx* ctx = context();
    check(bimsim_set_timing(ctx, 1), "timing");
    U.clear_gpu();
    int nk;
    double subA;
    if (planeWave) {
        const double w[3] = {k[0], k[1], k[2]};
        upload_waves(w, 1);
        nk = 1;
        subA = A;
    } else {
        nk = (int)inWaves.size();
        std::vector<double> w(3 * (size_t)nk);
        for (int i = 0; i < nk; ++i)
            for (int c = 0; c < 3; ++c) w[3 * i + c] = inWaves[i][c];
        upload_waves(w.data(), nk);
        subA = 2 * PI * A * ((1 - std::cos(std::asin(condenser[1]))) - (1 - std::cos(std::asin(condenser[0]))));
    }
    if (nk == 0) return;
    const bimsim_quad q = pos.abi();
    const double f[3] = {focus[0], focus[1], focus[2]};
    for (const sphere& s : sVector) {
        if (!s.Usp.x_hat || !s.Uip.x_hat)
            throw std::runtime_error("bimsim: sphere tables missing (calcUp first)");
        const bimsim_sphere sa = s.abi(0);
        check(bimsim_nf_us_lut(ctx, &q, (int)U.R[0], (int)U.R[1], f, kmag_of(lambda), &sa, s.Usp.x_hat,
                               (int)s.Usp.R[0], s.Uip.x_hat, (int)s.Uip.R[0], (int)s.Usp.R[1], s.d_min,
                               s.d_max, (const double*)d_waves, nk, subA, U.prec, U.x_hat),
              "scalarUpLut");
    }
    t_Us = stage_ms(ctx, 1);
}

void nearfieldStruct::sumUf() {
    bimsim_ctx* ctx = context();
    const bimsim_quad q = pos.abi();
    std::vector<bimsim_sphere> S;
    for (const sphere& s : sVector) S.push_back(s.abi(0));
    // each component of a vector field (gpuScalarUsp, nfSumUf.cu:6-48)
    void* const uf[3] = {Uf.x_hat, Uf.y_hat, Uf.z_hat};
    void* const u[3] = {U.x_hat, U.y_hat, U.z_hat};
    for (int c = 0; c < 3; ++c)
        if (uf[c] && u[c])
            check(bimsim_nf_sum_uf(ctx, &q, (int)U.R[0], (int)U.R[1], S.empty() ? nullptr : S.data(),
                                   (int)S.size(), U.prec, uf[c], u[c]),
                  "sumUf");
}

void nearfieldStruct::saveFieldsToTexture() {
    savedU = U;
    savedUf = Uf;
}

void nearfieldStruct::Interpolate() {
    if (!savedU.x_hat || !savedUf.x_hat)
        throw std::runtime_error("bimsim: Interpolate needs saveFieldsToTexture first");
    check(bimsim_nf_interpolate(context(), U.prec, rotationAngle, (int)U.R[0], (int)U.R[1],
                                savedU.x_hat, savedUf.x_hat, U.x_hat, Uf.x_hat),
          "Interpolate");
}

void nearfieldStruct::Simulate() {
    t_Uf = 0;
    t_Us = 0;
    calcWaves();
    calcUf();
    calcSpheres();
    calcUs();
    sumUf();
}
