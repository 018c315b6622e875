// microscopeStruct over the C ABI (reference microscope.cu:495-857).
#include <iostream>
#include <sstream>

#include "microscope.h"

using namespace bimsim;

microscopeStruct::microscopeStruct() {}

microscopeStruct::~microscopeStruct() {
    delete D;
    delete Di;
    delete D_ps1;
    delete Di_ps1;
}

void microscopeStruct::setZdPos(float d) { z_d = d; }

void microscopeStruct::init() {
    nf.scalarSim = scalarSim;
    nf.init();
    const int prec = settings().precision;
    Ud.prec = Ufd.prec = prec;
    const int w = (int)Ud.R[0] / ss, h = (int)Ud.R[1] / ss;
    delete D;
    delete Di;
    delete D_ps1;
    delete Di_ps1;
    if (interpolate == 2) {
        D = new scalarslice(w, h, nf.extSource[1]);
        Di = new scalarslice(w, h, nf.extSource[1]);
    } else {
        D = new scalarslice(w, h);
        Di = new scalarslice(w, h);
    }
    D_ps1 = new scalarslice(w, h);
    Di_ps1 = new scalarslice(w, h);
    clearDetector();
}

void microscopeStruct::destroy() {
    delete D;
    D = nullptr;
    delete Di;
    Di = nullptr;
    delete D_ps1;
    D_ps1 = nullptr;
    delete Di_ps1;
    Di_ps1 = nullptr;
    Ud.kill_gpu();
    Ufd.kill_gpu();
    nf.destroy();
}

void microscopeStruct::setPos(bsPoint pMin, bsPoint pMax, bsVector normal) {
    pos = bsRect(pMin, pMax, normal);
}

void microscopeStruct::setRes(int x_res, int y_res, int pad, int supersampling) {
    padding = pad;
    ss = supersampling;
    Ufd.R[0] = Ud.R[0] = x_res * ss;
    Ufd.R[1] = Ud.R[1] = y_res * ss;
}

void microscopeStruct::setNearfield() {
    nf.pos = pos * (ptype)(padding * 2 + 1);
    nf.setRes(Ud.R[0] * (padding * 2 + 1), Ud.R[1] * (padding * 2 + 1));
}

void microscopeStruct::applyBandpass() {
    nf.U.toAngularSpectrum();
    nf.Uf.toAngularSpectrum();
    const double du = 1.0 / nf.pos.X.len(), dv = 1.0 / nf.pos.Y.len();
    bimsim_ctx* ctx = context();
    check(bimsim_field_bandpass(ctx, nf.U.prec, settings().quirks, nf.U.x_hat, (int)nf.U.R[0],
                                (int)nf.U.R[1], du, dv, objective[0], objective[1], nf.lambda),
          "applyBandpass");
    check(bimsim_field_bandpass(ctx, nf.Uf.prec, settings().quirks, nf.Uf.x_hat, (int)nf.Uf.R[0],
                                (int)nf.Uf.R[1], du, dv, objective[0], objective[1], nf.lambda),
          "applyBandpass");
    nf.U.fromAngularSpectrum();
    nf.Uf.fromAngularSpectrum();
}

void microscopeStruct::getFarField() {
    Ud = nf.U.crop(padding * (int)Ud.R[0], padding * (int)Ud.R[1], (int)Ud.R[0], (int)Ud.R[1]);
    Ufd = nf.Uf.crop(padding * (int)Ufd.R[0], padding * (int)Ufd.R[1], (int)Ufd.R[0], (int)Ufd.R[1]);
}

void microscopeStruct::integrateDetector() {
    scale = 1;
    Ud.IntegrateAndResample(D, ss, scale);
    Ufd.IntegrateAndResample(Di, ss, scale);
}

void microscopeStruct::clearDetector() {
    if (D) D->clear();
    if (Di) Di->clear();
    if (D_ps1) D_ps1->clear();
    if (Di_ps1) Di_ps1->clear();
    if (interpolate == 2) {
        if (Di) Di->clearLine(nf.extSource[1]);
        if (D) D->clearLine(nf.extSource[1]);
    }
}

void microscopeStruct::SimulateScattering() { nf.Simulate(); }

void microscopeStruct::saveToTexture() { nf.saveFieldsToTexture(); }

void microscopeStruct::SimulateInterpolation() { nf.Interpolate(); }

void microscopeStruct::SimulateImaging() {
    const int uR = (int)nf.U.R[0], vR = (int)nf.U.R[1];
    const bool square = uR == vR && Ud.R[0] == Ud.R[1];
    if (fused_detector && square) {
        // band-pass, crop and |.|^2 integration in one pass; the filtered
        // fields stay in nf.U / nf.Uf as after applyBandpass
        const double du = 1.0 / nf.pos.X.len(), dv = 1.0 / nf.pos.Y.len();
        const int rc = bimsim_detector(context(), nf.U.prec, settings().quirks, nf.U.x_hat, nf.Uf.x_hat,
                                       uR, vR, du, dv, objective[0], objective[1], nf.lambda, padding,
                                       D->R[0], ss, D->S, Di->S, nullptr, 1);
        if (rc == BIMSIM_OK) {
            getFarField();
            return;
        }
        if (rc != BIMSIM_E_UNSUPPORTED) check(rc, "SimulateImaging");
    }
    applyBandpass();
    getFarField();
    integrateDetector();
}

void microscopeStruct::Simulate() {
    SimulateScattering();
    SimulateImaging();
}

void microscopeStruct::LoadExtendedSource(std::string filename) {
    int Rx = 0, Ry = 0;
    const std::vector<unsigned char> img = read_bmp_channel0(filename, &Rx, &Ry);
    // x outer, y inner as the reference, so the detector sums run in its order
    for (int x = 0; x < Rx; ++x)
        for (int y = 0; y < Ry; ++y) {
            const unsigned char s = img[(size_t)y * Rx + x];
            if (s == 0) continue;
            // the reference keeps u, v and the amplitude in float
            const float u = (float)((ptype)x / (ptype)Rx), v = (float)((ptype)y / (ptype)Ry);
            sourcePoint p;
            p.f = pos(u, v);
            p.A = (ptype)((float)s / 255.0f);
            focalPoints.push_back(p);
        }
    if (settings().verbose) std::cout << "Number of point sources: " << focalPoints.size() << std::endl;
}

void microscopeStruct::SimulateExtendedSource() {
    clearDetector();
    scale = 1.0f;
    for (const sourcePoint& p : focalPoints) {
        nf.focus = p.f;
        nf.A = p.A;
        Simulate();   // the detector images accumulate
    }
}

void microscopeStruct::SimulateAndSaveTexture() {
    SimulateScattering();
    saveToTexture();
    SimulateImaging();
}

void microscopeStruct::SimulateWithInterpolation() {
    SimulateInterpolation();
    SimulateImaging();
}

static scalarslice ratio(scalarslice* D, scalarslice* Di, int op) {
    scalarslice out;
    out.init(D->R[0], D->R[1], D->prec);
    check(bimsim_slice_ratio(context(), D->prec, op, D->S, Di->S, (long long)D->R[0] * D->R[1], out.S),
          "detector ratio");
    return out;
}

scalarslice microscopeStruct::getAbsorbance() { return ratio(D, Di, BIMSIM_RATIO_ABSORBANCE); }

ptype microscopeStruct::getAbsorbanceSpectrum() { return 0; }

scalarslice microscopeStruct::getTransmittance() { return ratio(D, Di, BIMSIM_RATIO_TRANSMITTANCE); }

scalarslice microscopeStruct::getIntensity() { return scalarslice(*D); }

scalarslice microscopeStruct::getIncidentFieldImage() { return scalarslice(*Di); }

std::string microscopeStruct::toStr() {
    std::stringstream ss_;
    ss_ << nf.toStr();
    ss_ << "----------Optics--------------" << std::endl << std::endl;
    ss_ << "Objective NA: " << objective[0] << " to " << objective[1] << std::endl;
    return ss_.str();
}
