// microscopeStruct over the C ABI (reference microscope.cu:495-857).
#include <cmath>
#include <iostream>
#include <sstream>
#include <stdexcept>

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

// microscope.cu:592-605: U (not Uf) propagated by z_d in its angular
// spectrum; the reference does not call it from Simulate.
void microscopeStruct::applyBackpropagation() {
    nf.U.toAngularSpectrum();
    const double du = 1.0 / nf.pos.X.len(), dv = 1.0 / nf.pos.Y.len();
    check(bimsim_field_backpropagate(context(), nf.U.prec, settings().quirks, nf.U.x_hat,
                                     (int)nf.U.R[0], (int)nf.U.R[1], du, dv, nf.lambda, z_d),
          "applyBackpropagation");
    nf.U.fromAngularSpectrum();
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
    // the fused detector sums |x|^2 only: vector fields take the separate
    // steps (band-pass on x as the reference, crop and |x|^2+|y|^2+|z|^2)
    if (fused_detector && square && nf.U.scalarField) {
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

// Ring of numPoints sources at `radius` (microscope.cu:1298-1325): the angle
// steps by 2 PI/numPoints in float starting at one step, and each source
// keeps its own angle as its rotation.
void microscopeStruct::generatePointSources(int numPoints, float radius, float amplitude) {
    focalPointsES.clear();
    const float step = 2 * PI / (float)numPoints;
    float theta = step;
    for (int i = 0; i < numPoints; ++i) {
        sourcePointES p;
        p.rotationAngle = theta;
        const float u = std::cos(theta) * radius, v = std::sin(theta) * radius;
        p.f = bsPoint(u, v, 0);
        p.A = amplitude;
        focalPointsES.push_back(p);
        theta += step;
    }
}

void microscopeStruct::saveDetectorForOnePs() {
    Ud.ResampleAndSave(D_ps1, ss);
    Ufd.ResampleAndSave(Di_ps1, ss);
}

// The reference copies D_ps1 / Di_ps1 into textures; the samplers here read
// the images themselves, which stay unchanged until the next
// saveDetectorForOnePs.
void microscopeStruct::SaveDetectorToTexture() {}

// Likewise for the radial lines (read in place by InterpLineAtDetector).
void microscopeStruct::SaveLineToTexture(int radiusLength) { line_len = radiusLength; }

void microscopeStruct::InterpAtDetector() {
    bimsim_ctx* ctx = context();
    check(bimsim_detector_rotate_add(ctx, Di->prec, Di->rotationAngle, Di->R[0], Di->R[1], scale,
                                     Di_ps1->S, Di->S),
          "InterpAtDetector");
    check(bimsim_detector_rotate_add(ctx, D->prec, D->rotationAngle, D->R[0], D->R[1], scale,
                                     D_ps1->S, D->S),
          "InterpAtDetector");
}

void microscopeStruct::InterpLineAtDetector() {
    bimsim_ctx* ctx = context();
    check(bimsim_detector_from_line(ctx, Di->prec, Di->R[0], Di->R[1], scale, Di->line, line_len,
                                    Di->S),
          "InterpLineAtDetector");
    check(bimsim_detector_from_line(ctx, D->prec, D->R[0], D->R[1], scale, D->line, line_len, D->S),
          "InterpLineAtDetector");
}

void microscopeStruct::InterpLineForAllPointSources(int numSamples, int numRings) {
    if (!D->line || !Di->line)
        throw std::runtime_error("bimsim: the radial lines exist only with --interpolate 2");
    bimsim_ctx* ctx = context();
    check(bimsim_detector_ring_sums(ctx, Di->prec, Di_ps1->R[0], Di_ps1->R[1], numSamples, numRings,
                                    Di_ps1->S, Di->line),
          "InterpLineForAllPointSources");
    check(bimsim_detector_ring_sums(ctx, D->prec, D_ps1->R[0], D_ps1->R[1], numSamples, numRings,
                                    D_ps1->S, D->line),
          "InterpLineForAllPointSources");
}

// microscope.cu:1004-1095
void microscopeStruct::SimulateExtendedSourceInterpAtDetector(int* pointsRings) {
    clearDetector();
    const int numRings = pointsRings[1];
    scale = 1.0f;
    Simulate();   // the source at the current focus
    const int npts = pointsRings[0];
    const float spacing = 1;
    float radius = spacing;
    for (float j = spacing; j < numRings; j += spacing) {
        generatePointSources(npts, radius, 1);
        nf.focus = focalPointsES[0].f;
        nf.A = focalPointsES[0].A;
        Simulate();
        saveDetectorForOnePs();
        SaveDetectorToTexture();
        for (int i = 1; i < npts; ++i) {
            nf.focus = focalPointsES[i].f;
            nf.A = focalPointsES[i].A;
            nf.rotationAngle = focalPointsES[i].rotationAngle;
            D->rotationAngle = focalPointsES[i].rotationAngle;
            Di->rotationAngle = focalPointsES[i].rotationAngle;
            InterpAtDetector();
        }
        radius += spacing;
    }
}

// microscope.cu:1097-1187.  The detector images are replaced by the radial
// lines at the end, so the sources' own images do not reach them.
void microscopeStruct::SimulateExtendedSourceInterpLineAtDetector(int* pointsRings) {
    clearDetector();
    const int numRings = pointsRings[1] - 1;
    scale = 1.0f;
    nf.A = 1;
    Simulate();   // the source at the current focus
    saveDetectorForOnePs();
    SaveDetectorToTexture();
    InterpLineForAllPointSources(pointsRings[0], pointsRings[1]);
    for (int j = 1; j <= numRings; ++j) {
        nf.focus = bsPoint((ptype)j, 0, 0);
        nf.A = 1;
        Simulate();
        saveDetectorForOnePs();
        SaveDetectorToTexture();
        InterpLineForAllPointSources(pointsRings[0], pointsRings[1]);
    }
    SaveLineToTexture(pointsRings[1]);
    InterpLineAtDetector();
}

// microscope.cu:1189-1256
void microscopeStruct::SimulateExtendedSourceWithInterpolation(float, int* pointsRings) {
    scale = 1;
    clearDetector();
    const int numRings = pointsRings[1];
    Simulate();   // the source at the current focus
    const float spacing = 1;
    float radius = spacing;
    for (int j = 1; j < numRings; ++j) {
        const int npts = pointsRings[0];
        generatePointSources(npts, radius, 1);
        radius += spacing;
        nf.focus = focalPointsES[0].f;
        nf.A = focalPointsES[0].A;
        SimulateAndSaveTexture();
        for (int i = 1; i < npts; ++i) {
            nf.focus = focalPointsES[i].f;
            nf.A = focalPointsES[i].A;
            nf.rotationAngle = focalPointsES[i].rotationAngle;
            SimulateWithInterpolation();
        }
    }
}

// microscope.cu:1327-1410.  After the source at the current focus the
// reference steps its ring counter, an int, by 0.5 from 0, so for one ring
// or more it never leaves the loop; that case is an error here.
void microscopeStruct::SimulateExtendedGausssianSourceNoInterp(int* pointsRings) {
    clearDetector();
    const int numRings = pointsRings[1];
    Simulate();
    if (numRings > 0)
        throw std::runtime_error(
            "bimsim: --interpolate 0 with one ring or more does not terminate in the reference "
            "(its ring counter is an int stepped by 0.5); use --points-rings <sigma> 0");
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
