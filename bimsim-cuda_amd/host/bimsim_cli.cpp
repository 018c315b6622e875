// The bimsim command line (reference main.cpp:33-78, arguments.h:27-505,
// defaults.h): the same option names, defaults and "a default counts as
// set" semantics, loading one microscopeStruct and writing its images.
//
// Options beyond the reference's (device arithmetic and generator choice;
// they default to the reference's behaviour where one exists):
//   --precision fp32|fp64|parity   device arithmetic (default fp32)
//   --quirks 1|0                   reference numerics on/off (default 1)
//   --mc-rng glibc|splitmix64      Monte-Carlo directions (default glibc:
//                                  rand() after srand(0), nearfield.cpp:102)
//   --mc-seed N                    seed of the splitmix64 generator
//   --print-config                 print the loaded parameters and exit
//                                  without touching the GPU
#include <cmath>
#include <iomanip>
#include <iostream>
#include <limits>
#include <sstream>

#include "arglist.h"
#include "fileout.h"
#include "microscope.h"

using bimsim::settings;

namespace {

// defaults.h
const char* DEFAULT_LAMBDA = "1";
const char* DEFAULT_AMPLITUDE = "1";
const char* DEFAULT_MATERIAL = "1.4 0.05";
const char* DEFAULT_FOCUS = "0 0 0";
const char* DEFAULT_CONDENSER = "0 1";
const char* DEFAULT_POINTS_RINGS = "1 1";
const char* DEFAULT_OBJECTIVE = "0 1";
const char* DEFAULT_PLANE_MIN = "-5 0 -5";
const char* DEFAULT_PLANE_MAX = "5 0 5";
const char* DEFAULT_PLANE_NORM = "0 1 0";
const float DEFAULT_PLANE_SIZE = 40;
const float DEFAULT_PLANE_POSITION = 0;
const char* DEFAULT_FIELD_ORDER = "10";
const char* DEFAULT_SAMPLES = "400";
const char* DEFAULT_SLICE_RES = "256";
const char* DEFAULT_PADDING = "1";
const char* DEFAULT_SUPERSAMPLE = "1";
const char* DEFAULT_INTENSITY_FILE = "out_i.bmp";
const char* DEFAULT_INCIDENT_FILE = "out_inc.bmp";
const char* DEFAULT_TRANSMITTANCE_FILE = "";
const char* DEFAULT_ABSORBANCE_FILE = "out_a.bmp";
const char* DEFAULT_ABSORBANCE_SPECTRUM_FILE = "out_as.bmp";
const char* DEFAULT_NEAR_FILE = "out_n.bmp";
const char* DEFAULT_FAR_FILE = "out_f.bmp";
const char* DEFAULT_EXTENDED_SOURCE = "";
const char* DEFAULT_INTERPOLATE = "-1";
const char* DEFAULT_FIELD_TYPE = "magnitude";
const char* DEFAULT_COLORMAP = "brewer";

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void SetArguments(rts::arglist& args) {
    args.section("Interface Flags");
    args.add("help", "prints this help");
    args.add("verbose", "verbose output");

    args.section("Output Parameters");
    args.add("vector", "run a vector field simulation");
    args.add("intensity", "output measured intensity (filename)", DEFAULT_INTENSITY_FILE);
    args.add("incident", "output measured incident field (filename)", DEFAULT_INCIDENT_FILE);
    args.add("absorbance", "output measured absorbance (filename)", DEFAULT_ABSORBANCE_FILE);
    args.add("absorbance-spec", "output measured absorbance spectrum (filename)",
             DEFAULT_ABSORBANCE_SPECTRUM_FILE);
    args.add("transmittance", "output measured transmittance (filename)", DEFAULT_TRANSMITTANCE_FILE);
    args.add("far-field", "output far-field at detector (filename)", DEFAULT_FAR_FILE);
    args.add("near-field", "output field at focal plane (filename)", DEFAULT_NEAR_FILE);
    args.add("extended-source", "image of source at focus (filename)", DEFAULT_EXTENDED_SOURCE);
    args.add("interpolate", "option to interpolate the point sources", DEFAULT_INTERPOLATE);
    args.add("output-type", "output field value", DEFAULT_FIELD_TYPE,
             "magnitude, polarization, real, imaginary");
    args.add("colormap", "colormap", DEFAULT_COLORMAP, "gray, brewer");
    args.add("append", "append result to an existing (binary) file");

    args.section("Sphere Parameters");
    args.add("spheres", "sphere position", "", "--spheres x y z a m");
    args.add("sphere-file", "sphere file:", "", "[x y z radius material]");
    args.add("materials", "refractive indices as n, k pairs", DEFAULT_MATERIAL,
             "--materials n0 k0 n1 k1 n2 k2");
    args.add("material-file", "material file", "", "[lambda n k]");

    args.section("Optics");
    args.add("lambda", "incident wavelength (micrometers)", DEFAULT_LAMBDA);
    args.add("nu", "incident frequency (in cm^-1)\n(if specified, lambda is ignored)");
    args.add("k", "k-vector direction in spherical coordinates", "",
             "--k theta phi; theta = [0 2*pi], phi = [0 pi]");
    args.add("amplitude", "incident field amplitude", DEFAULT_AMPLITUDE);
    args.add("condenser", "condenser numerical aperature\nA pair of values specify an inner obscuration",
             DEFAULT_CONDENSER);
    args.add("objective", "objective numerical aperature\nA pair of values specify an inner obscuration",
             DEFAULT_OBJECTIVE);
    args.add("focus", "focal position for the incident point source", DEFAULT_FOCUS);
    args.add("plane-wave", "simulates an incident plane wave");
    args.add("points-rings",
             "pair of values specying the number of rings and the number of point sources per ring",
             DEFAULT_POINTS_RINGS);

    args.section("Imaging Parameters");
    args.add("resolution", "resolution of the detector", DEFAULT_SLICE_RES);
    args.add("plane-lower-left", "lower-left position of the image plane", DEFAULT_PLANE_MIN);
    args.add("plane-upper-right", "upper-right position of the image plane", DEFAULT_PLANE_MAX);
    args.add("plane-normal", "normal for the image plane", DEFAULT_PLANE_NORM);
    args.add("xy", "specify an x-y axis-aligned image (standard microscope)");
    args.add("xz", "specify a x-z axis-aligned image");
    args.add("yz", "specify a y-z axis-aligned image");

    args.section("Sampling Parameters");
    args.add("samples", "Monte-Carlo samples used to compute Us", DEFAULT_SAMPLES);
    args.add("padding", "FFT padding for the objective bandpass", DEFAULT_PADDING);
    args.add("supersample", "super-sampling rate for the detector field", DEFAULT_SUPERSAMPLE);
    args.add("field-order", "order of the incident field", DEFAULT_FIELD_ORDER);
    args.add("seed", "seed for the Monte-Carlo random number generator");
    args.add("recursive", "evaluate all Bessel functions recursively");
    args.add("recursive-us", "evaluate scattered-field Bessel functions recursively");
    args.add("lut-uf", "evaluate the focused-field using a look-up table");

    args.section("MI355X Build");
    args.add("precision", "device arithmetic of the fields", "fp32", "fp32, fp64, parity");
    args.add("quirks", "reproduce the reference's numerical defects", "1", "1, 0");
    args.add("mc-rng", "Monte-Carlo direction generator", "glibc", "glibc, splitmix64");
    args.add("mc-seed", "seed of the splitmix64 generator", "0");
    args.add("print-config", "print the loaded parameters and exit");
}

void lFlags(const rts::arglist& args, microscopeStruct* S) {
    if (args("verbose")) settings().verbose = true;
    if (args("recursive")) {
        S->nf.lut_us = false;
        S->nf.lut_uf = false;
    } else if (args("recursive-us")) {
        S->nf.lut_us = false;
    } else if (args("lut-uf")) {
        S->nf.lut_uf = true;
    }
}

void lWavelength(const rts::arglist& args, microscopeStruct* S, fileoutStruct* out) {
    if (args("nu")) {
        S->nf.lambda = (float)(10000 / args["nu"].as_float());
        out->wavenumber = true;
    } else {
        S->nf.lambda = args["lambda"].as_float();
        out->wavenumber = false;
    }
}

void lMaterials(const rts::arglist& args, microscopeStruct* S) {
    if (args("material-file")) {
        const rts::argument& files = args["material-file"];
        for (int i = 0; i < files.nargs(); ++i) S->nf.mVector.push_back(rts::material<ptype>(files.as_text(i)));
        return;
    }
    const rts::argument& mats = args["materials"];
    const int n = mats.nargs();
    if (n == 1) {
        S->nf.mVector.push_back(rts::material<ptype>(S->nf.lambda, mats.as_float(), 0));
    } else if (n % 2 != 0) {
        throw Fail("BIMSim Error: materials must be specified in n, k pairs");
    } else {
        for (int i = 0; i < n; i += 2)
            S->nf.mVector.push_back(rts::material<ptype>(S->nf.lambda, mats.as_float(i), mats.as_float(i + 1)));
    }
}

// "x y z a [m]" rows (arguments.h:193-230)
void lSpheresText(const std::string& list, microscopeStruct* S) {
    std::stringstream ss(list);
    while (!ss.eof()) {
        sphere s;
        float x = 0, y = 0, z = 0, a = 0;
        ss >> x >> y >> z >> a;
        s.p = bsPoint(x, y, z);
        s.a = a;
        if (ss.peek() != '\n') {
            unsigned int m = 0;
            ss >> m;
            s.iMaterial = m;
        }
        S->nf.sVector.push_back(s);
        ss.ignore(1000, '\n');
        ss.peek();
    }
}

void lSpheres(const rts::arglist& args, microscopeStruct* S) {
    if (args("spheres")) {
        const rts::argument& sa = args["spheres"];
        const int nArgs = sa.nargs();
        int nS;
        if (nArgs <= 5) nS = 1;
        else if (nArgs % 5 == 0) nS = nArgs / 5;
        else throw Fail("BIMSIM Error: Invalid number of sphere parameters.");
        // the values pass through a default-formatted stream (6 significant
        // digits), as in the reference
        std::stringstream ss;
        for (int s = 0; s < nS; ++s) {
            const int nP = nS == 1 ? nArgs : 5;
            for (int i = 0; i < nP; ++i) ss << sa.as_float(s * 5 + i) << " ";
            ss << std::endl;
        }
        lSpheresText(ss.str(), S);
    }
    if (args("sphere-file")) {
        const rts::argument& files = args["sphere-file"];
        for (int i = 0; i < files.nargs(); ++i) {
            std::ifstream ifs(files.as_text(i).c_str());
            if (!ifs) throw Fail("Error loading sphere file.");
            const std::string text((std::istreambuf_iterator<char>(ifs)), std::istreambuf_iterator<char>());
            lSpheresText(text, S);
        }
    }
    for (size_t s = 0; s < S->nf.sVector.size(); ++s)
        if (S->nf.sVector[s].iMaterial + 1 > S->nf.mVector.size()) {
            std::stringstream msg;
            msg << "BIMSIM Error - A material is not loaded for sphere " << s + 1 << "." << std::endl
                << "Material requested: " << S->nf.sVector[s].iMaterial + 1 << std::endl
                << "Number of materials: " << S->nf.mVector.size();
            throw Fail(msg.str());
        }
}

void lOptics(const rts::arglist& args, microscopeStruct* S) {
    if (args("objective")) {
        if (args["objective"].nargs() == 1) {
            S->objective[0] = 0.0;
            S->objective[1] = args["objective"].as_float();
        } else {
            S->objective[0] = args["objective"].as_float(0);
            S->objective[1] = args["objective"].as_float(1);
        }
    }
}

void lImagePlane(const rts::arglist& args, microscopeStruct* S) {
    bsPoint pMin, pMax;
    bsVector normal;
    float s = DEFAULT_PLANE_SIZE;
    float pos = DEFAULT_PLANE_POSITION;
    if (args("xy")) {
        if (args["xy"].nargs() >= 1) s = args["xy"].as_float(0);
        if (args["xy"].nargs() == 2) pos = args["xy"].as_float(1);
        pMin = bsPoint(-s / 2, -s / 2, pos);
        pMax = bsPoint(s / 2, s / 2, pos);
        normal = bsVector(0, 0, 1);
    } else if (args("xz")) {
        float size = DEFAULT_PLANE_SIZE;
        if (args["xz"].nargs() >= 1) size = args["xz"].as_float(0);
        if (args["xz"].nargs() >= 2) pos = args["xz"].as_float(1);
        pMin = bsPoint(-size / 2, pos, -size / 2);
        pMax = bsPoint(size / 2, pos, size / 2);
        normal = bsVector(0, -1, 0);
    } else if (args("yz")) {
        float size = DEFAULT_PLANE_SIZE;
        if (args["yz"].nargs() >= 1) size = args["yz"].as_float(0);
        if (args["yz"].nargs() >= 2) pos = args["yz"].as_float(1);
        pMin = bsPoint(pos, -size / 2, -size / 2);
        pMax = bsPoint(pos, size / 2, size / 2);
        normal = bsVector(1, 0, 0);
    } else {
        const rts::argument &ll = args["plane-lower-left"], &ur = args["plane-upper-right"],
                            &nn = args["plane-normal"];
        pMin = bsPoint(ll.as_float(0), ll.as_float(1), ll.as_float(2));
        pMax = bsPoint(ur.as_float(0), ur.as_float(1), ur.as_float(2));
        normal = bsVector(nn.as_float(0), nn.as_float(1), nn.as_float(2));
    }
    S->setPos(pMin, pMax, normal);
    S->setZdPos(pos);
    // detector supersampling from the plane size (arguments.h:408-416)
    const float res = args["resolution"].as_float();
    const float dx = s / res;
    const float lambda = (float)S->nf.lambda;
    int ss = 1;
    if (dx > lambda / 2.0) ss = (int)std::ceil(2 * s / (res * lambda));
    S->setRes((int)res, (int)res, (int)args["padding"].as_float(), ss);
    S->setNearfield();
}

void lNearfield(const rts::arglist& args, microscopeStruct* S) {
    S->scalarSim = !args("vector");
    S->nf.planeWave = args("plane-wave");
    S->nf.A = args["amplitude"].as_float();
    if (args["condenser"].nargs() == 1) {
        S->nf.condenser[0] = 0;
        S->nf.condenser[1] = args["condenser"].as_float(0);
    } else {
        S->nf.condenser[0] = args["condenser"].as_float(0);
        S->nf.condenser[1] = args["condenser"].as_float(1);
    }
    S->nf.extSource[0] = args["points-rings"].as_int(0);
    S->nf.extSource[1] = args["points-rings"].as_int(1);
    for (int i = 0; i < 3; ++i) S->nf.focus[i] = args["focus"].as_float(i);
    bsVector spherical(1, 0, 0);
    if (args("k")) {
        spherical[1] = args["k"].as_float(0);
        spherical[2] = args["k"].as_float(1);
    }
    // sph2cart in float, as the reference's ptype
    const float r = (float)spherical[0], th = (float)spherical[1], ph = (float)spherical[2];
    S->nf.k = bsVector(r * std::cos(th) * std::sin(ph), r * std::sin(th) * std::sin(ph), r * std::cos(ph));
    S->nf.m = args["field-order"].as_int();
    S->nf.nWaves = args["samples"].as_int();
}

void lOutputParams(const rts::arglist& args, fileoutStruct* out) {
    out->append = args("append");
    const std::string f = args["output-type"].as_text();
    if (f == "magnitude") out->field = fileoutStruct::fieldMag;
    else if (f == "intensity") out->field = fileoutStruct::fieldIntensity;
    else if (f == "incident") out->field = fileoutStruct::fieldIncident;
    else if (f == "polarization") out->field = fileoutStruct::fieldPolar;
    else if (f == "imaginary") out->field = fileoutStruct::fieldImag;
    else if (f == "real") out->field = fileoutStruct::fieldReal;
    else if (f == "angular-spectrum") out->field = fileoutStruct::fieldAngularSpectrum;
    out->intFile = args["intensity"].as_text();
    out->incFile = args["incident"].as_text();
    out->absFile = args["absorbance"].as_text();
    out->absSpecFile = args["absorbance-spec"].as_text();
    if (args("transmittance")) out->transFile = args["transmittance"].as_text();
    out->nearFile = args["near-field"].as_text();
    out->farFile = args["far-field"].as_text();
    const std::string c = args["colormap"].as_text();
    if (c == "brewer") out->colormap = stim::cmBrewer;
    else if (c == "gray") out->colormap = stim::cmGrayscale;
    else std::cout << "color-map value not recognized (using default): " << c << std::endl;
}

void lBuild(const rts::arglist& args) {
    const std::string p = args["precision"].as_text();
    if (p == "fp32") settings().precision = BIMSIM_FP32_FAST;
    else if (p == "fp64") settings().precision = BIMSIM_FP64;
    else if (p == "parity") settings().precision = BIMSIM_FP64_PARITY;
    else throw Fail("unknown --precision " + p + " (fp32, fp64, parity)");
    settings().quirks = args["quirks"].as_int() ? BIMSIM_QUIRKS_REF : BIMSIM_QUIRKS_OFF;
    const std::string g = args["mc-rng"].as_text();
    if (g == "glibc") settings().mc_rng = BIMSIM_RNG_GLIBC;
    else if (g == "splitmix64") settings().mc_rng = BIMSIM_RNG_SPLITMIX64;
    else throw Fail("unknown --mc-rng " + g + " (glibc, splitmix64)");
    settings().mc_seed = std::stoull(args["mc-seed"].as_text());
    // calcWaves re-seeds glibc with srand(NULL), so --seed does not reach
    // the directions (nearfield.cpp:102); the glibc stream keeps seed 0
    if (settings().mc_rng == BIMSIM_RNG_GLIBC) settings().mc_seed = 0;
}

void LoadParameters(const rts::arglist& args, microscopeStruct* S, fileoutStruct* out) {
    lBuild(args);
    lFlags(args, S);
    lWavelength(args, S, out);
    lMaterials(args, S);
    lSpheres(args, S);
    lOptics(args, S);
    lImagePlane(args, S);
    lNearfield(args, S);
    lOutputParams(args, out);
    // an extended source, or else the interpolation mode (arguments.h:576-587)
    if (args("extended-source"))
        S->LoadExtendedSource(args["extended-source"].as_text());
    else if (args("interpolate"))
        S->interpolate = args["interpolate"].as_int();
}

void print_config(const microscopeStruct& S, const fileoutStruct& out) {
    std::cout << std::setprecision(17);
    const nearfieldStruct& nf = S.nf;
    std::cout << "lambda=" << nf.lambda << "\n";
    std::cout << "amplitude=" << nf.A << "\n";
    std::cout << "condenser=" << nf.condenser[0] << "," << nf.condenser[1] << "\n";
    std::cout << "objective=" << S.objective[0] << "," << S.objective[1] << "\n";
    std::cout << "focus=" << nf.focus[0] << "," << nf.focus[1] << "," << nf.focus[2] << "\n";
    std::cout << "k=" << nf.k[0] << "," << nf.k[1] << "," << nf.k[2] << "\n";
    std::cout << "m=" << nf.m << "\n";
    std::cout << "samples=" << nf.nWaves << "\n";
    std::cout << "plane_wave=" << nf.planeWave << "\n";
    std::cout << "lut_uf=" << nf.lut_uf << "\n";
    std::cout << "lut_us=" << nf.lut_us << "\n";
    std::cout << "padding=" << S.padding << "\n";
    std::cout << "ss=" << S.ss << "\n";
    std::cout << "detector=" << S.Ud.R[0] / S.ss << "," << S.Ud.R[1] / S.ss << "\n";
    std::cout << "nf_res=" << nf.U.R[0] << "," << nf.U.R[1] << "\n";
    for (int i = 0; i < 3; ++i) {
        const bsVector& v = i == 0 ? nf.pos.A : i == 1 ? nf.pos.X : nf.pos.Y;
        std::cout << "nf_pos_" << "AXY"[i] << "=" << v[0] << "," << v[1] << "," << v[2] << "\n";
    }
    std::cout << "z_d=" << S.z_d << "\n";
    for (size_t i = 0; i < nf.mVector.size(); ++i) {
        const bsComplex n = nf.mVector[i](nf.lambda);
        std::cout << "material" << i << "=" << n.real() << "," << n.imag() << "\n";
    }
    for (size_t i = 0; i < nf.sVector.size(); ++i) {
        const sphere& s = nf.sVector[i];
        std::cout << "sphere" << i << "=" << s.p[0] << "," << s.p[1] << "," << s.p[2] << "," << s.a << ","
                  << s.iMaterial << "\n";
    }
    std::cout << "interpolate=" << S.interpolate << "\n";
    std::cout << "sources=" << S.focalPoints.size() << "\n";
    for (size_t i = 0; i < S.focalPoints.size() && i < 8; ++i) {
        const sourcePoint& p = S.focalPoints[i];
        std::cout << "source" << i << "=" << p.f[0] << "," << p.f[1] << "," << p.f[2] << "," << p.A
                  << "\n";
    }
    std::cout << "files=" << out.intFile << "|" << out.incFile << "|" << out.absFile << "|"
              << out.transFile << "|" << out.nearFile << "|" << out.farFile << "\n";
    std::cout << "wavenumber=" << out.wavenumber << "\n";
    std::cout << "precision=" << settings().precision << "\n";
}

}  // namespace

int main(int argc, char* argv[]) {
    rts::arglist args;
    SetArguments(args);
    args.parse(argc, argv);
    if (args("help")) {
        std::cout << args.toStr() << std::endl;
        return 1;
    }
    int rc = 0;
    try {
        microscopeStruct* SCOPE = new microscopeStruct();
        fileoutStruct out;
        LoadParameters(args, SCOPE, &out);
        if (args("print-config")) {
            print_config(*SCOPE, out);
            delete SCOPE;
            return 0;
        }
        SCOPE->init();
        out.Save(SCOPE);
        if (settings().verbose) std::cout << SCOPE->toStr();
        SCOPE->destroy();
        delete SCOPE;
        bimsim::release_context();
    } catch (const std::exception& e) {
        std::cout << e.what() << std::endl;
        rc = 1;
    }
    return rc;
}
