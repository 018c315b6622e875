// fileoutStruct (reference fileout.cu:3-271).
#include "fileout.h"

#include <fstream>
#include <iostream>

bool fileoutStruct::is_binary(const std::string& filename) {
    const size_t i = filename.find_last_of('.');
    if (i == std::string::npos) return true;
    const std::string ext = filename.substr(i + 1);
    return !(ext == "bmp" || ext == "jpg" || ext == "tif" || ext == "gif" || ext == "png");
}

void fileoutStruct::saveNearField(nearfieldStruct* nf) {
    if (nearFile.empty()) return;
    if (field == fieldReal) nf->U.Real().toImage(nearFile, false, colormap);
    if (field == fieldImag) nf->U.Imag().toImage(nearFile, false, colormap);
    if (field == fieldMag) nf->U.Mag().toImage(nearFile, true, colormap);
}

void fileoutStruct::saveFarField(microscopeStruct* scope) {
    if (farFile.empty()) return;
    if (field == fieldReal) scope->Ud.Real().toImage(farFile, false, colormap);
    if (field == fieldImag) scope->Ud.Imag().toImage(farFile, false, colormap);
    if (field == fieldMag) scope->Ud.Mag().toImage(farFile, true, colormap);
}

static void save_slice(scalarslice I, const std::string& file, const microscopeStruct* scope,
                       bool wavenumber, bool append) {
    if (fileoutStruct::is_binary(file))
        I.toEnvi(file, wavenumber ? 10000.0f / scope->nf.lambda : scope->nf.lambda, append);
    else
        I.toImage(file);
}

void fileoutStruct::saveDetector(microscopeStruct* scope) {
    if (!intFile.empty()) save_slice(scope->getIntensity(), intFile, scope, wavenumber, append);
    if (!incFile.empty()) save_slice(scope->getIncidentFieldImage(), incFile, scope, wavenumber, append);
    if (!absFile.empty()) save_slice(scope->getAbsorbance(), absFile, scope, wavenumber, append);
    if (!absSpecFile.empty() && is_binary(absSpecFile)) {
        const ptype a = scope->getAbsorbanceSpectrum();
        std::ofstream f(absSpecFile, std::ios::app);
        f << (wavenumber ? 10000.0f / scope->nf.lambda : scope->nf.lambda) << "," << a << std::endl;
    }
    if (!transFile.empty()) save_slice(scope->getTransmittance(), transFile, scope, wavenumber, append);
}

void fileoutStruct::Save(microscopeStruct* scope) {
    if (!scope->focalPoints.empty()) {
        // extended source: the summed detector images only (fileout.cu:199-214)
        scope->SimulateExtendedSource();
        saveDetector(scope);
        return;
    }
    // rings of point sources (fileout.cu:215-241)
    switch (scope->interpolate) {
        case 3: scope->SimulateExtendedSourceWithInterpolation(0, scope->nf.extSource); break;
        case 1: scope->SimulateExtendedSourceInterpAtDetector(scope->nf.extSource); break;
        case 0: scope->SimulateExtendedGausssianSourceNoInterp(scope->nf.extSource); break;
        case 2: scope->SimulateExtendedSourceInterpLineAtDetector(scope->nf.extSource); break;
        default: break;
    }
    if (scope->interpolate >= 0 && scope->interpolate <= 3) {
        saveDetector(scope);
        return;
    }
    scope->Simulate();
    saveNearField(&scope->nf);
    saveFarField(scope);
    saveDetector(scope);
}
