// fileoutStruct (reference fileout.h:13-55, fileout.cu): which images the
// command line writes and in which format (ENVI BSQ float32 for names
// without an image extension, BMP otherwise).
#pragma once

#include <string>

#include "microscope.h"

struct fileoutStruct {
    std::string nearFile;
    std::string farFile;
    std::string intFile;
    std::string incFile;
    std::string absFile;
    std::string absSpecFile;
    std::string transFile;

    bool append = false;

    enum field_type {
        fieldMag,
        fieldIntensity,
        fieldIncident,
        fieldAbsorbance,
        fieldAbsorbanceSpectrum,
        fieldPolar,
        fieldImag,
        fieldReal,
        fieldAngularSpectrum
    };
    enum image_source { imageNearField, imageObjective, imageDetector, imageExtendedSource };

    field_type field = fieldMag;
    bool wavenumber = false;
    stim::colormapType colormap = stim::cmBrewer;
    ptype colorMax = 0;

    // Runs the simulation the options ask for and writes its outputs
    // (fileout.cu:180-271).  Only the single-source simulation is built.
    void Save(microscopeStruct* scope);

    static bool is_binary(const std::string& filename);
    void saveNearField(nearfieldStruct* nf);
    void saveFarField(microscopeStruct* scope);
    void saveDetector(microscopeStruct* scope);
};
