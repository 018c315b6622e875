// microscopeStruct (reference microscope.h:23-142, microscope.cu:495-857):
// the near field of a padded slice, the objective band-pass, the crop to the
// detector, the detector integration and the derived images.
#pragma once

#include <string>
#include <vector>

#include "fieldslice.h"
#include "nearfield.h"
#include "scalarslice.h"

struct sourcePoint {
    bsPoint f;
    ptype A;
};

struct sourcePointES {
    bsPoint f;
    ptype A;
    ptype rotationAngle;
};

struct microscopeStruct {
    int interpolate = -1;
    int padding = 1;
    int ss = 1;
    float scale = 1.0f;

    fieldslice Ufd;                 // cropped incident field at the detector
    fieldslice Ud;                  // cropped total field at the detector
    bool scalarSim = true;

    scalarslice* D = nullptr;       // detector image of the total field
    scalarslice* D_ps1 = nullptr;
    scalarslice* Di_ps1 = nullptr;
    ptype sumD = 0;
    scalarslice* Di = nullptr;      // detector image of the incident field
    ptype sumDi = 0;

    ptype objective[2] = {0, 1};
    bsRect pos;
    float z_d = 0;

    std::vector<sourcePoint> focalPoints;
    std::vector<sourcePointES> focalPointsES;
    int line_len = 0;               // samples of the radial lines (--interpolate 2)

    // The band-pass, crop and integration of SimulateImaging run as one
    // fused band-limited FFT pass (true) or as the reference's separate
    // steps (false); both leave the same fields and images.
    bool fused_detector = true;

    microscopeStruct();
    ~microscopeStruct();
    microscopeStruct(const microscopeStruct&) = delete;
    microscopeStruct& operator=(const microscopeStruct&) = delete;

    nearfieldStruct nf;

    void setZdPos(float d);

    void init();
    void destroy();

    void SimulateScattering();
    void SimulateInterpolation();
    void SimulateImaging();
    void Simulate();
    void SimulateAndSaveTexture();
    void SimulateWithInterpolation();
    void saveToTexture();

    // Extended sources (microscope.cu:873-914, 1258-1296): one point source
    // per non-zero pixel of a source image, focus at pos(x / Rx, y / Ry) and
    // amplitude value / 255; the detector images sum the sources' images.
    void LoadExtendedSource(std::string filename);
    void SimulateExtendedSource();

    // Extended sources from rings of point sources (microscope.cu:37-420,
    // 1004-1465), selected by --interpolate with --points-rings
    // (points per ring, rings); pointsRings is nf.extSource:
    //   1: each ring's first source is simulated; the others add its detector
    //      images rotated by their angles (SimulateExtendedSourceInterpAtDetector);
    //   2: ring sums of each simulated source's detector images into a radial
    //      line, then the detector images are that line as a function of the
    //      distance from the centre (SimulateExtendedSourceInterpLineAtDetector);
    //   3: each ring's first near field is rotated for the other sources and
    //      imaged (SimulateExtendedSourceWithInterpolation);
    //   0: the Gaussian ring source without interpolation
    //      (SimulateExtendedGausssianSourceNoInterp).
    void generatePointSources(int numPoints, float radius, float amplitude);
    void saveDetectorForOnePs();
    void SaveDetectorToTexture();
    void SaveLineToTexture(int radiusLength);
    void InterpAtDetector();
    void InterpLineAtDetector();
    void InterpLineForAllPointSources(int numSamples, int numRings);
    void SimulateExtendedSourceInterpAtDetector(int* pointsRings);
    void SimulateExtendedSourceInterpLineAtDetector(int* pointsRings);
    void SimulateExtendedSourceWithInterpolation(float sphereRadius, int* pointsRings);
    void SimulateExtendedGausssianSourceNoInterp(int* pointsRings);

    void setNearfield();
    void setRes(int x_res, int y_res, int pad, int supersampling);
    void setPos(bsPoint pMin, bsPoint pMax, bsVector normal);

    void applyBandpass();
    void applyBackpropagation();
    void getFarField();
    void integrateDetector();
    void clearDetector();

    scalarslice getAbsorbance();
    // always 0, as in the reference (microscope.cu:749-762)
    ptype getAbsorbanceSpectrum();
    scalarslice getTransmittance();
    scalarslice getIntensity();
    scalarslice getIncidentFieldImage();
    std::string toStr();
};
