// nearfieldStruct (reference nearfield.h:14-120, nearfield.cpp,
// nfScalarUf.cu, nfScalarUfLut.cu, nfScalarUs.cu, nfScalarUpLut.cu,
// nfSumUf.cu, nfInterpolate.cu): the incident field Uf and the total field
// U = Uf + sum of scattered fields over a slice, computed by the HIP library.
#pragma once

#include <string>
#include <vector>

#include "dataTypes.h"
#include "fieldslice.h"
#include "material.h"
#include "sphere.h"

#define EPSILON_FLOAT 0.000001f

struct nearfieldStruct {
    ptype lambda = 1;
    ptype condenser[2] = {0, 1};
    int extSource[2] = {1, 1};
    ptype A = 1;                 // incident amplitude
    bsVector E;                  // polarisation (vector fields only)
    bsVector k{0, 0, 1};         // incident direction
    bsPoint focus;
    ptype rotationAngle = 0;

    bsRect pos;                  // slice position in world space

    fieldslice Uf;
    ptype d_min = 0, d_max = 0;
    fieldslice U;

    int m = 10;                  // incident-field order
    bool scalarSim = true;
    bool planeWave = false;
    bool lut_uf = false;
    bool lut_us = true;

    float t_Uf = 0;              // milliseconds
    float t_Us = 0;

    unsigned int angRes = 0;
    std::vector<rts::material<ptype>> mVector;
    std::vector<sphere> sVector;

    nearfieldStruct();
    ~nearfieldStruct();
    nearfieldStruct(const nearfieldStruct&) = delete;
    nearfieldStruct& operator=(const nearfieldStruct&) = delete;

    void saveFieldsToTexture();
    void Interpolate();

    void init();
    void destroy();

    // calcWaves, calcUf, calcSpheres, calcUs, sumUf (nearfield.cpp:194-218)
    void Simulate();

    void calcSpheres();

    std::vector<bsVector> inWaves;
    int nWaves = 0;
    void calcWaves();

    std::string toStr();

    void setRes(int x_res, int y_res);
    void setPos(bsPoint pMin, bsPoint pMax, bsVector normal);

    void calcUf();
    void scalarUf();
    void scalarUfLut();
    void vectorUf();
    void vectorUfLut();

    // j[id*(m+1) + l] = j_l(k d), d over [d_min, d_max] in dR samples
    // (nearfield.cpp:222-283); host array
    void calcBesselLut(ptype* j, ptype d_min, ptype d_max, int dR);

    void calcUs();
    void scalarUs();
    void scalarUpLut();

    void sumUf();

    // device copies of the waves and the packed coefficients of all spheres
   private:
    void* d_waves = nullptr;
    size_t d_waves_cap = 0;
    void* d_A = nullptr;
    void* d_B = nullptr;
    size_t d_coef_cap = 0;
    fieldslice savedU, savedUf;   // saveFieldsToTexture copies
    void upload_waves(const double* w, int nk);
    void upload_coefficients();
    void free_device();
    friend struct microscopeStruct;
};
