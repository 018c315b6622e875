// fieldslice (reference fieldslice.h:13-68, fieldslice.cpp, fieldslice.cu):
// an owning device field of R[0] x R[1] complex samples, row-major with u
// fastest (index iv*R[0] + iu), complex64 for the FP32 form and complex128
// for the FP64 forms.  Scalar fields use x_hat only; vector fields
// (scalarField false) hold x_hat, y_hat and z_hat.
#pragma once

#include <string>

#include "dataTypes.h"
#include "scalarslice.h"

struct fieldslice {
    void* x_hat = nullptr;
    void* y_hat = nullptr;
    void* z_hat = nullptr;
    unsigned int R[2] = {0, 0};
    bool scalarField = true;
    int prec;   // element type: bimsim_precision (taken from the settings)

    fieldslice();
    // The reference sets R[1] = x here (fieldslice.cpp:16); this one keeps y.
    fieldslice(unsigned int x_res, unsigned int y_res);
    fieldslice(const fieldslice& rhs);
    fieldslice& operator=(const fieldslice& rhs);
    ~fieldslice();

    scalarslice Mag();
    scalarslice Real();
    scalarslice Imag();
    scalarslice Intensity();
    // detector->S += scale * |field|^2 averaged over supersample^2 samples
    void IntegrateAndResample(scalarslice* detector, unsigned int supersample, float scale);
    // detector->S = the same average
    void ResampleAndSave(scalarslice* detector, unsigned int supersample);

    void ScaleField(ptype v);
    // zero y_hat and z_hat (vector fields)
    void clear_yz();

    void toAngularSpectrum();
    void fromAngularSpectrum();

    fieldslice crop(int u, int v, int su, int sv);

    void init_gpu();
    void kill_gpu();
    void clear_gpu();

    // Host copy of x_hat as interleaved (re, im) doubles.
    void download(double* out) const;
    size_t samples() const { return (size_t)R[0] * R[1]; }

    std::string toStr();
};
