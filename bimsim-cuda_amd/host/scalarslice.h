// scalarslice (reference scalarslice.h:7-38, scalarslice.cu): an owning
// device image of R[0] x R[1] real samples (row-major, u fastest), float for
// the FP32 form and double for the FP64 forms.
#pragma once

#include <string>
#include <vector>

#include "dataTypes.h"

namespace stim {
// Colour maps of the image writer (the reference takes them from STIM's
// colormap.h; only the two the CLI selects are provided).
enum colormapType { cmBrewer, cmGrayscale };
}  // namespace stim

struct scalarslice {
    void* S = nullptr;        // device samples
    void* line = nullptr;     // radial buffer of the extended-source drivers
    int R[2] = {0, 0};
    ptype rotationAngle = 0;
    int prec;                 // element type: bimsim_precision

    scalarslice();
    scalarslice(int x, int y);
    scalarslice(int x, int y, int psSamples);
    explicit scalarslice(ptype s);
    scalarslice(const scalarslice& rhs);
    scalarslice& operator=(const scalarslice& rhs);
    ~scalarslice();

    void clear();
    void clearLine(int radiusLength);
    // Re-allocate as x by y samples of precision prec, cleared.
    void init(int x, int y, int precision);

    // Host copy of the samples as double (row-major).
    void download(double* out) const;

    // 24-bit BMP through a colour map over [vmin, vmax]
    // (scalarslice.cu:89-144); positive = true maps [0, max|S|], otherwise
    // [-max|S|, max|S|].
    void toImage(std::string filename, ptype vmin, ptype vmax,
                 stim::colormapType cmap = stim::cmBrewer);
    void toImage(std::string filename, bool positive = true,
                 stim::colormapType cmap = stim::cmBrewer);
    // One float32 band of an ENVI BSQ file (scalarslice.cu:146-166).
    void toEnvi(std::string filename, ptype wavelength = 0, bool append = false);

   private:
    int line_len = 0;
    void alloc();
    void release();
};

// ENVI BSQ float32 writer with the reference header layout
// (rts/envi/envi.h:196-343, envi_header.h:48-70,318-364).
void envi_add_band(const std::string& filename, const float* band, int samples, int lines,
                   double wavelength, bool append);

// Source images of the extended-source driver: an uncompressed BMP (8-bit
// palette, 24- or 32-bit) read as the first stored channel of each pixel
// (blue for colour files, the grey value for grey ones; the reference reads
// channel 0 of its image class), row 0 at the top, *w x *h bytes.
std::vector<unsigned char> read_bmp_channel0(const std::string& filename, int* w, int* h);
