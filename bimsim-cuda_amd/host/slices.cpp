// Settings, the library context, fieldslice and scalarslice over the C ABI,
// and the ENVI / BMP writers.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <vector>

#include "fieldslice.h"
#include "scalarslice.h"

namespace bimsim {

Settings& settings() {
    static Settings s;
    return s;
}

static bimsim_ctx* g_ctx = nullptr;

bimsim_ctx* context() {
    if (!g_ctx) {
        const int rc = bimsim_ctx_create(settings().device, nullptr, &g_ctx);
        if (rc != BIMSIM_OK) {
            g_ctx = nullptr;
            throw std::runtime_error("bimsim: no usable gfx950 device (bimsim_ctx_create failed)");
        }
    }
    return g_ctx;
}

void release_context() {
    if (g_ctx) bimsim_ctx_destroy(g_ctx);
    g_ctx = nullptr;
}

void check(int rc, const char* what) {
    if (rc == BIMSIM_OK) return;
    std::string msg = std::string("bimsim: ") + what + " failed";
    if (g_ctx) msg += std::string(": ") + bimsim_last_error(g_ctx);
    throw std::runtime_error(msg);
}

static void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("bimsim: ") + what + ": " + hipGetErrorString(e));
}

static void* dev_alloc(size_t bytes) {
    context();
    void* p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc");
    return p;
}

static void dev_free(void* p) {
    if (p) (void)hipFree(p);
}

static void dev_copy(void* dst, const void* src, size_t bytes) {
    hip_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "device copy");
}

static void dev_zero(void* p, size_t bytes) {
    hip_check(hipMemset(p, 0, bytes), "hipMemset");
}

}  // namespace bimsim

using namespace bimsim;

// ----------------------------------------------------------------- scalarslice

scalarslice::scalarslice() : prec(settings().precision) {}

scalarslice::scalarslice(int x, int y) : prec(settings().precision) {
    R[0] = x;
    R[1] = y;
    alloc();
    clear();
}

scalarslice::scalarslice(int x, int y, int psSamples) : scalarslice(x, y) {
    line_len = psSamples;
    line = dev_alloc(scalar_bytes(prec) * (size_t)std::max(psSamples, 1));
    clearLine(psSamples);
}

scalarslice::scalarslice(ptype s) : prec(settings().precision) {
    R[0] = R[1] = 1;
    alloc();
    const float f = (float)s;
    const double d = s;
    hip_check(hipMemcpy(S, prec == BIMSIM_FP32_FAST ? (const void*)&f : (const void*)&d,
                        scalar_bytes(prec), hipMemcpyHostToDevice),
              "upload");
}

scalarslice::scalarslice(const scalarslice& rhs) : rotationAngle(rhs.rotationAngle), prec(rhs.prec) {
    R[0] = rhs.R[0];
    R[1] = rhs.R[1];
    alloc();
    if (S && rhs.S) dev_copy(S, rhs.S, scalar_bytes(prec) * (size_t)R[0] * R[1]);
}

scalarslice& scalarslice::operator=(const scalarslice& rhs) {
    if (this == &rhs) return *this;
    release();
    R[0] = rhs.R[0];
    R[1] = rhs.R[1];
    prec = rhs.prec;
    rotationAngle = rhs.rotationAngle;
    alloc();
    if (S && rhs.S) dev_copy(S, rhs.S, scalar_bytes(prec) * (size_t)R[0] * R[1]);
    return *this;
}

scalarslice::~scalarslice() { release(); }

void scalarslice::alloc() {
    if (R[0] > 0 && R[1] > 0) S = dev_alloc(scalar_bytes(prec) * (size_t)R[0] * R[1]);
}

void scalarslice::release() {
    dev_free(S);
    dev_free(line);
    S = line = nullptr;
}

void scalarslice::init(int x, int y, int precision) {
    release();
    R[0] = x;
    R[1] = y;
    prec = precision;
    alloc();
    clear();
}

void scalarslice::clear() {
    if (S) dev_zero(S, scalar_bytes(prec) * (size_t)R[0] * R[1]);
}

void scalarslice::clearLine(int radiusLength) {
    if (line) dev_zero(line, scalar_bytes(prec) * (size_t)std::min(std::max(radiusLength, 0), line_len));
}

void scalarslice::download(double* out) const {
    const size_t n = (size_t)R[0] * R[1];
    if (!S || n == 0) return;
    check(bimsim_sync(context()), "sync");
    if (prec == BIMSIM_FP32_FAST) {
        std::vector<float> h(n);
        hip_check(hipMemcpy(h.data(), S, 4 * n, hipMemcpyDeviceToHost), "download");
        for (size_t i = 0; i < n; ++i) out[i] = h[i];
    } else {
        hip_check(hipMemcpy(out, S, 8 * n, hipMemcpyDeviceToHost), "download");
    }
}

// --------------------------------------------------------------- image output

namespace {

// ColorBrewer 11-class RdYlBu, reversed (blue = low, red = high).
const unsigned char kBrewer[11][3] = {
    {49, 54, 149},  {69, 117, 180}, {116, 173, 209}, {171, 217, 233}, {224, 243, 248}, {255, 255, 191},
    {254, 224, 144}, {253, 174, 97}, {244, 109, 67},  {215, 48, 39},   {165, 0, 38}};

void colour(double t, stim::colormapType cmap, unsigned char* rgb) {
    if (!(t >= 0)) t = 0;   // NaN -> low end
    if (t > 1) t = 1;
    if (cmap == stim::cmGrayscale) {
        rgb[0] = rgb[1] = rgb[2] = (unsigned char)std::lround(t * 255.0);
        return;
    }
    const double x = t * 10.0;
    const int i = std::min(9, (int)x);
    const double w = x - i;
    for (int c = 0; c < 3; ++c)
        rgb[c] = (unsigned char)std::lround((1 - w) * kBrewer[i][c] + w * kBrewer[i + 1][c]);
}

void write_bmp(const std::string& filename, const std::vector<unsigned char>& rgb, int w, int h) {
    const int row = (3 * w + 3) & ~3;
    const unsigned size = 54 + (unsigned)row * h;
    unsigned char hdr[54] = {'B', 'M'};
    auto put32 = [&](int off, unsigned v) {
        for (int i = 0; i < 4; ++i) hdr[off + i] = (unsigned char)(v >> (8 * i));
    };
    put32(2, size);
    put32(10, 54);
    put32(14, 40);
    put32(18, (unsigned)w);
    put32(22, (unsigned)h);
    hdr[26] = 1;
    hdr[28] = 24;
    put32(34, (unsigned)row * h);
    std::ofstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("bimsim: cannot write image " + filename);
    f.write((const char*)hdr, 54);
    std::vector<unsigned char> line(row, 0);
    // image row iv = 0 at the top: BMP stores rows bottom-up
    for (int y = h - 1; y >= 0; --y) {
        for (int x = 0; x < w; ++x) {
            const unsigned char* p = &rgb[3 * ((size_t)y * w + x)];
            line[3 * x + 0] = p[2];
            line[3 * x + 1] = p[1];
            line[3 * x + 2] = p[0];
        }
        f.write((const char*)line.data(), row);
    }
}

}  // namespace

std::vector<unsigned char> read_bmp_channel0(const std::string& filename, int* w, int* h) {
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("bimsim: cannot read source image " + filename);
    std::vector<unsigned char> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    auto u16 = [&](size_t o) { return (unsigned)data[o] | ((unsigned)data[o + 1] << 8); };
    auto u32 = [&](size_t o) { return u16(o) | (u16(o + 2) << 16); };
    if (data.size() < 54 || data[0] != 'B' || data[1] != 'M')
        throw std::runtime_error("bimsim: " + filename + " is not a BMP image (only BMP sources are read)");
    const unsigned off = u32(10), hsize = u32(14);
    const int width = (int)u32(18), height_raw = (int)u32(22);
    const unsigned bpp = u16(28), comp = u32(30);
    const bool top_down = height_raw < 0;
    const int height = top_down ? -height_raw : height_raw;
    if (width <= 0 || height <= 0 || width > (1 << 16) || height > (1 << 16))
        throw std::runtime_error("bimsim: bad BMP size in " + filename);
    if (!((bpp == 8 || bpp == 24) && comp == 0) && !(bpp == 32 && (comp == 0 || comp == 3)))
        throw std::runtime_error("bimsim: unsupported BMP format in " + filename +
                                 " (uncompressed 8-, 24- or 32-bit only)");
    const size_t row = ((size_t)bpp * width + 31) / 32 * 4;
    if ((size_t)off + row * height > data.size())
        throw std::runtime_error("bimsim: truncated BMP " + filename);
    std::vector<unsigned char> pal;
    if (bpp == 8) {
        unsigned n = u32(46);
        if (n == 0) n = 256;
        const size_t p0 = 14 + (size_t)hsize;
        if (p0 + 4 * (size_t)n > off) throw std::runtime_error("bimsim: bad BMP palette in " + filename);
        for (unsigned i = 0; i < n; ++i) pal.push_back(data[p0 + 4 * i]);   // blue
        pal.resize(256, 0);
    }
    std::vector<unsigned char> out((size_t)width * height);
    for (int y = 0; y < height; ++y) {
        const size_t src = off + row * (size_t)(top_down ? y : height - 1 - y);
        for (int x = 0; x < width; ++x) {
            unsigned char v;
            if (bpp == 8) v = pal[data[src + x]];
            else v = data[src + (size_t)x * (bpp / 8)];
            out[(size_t)y * width + x] = v;
        }
    }
    *w = width;
    *h = height;
    return out;
}

void scalarslice::toImage(std::string filename, ptype vmin, ptype vmax, stim::colormapType cmap) {
    const size_t n = (size_t)R[0] * R[1];
    std::vector<double> v(n);
    download(v.data());
    std::vector<unsigned char> rgb(3 * n);
    const double span = vmax - vmin;
    for (size_t i = 0; i < n; ++i) colour(span != 0 ? (v[i] - vmin) / span : 0.0, cmap, &rgb[3 * i]);
    write_bmp(filename, rgb, R[0], R[1]);
}

void scalarslice::toImage(std::string filename, bool positive, stim::colormapType cmap) {
    const size_t n = (size_t)R[0] * R[1];
    std::vector<double> v(n);
    download(v.data());
    // first index of the largest magnitude (cublasIsamax), then its value
    size_t im = 0;
    for (size_t i = 1; i < n; ++i)
        if (std::fabs(v[i]) > std::fabs(v[im])) im = i;
    const double maxVal = n ? v[im] : 0.0;
    if (positive) toImage(filename, 0, maxVal, cmap);
    else toImage(filename, -std::fabs(maxVal), std::fabs(maxVal), cmap);
}

void scalarslice::toEnvi(std::string filename, ptype wavelength, bool append) {
    const size_t n = (size_t)R[0] * R[1];
    std::vector<double> v(n);
    download(v.data());
    std::vector<float> f(v.begin(), v.end());
    envi_add_band(filename, f.data(), R[0], R[1], wavelength, append);
}

namespace {

struct EnviHeader {
    unsigned samples = 0, lines = 0, bands = 0;
    std::vector<double> wavelength;
};

std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

bool load_header(const std::string& name, EnviHeader* h) {
    std::ifstream f(name);
    if (!f) return false;
    std::stringstream all;
    all << f.rdbuf();
    const std::string text = all.str();
    auto value = [&](const std::string& key, std::string* out) {
        const size_t k = text.find("\n" + key);
        if (k == std::string::npos) return false;
        const size_t eq = text.find('=', k);
        const size_t nl = text.find('\n', eq);
        *out = trim(text.substr(eq + 1, nl - eq - 1));
        return true;
    };
    std::string s;
    if (value("samples", &s)) h->samples = (unsigned)std::stoul(s);
    if (value("lines", &s)) h->lines = (unsigned)std::stoul(s);
    if (value("bands", &s)) h->bands = (unsigned)std::stoul(s);
    const size_t w = text.find("\nwavelength = {");
    if (w != std::string::npos) {
        const size_t a = text.find('{', w), b = text.find('}', a);
        std::stringstream ws(text.substr(a + 1, b - a - 1));
        std::string item;
        while (std::getline(ws, item, ','))
            if (!trim(item).empty()) h->wavelength.push_back(std::stod(trim(item)));
    }
    return true;
}

void save_header(const std::string& name, const EnviHeader& h) {
    std::ofstream o(name);
    o << "ENVI" << std::endl;
    o << "description = {" << std::endl;
    o << "  " << "}" << std::endl;
    o << "samples = " << h.samples << std::endl;
    o << "lines = " << h.lines << std::endl;
    o << "bands = " << h.bands << std::endl;
    o << "header offset = " << 0 << std::endl;
    o << "file type = " << "ENVI Standard" << std::endl;
    o << "data type = " << 4 << std::endl;
    o << "interleave = bsq" << std::endl;
    o << "sensor type = " << "Unknown" << std::endl;
    o << "byte order = " << 0 << std::endl;
    o << "x start = " << 0 << std::endl;
    o << "y start = " << 0 << std::endl;
    o << "wavelength units = " << "Unknown" << std::endl;
    o << "z plot titles = {Unknown, Unknown}" << std::endl;
    o << "pixel size = {1, 1, units=Meters}" << std::endl;
    o << "wavelength = {" << std::endl;
    for (size_t i = 0; i + 1 < h.wavelength.size(); ++i) o << h.wavelength[i] << ", ";
    if (!h.wavelength.empty()) o << h.wavelength.back();
    o << "}" << std::endl;
}

}  // namespace

void envi_add_band(const std::string& filename, const float* band, int samples, int lines,
                   double wavelength, bool append) {
    EnviHeader h;
    if (append) load_header(filename + ".hdr", &h);
    FILE* f = std::fopen(filename.c_str(), append ? "ab" : "wb");
    if (!f) throw std::runtime_error("bimsim: unable to open binary file " + filename);
    const size_t n = (size_t)samples * lines;
    const size_t wrote = std::fwrite(band, sizeof(float), n, f);
    std::fclose(f);
    if (wrote != n) throw std::runtime_error("bimsim: short write to " + filename);
    h.bands += 1;
    h.samples = samples;
    h.lines = lines;
    h.wavelength.push_back(wavelength);
    save_header(filename + ".hdr", h);
}

// ----------------------------------------------------------------- fieldslice

fieldslice::fieldslice() : prec(settings().precision) {}

fieldslice::fieldslice(unsigned int x_res, unsigned int y_res) : prec(settings().precision) {
    R[0] = x_res;
    R[1] = y_res;
    init_gpu();
}

fieldslice::fieldslice(const fieldslice& rhs) : scalarField(rhs.scalarField), prec(rhs.prec) {
    R[0] = rhs.R[0];
    R[1] = rhs.R[1];
    const size_t bytes = field_bytes(prec) * samples();
    void* const src[3] = {rhs.x_hat, rhs.y_hat, rhs.z_hat};
    void** dst[3] = {&x_hat, &y_hat, &z_hat};
    for (int c = 0; c < 3; ++c)
        if (src[c]) {
            *dst[c] = dev_alloc(bytes);
            dev_copy(*dst[c], src[c], bytes);
        }
}

fieldslice& fieldslice::operator=(const fieldslice& rhs) {
    if (this == &rhs) return *this;
    R[0] = rhs.R[0];
    R[1] = rhs.R[1];
    scalarField = rhs.scalarField;
    prec = rhs.prec;
    init_gpu();
    const size_t bytes = field_bytes(prec) * samples();
    if (x_hat && rhs.x_hat) dev_copy(x_hat, rhs.x_hat, bytes);
    if (y_hat && rhs.y_hat) dev_copy(y_hat, rhs.y_hat, bytes);
    if (z_hat && rhs.z_hat) dev_copy(z_hat, rhs.z_hat, bytes);
    return *this;
}

fieldslice::~fieldslice() { kill_gpu(); }

void fieldslice::init_gpu() {
    if (R[0] == 0 || R[1] == 0) return;
    kill_gpu();
    const size_t bytes = field_bytes(prec) * samples();
    x_hat = dev_alloc(bytes);
    if (!scalarField) {
        y_hat = dev_alloc(bytes);
        z_hat = dev_alloc(bytes);
    }
    clear_gpu();
}

void fieldslice::kill_gpu() {
    dev_free(x_hat);
    dev_free(y_hat);
    dev_free(z_hat);
    x_hat = y_hat = z_hat = nullptr;
}

void fieldslice::clear_gpu() {
    const size_t bytes = field_bytes(prec) * samples();
    if (x_hat) dev_zero(x_hat, bytes);
    if (y_hat) dev_zero(y_hat, bytes);
    if (z_hat) dev_zero(z_hat, bytes);
}

void fieldslice::download(double* out) const {
    const size_t n = samples();
    if (!x_hat || n == 0) return;
    check(bimsim_sync(context()), "sync");
    if (prec == BIMSIM_FP32_FAST) {
        std::vector<float> h(2 * n);
        hip_check(hipMemcpy(h.data(), x_hat, 8 * n, hipMemcpyDeviceToHost), "download");
        for (size_t i = 0; i < 2 * n; ++i) out[i] = h[i];
    } else {
        hip_check(hipMemcpy(out, x_hat, 16 * n, hipMemcpyDeviceToHost), "download");
    }
}

// Vector fields (scalarField false) carry x_hat, y_hat and z_hat; the
// intensity and magnitude views and the detector sums use all three
// (field_intensity / resample_intensity, fieldslice.cu:6-108).
static scalarslice field_view(fieldslice& f, int op) {
    scalarslice s;
    s.init((int)f.R[0], (int)f.R[1], f.prec);
    check(bimsim_field3_to_scalar(context(), f.prec, op, f.x_hat, f.y_hat, f.z_hat,
                                  (long long)f.samples(), s.S),
          "field view");
    return s;
}

scalarslice fieldslice::Mag() { return field_view(*this, BIMSIM_SLICE_MAGNITUDE); }
scalarslice fieldslice::Real() { return field_view(*this, BIMSIM_SLICE_REAL); }
scalarslice fieldslice::Imag() { return field_view(*this, BIMSIM_SLICE_IMAG); }
scalarslice fieldslice::Intensity() { return field_view(*this, BIMSIM_SLICE_INTENSITY); }

void fieldslice::IntegrateAndResample(scalarslice* detector, unsigned int supersample, float scale) {
    check(bimsim_detector_resample3(context(), prec, x_hat, y_hat, z_hat, detector->R[0],
                                    detector->R[1], (int)supersample, scale, 1, detector->S),
          "IntegrateAndResample");
}

void fieldslice::ResampleAndSave(scalarslice* detector, unsigned int supersample) {
    check(bimsim_detector_resample3(context(), prec, x_hat, y_hat, z_hat, detector->R[0],
                                    detector->R[1], (int)supersample, 1.0, 0, detector->S),
          "ResampleAndSave");
}

// field_scale scales every component (fieldslice.cu:141-153)
void fieldslice::ScaleField(ptype v) {
    for (void* c : {x_hat, y_hat, z_hat})
        if (c) check(bimsim_field_scale(context(), prec, c, (long long)samples(), v), "ScaleField");
}

void fieldslice::clear_yz() {
    const size_t bytes = field_bytes(prec) * samples();
    if (y_hat) dev_zero(y_hat, bytes);
    if (z_hat) dev_zero(z_hat, bytes);
}

void fieldslice::toAngularSpectrum() {
    check(bimsim_field_fft(context(), prec, x_hat, (int)R[0], (int)R[1], 0), "toAngularSpectrum");
}

void fieldslice::fromAngularSpectrum() {
    check(bimsim_field_fft(context(), prec, x_hat, (int)R[0], (int)R[1], 1), "fromAngularSpectrum");
}

fieldslice fieldslice::crop(int u, int v, int su, int sv) {
    fieldslice r;
    r.prec = prec;
    r.scalarField = scalarField;
    r.R[0] = su;
    r.R[1] = sv;
    r.init_gpu();
    // every component (the reference's crop result holds x_hat only, then
    // copies y_hat and z_hat of a vector field into it: fieldslice.cu:345-359)
    const void* src[3] = {x_hat, y_hat, z_hat};
    void* dst[3] = {r.x_hat, r.y_hat, r.z_hat};
    for (int c = 0; c < 3; ++c)
        if (src[c] && dst[c])
            check(bimsim_field_crop(context(), prec, src[c], (int)R[0], (int)R[1], u, v, su, sv,
                                    dst[c]),
                  "crop");
    return r;
}

std::string fieldslice::toStr() {
    std::stringstream ss;
    ss << "Resolution: " << R[0] << " x " << R[1] << (scalarField ? " (scalar)" : " (vector)");
    return ss.str();
}
