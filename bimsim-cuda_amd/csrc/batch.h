// Device tables and launchers of the batched pipeline (many instances of
// slice, spheres, wavelength and waves in one pass).
#pragma once

#include <algorithm>
#include <cmath>

#include <vector>

#include "internal.h"

namespace bimsim {

// Incident field of one instance.
struct UfInst {
    QuadArg q;
    d3 f, k;
    double kmag, amp, ca, cb;
    int m, uR, vR;
    int mode;        // bimsim_uf_mode: series, plane wave or Monte-Carlo sum
    long long off;   // element offset of the instance's field
    long long tab;   // complex64 fields: offset of its uf_fast_table, else -1
    int flat;        // k.r = 0 at every pixel (uf_flat)
    int nk;          // Monte-Carlo sum: waves [wave_off, wave_off + nk)
    long long wave_off;
    double sA;       // Monte-Carlo sum: weight (1.0f/nk) subA of every wave
    // Mirror symmetry of the series field about the focus (uf_mirror): with
    // ncx > 0, columns iu and mx - iu (rows iv and my - iv) hold the same
    // values, and the ncx x ncy canonical pixels cover the slice: column
    // index c < nlx is column c, else column hx + c - nlx (rows alike).
    // diag: pixel (iu, iv) and the transposed (iv, iu) hold the same values
    // too (ncx = ncy), and the canonical pixels with cx <= cy cover the slice.
    int ncx, ncy, mx, my, nlx, nly, hx, hy, diag;
};

// The mirror axis of one slice direction: positions r(i) = (a + x i/n) - f
// (as quad_at and sub form them), i in [0, n).  Returns m with
// r(m - i) == -r(i) for every pair in range and at least one pair, or -1.
inline int uf_mirror_axis(double a, double x, double f, int n) {
    if (!(x != 0.0) || n < 2) return -1;
    const double c = 2.0 * (f - a) / x * (double)n;
    if (!(c > 0.5 && c < 2.0 * n - 2.5)) return -1;
    const long long m = (long long)std::llround(c);
    auto r = [&](long long i) { return (a + x * ((double)i / (double)n)) - f; };
    for (long long i = 0; i < n; ++i) {
        const long long j = m - i;
        if (j < 0 || j >= n || j == i) continue;
        if (!(r(j) == -r(i))) return -1;
    }
    return (int)m;
}

// Fills I's mirror fields when its field is the series form, the slice is
// normal to z with edges along x and y, k is along z and the pixel
// positions mirror exactly about the focus on both axes (the value then
// depends on the pixel only through r.x^2 + r.y^2 and r.z, so mirrored
// pixels get the same bits).
inline void uf_mirror(UfInst& I) {
    I.ncx = I.ncy = 0;
    I.diag = 0;
    const QuadArg& q = I.q;
    if (I.mode != BIMSIM_UF_SERIES || q.X.y != 0.0 || q.X.z != 0.0 || q.Y.x != 0.0 ||
        q.Y.z != 0.0 || I.k.x != 0.0 || I.k.y != 0.0)
        return;
    const int mx = uf_mirror_axis(q.A.x, q.X.x, I.f.x, I.uR);
    const int my = uf_mirror_axis(q.A.y, q.Y.y, I.f.y, I.vR);
    if (mx < 0 || my < 0) return;
    auto axis = [](int m, int n, int& nl, int& h) {
        nl = std::max(0, m - n + 1);   // [0, m - n]: partners past n - 1
        h = (m + 1) / 2;               // [h, n): the larger of each pair, or unpaired
        return nl + (n - h);
    };
    I.mx = mx;
    I.my = my;
    I.ncx = axis(mx, I.uR, I.nlx, I.hx);
    I.ncy = axis(my, I.vR, I.nly, I.hy);
    // the transposed pixel's offsets from the focus are this one's swapped,
    // formed from the same numbers: the value, a function of r.x^2 + r.y^2
    // and r.z, keeps its bits
    I.diag = I.uR == I.vR && mx == my && q.A.x == q.A.y && q.X.x == q.Y.y && I.f.x == I.f.y ? 1 : 0;
}

// Work items of instance I's incident field: its pixels, or with mirror
// symmetry its canonical pixels, or with the diagonal symmetry too 64 per
// 8 x 8 tile of canonical pixels on and below the diagonal (so that the
// transposed pixels a wave stores lie in eight runs of eight, as its own
// do).
__host__ __device__ inline int uf_work_items(const UfInst& I) {
    if (I.ncx <= 0) return I.uR * I.vR;
    if (!I.diag) return I.ncx * I.ncy;
    const int t = (I.ncx + 7) / 8;
    return 64 * (t * (t + 1) / 2);
}

// The incident field's plane case: k along z and a slice normal to z
// through the focus, so that k.(p - f) = 0 exactly at every pixel.
inline bool uf_flat(const QuadArg& q, const d3& f, const d3& k) {
    return k.x == 0.0 && k.y == 0.0 && (k.z == 1.0 || k.z == -1.0) && q.X.z == 0.0 &&
           q.Y.z == 0.0 && q.A.z == f.z;
}

// Uf + Us sum of one instance, and the clearing of its detector images.
struct SumInst {
    QuadArg q;
    int uR, vR, s_begin, ns;
    long long off;
    long long image_off;   // its Nd x Nd images in D and Di
    int nd2;               // Nd * Nd (<= uR * vR) where the pass clears the images, else 0
};

// Mie coefficients of n spheres, sphere i at wavenumber kmag[i]; writes
// A/B at each sphere's coeff_off.  work: n * 8 * (lmax + 2) complex doubles.
int mie_batch_launch(bimsim_ctx* ctx, const SphereDev* d_sph, const double* d_kmag, int n,
                     int lmax, int quirks, void* work, double2* A, double2* B);

// Pixel-invariant table of the incident-field series for the complex64
// form: for l = 0..m the four values (i^l sign) * D_l, (2l-1)/l, (l-1)/l
// and 2l-1 (the ratios unused at l = 0); D_l from the same FP64 Legendre
// recurrences as the per-pixel reference form, so the values are
// identical.  flat (uf_flat: k.r = 0 at every pixel): the first value of
// each even order l >= 2 is D_l times P_l(0) instead (P_l(0) =
// -((l-1)/l) P_{l-2}(0), in that order), the whole weight of its term.
// Appends 4 (m + 1) doubles.
void uf_fast_table(double ca, double cb, int m, std::vector<double>* out, bool flat = false);

// Incident fields of n instances into Uf (complex64 when f32, else
// complex128); tabs holds the instances' uf_fast_table rows (f32 only).
// Monte-Carlo instances read their waves from `waves` (complex128 only).
// With sum (one SumInst per instance), the launch also clears the
// instances' images in D and Di (float when f32, else double) where its
// nd2 asks.  items: at least every instance's uf_work_items and nd2 (the
// launch's pixel blocks per instance; it strides over more).
struct SumInst;
int uf_batch_launch(bimsim_ctx* ctx, const UfInst* d_inst, int n, int items, int f32,
                    const double* tabs, const double* waves, void* Uf,
                    const SumInst* sum = nullptr, void* D = nullptr, void* Di = nullptr);

// Incident fields of the Monte-Carlo instances list[0..n) of d_inst into the
// complex64 field Uf, in tiles of 64 x 64 pixels (the per-pixel batch
// kernels leave those instances alone in complex64 fields).
int uf_mc_batch_launch(bimsim_ctx* ctx, const UfInst* d_inst, const int* list, int n,
                       int max_tiles, const double* waves, float2* Uf);

// Workgroups of uf_mc_batch_launch for a uR x vR field.
int mc_tiles(int uR, int vR);

// U = Uf or U += Uf outside the spheres, for n instances; d_sph holds
// (x, y, z, a) per sphere.  Also zeroes each instance's images in D and Di
// (float when f32, else double), which the detector then accumulates into.
int sum_batch_launch(bimsim_ctx* ctx, const SumInst* d_inst, int n, int max_npix,
                     const double4* d_sph, int f32, const void* Uf, void* U, void* D, void* Di);

// A batch instance's Uf + Us sum for the detector's first row pass to
// apply as it loads the fields (instead of sum_batch_launch): its slice and
// its ns spheres as (x, y, z, a) at sph.
struct SumFuse {
    QuadArg q;
    int uR, vR, ns;
    const double4* sph;
    // pixels inside a sphere as bits (pixel p: bit p % 32 of word p / 32),
    // formed once by sum_mask_launch from the same test; null: the test runs
    // per pixel as the rows load
    const unsigned* inside = nullptr;
};

// fz's inside bits into mask (ceil(uR vR / 32) words), on ctx->stream.
int sum_mask_launch(bimsim_ctx* ctx, const SumFuse& fz, unsigned* mask);

// bimsim_detector's work; with fz (complex64 fields on the band-limited
// path, detector_fusable), U holds the scattered field alone and the sum
// with Uf is formed as the rows load.  The images are accumulated into as
// always, so they must be cleared before.
bool detector_fusable(const bimsim_ctx* ctx, int uR, int vR);
int detector_run(bimsim_ctx* ctx, int prec, int quirks, void* U, void* Uf, int uR, int vR,
                 double du, double dv, double NAin, double NAout, double lambda, int padding,
                 int Nd, int ss, void* D, void* Di, void* Abs, int write_filtered,
                 const SumFuse* fz);

// The band-limited detector over many instances of one FFT size N (uR = vR
// = N) in three launches per pass.  Instances qualify when det_batchable:
// ss = 1, no zero fill, the N-point row pair fits LDS.
struct DetParams {
    long long uoff;        // element offset of U and Uf
    long long image_off;   // its D, Di, Abs images
    int padding, Nd;
    double du, dv, NAin, NAout, lambda;
    SumFuse fz;            // Uf + Us formed as the rows load (fused batches)
};
struct DetBatch {
    int N = 0, n = 0, B = 0, B2 = 0, Bc = 0, max_nc = 0, max_nd = 0;
    bool f32 = false, fused = false;
    void* d_inst = nullptr;
    int* d_cols = nullptr;
    void* d_G = nullptr;
    const void* tw = nullptr;
};
bool det_batchable(const bimsim_ctx* ctx, int prec, int N, int ss, int zero_fill);
int det_batch_create(bimsim_ctx* ctx, int prec, int quirks, int N, const std::vector<DetParams>& p,
                     bool fused, DetBatch* out);
int det_batch_launch(bimsim_ctx* ctx, const DetBatch& b, const void* U, const void* Uf, void* D,
                     void* Di, void* Abs);
void det_batch_destroy(DetBatch* b);

// bimsim_detector_zero_filled's work (the images accumulate, like
// detector_run's).
int detector_zero_filled_run(bimsim_ctx* ctx, int prec, int quirks, const void* U, const void* Uf,
                             int uR, int vR, int fill, double du, double dv, double NAin,
                             double NAout, double lambda, void* D, void* Di, void* Abs);

// Zeroes `bytes` (a multiple of 4) at p with a kernel (graph-safe).
int zero_launch(bimsim_ctx* ctx, void* p, size_t bytes);

}  // namespace bimsim
