// Look-up-table paths and the rotation resample:
//   incident-field Bessel LUT  nearfieldStruct::calcBesselLut (nearfield.cpp:222-283)
//   Uf from the LUT            gpuScalarUfLut (nfScalarUfLut.cu:11-123)
//   sphere radial LUTs         sphere::calcHankelLut / calcBesselLut (sphere.cpp:99-203)
//   angular tables Usp / Uip   gpuScalarUsp / gpuScalarUip (sphere.cu:4-99)
//   scattered field lookups    gpuScalarUpLut (nfScalarUpLut.cu:11-97)
//   rotation resample          gpuInterpolateU / gpuInterpolateUf (nfInterpolate.cu:13-66)
// Tables are built in FP64 on the GPU and stored as double (FP64 precisions)
// or float (FP32_FAST, like the reference's float textures).  Texture
// fetches are bilinear with exact weights on unnormalised coordinates with
// the +0.5 texel offset (clamp addressing for the LUTs, zero border for the
// rotation).  Built with -ffp-contract=off.
#include "geometry.h"
#include "internal.h"
#include "refmath.h"
#include "zj_device.h"

namespace bimsim {
namespace {

constexpr int kThreads = 256;

template <class T> struct Cx;
template <> struct Cx<double> { using C = double2; };
template <> struct Cx<float> { using C = float2; };

__device__ inline void pixel_of(int pix, int uR, int& iu, int& iv) {
    iv = pix / uR;
    iu = pix - iv * uR;
}

__device__ inline dc ld(const double2* p, size_t i) {
    const double2 v = p[i];
    return dc{v.x, v.y};
}
__device__ inline dc ld(const float2* p, size_t i) {
    const float2 v = p[i];
    return dc{(double)v.x, (double)v.y};
}
__device__ inline double ldr(const double* p, size_t i) { return p[i]; }
__device__ inline double ldr(const float* p, size_t i) { return (double)p[i]; }
__device__ inline void st(double2* p, size_t i, dc v) { p[i] = make_double2(v.r, v.i); }
__device__ inline void st(float2* p, size_t i, dc v) { p[i] = make_float2((float)v.r, (float)v.i); }

// ------------------------------------------------- incident-field Bessel LUT
template <class T>
__global__ void uf_bessel_lut_kernel(int m, double kmag, double dmin, double dr, int dR,
                                     double* __restrict__ work, int wstride, T* __restrict__ lut) {
    const int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= dR) return;
    const double d = id * dr + dmin;
    const double kd = kmag * d;
    double* j = work + (size_t)id * 4 * wstride;
    zj::sph_jy_real(m, kd, true, j, j + wstride, j + 2 * wstride, j + 3 * wstride);
    for (int l = 0; l <= m; ++l) {
        T jv = (T)j[l];
        if (isnan(jv) || isinf(jv)) jv = (kd == 0.0 && l == 0) ? (T)1 : (T)0;
        lut[(size_t)id * (m + 1) + l] = jv;
    }
}

// gpuScalarUfLut: j_l from tex2D(texJ, l + 0.5, di + 0.5): exact column l,
// linear in di, clamped.
template <class T, class Out>
__global__ __launch_bounds__(kThreads) void uf_lut_kernel(QuadArg q, int uR, int vR, d3 f, d3 k,
                                                          double amp, double ca, double cb, int m,
                                                          const T* __restrict__ lut, double dmin,
                                                          double dmax, int dR,
                                                          Out* __restrict__ Uf) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= uR * vR) return;
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const d3 p = quad_at(q, (double)iu / (double)uR, (double)iv / (double)vR);
    d3 r = sub(p, f);
    const double d = len(r);
    if (d == 0.0) {
        st(Uf, pix, dc{amp * 2.0 * (double)kPIf * (ca - cb), 0.0});
        return;
    }
    r = norm(r);
    const dc im = {0.0, 1.0};
    dc il = {1.0, 0.0};
    const double ct = dot(k, r);
    double P0 = 1.0, P1 = ct;
    double Pa[3] = {1.0, ca, 1.0}, Pb[3] = {1.0, cb, 1.0};
    const double di = ((d - dmin) / (dmax - dmin)) * (double)(dR - 1);
    const double fy = floor(di);
    const double wy = di - fy;
    const int i0 = min(max((int)fy, 0), dR - 1), i1 = min(max((int)fy + 1, 0), dR - 1);
    dc sum = {0.0, 0.0};
    for (int l = 0; l <= m; ++l) {
        const double jl = (1.0 - wy) * ldr(lut, (size_t)i0 * (m + 1) + l) +
                          wy * ldr(lut, (size_t)i1 * (m + 1) + l);
        double Pl;
        if (l == 0) {
            Pl = P0;
        } else if (l == 1) {
            Pl = P1;
            Pa[2] = Pa[0];
            ref::shift_legendre(l + 1, ca, Pa[0], Pa[1]);
            Pb[2] = Pb[0];
            ref::shift_legendre(l + 1, cb, Pb[0], Pb[1]);
        } else {
            ref::shift_legendre(l, ct, P0, P1);
            Pl = P1;
            Pa[2] = Pa[0];
            ref::shift_legendre(l + 1, ca, Pa[0], Pa[1]);
            Pb[2] = Pb[0];
            ref::shift_legendre(l + 1, cb, Pb[0], Pb[1]);
        }
        const double D = ((Pa[1] - Pa[2]) - Pb[1]) + Pb[2];
        sum = cadd(sum, cscl(cscl(cscl(il, jl), Pl), D));
        il = cmul(il, im);
    }
    st(Uf, pix, cscl(cscl(cscl(sum, 2.0), (double)kPIf), amp));
}

// -------------------------------------------------------- sphere radial LUTs
// h[ir][l] = (j_l, y_l)(k r), r in [max(a, dmin), dmax] over dR samples
__global__ void hankel_lut_kernel(int nl, double kmag, double lo, double dr, int dR,
                                  double* __restrict__ work, int wstride,
                                  double2* __restrict__ h) {
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir >= dR) return;
    const double r = ir * dr + lo;
    const double kr = kmag * r;
    double* j = work + (size_t)ir * 4 * wstride;
    zj::sph_jy_real(nl, kr, true, j, j + wstride, j + 2 * wstride, j + 3 * wstride);
    for (int l = 0; l <= nl; ++l) h[(size_t)ir * (nl + 1) + l] = make_double2(j[l], j[wstride + l]);
}

// j[ir][l] = j_l(k n r), r = ir * a/(aR-1); r = 0 -> delta_l0
__global__ void bessel_lut_kernel(int nl, double kmag, double nr, double ni, double a, int aR,
                                  zj::zc* __restrict__ work, int wstride,
                                  double2* __restrict__ jl) {
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    if (ir >= aR) return;
    if (ir == 0) {
        for (int l = 0; l <= nl; ++l) jl[l] = make_double2(l == 0 ? 1.0 : 0.0, 0.0);
        return;
    }
    const double dr = a / (aR - 1);
    const double r = ir * dr;
    // k*n*r as (k*n)*r with k*n scaling each component
    const zj::zc knr = {(kmag * nr) * r, (kmag * ni) * r};
    zj::zc* j = work + (size_t)ir * 4 * wstride;
    zj::sph_jy_cplx(nl, knr, true, j, j + wstride, j + 2 * wstride, j + 3 * wstride, false);
    for (int l = 0; l <= nl; ++l) jl[(size_t)ir * (nl + 1) + l] = make_double2(j[l].r, j[l].i);
}

// L[it][ir] = sum_l C_l rad[ir][l] P_l(cos(it * dtheta)) (sphere.cu:4-99)
template <class Out>
__global__ void lut_eval_kernel(int nl, const double2* __restrict__ C, const double2* __restrict__ rad,
                                int nr, int thetaR, double dtheta, Out* __restrict__ L) {
    const int ir = blockIdx.x * blockDim.x + threadIdx.x;
    const int it = blockIdx.y;
    if (ir >= nr) return;
    const double theta = dtheta * it;
    const double ct = cos(theta);
    double P0 = 1.0, P1 = ct;
    dc acc = {0.0, 0.0};
    for (int l = 0; l <= nl; ++l) {
        const dc h = ld(rad, (size_t)ir * (nl + 1) + l);
        if (l == 0) {
            acc = cadd(acc, cscl(cmul(ld(C, 0), h), P0));
        } else {
            if (l > 1) ref::shift_legendre(l, ct, P0, P1);
            acc = cadd(acc, cscl(cmul(ld(C, l), h), P1));
        }
    }
    st(L, (size_t)it * nr + ir, acc);
}

// lut_eval_kernel for kEvalRows angle rows per workgroup, with the same
// operations in the same order (identical results): the rows' Legendre
// values are computed once into LDS, and C_l rad[ir][l] once per (ir, l)
// rather than once per (angle, ir, l).
constexpr int kEvalRows = 16;
constexpr int kEvalMaxOrders = 512;   // LDS: kEvalRows doubles per order

template <class Out>
__global__ __launch_bounds__(kThreads) void lut_eval_rows_kernel(
    int nl, const double2* __restrict__ C, const double2* __restrict__ rad, int nr, int thetaR,
    double dtheta, Out* __restrict__ L) {
    extern __shared__ double Pl[];   // [l][row], (nl + 1) kEvalRows doubles
    const int it0 = blockIdx.y * kEvalRows;
    if (threadIdx.x < kEvalRows) {
        const int r = threadIdx.x;
        const double ct = cos(dtheta * (it0 + r));
        double P0 = 1.0, P1 = ct;
        Pl[r] = P0;
        Pl[kEvalRows + r] = P1;
        for (int l = 2; l <= nl; ++l) {
            ref::shift_legendre(l, ct, P0, P1);
            Pl[l * kEvalRows + r] = P1;
        }
    }
    __syncthreads();
    const int ir = blockIdx.x * kThreads + threadIdx.x;
    if (ir >= nr) return;
    dc acc[kEvalRows];
#pragma unroll
    for (int r = 0; r < kEvalRows; ++r) acc[r] = dc{0.0, 0.0};
    const double2* h = rad + (size_t)ir * (nl + 1);
    for (int l = 0; l <= nl; ++l) {
        const dc c = cmul(ld(C, l), ld(h, l));
        const double* P = Pl + l * kEvalRows;
#pragma unroll
        for (int r = 0; r < kEvalRows; ++r) acc[r] = cadd(acc[r], cscl(c, P[r]));
    }
#pragma unroll
    for (int r = 0; r < kEvalRows; ++r)
        if (it0 + r < thetaR) st(L, (size_t)(it0 + r) * nr + ir, acc[r]);
}

// bilinear fetch with clamp addressing at data coordinates (x, y)
template <class CT>
__device__ inline dc fetch_clamp(const CT* L, int W, int H, float xf, float yf) {
    const double x = xf, y = yf;
    const double fx = floor(x), fy = floor(y);
    const double wx = x - fx, wy = y - fy;
    const int x0 = (int)fx, y0 = (int)fy;
    const int xa = min(max(x0, 0), W - 1), xb = min(max(x0 + 1, 0), W - 1);
    const int ya = min(max(y0, 0), H - 1), yb = min(max(y0 + 1, 0), H - 1);
    const dc a = cadd(cscl(ld(L, (size_t)ya * W + xa), 1.0 - wx), cscl(ld(L, (size_t)ya * W + xb), wx));
    const dc b = cadd(cscl(ld(L, (size_t)yb * W + xa), 1.0 - wx), cscl(ld(L, (size_t)yb * W + xb), wx));
    return cadd(cscl(a, 1.0 - wy), cscl(b, wy));
}

// gpuScalarUpLut for one sphere (U += sum over waves)
template <class CT, class Out>
__global__ __launch_bounds__(kThreads) void uplut_kernel(
    QuadArg q, int uR, int vR, d3 f, double kmag, d3 ps, double a, const CT* __restrict__ Usp,
    int dR, const CT* __restrict__ Uip, int aR, int thetaR, double dmin, double dmax,
    const double* __restrict__ waves, int nk, double sA, Out* __restrict__ U) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= uR * vR) return;
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const d3 p = quad_at(q, (double)iu / (double)uR, (double)iv / (double)vR);
    d3 r = sub(p, ps);
    const double d = len(r);
    const double lo = fmax(a, dmin);
    // the reference keeps the texture coordinates in float
    const float di = (float)(((d - lo) / (dmax - lo)) * (double)(dR - 1));
    const float ai = (float)(((d - dmin) / (a - dmin)) * (double)(aR - 1));
    const d3 c = sub(ps, f);
    dc sum = {0.0, 0.0};
    for (int w = 0; w < nk; ++w) {
        r = norm(r);
        const d3 k = mk(waves + 3 * w);
        double ct = dot(k, r);
        if (ct < -1.0) ct = -1.0;
        if (ct > 1.0) ct = 1.0;
        // r-hat is 0/0 at the sphere centre (the reference then fetches at a
        // NaN coordinate); the Uip column there is A_0 for every angle
        if (ct != ct) ct = 1.0;
        const float ti = (float)((acos(ct) / (double)kPIf) * (double)(thetaR - 1));
        const dc phase = cexp_i(kmag * dot(k, c));
        const dc val = d < a ? fetch_clamp(Uip, aR, thetaR, ai, ti) : fetch_clamp(Usp, dR, thetaR, di, ti);
        sum = cadd(sum, cmul(cscl(phase, sA), val));
    }
    const dc o = ld(U, pix);
    st(U, pix, cadd(o, sum));
}

// FP32 form of gpuScalarUpLut (nfScalarUpLut.cu:11-97) for float tables of
// width >= 2.  What depends on the pixel only (d, r-hat, the table row, the
// radial column pair and weight) is computed once per pixel in FP64; what
// depends on the wave only (direction and the phase (1/nk) subA
// exp(i kmag k.c)) is computed once per workgroup in FP64 into LDS.  The
// per-(pixel, wave) work is the reference's float arithmetic: the angle
// acos(k.r-hat) and a bilinear fetch as two 16-byte loads of neighbouring
// texels.  Clamp addressing is kept by clamping the coordinates into
// [0, W-1] x [0, thetaR-1]: at a clamped edge the weight of the outer texel
// is 0 or 1, which reproduces the edge texel.
constexpr int kLutWaves = 1024;   // waves per LDS chunk
constexpr int kXcds = 8;          // XCDs of the chip
static_assert(kThreads == 256, "a workgroup is four waves over a 16 x 16 tile");

typedef float f4_a8 __attribute__((ext_vector_type(4), aligned(8)));

__global__ __launch_bounds__(kThreads) void uplut_f32_kernel(
    QuadArg q, int uR, int vR, d3 f, double kmag, d3 ps, double a, const float2* __restrict__ Usp,
    int dR, const float2* __restrict__ Uip, int aR, int thetaR, double dmin, double dmax,
    const double* __restrict__ waves, int nk, double sA, float2* __restrict__ U) {
    __shared__ float4 wk[kLutWaves];   // (kx, ky, kz, phase re)
    __shared__ float wpi[kLutWaves];   // phase im
    // Workgroups take 16 x 16 pixel tiles and each wave an 8 x 8 quarter, so
    // that the lanes of a wave fetch nearby texels.  Workgroups are
    // dispatched to the XCDs in turn; tiles are renumbered so that each XCD
    // (and its L2) holds a contiguous band of the image.
    const int tiles_u = (uR + 15) >> 4;
    const int ntiles = tiles_u * ((vR + 15) >> 4);
    const int xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    const int tile = xcd * (ntiles / kXcds) + min(xcd, ntiles % kXcds) + slot;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int iu0 = (tile % tiles_u) * 16 + (wave & 1) * 8 + (lane & 7);
    const int iv0 = (tile / tiles_u) * 16 + (wave >> 1) * 8 + (lane >> 3);
    const bool live = iu0 < uR && iv0 < vR;
    const int iu = min(iu0, uR - 1), iv = min(iv0, vR - 1);
    const d3 p = quad_at(q, (double)iu / (double)uR, (double)iv / (double)vR);
    const d3 r = sub(p, ps);
    const double d = len(r);
    // r-hat is 0/0 at the sphere centre; the angle there is taken as 0 (as
    // in the FP64 kernel) through a bias of the cosine with r-hat = 0
    const d3 rn = norm(r);
    const bool centre = rn.x != rn.x || rn.y != rn.y || rn.z != rn.z;
    const float rx = centre ? 0.0f : (float)rn.x, ry = centre ? 0.0f : (float)rn.y;
    const float rz = centre ? 0.0f : (float)rn.z, cbias = centre ? 1.0f : 0.0f;
    const bool inside = d < a;
    const double lo = fmax(a, dmin);
    const float di = (float)(((d - lo) / (dmax - lo)) * (double)(dR - 1));
    const float ai = (float)(((d - dmin) / (a - dmin)) * (double)(aR - 1));
    const int W = inside ? aR : dR;
    const float2* L = inside ? Uip : Usp;
    float x = inside ? ai : di;
    x = x >= 0.0f ? x : 0.0f;                    // also maps NaN to the first column
    x = x <= (float)(W - 1) ? x : (float)(W - 1);
    const int xa = min((int)floorf(x), W - 2);
    const float wx = x - (float)xa;
    const float* Lx = (const float*)(L + xa);
    const int W2 = 2 * W;   // floats per table row
    const float tscale = (float)(thetaR - 1) / kPIf;
    const float tmax = (float)(thetaR - 1);
    const d3 c = sub(ps, f);
    // waves summed in groups of 8 whose sums enter a Kahan accumulator
    float sr = 0.0f, si = 0.0f, cr = 0.0f, ci = 0.0f;
    for (int w0 = 0; w0 < nk; w0 += kLutWaves) {
        const int nw = min(kLutWaves, nk - w0);
        __syncthreads();
        for (int i = threadIdx.x; i < nw; i += kThreads) {
            const d3 k = mk(waves + 3 * (size_t)(w0 + i));
            const dc ph = cscl(cexp_i(kmag * dot(k, c)), sA);
            wk[i] = make_float4((float)k.x, (float)k.y, (float)k.z, (float)ph.r);
            wpi[i] = (float)ph.i;
        }
        __syncthreads();
        for (int g = 0; g < nw; g += 8) {
            float gr = 0.0f, gi = 0.0f;
#pragma unroll 4
            for (int j = 0; j < 8; ++j) {
                const int w = g + j;
                if (w >= nw) break;
                const float4 k = wk[w];
                float ct = fmaf(k.z, rz, fmaf(k.y, ry, fmaf(k.x, rx, cbias)));
                ct = fminf(fmaxf(ct, -1.0f), 1.0f);
                float t = acosf(ct) * tscale;
                t = t <= tmax ? t : tmax;
                const int ya = min((int)t, thetaR - 2);
                const float wy = t - (float)ya;
                const float* row = Lx + __umul24(ya, W2);
                const f4_a8 A = *(const f4_a8*)row;
                const f4_a8 B = *(const f4_a8*)(row + W2);
                const float ar = fmaf(wx, A.z - A.x, A.x), aim = fmaf(wx, A.w - A.y, A.y);
                const float br = fmaf(wx, B.z - B.x, B.x), bim = fmaf(wx, B.w - B.y, B.y);
                const float vr = fmaf(wy, br - ar, ar), vi = fmaf(wy, bim - aim, aim);
                const float pr = k.w, pi = wpi[w];
                gr = fmaf(pr, vr, fmaf(-pi, vi, gr));
                gi = fmaf(pr, vi, fmaf(pi, vr, gi));
            }
            kahan(sr, si, cr, ci, gr, gi);
        }
    }
    if (live) {
        const size_t pix = (size_t)iv * uR + iu;
        const float2 o = U[pix];
        U[pix] = make_float2(o.x + sr, o.y + si);
    }
}

// gpuInterpolateU (integer centre) / gpuInterpolateUf (float centre), zero border
template <class C, class T>
__global__ void rotate_kernel(const C* __restrict__ src, C* __restrict__ dst, int uR, int vR,
                              double deg, int centre_float) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= uR * vR) return;
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const double cs = cos(deg), sn = sin(deg);
    float nx, ny;
    if (centre_float) {
        const double uc = (double)(float)(uR / 2.0), vc = (double)(float)(vR / 2.0);
        nx = (float)(uc + ((double)iu - uc) * cs - ((double)iv - vc) * sn);
        ny = (float)(vc + ((double)iu - uc) * sn + ((double)iv - vc) * cs);
    } else {
        const int uc = (int)(uR / 2.0), vc = (int)(vR / 2.0);
        nx = (float)((double)uc + (double)(iu - uc) * cs - (double)(iv - vc) * sn);
        ny = (float)((double)vc + (double)(iu - uc) * sn + (double)(iv - vc) * cs);
    }
    const double x = nx, y = ny;
    const double fx = floor(x), fy = floor(y);
    const double wx = x - fx, wy = y - fy;
    const int x0 = (int)fx, y0 = (int)fy;
    auto at = [&](int xx, int yy) -> dc {
        if (xx < 0 || yy < 0 || xx >= uR || yy >= vR) return dc{0.0, 0.0};
        return ld(src, (size_t)yy * uR + xx);
    };
    const dc A = cadd(cscl(at(x0, y0), 1.0 - wx), cscl(at(x0 + 1, y0), wx));
    const dc B = cadd(cscl(at(x0, y0 + 1), 1.0 - wx), cscl(at(x0 + 1, y0 + 1), wx));
    st(dst, pix, cadd(cscl(A, 1.0 - wy), cscl(B, wy)));
}

}  // namespace

// sphere::calcUp table sizes (sphere.cpp:222-241): aR from the cross-section
// radius sqrt(a - d_min) (sic), dR = 2R unless overridden.
void sphere_lut_dims_host(const QuadArg& q, const bimsim_sphere& s, int R, int dR_override,
                          int* aR, int* dR, double* dmin, double* dmax) {
    quad_dist(q, mk(s.p), dmin, dmax);
    double a_inter = 0.0;
    if (*dmin < s.a) a_inter = std::sqrt(s.a - *dmin);
    const d3 diag = sub(quad_at(q, 0.0, 0.0), quad_at(q, 1.0, 1.0));
    *aR = (int)(1.0 + (double)(2u * (unsigned)R) * a_inter / len(diag));
    *dR = dR_override > 0 ? dR_override : 2 * R;
}

// L[it][ir] = sum_l C_l rad[ir][l] P_l(cos(it dtheta)) for nr radii
static int lut_eval_launch(bimsim_ctx* ctx, bool f32, int nl, const double2* C,
                           const double2* rad, int nr, int thetaR, double dtheta, void* L) {
    const unsigned gx = (unsigned)((nr + kThreads - 1) / kThreads);
    if (nl + 1 <= kEvalMaxOrders) {
        const dim3 grid(gx, (unsigned)((thetaR + kEvalRows - 1) / kEvalRows));
        const size_t lds = sizeof(double) * kEvalRows * (size_t)(nl + 1);
        if (f32)
            hipLaunchKernelGGL(lut_eval_rows_kernel<float2>, grid, dim3(kThreads), lds,
                               ctx->stream, nl, C, rad, nr, thetaR, dtheta, (float2*)L);
        else
            hipLaunchKernelGGL(lut_eval_rows_kernel<double2>, grid, dim3(kThreads), lds,
                               ctx->stream, nl, C, rad, nr, thetaR, dtheta, (double2*)L);
    } else if (f32) {
        hipLaunchKernelGGL(lut_eval_kernel<float2>, dim3(gx, thetaR), dim3(kThreads), 0,
                           ctx->stream, nl, C, rad, nr, thetaR, dtheta, (float2*)L);
    } else {
        hipLaunchKernelGGL(lut_eval_kernel<double2>, dim3(gx, thetaR), dim3(kThreads), 0,
                           ctx->stream, nl, C, rad, nr, thetaR, dtheta, (double2*)L);
    }
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // namespace bimsim

using namespace bimsim;

extern "C" {

int bimsim_sphere_lut_dims(const bimsim_quad* pos, const bimsim_sphere* s, int R, int dR_override,
                           int* aR, int* dR, double* dmin, double* dmax) {
    if (!pos || !s || !aR || !dR || !dmin || !dmax || R <= 0) return BIMSIM_E_INVALID;
    sphere_lut_dims_host(to_quad(*pos), *s, R, dR_override, aR, dR, dmin, dmax);
    return BIMSIM_OK;
}

int bimsim_uf_bessel_lut(bimsim_ctx* ctx, int m, double kmag, double dmin, double dmax, int dR,
                         int prec, void* lut) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, lut && m >= 1 && m <= 4096 && dR >= 2 && dR <= (1 << 24),
                     "bad Bessel LUT arguments");
    BIMSIM_CHECK_ARG(ctx, dmax > dmin && dmin >= 0.0 && kmag > 0.0, "bad LUT range");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int wstride = m + 2;
    void* work = nullptr;
    BIMSIM_TRY(scratch_get(ctx, kSlotMisc, sizeof(double) * 4 * (size_t)wstride * dR, &work));
    const double dr = (dmax - dmin) / (dR - 1);
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL(uf_bessel_lut_kernel<float>, dim3((dR + 63) / 64), dim3(64), 0,
                           ctx->stream, m, kmag, dmin, dr, dR, (double*)work, wstride, (float*)lut);
    else
        hipLaunchKernelGGL(uf_bessel_lut_kernel<double>, dim3((dR + 63) / 64), dim3(64), 0,
                           ctx->stream, m, kmag, dmin, dr, dR, (double*)work, wstride, (double*)lut);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int bimsim_nf_uf_lut(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR, const double focus[3],
                     const double kdir[3], double amp, double cos_alpha, double cos_beta, int m,
                     const void* lut, int dR, double dmin, double dmax, int prec, void* Uf) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, pos && focus && kdir && lut && Uf, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 28) && m >= 1 && dR >= 2,
                     "bad sizes");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int npix = uR * vR;
    const dim3 grid((npix + kThreads - 1) / kThreads);
    const QuadArg q = to_quad(*pos);
    const int tmr = timer_begin(ctx, kStageUf);
    if (prec == BIMSIM_FP32_FAST)
        hipLaunchKernelGGL((uf_lut_kernel<float, float2>), grid, dim3(kThreads), 0, ctx->stream, q,
                           uR, vR, mk(focus), mk(kdir), amp, cos_alpha, cos_beta, m,
                           (const float*)lut, dmin, dmax, dR, (float2*)Uf);
    else
        hipLaunchKernelGGL((uf_lut_kernel<double, double2>), grid, dim3(kThreads), 0, ctx->stream,
                           q, uR, vR, mk(focus), mk(kdir), amp, cos_alpha, cos_beta, m,
                           (const double*)lut, dmin, dmax, dR, (double2*)Uf);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    timer_end(ctx, tmr);
    return BIMSIM_OK;
}

int bimsim_sphere_lut(bimsim_ctx* ctx, const bimsim_sphere* s, const void* A, const void* B,
                      double kmag, double dmin, double dmax, int dR, int aR, int thetaR, int prec,
                      void* Usp, void* Uip) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, s && A && B && Usp, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, dR >= 2 && aR >= 1 && thetaR >= 2 && s->Nl >= 1 && s->Nl <= 4096,
                     "bad LUT sizes");
    BIMSIM_CHECK_ARG(ctx, (long long)dR * thetaR <= (1LL << 30), "LUT too large");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int nl = s->Nl;
    const int wstride = nl + 2;
    const int nmax = std::max(dR, aR);
    void *work = nullptr, *rad = nullptr;
    BIMSIM_TRY(scratch_get(ctx, kSlotMisc, sizeof(zj::zc) * 4 * (size_t)wstride * nmax, &work));
    BIMSIM_TRY(scratch_get(ctx, kSlotCoef, sizeof(double2) * (size_t)(nl + 1) * nmax, &rad));
    const double2* Ad = (const double2*)A + s->coeff_off;
    const double2* Bd = (const double2*)B + s->coeff_off;
    // (PI) / (thetaR - 1) is float arithmetic in the reference (sphere.cu:21)
    const double dtheta = (double)(kPIf / (float)(thetaR - 1));
    const bool f32 = prec == BIMSIM_FP32_FAST;
    // Usp from the Hankel table
    const double lo = std::max(s->a, dmin);
    hipLaunchKernelGGL(hankel_lut_kernel, dim3((dR + 63) / 64), dim3(64), 0, ctx->stream, nl, kmag,
                       lo, (dmax - lo) / (dR - 1), dR, (double*)work, wstride, (double2*)rad);
    BIMSIM_TRY(lut_eval_launch(ctx, f32, nl, Bd, (const double2*)rad, dR, thetaR, dtheta, Usp));
    if (Uip) {
        hipLaunchKernelGGL(bessel_lut_kernel, dim3((aR + 63) / 64), dim3(64), 0, ctx->stream, nl,
                           kmag, s->n[0], s->n[1], s->a, aR, (zj::zc*)work, wstride, (double2*)rad);
        BIMSIM_TRY(lut_eval_launch(ctx, f32, nl, Ad, (const double2*)rad, aR, thetaR, dtheta, Uip));
    }
    return BIMSIM_OK;
}

int bimsim_sphere_radial_lut(bimsim_ctx* ctx, const bimsim_sphere* s, double kmag, double dmin,
                             double dmax, int dR, int aR, void* h, void* j) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, s && (h || j), "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, s->Nl >= 1 && s->Nl <= 4096 && (!h || dR >= 2) && (!j || aR >= 2),
                     "bad radial table sizes");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int nl = s->Nl;
    const int wstride = nl + 2;
    const int nmax = std::max(h ? dR : 0, j ? aR : 0);
    void* work = nullptr;
    BIMSIM_TRY(scratch_get(ctx, kSlotMisc, sizeof(zj::zc) * 4 * (size_t)wstride * nmax, &work));
    if (h) {
        const double lo = std::max(s->a, dmin);
        hipLaunchKernelGGL(hankel_lut_kernel, dim3((dR + 63) / 64), dim3(64), 0, ctx->stream, nl,
                           kmag, lo, (dmax - lo) / (dR - 1), dR, (double*)work, wstride, (double2*)h);
        BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    }
    if (j) {
        hipLaunchKernelGGL(bessel_lut_kernel, dim3((aR + 63) / 64), dim3(64), 0, ctx->stream, nl,
                           kmag, s->n[0], s->n[1], s->a, aR, (zj::zc*)work, wstride, (double2*)j);
        BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    }
    return BIMSIM_OK;
}

int bimsim_nf_us_lut(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                     const double focus[3], double kmag, const bimsim_sphere* s, const void* Usp,
                     int dR, const void* Uip, int aR, int thetaR, double dmin, double dmax,
                     const double* waves, int nk, double subA, int prec, void* U) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, pos && focus && s && Usp && Uip && waves && U, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 28) && nk > 0 &&
                              dR >= 2 && aR >= 1 && thetaR >= 2,
                     "bad sizes");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int npix = uR * vR;
    const dim3 grid((npix + kThreads - 1) / kThreads);
    const QuadArg q = to_quad(*pos);
    const double sA = (double)(1.0f / (float)nk) * subA;   // (1.0f/nk) * A
    const int tmr = timer_begin(ctx, kStageUs);
    // the FP32 kernel addresses rows with 24-bit products
    const bool rows24 = (long long)thetaR * 2 * std::max(dR, aR) < (1LL << 24);
    const dim3 tiles((unsigned)((uR + 15) / 16) * (unsigned)((vR + 15) / 16));
    if (prec == BIMSIM_FP32_FAST && aR >= 2 && rows24)
        hipLaunchKernelGGL(uplut_f32_kernel, tiles, dim3(kThreads), 0, ctx->stream, q, uR, vR,
                           mk(focus), kmag, mk(s->p), s->a, (const float2*)Usp, dR,
                           (const float2*)Uip, aR, thetaR, dmin, dmax, waves, nk, sA, (float2*)U);
    else if (prec == BIMSIM_FP32_FAST)   // one-column inner table, or very large tables
        hipLaunchKernelGGL((uplut_kernel<float2, float2>), grid, dim3(kThreads), 0, ctx->stream, q,
                           uR, vR, mk(focus), kmag, mk(s->p), s->a, (const float2*)Usp, dR,
                           (const float2*)Uip, aR, thetaR, dmin, dmax, waves, nk, sA, (float2*)U);
    else
        hipLaunchKernelGGL((uplut_kernel<double2, double2>), grid, dim3(kThreads), 0, ctx->stream,
                           q, uR, vR, mk(focus), kmag, mk(s->p), s->a, (const double2*)Usp, dR,
                           (const double2*)Uip, aR, thetaR, dmin, dmax, waves, nk, sA, (double2*)U);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    timer_end(ctx, tmr);
    return BIMSIM_OK;
}

int bimsim_nf_interpolate(bimsim_ctx* ctx, int prec, double deg, int uR, int vR, const void* U_src,
                          const void* Uf_src, void* U, void* Uf) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 28), "bad sizes");
    BIMSIM_CHECK_ARG(ctx, (U_src && U) || (Uf_src && Uf), "nothing to resample");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int npix = uR * vR;
    const dim3 grid((npix + kThreads - 1) / kThreads);
    const bool f32 = prec == BIMSIM_FP32_FAST;
    // nfInterpolate.cu:112-113: Uf with a float centre, U with an integer centre
    if (Uf_src && Uf) {
        if (f32)
            hipLaunchKernelGGL((rotate_kernel<float2, float>), grid, dim3(kThreads), 0, ctx->stream,
                               (const float2*)Uf_src, (float2*)Uf, uR, vR, deg, 1);
        else
            hipLaunchKernelGGL((rotate_kernel<double2, double>), grid, dim3(kThreads), 0,
                               ctx->stream, (const double2*)Uf_src, (double2*)Uf, uR, vR, deg, 1);
    }
    if (U_src && U) {
        if (f32)
            hipLaunchKernelGGL((rotate_kernel<float2, float>), grid, dim3(kThreads), 0, ctx->stream,
                               (const float2*)U_src, (float2*)U, uR, vR, deg, 0);
        else
            hipLaunchKernelGGL((rotate_kernel<double2, double>), grid, dim3(kThreads), 0,
                               ctx->stream, (const double2*)U_src, (double2*)U, uR, vR, deg, 0);
    }
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // extern "C"
