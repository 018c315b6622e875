// Incident field Uf and the Uf + Us sum (reference nfScalarUf.cu:9-217,
// nfSumUf.cu:6-117).  Built with -ffp-contract=off.  The FP64 forms run the
// series in reference operation order.  FP32_FAST runs it in FP64 too, with
// the pixel-invariant factors i^l D_l from a per-instance table and the
// recurrences multiplying by reciprocals (explicit FMAs), and stores float
// pairs.
#include <type_traits>

#include "geometry.h"
#include "internal.h"
#include "batch.h"
#include "refmath.h"
#include "uf_device.h"

namespace bimsim {
namespace {

constexpr int kThreads = 256;

using namespace ufd;

template <class Out>
__global__ __launch_bounds__(kThreads) void uf_series_kernel(QuadArg q, int uR, int vR, d3 f,
                                                             d3 k, double kmag, double amp,
                                                             double ca, double cb, int m,
                                                             Out* __restrict__ Uf) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= uR * vR) return;
    store(Uf, pix, uf_series_value(q, uR, vR, pix, f, k, kmag, amp, ca, cb, m));
}

__global__ __launch_bounds__(kThreads) void uf_fast_kernel(QuadArg q, int uR, int vR, d3 f, d3 k,
                                                           double kmag, double amp, double ca,
                                                           double cb, int m,
                                                           const double* __restrict__ T, int flat,
                                                           float2* __restrict__ Uf) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= uR * vR) return;
    store(Uf, pix, uf_fast_value(q, uR, vR, pix, f, k, kmag, amp, ca, cb, m, T, flat != 0));
}

template <class Out>
__global__ __launch_bounds__(kThreads) void uf_plane_kernel(QuadArg q, int uR, int vR, d3 f, d3 k,
                                                            double kmag, double amp,
                                                            Out* __restrict__ Uf) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= uR * vR) return;
    store(Uf, pix, uf_plane_value(q, uR, vR, pix, f, k, kmag, amp));
}

// Incident fields of a batch of instances.  Grid: (pixel blocks, instances).
// With sum given, also clears each instance's detector images (sum[i],
// nd2 <= uR vR), which the detectors then accumulate into.
template <class Out, class R>
__global__ __launch_bounds__(kThreads) void uf_batch_kernel(const UfInst* __restrict__ insts, int n,
                                                            const double* __restrict__ tabs,
                                                            const double* __restrict__ waves,
                                                            Out* __restrict__ Uf,
                                                            const SumInst* __restrict__ sum,
                                                            R* __restrict__ D, R* __restrict__ Di) {
    for (int i = blockIdx.y; i < n; i += gridDim.y) {
        const UfInst& I = insts[i];
        const int nw = uf_batch_work(I), nd2 = sum ? sum[i].nd2 : 0;   // nd2 <= uR vR
        for (int pix = blockIdx.x * kThreads + threadIdx.x; pix < max(nw, nd2);
             pix += gridDim.x * kThreads) {
            if (pix < nd2) {
                D[sum[i].image_off + pix] = R(0);
                Di[sum[i].image_off + pix] = R(0);
            }
            if (pix < nw) uf_batch_item(I, pix, tabs, waves, Uf);
        }
    }
}

// Monte-Carlo incident field, FP64 forms: one pixel per thread, the waves
// summed in order (uf_mc_value).
template <class Out>
__global__ __launch_bounds__(kThreads) void uf_mc_kernel(QuadArg q, int uR, int vR, d3 f,
                                                         double kmag,
                                                         const double* __restrict__ W, int nk,
                                                         double sA, Out* __restrict__ Uf) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= uR * vR) return;
    store(Uf, pix, uf_mc_value(q, uR, vR, pix, f, kmag, W, nk, sA));
}

// Monte-Carlo incident field, complex64 form.  With p = A + X iu/uR +
// Y iv/vR each wave's phase splits into a row and a column factor,
//   sA e^{i k K.(p - f)} = [sA e^{i k (K.(A - f) + K.Y iv/vR)}] e^{i k K.X iu/uR},
// so a tile of kMcTile x kMcTile pixels is a complex rank-nk product: per
// chunk of kMcChunk waves the workgroup evaluates the tile's row and column
// factors (phases in FP64, reduced to [-pi, pi], then sincosf) into LDS and
// every thread accumulates its 4 x 4 pixels, 4 FMAs per pixel and wave.
constexpr int kMcTile = 64;
constexpr int kMcChunk = 32;
constexpr double kTwoPi = 6.283185307179586476925;

__device__ __forceinline__ float2 cis_reduced(double x) {
    const double r = fma(-kTwoPi, rint(x * (1.0 / kTwoPi)), x);
    float sn, cs;
    sincosf((float)r, &sn, &cs);
    return make_float2(cs, sn);
}

// Tile `tile` of instance I's Monte-Carlo field.
__device__ __forceinline__ void uf_mc_tile(const UfInst& I, int tile,
                                           const double* __restrict__ waves,
                                           float2* __restrict__ Uf) {
    static_assert(kThreads == 256 && kMcTile == 64, "16 x 16 threads of 4 x 4 pixels");
    __shared__ __align__(16) float2 sa[kMcChunk][kMcTile];   // column factors
    __shared__ __align__(16) float2 sb[kMcChunk][kMcTile];   // row factors times the weight
    __shared__ double sw[kMcChunk][3];          // k K.X/uR, k K.Y/vR, k K.(A - f)
    const int tu = (I.uR + kMcTile - 1) / kMcTile, tv = (I.vR + kMcTile - 1) / kMcTile;
    if (tile >= tu * tv) return;   // uniform per workgroup
    const int u0 = (tile % tu) * kMcTile, v0 = (tile / tu) * kMcTile;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const double* W = waves + 3 * I.wave_off;
    const d3 af = sub(I.q.A, I.f);
    float ar[4][4], ai[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) ar[i][j] = ai[i][j] = 0.f;
    for (int w0 = 0; w0 < I.nk; w0 += kMcChunk) {
        const int nw = min(kMcChunk, I.nk - w0);
        __syncthreads();   // the previous chunk's factors are consumed
        if (threadIdx.x < nw) {
            const int w = w0 + threadIdx.x;
            const d3 K = {W[3 * w], W[3 * w + 1], W[3 * w + 2]};
            sw[threadIdx.x][0] = I.kmag * dot(K, I.q.X) / (double)I.uR;
            sw[threadIdx.x][1] = I.kmag * dot(K, I.q.Y) / (double)I.vR;
            sw[threadIdx.x][2] = I.kmag * dot(K, af);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < nw * kMcTile; e += kThreads) {
            const int wl = e / kMcTile, c = e - wl * kMcTile;
            sa[wl][c] = cis_reduced(sw[wl][0] * (double)(u0 + c));
            const float2 b = cis_reduced(fma(sw[wl][1], (double)(v0 + c), sw[wl][2]));
            sb[wl][c] = make_float2((float)I.sA * b.x, (float)I.sA * b.y);
        }
        __syncthreads();
        for (int wl = 0; wl < nw; ++wl) {
            const float4 a01 = *reinterpret_cast<const float4*>(&sa[wl][4 * tx]);
            const float4 a23 = *reinterpret_cast<const float4*>(&sa[wl][4 * tx + 2]);
            const float4 b01 = *reinterpret_cast<const float4*>(&sb[wl][4 * ty]);
            const float4 b23 = *reinterpret_cast<const float4*>(&sb[wl][4 * ty + 2]);
            const float2 a[4] = {{a01.x, a01.y}, {a01.z, a01.w}, {a23.x, a23.y}, {a23.z, a23.w}};
            const float2 b[4] = {{b01.x, b01.y}, {b01.z, b01.w}, {b23.x, b23.y}, {b23.z, b23.w}};
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    ar[i][j] = fmaf(b[i].x, a[j].x, fmaf(-b[i].y, a[j].y, ar[i][j]));
                    ai[i][j] = fmaf(b[i].x, a[j].y, fmaf(b[i].y, a[j].x, ai[i][j]));
                }
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int iv = v0 + 4 * ty + i;
        if (iv >= I.vR) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int iu = u0 + 4 * tx + j;
            if (iu < I.uR) Uf[I.off + (long long)iv * I.uR + iu] = make_float2(ar[i][j], ai[i][j]);
        }
    }
}

// Grid: (tiles, Monte-Carlo instances of the batch).
__global__ __launch_bounds__(kThreads) void uf_mc_tile_kernel(const UfInst* __restrict__ insts,
                                                              const int* __restrict__ list,
                                                              const double* __restrict__ waves,
                                                              float2* __restrict__ Uf) {
    uf_mc_tile(insts[list[blockIdx.y]], (int)blockIdx.x, waves, Uf);
}

__global__ __launch_bounds__(kThreads) void uf_mc_one_kernel(UfInst I,
                                                             const double* __restrict__ waves,
                                                             float2* __restrict__ Uf) {
    uf_mc_tile(I, (int)blockIdx.x, waves, Uf);
}

// U = Uf (no spheres) or U += Uf outside every sphere (d < a is inside,
// nfSumUf.cu:31-48) for a batch of instances.  Spheres as (x, y, z, a).
template <class C, class R>
__global__ __launch_bounds__(kThreads) void sum_batch_kernel(const SumInst* __restrict__ insts, int n,
                                                             const double4* __restrict__ sph,
                                                             const C* __restrict__ Uf,
                                                             C* __restrict__ U, R* __restrict__ D,
                                                             R* __restrict__ Di) {
    for (int i = blockIdx.y; i < n; i += gridDim.y) {
        const SumInst& I = insts[i];
        const int npix = I.uR * I.vR;
        for (int pix = blockIdx.x * kThreads + threadIdx.x; pix < npix; pix += gridDim.x * kThreads) {
            if (pix < I.nd2) {
                D[I.image_off + pix] = R(0);
                Di[I.image_off + pix] = R(0);
            }
            const C b = Uf[I.off + pix];
            if (I.ns == 0) {
                U[I.off + pix] = b;
                continue;
            }
            int iu, iv;
            pixel_of(pix, I.uR, iu, iv);
            const d3 p = quad_at(I.q, (double)iu / (double)I.uR, (double)iv / (double)I.vR);
            bool in = false;
            for (int s = I.s_begin; s < I.s_begin + I.ns && !in; ++s) {
                const double4 S = sph[s];
                if (len(sub(p, d3{S.x, S.y, S.z})) < S.w) in = true;
            }
            if (in) continue;
            C a = U[I.off + pix];
            a.x = a.x + b.x;
            a.y = a.y + b.y;
            U[I.off + pix] = a;
        }
    }
}

struct MaskBatch {
    double p[32][3];
    double a[32];
};

// Marks pixels inside any sphere of the batch (d < a, nfSumUf.cu:31-39).
__global__ __launch_bounds__(kThreads) void inside_mask_kernel(QuadArg q, int uR, int vR,
                                                               MaskBatch b, int cnt, int first,
                                                               unsigned char* __restrict__ mask) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= uR * vR) return;
    int iu, iv;
    pixel_of(pix, uR, iu, iv);
    const d3 p = quad_at(q, (double)iu / (double)uR, (double)iv / (double)vR);
    unsigned char in = first ? 0 : mask[pix];
    for (int s = 0; s < cnt && !in; ++s)
        if (len(sub(p, d3{b.p[s][0], b.p[s][1], b.p[s][2]})) < b.a[s]) in = 1;
    mask[pix] = in;
}

template <class C>
__global__ void add_outside_kernel(int npix, const unsigned char* __restrict__ mask,
                                   const C* __restrict__ Uf, C* __restrict__ U) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= npix || (mask && mask[pix])) return;
    C a = U[pix];
    const C b = Uf[pix];
    a.x = a.x + b.x;
    a.y = a.y + b.y;
    U[pix] = a;
}

}  // namespace

void uf_fast_table(double ca, double cb, int m, std::vector<double>* out, bool flat) {
    const size_t base = out->size();
    out->resize(base + 4 * (size_t)(m + 1), 0.0);
    double* T = out->data() + base;
    // P_{n}(x) from P_{n-1}, P_{n-2} as ref::shift_legendre (legendre.h:20)
    auto leg = [](int n, double x, double& P0, double& P1) {
        const double Pn = ((double)(2 * n - 1) * x * P1 - (double)(n - 1) * P0) / (double)n;
        P0 = P1;
        P1 = Pn;
    };
    double Pa[3] = {1.0, ca, 1.0}, Pb[3] = {1.0, cb, 1.0};
    for (int l = 0; l <= m; ++l) {
        if (l >= 1) {
            Pa[2] = Pa[0];
            leg(l + 1, ca, Pa[0], Pa[1]);
            Pb[2] = Pb[0];
            leg(l + 1, cb, Pb[0], Pb[1]);
        }
        const double D = ((Pa[1] - Pa[2]) - Pb[1]) + Pb[2];
        T[4 * l] = (l & 2) ? -D : D;   // i^l = 1, i, -1, -i
        if (l > 0) {
            T[4 * l + 1] = (2.0 * l - 1.0) / l;
            T[4 * l + 2] = (l - 1.0) / l;
        }
        T[4 * l + 3] = 2.0 * l - 1.0;
    }
    if (flat) {
        double P = 1.0;
        for (int l = 2; l <= m; l += 2) {
            P = -(T[4 * l + 2] * P);
            T[4 * l] = P * T[4 * l];
        }
    }
}

int uf_batch_launch(bimsim_ctx* ctx, const UfInst* d_inst, int n, int items, int f32,
                    const double* tabs, const double* waves, void* Uf, const SumInst* sum,
                    void* D, void* Di) {
    // four work items a thread (the launch strides over them): a block's
    // items are short, and fewer blocks spend less on starting them
    const dim3 grid((items + 4 * kThreads - 1) / (4 * kThreads), std::min(n, 65535));
    if (f32)
        hipLaunchKernelGGL((uf_batch_kernel<float2, float>), grid, dim3(kThreads), 0, ctx->stream,
                           d_inst, n, tabs, waves, (float2*)Uf, sum, (float*)D, (float*)Di);
    else
        hipLaunchKernelGGL((uf_batch_kernel<double2, double>), grid, dim3(kThreads), 0, ctx->stream,
                           d_inst, n, tabs, waves, (double2*)Uf, sum, (double*)D, (double*)Di);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int uf_mc_batch_launch(bimsim_ctx* ctx, const UfInst* d_inst, const int* list, int n,
                       int max_tiles, const double* waves, float2* Uf) {
    if (n <= 0) return BIMSIM_OK;
    if (n > 65535) return set_error(ctx, BIMSIM_E_UNSUPPORTED, "too many Monte-Carlo instances");
    hipLaunchKernelGGL(uf_mc_tile_kernel, dim3(max_tiles, n), dim3(kThreads), 0, ctx->stream,
                       d_inst, list, waves, Uf);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int mc_tiles(int uR, int vR) {
    return ((uR + kMcTile - 1) / kMcTile) * ((vR + kMcTile - 1) / kMcTile);
}

int sum_batch_launch(bimsim_ctx* ctx, const SumInst* d_inst, int n, int max_npix,
                     const double4* d_sph, int f32, const void* Uf, void* U, void* D, void* Di) {
    const dim3 grid((max_npix + kThreads - 1) / kThreads, std::min(n, 65535));
    if (f32)
        hipLaunchKernelGGL((sum_batch_kernel<float2, float>), grid, dim3(kThreads), 0, ctx->stream,
                           d_inst, n, d_sph, (const float2*)Uf, (float2*)U, (float*)D, (float*)Di);
    else
        hipLaunchKernelGGL((sum_batch_kernel<double2, double>), grid, dim3(kThreads), 0, ctx->stream,
                           d_inst, n, d_sph, (const double2*)Uf, (double2*)U, (double*)D,
                           (double*)Di);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // namespace bimsim

using namespace bimsim;

extern "C" int bimsim_nf_uf(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                            const double focus[3], const double kdir[3], double kmag,
                            double amp, double cos_alpha, double cos_beta, int m, int mode,
                            int prec, int quirks, void* Uf) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, pos && focus && kdir && Uf, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 28),
                     "slice resolution must be positive and at most 2^28 pixels");
    BIMSIM_CHECK_ARG(ctx, prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST, "unknown precision");
    BIMSIM_CHECK_ARG(ctx, m >= 1 && m <= 100000, "incident order m must be in [1, 100000]");
    BIMSIM_CHECK_ARG(ctx, kmag > 0.0 && std::isfinite(kmag), "kmag must be positive");
    BIMSIM_CHECK_ARG(ctx, mode == BIMSIM_UF_SERIES || mode == BIMSIM_UF_PLANE || mode == BIMSIM_UF_LUT,
                     "unknown Uf mode");
    if (mode == BIMSIM_UF_LUT) {
        // nearfieldStruct::scalarUfLut (nfScalarUfLut.cu:125-188): table over
        // [d_min, d_max] from the focus with 2 max(uR, vR) samples
        double dmin, dmax;
        quad_dist(to_quad(*pos), mk(focus), &dmin, &dmax);
        const int dR = 2 * std::max(uR, vR);
        const size_t esz = prec == BIMSIM_FP32_FAST ? sizeof(float) : sizeof(double);
        void* lut = nullptr;
        BIMSIM_TRY(scratch_get(ctx, kSlotUfLut, esz * (size_t)dR * (m + 1), &lut));
        BIMSIM_TRY(bimsim_uf_bessel_lut(ctx, m, kmag, dmin, dmax, dR, prec, lut));
        return bimsim_nf_uf_lut(ctx, pos, uR, vR, focus, kdir, amp, cos_alpha, cos_beta, m, lut, dR,
                                dmin, dmax, prec, Uf);
    }
    if (!quirks && mode == BIMSIM_UF_SERIES)
        return set_error(ctx, BIMSIM_E_UNSUPPORTED,
                         "incident-field series without the reference recurrences is not built");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int npix = uR * vR;
    const dim3 grid((npix + kThreads - 1) / kThreads);
    const QuadArg q = to_quad(*pos);
    const d3 f = mk(focus), k = mk(kdir);
    const int tmr = timer_begin(ctx, kStageUf);
    const bool f32 = prec == BIMSIM_FP32_FAST;
    if (mode == BIMSIM_UF_PLANE) {
        if (f32)
            hipLaunchKernelGGL(uf_plane_kernel<float2>, grid, dim3(kThreads), 0, ctx->stream, q, uR,
                               vR, f, k, kmag, amp, (float2*)Uf);
        else
            hipLaunchKernelGGL(uf_plane_kernel<double2>, grid, dim3(kThreads), 0, ctx->stream, q,
                               uR, vR, f, k, kmag, amp, (double2*)Uf);
    } else {
        if (f32) {
            std::vector<double> tab;
            uf_fast_table(cos_alpha, cos_beta, m, &tab, uf_flat(q, f, k));
            const double* T = nullptr;
            BIMSIM_TRY(dbl_table_get(ctx, tab, &T));
            hipLaunchKernelGGL(uf_fast_kernel, grid, dim3(kThreads), 0, ctx->stream, q, uR, vR, f, k,
                               kmag, amp, cos_alpha, cos_beta, m, T, uf_flat(q, f, k) ? 1 : 0,
                               (float2*)Uf);
        } else
            hipLaunchKernelGGL(uf_series_kernel<double2>, grid, dim3(kThreads), 0, ctx->stream, q,
                               uR, vR, f, k, kmag, amp, cos_alpha, cos_beta, m, (double2*)Uf);
    }
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    timer_end(ctx, tmr);
    return BIMSIM_OK;
}

extern "C" int bimsim_nf_uf_mc(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                               const double focus[3], double kmag, const double* waves, int nk,
                               double subA, int prec, void* Uf) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, pos && focus && waves && Uf, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 28),
                     "slice resolution must be positive and at most 2^28 pixels");
    BIMSIM_CHECK_ARG(ctx, nk >= 1, "the Monte-Carlo sum needs at least one wave");
    BIMSIM_CHECK_ARG(ctx, prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST, "unknown precision");
    BIMSIM_CHECK_ARG(ctx, kmag > 0.0 && std::isfinite(kmag), "kmag must be positive");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const QuadArg q = to_quad(*pos);
    const d3 f = mk(focus);
    // the weight of every wave, (1.0f/nk) subA as the scattered field
    const double sA = (double)(1.0f / (float)nk) * subA;
    const int tmr = timer_begin(ctx, kStageUf);
    if (prec == BIMSIM_FP32_FAST) {
        UfInst I{};
        I.q = q;
        I.f = f;
        I.k = d3{0.0, 0.0, 1.0};
        I.kmag = kmag;
        I.uR = uR;
        I.vR = vR;
        I.mode = BIMSIM_UF_MC;
        I.off = 0;
        I.nk = nk;
        I.wave_off = 0;
        I.sA = sA;
        hipLaunchKernelGGL(uf_mc_one_kernel, dim3(mc_tiles(uR, vR)), dim3(kThreads), 0, ctx->stream,
                           I, waves, (float2*)Uf);
        BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    } else {
        const dim3 grid((uR * vR + kThreads - 1) / kThreads);
        hipLaunchKernelGGL(uf_mc_kernel<double2>, grid, dim3(kThreads), 0, ctx->stream, q, uR, vR,
                           f, kmag, waves, nk, sA, (double2*)Uf);
        BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    }
    timer_end(ctx, tmr);
    return BIMSIM_OK;
}

extern "C" int bimsim_nf_sum_uf(bimsim_ctx* ctx, const bimsim_quad* pos, int uR, int vR,
                                const bimsim_sphere* spheres, int ns, int prec, const void* Uf,
                                void* U) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, pos && Uf && U, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && (long long)uR * vR <= (1LL << 28),
                     "slice resolution must be positive and at most 2^28 pixels");
    BIMSIM_CHECK_ARG(ctx, ns >= 0 && (ns == 0 || spheres), "bad sphere list");
    BIMSIM_CHECK_ARG(ctx, prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST, "unknown precision");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int npix = uR * vR;
    const bool f32 = prec == BIMSIM_FP32_FAST;
    const size_t esz = f32 ? sizeof(float2) : sizeof(double2);
    if (ns == 0) {
        // U = Uf (nfSumUf.cu:59-68)
        BIMSIM_HIP_CHECK(ctx, hipMemcpyAsync(U, Uf, esz * npix, hipMemcpyDeviceToDevice, ctx->stream));
        return BIMSIM_OK;
    }
    void* mask = nullptr;
    BIMSIM_TRY(scratch_get(ctx, kSlotMisc, (size_t)npix, &mask));
    const QuadArg q = to_quad(*pos);
    const dim3 grid((npix + kThreads - 1) / kThreads);
    for (int s0 = 0; s0 < ns; s0 += 32) {
        const int cnt = std::min(32, ns - s0);
        MaskBatch b{};
        for (int i = 0; i < cnt; ++i) {
            for (int c = 0; c < 3; ++c) b.p[i][c] = spheres[s0 + i].p[c];
            b.a[i] = spheres[s0 + i].a;
        }
        hipLaunchKernelGGL(inside_mask_kernel, grid, dim3(kThreads), 0, ctx->stream, q, uR, vR, b,
                           cnt, s0 == 0 ? 1 : 0, (unsigned char*)mask);
    }
    if (f32)
        hipLaunchKernelGGL(add_outside_kernel<float2>, grid, dim3(kThreads), 0, ctx->stream, npix,
                           (const unsigned char*)mask, (const float2*)Uf, (float2*)U);
    else
        hipLaunchKernelGGL(add_outside_kernel<double2>, grid, dim3(kThreads), 0, ctx->stream, npix,
                           (const unsigned char*)mask, (const double2*)Uf, (double2*)U);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}
