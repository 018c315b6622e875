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
template <class Out>
__global__ __launch_bounds__(kThreads) void uf_batch_kernel(const UfInst* __restrict__ insts, int n,
                                                            const double* __restrict__ tabs,
                                                            Out* __restrict__ Uf) {
    for (int i = blockIdx.y; i < n; i += gridDim.y) {
        const UfInst& I = insts[i];
        const int npix = I.uR * I.vR;
        for (int pix = blockIdx.x * kThreads + threadIdx.x; pix < npix; pix += gridDim.x * kThreads)
            uf_batch_pixel(I, pix, tabs, Uf);
    }
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

void uf_fast_table(double ca, double cb, int m, std::vector<double>* out) {
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
}

int uf_batch_launch(bimsim_ctx* ctx, const UfInst* d_inst, int n, int max_npix, int f32,
                    const double* tabs, void* Uf) {
    const dim3 grid((max_npix + kThreads - 1) / kThreads, std::min(n, 65535));
    if (f32)
        hipLaunchKernelGGL(uf_batch_kernel<float2>, grid, dim3(kThreads), 0, ctx->stream, d_inst, n,
                           tabs, (float2*)Uf);
    else
        hipLaunchKernelGGL(uf_batch_kernel<double2>, grid, dim3(kThreads), 0, ctx->stream, d_inst,
                           n, tabs, (double2*)Uf);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
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
            uf_fast_table(cos_alpha, cos_beta, m, &tab);
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
