// FP32 scattered-field series (nfScalarUs hot loop, reference
// nfScalarUs.cu:449-491 with calc_Us/calc_Ui :50-95,201-245), evaluated as
//
//   U(p) += sum_s sum_w phi_{s,w} sum_l c'_{s,l}(p) Q_l(x_{s,w}(p))
//
// with c'_l = s_l * B_l h_l(kd) (or s_l * A_l j_l(knd) inside the sphere)
// computed once per (pixel, sphere) in FP64 by nf_us64.hip, and the
// rescaled Legendre polynomials Q_l = P_l / s_l obeying
// Q_l = 2x Q_{l-1} - g_l Q_{l-2}.  Per term that is one multiply and three
// FMAs with the recurrence constants g_l folded into the instruction stream
// (the l loop is unrolled, c'_l live in VGPRs).  One thread per pixel; all
// lanes of a wavefront walk the same sphere, wave and order, so the wave
// records are scalar loads.  Waves are summed in chunks of 8 whose partial
// sums enter a Kahan-compensated accumulator.
#include "nf_us.h"

namespace bimsim {
namespace {

constexpr int kThreads = 256;
constexpr int kW = 2;   // waves in flight per thread

// One step of the series for order l on kW waves (l >= 2).
template <int l>
__device__ __forceinline__ void series_step(const float* cr, const float* ci, const float* x2,
                                            float* q0, float* q1, float* sr, float* si) {
    const float g = -(float)legq_g(l);
#pragma unroll
    for (int u = 0; u < kW; ++u) {
        const float qn = fmaf(g, q0[u], x2[u] * q1[u]);
        q0[u] = q1[u];
        q1[u] = qn;
        sr[u] = fmaf(cr[l], qn, sr[u]);
        si[u] = fmaf(ci[l], qn, si[u]);
    }
}

// Step j (0..7) of the block starting at order l0, when that order is >= 2
// (orders 0 and 1 are folded into the initial sum).
template <int l0, int j>
__device__ __forceinline__ void block_step(const float* cr, const float* ci, const float* x2,
                                           float* q0, float* q1, float* sr, float* si) {
    if constexpr (l0 + j >= 2) series_step<l0 + j>(cr, ci, x2, q0, q1, sr, si);
}

// Orders in blocks of 8: blocks below nfull run whole, block nfull runs its
// first rem orders, each behind a wavefront-uniform test.
template <int L, int B = 0>
__device__ __forceinline__ void series_blocks(const float* cr, const float* ci, const float* x2,
                                              float* q0, float* q1, float* sr, float* si,
                                              int nfull, int rem) {
    if constexpr (B < L / 8) {
        constexpr int l0 = 8 * B;
        if (B < nfull) {
            block_step<l0, 0>(cr, ci, x2, q0, q1, sr, si);
            block_step<l0, 1>(cr, ci, x2, q0, q1, sr, si);
            block_step<l0, 2>(cr, ci, x2, q0, q1, sr, si);
            block_step<l0, 3>(cr, ci, x2, q0, q1, sr, si);
            block_step<l0, 4>(cr, ci, x2, q0, q1, sr, si);
            block_step<l0, 5>(cr, ci, x2, q0, q1, sr, si);
            block_step<l0, 6>(cr, ci, x2, q0, q1, sr, si);
            block_step<l0, 7>(cr, ci, x2, q0, q1, sr, si);
            series_blocks<L, B + 1>(cr, ci, x2, q0, q1, sr, si, nfull, rem);
        } else if (B == nfull) {
            if (0 < rem) block_step<l0, 0>(cr, ci, x2, q0, q1, sr, si);
            if (1 < rem) block_step<l0, 1>(cr, ci, x2, q0, q1, sr, si);
            if (2 < rem) block_step<l0, 2>(cr, ci, x2, q0, q1, sr, si);
            if (3 < rem) block_step<l0, 3>(cr, ci, x2, q0, q1, sr, si);
            if (4 < rem) block_step<l0, 4>(cr, ci, x2, q0, q1, sr, si);
            if (5 < rem) block_step<l0, 5>(cr, ci, x2, q0, q1, sr, si);
            if (6 < rem) block_step<l0, 6>(cr, ci, x2, q0, q1, sr, si);
        }
    }
}

template <int L>
__global__ __launch_bounds__(kThreads) void us_series_f32(
    const float2* __restrict__ cp, const float4* __restrict__ rh, const WaveF* __restrict__ wt,
    SeriesMeta meta, int ns, int nk_pad, int wchunk, int npix, float2* __restrict__ out,
    int accumulate, int nsplit) {
    const int pix = blockIdx.x * kThreads + threadIdx.x;
    if (pix >= npix) return;
    const int w_begin = blockIdx.y * wchunk;
    const int w_end = min(nk_pad, w_begin + wchunk);
    float acc_r = 0.f, acc_i = 0.f, cmp_r = 0.f, cmp_i = 0.f;
    for (int s = 0; s < ns; ++s) {
        const int nb = meta.nb[s];
        const int nl1 = meta.nl1[s];        // Nl + 1 orders
        const int nfull = nl1 >> 3, rem = nl1 & 7;
        float cr[L], ci[L];
        const float2* c = cp + meta.coff[s] + pix;
#pragma unroll
        for (int b = 0; b < L / 8; ++b) {
            if (b < nb) {
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float2 v = c[(size_t)(8 * b + j) * npix];
                    cr[8 * b + j] = v.x;
                    ci[8 * b + j] = v.y;
                }
            } else {
#pragma unroll
                for (int j = 0; j < 8; ++j) cr[8 * b + j] = ci[8 * b + j] = 0.f;
            }
        }
        const float4 r = rh[meta.roff[s] + pix];
        const WaveF* W = wt + meta.woff[s];
        float ch_r = 0.f, ch_i = 0.f;
        int in_chunk = 0;
        if (w_begin >= w_end) continue;
        WaveF nxt[kW];
#pragma unroll
        for (int u = 0; u < kW; ++u) nxt[u] = W[w_begin + u];
        for (int w = w_begin; w < w_end; w += kW) {
            float x2[kW], q0[kW], q1[kW], sr[kW], si[kW], pr[kW], pi[kW];
#pragma unroll
            for (int u = 0; u < kW; ++u) {
                const WaveF e = nxt[u];
                x2[u] = e.kx2 * r.x + e.ky2 * r.y + e.kz2 * r.z;
                pr[u] = e.pr;
                pi[u] = e.pi;
                q0[u] = 1.f;
                q1[u] = x2[u];
                sr[u] = fmaf(cr[1], x2[u], cr[0]);
                si[u] = fmaf(ci[1], x2[u], ci[0]);
            }
            // prefetch the next wave records (the last iteration re-reads its own)
            const int wn = w + kW < w_end ? w + kW : w;
#pragma unroll
            for (int u = 0; u < kW; ++u) nxt[u] = W[wn + u];
            series_blocks<L>(cr, ci, x2, q0, q1, sr, si, nfull, rem);
#pragma unroll
            for (int u = 0; u < kW; ++u) {
                ch_r = fmaf(pr[u], sr[u], ch_r);
                ch_r = fmaf(-pi[u], si[u], ch_r);
                ch_i = fmaf(pr[u], si[u], ch_i);
                ch_i = fmaf(pi[u], sr[u], ch_i);
            }
            in_chunk += kW;
            if (in_chunk >= 8) {
                const float yr = ch_r - cmp_r, yi = ch_i - cmp_i;
                const float tr = acc_r + yr, ti = acc_i + yi;
                cmp_r = (tr - acc_r) - yr;
                cmp_i = (ti - acc_i) - yi;
                acc_r = tr;
                acc_i = ti;
                ch_r = ch_i = 0.f;
                in_chunk = 0;
            }
        }
        {
            const float yr = ch_r - cmp_r, yi = ch_i - cmp_i;
            const float tr = acc_r + yr, ti = acc_i + yi;
            cmp_r = (tr - acc_r) - yr;
            cmp_i = (ti - acc_i) - yi;
            acc_r = tr;
            acc_i = ti;
        }
    }
    if (nsplit == 1) {
        float2 o = make_float2(0.f, 0.f);
        if (accumulate) o = out[pix];
        out[pix] = make_float2(o.x + acc_r, o.y + acc_i);
    } else {
        out[(size_t)blockIdx.y * npix + pix] = make_float2(acc_r, acc_i);
    }
}

// U[pix] (+)= sum over splits of partial[split][pix], in split order.
__global__ void us_reduce_splits(const float2* __restrict__ part, int nsplit, int npix,
                                 float2* __restrict__ U, int accumulate) {
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    float sr = 0.f, si = 0.f;
    for (int k = 0; k < nsplit; ++k) {
        const float2 v = part[(size_t)k * npix + pix];
        sr += v.x;
        si += v.y;
    }
    float2 o = make_float2(0.f, 0.f);
    if (accumulate) o = U[pix];
    U[pix] = make_float2(o.x + sr, o.y + si);
}

template <int L>
void launch_L(dim3 grid, hipStream_t st, const float2* cp, const float4* rh, const WaveF* wt,
              const SeriesMeta& meta, int ns, int nk_pad, int wchunk, int npix, float2* out,
              int accumulate, int nsplit) {
    hipLaunchKernelGGL(us_series_f32<L>, grid, dim3(kThreads), 0, st, cp, rh, wt, meta, ns,
                       nk_pad, wchunk, npix, out, accumulate, nsplit);
}

}  // namespace

int launch_us_series_f32(bimsim_ctx* ctx, int L, const float2* cp, const float4* rh,
                         const WaveF* wt, const SeriesMeta& meta, int ns, int nk_pad, int npix,
                         float2* U, int accumulate, int first_batch) {
    (void)first_batch;
    const int pix_blocks = (npix + kThreads - 1) / kThreads;
    // Split the wave range across blockIdx.y until the grid holds several
    // workgroups per CU; each split keeps at least 32 waves.
    const int target = 4 * ctx->cu_count;
    int nsplit = 1;
    while (pix_blocks * nsplit < target && nk_pad / (2 * nsplit) >= 32) nsplit *= 2;
    int wchunk = (nk_pad + nsplit - 1) / nsplit;
    wchunk = (wchunk + kW - 1) / kW * kW;
    nsplit = (nk_pad + wchunk - 1) / wchunk;
    float2* out = U;
    if (nsplit > 1) {
        void* part = nullptr;
        BIMSIM_TRY(scratch_get(ctx, kSlotPartial, sizeof(float2) * (size_t)nsplit * npix, &part));
        out = (float2*)part;
    }
    dim3 grid(pix_blocks, nsplit);
    const int tmr = timer_begin(ctx, kStageSeries);
    switch (L) {
#define BIMSIM_CASE(LL) \
    case LL: launch_L<LL>(grid, ctx->stream, cp, rh, wt, meta, ns, nk_pad, wchunk, npix, out, accumulate, nsplit); break;
        BIMSIM_CASE(16)
        BIMSIM_CASE(24)
        BIMSIM_CASE(32)
        BIMSIM_CASE(40)
        BIMSIM_CASE(48)
        BIMSIM_CASE(64)
        BIMSIM_CASE(80)
        BIMSIM_CASE(96)
        BIMSIM_CASE(112)
#undef BIMSIM_CASE
        default:
            return set_error(ctx, BIMSIM_E_UNSUPPORTED, "no FP32 series kernel for %d orders", L);
    }
    timer_end(ctx, tmr);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    if (nsplit > 1) {
        hipLaunchKernelGGL(us_reduce_splits, dim3(pix_blocks), dim3(kThreads), 0, ctx->stream,
                           (const float2*)out, nsplit, npix, U, accumulate);
        BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    }
    return BIMSIM_OK;
}

}  // namespace bimsim
