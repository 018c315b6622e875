// FP32 scattered-field series (nfScalarUs hot loop, reference
// nfScalarUs.cu:449-491 with calc_Us/calc_Ui :50-95,201-245), evaluated as
//
//   U(p) += sum_s sum_w phi_{s,w} sum_l c'_{s,l}(p) Q_l(x_{s,w}(p))
//
// with c'_l = s_l * B_l h_l(kd) (or s_l * A_l j_l(knd) inside the sphere),
// computed once per (pixel, sphere) in FP64 with the reference recurrences,
// and the rescaled Legendre polynomials Q_l = P_l / s_l obeying
// Q_l = 2x Q_{l-1} - g_l Q_{l-2}, s_l = (2l)!/(4^l (l!)^2).  Per term: one
// multiply and three FMAs with g_l as an instruction literal (the order loop
// is unrolled, c'_l live in VGPRs).  All lanes of a wavefront walk the same
// sphere, wave and order, so wave records are scalar loads.  Waves are
// summed in chunks of 8 whose sums enter a Kahan accumulator.
//
// Built with -ffp-contract=off so that the FP64 recurrences round like the
// reference's; every FP32 FMA of the series is written out with fmaf.
#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "mie_device.h"
#include "refmath.h"
#include "uf_device.h"
#include "us_series.h"

namespace bimsim {
namespace {

constexpr int kThreads = 256;
constexpr int kW = 2;   // waves in flight per thread
constexpr int kBuckets[] = {16, 24, 32, 40, 48, 64, 80, 96, 112};

__device__ inline dc ld2(const double2* p, int i) {
    const double2 v = p[i];
    return dc{v.x, v.y};
}

// Phase 1: c'_l(p) (orders padded with zeros to 8 nb) and r-hat(p) for
// every sphere of a chunk.  Grid: (pixel blocks, spheres of the chunk).
__global__ __launch_bounds__(kThreads) void series_coef_kernel(
    const SeriesInst* __restrict__ insts, const SeriesSph* __restrict__ sph, int sph_begin,
    int sph_count, const double2* __restrict__ A, const double2* __restrict__ B,
    float2* __restrict__ cp, float4* __restrict__ rh) {
    for (int si = blockIdx.y; si < sph_count; si += gridDim.y) {
        const SeriesSph& SS = sph[sph_begin + si];
        const SeriesInst& I = insts[SS.inst];
        const SphereDev S = SS.s;
        const int npix = I.uR * I.vR;
        for (int pix = blockIdx.x * kThreads + threadIdx.x; pix < npix;
             pix += gridDim.x * kThreads) {
            const int iv = pix / I.uR, iu = pix - iv * I.uR;
            const d3 p = quad_at(I.q, (double)iu / (double)I.uR, (double)iv / (double)I.vR);
            d3 r = sub(p, mk(S.p));
            const double d = len(r);
            r = norm(r);
            float2* out = cp + SS.coff + pix;
            const double2* As = A + S.coeff_off;
            const double2* Bs = B + S.coeff_off;
            const int Nl = S.Nl;
            double sl = 1.0;   // s_l = (2l)! / (4^l (l!)^2)
            auto put = [&](int l, dc c) {
                if (l > 0) sl = sl * (2.0 * l - 1.0) / (2.0 * l);
                out[(size_t)l * npix] = make_float2((float)(c.r * sl), (float)(c.i * sl));
            };
            bool centre = false;
            if (d <= S.a) {
                const dc knd = cscl(dc{S.n[0], S.n[1]}, I.kmag * d);
                if (knd.r < (double)kEpsFloat) {
                    // calc_Ui returns A_0 for every wave: the series reduces to c'_0
                    put(0, ld2(As, 0));
                    for (int l = 1; l <= Nl; ++l) put(l, dc{0.0, 0.0});
                    centre = true;
                } else {
                    dc j[2];
                    ref::init_jc(knd, j);
                    put(0, cmul(ld2(As, 0), j[0]));
                    if (Nl >= 1) put(1, cmul(ld2(As, 1), j[1]));
                    for (int l = 2; l <= Nl; ++l) {
                        ref::shift_jc(l, knd, j);
                        put(l, cmul(ld2(As, l), j[1]));
                    }
                }
            } else {
                const double kd = I.kmag * d;
                double j[2], y[2];
                ref::init_j(kd, j);
                ref::init_y(kd, y);
                put(0, cmul(ld2(Bs, 0), dc{j[0], y[0]}));
                if (Nl >= 1) put(1, cmul(ld2(Bs, 1), dc{j[1], y[1]}));
                for (int l = 2; l <= Nl; ++l) {
                    ref::shift_j(l, kd, j);
                    ref::shift_y(l, kd, y);
                    put(l, cmul(ld2(Bs, l), dc{j[1], y[1]}));
                }
            }
            for (int l = Nl + 1; l < 8 * SS.nb; ++l) out[(size_t)l * npix] = make_float2(0.f, 0.f);
            if (centre) r = d3{0.0, 0.0, 0.0};
            rh[SS.roff + pix] = make_float4((float)r.x, (float)r.y, (float)r.z, 0.f);
        }
    }
}

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

// Phase 2: persistent work-queue kernel over the items of one bucket.
template <int L>
__global__ __launch_bounds__(kThreads) void series_f32_kernel(
    const SeriesInst* __restrict__ insts, const SeriesSph* __restrict__ sph,
    const SeriesItem* __restrict__ items, int n_items, const float2* __restrict__ cp,
    const float4* __restrict__ rh, const WaveF* __restrict__ wt, float2* __restrict__ U,
    float2* __restrict__ part, int* __restrict__ counter) {
    __shared__ int s_item;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= n_items) break;
        const SeriesItem item = items[it];
        const SeriesInst& I = insts[item.inst];
        const int npix = I.uR * I.vR;
        const int pix = item.pix0 + (int)threadIdx.x;
        if (pix >= npix) continue;
        const int w_begin = item.split * I.wchunk;
        const int w_end = min(I.nk_pad, w_begin + I.wchunk);
        float acc_r = 0.f, acc_i = 0.f, cmp_r = 0.f, cmp_i = 0.f;
        for (int s = I.s_begin; s < I.s_begin + I.ns; ++s) {
            const SeriesSph& SS = sph[s];
            const int nb = SS.nb;
            const int nl1 = SS.s.Nl + 1;
            const int nfull = nl1 >> 3, rem = nl1 & 7;
            float cr[L], ci[L];
            const float2* c = cp + SS.coff + pix;
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
            const float4 r = rh[SS.roff + pix];
            const WaveF* W = wt + SS.woff;
            if (w_begin >= w_end) continue;
            float ch_r = 0.f, ch_i = 0.f;
            int in_chunk = 0;
            WaveF nxt[kW];
#pragma unroll
            for (int u = 0; u < kW; ++u) nxt[u] = W[w_begin + u];
            for (int w = w_begin; w < w_end; w += kW) {
                float x2[kW], q0[kW], q1[kW], sr[kW], si[kW], pr[kW], pi[kW];
#pragma unroll
                for (int u = 0; u < kW; ++u) {
                    const WaveF e = nxt[u];
                    x2[u] = fmaf(e.kz2, r.z, fmaf(e.ky2, r.y, e.kx2 * r.x));
                    pr[u] = e.pr;
                    pi[u] = e.pi;
                    q0[u] = 1.f;
                    q1[u] = x2[u];
                    sr[u] = fmaf(cr[1], x2[u], cr[0]);
                    si[u] = fmaf(ci[1], x2[u], ci[0]);
                }
                // prefetch the next records (the last iteration re-reads its own)
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
                    kahan(acc_r, acc_i, cmp_r, cmp_i, ch_r, ch_i);
                    ch_r = ch_i = 0.f;
                    in_chunk = 0;
                }
            }
            kahan(acc_r, acc_i, cmp_r, cmp_i, ch_r, ch_i);
        }
        if (I.nsplit == 1) {
            float2* o = U + I.u_off + pix;
            float2 v = make_float2(0.f, 0.f);
            if (I.accumulate) v = *o;
            *o = make_float2(v.x + acc_r, v.y + acc_i);
        } else {
            part[I.part_off + (long long)item.split * npix + pix] = make_float2(acc_r, acc_i);
        }
    }
}

// WaveF records: 2k and (1/nk) subA exp(i kmag k.c), c = p_s - f (FP64).
__global__ void series_waves_kernel(const SeriesInst* __restrict__ insts,
                                    const SeriesSph* __restrict__ sph, int n_sph,
                                    const double* __restrict__ waves, WaveF* __restrict__ wt,
                                    int* __restrict__ counters, int n_counters) {
    if (blockIdx.x == 0 && blockIdx.y == 0)   // the work-queue counters of the sums
        for (int i = threadIdx.x; i < n_counters; i += blockDim.x) counters[i] = 0;
    for (int s = blockIdx.y; s < n_sph; s += gridDim.y) {
        const SeriesSph& SS = sph[s];
        const SeriesInst& I = insts[SS.inst];
        for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < I.nk_pad;
             w += gridDim.x * blockDim.x) {
            WaveF e = {};
            if (w < I.nk) {
                const d3 k = mk(waves + 3 * (I.wave_off + w));
                const d3 c = sub(mk(SS.s.p), I.f);
                const dc ph = cscl(cexp_i(I.kmag * dot(k, c)), I.sA);
                e.kx2 = (float)(2.0 * k.x);
                e.ky2 = (float)(2.0 * k.y);
                e.kz2 = (float)(2.0 * k.z);
                e.pr = (float)ph.r;
                e.pi = (float)ph.i;
            }
            wt[SS.woff + w] = e;
        }
    }
}

// U (+)= sum over splits, in split order, for the listed instances.
__global__ void series_reduce_kernel(const SeriesInst* __restrict__ insts,
                                     const int* __restrict__ list, int n,
                                     const float2* __restrict__ part, float2* __restrict__ U) {
    for (int k = blockIdx.y; k < n; k += gridDim.y) {
        const SeriesInst& I = insts[list[k]];
        const int npix = I.uR * I.vR;
        for (int pix = blockIdx.x * blockDim.x + threadIdx.x; pix < npix;
             pix += gridDim.x * blockDim.x) {
            float sr = 0.f, si = 0.f;
            for (int t = 0; t < I.nsplit; ++t) {
                const float2 v = part[I.part_off + (long long)t * npix + pix];
                sr += v.x;
                si += v.y;
            }
            float2* o = U + I.u_off + pix;
            float2 v = make_float2(0.f, 0.f);
            if (I.accumulate) v = *o;
            *o = make_float2(v.x + sr, v.y + si);
        }
    }
}

// ----------------------------------------------------------- harmonic form
// By the addition theorem P_l(k.r) = 4 pi/(2l+1) sum_m Y_lm(k) Y*_lm(r), the
// wave sum moves onto the sphere:
//   S(p) = sum_l c_l(p) sum_{m=0..l} Pn_lm(cos th_r) G_lm(phi_r),
//   G_lm = e^{-i m phi_r} T+_lm + e^{i m phi_r} T-_lm  (m > 0),  G_l0 = T_l0,
//   T+-_lm = 4 pi/(2l+1) sum_w phi_w Pn_lm(cos th_w) e^{+-i m phi_w},
// Pn_lm the orthonormal associated Legendre functions (the sign convention
// cancels in each product).  T is formed once per sphere in FP64 and kept as
// U1 = T+ + T-, U2 = (Im T+ - Im T-, Re T- - Re T+) (U1 = T, U2 = 0 at
// m = 0), times the scale K_lm of the pixel kernel's recurrence.  The
// radial factors c_l are formed per pixel inside the pixel kernel, so this
// form keeps no per-pixel coefficient buffer.
constexpr int kCS = 112;   // row stride of the recurrence-constant table (m, l < 112)
// FP64 tables of the wave sums (host-built, so the device loops carry no
// square roots): a_lm, b_lm at h64[2 (m kH + l)], sqrt((2k+1)/(2k)) at
// h64[kHDiag + k], K_lm = prod_{j=m+1..l} a_jm / 2 at h64[kHK + m kH + l].
constexpr int kH = 113;
constexpr int kHDiag = 2 * kH * kH;
constexpr int kHK = kHDiag + kH;
constexpr int kHSize = kHK + kH * kH;

// Orthonormal Pn_lm(x) for l = m .. lh-1 of one wave into out[l - m]
// (FP64): Pn_00 = 1/sqrt(4 pi), Pn_mm = sqrt((2m+1)/(2m)) s Pn_{m-1,m-1},
// Pn_lm = a_lm (x Pn_{l-1,m} - b_lm Pn_{l-2,m}); constants from h64.
__device__ inline void legendre_column(int m, int lh, double x, double* out, long long stride,
                                       const double* __restrict__ h64) {
    const double st = sqrt(fmax(0.0, 1.0 - x * x));
    double pmm = 0.28209479177387814;   // 1/sqrt(4 pi)
    for (int k = 1; k <= m; ++k) pmm = h64[kHDiag + k] * st * pmm;
    double p0 = 0.0, p1 = pmm;
    out[0] = pmm;
    const double* ab = h64 + 2 * (m * kH);
    for (int l = m + 1; l < lh; ++l) {
        const double a = ab[2 * l], b = ab[2 * l + 1];
        const double p = a * (x * p1 - b * p0);
        p0 = p1;
        p1 = p;
        out[(long long)(l - m) * stride] = p;
    }
}

// Offset of column m in an instance's triangular Legendre table.
__host__ __device__ inline long long tri_off(int m, int lh) {
    return (long long)m * lh - (long long)m * (m - 1) / 2;
}

// Pn_lm(cos th_w) in pl[pl_off + (tri_off(m) + (l - m)) * nk + w],
// e^{i m phi_w} in e[e_off + m * nk + w] and, at m = 0, the wave weights
// phi_w = (1/nk) subA exp(i kmag k.c) of every sphere of the instance in
// phi[woff + w] (FP64); block 0 also clears the work-queue counters.
// Grid: (wave blocks, instances, m).
__global__ void harm_legendre_kernel(const SeriesInst* __restrict__ insts,
                                     const SeriesSph* __restrict__ sph, int n_inst,
                                     const double* __restrict__ waves,
                                     const double* __restrict__ h64, double* __restrict__ pl,
                                     double2* __restrict__ e, double2* __restrict__ phi,
                                     int* __restrict__ counters, int n_counters) {
    const int m = blockIdx.z;
    if (blockIdx.x == 0 && blockIdx.y == 0 && m == 0)
        for (int i = threadIdx.x; i < n_counters; i += blockDim.x) counters[i] = 0;
    for (int i = blockIdx.y; i < n_inst; i += gridDim.y) {
        const SeriesInst& I = insts[i];
        if (m >= I.lh) continue;
        for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < I.nk; w += gridDim.x * blockDim.x) {
            const d3 k = mk(waves + 3 * (I.wave_off + w));
            legendre_column(m, I.lh, k.z, pl + I.pl_off + tri_off(m, I.lh) * I.nk + w, I.nk, h64);
            const double az = atan2(k.y, k.x);
            e[I.e_off + (long long)m * I.nk + w] = make_double2(cos(m * az), sin(m * az));
            if (m == 0) {
                for (int s = I.s_begin; s < I.s_begin + I.ns; ++s) {
                    const SeriesSph& SS = sph[s];
                    const d3 c = sub(mk(SS.s.p), I.f);
                    const dc ph = cscl(cexp_i(I.kmag * dot(k, c)), I.sA);
                    phi[SS.woff + w] = make_double2(ph.r, ph.i);
                }
            }
        }
    }
}

// Records per (sphere, m, l) from the wave tables.  Block (m, instance)
// forms row m of the records of every sphere of the instance: with
// G_w = Pn_lm(cos th_w) e^{i m phi_w} and the sphere's wave weights phi_w,
//   S = sum_w phi_w G_w,  S' = sum_w phi_w conj(G_w),
// i.e. per (order, sphere) four real sums a = sum Re phi Re G,
// b = sum Im phi Im G, c = sum Re phi Im G, d = sum Im phi Re G, with
// S = (a - b, c + d) and S' = (a + b, d - c); the records hold
//   U1 = S + S' = (2a, 2d),  U2 = (Im S - Im S', Re S' - Re S) = (2c, 2b)
// (at m = 0, G is real: U1 = S = (a, d), U2 = 0).  A record is 4 pi/(2l+1)
// K_lm (or Pn_lm(0) for a planar sphere) times (U1, U2).  The waves run in
// chunks of kRecW staged in LDS (the Legendre values of the row's orders,
// the azimuth phases, the weights of the instance's spheres); each thread
// owns up to kRecPer (order, four spheres) combinations and sums them in
// FP64, so the wave tables are read once per row for all spheres.  Planar-
// only instances form only the orders l = m, m + 2, ...; the orders a
// sphere does not need (past its Nl, or l + m odd when planar) are not
// written (the plan clears the record tables once).
// Grid: (m < the largest Legendre order count, instances).
constexpr int kRecThreads = 256;
// Limits of the harmonic form's per-instance pixel orders: pixels in the
// plan and pixel-sphere tests on the host to build them.
constexpr long long kPermMaxPixels = 16LL << 20;
constexpr long long kPermMaxTests = 128LL << 20;
constexpr int kRecPer = 2;        // combinations per thread and pass

// NT threads per block, tiles of tl orders and ts spheres (a multiple of
// four), kRecW waves per LDS chunk: 256 threads, tiles of up to 128 orders
// and 64 spheres, or 64 threads, up to 64 and 32, where a row has few
// combinations (a 256-thread block would leave most of its threads idle;
// on the sweep's rows of at most 25 orders and five groups the 64-thread
// form took 0.82 against 1.02 ms, and chunks of 32 waves were slower,
// 0.99 ms).  The tiles are the plan's largest row and sphere count within
// those limits (rec_lds_bytes), so that the LDS a block takes, and with it
// the blocks a CU holds, follows the batch (the sweep's 25 orders and 20
// spheres: 8.6 KB, against 16.3 KB for fixed tiles of 64 and 32).
template <int NT, int kRecW>
__device__ __forceinline__ void harm_records_row(
    int m, int i, const SeriesInst* __restrict__ insts, const SeriesSph* __restrict__ sph,
    const double* __restrict__ pl, const double2* __restrict__ e, const double2* __restrict__ phi,
    const double* __restrict__ h64, float4* __restrict__ rec, int tl, int ts) {
    extern __shared__ double2 rec_lds[];
    double2* const sE = rec_lds;                                       // e^{i m phi_w}
    double2* const sF = sE + kRecW;                                    // [w][ts] sphere weights
    double* const sP = reinterpret_cast<double*>(sF + kRecW * ts);  // [w][tl] Pn_lm(cos th_w)
    const SeriesInst& I = insts[i];
    if (m >= I.lh) return;
    const int t = (int)threadIdx.x;
    bool all_planar = true;
    for (int s = 0; s < I.ns; ++s) all_planar = all_planar && sph[I.s_begin + s].planar != 0;
    const int lstep = all_planar ? 2 : 1;
    const int nlv = (I.lh - m + lstep - 1) / lstep;   // orders l = m + j lstep
    const double* Prow = pl + I.pl_off + tri_off(m, I.lh) * I.nk;
    const double2* E = e + I.e_off + (long long)m * I.nk;
    // the wave weights of the instance's spheres follow each other, nk_pad
    // apart
    const long long wbase = sph[I.s_begin].woff;
    static_assert((kRecW & (kRecW - 1)) == 0, "chunks of a power of two waves");
    constexpr int kRecWLog = __builtin_ctz(kRecW);
    for (int s0 = 0; s0 < I.ns; s0 += ts) {
        const int nst = min(ts, I.ns - s0);
        const int nq = (nst + 3) >> 2;
        for (int j0 = 0; j0 < nlv; j0 += tl) {
            const int nlt = min(tl, nlv - j0);
            const int combos = nq * nlt;
            for (int c0 = 0; c0 < combos; c0 += kRecPer * NT) {
                // combination c: order index c % nlt, sphere group c / nlt
                double acc[kRecPer][4][4];
#pragma unroll
                for (int u = 0; u < kRecPer; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
#pragma unroll
                        for (int q = 0; q < 4; ++q) acc[u][v][q] = 0.0;
                // the azimuth phase a thread stages, loaded a chunk ahead
                double2 ev = E[min(t, I.nk - 1)];
                for (int w0 = 0; w0 < I.nk; w0 += kRecW) {
                    const int nw = min(kRecW, I.nk - w0);
                    __syncthreads();   // the previous chunk is consumed
                    // (consecutive threads take consecutive waves: the
                    // element's wave and order or sphere by a mask and a
                    // shift, no division.  Four elements of a table a step,
                    // all loaded before any is stored, so that a step waits
                    // on its loads once; the stores take no condition (one
                    // would draw each load down beside its store): an index
                    // past the table repeats its last element, and the rows
                    // of waves past the chunk's, which no product reads, get
                    // the chunk's last wave)
                    for (int x0 = t; x0 < kRecW * nlt; x0 += 4 * NT) {
                        double v[4];
                        int xs[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            xs[k] = min(x0 + k * NT, kRecW * nlt - 1);
                            const int w = min(xs[k] & (kRecW - 1), nw - 1), jj = xs[k] >> kRecWLog;
                            v[k] = Prow[(long long)((j0 + jj) * lstep) * I.nk + w0 + w];
                        }
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            sP[(xs[k] & (kRecW - 1)) * tl + (xs[k] >> kRecWLog)] = v[k];
                    }
                    for (int x0 = t; x0 < kRecW * nst; x0 += 4 * NT) {
                        double2 v[4];
                        int xs[4];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            xs[k] = min(x0 + k * NT, kRecW * nst - 1);
                            const int w = min(xs[k] & (kRecW - 1), nw - 1), ss = xs[k] >> kRecWLog;
                            v[k] = phi[wbase + (long long)(s0 + ss) * I.nk_pad + w0 + w];
                        }
#pragma unroll
                        for (int k = 0; k < 4; ++k)
                            sF[(xs[k] & (kRecW - 1)) * ts + (xs[k] >> kRecWLog)] = v[k];
                    }
                    if (t < nw) sE[t] = ev;
                    __syncthreads();
                    ev = E[min(w0 + kRecW + t, I.nk - 1)];   // the next chunk's
#pragma unroll
                    for (int u = 0; u < kRecPer; ++u) {
                        const int c = c0 + u * NT + t;
                        if (c >= combos) continue;
                        const int jj = c % nlt, g = c / nlt;
                        for (int w = 0; w < nw; ++w) {
                            const double p = sP[w * tl + jj];
                            const double2 ew = sE[w];
                            const double gr = p * ew.x, gi = p * ew.y;
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const double2 f = sF[w * ts + min(4 * g + v, nst - 1)];
                                acc[u][v][0] = fma(f.x, gr, acc[u][v][0]);
                                acc[u][v][1] = fma(f.y, gi, acc[u][v][1]);
                                acc[u][v][2] = fma(f.x, gi, acc[u][v][2]);
                                acc[u][v][3] = fma(f.y, gr, acc[u][v][3]);
                            }
                        }
                    }
                }
                // the records of this pass's combinations
#pragma unroll
                for (int u = 0; u < kRecPer; ++u) {
                    const int c = c0 + u * NT + t;
                    if (c >= combos) continue;
                    const int jj = c % nlt, g = c / nlt;
                    const int l = m + (j0 + jj) * lstep;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int ss = 4 * g + v;
                        if (ss >= nst) continue;
                        const SeriesSph& SS = sph[I.s_begin + s0 + ss];
                        const bool planar = SS.planar != 0;
                        if (l > SS.s.Nl || l >= SS.tl || (planar && ((l + m) & 1))) continue;
                        double f;
                        if (planar) {
                            // Pn_mm(0) = Pn_00 prod_q sqrt((2q+1)/(2q)),
                            // Pn_lm(0) = -a_lm b_lm Pn_{l-2,m}(0)
                            double pn = 0.28209479177387814;
                            for (int q = 1; q <= m; ++q) pn *= h64[kHDiag + q];
                            const double* ab = h64 + 2 * (m * kH);
                            for (int q = m + 2; q <= l; q += 2) pn = -(ab[2 * q] * ab[2 * q + 1]) * pn;
                            f = 4.0 * M_PI / (2.0 * l + 1.0) * pn;
                        } else {
                            f = 4.0 * M_PI / (2.0 * l + 1.0) * h64[kHK + m * kH + l];
                        }
                        const double a = acc[u][v][0], b = acc[u][v][1];
                        const double cc = acc[u][v][2], d = acc[u][v][3];
                        float4 r;
                        if (m == 0)
                            r = make_float4((float)(f * a), (float)(f * d), 0.f, 0.f);
                        else
                            r = make_float4((float)(f * 2.0 * a), (float)(f * 2.0 * d),
                                            (float)(f * 2.0 * cc), (float)(f * 2.0 * b));
                        const long long o = planar ? (long long)m * (SS.tl / 2) + (l - (m & 1)) / 2
                                                   : (long long)m * SS.tl + l;
                        rec[SS.toff + o] = r;
                    }
                }
            }
        }
    }
}

// Records of few instances: block (m, tile, s) forms four records of row m
// of sphere s (orders l = 4 tile + j, or for a planar sphere k = 4 tile + j
// with l = 2k + (m & 1)), its threads running over the instance's waves
// (coalesced reads of the Legendre, azimuth-phase and wave-weight tables)
// and a block reduction forming S and S' as harm_records_row's; records
// outside m <= l <= Nl are written as zeros.  More blocks than the row form
// when a batch holds few instances, so that the waves spread over the chip.
// Grid: (m < largest record stride, tiles, spheres).
constexpr int kRecTile = 4;

__device__ __forceinline__ void harm_records_tile(
    int m, int tile, int s_first, int s_step, const SeriesInst* __restrict__ insts,
    const SeriesSph* __restrict__ sph, int n_sph, const double* __restrict__ pl,
    const double2* __restrict__ e, const double2* __restrict__ phi,
    const double* __restrict__ h64, float4* __restrict__ rec) {
    __shared__ double red[kRecThreads / 64][4 * kRecTile];
    for (int s = s_first; s < n_sph; s += s_step) {
        const SeriesSph& SS = sph[s];
        const int tl = SS.tl;
        const bool planar = SS.planar != 0;
        const int width = planar ? tl / 2 : tl;   // records per row
        if (m >= tl || kRecTile * tile >= width) continue;
        const SeriesInst& I = insts[SS.inst];
        const int nl1 = SS.s.Nl + 1;
        int lj[kRecTile];
        bool need[kRecTile];
        bool any = false;
#pragma unroll
        for (int j = 0; j < kRecTile; ++j) {
            const int k = kRecTile * tile + j;
            lj[j] = planar ? 2 * k + (m & 1) : k;
            need[j] = k < width && m < nl1 && lj[j] >= m && lj[j] < nl1;
            any = any || need[j];
        }
        // (T+ re, T+ im, T- re, T- im) of each record
        double acc[4 * kRecTile];
#pragma unroll
        for (int q = 0; q < 4 * kRecTile; ++q) acc[q] = 0.0;
        if (any) {
            const double* P[kRecTile];
#pragma unroll
            for (int j = 0; j < kRecTile; ++j)
                P[j] = pl + I.pl_off + (tri_off(m, I.lh) + (need[j] ? lj[j] - m : 0)) * I.nk;
            const double2* E = e + I.e_off + (long long)m * I.nk;
            const double2* F = phi + SS.woff;
            for (int w = threadIdx.x; w < I.nk; w += kRecThreads) {
                const double2 ew = E[w], fw = F[w];
#pragma unroll
                for (int j = 0; j < kRecTile; ++j) {
                    if (!need[j]) continue;
                    const double p = P[j][w];
                    const double ar = fw.x * p, ai = fw.y * p;
                    acc[4 * j] += ar * ew.x - ai * ew.y;
                    acc[4 * j + 1] += ar * ew.y + ai * ew.x;
                    acc[4 * j + 2] += ar * ew.x + ai * ew.y;
                    acc[4 * j + 3] += ai * ew.x - ar * ew.y;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4 * kRecTile; ++q) {
            double v = acc[q];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            acc[q] = v;
        }
        __syncthreads();   // red is reused for the next sphere
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int q = 0; q < 4 * kRecTile; ++q) red[threadIdx.x >> 6][q] = acc[q];
        __syncthreads();
        if (threadIdx.x < kRecTile) {
            const int j = threadIdx.x;
            const int k = kRecTile * tile + j;
            if (k < width) {
                float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
                if (need[j]) {
                    dc tp = {0.0, 0.0}, tm = {0.0, 0.0};
                    for (int q = 0; q < kRecThreads / 64; ++q) {
                        tp = cadd(tp, dc{red[q][4 * j], red[q][4 * j + 1]});
                        tm = cadd(tm, dc{red[q][4 * j + 2], red[q][4 * j + 3]});
                    }
                    const int l = lj[j];
                    double f;
                    if (planar) {
                        // Pn_lm(0): Pn_mm(0) = Pn_00 prod_q sqrt((2q+1)/(2q)),
                        // Pn_lm(0) = -a_lm b_lm Pn_{l-2,m}(0)
                        double pn = 0.28209479177387814;
                        for (int q = 1; q <= m; ++q) pn *= h64[kHDiag + q];
                        const double* ab = h64 + 2 * (m * kH);
                        for (int q = m + 2; q <= l; q += 2) pn = -(ab[2 * q] * ab[2 * q + 1]) * pn;
                        f = 4.0 * M_PI / (2.0 * l + 1.0) * pn;
                    } else {
                        f = 4.0 * M_PI / (2.0 * l + 1.0) * h64[kHK + m * kH + l];
                    }
                    if (m == 0)
                        r = make_float4((float)(f * tp.r), (float)(f * tp.i), 0.f, 0.f);
                    else
                        r = make_float4((float)(f * (tp.r + tm.r)), (float)(f * (tp.i + tm.i)),
                                        (float)(f * (tp.i - tm.i)), (float)(f * (tm.r - tp.r)));
                }
                rec[SS.toff + (long long)m * width + k] = r;
            }
        }
    }
}

template <int NT, int kRecW>
__global__ __launch_bounds__(NT) void harm_records_kernel(
    const SeriesInst* __restrict__ insts, const SeriesSph* __restrict__ sph, int n_inst,
    const double* __restrict__ pl, const double2* __restrict__ e, const double2* __restrict__ phi,
    const double* __restrict__ h64, float4* __restrict__ rec, int tl, int ts) {
    for (int i = blockIdx.y; i < n_inst; i += gridDim.y)
        harm_records_row<NT, kRecW>(blockIdx.x, i, insts, sph, pl, e, phi, h64, rec, tl, ts);
}

// Tiles of harm_records_row for the plan's rows (orders) and spheres, and
// the LDS they take with chunks of kW waves.
struct RecTiles {
    int tl, ts;
    size_t lds;
};
RecTiles rec_tiles(int max_nlv, int max_ns, int l_cap, int s_cap, int kW) {
    RecTiles r;
    r.tl = std::max(1, std::min(max_nlv, l_cap));
    r.ts = std::min((std::max(1, max_ns) + 3) / 4 * 4, s_cap);
    r.lds = sizeof(double2) * (size_t)kW * (1 + r.ts) + sizeof(double) * (size_t)kW * r.tl;
    return r;
}

__global__ __launch_bounds__(kRecThreads) void harm_records_tile_kernel(
    const SeriesInst* __restrict__ insts, const SeriesSph* __restrict__ sph, int n_sph,
    const double* __restrict__ pl, const double2* __restrict__ e, const double2* __restrict__ phi,
    const double* __restrict__ h64, float4* __restrict__ rec) {
    harm_records_tile(blockIdx.x, blockIdx.y, blockIdx.z, gridDim.z, insts, sph, n_sph, pl, e, phi,
                      h64, rec);
}

// The records times the sphere's coefficients, once both are formed:
// rec (in place) times B_l, rec_in times A_l, l the record's order; the
// records of orders below the row or past Nl stay zero.  Rows 0 .. tl - 1
// of each table (tl / 2 records each for a planar sphere, else tl) hold
// every record the pixel kernel reads.  Grid: (record blocks, spheres).
__global__ __launch_bounds__(kRecThreads) void harm_fold_kernel(const SeriesSph* __restrict__ sph,
                                                                int n_sph,
                                                                const double2* __restrict__ A,
                                                                const double2* __restrict__ B,
                                                                float4* __restrict__ rec,
                                                                float4* __restrict__ rec_in) {
    for (int s = blockIdx.y; s < n_sph; s += gridDim.y) {
        const SeriesSph& SS = sph[s];
        const int tl = SS.tl;
        const bool planar = SS.planar != 0;
        const int width = planar ? tl / 2 : tl;
        const int n = tl * width;
        for (int o = blockIdx.x * kRecThreads + (int)threadIdx.x; o < n;
             o += gridDim.x * kRecThreads) {
            const int m = o / width, k = o - m * width;
            const int l = planar ? 2 * k + (m & 1) : k;
            float4 ob = make_float4(0.f, 0.f, 0.f, 0.f), oa = ob;
            if (l >= m && l <= SS.s.Nl) {
                const float4 r = rec[SS.toff + o];
                const double2 b = B[SS.s.coeff_off + l], a = A[SS.s.coeff_off + l];
                // U1 = (r.x, r.y) and U2 = (r.z, r.w), each times the coefficient
                auto mul = [](double x, double y, double2 c) {
                    return make_float2((float)(x * c.x - y * c.y), (float)(x * c.y + y * c.x));
                };
                const float2 b1 = mul(r.x, r.y, b), b2 = mul(r.z, r.w, b);
                const float2 a1 = mul(r.x, r.y, a), a2 = mul(r.z, r.w, a);
                ob = make_float4(b1.x, b1.y, b2.x, b2.y);
                oa = make_float4(a1.x, a1.y, a2.x, a2.y);
            }
            rec[SS.toff + o] = ob;
            rec_in[SS.toff + o] = oa;
        }
    }
}

// The FP32 batch's pre-pass as one launch: Mie coefficients (blocks
// [0, n_mie), mie_block over spb spheres each, Bessel arrays in dynamic
// LDS), the harmonic records (the next n_rec blocks, one per (m, instance)
// as harm_records_kernel) and the incident fields (the rest, 256
// pixels of one instance each, which also clear its detector images when
// a.sum is given).  None reads what another writes, and the
// records read only the Legendre tables of the launch before, so the long
// dependent chains of the coefficients run on a few compute units while
// the records and fields fill the others.  The role is uniform per block.
__global__ __launch_bounds__(kRecThreads) void prepass_kernel(
    PrepassArgs a, int n_mie, int n_rec, int rec_m, int rec_tiles,
    const SeriesInst* __restrict__ insts, const SeriesSph* __restrict__ sph, int n_sph,
    const double* __restrict__ pl, const double2* __restrict__ e, const double2* __restrict__ phi,
    const double* __restrict__ h64, float4* __restrict__ rec) {
    static_assert(kRecThreads >= mie::kMieThreads, "the coefficient waves fit the block");
    extern __shared__ mie::zc mie_lds[];
    __shared__ SphereDev msph[mie::kSphPerLaunch];
    __shared__ double mk[mie::kSphPerLaunch];
    int b = (int)blockIdx.x;
    const int t = (int)threadIdx.x;
    if (b < n_mie) {
        const int s0 = b * a.spb;
        const int cnt = min(a.spb, a.n_mie_sph - s0);
        if (t < cnt) {
            msph[t] = a.mie_sph[s0 + t];
            mk[t] = a.mie_kmag[s0 + t];
        }
        __syncthreads();
        mie::mie_block(msph, mk, cnt, a.quirks, mie_lds, a.wstride, a.A, a.B);
        return;
    }
    b -= n_mie;
    if (b < n_rec) {
        const int per_s = rec_m * rec_tiles;
        const int s = b / per_s, r = b - s * per_s;
        harm_records_tile(r % rec_m, r / rec_m, s, n_sph, insts, sph, n_sph, pl, e, phi, h64, rec);
        return;
    }
    b -= n_rec;
    const int i = b / a.uf_blocks, pix = (b - i * a.uf_blocks) * kRecThreads + t;
    const UfInst& I = a.uf[i];
    if (pix < ufd::uf_batch_work(I)) ufd::uf_batch_item(I, pix, a.uf_tabs, a.uf_waves, a.Uf);
    if (a.sum) {
        const SumInst& S = a.sum[i];   // nd2 <= uR vR
        if (pix < S.nd2) {
            a.D[S.image_off + pix] = 0.f;
            a.Di[S.image_off + pix] = 0.f;
        }
    }
}

template <int L>
void launch_bucket(dim3 grid, hipStream_t st, const SeriesPlan& P, const SeriesPlan::Bucket& K,
                   float2* U) {
    hipLaunchKernelGGL(series_f32_kernel<L>, grid, dim3(kThreads), 0, st, P.d_inst, P.d_sph,
                       P.d_items + K.item_begin, K.item_count, P.d_coef, P.d_rhat, P.d_wt, U,
                       P.d_part, P.d_counters + K.counter);
}

template <class T>
int upload(bimsim_ctx* ctx, T** dst, const std::vector<T>& v) {
    if (v.empty()) return BIMSIM_OK;
    if (hipMalloc((void**)dst, sizeof(T) * v.size()) != hipSuccess)
        return set_error(ctx, BIMSIM_E_NOMEM, "out of device memory for the series tables");
    BIMSIM_HIP_CHECK(ctx, hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return BIMSIM_OK;
}

template <class T>
int alloc(bimsim_ctx* ctx, T** dst, long long n) {
    if (n <= 0) return BIMSIM_OK;
    if (hipMalloc((void**)dst, sizeof(T) * (size_t)n) != hipSuccess)
        return set_error(ctx, BIMSIM_E_NOMEM, "out of device memory for the series buffers");
    return BIMSIM_OK;
}

// Plans of per-call scattered fields, keyed by their tables' bytes.  A plan
// used while the stream is being captured is pinned: the captured graph keeps
// its device pointers, so it stays until the context ends.  Other plans are
// evicted least recently used beyond kPlanCacheSize.
struct SeriesCacheEntry {
    std::vector<char> key;
    SeriesPlan plan;
    bool pinned = false;
    unsigned long long last_use = 0;
};
struct SeriesCache {
    std::vector<SeriesCacheEntry*> entries;
    unsigned long long clock = 0;
};
constexpr int kPlanCacheSize = 4;

}  // namespace

int series_capacity(int nl1) {
    for (int L : kBuckets)
        if (L >= nl1) return L;
    return -1;
}

void series_plan_destroy(SeriesPlan* P) {
    (void)hipFree(P->d_inst);
    (void)hipFree(P->d_sph);
    (void)hipFree(P->d_items);
    (void)hipFree(P->d_perm);
    (void)hipFree(P->d_grp);
    (void)hipFree(P->d_split_insts);
    (void)hipFree(P->d_counters);
    (void)hipFree(P->d_wt);
    (void)hipFree(P->d_part);
    (void)hipFree(P->d_coef);
    (void)hipFree(P->d_rhat);
    void* harm[] = {P->d_pl, P->d_e, P->d_phi, P->d_rec, P->d_rec_in, P->d_cst, P->d_h64};
    for (void* q : harm) (void)hipFree(q);
    *P = SeriesPlan();
}

// The harmonic form's planar case: pixels p = A + u X + v Y with X.z = Y.z
// = 0 and A.z = c.z have r.z = 0 exactly, so x = 0 at every pixel.
static bool harm_planar(const QuadArg& q, const SphereDev& S) {
    return q.X.z == 0.0 && q.Y.z == 0.0 && q.A.z == S.p[2];
}

int series_plan_create(bimsim_ctx* ctx, SeriesBatch& batch, size_t coef_budget, SeriesPlan* P,
                       int split_mult, int groups_per_cu) {
    *P = SeriesPlan();
    P->groups_per_cu = std::max(1, groups_per_cu);
    const int ni = (int)batch.inst.size();
    P->n_inst = ni;
    P->n_sph = (int)batch.sph.size();
    std::vector<int> bucket(ni);
    std::vector<double> cost(ni);
    long long blocks_all = 0;
    int next_sph = 0;
    for (int i = 0; i < ni; ++i) {
        SeriesInst& I = batch.inst[i];
        if (I.s_begin != next_sph || I.ns < 0 || I.s_begin + I.ns > P->n_sph)
            return set_error(ctx, BIMSIM_E_INVALID, "instance spheres must be consecutive ranges");
        next_sph += I.ns;
        int nl1 = 1;
        for (int s = I.s_begin; s < I.s_begin + I.ns; ++s) {
            SeriesSph& S = batch.sph[s];
            if (S.inst != i) return set_error(ctx, BIMSIM_E_INVALID, "sphere/instance mismatch");
            S.nb = (S.s.Nl + 1 + 7) / 8;
            nl1 = std::max(nl1, S.s.Nl + 1);
        }
        bucket[i] = series_capacity(nl1);
        if (bucket[i] < 0)
            return set_error(ctx, BIMSIM_E_UNSUPPORTED, "FP32 series supports Nl <= %d (got %d)",
                             kBuckets[8] - 1, nl1 - 1);
        I.nk_pad = (I.nk + kW - 1) / kW * kW;
        const int npix = I.uR * I.vR;
        blocks_all += (npix + kThreads - 1) / kThreads;
        cost[i] = (double)npix * I.ns * I.nk_pad * bucket[i];
        P->nk_max = std::max(P->nk_max, I.nk_pad);
        P->max_npix = std::max(P->max_npix, npix);
        P->field_elems = std::max(P->field_elems, I.u_off + (long long)npix);
    }
    if (next_sph != P->n_sph) return set_error(ctx, BIMSIM_E_INVALID, "unused spheres in batch");
    // the harmonic form replaces the wave loop by (Nl+1)(Nl+2)/2 terms per
    // pixel and sphere: chosen unless some instance has very few waves
    int min_nk = 1 << 30;
    for (const SeriesInst& I : batch.inst) min_nk = std::min(min_nk, I.nk);
    P->harmonic = ctx->series_algo == kSeriesHarmonic ||
                  (ctx->series_algo == kSeriesAuto && min_nk >= 16);
    if (P->harmonic) {
        split_mult = 1;
        // each instance's spheres sorted by their own order bucket, so that
        // a bucket's spheres form one consecutive group per instance
        for (const SeriesInst& I : batch.inst)
            std::stable_sort(batch.sph.begin() + I.s_begin, batch.sph.begin() + I.s_begin + I.ns,
                             [](const SeriesSph& a, const SeriesSph& b) {
                                 return series_capacity(a.s.Nl + 1) < series_capacity(b.s.Nl + 1);
                             });
    }
    // harmonic form: the sphere groups of each instance, in bucket order
    std::vector<SeriesGroup> groups;
    std::vector<std::vector<int>> inst_groups(ni);
    if (P->harmonic) {
        for (int i = 0; i < ni; ++i) {
            const SeriesInst& I = batch.inst[i];
            for (int s = I.s_begin; s < I.s_begin + I.ns;) {
                const int L = series_capacity(batch.sph[s].s.Nl + 1);
                int e = s + 1;
                while (e < I.s_begin + I.ns && series_capacity(batch.sph[e].s.Nl + 1) == L) ++e;
                inst_groups[i].push_back((int)groups.size());
                groups.push_back(SeriesGroup{s, e - s,
                                             inst_groups[i].size() == 1 ? I.accumulate : 1, L});
                s = e;
            }
        }
    }
    // split the waves until the queue holds several items per CU (each
    // split keeps at least 32 waves)
    const long long target = 4LL * ctx->cu_count;
    int mult = 1;
    if (split_mult > 0) {
        mult = split_mult;
    } else {
        for (;;) {
            bool can = false;
            for (int i = 0; i < ni; ++i)
                if (batch.inst[i].nk_pad / (2 * mult) >= 32) can = true;
            if (blocks_all * mult >= target || !can) break;
            mult *= 2;
        }
    }
    P->split_mult = mult;
    long long woff = 0, poff = 0;
    for (int s = 0; s < P->n_sph; ++s) {
        batch.sph[s].woff = woff;
        woff += batch.inst[batch.sph[s].inst].nk_pad;
    }
    P->n_wave_rec = woff;
    for (int i = 0; i < ni; ++i) {
        SeriesInst& I = batch.inst[i];
        int nsplit = 1;
        while (nsplit < mult && I.nk_pad / (2 * nsplit) >= 32) nsplit *= 2;
        int wchunk = (I.nk_pad + nsplit - 1) / nsplit;
        wchunk = std::max(kW, (wchunk + kW - 1) / kW * kW);
        I.wchunk = wchunk;
        I.nsplit = std::max(1, (I.nk_pad + wchunk - 1) / wchunk);
        I.part_off = 0;
        if (I.nsplit > 1) {
            I.part_off = poff;
            poff += (long long)I.nsplit * I.uR * I.vR;
            P->split_insts.push_back(i);
        }
    }
    P->n_part = poff;
    // chunks of whole instances within the coefficient budget
    std::vector<SeriesItem> items;
    int i0 = 0;
    while (i0 < ni) {
        long long celems = 0, relems = 0;
        int i1 = i0;
        for (; i1 < ni; ++i1) {
            const SeriesInst& I = batch.inst[i1];
            const long long npix = (long long)I.uR * I.vR;
            // the harmonic form forms its coefficients in the pixel kernel
            if (P->harmonic) continue;
            long long c = 0;
            for (int s = I.s_begin; s < I.s_begin + I.ns; ++s) c += 8LL * batch.sph[s].nb * npix;
            if (i1 > i0 && (size_t)(celems + c) * sizeof(float2) > coef_budget) break;
            for (int s = I.s_begin; s < I.s_begin + I.ns; ++s) {
                batch.sph[s].coff = celems;
                celems += 8LL * batch.sph[s].nb * npix;
                batch.sph[s].roff = relems;
                relems += npix;
            }
        }
        P->coef_elems = std::max(P->coef_elems, celems);
        P->rhat_elems = std::max(P->rhat_elems, relems);
        SeriesPlan::Chunk ch;
        ch.sph_begin = batch.inst[i0].s_begin;
        ch.sph_count = batch.inst[i1 - 1].s_begin + batch.inst[i1 - 1].ns - ch.sph_begin;
        std::vector<int> order;
        for (int i = i0; i < i1; ++i) order.push_back(i);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return cost[a] > cost[b]; });
        for (int L : kBuckets) {
            const int begin = (int)items.size();
            int all_planar = 1;
            for (int i : order) {
                const SeriesInst& I = batch.inst[i];
                const int npix = I.uR * I.vR;
                if (P->harmonic) {
                    // the instance's group of this bucket, if any
                    for (int g : inst_groups[i]) {
                        const SeriesGroup& G = groups[g];
                        if (G.L != L) continue;
                        for (int t = G.s_begin; t < G.s_begin + G.ns; ++t)
                            if (!harm_planar(I.q, batch.sph[t].s)) all_planar = 0;
                        for (int p0 = 0; p0 < npix; p0 += kThreads)
                            items.push_back(SeriesItem{i, p0, 0, g});
                    }
                    continue;
                }
                if (bucket[i] != L) continue;
                for (int t = I.s_begin; t < I.s_begin + I.ns; ++t)
                    if (!harm_planar(I.q, batch.sph[t].s)) all_planar = 0;
                // the splits of a pixel block follow each other, so that they
                // run side by side and share its coefficients in L2
                for (int p0 = 0; p0 < npix; p0 += kThreads)
                    for (int t = 0; t < I.nsplit; ++t) items.push_back(SeriesItem{i, p0, t, -1});
            }
            const int count = (int)items.size() - begin;
            if (count > 0)
                ch.buckets.push_back(SeriesPlan::Bucket{L, begin, count, P->n_counters++,
                                                        all_planar,
                                                        harm_default_variant(all_planar != 0)});
        }
        P->chunks.push_back(ch);
        i0 = i1;
    }
    std::vector<float> cst;
    std::vector<double> h64;
    if (P->harmonic) {
        long long npl = 0, ne = 0, nrec = 0;
        for (int i = 0; i < ni; ++i) {
            SeriesInst& I = batch.inst[i];
            int lh = 1;
            for (int t = I.s_begin; t < I.s_begin + I.ns; ++t)
                lh = std::max(lh, batch.sph[t].s.Nl + 1);
            I.lh = lh;
            I.pl_off = npl;
            npl += (long long)lh * (lh + 1) / 2 * I.nk;
            I.e_off = ne;
            ne += (long long)lh * I.nk;
            P->max_lh = std::max(P->max_lh, lh);
            // combinations (order, group of four spheres) of the instance's row 0
            bool planar_all = true;
            for (int t = I.s_begin; t < I.s_begin + I.ns; ++t)
                planar_all = planar_all && harm_planar(I.q, batch.sph[t].s);
            const int nlv = planar_all ? (lh + 1) / 2 : lh;
            P->max_rec_combos = std::max(P->max_rec_combos, nlv * ((I.ns + 3) / 4));
            P->max_rec_nlv = std::max(P->max_rec_nlv, nlv);
            P->max_rec_ns = std::max(P->max_rec_ns, I.ns);
        }
        for (int t = 0; t < P->n_sph; ++t) {
            SeriesSph& S = batch.sph[t];
            const SeriesInst& I = batch.inst[S.inst];
            S.tl = series_capacity(S.s.Nl + 1);
            P->max_tl = std::max(P->max_tl, S.tl);
            S.planar = harm_planar(I.q, S.s) ? 1 : 0;
            S.toff = nrec;
            nrec += (long long)S.tl * S.tl;
        }
        P->n_pl = npl;
        P->n_e = ne;
        P->n_rec = nrec;
        // c_lm = 4 b_lm / a_{l-1,m} (l >= m + 2) and the diagonal factors,
        // a_lm = sqrt((4l^2-1)/(l^2-m^2)), b_lm = sqrt(((l-1)^2-m^2)/(4(l-1)^2-1))
        auto a_ = [](int l, int m) {
            return std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
        };
        h64.assign((size_t)kHSize, 0.0);
        for (int m = 0; m < kH; ++m) {
            if (m >= 1) h64[kHDiag + m] = std::sqrt((2.0 * m + 1.0) / (2.0 * m));
            double K = 1.0;
            h64[kHK + (size_t)m * kH + m] = K;
            for (int l = m + 1; l < kH; ++l) {
                const double a = std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
                const double b = std::sqrt((((double)l - 1.0) * (l - 1.0) - (double)m * m) /
                                           (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                h64[2 * ((size_t)m * kH + l)] = a;
                h64[2 * ((size_t)m * kH + l) + 1] = b;
                K *= 0.5 * std::sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
                h64[kHK + (size_t)m * kH + l] = K;
            }
        }
        // (one row more: the kernel reads the diagonal factor of row m + 1)
        cst.assign((size_t)2 * (kCS + 2) * kCS, 0.f);
        for (int m = 0; m <= kCS; ++m) {
            if (m >= 1)
                cst[2 * ((size_t)m * kCS + m)] = (float)std::sqrt((2.0 * m + 1.0) / (2.0 * m));
            if (m == kCS) break;
            cst[2 * ((size_t)m * kCS + m) + 1] = 1.f;
            for (int l = m + 2; l < kCS; ++l) {
                const double b = std::sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) /
                                           (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                cst[2 * ((size_t)m * kCS + l)] = (float)(4.0 * b / a_(l - 1, m));
            }
        }
    }
    // harmonic form: the pixels of each instance with pixels inside a sphere
    // in the order inside each sphere, then outside every sphere (stable
    // within a class), so that the pixel kernel's waves mostly take one of
    // the two radial paths (a wave across a sphere's edge takes both); the
    // order changes no value, each pixel being computed on its own.  Left
    // out for very large plans (host time and table size).
    std::vector<int> perm;
    for (SeriesInst& I : batch.inst) I.perm_off = -1;
    if (P->harmonic) {
        long long npx = 0, tests = 0;
        for (const SeriesInst& I : batch.inst) {
            npx += (long long)I.uR * I.vR;
            tests += (long long)I.uR * I.vR * I.ns;
        }
        if (npx <= kPermMaxPixels && tests <= kPermMaxTests) {
            std::vector<int> cls, start;
            for (SeriesInst& I : batch.inst) {
                const int npix = I.uR * I.vR;
                cls.assign(npix, 0);
                bool any = false;
                for (int pix = 0; pix < npix; ++pix) {
                    const int iv = pix / I.uR, iu = pix - iv * I.uR;
                    const d3 p = quad_at(I.q, (double)iu / (double)I.uR, (double)iv / (double)I.vR);
                    for (int t = 0; t < I.ns; ++t) {
                        const SphereDev& S = batch.sph[I.s_begin + t].s;
                        if (len(sub(p, mk(S.p))) <= S.a) {
                            cls[pix] = t + 1;
                            any = true;
                            break;
                        }
                    }
                }
                if (!any) continue;
                // the inside classes (slower waves) first, so that the
                // work queue starts them first; the outside class last
                for (int& c : cls) c = c == 0 ? I.ns : c - 1;
                I.perm_off = (long long)perm.size();
                start.assign(I.ns + 2, 0);
                for (int c : cls) ++start[c + 1];
                for (int c = 1; c <= I.ns + 1; ++c) start[c] += start[c - 1];
                perm.resize(perm.size() + npix);
                int* o = perm.data() + I.perm_off;
                for (int pix = 0; pix < npix; ++pix) o[start[cls[pix]]++] = pix;
            }
        }
    }
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    BIMSIM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    int rc;
    if ((!perm.empty() && (rc = upload(ctx, &P->d_perm, perm)) != BIMSIM_OK) ||
        (rc = upload(ctx, &P->d_grp, groups)) != BIMSIM_OK) {
        series_plan_destroy(P);
        return rc;
    }
    if (P->harmonic &&
        ((rc = upload(ctx, &P->d_cst, cst)) != BIMSIM_OK ||
         (rc = upload(ctx, &P->d_h64, h64)) != BIMSIM_OK ||
         (rc = alloc(ctx, &P->d_pl, P->n_pl)) != BIMSIM_OK ||
         (rc = alloc(ctx, &P->d_e, P->n_e)) != BIMSIM_OK ||
         (rc = alloc(ctx, &P->d_phi, std::max<long long>(1, P->n_wave_rec))) != BIMSIM_OK ||
         (rc = alloc(ctx, &P->d_rec, P->n_rec)) != BIMSIM_OK ||
         (rc = alloc(ctx, &P->d_rec_in, P->n_rec)) != BIMSIM_OK)) {
        series_plan_destroy(P);
        return rc;
    }
    // the records kernel writes only the records the spheres use: the rest
    // of each table stays zero, as the pixel kernel's blocks may read past Nl
    if (P->harmonic && P->n_rec > 0) {
        BIMSIM_HIP_CHECK(ctx, hipMemset(P->d_rec, 0, sizeof(float4) * (size_t)P->n_rec));
        BIMSIM_HIP_CHECK(ctx, hipMemset(P->d_rec_in, 0, sizeof(float4) * (size_t)P->n_rec));
    }
    if ((rc = upload(ctx, &P->d_inst, batch.inst)) != BIMSIM_OK ||
        (rc = upload(ctx, &P->d_sph, batch.sph)) != BIMSIM_OK ||
        (rc = upload(ctx, &P->d_items, items)) != BIMSIM_OK ||
        (rc = upload(ctx, &P->d_split_insts, P->split_insts)) != BIMSIM_OK ||
        (rc = alloc(ctx, &P->d_counters, std::max(1, P->n_counters))) != BIMSIM_OK ||
        (rc = alloc(ctx, &P->d_wt, std::max<long long>(1, P->n_wave_rec))) != BIMSIM_OK ||
        (rc = alloc(ctx, &P->d_part, P->n_part)) != BIMSIM_OK ||
        (rc = alloc(ctx, &P->d_coef, std::max<long long>(1, P->coef_elems))) != BIMSIM_OK ||
        (rc = alloc(ctx, &P->d_rhat, std::max<long long>(1, P->rhat_elems))) != BIMSIM_OK) {
        series_plan_destroy(P);
        return rc;
    }
    return BIMSIM_OK;
}

int series_plan_launch(bimsim_ctx* ctx, const SeriesPlan& P, const double* waves, const double2* A,
                       const double2* B, float2* U) {
    BIMSIM_TRY(series_plan_launch_waves(ctx, P, waves));
    return series_plan_launch_sums(ctx, P, A, B, U);
}

int series_plan_launch_waves(bimsim_ctx* ctx, const SeriesPlan& P, const double* waves) {
    if (P.n_inst == 0 || P.n_sph == 0) return BIMSIM_OK;
    const hipStream_t st = ctx->stream;
    const dim3 wgrid((P.nk_max + 255) / 256, std::min(P.n_sph, 65535));
    if (P.harmonic) {
        // per-wave Legendre values, azimuth phases and wave weights (and the
        // counters cleared), then the records from them
        const dim3 lgrid((P.nk_max + 63) / 64, std::min(P.n_inst, 65535), P.max_lh);
        hipLaunchKernelGGL(harm_legendre_kernel, lgrid, dim3(64), 0, st, P.d_inst, P.d_sph,
                           P.n_inst, waves, P.d_h64, P.d_pl, P.d_e, P.d_phi, P.d_counters,
                           P.n_counters);
        if ((long long)P.max_lh * P.n_inst >= 2LL * ctx->cu_count) {
            // one block per (row, instance) over all its spheres
            const dim3 rgrid(P.max_lh, std::min(P.n_inst, 65535));
            if (P.max_rec_combos <= 2 * kRecPer * 64) {
                const RecTiles T = rec_tiles(P.max_rec_nlv, P.max_rec_ns, 64, 32, 16);
                hipLaunchKernelGGL((harm_records_kernel<64, 16>), rgrid, dim3(64), T.lds, st,
                                   P.d_inst, P.d_sph, P.n_inst, P.d_pl, P.d_e, P.d_phi, P.d_h64,
                                   P.d_rec, T.tl, T.ts);
            } else {
                const RecTiles T = rec_tiles(P.max_rec_nlv, P.max_rec_ns, 128, 64, 16);
                hipLaunchKernelGGL((harm_records_kernel<kRecThreads, 16>), rgrid,
                                   dim3(kRecThreads), T.lds, st, P.d_inst, P.d_sph, P.n_inst,
                                   P.d_pl, P.d_e, P.d_phi, P.d_h64, P.d_rec, T.tl, T.ts);
            }
        } else {
            // few instances: blocks per (row, four records, sphere)
            const dim3 rgrid(P.max_tl, (P.max_tl + kRecTile - 1) / kRecTile,
                             std::min(P.n_sph, 65535));
            hipLaunchKernelGGL(harm_records_tile_kernel, rgrid, dim3(kRecThreads), 0, st,
                               P.d_inst, P.d_sph, P.n_sph, P.d_pl, P.d_e, P.d_phi, P.d_h64,
                               P.d_rec);
        }
    } else {
        hipLaunchKernelGGL(series_waves_kernel, wgrid, dim3(256), 0, st, P.d_inst, P.d_sph, P.n_sph,
                           waves, P.d_wt, P.d_counters, P.n_counters);
    }
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

size_t prepass_lds(const PrepassArgs& a) {
    return sizeof(mie::zc) * 8 * (size_t)a.wstride * (size_t)std::min(a.spb, a.n_mie_sph);
}

int series_plan_launch_prepass(bimsim_ctx* ctx, const SeriesPlan& P, const double* waves,
                               const PrepassArgs& a) {
    if (!P.harmonic || P.n_inst == 0 || P.n_sph == 0 || a.n_mie_sph <= 0 || a.n_uf <= 0 ||
        a.spb < 1 || a.spb > mie::kSphPerLaunch || a.uf_blocks < 1 || prepass_lds(a) > kPrepassLds)
        return set_error(ctx, BIMSIM_E_INVALID, "pre-pass launch outside its limits");
    const hipStream_t st = ctx->stream;
    const dim3 lgrid((P.nk_max + 63) / 64, std::min(P.n_inst, 65535), P.max_lh);
    hipLaunchKernelGGL(harm_legendre_kernel, lgrid, dim3(64), 0, st, P.d_inst, P.d_sph, P.n_inst,
                       waves, P.d_h64, P.d_pl, P.d_e, P.d_phi, P.d_counters, P.n_counters);
    const int n_mie = (a.n_mie_sph + a.spb - 1) / a.spb;
    const int tiles = (P.max_tl + kRecTile - 1) / kRecTile;
    const long long n_rec = (long long)P.max_tl * tiles * P.n_sph;
    const long long total = n_mie + n_rec + (long long)a.uf_blocks * a.n_uf;
    if (total > INT_MAX)
        return set_error(ctx, BIMSIM_E_INVALID, "pre-pass launch outside its limits");
    hipLaunchKernelGGL(prepass_kernel, dim3((unsigned)total), dim3(kRecThreads), prepass_lds(a), st,
                       a, n_mie, (int)n_rec, P.max_tl, tiles, P.d_inst, P.d_sph, P.n_sph, P.d_pl,
                       P.d_e, P.d_phi, P.d_h64, P.d_rec);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

namespace {

// One bucket's persistent launch (the counter must be zero).
// The coefficients into the records (harm_fold_kernel).
int fold_launch(bimsim_ctx* ctx, const SeriesPlan& P, const double2* A, const double2* B) {
    const dim3 grid((P.max_tl * P.max_tl / 2 + kRecThreads - 1) / kRecThreads,
                    std::min(P.n_sph, 65535));
    hipLaunchKernelGGL(harm_fold_kernel, grid, dim3(kRecThreads), 0, ctx->stream, P.d_sph,
                       P.n_sph, A, B, P.d_rec, P.d_rec_in);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

int launch_one_bucket(bimsim_ctx* ctx, const SeriesPlan& P, const SeriesPlan::Bucket& K,
                      const double2* A, const double2* B, float2* U) {
    const hipStream_t st = ctx->stream;
    // resident workgroups: the queue keeps them busy until it drains
    const dim3 grid(std::min(K.item_count, P.groups_per_cu * ctx->cu_count));
    const HarmLaunch hl{grid, st, P.d_inst, P.d_sph, P.d_grp, P.d_items + K.item_begin,
                        K.item_count, P.d_rec, P.d_rec_in, P.d_cst, P.d_perm, U,
                        P.d_counters + K.counter, K.planar};
    int hv = 0;
    switch (K.L) {
#define BIMSIM_CASE(LL) \
    case LL:                                                \
        if (P.harmonic) hv = harm_launch<LL>(hl, K.variant); \
        else launch_bucket<LL>(grid, st, P, K, U);          \
        break;
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
            return set_error(ctx, BIMSIM_E_UNSUPPORTED, "no FP32 series kernel for %d orders", K.L);
    }
    if (hv != 0)
        return set_error(ctx, BIMSIM_E_INVALID, "no variant %d of the %d-order kernel", K.variant,
                         K.L);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // namespace

int series_plan_launch_sums(bimsim_ctx* ctx, const SeriesPlan& P, const double2* A,
                            const double2* B, float2* U, hipEvent_t after_fold) {
    const hipStream_t st = ctx->stream;
    if (P.n_inst == 0 || P.n_sph == 0) {
        if (after_fold) BIMSIM_HIP_CHECK(ctx, hipStreamWaitEvent(st, after_fold, 0));
        return BIMSIM_OK;
    }
    if (P.harmonic) BIMSIM_TRY(fold_launch(ctx, P, A, B));
    if (after_fold) BIMSIM_HIP_CHECK(ctx, hipStreamWaitEvent(st, after_fold, 0));
    for (const SeriesPlan::Chunk& ch : P.chunks) {
        if (!P.harmonic) {
            const dim3 cgrid((P.max_npix + kThreads - 1) / kThreads, std::min(ch.sph_count, 65535));
            hipLaunchKernelGGL(series_coef_kernel, cgrid, dim3(kThreads), 0, st, P.d_inst, P.d_sph,
                               ch.sph_begin, ch.sph_count, A, B, P.d_coef, P.d_rhat);
            BIMSIM_HIP_CHECK(ctx, hipGetLastError());
        }
        const int tmr = timer_begin(ctx, kStageSeries);
        for (const SeriesPlan::Bucket& K : ch.buckets) BIMSIM_TRY(launch_one_bucket(ctx, P, K, A, B, U));
        timer_end(ctx, tmr);
    }
    if (!P.split_insts.empty()) {
        const dim3 rgrid((P.max_npix + 255) / 256, std::min((int)P.split_insts.size(), 65535));
        hipLaunchKernelGGL(series_reduce_kernel, rgrid, dim3(256), 0, st, P.d_inst, P.d_split_insts,
                           (int)P.split_insts.size(), P.d_part, U);
        BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    }
    return BIMSIM_OK;
}

namespace {

// Shape of a batch for the tuning cache: per instance (bucket, pixels,
// spheres, waves) and the total.
std::string shape_key(const bimsim_ctx* ctx, const SeriesBatch& batch) {
    std::string k = "algo" + std::to_string(ctx->series_algo) + ";";
    char buf[96];
    for (const SeriesInst& I : batch.inst) {
        int nl1 = 1;
        for (int s = I.s_begin; s < I.s_begin + I.ns; ++s) nl1 = std::max(nl1, batch.sph[s].s.Nl + 1);
        std::snprintf(buf, sizeof buf, "%d:%d:%d:%d;", series_capacity(nl1), I.uR * I.vR, I.ns, I.nk);
        k += buf;
    }
    return k;
}

}  // namespace

namespace {

// The register budget of each bucket of a harmonic plan, chosen by timing
// its variants on the plan's own inputs (into a scratch field); the choice
// is kept in the context per batch shape.
int harm_tune(bimsim_ctx* ctx, const std::string& shape, const double* waves, const double2* A,
              const double2* B, SeriesPlan* P) {
    std::vector<SeriesPlan::Bucket*> bks;
    for (SeriesPlan::Chunk& ch : P->chunks)
        for (SeriesPlan::Bucket& K : ch.buckets) bks.push_back(&K);
    // a fixed variant for every bucket that has it (for comparing variants)
    if (const char* fixed = std::getenv("BIMSIM_HARM_VARIANT")) {
        const int v = std::atoi(fixed);
        for (SeriesPlan::Bucket* K : bks)
            if (harm_variant_exists(K->L, v, K->planar)) K->variant = v;
        return BIMSIM_OK;
    }
    // the key holds the buckets' orders and planar flags: the variants that
    // exist differ with them
    std::string key = shape;
    for (const SeriesPlan::Bucket* K : bks) key += "|" + std::to_string(K->L) + (K->planar ? "p" : "g");
    auto hit = ctx->tune_harm.find(key);
    if (hit != ctx->tune_harm.end() && hit->second.size() == bks.size()) {
        for (size_t i = 0; i < bks.size(); ++i) bks[i]->variant = hit->second[i];
        return BIMSIM_OK;
    }
    const long long field = P->field_elems;
    float2* tmp = nullptr;
    if (hipMalloc((void**)&tmp, sizeof(float2) * (size_t)std::max<long long>(1, field)) != hipSuccess)
        return set_error(ctx, BIMSIM_E_NOMEM, "out of device memory for tuning");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const bool timing = ctx->timing;
    ctx->timing = false;
    int rc = zero_launch(ctx, tmp, sizeof(float2) * (size_t)std::max<long long>(1, field));
    if (rc == BIMSIM_OK) rc = series_plan_launch_waves(ctx, *P, waves);
    if (rc == BIMSIM_OK) rc = fold_launch(ctx, *P, A, B);
    std::vector<int> chosen;
    for (SeriesPlan::Bucket* K : bks) {
        if (rc != BIMSIM_OK) break;
        float best = 1e30f;
        int best_v = K->variant;
        for (int v = 0; v < kHarmVariants && rc == BIMSIM_OK; ++v) {
            if (!harm_variant_tuned(K->L, v, K->planar)) continue;
            K->variant = v;
            // a warm run, then the best of three timed ones (variants a few
            // per cent apart were ranked wrongly by one or two)
            float ms = 1e30f;
            for (int r = 0; r < 4 && rc == BIMSIM_OK; ++r) {
                rc = zero_launch(ctx, P->d_counters + K->counter, sizeof(int));
                if (rc != BIMSIM_OK) break;
                (void)hipEventRecord(e0, ctx->stream);
                rc = launch_one_bucket(ctx, *P, *K, A, B, tmp);
                (void)hipEventRecord(e1, ctx->stream);
                (void)hipEventSynchronize(e1);
                float t = 0.f;
                (void)hipEventElapsedTime(&t, e0, e1);
                if (r > 0) ms = std::min(ms, t);
            }
            if (rc == BIMSIM_OK && ms < best) {
                best = ms;
                best_v = v;
            }
        }
        K->variant = best_v;
        chosen.push_back(best_v);
    }
    ctx->timing = timing;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(tmp);
    BIMSIM_TRY(rc);
    ctx->tune_harm[key] = chosen;
    if (std::getenv("BIMSIM_TUNE_LOG")) {
        std::string msg;
        for (size_t i = 0; i < bks.size(); ++i)
            msg += " L" + std::to_string(bks[i]->L) + ":" +
                   std::to_string(harm_variant_waves(bks[i]->L, chosen[i])) +
                   (chosen[i] >= 3 ? "/c" + std::to_string(harm_variant_chunk(bks[i]->L, chosen[i]))
                                   : std::string()) +
                   (harm_variant_mfma(chosen[i]) ? "m" : "");
        std::fprintf(stderr, "bimsim: harmonic waves per SIMD%s\n", msg.c_str());
    }
    return BIMSIM_OK;
}

}  // namespace

int series_plan_create_tuned(bimsim_ctx* ctx, SeriesBatch& batch, size_t coef_budget,
                             const double* waves, const double2* A, const double2* B, SeriesPlan* P) {
    // the heuristic plan first: long queues keep it
    SeriesBatch work = batch;
    BIMSIM_TRY(series_plan_create(ctx, work, coef_budget, P));
    long long items = 0;
    for (const SeriesPlan::Chunk& ch : P->chunks)
        for (const SeriesPlan::Bucket& K : ch.buckets) items += K.item_count;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(ctx->stream, &cap);
    if (P->harmonic) {
        // the harmonic form tunes the register budget of each bucket
        batch = work;
        if (!ctx->autotune || cap != hipStreamCaptureStatusNone) return BIMSIM_OK;
        return harm_tune(ctx, shape_key(ctx, batch) + "|harm", waves, A, B, P);
    }
    if (!ctx->autotune || items > 16LL * ctx->cu_count || cap != hipStreamCaptureStatusNone) {
        batch = work;
        return BIMSIM_OK;
    }
    const std::string key = shape_key(ctx, batch);
    auto hit = ctx->tune.find(key);
    int best_mult = P->split_mult, best_gpc = P->groups_per_cu;
    if (hit == ctx->tune.end()) {
        // time the candidates into a temporary field
        long long field = 0;
        for (const SeriesInst& I : batch.inst) field = std::max(field, I.u_off + (long long)I.uR * I.vR);
        float2* tmp = nullptr;
        if (hipMalloc((void**)&tmp, sizeof(float2) * (size_t)std::max<long long>(1, field)) != hipSuccess)
            return set_error(ctx, BIMSIM_E_NOMEM, "out of device memory for tuning");
        hipEvent_t e0, e1;
        (void)hipEventCreate(&e0);
        (void)hipEventCreate(&e1);
        const bool timing = ctx->timing;
        ctx->timing = false;
        float best = 1e30f;
        int rc = BIMSIM_OK;
        for (int mult : {1, 2, 4, 8, 16}) {
            for (int gpc : {2, 4, 8}) {
                SeriesBatch cand = batch;
                for (SeriesInst& I : cand.inst) I.accumulate = 0;
                SeriesPlan T;
                rc = series_plan_create(ctx, cand, coef_budget, &T, mult, gpc);
                if (rc != BIMSIM_OK) break;
                if (mult > 1 && T.split_insts.empty()) {   // no instance splits further
                    series_plan_destroy(&T);
                    continue;
                }
                rc = series_plan_launch(ctx, T, waves, A, B, tmp);   // warm
                (void)hipEventRecord(e0, ctx->stream);
                for (int r = 0; r < 2 && rc == BIMSIM_OK; ++r) rc = series_plan_launch(ctx, T, waves, A, B, tmp);
                (void)hipEventRecord(e1, ctx->stream);
                (void)hipEventSynchronize(e1);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                series_plan_destroy(&T);
                if (rc != BIMSIM_OK) break;
                if (ms < best) {
                    best = ms;
                    best_mult = mult;
                    best_gpc = gpc;
                }
            }
            if (rc != BIMSIM_OK) break;
        }
        ctx->timing = timing;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipFree(tmp);
        BIMSIM_TRY(rc);
        ctx->tune[key] = std::make_pair(best_mult, best_gpc);
    } else {
        best_mult = hit->second.first;
        best_gpc = hit->second.second;
    }
    if (best_mult == P->split_mult && best_gpc == P->groups_per_cu) {
        batch = work;
        return BIMSIM_OK;
    }
    series_plan_destroy(P);
    return series_plan_create(ctx, batch, coef_budget, P, best_mult, best_gpc);
}

int series_cached_plan(bimsim_ctx* ctx, SeriesBatch& batch, const double* waves, const double2* A,
                       const double2* B, const SeriesPlan** out) {
    std::vector<char> key(sizeof(SeriesInst) * batch.inst.size() + sizeof(SeriesSph) * batch.sph.size() +
                          sizeof(int));
    std::memcpy(key.data() + key.size() - sizeof(int), &ctx->series_algo, sizeof(int));
    if (!batch.inst.empty())
        std::memcpy(key.data(), batch.inst.data(), sizeof(SeriesInst) * batch.inst.size());
    if (!batch.sph.empty())
        std::memcpy(key.data() + sizeof(SeriesInst) * batch.inst.size(), batch.sph.data(),
                    sizeof(SeriesSph) * batch.sph.size());
    SeriesCache* c = (SeriesCache*)ctx->series_cache;
    if (!c) {
        c = new SeriesCache();
        ctx->series_cache = c;
    }
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    BIMSIM_HIP_CHECK(ctx, hipStreamIsCapturing(ctx->stream, &cap));
    const bool capturing = cap != hipStreamCaptureStatusNone;
    for (SeriesCacheEntry* e : c->entries)
        if (e->key == key) {
            e->last_use = ++c->clock;
            e->pinned = e->pinned || capturing;
            *out = &e->plan;
            return BIMSIM_OK;
        }
    // a new plan allocates and times launches: not possible inside a capture
    if (capturing)
        return set_error(ctx, BIMSIM_E_UNSUPPORTED,
                         "scattered-field plan for this call is not cached: run the call once "
                         "outside the capture first");
    int unpinned = 0;
    for (SeriesCacheEntry* e : c->entries) unpinned += e->pinned ? 0 : 1;
    if (unpinned >= kPlanCacheSize) {
        size_t victim = c->entries.size();
        for (size_t i = 0; i < c->entries.size(); ++i)
            if (!c->entries[i]->pinned &&
                (victim == c->entries.size() || c->entries[i]->last_use < c->entries[victim]->last_use))
                victim = i;
        BIMSIM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        series_plan_destroy(&c->entries[victim]->plan);
        delete c->entries[victim];
        c->entries.erase(c->entries.begin() + victim);
    }
    SeriesCacheEntry* e = new SeriesCacheEntry();
    const int rc = series_plan_create_tuned(ctx, batch, (size_t)8 << 30, waves, A, B, &e->plan);
    if (rc != BIMSIM_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        series_plan_destroy(&e->plan);
        delete e;
        return rc;
    }
    e->key = std::move(key);
    e->last_use = ++c->clock;
    c->entries.push_back(e);
    *out = &e->plan;
    return BIMSIM_OK;
}

void series_cache_release(bimsim_ctx* ctx) {
    SeriesCache* c = (SeriesCache*)ctx->series_cache;
    if (!c) return;
    for (SeriesCacheEntry* e : c->entries) {
        series_plan_destroy(&e->plan);
        delete e;
    }
    delete c;
    ctx->series_cache = nullptr;
}

}  // namespace bimsim
