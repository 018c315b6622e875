// Device code of the FP32 harmonic scattered-field series (the per-pixel
// kernel of us_series.hip's plans), compiled once per order bucket L by
// us_harm.hip, each bucket in several register budgets (waves per SIMD)
// that the plan's tuner chooses between.  Every file that includes it is
// built with -ffp-contract=off; the FP32 FMAs are written out.
#pragma once

#include "sincos_fast.h"
#include "us_series.h"

namespace bimsim {
namespace harm {

constexpr int kHarmThreads = 256;
// row stride of the recurrence-constant table of the general form (m, l < 112)
constexpr int kCS = 112;

__device__ inline dc ld2(const double2* p, int i) {
    const double2 v = p[i];
    return dc{v.x, v.y};
}

// Trigonometric start values of the radial recurrences of one pixel and
// sphere, evaluated before the order sums so that their temporaries do not
// coincide with the sums' registers: sin, cos of kd outside the sphere;
// sin, cos of Re knd and sinh, cosh of Im knd inside.
struct RadialTrig {
    double t0, t1, t2, t3;
};

__device__ __forceinline__ RadialTrig harm_trig(const SphereDev& S, double kmag, double d) {
    RadialTrig T;
    if (d <= S.a) {
        const dc knd = cscl(dc{S.n[0], S.n[1]}, kmag * d);
        T.t0 = sin(knd.r);
        T.t1 = cos(knd.r);
        T.t2 = sinh(knd.i);
        T.t3 = cosh(knd.i);
    } else {
        const double kd = kmag * d;
        if (__all(kd < 1048576.0)) sincos_reduced(kd, &T.t0, &T.t1);
        else sincos(kd, &T.t0, &T.t1);
        T.t2 = T.t3 = 0.0;
    }
    return T;
}

// Radial factors of one pixel and sphere applied to its order sums:
// acc += sum_{l <= Nl} c_l (ar_l + i ai_l), c_l = B_l (j_l + i y_l)(kd)
// outside the sphere, A_l j_l(knd) inside (calc_Us / calc_Ui,
// nfScalarUs.cu:50-95,201-245), in FP64 with the reference's upward
// recurrences and threshold rules (spherical_bessel.h:85-141).  The steps
// multiply by (2l - 1)/kd (or /knd) formed from the reciprocal where the
// reference divides, with FMAs.  The threshold rules act only where l > kd
// (the y lower bound only where y is that large, which needs l > kd as
// well): an order skips its selects when no lane of the wave is past kd.
// The recurrences' state carries from one range of orders to the next.
// One set of registers serves both cases: outside, cj0 = (j_{l-2}, y_{l-2}),
// cj1 = (j_{l-1}, y_{l-1}) and cinv.r = 1/kd; inside, cj0 and cj1 are the
// complex j of the last two orders and cinv = 1/knd.
struct RadialState {
    int mode;          // 0 outside the sphere, 1 inside, 2 at the centre (l = 0 only)
    dc cj0, cj1, cinv;
    double kx;         // kd (outside) or Re knd (inside)
    int kfl;           // floor(kx), at most 2^30: order l is past kx where l > kfl
};

// whether the condition holds on some lane of the wave: the ballot of the
// lane mask itself (HIP's __any(int) turns the mask into 0/1 values and
// compares them, two VALU instructions)
__device__ __forceinline__ bool wave_any(bool b) { return __builtin_amdgcn_ballot_w64(b) != 0; }

__device__ __forceinline__ int harm_floor_order(double kx) {
    return kx < 1073741824.0 ? (int)kx : 1073741824;
}

__device__ __forceinline__ RadialState harm_radial_init(const SphereDev& S, double kmag, double d,
                                                        double inv_d, double inv_kmag,
                                                        const RadialTrig& T) {
    RadialState st;
    if (d <= S.a) {
        const dc knd = cscl(dc{S.n[0], S.n[1]}, kmag * d);
        st.kx = knd.r;
        st.kfl = harm_floor_order(knd.r);
        // calc_Ui returns A_0 for every wave at the centre: only l = 0
        st.mode = knd.r < (double)kEpsFloat ? 2 : 1;
        // init_jc / shift_jc (csin, ccos as csin_r / ccos_r)
        st.cinv = cdiv_naive(dc{1.0, 0.0}, knd);
        const dc sn = {T.t0 * T.t3, T.t1 * T.t2}, cs = {T.t1 * T.t3, -(T.t0 * T.t2)};
        st.cj0 = cmul(sn, st.cinv);
        st.cj1 = csub(cmul(st.cj0, st.cinv), cmul(cs, st.cinv));
        return st;
    }
    st.mode = 0;
    st.kx = kmag * d;
    st.kfl = harm_floor_order(st.kx);
    const double inv = inv_d * inv_kmag;
    const double sn = T.t0, cs = T.t1;
    const double j0 = sn * inv, y0 = -(cs * inv);
    st.cinv = dc{inv, 0.0};
    st.cj0 = dc{j0, y0};
    st.cj1 = dc{fma(j0, inv, -(cs * inv)), fma(y0, inv, -(sn * inv))};
    return st;
}

// The outside-sphere steps of harm_radial_range.  The recurrences keep
// even orders in (je, ye) and odd ones in (jo, yo), each step overwriting
// the order two below (the registers of a skipped order then stay as they
// are, with no moves).  kCheck: the threshold rules (shift_sbesselj /
// shift_sbessely, spherical_bessel.h:107-141: where l > kd, j outside
// [kJMin, kJMax] becomes 0; y below kYMin, or positive past kd, becomes 0
// with the order below it), on ranges where some lane of the wave is past
// kd at one of their orders; their conditions are combined bitwise (no
// branches per lane), and y's selects run only where a lane needs them (in
// practice none).  kAll (only without the rules): every order of the
// range is stepped and summed, also those past Nl, whose order sums are
// zero there (the record rows hold zeros past Nl), so that they add
// nothing; the values stay moderate at orders no lane is past kd at, and
// the range takes no branch per order (nor the register copies at their
// merges).
template <int L0, int C, bool kCheck, bool kAll = false>
__device__ __forceinline__ void harm_radial_out(RadialState& st, int nl1, const float* ar,
                                                const float* ai, float& acc_r, float& acc_i) {
    static_assert(L0 % 2 == 0 && C % 2 == 0, "ranges start at even orders");
    static_assert(!(kCheck && kAll), "the threshold rules keep the orders past Nl out");
    const double inv = st.cinv.r;
    double je = st.cj0.r, ye = st.cj0.i, jo = st.cj1.r, yo = st.cj1.i;
#pragma clang loop unroll(full)
    for (int i = 0; i < C; ++i) {
        const int l = L0 + i;
        if (!kAll && l >= nl1) continue;
        const bool even = (l & 1) == 0;
        double& jn = even ? je : jo;   // order l - 2, then l
        double& yn = even ? ye : yo;
        double& jp = even ? jo : je;   // order l - 1
        double& yp = even ? yo : ye;
        if (l >= 2) {
            const double t = (double)(2 * l - 1) * inv;
            jn = fma(t, jp, -jn);
            yn = fma(t, yp, -yn);
            if constexpr (kCheck) {
                // j: a compare and a select on every order of the range (a
                // branch around them cost more in register moves at its
                // merge); y: the lane masks of the compares combined as
                // ballots (scalar), the selects in a branch of their own
                const bool past = l > st.kfl;
                jn = past & ((jn < kJMin) | (jn > kJMax)) ? 0.0 : jn;
                const bool ylo = yn < kYMin, ypos = yn > 0.0;
                const uint64_t yz = __builtin_amdgcn_ballot_w64(ylo) |
                                    (__builtin_amdgcn_ballot_w64(past) &
                                     __builtin_amdgcn_ballot_w64(ypos));
                if (yz != 0) {
                    // (kept a branch: the compiler would run these selects
                    // on every order)
                    asm volatile("");
                    const bool z = ylo | (past & ypos);
                    yn = z ? 0.0 : yn;
                    yp = z ? 0.0 : yp;
                }
            }
        }
        const float fr = (float)jn, fi = (float)yn;
        acc_r = fmaf(fr, ar[i], fmaf(-fi, ai[i], acc_r));
        acc_i = fmaf(fr, ai[i], fmaf(fi, ar[i], acc_i));
    }
    st.cj0 = dc{je, ye};
    st.cj1 = dc{jo, yo};
}

// Orders L0 .. L0 + C - 1 (those <= Nl) of the radial sum; ar / ai hold
// the order sums of these orders at [l - L0], formed from records that
// already carry the coefficient: B_l for the pixels outside the sphere
// (pass 0), A_l for those inside it (pass 1).  A lane takes part only in
// its own pass: outside, acc += (j_l + i y_l)(kd) times the sum; inside,
// acc += j_l(knd) times the sum; at the centre (knd below float epsilon)
// only the l = 0 sum, A_0 being in it (calc_Ui returns A_0 there).
template <int L0, int C>
__device__ __forceinline__ void harm_radial_range(int pass, RadialState& st, const SphereDev& S,
                                                  const float* ar, const float* ai, float& acc_r,
                                                  float& acc_i) {
    const int nl1 = S.Nl + 1;
    auto mac = [&](int i, double cr, double ci) {
        const float fr = (float)cr, fi = (float)ci;
        acc_r = fmaf(fr, ar[i], fmaf(-fi, ai[i], acc_r));
        acc_i = fmaf(fr, ai[i], fmaf(fi, ar[i], acc_i));
    };
    if (st.mode == 2) {
        if (pass == 1 && L0 == 0) {
            acc_r += ar[0];
            acc_i += ai[0];
        }
        return;
    }
    if (st.mode == 1) {
        if (pass != 1) return;
        const dc inv = st.cinv;
#pragma clang loop unroll(full)
        for (int i = 0; i < C; ++i) {
            const int l = L0 + i;
            if (l >= nl1) continue;
            dc jl = l == 0 ? st.cj0 : st.cj1;
            if (l >= 2) {
                // (2l - 1)/knd times j_{l-1}, minus j_{l-2}
                const double tr = (double)(2 * l - 1) * inv.r, ti = (double)(2 * l - 1) * inv.i;
                dc jn = {fma(tr, st.cj1.r, fma(-ti, st.cj1.i, -st.cj0.r)),
                         fma(tr, st.cj1.i, fma(ti, st.cj1.r, -st.cj0.i))};
                // the threshold rule can only act where l > Re knd: skip its
                // selects when no lane of the wave is there
                const bool past = (double)l > st.kx;
                if (__any(past))
                    if (past && (jn.r < kJMin || jn.r > kJMax)) jn = dc{0.0, 0.0};
                st.cj0 = st.cj1;
                st.cj1 = jn;
                jl = jn;
            }
            mac(i, jl.r, jl.i);
        }
        return;
    }
    if (pass != 0) return;
    // the rules only where some lane of the wave is past kd at an order of
    // the range (orders l <= kfl of every lane: none)
    if (wave_any(L0 + C - 1 > st.kfl))
        harm_radial_out<L0, C, true>(st, nl1, ar, ai, acc_r, acc_i);
    else
        harm_radial_out<L0, C, false>(st, nl1, ar, ai, acc_r, acc_i);
}

// All orders at once (the sums of orders 0 .. L - 1 in ar / ai).
template <int L>
__device__ __forceinline__ void harm_radial(int pass, const SphereDev& S, double kmag, double d,
                                            double inv_d, double inv_kmag, const RadialTrig& T,
                                            const float* ar, const float* ai, float& acc_r,
                                            float& acc_i) {
    RadialState st = harm_radial_init(S, kmag, d, inv_d, inv_kmag, T);
    harm_radial_range<0, L>(pass, st, S, ar, ai, acc_r, acc_i);
}

// Order sums of a planar sphere (x = 0 at every pixel): the Legendre
// values are in the records, so row m only rotates them by e^{i m phi}:
//   ar_l + i ai_l += cos(m phi) U1_lm + sin(m phi) U2_lm,  l = m, m+2, ...
// Rows in pairs (m even: even l; m + 1: odd l); records at
// R[m (L/2) + k], l = 2k + (m & 1); one scalar load of two records per
// row and block of two k, blocks below the diagonal or past Nl skipped.
template <int L>
__device__ __forceinline__ void harm_planar_sums(const float4* __restrict__ R, int nl1, float z1r,
                                                 float z1i, float* ar, float* ai) {
    constexpr int H = L / 2;
    constexpr int KB = 2;            // k per block
    const int ke = (nl1 + 1) >> 1;   // k past the last of row 0
    float zr = 1.f, zi = 0.f;
#pragma clang loop unroll(disable)
    for (int m = 0; m < nl1; m += 2) {
        const float4* Ra = R + m * H;
        const float4* Rb = Ra + H;
        const float zar = zr, zai = zi;
        const float zbr = fmaf(zr, z1r, -(zi * z1i)), zbi = fmaf(zr, z1i, zi * z1r);
        const int bf = (m >> 1) / KB;   // block of k = m / 2
#pragma clang loop unroll(full)
        for (int b = 0; b < H / KB; ++b) {
            if (b < bf || KB * b >= ke) continue;
            float4 ra[KB], rb[KB];
#pragma clang loop unroll(full)
            for (int j = 0; j < KB; ++j) {
                ra[j] = Ra[KB * b + j];
                rb[j] = Rb[KB * b + j];
            }
#pragma clang loop unroll(full)
            for (int j = 0; j < KB; ++j) {
                const int le = 2 * (KB * b + j);
                ar[le] = fmaf(zai, ra[j].z, fmaf(zar, ra[j].x, ar[le]));
                ai[le] = fmaf(zai, ra[j].w, fmaf(zar, ra[j].y, ai[le]));
                ar[le + 1] = fmaf(zbi, rb[j].z, fmaf(zbr, rb[j].x, ar[le + 1]));
                ai[le + 1] = fmaf(zbi, rb[j].w, fmaf(zbr, rb[j].y, ai[le + 1]));
            }
        }
        zr = fmaf(zbr, z1r, -(zbi * z1i));
        zi = fmaf(zbr, z1i, zbi * z1r);
    }
}


// The planar order sums of orders L0 .. L0 + C - 1 only (L0 a multiple of
// 4): rows m < L0 + C (phases from e^{i 0 phi} again), the records of
// k = L0/2 .. (L0 + C)/2 - 1 in each, in the blocks of two k of
// harm_planar_sums; ar / ai at [l - L0].  Every sum meets the same terms
// in the same order as in harm_planar_sums.
template <int L, int C, int L0>
__device__ __forceinline__ void harm_planar_chunk(const float4* __restrict__ R, int nl1, float z1r,
                                                  float z1i, float* ar, float* ai) {
    constexpr int H = L / 2;
    constexpr int KB = 2;
    constexpr int K0 = L0 / 2;
    constexpr int NB = ((L0 + C <= L ? C : L - L0) / 2 + KB - 1) / KB;   // blocks in the chunk
    const int ke = (nl1 + 1) >> 1;
    const int mend = min(nl1, L0 + C);
    float zr = 1.f, zi = 0.f;
#pragma clang loop unroll(disable)
    for (int m = 0; m < mend; m += 2) {
        const float4* Ra = R + m * H;
        const float4* Rb = Ra + H;
        const float zar = zr, zai = zi;
        const float zbr = fmaf(zr, z1r, -(zi * z1i)), zbi = fmaf(zr, z1i, zi * z1r);
        const int bf = (m >> 1) / KB;   // block of k = m / 2
#pragma clang loop unroll(full)
        for (int b = 0; b < NB; ++b) {
            constexpr int gb0 = K0 / KB;
            const int gb = gb0 + b;
            if (gb < bf || KB * gb >= ke) continue;
            float4 ra[KB], rb[KB];
#pragma clang loop unroll(full)
            for (int j = 0; j < KB; ++j) {
                ra[j] = Ra[KB * gb + j];
                rb[j] = Rb[KB * gb + j];
            }
#pragma clang loop unroll(full)
            for (int j = 0; j < KB; ++j) {
                const int le = 2 * (KB * b + j);
                ar[le] = fmaf(zai, ra[j].z, fmaf(zar, ra[j].x, ar[le]));
                ai[le] = fmaf(zai, ra[j].w, fmaf(zar, ra[j].y, ai[le]));
                ar[le + 1] = fmaf(zbi, rb[j].z, fmaf(zbr, rb[j].x, ar[le + 1]));
                ai[le + 1] = fmaf(zbi, rb[j].w, fmaf(zbr, rb[j].y, ai[le + 1]));
            }
        }
        zr = fmaf(zbr, z1r, -(zbi * z1i));
        zi = fmaf(zbr, z1i, zbi * z1r);
    }
}

// Chunk c of the chunked planar form and the ones after it (one pass).
template <int L, int C, int c>
__device__ __forceinline__ void harm_planar_chunks(const float4* __restrict__ R, const SphereDev& S,
                                                   float z1r, float z1i, int pass, RadialState& st,
                                                   float& acc_r, float& acc_i) {
    if constexpr (c * C < L) {
        const int nl1 = S.Nl + 1;
        if (c * C >= nl1) return;
        float ar[C], ai[C];
#pragma unroll
        for (int i = 0; i < C; ++i) ar[i] = ai[i] = 0.f;
        harm_planar_chunk<L, C, c * C>(R, nl1, z1r, z1i, ar, ai);
        harm_radial_range<c * C, C>(pass, st, S, ar, ai, acc_r, acc_i);
        harm_planar_chunks<L, C, c + 1>(R, S, z1r, z1i, pass, st, acc_r, acc_i);
    }
}


// ---- The planar order sums on the matrix cores ----
// v_mfma_f32_4x4x1_16b_f32 with its A operand broadcast from one block of
// four lanes (CBSZ 4, ABID a) computes in every lane
//   d[i] = fmaf(A[4a + i], b, d[i]),  i = 0..3,
// A[4a + i] being lane 4a + i's value and b the lane's own: the four FMAs
// of one record quad and one pixel, rounded as fmaf, in one instruction at
// twice the issue rate of four VALU FMAs.  A lane keeps its pixel: b is its
// cos(m phi) or sin(m phi), d four of its order sums.  The instruction
// holds the SIMD for its 8 cycles (VALU work of other waves does not run
// beside it), so only the products get faster; the radial factors and the
// phase steps stay on the VALU, and the tuner takes these variants where
// they win.
typedef float f32x4 __attribute__((ext_vector_type(4)));

// One row pair (m, m + 1) of the products for n blocks of a chunk, as one
// asm statement: blocks n - 1 .. 0 in turn, each adding its cos terms
// (records of slots 2b in ae for row m, in ao for row m + 1, times zar,
// zbr), then the same blocks their sin terms (slots 2b + 1, times zai,
// zbi), so that each sum takes its cos term before its sin term as
// harm_planar_chunk does.  Block b runs while first <= b <= last (first:
// the row pair's first block of the chunk, below which the records are
// zero; last: the block of the sphere's order Nl, past which they are zero
// too): each cascade enters at block min(last, n - 1) by a scalar compare
// per block above it, and leaves at the first block below first.  The
// sums stay in place in AGPRs: the MFMAs are written with their
// accumulators tied, which the compiler's forms of them are not (it then
// moves the sums through other registers with copies every row).  The
// compiler does not see into the asm, so the statement keeps its own wait
// states: two before the first product (a VALU write of an operand just
// before it), two before the sin terms (a sum's cos term written just
// before), one after the last (a VALU write of an operand read by the
// last product); the sums are first read after `settle` (eight more), and
// tests/test_mfma_wait_states.py checks in every bucket's compiled code
// that nothing else touches them sooner (the compiler did, where it
// spilled them).  Slot layouts: packed (both rows in one register, row
// m + 1 from lane 32) or one register per row.
#define QUAD_C(b, ce, co)                                                                     \
    ".Lqcb" #b "_%=:\n\t"                                                                      \
    "s_cmp_gt_i32 %[first], " #b "\n\ts_cbranch_scc1 .Lqc%=\n\t"                               \
    "v_mfma_f32_4x4x1_16b_f32 %[e" #b "], %[ae], %[zar], %[e" #b "] cbsz:4 abid:" #ce "\n\t" \
    "v_mfma_f32_4x4x1_16b_f32 %[o" #b "], %[ao], %[zbr], %[o" #b "] cbsz:4 abid:" #co "\n\t"
#define QUAD_S(b, ce, co)                                                                     \
    ".Lqsb" #b "_%=:\n\t"                                                                      \
    "s_cmp_gt_i32 %[first], " #b "\n\ts_cbranch_scc1 .Lqs%=\n\t"                               \
    "v_mfma_f32_4x4x1_16b_f32 %[e" #b "], %[ae], %[zai], %[e" #b "] cbsz:4 abid:" #ce "\n\t" \
    "v_mfma_f32_4x4x1_16b_f32 %[o" #b "], %[ao], %[zbi], %[o" #b "] cbsz:4 abid:" #co "\n\t"
// the entry of a cascade (x: c or s): the highest block b <= last, or
// none (last < 0)
#define QUAD_DB(x, b) "s_cmp_le_i32 " #b ", %[last]\n\ts_cbranch_scc1 .Lq" #x "b" #b "_%=\n\t"
#define QUAD_D1(x) QUAD_DB(x, 0) "s_branch .Lq" #x "%=\n\t"
#define QUAD_D2(x) QUAD_DB(x, 1) QUAD_D1(x)
#define QUAD_D3(x) QUAD_DB(x, 2) QUAD_D2(x)
#define QUAD_D4(x) QUAD_DB(x, 3) QUAD_D3(x)
#define QUAD_D5(x) QUAD_DB(x, 4) QUAD_D4(x)
#define QUAD_D6(x) QUAD_DB(x, 5) QUAD_D5(x)
#define QUAD_D7(x) QUAD_DB(x, 6) QUAD_D6(x)
#define QUAD_D8(x) QUAD_DB(x, 7) QUAD_D7(x)
// the first row pair of a chunk (every block; the cos products start the
// sums: accumulator 0)
#define QUAD_C0(b, ce, co)                                                                    \
    "v_mfma_f32_4x4x1_16b_f32 %[e" #b "], %[ae], %[zar], 0 cbsz:4 abid:" #ce "\n\t"           \
    "v_mfma_f32_4x4x1_16b_f32 %[o" #b "], %[ao], %[zbr], 0 cbsz:4 abid:" #co "\n\t"
#define QUAD_S0(b, ce, co)                                                                    \
    "v_mfma_f32_4x4x1_16b_f32 %[e" #b "], %[ae], %[zai], %[e" #b "] cbsz:4 abid:" #ce "\n\t" \
    "v_mfma_f32_4x4x1_16b_f32 %[o" #b "], %[ao], %[zbi], %[o" #b "] cbsz:4 abid:" #co "\n\t"
// packed slots: row m's slot s at broadcast block s, row m + 1's at 8 + s
#define QUAD_CP1 QUAD_C(0, 0, 8)
#define QUAD_CP2 QUAD_C(1, 2, 10) QUAD_CP1
#define QUAD_CP3 QUAD_C(2, 4, 12) QUAD_CP2
#define QUAD_CP4 QUAD_C(3, 6, 14) QUAD_CP3
#define QUAD_SP1 QUAD_S(0, 1, 9)
#define QUAD_SP2 QUAD_S(1, 3, 11) QUAD_SP1
#define QUAD_SP3 QUAD_S(2, 5, 13) QUAD_SP2
#define QUAD_SP4 QUAD_S(3, 7, 15) QUAD_SP3
// a register per row: slot s at broadcast block s of each
#define QUAD_CU1 QUAD_C(0, 0, 0)
#define QUAD_CU2 QUAD_C(1, 2, 2) QUAD_CU1
#define QUAD_CU3 QUAD_C(2, 4, 4) QUAD_CU2
#define QUAD_CU4 QUAD_C(3, 6, 6) QUAD_CU3
#define QUAD_CU5 QUAD_C(4, 8, 8) QUAD_CU4
#define QUAD_CU6 QUAD_C(5, 10, 10) QUAD_CU5
#define QUAD_CU7 QUAD_C(6, 12, 12) QUAD_CU6
#define QUAD_CU8 QUAD_C(7, 14, 14) QUAD_CU7
#define QUAD_SU1 QUAD_S(0, 1, 1)
#define QUAD_SU2 QUAD_S(1, 3, 3) QUAD_SU1
#define QUAD_SU3 QUAD_S(2, 5, 5) QUAD_SU2
#define QUAD_SU4 QUAD_S(3, 7, 7) QUAD_SU3
#define QUAD_SU5 QUAD_S(4, 9, 9) QUAD_SU4
#define QUAD_SU6 QUAD_S(5, 11, 11) QUAD_SU5
#define QUAD_SU7 QUAD_S(6, 13, 13) QUAD_SU6
#define QUAD_SU8 QUAD_S(7, 15, 15) QUAD_SU7
#define QUAD_CP01 QUAD_C0(0, 0, 8)
#define QUAD_SP01 QUAD_S0(0, 1, 9)
#define QUAD_CP02 QUAD_C0(1, 2, 10) QUAD_CP01
#define QUAD_SP02 QUAD_S0(1, 3, 11) QUAD_SP01
#define QUAD_CP03 QUAD_C0(2, 4, 12) QUAD_CP02
#define QUAD_SP03 QUAD_S0(2, 5, 13) QUAD_SP02
#define QUAD_CP04 QUAD_C0(3, 6, 14) QUAD_CP03
#define QUAD_SP04 QUAD_S0(3, 7, 15) QUAD_SP03
#define QUAD_CU01 QUAD_C0(0, 0, 0)
#define QUAD_SU01 QUAD_S0(0, 1, 1)
#define QUAD_CU02 QUAD_C0(1, 2, 2) QUAD_CU01
#define QUAD_SU02 QUAD_S0(1, 3, 3) QUAD_SU01
#define QUAD_CU03 QUAD_C0(2, 4, 4) QUAD_CU02
#define QUAD_SU03 QUAD_S0(2, 5, 5) QUAD_SU02
#define QUAD_CU04 QUAD_C0(3, 6, 6) QUAD_CU03
#define QUAD_SU04 QUAD_S0(3, 7, 7) QUAD_SU03
#define QUAD_CU05 QUAD_C0(4, 8, 8) QUAD_CU04
#define QUAD_SU05 QUAD_S0(4, 9, 9) QUAD_SU04
#define QUAD_CU06 QUAD_C0(5, 10, 10) QUAD_CU05
#define QUAD_SU06 QUAD_S0(5, 11, 11) QUAD_SU05
#define QUAD_CU07 QUAD_C0(6, 12, 12) QUAD_CU06
#define QUAD_SU07 QUAD_S0(6, 13, 13) QUAD_SU06
#define QUAD_CU08 QUAD_C0(7, 14, 14) QUAD_CU07
#define QUAD_SU08 QUAD_S0(7, 15, 15) QUAD_SU07
#define QUAD_OUT1 [e0] "=a"(E[0]), [o0] "=a"(O[0])
#define QUAD_OUT2 QUAD_OUT1, [e1] "=a"(E[1]), [o1] "=a"(O[1])
#define QUAD_OUT3 QUAD_OUT2, [e2] "=a"(E[2]), [o2] "=a"(O[2])
#define QUAD_OUT4 QUAD_OUT3, [e3] "=a"(E[3]), [o3] "=a"(O[3])
#define QUAD_OUT5 QUAD_OUT4, [e4] "=a"(E[4]), [o4] "=a"(O[4])
#define QUAD_OUT6 QUAD_OUT5, [e5] "=a"(E[5]), [o5] "=a"(O[5])
#define QUAD_OUT7 QUAD_OUT6, [e6] "=a"(E[6]), [o6] "=a"(O[6])
#define QUAD_OUT8 QUAD_OUT7, [e7] "=a"(E[7]), [o7] "=a"(O[7])
#define QUAD_OPS1 [e0] "+a"(E[0]), [o0] "+a"(O[0])
#define QUAD_OPS2 QUAD_OPS1, [e1] "+a"(E[1]), [o1] "+a"(O[1])
#define QUAD_OPS3 QUAD_OPS2, [e2] "+a"(E[2]), [o2] "+a"(O[2])
#define QUAD_OPS4 QUAD_OPS3, [e3] "+a"(E[3]), [o3] "+a"(O[3])
#define QUAD_OPS5 QUAD_OPS4, [e4] "+a"(E[4]), [o4] "+a"(O[4])
#define QUAD_OPS6 QUAD_OPS5, [e5] "+a"(E[5]), [o5] "+a"(O[5])
#define QUAD_OPS7 QUAD_OPS6, [e6] "+a"(E[6]), [o6] "+a"(O[6])
#define QUAD_OPS8 QUAD_OPS7, [e7] "+a"(E[7]), [o7] "+a"(O[7])

template <int N, bool kPacked>
struct QuadRow;
#define QUAD_ROW(N, P, CASC_C, CASC_S, SEED_C, SEED_S)                                           \
    template <>                                                                                  \
    struct QuadRow<N, P> {                                                                       \
        __device__ static __forceinline__ void run(f32x4* E, f32x4* O, float ae, float ao,       \
                                                   float zar, float zai, float zbr, float zbi,   \
                                                   int first, int last) {                        \
            asm volatile("s_nop 1\n\t" QUAD_D##N(c) CASC_C ".Lqc%=:\n\ts_nop 1\n\t" QUAD_D##N(s)  \
                         CASC_S ".Lqs%=:\n\ts_nop 0"                                             \
                         : QUAD_OPS##N                                                           \
                         : [ae] "v"(ae), [ao] "v"(ao), [zar] "v"(zar), [zai] "v"(zai),           \
                           [zbr] "v"(zbr), [zbi] "v"(zbi), [first] "s"(first), [last] "s"(last)); \
        }                                                                                        \
        /* the chunk's first row pair (rows 0, 1: every block) */                                \
        __device__ static __forceinline__ void seed(f32x4* E, f32x4* O, float ae, float ao,      \
                                                    float zar, float zai, float zbr, float zbi) { \
            asm volatile("s_nop 1\n\t" SEED_C "s_nop 1\n\t" SEED_S "s_nop 0"                   \
                         : QUAD_OUT##N                                                           \
                         : [ae] "v"(ae), [ao] "v"(ao), [zar] "v"(zar), [zai] "v"(zai),           \
                           [zbr] "v"(zbr), [zbi] "v"(zbi));                                      \
        }                                                                                        \
        /* wait states between the sums' last products and their first read */                 \
        __device__ static __forceinline__ void settle(f32x4* E, f32x4* O) {                      \
            asm volatile("s_nop 7" : QUAD_OPS##N);                                                \
        }                                                                                        \
    };
QUAD_ROW(1, true, QUAD_CP1, QUAD_SP1, QUAD_CP01, QUAD_SP01)
QUAD_ROW(2, true, QUAD_CP2, QUAD_SP2, QUAD_CP02, QUAD_SP02)
QUAD_ROW(3, true, QUAD_CP3, QUAD_SP3, QUAD_CP03, QUAD_SP03)
QUAD_ROW(4, true, QUAD_CP4, QUAD_SP4, QUAD_CP04, QUAD_SP04)
QUAD_ROW(1, false, QUAD_CU1, QUAD_SU1, QUAD_CU01, QUAD_SU01)
QUAD_ROW(2, false, QUAD_CU2, QUAD_SU2, QUAD_CU02, QUAD_SU02)
QUAD_ROW(3, false, QUAD_CU3, QUAD_SU3, QUAD_CU03, QUAD_SU03)
QUAD_ROW(4, false, QUAD_CU4, QUAD_SU4, QUAD_CU04, QUAD_SU04)
QUAD_ROW(5, false, QUAD_CU5, QUAD_SU5, QUAD_CU05, QUAD_SU05)
QUAD_ROW(6, false, QUAD_CU6, QUAD_SU6, QUAD_CU06, QUAD_SU06)
QUAD_ROW(7, false, QUAD_CU7, QUAD_SU7, QUAD_CU07, QUAD_SU07)
QUAD_ROW(8, false, QUAD_CU8, QUAD_SU8, QUAD_CU08, QUAD_SU08)
#undef QUAD_ROW
#undef QUAD_C
#undef QUAD_S
#undef QUAD_C0
#undef QUAD_S0
#undef QUAD_DB

// Record operands of a chunk of C orders.  Row m's part of the chunk is
// F = 2C floats: float4 (U1, U2) per order l = L0 + (m & 1) + 2k.  Its
// quads are slots s = 2b + sigma, b the block of the orders k = 2b, 2b + 1
// (l and l + 2 of the row's parity), sigma 0 for the cos records (U1 of
// both orders), 1 for the sin records (U2); slot s holds floats
// 4 (2b + i/2) + 2 sigma + (i & 1), i = 0..3, in lanes 4s + i.  Rows m
// and m + 1 load together, into one register (row m + 1 from lane 32) when
// F <= 32, else into NV registers each.
template <int L, int C>
struct QuadRegs {
    static constexpr int F = 2 * C;
    static constexpr bool kPacked = F <= 32;
    static constexpr int NV = kPacked ? 1 : (F + 63) / 64;
    static constexpr int NR = kPacked ? 1 : 2 * NV;   // registers per row pair
    // register and broadcast block of slot (row r of the pair, block b, sigma)
    static constexpr int reg(int r, int b, int sg) {
        return kPacked ? 0 : r * NV + (2 * b + sg) / 16;
    }
    static constexpr int abid(int r, int b, int sg) {
        return kPacked ? 8 * r + 2 * b + sg : (2 * b + sg) % 16;
    }
};

// The byte offsets (from row m's part of the chunk) that a lane loads into
// each register of a row pair; lanes past F load the part's first float,
// which no slot broadcasts.
template <int L, int C>
__device__ __forceinline__ void quad_offsets(int lane, int* off) {
    using G = QuadRegs<L, C>;
    auto perm = [](int u) {
        return 4 * (2 * (u >> 3) + ((u & 3) >> 1)) + 2 * ((u >> 2) & 1) + (u & 1);
    };
    if constexpr (G::kPacked) {
        const int r = lane >> 5, u = lane & 31;
        off[0] = 4 * (u < G::F ? r * 2 * L + perm(u) : 0);
    } else {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int v = 0; v < G::NV; ++v) {
                const int u = 64 * v + lane;
                off[r * G::NV + v] = 4 * (u < G::F ? r * 2 * L + perm(u) : 0);
            }
    }
}

// The record table of one sphere (L x L float4) as a buffer resource: the
// row offset of a load is a scalar operand, so a row's loads take no VALU
// address arithmetic (and a load past the table returns zero).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t quad_table(const float* Rf, int L) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(Rf), 0, L * L * 16, 0x00020000);
}

template <int NR>
__device__ __forceinline__ void quad_load(__amdgpu_buffer_rsrc_t rs, const int* off, int row_bytes,
                                          float* a) {
#pragma unroll
    for (int q = 0; q < NR; ++q)
        a[q] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off[q], row_bytes, 0));
}

// The blocks of a row pair in asm statements of up to eight blocks each
// (group k: blocks 8k .. of the chunk; its registers: row m's register k
// and row m + 1's, or the one packed register).  Block gb of the bucket
// (= L0/4 + b) runs from row 4 gb on and up to the block of order Nl, as
// in harm_planar_chunk.  The chunk's first row pair runs every block (the
// blocks past Nl there take the zero records the fold writes, which start
// their sums at zero; no later product or read touches them).
template <int L, int C, int L0, int k, bool kSeed>
__device__ __forceinline__ void quad_groups(const float* a, float zar, float zai, float zbr,
                                            float zbi, int m, int nl1, f32x4* De, f32x4* Do) {
    constexpr int Qv = (L0 + C <= L ? C : L - L0) / 4;
    if constexpr (8 * k < Qv) {
        using G = QuadRegs<L, C>;
        constexpr int n = Qv - 8 * k < 8 ? Qv - 8 * k : 8;
        const float ae = a[G::kPacked ? 0 : k], ao = a[G::kPacked ? 0 : G::NV + k];
        if constexpr (kSeed) {
            QuadRow<n, G::kPacked>::seed(De + 8 * k, Do + 8 * k, ae, ao, zar, zai, zbr, zbi);
        } else {
            const int first = (m >> 2) - (L0 / 4 + 8 * k);
            const int last = ((nl1 - 1) >> 2) - (L0 / 4 + 8 * k);
            QuadRow<n, G::kPacked>::run(De + 8 * k, Do + 8 * k, ae, ao, zar, zai, zbr, zbi,
                                        first, last);
        }
        quad_groups<L, C, L0, k + 1, kSeed>(a, zar, zai, zbr, zbi, m, nl1, De, Do);
    }
}

template <int L, int C, int L0, int k>
__device__ __forceinline__ void quad_settle(f32x4* De, f32x4* Do) {
    constexpr int Qv = (L0 + C <= L ? C : L - L0) / 4;
    if constexpr (8 * k < Qv) {
        constexpr int n = Qv - 8 * k < 8 ? Qv - 8 * k : 8;
        QuadRow<n, QuadRegs<L, C>::kPacked>::settle(De + 8 * k, Do + 8 * k);
        quad_settle<L, C, L0, k + 1>(De, Do);
    }
}

// Phases of the row pairs kept in LDS from a pixel's first chunk for its
// later ones, which would otherwise step them again: per lane, pair
// m = 2, 4, ... at slot m/2 - 1 as (cos, sin) of rows m and m + 1, slots
// kHarmThreads apart (a wave's 64 lanes read 1 KB in a row).  As many pairs
// as the chunk's first rows hold (C/2 - 1) and as leave the kernel's waves
// per SIMD in the 160 KB of LDS of a CU (a block of 256 lanes per wave per
// SIMD, 4 KB per pair, 1 KB kept for the rest).
constexpr int quad_phase_pairs(int C, int waves) {
    const int budget = waves > 0 ? (163840 / waves - 1024) / 4096 : 0;
    const int need = C / 2 - 1;
    return need < budget ? need : budget;
}

// harm_planar_chunk on the matrix cores: De[b] holds the sums of orders
// L0 + 4b and L0 + 4b + 2 (re, im, re, im), Do[b] those of L0 + 4b + 1 and
// L0 + 4b + 3.  The next row pair's records load while this one's
// products run.  The first chunk (L0 = 0) leaves the phases of its first
// P row pairs past rows 0, 1 in ph; the later ones read them there and
// step the phases on from the last (the same values as stepping them
// again).
template <int L, int C, int L0, int P>
__device__ __forceinline__ void quad_chunk(__amdgpu_buffer_rsrc_t rs, const int* off, int nl1,
                                           float z1r, float z1i, float4* ph, f32x4* De,
                                           f32x4* Do) {
    using G = QuadRegs<L, C>;
    static_assert(!G::kPacked || C <= 16, "packed record registers hold four blocks");
    static_assert(2 + 2 * P <= C, "the first chunk holds the kept pairs");
    const int mend = min(nl1, L0 + C);
    // bytes of row m's part of the chunk
    auto row = [](int m) { return 4 * (m * (2 * L) + 2 * L0); };
    float a0[G::NR], a1[G::NR];
    quad_load<G::NR>(rs, off, row(0), a0);
    // (a second register set for alternate row pairs, instead of the copies
    // below, doubled the asm and raised the register count)
    // rows 0 and 1 (every block runs there, and starts its sums): phases
    // 1 and e^{i phi}
    quad_load<G::NR>(rs, off, row(2), a1);
    quad_groups<L, C, L0, 0, true>(a0, 1.f, 0.f, z1r, z1i, 0, nl1, De, Do);
    float zr, zi;
    // rows 2, 3 run even where Nl < 2: their records are zero there (rows
    // past Nl hold zeros), so they leave the sums as they are, and with one
    // path out of each loop the sums keep their registers (an exit that
    // could skip a loop cost a shuffle of the accumulators after it)
    int m = 2;
    if constexpr (L0 > 0 && P > 0) {
        // (a later chunk runs past row C - 1 >= 2P + 1: both loops run; the
        // first unrolled, so that each read has its own registers and a
        // constant offset)
        float4 p = ph[0];
#pragma unroll
        do {
#pragma unroll
            for (int q = 0; q < G::NR; ++q) a0[q] = a1[q];
            quad_load<G::NR>(rs, off, row(m + 2), a1);
            const float4 z = p;
            p = ph[min(m >> 1, P - 1) * kHarmThreads];
            quad_groups<L, C, L0, 0, false>(a0, z.x, z.y, z.z, z.w, m, nl1, De, Do);
            zr = fmaf(z.z, z1r, -(z.w * z1i));
            zi = fmaf(z.z, z1i, z.w * z1r);
            m += 2;
        } while (m < 2 + 2 * P);
    } else {
        zr = fmaf(z1r, z1r, -(z1i * z1i));
        zi = fmaf(z1r, z1i, z1i * z1r);
    }
#pragma clang loop unroll(disable)
    do {
#pragma unroll
        for (int q = 0; q < G::NR; ++q) a0[q] = a1[q];
        quad_load<G::NR>(rs, off, row(m + 2), a1);
        const float zar = zr, zai = zi;
        const float zbr = fmaf(zr, z1r, -(zi * z1i)), zbi = fmaf(zr, z1i, zi * z1r);
        if constexpr (L0 == 0 && P > 0) {
            if (m < 2 + 2 * P) ph[((m >> 1) - 1) * kHarmThreads] = make_float4(zar, zai, zbr, zbi);
        }
        quad_groups<L, C, L0, 0, false>(a0, zar, zai, zbr, zbi, m, nl1, De, Do);
        zr = fmaf(zbr, z1r, -(zbi * z1i));
        zi = fmaf(zbr, z1i, zbi * z1r);
        m += 2;
    } while (m < mend);
    quad_settle<L, C, L0, 0>(De, Do);
}

// The radial factors of the chunk's orders, block by block (orders
// L0 + 4b .. L0 + 4b + 3 from De[b] and Do[b]), the recurrences carried
// as in harm_radial_range over the whole chunk.  Outside the sphere (the
// lanes of pass 0 whose pixel is outside, tested once per chunk): the
// steps of harm_radial_out, the threshold rules on the blocks some lane
// is past kd at; a block with no such lane steps all four orders.
template <int L, int C, int L0, int b>
__device__ __forceinline__ void quad_radial_out(const f32x4* De, const f32x4* Do,
                                                RadialState& st, int nl1, float& acc_r,
                                                float& acc_i) {
    constexpr int Qv = (L0 + C <= L ? C : L - L0) / 4;
    if constexpr (b < Qv) {
        constexpr int l0 = L0 + 4 * b;
        if (l0 >= nl1) return;
        const float ar[4] = {De[b][0], Do[b][0], De[b][2], Do[b][2]};
        const float ai[4] = {De[b][1], Do[b][1], De[b][3], Do[b][3]};
        if (wave_any(l0 + 3 > st.kfl))
            harm_radial_out<l0, 4, true>(st, nl1, ar, ai, acc_r, acc_i);
        else
            harm_radial_out<l0, 4, false, true>(st, nl1, ar, ai, acc_r, acc_i);
        quad_radial_out<L, C, L0, b + 1>(De, Do, st, nl1, acc_r, acc_i);
    }
}

// The same for the lanes of pass 1 (inside the sphere, or at its centre).
template <int L, int C, int L0, int b>
__device__ __forceinline__ void quad_radial_in(const f32x4* De, const f32x4* Do,
                                               RadialState& st, const SphereDev& S,
                                               float& acc_r, float& acc_i) {
    constexpr int Qv = (L0 + C <= L ? C : L - L0) / 4;
    if constexpr (b < Qv) {
        if (L0 + 4 * b > S.Nl) return;
        const float ar[4] = {De[b][0], Do[b][0], De[b][2], Do[b][2]};
        const float ai[4] = {De[b][1], Do[b][1], De[b][3], Do[b][3]};
        harm_radial_range<L0 + 4 * b, 4>(1, st, S, ar, ai, acc_r, acc_i);
        quad_radial_in<L, C, L0, b + 1>(De, Do, st, S, acc_r, acc_i);
    }
}

// harm_planar_chunks on the matrix cores: chunk c and the ones after it.
template <int L, int C, int P, int c>
__device__ __forceinline__ void quad_chunks(__amdgpu_buffer_rsrc_t rs, const int* off,
                                            const SphereDev& S, float z1r, float z1i, int pass,
                                            RadialState& st, float4* ph, float& acc_r,
                                            float& acc_i) {
    if constexpr (c * C < L) {
        const int nl1 = S.Nl + 1;
        if (c * C >= nl1) return;
        constexpr int Q = C / 4;
        f32x4 De[Q], Do[Q];   // set by the chunk's first row pair
        quad_chunk<L, C, c * C, P>(rs, off, nl1, z1r, z1i, ph, De, Do);
        if (pass == 0) {
            if (st.mode == 0) quad_radial_out<L, C, c * C, 0>(De, Do, st, nl1, acc_r, acc_i);
        } else {
            quad_radial_in<L, C, c * C, 0>(De, Do, st, S, acc_r, acc_i);
        }
        quad_chunks<L, C, P, c + 1>(rs, off, S, z1r, z1i, pass, st, ph, acc_r, acc_i);
    }
}

// Per-pixel harmonic sum, persistent over (instance, pixel block) items.
// Row m of the order sums runs the scaled recurrence
//   q_mm = Pn_mm,  q_lm = 2x q_{l-1,m} - c_lm q_{l-2,m}
// (Pn_lm = K_lm q_lm, K_lm = prod_{j=m+1..l} a_jm / 2, folded into the
// records) in blocks of four orders: one scalar load of four records and
// one of their constants per block, then straight-line FP32 code (blocks
// wholly below the diagonal or past Nl are skipped).  cst[2 (m kCS + l)]
// holds c_lm (l >= m + 2) and the diagonal factor sqrt((2m+1)/(2m)) at
// l = m (where the recurrence does not read c); cst[2 (m kCS + l) + 1] is 1
// at l = m, else 0, and injects the diagonal value (q is 0 below it).
// rec holds the records times B_l, rec_in the same times A_l: a wave with
// pixels outside the sphere sums the first (pass 0), one with pixels inside
// the second (pass 1); a wave with both does both passes, each lane
// keeping its own.  kMfma (planar buckets, orders in chunks of kChunk):
// the order sums on the matrix cores (quad_chunks), which give the same
// bits as the VALU chunks.
template <int L, bool kPlanar, int kWaves, int kChunk, bool kMfma = false>
__global__ __launch_bounds__(kHarmThreads, kWaves) void series_harm_kernel(
    const SeriesInst* __restrict__ insts, const SeriesSph* __restrict__ sph,
    const SeriesGroup* __restrict__ grps, const SeriesItem* __restrict__ items, int n_items,
    const float4* __restrict__ rec, const float4* __restrict__ rec_in,
    const float* __restrict__ cst, const int* __restrict__ perm, float2* __restrict__ U,
    int* __restrict__ counter) {
    static_assert(!kMfma || (kPlanar && kChunk > 0), "matrix-core sums: planar chunks only");
    __shared__ int s_item;
    int qoff[QuadRegs<L, (kChunk > 0 ? kChunk : 4)>::NR];
    if constexpr (kMfma) quad_offsets<L, kChunk>((int)(threadIdx.x & 63), qoff);
    // the row pairs' phases that a pixel's first chunk keeps for its later ones
    constexpr int kPhasePairs = kMfma ? quad_phase_pairs(kChunk, kWaves) : 0;
    __shared__ float4 s_phase[(kPhasePairs > 0 ? kPhasePairs : 1) * kHarmThreads];
    float4* const ph = s_phase + threadIdx.x;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = __builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (it >= n_items) break;
        const SeriesItem item = items[it];
        const SeriesInst& I = insts[item.inst];
        const int npix = I.uR * I.vR;
        const int j0 = item.pix0 + (int)threadIdx.x;
        // lanes past the instance's pixels repeat its last one and store
        // nothing: every lane stays active (the matrix-core sums broadcast
        // record operands from lanes of the wave)
        const bool live = j0 < npix;
        const int j = live ? j0 : npix - 1;
        // pixels in the instance's order of uniform waves (inside each
        // sphere, then outside every sphere), so that a wave takes one
        // radial path
        const int pix = I.perm_off >= 0 ? perm[I.perm_off + j] : j;
        const int iv = pix / I.uR, iu = pix - iv * I.uR;
        const d3 p = quad_at(I.q, (double)iu / (double)I.uR, (double)iv / (double)I.vR);
        const SeriesGroup G = grps[item.grp];
        const double inv_kmag = 1.0 / I.kmag;
        float acc_r = 0.f, acc_i = 0.f;
        for (int s = G.s_begin; s < G.s_begin + G.ns; ++s) {
            const SeriesSph& SS = sph[s];
            const SphereDev S = SS.s;
            const int nl1 = S.Nl + 1;
            d3 r = sub(p, mk(S.p));
            double d, inv_d;
            dist_rsq(dot(r, r), d, inv_d);
            // direction; the centre (knd below float epsilon) keeps only l = 0
            const bool centre = d <= S.a && S.n[0] * (I.kmag * d) < (double)kEpsFloat;
            r = centre ? d3{0.0, 0.0, 0.0} : scale(r, inv_d);
            const bool inside = d <= S.a;
            const int pass0 = __any(!inside) ? 0 : 1, pass1 = __any(inside) ? 1 : 0;
            if constexpr (kMfma) {
                const RadialTrig trig = harm_trig(S, I.kmag, d);
                RadialState st = harm_radial_init(S, I.kmag, d, inv_d, inv_kmag, trig);
                if (pass0 == 0)
                    quad_chunks<L, kChunk, kPhasePairs, 0>(
                        quad_table(reinterpret_cast<const float*>(rec + SS.toff), L), qoff, S,
                        (float)r.x, (float)r.y, 0, st, ph, acc_r, acc_i);
                if (pass1 == 1)
                    quad_chunks<L, kChunk, kPhasePairs, 0>(
                        quad_table(reinterpret_cast<const float*>(rec_in + SS.toff), L), qoff, S,
                        (float)r.x, (float)r.y, 1, st, ph, acc_r, acc_i);
                continue;
            } else if constexpr (kChunk > 0) {
                // planar buckets only; orders in chunks of kChunk
                const RadialTrig trig = harm_trig(S, I.kmag, d);
                RadialState st = harm_radial_init(S, I.kmag, d, inv_d, inv_kmag, trig);
                if (pass0 == 0)
                    harm_planar_chunks<L, kChunk, 0>(rec + SS.toff, S, (float)r.x, (float)r.y, 0,
                                                     st, acc_r, acc_i);
                if (pass1 == 1)
                    harm_planar_chunks<L, kChunk, 0>(rec_in + SS.toff, S, (float)r.x, (float)r.y,
                                                     1, st, acc_r, acc_i);
                continue;
            }
            if (kPlanar || SS.planar) {
                // r.z = 0, so sin(theta) = 1 and e^{i phi} = r.x + i r.y
                const RadialTrig trig = harm_trig(S, I.kmag, d);
                auto sums = [&](const float4* R, int pass) {
                    float ar[L], ai[L];
#pragma unroll
                    for (int l = 0; l < L; ++l) ar[l] = ai[l] = 0.f;
                    harm_planar_sums<L>(R + SS.toff, nl1, (float)r.x, (float)r.y, ar, ai);
                    harm_radial<L>(pass, S, I.kmag, d, inv_d, inv_kmag, trig, ar, ai, acc_r,
                                   acc_i);
                };
                // (two calls rather than a loop over the passes: a loop here
                // keeps the order sums in scratch)
                if (pass0 == 0) sums(rec, 0);
                if (pass1 == 1) sums(rec_in, 1);
                continue;
            }
            const float x = (float)r.z;
            const float st = sqrtf(fmaxf(0.f, fmaf(-x, x, 1.f)));
            float z1r = 1.f, z1i = 0.f;
            if (st > 1e-30f) {
                z1r = (float)r.x / st;
                z1i = (float)r.y / st;
            }
            const float x2 = 2.f * x;
            const RadialTrig trig = harm_trig(S, I.kmag, d);
            auto sums = [&](const float4* Rp, int pass) {
                float ar[L], ai[L];
#pragma unroll
                for (int l = 0; l < L; ++l) ar[l] = ai[l] = 0.f;
                // row stride L: the instance's bucket
                const float4* R = Rp + SS.toff;
                const int be = (nl1 + 3) >> 2;
                float zr = 1.f, zi = 0.f;
                float pmm = 0.28209479177387814f;   // 1/sqrt(4 pi)
                // rows in pairs (m, m + 1), m even: both start in block m / 4, so
                // one wait covers two rows' loads and the two recurrences
                // interleave (a row m + 1 = Nl + 1 has zero records)
#pragma clang loop unroll(disable)
                for (int m = 0; m < nl1; m += 2) {
                    const float* Ca = cst + 2 * m * kCS;
                    const float* Cb = Ca + 2 * kCS;
                    const float4* Ra = R + m * L;
                    const float4* Rb = Ra + L;
                    const float pa = pmm, pb = Cb[2 * (m + 1)] * st * pmm;
                    const float zar = zr, zai = zi;
                    const float zbr = fmaf(zr, z1r, -(zi * z1i)), zbi = fmaf(zr, z1i, zi * z1r);
                    const int bf = m >> 2;
                    float qa0 = 0.f, qa1 = 0.f, qb0 = 0.f, qb1 = 0.f;
#pragma clang loop unroll(full)
                    for (int b = 0; b < L / 4; ++b) {
                        if (b < bf || b >= be) continue;
                        const float4 ca0 = *reinterpret_cast<const float4*>(Ca + 8 * b);
                        const float4 ca1 = *reinterpret_cast<const float4*>(Ca + 8 * b + 4);
                        const float4 cb0 = *reinterpret_cast<const float4*>(Cb + 8 * b);
                        const float4 cb1 = *reinterpret_cast<const float4*>(Cb + 8 * b + 4);
                        const float cca[4] = {ca0.x, ca0.z, ca1.x, ca1.z};
                        const float eea[4] = {ca0.y, ca0.w, ca1.y, ca1.w};
                        const float ccb[4] = {cb0.x, cb0.z, cb1.x, cb1.z};
                        const float eeb[4] = {cb0.y, cb0.w, cb1.y, cb1.w};
                        float4 ra[4], rb[4];
#pragma clang loop unroll(full)
                        for (int j = 0; j < 4; ++j) {
                            ra[j] = Ra[4 * b + j];
                            rb[j] = Rb[4 * b + j];
                        }
#pragma clang loop unroll(full)
                        for (int j = 0; j < 4; ++j) {
                            // the diagonal value enters at l = m (ee = 1 there, else
                            // 0); below it q stays 0
                            const int l = 4 * b + j;
                            const float qa = fmaf(x2, qa1, fmaf(-cca[j], qa0, eea[j] * pa));
                            const float qb = fmaf(x2, qb1, fmaf(-ccb[j], qb0, eeb[j] * pb));
                            qa0 = qa1;
                            qa1 = qa;
                            qb0 = qb1;
                            qb1 = qb;
                            const float ta1 = qa * zar, ta2 = qa * zai;
                            const float tb1 = qb * zbr, tb2 = qb * zbi;
                            ar[l] = fmaf(ta2, ra[j].z, fmaf(ta1, ra[j].x, ar[l]));
                            ai[l] = fmaf(ta2, ra[j].w, fmaf(ta1, ra[j].y, ai[l]));
                            ar[l] = fmaf(tb2, rb[j].z, fmaf(tb1, rb[j].x, ar[l]));
                            ai[l] = fmaf(tb2, rb[j].w, fmaf(tb1, rb[j].y, ai[l]));
                        }
                    }
                    // the next pair's diagonal value and phase
                    pmm = Cb[2 * (kCS + m + 2)] * st * pb;
                    zr = fmaf(zbr, z1r, -(zbi * z1i));
                    zi = fmaf(zbr, z1i, zbi * z1r);
                }
                harm_radial<L>(pass, S, I.kmag, d, inv_d, inv_kmag, trig, ar, ai, acc_r, acc_i);
            };
            if (pass0 == 0) sums(rec, 0);
            if (pass1 == 1) sums(rec_in, 1);
        }
        if (live) {
            float2* o = U + I.u_off + pix;
            float2 v = make_float2(0.f, 0.f);
            if (G.accumulate) v = *o;
            *o = make_float2(v.x + acc_r, v.y + acc_i);
        }
    }
}


}  // namespace harm
}  // namespace bimsim
