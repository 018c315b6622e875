// Detector: microscopeStruct::SimulateImaging (reference microscope.cu:
// 608-662, 722-791; fieldslice.cpp:25-94; fieldslice.cu:71-153,319-362).
//
// 2D FFT as two passes of batched 1D Stockham FFTs staged in LDS (one
// workgroup owns B lines of length N; radix 4/2/3/5 stages ping-pong between
// two LDS buffers).  The objective band-pass is applied in the store phase of
// the forward column pass, the 1/(uR vR) scale in the store phase of the
// inverse row pass.  A final kernel crops, forms |U|^2 averaged over ss x ss
// into D and Di (accumulating, like resample_intensity) and -log10(D/Di).
// FP64 instantiations run in FP64 end to end (-ffp-contract=off for the
// whole translation unit); FP32 ones in float with FP64-accurate twiddles.
#include <mutex>
#include <type_traits>

#include "batch.h"
#include "geometry.h"
#include "internal.h"

namespace bimsim {
namespace {

constexpr int kThreads = 256;
constexpr int kMaxElems = 8192;   // B*N per workgroup
constexpr int kMaxStages = 16;

struct FftPlan {
    int n;
    int nstages;
    int radix[kMaxStages];
};

int make_plan(int n, FftPlan* p) {
    p->n = n;
    p->nstages = 0;
    int m = n;
    while (m > 1) {
        int r;
        if (m % 4 == 0) r = 4;
        else if (m % 2 == 0) r = 2;
        else if (m % 3 == 0) r = 3;
        else if (m % 5 == 0) r = 5;
        else return -1;
        if (p->nstages >= kMaxStages) return -1;
        p->radix[p->nstages++] = r;
        m /= r;
    }
    return 0;
}

template <class R> struct C2;
template <> struct C2<float> { using T = float2; };
template <> struct C2<double> { using T = double2; };

template <class C>
__device__ inline C cmk(decltype(C::x) a, decltype(C::x) b) {
    C c;
    c.x = a;
    c.y = b;
    return c;
}

// Post-operation applied when a pass stores its results.
struct PostOp {
    int kind;        // 0 none, 1 objective band-pass (freq. index = (elem, line)), 2 scale
    int uR, vR;
    int quirks;
    double du, dv, fmin, fmax;   // fmin = NAin/lambda, fmax = NAout/lambda
    double scale;
};

// bandpass kernel of microscope.cu:495-526 (frequencies in double)
__device__ inline bool passband(const PostOp& op, int iu, int iv) {
    double u, v;
    if (iu <= op.uR / 2) u = (double)iu * op.du;
    else u = -(double)(op.quirks ? op.uR - 1 - iu : op.uR - iu) * op.du;
    if (iv <= op.vR / 2) v = (double)iv * op.dv;
    else v = -(double)(op.quirks ? op.vR - 1 - iv : op.vR - iv) * op.dv;
    const double fmag = sqrt(u * u + v * v);
    return !(fmag < op.fmin || fmag > op.fmax);
}

// One instance of a batched detector pass (band-limited path, ss = 1, the
// images accumulate): its fields at uoff, pass-band columns at cols_off in
// the batch's column table, compact spectrum (2 Nc N values) at g_off,
// images at image_off; crop rows and columns [off, off + Nd).
struct DetInst {
    long long uoff, g_off, image_off;
    int cols_off, Nc, off, Nd;
    PostOp bp;
    SumFuse fz;
};

template <class R>
__device__ inline void butterfly(typename C2<R>::T* x, int r, R sgn) {
    using C = typename C2<R>::T;
    if (r == 2) {
        const C a = x[0], b = x[1];
        x[0] = cmk<C>(a.x + b.x, a.y + b.y);
        x[1] = cmk<C>(a.x - b.x, a.y - b.y);
    } else if (r == 4) {
        const C a = cmk<C>(x[0].x + x[2].x, x[0].y + x[2].y);
        const C b = cmk<C>(x[0].x - x[2].x, x[0].y - x[2].y);
        const C c = cmk<C>(x[1].x + x[3].x, x[1].y + x[3].y);
        const C t = cmk<C>(x[1].x - x[3].x, x[1].y - x[3].y);
        const C d = cmk<C>(-sgn * t.y, sgn * t.x);   // (x1 - x3) * (sgn i)
        x[0] = cmk<C>(a.x + c.x, a.y + c.y);
        x[2] = cmk<C>(a.x - c.x, a.y - c.y);
        x[1] = cmk<C>(b.x + d.x, b.y + d.y);
        x[3] = cmk<C>(b.x - d.x, b.y - d.y);
    } else if (r == 3) {
        const R h = (R)0.86602540378443864676;   // sqrt(3)/2
        const C t1 = cmk<C>(x[1].x + x[2].x, x[1].y + x[2].y);
        const C t = cmk<C>(x[1].x - x[2].x, x[1].y - x[2].y);
        const C t2 = cmk<C>(-sgn * h * t.y, sgn * h * t.x);
        const C m = cmk<C>(x[0].x - (R)0.5 * t1.x, x[0].y - (R)0.5 * t1.y);
        x[0] = cmk<C>(x[0].x + t1.x, x[0].y + t1.y);
        x[1] = cmk<C>(m.x + t2.x, m.y + t2.y);
        x[2] = cmk<C>(m.x - t2.x, m.y - t2.y);
    } else {   // r == 5, direct DFT with exact constants
        const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
        const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
        C y[5];
        y[0] = cmk<C>(x[0].x + x[1].x + x[2].x + x[3].x + x[4].x,
                      x[0].y + x[1].y + x[2].y + x[3].y + x[4].y);
        const C a1 = cmk<C>(x[1].x + x[4].x, x[1].y + x[4].y);
        const C b1 = cmk<C>(x[1].x - x[4].x, x[1].y - x[4].y);
        const C a2 = cmk<C>(x[2].x + x[3].x, x[2].y + x[3].y);
        const C b2 = cmk<C>(x[2].x - x[3].x, x[2].y - x[3].y);
        const C m1 = cmk<C>(x[0].x + c1 * a1.x + c2 * a2.x, x[0].y + c1 * a1.y + c2 * a2.y);
        const C m2 = cmk<C>(x[0].x + c2 * a1.x + c1 * a2.x, x[0].y + c2 * a1.y + c1 * a2.y);
        // sgn i (s1 b1 + s2 b2) and sgn i (s2 b1 - s1 b2)
        const C n1 = cmk<C>(-sgn * (s1 * b1.y + s2 * b2.y), sgn * (s1 * b1.x + s2 * b2.x));
        const C n2 = cmk<C>(-sgn * (s2 * b1.y - s1 * b2.y), sgn * (s2 * b1.x - s1 * b2.x));
        y[1] = cmk<C>(m1.x + n1.x, m1.y + n1.y);
        y[4] = cmk<C>(m1.x - n1.x, m1.y - n1.y);
        y[2] = cmk<C>(m2.x + n2.x, m2.y + n2.y);
        y[3] = cmk<C>(m2.x - n2.x, m2.y - n2.y);
        for (int i = 0; i < 5; ++i) x[i] = y[i];
    }
}

// Batched 1D FFTs: line `g` of the array has its element e at
// base[g*line_stride + e*elem_stride].  Each workgroup owns B lines.
template <class R>
__global__ __launch_bounds__(kThreads) void fft_lines_kernel(
    typename C2<R>::T* __restrict__ data, int nlines, int B, int P, long long line_stride,
    long long elem_stride, FftPlan plan, const typename C2<R>::T* __restrict__ tw, int inverse,
    PostOp op) {
    using C = typename C2<R>::T;
    extern __shared__ unsigned char smem_raw[];
    const int N = plan.n;   // P >= N: line stride in LDS (N + 1 breaks bank conflicts)
    C* buf0 = reinterpret_cast<C*>(smem_raw);
    C* buf1 = buf0 + (size_t)B * P;
    const int g0 = blockIdx.x * B;
    const int nb = min(B, nlines - g0);
    const int tot = nb * N;
    // load: consecutive threads walk the unit-stride dimension
    const bool lines_fast = line_stride == 1;
    for (int t = threadIdx.x; t < tot; t += kThreads) {
        int g, e;
        if (lines_fast) { g = t % nb; e = t / nb; }
        else { e = t % N; g = t / N; }
        buf0[g * P + e] = data[(long long)(g0 + g) * line_stride + (long long)e * elem_stride];
    }
    __syncthreads();
    const R sgn = inverse ? (R)1 : (R)-1;
    C* src = buf0;
    C* dst = buf1;
    int Ns = 1;
    for (int st = 0; st < plan.nstages; ++st) {
        const int r = plan.radix[st];
        const int nq = N / r;
        const int tstep = N / (Ns * r);
        for (int t = threadIdx.x; t < nb * nq; t += kThreads) {
            const int g = t / nq, j = t - g * nq;
            const int k = j % Ns;
            C x[5];
            for (int q = 0; q < r; ++q) {
                C v = src[g * P + j + q * nq];
                if (q > 0 && k > 0) {
                    C w = tw[k * q * tstep];   // k*q*tstep < N
                    if (inverse) w.y = -w.y;
                    v = cmk<C>(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
                }
                x[q] = v;
            }
            butterfly<R>(x, r, sgn);
            const int o = (j / Ns) * Ns * r + k;
            for (int q = 0; q < r; ++q) dst[g * P + o + q * Ns] = x[q];
        }
        __syncthreads();
        C* tmp = src;
        src = dst;
        dst = tmp;
        Ns *= r;
    }
    for (int t = threadIdx.x; t < tot; t += kThreads) {
        int g, e;
        if (lines_fast) { g = t % nb; e = t / nb; }
        else { e = t % N; g = t / N; }
        C v = src[g * P + e];
        if (op.kind == 1) {
            // frequency indices: elements run along v (column pass), lines along u
            const int iu = g0 + g, iv = e;
            if (!passband(op, iu, iv)) v = cmk<C>((R)0, (R)0);
        } else if (op.kind == 2) {
            v = cmk<C>(v.x * (R)op.scale, v.y * (R)op.scale);
        }
        data[(long long)(g0 + g) * line_stride + (long long)e * elem_stride] = v;
    }
}

// crop + resample_intensity (+=) + calc_absorbance, per detector pixel
template <class R>
__global__ void detect_kernel(const typename C2<R>::T* __restrict__ U,
                              const typename C2<R>::T* __restrict__ Uf, int uR, int off, int Nd,
                              int ss, R* __restrict__ D, R* __restrict__ Di, R* __restrict__ Abs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nd * Nd) return;
    const int iv = i / Nd, iu = i - iv * Nd;
    R I = 0, Ii = 0;
    for (int ix = 0; ix < ss; ++ix)
        for (int iy = 0; iy < ss; ++iy) {
            const long long fi = (long long)(off + iv * ss + iy) * uR + (off + iu * ss + ix);
            const auto a = U[fi], b = Uf[fi];
            const R am = sqrt(a.x * a.x + a.y * a.y), bm = sqrt(b.x * b.x + b.y * b.y);
            I += am * am;
            Ii += bm * bm;
        }
    const R d = D[i] + I / (R)(ss * ss);
    const R di = Di[i] + Ii / (R)(ss * ss);
    D[i] = d;
    Di[i] = di;
    if (Abs) Abs[i] = -log10(d / di);
}

// ------------------------------------------------ fixed-length line FFTs
// Stage plan of a length-N Stockham FFT known at compile time: radix 4 while
// possible, then 2, 3, 5 (the same order as make_plan).
constexpr int first_radix(int m) {
    return m % 4 == 0 ? 4 : (m % 2 == 0 ? 2 : (m % 3 == 0 ? 3 : (m % 5 == 0 ? 5 : 0)));
}
template <int N>
struct Shape {
    static constexpr int nst() {
        int m = N, s = 0;
        while (m > 1) {
            m /= first_radix(m);
            ++s;
        }
        return s;
    }
    static constexpr int rad(int st) {
        int m = N;
        for (int i = 0; i < st; ++i) m /= first_radix(m);
        return first_radix(m);
    }
    static constexpr int span(int st) {   // product of the radices before stage st
        int p = 1, m = N;
        for (int i = 0; i < st; ++i) {
            const int r = first_radix(m);
            p *= r;
            m /= r;
        }
        return p;
    }
};

template <class R, int r>
__device__ __forceinline__ void bfly_fixed(typename C2<R>::T* x, R sgn) {
    butterfly<R>(x, r, sgn);
}

template <class R, int N, int ST>
__device__ __forceinline__ void fixed_stage(const typename C2<R>::T* __restrict__ src,
                                            typename C2<R>::T* __restrict__ dst, int nb,
                                            const typename C2<R>::T* __restrict__ tw, R sgn,
                                            bool inverse) {
    using C = typename C2<R>::T;
    constexpr int r = Shape<N>::rad(ST);
    constexpr int Ns = Shape<N>::span(ST);
    constexpr int nq = N / r;
    constexpr int tstep = N / (Ns * r);
    for (int t = threadIdx.x; t < nb * nq; t += kThreads) {
        const int g = t / nq, j = t - g * nq;
        const int k = j % Ns;
        C x[r];
#pragma unroll
        for (int q = 0; q < r; ++q) {
            C v = src[g * N + j + q * nq];
            if (Ns > 1 && q > 0) {
                C w = tw[k * q * tstep];
                if (inverse) w.y = -w.y;
                v = cmk<C>(v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x);
            }
            x[q] = v;
        }
        bfly_fixed<R, r>(x, sgn);
        const int o = (j / Ns) * Ns * r + k;
#pragma unroll
        for (int q = 0; q < r; ++q) dst[g * N + o + q * Ns] = x[q];
    }
}

// All stages; returns the buffer holding the transform (natural order).
template <class R, int N, int ST = 0>
__device__ __forceinline__ typename C2<R>::T* fixed_fft(typename C2<R>::T* a,
                                                       typename C2<R>::T* b, int nb,
                                                       const typename C2<R>::T* tw, bool inverse) {
    if constexpr (ST == Shape<N>::nst()) {
        return a;
    } else {
        fixed_stage<R, N, ST>(a, b, nb, tw, inverse ? (R)1 : (R)-1, inverse);
        __syncthreads();
        return fixed_fft<R, N, ST + 1>(b, a, nb, tw, inverse);
    }
}

// Whether pixel pix of a batch instance lies inside one of its spheres, by
// the rule of the batch's sum kernel (d < a is inside, nfSumUf.cu:31-48).
// Only the spheres of `near` (bit s: sphere s) are tested when fz.ns <= 64;
// the others hold no pixel of the caller's rows.
__device__ __forceinline__ bool fused_inside(const SumFuse& fz, int pix,
                                             unsigned long long near = ~0ull) {
    const int iv = pix / fz.uR, iu = pix - iv * fz.uR;
    const d3 p = quad_at(fz.q, (double)iu / (double)fz.uR, (double)iv / (double)fz.vR);
    const bool masked = fz.ns <= 64;
    unsigned long long left = near;
    for (int s = 0; s < fz.ns; ++s) {
        if (masked) {
            if (left == 0) break;
            s = __builtin_ctzll(left);
            left &= left - 1;
        }
        const double4 S = fz.sph[s];
        // |p - c| < a from the squared distance, the square root taken only
        // within a relative 1e-12 of the boundary, where the rounded distance
        // could decide otherwise than its square
        const d3 r = sub(p, d3{S.x, S.y, S.z});
        const double d2 = dot(r, r), a2 = S.w * S.w;
        if (d2 < a2 * (1.0 - 1e-12)) return true;
        if (d2 <= a2 * (1.0 + 1e-12) && sqrt(d2) < S.w) return true;
    }
    return false;
}

// A batch instance's Uf + Us at pixel pix from U = Us (a) and Uf (b): Uf
// without spheres, U inside a sphere, else U + Uf; inside from fz's bits
// where it has them.
template <class C>
__device__ __forceinline__ C fused_sum(const SumFuse& fz, int pix, C a, C b,
                                       unsigned long long near = ~0ull) {
    if (fz.ns == 0) return b;
    const bool in = fz.inside ? ((fz.inside[pix >> 5] >> (pix & 31)) & 1u) != 0
                              : fused_inside(fz, pix, near);
    if (in) return a;
    a.x = a.x + b.x;
    a.y = a.y + b.y;
    return a;
}

// fz's inside bits: one word of 32 pixels per lane.
__global__ __launch_bounds__(kThreads) void sum_mask_kernel(SumFuse fz, unsigned* __restrict__ mask) {
    const int npix = fz.uR * fz.vR;
    const int w = blockIdx.x * kThreads + threadIdx.x;
    if (w >= (npix + 31) / 32) return;
    unsigned bits = 0u;
    for (int i = 0; i < 32; ++i) {
        const int pix = 32 * w + i;
        if (pix < npix && fz.ns > 0 && fused_inside(fz, pix)) bits |= 1u << i;
    }
    mask[w] = bits;
}

// A thread's share of tw[0, N), loaded into registers at a kernel's start
// (beside its other first loads) and stored to a copy in LDS later.
template <class C, int N>
struct TwShare {
    static constexpr int kPer = (N + kThreads - 1) / kThreads;
    C v[kPer];
    __device__ __forceinline__ void load(const C* __restrict__ tw) {
#pragma unroll
        for (int k = 0; k < kPer; ++k) v[k] = tw[min((int)threadIdx.x + k * kThreads, N - 1)];
    }
    __device__ __forceinline__ void store(C* T) const {
#pragma unroll
        for (int k = 0; k < kPer; ++k)
            if ((int)threadIdx.x + k * kThreads < N) T[threadIdx.x + k * kThreads] = v[k];
    }
};

// Forward FFT of B rows (length N) of U (f = 0) or Uf (f = 1); the
// pass-band columns cols[0..Nc) go to the compact spectrum
// G[(field*Nc + c)*nrows + row].  kSum: U is the scattered field alone and
// the rows of U load as fused_sum(fz, ...).  kBoth: the rows of both
// fields in one transform of 2 nb rows (U's, then Uf's; f unused), each Uf
// element loaded once for both.  kTwLds: the twiddles read from a copy in
// LDS (N more elements past the transform's buffers), not through the
// vector caches at every stage.
template <class R, int N, bool kSum, bool kBoth = false, bool kTwLds = false>
__device__ __forceinline__ void rows_fwd_body(
    const typename C2<R>::T* __restrict__ U, const typename C2<R>::T* __restrict__ Uf, int nrows,
    int B, const int* __restrict__ cols, int Nc, typename C2<R>::T* __restrict__ G,
    const typename C2<R>::T* __restrict__ tw, const SumFuse& fz, int blk, int f) {
    using C = typename C2<R>::T;
    constexpr int kF = kBoth ? 2 : 1;   // fields in the transform
    extern __shared__ unsigned char smem_raw[];
    C* a = reinterpret_cast<C*>(smem_raw);
    C* b = a + (size_t)kF * B * N;
    if (kBoth) f = 0;
    const C* F = f ? Uf : U;
    const int row0 = blk * B;
    const int nb = min(B, nrows - row0);
    const C* twp = tw;
    TwShare<C, N> tws;
    if constexpr (kTwLds) {
        tws.load(tw);
        twp = b + (size_t)kF * B * N;
    }
    if (kBoth && !(kSum && fz.ns > 0 && !fz.inside)) {
        // fused_sum with the inside bits (or no spheres), written out so that
        // the loop body has no branch: four elements of each field (and
        // their mask words) a step, every load issued before the first
        // value is used
        const int n = nb * N;
        const bool none = !kSum || fz.ns == 0;
        for (int t0 = threadIdx.x; t0 < n; t0 += 4 * kThreads) {
            C u[4], w[4];
            unsigned mk[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int t = min(t0 + k * kThreads, n - 1);
                const size_t i = (size_t)row0 * N + t;
                u[k] = U[i];
                w[k] = Uf[i];
                mk[k] = kSum && !none ? fz.inside[i >> 5] : 0u;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int t = t0 + k * kThreads;
                if (t >= n) break;
                const int i = row0 * N + t;
                C v = u[k];
                if (kSum) {
                    const bool in = ((mk[k] >> (i & 31)) & 1u) != 0;
                    C s = v;
                    s.x = v.x + w[k].x;
                    s.y = v.y + w[k].y;
                    v = none ? w[k] : (in ? v : s);
                }
                a[t] = v;
                a[n + t] = w[k];
            }
        }
    } else if (kSum && f == 0 && fz.inside) {
        for (int t = threadIdx.x; t < nb * N; t += kThreads) {
            const size_t i = (size_t)row0 * N + t;
            a[t] = fused_sum(fz, (int)i, U[i], Uf[i]);
        }
    } else if (kSum && f == 0) {
        // the spheres within their radius of these rows' bounding box (the
        // box of the four corner pixels holds every pixel of the rows): the
        // others cannot hold one of their pixels
        __shared__ unsigned long long s_near;
        if (threadIdx.x < 64) {
            bool near = false;
            if ((int)threadIdx.x < fz.ns && fz.ns <= 64) {
                const double v0 = (double)row0 / (double)fz.vR;
                const double v1 = (double)(row0 + nb - 1) / (double)fz.vR;
                const double u1 = (double)(fz.uR - 1) / (double)fz.uR;
                const d3 c0 = quad_at(fz.q, 0.0, v0), c1 = quad_at(fz.q, u1, v0);
                const d3 c2 = quad_at(fz.q, 0.0, v1), c3 = quad_at(fz.q, u1, v1);
                const double4 S = fz.sph[threadIdx.x];
                const double cc[3] = {S.x, S.y, S.z};
                const double q0[3] = {c0.x, c0.y, c0.z}, q1[3] = {c1.x, c1.y, c1.z};
                const double q2[3] = {c2.x, c2.y, c2.z}, q3[3] = {c3.x, c3.y, c3.z};
                double d2 = 0.0;
                for (int x = 0; x < 3; ++x) {
                    const double lo = fmin(fmin(q0[x], q1[x]), fmin(q2[x], q3[x]));
                    const double hi = fmax(fmax(q0[x], q1[x]), fmax(q2[x], q3[x]));
                    const double g = fmax(fmax(lo - cc[x], cc[x] - hi), 0.0);
                    d2 += g * g;
                }
                near = d2 <= S.w * S.w * (1.0 + 1e-9);
            }
            const unsigned long long m = __ballot(near);
            if (threadIdx.x == 0) s_near = m;
        }
        __syncthreads();
        const unsigned long long near = s_near;
        for (int t = threadIdx.x; t < nb * N; t += kThreads) {
            const size_t i = (size_t)row0 * N + t;
            const C w = Uf[i];
            a[t] = fused_sum(fz, (int)i, U[i], w, near);
            if (kBoth) a[nb * N + t] = w;
        }
    } else {
        for (int t = threadIdx.x; t < nb * N; t += kThreads) a[t] = F[(size_t)row0 * N + t];
    }
    // (the barrier below covers the copy)
    if constexpr (kTwLds) tws.store(b + (size_t)kF * B * N);
    // the column of the thread's first output, read before the transform
    // (the rest, if any, as they come)
    const int nout = kF * nb * Nc;
    const int col0 = (int)threadIdx.x < nout ? cols[(int)threadIdx.x % Nc] : 0;
    __syncthreads();
    const C* res = fixed_fft<R, N>(a, b, kF * nb, twp, false);
    for (int t = threadIdx.x; t < nout; t += kThreads) {
        const int h = kBoth ? t / (nb * Nc) : 0, u = t - h * (nb * Nc);
        const int g = u / Nc, c = u - g * Nc;
        const int col = t == (int)threadIdx.x ? col0 : cols[c];
        G[((size_t)(f + h) * Nc + c) * nrows + row0 + g] = res[(h * nb + g) * N + col];
    }
}

template <class R, int N, bool kSum>
__global__ __launch_bounds__(kThreads) void rows_fwd_kernel(
    const typename C2<R>::T* __restrict__ U, const typename C2<R>::T* __restrict__ Uf, int nrows,
    int B, const int* __restrict__ cols, int Nc, typename C2<R>::T* __restrict__ G,
    const typename C2<R>::T* __restrict__ tw, SumFuse fz) {
    rows_fwd_body<R, N, kSum>(U, Uf, nrows, B, cols, Nc, G, tw, fz, blockIdx.x, blockIdx.y);
}

// Pass-band columns of the compact spectrum, each: forward FFT along v,
// objective band-pass, inverse FFT, scale 1/(uR vR).  In place in G.
template <class R, int M>
__device__ __forceinline__ void cols_band_body(typename C2<R>::T* __restrict__ G, int Nc, int B,
                                               const int* __restrict__ cols, const PostOp& bp,
                                               const typename C2<R>::T* __restrict__ tw, int blk,
                                               int f) {
    using C = typename C2<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C* a = reinterpret_cast<C*>(smem_raw);
    C* b = a + (size_t)B * M;
    const int c0 = blk * B;
    const int nb = min(B, Nc - c0);
    C* line = G + ((size_t)f * Nc + c0) * M;
    for (int t = threadIdx.x; t < nb * M; t += kThreads) a[t] = line[t];
    __syncthreads();
    C* res = fixed_fft<R, M>(a, b, nb, tw, false);
    C* other = res == a ? b : a;
    for (int t = threadIdx.x; t < nb * M; t += kThreads) {
        const int g = t / M, iv = t - g * M;
        if (!passband(bp, cols[c0 + g], iv)) res[t] = cmk<C>((R)0, (R)0);
    }
    __syncthreads();
    res = fixed_fft<R, M>(res, other, nb, tw, true);
    const R sc = (R)bp.scale;
    for (int t = threadIdx.x; t < nb * M; t += kThreads) {
        const C v = res[t];
        line[t] = cmk<C>(v.x * sc, v.y * sc);
    }
}

template <class R, int M>
__global__ __launch_bounds__(kThreads) void cols_band_kernel(typename C2<R>::T* __restrict__ G,
                                                             int Nc, int B,
                                                             const int* __restrict__ cols,
                                                             PostOp bp,
                                                             const typename C2<R>::T* __restrict__ tw) {
    cols_band_body<R, M>(G, Nc, B, cols, bp, tw, blockIdx.x, blockIdx.y);
}

// Inverse FFT of rows [row_begin, row_begin + nrows) of U/Uf from the compact
// spectrum (zeros outside the pass band).  Writes |.|^2 (as |.| squared, like
// resample_intensity) into I[(field*nrows + row - row_begin)*N + iu] and, when
// Uw/Ufw are given, the band-passed field itself.
template <class R, int N>
__global__ __launch_bounds__(kThreads) void rows_inv_kernel(
    const typename C2<R>::T* __restrict__ G, int Nc, int vR, const int* __restrict__ cols,
    int row_begin, int nrows, int B, R* __restrict__ I, typename C2<R>::T* __restrict__ Uw,
    typename C2<R>::T* __restrict__ Ufw, const typename C2<R>::T* __restrict__ tw) {
    using C = typename C2<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C* a = reinterpret_cast<C*>(smem_raw);
    C* b = a + (size_t)B * N;
    const int f = blockIdx.y;
    const int r0 = row_begin + blockIdx.x * B;
    const int nb = min(B, row_begin + nrows - r0);
    for (int t = threadIdx.x; t < nb * N; t += kThreads) a[t] = cmk<C>((R)0, (R)0);
    __syncthreads();
    for (int t = threadIdx.x; t < nb * Nc; t += kThreads) {
        const int g = t / Nc, c = t - g * Nc;
        a[g * N + cols[c]] = G[((size_t)f * Nc + c) * vR + r0 + g];
    }
    __syncthreads();
    const C* res = fixed_fft<R, N>(a, b, nb, tw, true);
    C* W = f ? Ufw : Uw;
    for (int t = threadIdx.x; t < nb * N; t += kThreads) {
        const C v = res[t];
        if (W) W[(size_t)r0 * N + t] = v;
        const R m = sqrt(v.x * v.x + v.y * v.y);
        I[((size_t)f * nrows + (r0 - row_begin)) * N + t] = m * m;
    }
}

// rows_inv_kernel on the U and Uf rows of a block together (2 nb rows in
// one transform), then, for a detector without supersampling (ss = 1) or
// filtered-field output, detect_from_intensity on those rows: D += |U|^2,
// Di += |Uf|^2 (|.| squared, as the intensity rows), Abs = -log10(D / Di)
// over the crop columns [off, off + Nd); the rows are those of the crop.
// kAdd false: D = |U|^2 and Di = |Uf|^2 (the same bits as adding to
// cleared images), for the batches, whose passes then clear nothing.
// kTwLds: the twiddles read from a copy in LDS, as in rows_fwd_body.
template <class R, int N, bool kAdd = true, bool kTwLds = false>
__device__ __forceinline__ void rows_inv_detect_body(
    const typename C2<R>::T* __restrict__ G, int Nc, int vR, const int* __restrict__ cols,
    int row_begin, int nrows, int B, int off, int Nd, R* __restrict__ D, R* __restrict__ Di,
    R* __restrict__ Abs, const typename C2<R>::T* __restrict__ tw, int blk) {
    using C = typename C2<R>::T;
    extern __shared__ unsigned char smem_raw[];
    C* a = reinterpret_cast<C*>(smem_raw);
    C* b = a + (size_t)2 * B * N;
    const int r0 = row_begin + blk * B;
    const int nb = min(B, row_begin + nrows - r0);
    const C* twp = tw;
    TwShare<C, N> tws;
    if constexpr (kTwLds) {
        tws.load(tw);
        twp = b + (size_t)2 * B * N;
    }
    // the spectrum element of the thread's first scatter and its column,
    // loaded beside the twiddles before the clearing (the rest, if any, as
    // they come)
    const int nsc = 2 * nb * Nc;
    auto spec = [&](int t, int* col) {
        const int f = t / (nb * Nc), u = t - f * (nb * Nc);
        const int g = u / Nc, c = u - g * Nc;
        *col = (f * nb + g) * N + cols[c];
        return G[((size_t)f * Nc + c) * vR + r0 + g];
    };
    int col0 = 0;
    C v0 = cmk<C>((R)0, (R)0);
    if ((int)threadIdx.x < nsc) v0 = spec(threadIdx.x, &col0);
    for (int t = threadIdx.x; t < 2 * nb * N; t += kThreads) a[t] = cmk<C>((R)0, (R)0);
    if constexpr (kTwLds) tws.store(b + (size_t)2 * B * N);
    __syncthreads();
    for (int t = threadIdx.x; t < nsc; t += kThreads) {
        int col = col0;
        const C v = t == (int)threadIdx.x ? v0 : spec(t, &col);
        a[col] = v;
    }
    __syncthreads();
    const C* res = fixed_fft<R, N>(a, b, 2 * nb, twp, true);
    for (int t = threadIdx.x; t < nb * Nd; t += kThreads) {
        const int g = t / Nd, iu = t - g * Nd;
        const C v = res[g * N + off + iu], w = res[(nb + g) * N + off + iu];
        const R m = sqrt(v.x * v.x + v.y * v.y), mf = sqrt(w.x * w.x + w.y * w.y);
        const size_t i = (size_t)(r0 - row_begin + g) * Nd + iu;
        const R d = kAdd ? D[i] + m * m : m * m, di = kAdd ? Di[i] + mf * mf : mf * mf;
        D[i] = d;
        Di[i] = di;
        if (Abs) Abs[i] = -log10(d / di);
    }
}

template <class R, int N>
__global__ __launch_bounds__(kThreads) void rows_inv_detect_kernel(
    const typename C2<R>::T* __restrict__ G, int Nc, int vR, const int* __restrict__ cols,
    int row_begin, int nrows, int B, int off, int Nd, R* __restrict__ D, R* __restrict__ Di,
    R* __restrict__ Abs, const typename C2<R>::T* __restrict__ tw) {
    rows_inv_detect_body<R, N>(G, Nc, vR, cols, row_begin, nrows, B, off, Nd, D, Di, Abs, tw,
                               blockIdx.x);
}

// ------------------------------------------ batched band-limited detector
// The three passes above for many instances of one FFT size N at once: the
// block's instance from DetInst[blockIdx.z], its fields, columns, compact
// spectrum and images at their offsets.  Blocks past an instance's rows or
// columns return (uniformly per block).
template <class R, int N, bool kSum, bool kTwLds>
__global__ __launch_bounds__(kThreads) void rows_fwd_batch_kernel(
    const DetInst* __restrict__ insts, int n, const typename C2<R>::T* __restrict__ U,
    const typename C2<R>::T* __restrict__ Uf, int B, const int* __restrict__ cols,
    typename C2<R>::T* __restrict__ G, const typename C2<R>::T* __restrict__ tw) {
    for (int z = blockIdx.z; z < n; z += gridDim.z) {
        const DetInst& I = insts[z];
        if (I.Nc == 0) continue;
        rows_fwd_body<R, N, kSum, true, kTwLds>(U + I.uoff, Uf + I.uoff, N, B, cols + I.cols_off,
                                                I.Nc, G + I.g_off, tw, I.fz, blockIdx.x, 0);
        __syncthreads();   // the block's LDS is reused by its next instance
    }
}

template <class R, int M>
__global__ __launch_bounds__(kThreads) void cols_band_batch_kernel(
    const DetInst* __restrict__ insts, int n, int B, const int* __restrict__ cols,
    typename C2<R>::T* __restrict__ G, const typename C2<R>::T* __restrict__ tw) {
    for (int z = blockIdx.z; z < n; z += gridDim.z) {
        const DetInst& I = insts[z];
        if ((int)blockIdx.x * B >= I.Nc) continue;
        cols_band_body<R, M>(G + I.g_off, I.Nc, B, cols + I.cols_off, I.bp, tw, blockIdx.x,
                             blockIdx.y);
        __syncthreads();
    }
}

template <class R, int N, bool kTwLds>
__global__ __launch_bounds__(kThreads) void rows_inv_detect_batch_kernel(
    const DetInst* __restrict__ insts, int n, int B, const int* __restrict__ cols,
    const typename C2<R>::T* __restrict__ G, R* __restrict__ D, R* __restrict__ Di,
    R* __restrict__ Abs, const typename C2<R>::T* __restrict__ tw) {
    for (int z = blockIdx.z; z < n; z += gridDim.z) {
        const DetInst& I = insts[z];
        if ((int)blockIdx.x * B >= I.Nd) continue;
        rows_inv_detect_body<R, N, false, kTwLds>(G + I.g_off, I.Nc, N, cols + I.cols_off, I.off,
                                                  I.Nd, B, I.off, I.Nd, D + I.image_off,
                                                  Di + I.image_off,
                                                  Abs ? Abs + I.image_off : nullptr, tw, blockIdx.x);
        __syncthreads();
    }
}

template <class R>
__global__ void fill_zero(R* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (R)0;
}

// crop + resample_intensity (+=) + calc_absorbance from intensity rows
template <class R>
__global__ void detect_from_intensity(const R* __restrict__ I, int nrows, int uR, int off, int Nd,
                                      int ss, R* __restrict__ D, R* __restrict__ Di,
                                      R* __restrict__ Abs) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nd * Nd) return;
    const int iv = i / Nd, iu = i - iv * Nd;
    const R* I0 = I;
    const R* I1 = I + (size_t)nrows * uR;
    R s = 0, si = 0;
    for (int ix = 0; ix < ss; ++ix)
        for (int iy = 0; iy < ss; ++iy) {
            const size_t fi = (size_t)(iv * ss + iy) * uR + (off + iu * ss + ix);
            s += I0[fi];
            si += I1[fi];
        }
    const R d = D[i] + s / (R)(ss * ss);
    const R di = Di[i] + si / (R)(ss * ss);
    D[i] = d;
    Di[i] = di;
    if (Abs) Abs[i] = -log10(d / di);
}

#define BIMSIM_FFT_SIZES(X) X(64) X(96) X(128) X(192) X(256) X(384) X(512) X(768) X(1024) \
    X(1536) X(2048) X(3072) X(4096)

bool fixed_size(int n) {
    switch (n) {
#define BIMSIM_X(NN) case NN:
        BIMSIM_FFT_SIZES(BIMSIM_X)
#undef BIMSIM_X
        return true;
        default:
            return false;
    }
}

// The band-limited LDS FFT path: both sizes tabled and a row pair in LDS.
bool band_path(const bimsim_ctx* ctx, size_t csize, int uR, int vR) {
    return fixed_size(uR) && fixed_size(vR) && csize * 2 * (size_t)std::max(uR, vR) <= ctx->lds_max;
}

template <class R>
int launch_rows_fwd(bimsim_ctx* ctx, int N, dim3 grid, size_t lds, const typename C2<R>::T* U,
                    const typename C2<R>::T* Uf, int nrows, int B, const int* cols, int Nc,
                    typename C2<R>::T* G, const typename C2<R>::T* tw, const SumFuse* fz) {
    // the summing form exists for complex64 fields (the FP32 batches) only
    constexpr bool f32 = std::is_same<R, float>::value;
    if (fz && !f32) return set_error(ctx, BIMSIM_E_UNSUPPORTED, "fused sum needs complex64 fields");
    switch (N) {
#define BIMSIM_X(NN)                                                                          \
    case NN:                                                                                  \
        if (f32 && fz)                                                                        \
            hipLaunchKernelGGL((rows_fwd_kernel<R, NN, f32>), grid, dim3(kThreads), lds,      \
                               ctx->stream, U, Uf, nrows, B, cols, Nc, G, tw, *fz);           \
        else                                                                                  \
            hipLaunchKernelGGL((rows_fwd_kernel<R, NN, false>), grid, dim3(kThreads), lds,    \
                               ctx->stream, U, Uf, nrows, B, cols, Nc, G, tw, SumFuse{});     \
        break;
        BIMSIM_FFT_SIZES(BIMSIM_X)
#undef BIMSIM_X
    }
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

template <class R>
int launch_cols_band(bimsim_ctx* ctx, int M, dim3 grid, size_t lds, typename C2<R>::T* G, int Nc,
                     int B, const int* cols, const PostOp& bp, const typename C2<R>::T* tw) {
    switch (M) {
#define BIMSIM_X(NN)                                                                            \
    case NN:                                                                                    \
        hipLaunchKernelGGL((cols_band_kernel<R, NN>), grid, dim3(kThreads), lds, ctx->stream, G, \
                           Nc, B, cols, bp, tw);                                                \
        break;
        BIMSIM_FFT_SIZES(BIMSIM_X)
#undef BIMSIM_X
    }
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

template <class R>
int launch_rows_inv(bimsim_ctx* ctx, int N, dim3 grid, size_t lds, const typename C2<R>::T* G,
                    int Nc, int vR, const int* cols, int row_begin, int nrows, int B, R* I,
                    typename C2<R>::T* Uw, typename C2<R>::T* Ufw, const typename C2<R>::T* tw) {
    switch (N) {
#define BIMSIM_X(NN)                                                                            \
    case NN:                                                                                    \
        hipLaunchKernelGGL((rows_inv_kernel<R, NN>), grid, dim3(kThreads), lds, ctx->stream, G, \
                           Nc, vR, cols, row_begin, nrows, B, I, Uw, Ufw, tw);                  \
        break;
        BIMSIM_FFT_SIZES(BIMSIM_X)
#undef BIMSIM_X
    }
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

template <class R>
int launch_rows_inv_detect(bimsim_ctx* ctx, int N, dim3 grid, size_t lds,
                           const typename C2<R>::T* G, int Nc, int vR, const int* cols,
                           int row_begin, int nrows, int B, int off, int Nd, R* D, R* Di, R* Abs,
                           const typename C2<R>::T* tw) {
    switch (N) {
#define BIMSIM_X(NN)                                                                       \
    case NN:                                                                               \
        hipLaunchKernelGGL((rows_inv_detect_kernel<R, NN>), grid, dim3(kThreads), lds,     \
                           ctx->stream, G, Nc, vR, cols, row_begin, nrows, B, off, Nd, D, \
                           Di, Abs, tw);                                                   \
        break;
        BIMSIM_FFT_SIZES(BIMSIM_X)
#undef BIMSIM_X
    }
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

// Pass-band columns (|f_u| <= NAout/lambda with the band-pass's own index
// map): |u| grows with iu up to uR/2 and shrinks after it, so they are the
// first n1 columns and the last n2.
void band_columns(const PostOp& bp, int uR, int* n1, int* n2) {
    *n1 = *n2 = 0;
    for (int iu = 0; iu < uR; ++iu) {
        double u;
        if (iu <= uR / 2) u = (double)iu * bp.du;
        else u = -(double)(bp.quirks ? uR - 1 - iu : uR - iu) * bp.du;
        if (!(std::fabs(u) > bp.fmax)) ++(iu <= uR / 2 ? *n1 : *n2);
    }
}

std::vector<int> band_column_list(const PostOp& bp, int uR) {
    int n1, n2;
    band_columns(bp, uR, &n1, &n2);
    std::vector<int> cols;
    for (int c = 0; c < n1; ++c) cols.push_back(c);
    for (int c = 0; c < n2; ++c) cols.push_back(uR - n2 + c);
    return cols;
}

// Band-limited detector for FFT lengths with a fixed plan, on the given
// work buffers: the pass-band column list colsd (Nc entries), the compact
// spectrum G (2 Nc vR complex) and, unless the inverse row pass detects
// on its own rows, the intensity rows I (2 nrows uR real).
template <class R>
int detector_band(bimsim_ctx* ctx, const PostOp& bp, void* Uin, void* Ufin, int uR, int vR,
                  int padding, int Nd, int ss, void* D, void* Di, void* Abs, int write_filtered,
                  const typename C2<R>::T* tw_u, const typename C2<R>::T* tw_v,
                  const SumFuse* fz, const int* colsd, int Nc, void* G, void* I) {
    using C = typename C2<R>::T;
    const int S = Nd * ss, off = padding * S;
    const int row_begin = write_filtered ? 0 : off;
    const int nrows = write_filtered ? vR : S;
    // rows per workgroup: at most 1024 points, and at least 512 workgroups
    // per field where the rows allow it
    auto rows_per_group = [&](int nr) {
        int b = std::max(1, std::min(8, 1024 / uR));
        while (b > 1 && (nr + b - 1) / b < 512) b /= 2;
        return b;
    };
    if (Nc > 0) {
        int B = rows_per_group(vR);
        size_t lds = sizeof(C) * 2 * (size_t)B * uR;
        BIMSIM_TRY(launch_rows_fwd<R>(ctx, uR, dim3((vR + B - 1) / B, 2), lds, (const C*)Uin,
                                      (const C*)Ufin, vR, B, colsd, Nc, (C*)G, tw_u, fz));
        int Bc = std::max(1, std::min(8, 2048 / vR));
        Bc = std::min(Bc, Nc);
        // a narrow pass band leaves few columns: fewer per workgroup, so
        // that the column pass still spreads over the chip
        while (Bc > 1 && (Nc + Bc - 1) / Bc < 128) Bc /= 2;
        lds = sizeof(C) * 2 * (size_t)Bc * vR;
        BIMSIM_TRY(launch_cols_band<R>(ctx, vR, dim3((Nc + Bc - 1) / Bc, 2), lds, (C*)G, Nc, Bc,
                                       colsd, bp, tw_v));
    }
    int B = rows_per_group(nrows);
    size_t lds = sizeof(C) * 2 * (size_t)B * uR;
    if (Nc > 0 && ss == 1 && !write_filtered && 2 * lds <= ctx->lds_max) {
        // both fields' rows per workgroup, detection on them in the same pass
        return launch_rows_inv_detect<R>(ctx, uR, dim3((nrows + B - 1) / B), 2 * lds, (const C*)G,
                                         Nc, vR, colsd, row_begin, nrows, B, off, Nd, (R*)D,
                                         (R*)Di, (R*)Abs, tw_u);
    }
    if (Nc > 0) {
        BIMSIM_TRY(launch_rows_inv<R>(ctx, uR, dim3((nrows + B - 1) / B, 2), lds, (const C*)G, Nc,
                                      vR, colsd, row_begin, nrows, B, (R*)I,
                                      write_filtered ? (C*)Uin : nullptr,
                                      write_filtered ? (C*)Ufin : nullptr, tw_u));
    } else {
        // nothing passes the objective: zero intensities (and fields)
        const size_t ni = 2 * (size_t)nrows * uR;
        hipLaunchKernelGGL(fill_zero<R>, dim3((unsigned)((ni + 255) / 256)), dim3(256), 0,
                           ctx->stream, (R*)I, ni);
        if (write_filtered) {
            const size_t nf = 2 * (size_t)uR * vR;   // complex values as pairs of R
            hipLaunchKernelGGL(fill_zero<R>, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0,
                               ctx->stream, (R*)Uin, nf);
            hipLaunchKernelGGL(fill_zero<R>, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0,
                               ctx->stream, (R*)Ufin, nf);
        }
        BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    }
    const R* Ic = (const R*)I + (size_t)(off - row_begin) * uR;
    // detect_from_intensity indexes rows relative to the crop start
    hipLaunchKernelGGL(detect_from_intensity<R>, dim3((Nd * Nd + 255) / 256), dim3(256), 0,
                       ctx->stream, Ic, nrows, uR, off, Nd, ss, (R*)D, (R*)Di, (R*)Abs);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

template <class R>
int run_fft_pass(bimsim_ctx* ctx, typename C2<R>::T* data, int nlines, int n, long long line_stride,
                 long long elem_stride, const typename C2<R>::T* tw, int inverse, const PostOp& op) {
    using C = typename C2<R>::T;
    FftPlan plan;
    if (make_plan(n, &plan) != 0)
        return set_error(ctx, BIMSIM_E_UNSUPPORTED, "FFT length %d is not a product of 2, 3 and 5", n);
    // lines per workgroup: as many as fit the LDS ping-pong buffers
    int B = std::min(16, std::min(nlines, std::max(1, kMaxElems / n)));
    size_t lds = 0;
    int P = n;
    for (; B >= 1; --B) {
        P = B == 1 ? n : n + 1;
        lds = sizeof(C) * 2 * (size_t)B * P;
        if (lds <= ctx->lds_max) break;
    }
    if (B < 1)
        return set_error(ctx, BIMSIM_E_UNSUPPORTED, "FFT length %d does not fit in LDS", n);
    const dim3 grid((nlines + B - 1) / B);
    hipLaunchKernelGGL(fft_lines_kernel<R>, grid, dim3(kThreads), lds, ctx->stream, data, nlines,
                       B, P, line_stride, elem_stride, plan, tw, inverse, op);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

template <class R>
int detector_impl(bimsim_ctx* ctx, void* Uin, void* Ufin, int uR, int vR, double du, double dv,
                  double NAin, double NAout, double lambda, int padding, int Nd, int ss, void* D,
                  void* Di, void* Abs, int write_filtered, int quirks, const SumFuse* fz) {
    using C = typename C2<R>::T;
    const size_t npix = (size_t)uR * vR;
    C* U = (C*)Uin;
    C* Uf = (C*)Ufin;
    if (band_path(ctx, sizeof(C), uR, vR)) {
        const double *twd_u, *twd_v;
        const float *twf_u, *twf_v;
        BIMSIM_TRY(twiddles_get(ctx, uR, &twd_u, &twf_u));
        BIMSIM_TRY(twiddles_get(ctx, vR, &twd_v, &twf_v));
        PostOp bp{};
        bp.kind = 1;
        bp.uR = uR;
        bp.vR = vR;
        bp.quirks = quirks;
        bp.du = du;
        bp.dv = dv;
        bp.fmin = NAin / lambda;
        bp.fmax = NAout / lambda;
        bp.scale = 1.0 / ((unsigned)uR * (unsigned)vR);   // ScaleField(1.0/(R0*R1))
        const bool dbl = std::is_same<R, double>::value;
        const std::vector<int> cols = band_column_list(bp, uR);
        const int Nc = (int)cols.size();
        const int nrows = write_filtered ? vR : Nd * ss;
        const int* colsd = nullptr;
        void *G = nullptr, *I = nullptr;
        BIMSIM_TRY(int_table_get(ctx, cols, &colsd));
        BIMSIM_TRY(scratch_get(ctx, kSlotFftA, sizeof(C) * 2 * (size_t)std::max(Nc, 1) * vR, &G));
        BIMSIM_TRY(scratch_get(ctx, kSlotFftB, sizeof(R) * 2 * (size_t)nrows * uR, &I));
        return detector_band<R>(ctx, bp, Uin, Ufin, uR, vR, padding, Nd, ss, D, Di, Abs,
                                write_filtered, dbl ? (const C*)twd_u : (const C*)twf_u,
                                dbl ? (const C*)twd_v : (const C*)twf_v, fz, colsd, Nc, G, I);
    }
    if (fz) return set_error(ctx, BIMSIM_E_UNSUPPORTED, "fused sum needs the band-limited detector");
    if (!write_filtered) {
        void *a = nullptr, *b = nullptr;
        BIMSIM_TRY(scratch_get(ctx, kSlotFftA, sizeof(C) * npix, &a));
        BIMSIM_TRY(scratch_get(ctx, kSlotFftB, sizeof(C) * npix, &b));
        BIMSIM_HIP_CHECK(ctx, hipMemcpyAsync(a, U, sizeof(C) * npix, hipMemcpyDeviceToDevice, ctx->stream));
        BIMSIM_HIP_CHECK(ctx, hipMemcpyAsync(b, Uf, sizeof(C) * npix, hipMemcpyDeviceToDevice, ctx->stream));
        U = (C*)a;
        Uf = (C*)b;
    }
    const double *twd_u, *twd_v;
    const float *twf_u, *twf_v;
    BIMSIM_TRY(twiddles_get(ctx, uR, &twd_u, &twf_u));
    BIMSIM_TRY(twiddles_get(ctx, vR, &twd_v, &twf_v));
    const C* tw_u = std::is_same<R, double>::value ? (const C*)twd_u : (const C*)twf_u;
    const C* tw_v = std::is_same<R, double>::value ? (const C*)twd_v : (const C*)twf_v;
    PostOp none{};
    PostOp bp{};
    bp.kind = 1;
    bp.uR = uR;
    bp.vR = vR;
    bp.quirks = quirks;
    bp.du = du;
    bp.dv = dv;
    bp.fmin = NAin / lambda;
    bp.fmax = NAout / lambda;
    PostOp sc{};
    sc.kind = 2;
    sc.scale = 1.0 / ((unsigned)uR * (unsigned)vR);   // ScaleField(1.0/(R0*R1))
    C* fields[2] = {U, Uf};
    for (C* F : fields) {
        // forward: rows (length uR), then columns (length vR) + band-pass
        BIMSIM_TRY(run_fft_pass<R>(ctx, F, vR, uR, uR, 1, tw_u, 0, none));
        BIMSIM_TRY(run_fft_pass<R>(ctx, F, uR, vR, 1, uR, tw_v, 0, bp));
        // inverse: columns, then rows + 1/(uR vR)
        BIMSIM_TRY(run_fft_pass<R>(ctx, F, uR, vR, 1, uR, tw_v, 1, none));
        BIMSIM_TRY(run_fft_pass<R>(ctx, F, vR, uR, uR, 1, tw_u, 1, sc));
    }
    const int off = padding * Nd * ss;
    hipLaunchKernelGGL(detect_kernel<R>, dim3((Nd * Nd + 255) / 256), dim3(256), 0, ctx->stream,
                       (const C*)U, (const C*)Uf, uR, off, Nd, ss, (R*)D, (R*)Di, (R*)Abs);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

// dst (M x Mv) = src (uR x vR) at the corner, zeros elsewhere.
template <class C>
__global__ void embed_kernel(const C* __restrict__ src, int uR, int vR, C* __restrict__ dst, int M,
                             int Mv) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)M * Mv) return;
    const int iv = (int)(i / M), iu = (int)(i - (long long)iv * M);
    C v;
    v.x = 0;
    v.y = 0;
    if (iu < uR && iv < vR) v = src[(long long)iv * uR + iu];
    dst[i] = v;
}

// The detector on the near field zero-filled to (fill uR) x (fill vR): both
// fields embedded at the corner of scratch arrays, the FFT and band-pass on
// the finer frequency grid (du/fill, dv/fill), the crop of the field's own
// uR x uR samples at the corner.
template <class R>
int detector_zero_filled_impl(bimsim_ctx* ctx, const void* U, const void* Uf, int uR, int vR,
                              int fill, double du, double dv, double NAin, double NAout,
                              double lambda, void* D, void* Di, void* Abs, int quirks) {
    using C = typename C2<R>::T;
    const int M = fill * uR, Mv = fill * vR;
    const size_t n = (size_t)M * Mv;
    void *a = nullptr, *b = nullptr;
    BIMSIM_TRY(scratch_get(ctx, kSlotDet, sizeof(C) * 2 * n, &a));
    b = (char*)a + sizeof(C) * n;
    const dim3 grid((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(embed_kernel<C>, grid, dim3(256), 0, ctx->stream, (const C*)U, uR, vR, (C*)a,
                       M, Mv);
    hipLaunchKernelGGL(embed_kernel<C>, grid, dim3(256), 0, ctx->stream, (const C*)Uf, uR, vR,
                       (C*)b, M, Mv);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    // the embedded copies are ours: the transform may run in place on them
    return detector_impl<R>(ctx, a, b, M, Mv, du / fill, dv / fill, NAin, NAout, lambda, 0, uR, 1,
                            D, Di, Abs, 1, quirks, nullptr);
}

// fieldslice::toAngularSpectrum / fromAngularSpectrum (fieldslice.cpp:
// 25-94): in-place forward (e^{-i}) or inverse (e^{+i}) 2D FFT, the inverse
// followed by the 1/(R0 R1) of ScaleField.
template <class R>
int field_fft_impl(bimsim_ctx* ctx, void* Fin, int uR, int vR, int inverse) {
    using C = typename C2<R>::T;
    C* F = (C*)Fin;
    const double *twd_u, *twd_v;
    const float *twf_u, *twf_v;
    BIMSIM_TRY(twiddles_get(ctx, uR, &twd_u, &twf_u));
    BIMSIM_TRY(twiddles_get(ctx, vR, &twd_v, &twf_v));
    const bool dbl = std::is_same<R, double>::value;
    const C* tw_u = dbl ? (const C*)twd_u : (const C*)twf_u;
    const C* tw_v = dbl ? (const C*)twd_v : (const C*)twf_v;
    PostOp none{};
    if (!inverse) {
        BIMSIM_TRY(run_fft_pass<R>(ctx, F, vR, uR, uR, 1, tw_u, 0, none));
        BIMSIM_TRY(run_fft_pass<R>(ctx, F, uR, vR, 1, uR, tw_v, 0, none));
    } else {
        PostOp sc{};
        sc.kind = 2;
        sc.scale = 1.0 / ((unsigned)uR * (unsigned)vR);
        BIMSIM_TRY(run_fft_pass<R>(ctx, F, uR, vR, 1, uR, tw_v, 1, none));
        BIMSIM_TRY(run_fft_pass<R>(ctx, F, vR, uR, uR, 1, tw_u, 1, sc));
    }
    return BIMSIM_OK;
}

}  // namespace
}  // namespace bimsim

using namespace bimsim;

namespace bimsim {

int sum_mask_launch(bimsim_ctx* ctx, const SumFuse& fz, unsigned* mask) {
    SumFuse f = fz;
    f.inside = nullptr;
    const int words = (fz.uR * fz.vR + 31) / 32;
    hipLaunchKernelGGL(sum_mask_kernel, dim3((words + kThreads - 1) / kThreads), dim3(kThreads), 0,
                       ctx->stream, f, mask);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

bool detector_fusable(const bimsim_ctx* ctx, int uR, int vR) {
    return band_path(ctx, sizeof(float2), uR, vR);
}

int detector_run(bimsim_ctx* ctx, int prec, int quirks, void* U, void* Uf, int uR, int vR,
                 double du, double dv, double NAin, double NAout, double lambda, int padding,
                 int Nd, int ss, void* D, void* Di, void* Abs, int write_filtered,
                 const SumFuse* fz) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, U && Uf && D && Di, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && Nd > 0 && ss > 0 && padding >= 0, "bad sizes");
    BIMSIM_CHECK_ARG(ctx, (long long)(2 * padding + 1) * Nd * ss <= std::min(uR, vR),
                     "detector crop exceeds the near-field grid");
    BIMSIM_CHECK_ARG(ctx, lambda > 0.0 && du > 0.0 && dv > 0.0, "lambda, du and dv must be positive");
    BIMSIM_CHECK_ARG(ctx, prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST, "unknown precision");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int tmr = timer_begin(ctx, kStageDet);
    int rc;
    if (prec == BIMSIM_FP32_FAST)
        rc = detector_impl<float>(ctx, U, Uf, uR, vR, du, dv, NAin, NAout, lambda, padding, Nd, ss,
                                  D, Di, Abs, write_filtered, quirks, fz);
    else
        rc = detector_impl<double>(ctx, U, Uf, uR, vR, du, dv, NAin, NAout, lambda, padding, Nd,
                                   ss, D, Di, Abs, write_filtered, quirks, fz);
    if (rc != BIMSIM_OK) return rc;
    timer_end(ctx, tmr);
    return BIMSIM_OK;
}

// ------------------------------------------ batched band-limited detector
namespace {

int rows_per_wg(int n) { return std::max(1, std::min(8, 1024 / n)); }

template <class R>
int det_batch_launch_impl(bimsim_ctx* ctx, const DetBatch& b, const void* U, const void* Uf,
                          void* D, void* Di, void* Abs) {
    using C = typename C2<R>::T;
    const int N = b.N;
    const DetInst* I = (const DetInst*)b.d_inst;
    const C* tw = (const C*)b.tw;
    const unsigned nz = (unsigned)std::min(b.n, 65535);
    const size_t lds_r = sizeof(C) * 2 * 2 * (size_t)b.B * N;   // both fields' rows
    const size_t lds_c = sizeof(C) * 2 * (size_t)b.Bc * N;
    const size_t lds_i = sizeof(C) * 2 * 2 * (size_t)b.B2 * N;
    // the row passes read their twiddles from LDS where the N more elements fit
    const size_t lds_tw = sizeof(C) * (size_t)N;
    const bool tw_r = lds_r + lds_tw <= ctx->lds_max, tw_i = lds_i + lds_tw <= ctx->lds_max;
    const hipStream_t st = ctx->stream;
    constexpr bool f32 = std::is_same<R, float>::value;
    const bool sum = f32 && b.fused;
    const dim3 grid_r((N + b.B - 1) / b.B, 1, nz), grid_i((b.max_nd + b.B2 - 1) / b.B2, 1, nz);
    switch (N) {
#define BIMSIM_FWD(NN, S, T)                                                                       \
    hipLaunchKernelGGL((rows_fwd_batch_kernel<R, NN, S, T>), grid_r, dim3(kThreads),               \
                       lds_r + (T ? lds_tw : 0), st, I, b.n, (const C*)U, (const C*)Uf, b.B,       \
                       b.d_cols, (C*)b.d_G, tw)
#define BIMSIM_INV(NN, T)                                                                          \
    hipLaunchKernelGGL((rows_inv_detect_batch_kernel<R, NN, T>), grid_i, dim3(kThreads),           \
                       lds_i + (T ? lds_tw : 0), st, I, b.n, b.B2, b.d_cols, (const C*)b.d_G,      \
                       (R*)D, (R*)Di, (R*)Abs, tw)
#define BIMSIM_X(NN)                                                                               \
    case NN:                                                                                       \
        if (sum && tw_r) BIMSIM_FWD(NN, f32, true);                                                \
        else if (sum) BIMSIM_FWD(NN, f32, false);                                                  \
        else if (tw_r) BIMSIM_FWD(NN, false, true);                                                \
        else BIMSIM_FWD(NN, false, false);                                                         \
        if (b.max_nc > 0)                                                                          \
            hipLaunchKernelGGL((cols_band_batch_kernel<R, NN>),                                    \
                               dim3((b.max_nc + b.Bc - 1) / b.Bc, 2, nz), dim3(kThreads), lds_c,   \
                               st, I, b.n, b.Bc, b.d_cols, (C*)b.d_G, tw);                        \
        if (tw_i) BIMSIM_INV(NN, true);                                                            \
        else BIMSIM_INV(NN, false);                                                                \
        break;
        BIMSIM_FFT_SIZES(BIMSIM_X)
#undef BIMSIM_X
#undef BIMSIM_INV
#undef BIMSIM_FWD
        default:
            return set_error(ctx, BIMSIM_E_UNSUPPORTED, "no batched detector for N = %d", N);
    }
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return BIMSIM_OK;
}

}  // namespace

bool det_batchable(const bimsim_ctx* ctx, int prec, int N, int ss, int zero_fill) {
    const size_t cs = prec == BIMSIM_FP32_FAST ? sizeof(float2) : sizeof(double2);
    return ss == 1 && zero_fill <= 1 && band_path(ctx, cs, N, N) &&
           2 * cs * 2 * (size_t)rows_per_wg(N) * N <= ctx->lds_max;
}

int det_batch_create(bimsim_ctx* ctx, int prec, int quirks, int N, const std::vector<DetParams>& p,
                     bool fused, DetBatch* out) {
    *out = DetBatch();
    const bool f32 = prec == BIMSIM_FP32_FAST;
    const size_t cs = f32 ? sizeof(float2) : sizeof(double2);
    std::vector<DetInst> insts;
    std::vector<int> cols;
    long long g = 0;
    int max_nc = 0, max_nd = 0;
    for (const DetParams& q : p) {
        DetInst I{};
        I.uoff = q.uoff;
        I.image_off = q.image_off;
        I.Nd = q.Nd;
        I.off = q.padding * q.Nd;
        BIMSIM_CHECK_ARG(ctx, I.off + q.Nd <= N, "detector crop exceeds the near-field grid");
        PostOp& bp = I.bp;
        bp.kind = 1;
        bp.uR = N;
        bp.vR = N;
        bp.quirks = quirks;
        bp.du = q.du;
        bp.dv = q.dv;
        bp.fmin = q.NAin / q.lambda;
        bp.fmax = q.NAout / q.lambda;
        bp.scale = 1.0 / ((unsigned)N * (unsigned)N);
        I.fz = q.fz;
        // pass-band columns, as detector_band
        I.cols_off = (int)cols.size();
        const std::vector<int> c = band_column_list(bp, N);
        cols.insert(cols.end(), c.begin(), c.end());
        I.Nc = (int)c.size();
        I.g_off = g;
        g += 2LL * I.Nc * N;
        max_nc = std::max(max_nc, I.Nc);
        max_nd = std::max(max_nd, I.Nd);
        insts.push_back(I);
    }
    out->N = N;
    out->n = (int)insts.size();
    out->f32 = f32;
    out->fused = fused && f32;
    out->max_nc = max_nc;
    out->max_nd = max_nd;
    // rows (columns) per workgroup, fewer when the batch is small so that
    // each pass still spreads over the chip
    const int n = (int)insts.size();
    int B = rows_per_wg(N), B2 = B, Bc = std::max(1, std::min(8, 2048 / N));
    while (B > 1 && (long long)n * ((N + B - 1) / B) < 512) B /= 2;
    while (B2 > 1 && (long long)n * ((max_nd + B2 - 1) / B2) < 512) B2 /= 2;
    while (Bc > 1 && (long long)n * ((max_nc + Bc - 1) / Bc) < 128) Bc /= 2;
    out->B = B;
    out->B2 = B2;
    out->Bc = Bc;
    const double* twd;
    const float* twf;
    BIMSIM_TRY(twiddles_get(ctx, N, &twd, &twf));
    out->tw = f32 ? (const void*)twf : (const void*)twd;
    if (cols.empty()) cols.push_back(0);
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (hipMalloc(&out->d_inst, sizeof(DetInst) * insts.size()) != hipSuccess ||
        hipMalloc((void**)&out->d_cols, sizeof(int) * cols.size()) != hipSuccess ||
        hipMalloc(&out->d_G, cs * (size_t)std::max<long long>(1, g)) != hipSuccess) {
        det_batch_destroy(out);
        return set_error(ctx, BIMSIM_E_NOMEM, "out of device memory for the batched detector");
    }
    BIMSIM_HIP_CHECK(ctx, hipMemcpy(out->d_inst, insts.data(), sizeof(DetInst) * insts.size(),
                                    hipMemcpyHostToDevice));
    BIMSIM_HIP_CHECK(ctx, hipMemcpy(out->d_cols, cols.data(), sizeof(int) * cols.size(),
                                    hipMemcpyHostToDevice));
    return BIMSIM_OK;
}

int det_batch_launch(bimsim_ctx* ctx, const DetBatch& b, const void* U, const void* Uf, void* D,
                     void* Di, void* Abs) {
    if (b.n == 0) return BIMSIM_OK;
    const int tmr = timer_begin(ctx, kStageDet);
    const int rc = b.f32 ? det_batch_launch_impl<float>(ctx, b, U, Uf, D, Di, Abs)
                         : det_batch_launch_impl<double>(ctx, b, U, Uf, D, Di, Abs);
    if (rc != BIMSIM_OK) return rc;
    timer_end(ctx, tmr);
    return BIMSIM_OK;
}

void det_batch_destroy(DetBatch* b) {
    if (b->d_inst) (void)hipFree(b->d_inst);
    if (b->d_cols) (void)hipFree(b->d_cols);
    if (b->d_G) (void)hipFree(b->d_G);
    *b = DetBatch();
}

}  // namespace bimsim

// ---------------------------------------------- detector without allocation
// bimsim_detector_noalloc keeps its work memory (twiddles, pass-band column
// list, compact spectrum, intensity rows) in `scratch`, a file-scope device
// array that every GPU holds from the library's load until the process
// ends.  Its device address is taken once per device at context creation
// (noalloc_scratch_init, with the device current) and kept per device; the
// array's name is a host-side stand-in and never reaches a kernel.
namespace {

constexpr size_t kNoAllocBytes = 64ull << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kNoAllocDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kNoAllocDevices] = {};
bool g_noalloc_tried[kNoAllocDevices] = {};

// e^{-2 pi i t/n}, t < n, evaluated in double (as the context's tables)
template <class C>
__global__ void twiddle_fill_kernel(C* __restrict__ tw, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const double ang = -2.0 * M_PI * (double)t / (double)n;
    tw[t].x = cos(ang);
    tw[t].y = sin(ang);
}

// the first n1 and the last n2 of uR columns
__global__ void band_cols_kernel(int* __restrict__ cols, int n1, int n2, int uR) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n1 + n2) cols[c] = c < n1 ? c : uR - n2 + (c - n1);
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

template <class R>
int detector_noalloc_impl(bimsim_ctx* ctx, unsigned char* work, int quirks, const void* U,
                          const void* Uf, int uR, int vR, double du, double dv, double NAin,
                          double NAout, double lambda, int padding, int Nd, int ss, void* D,
                          void* Di, void* Abs) {
    using C = typename C2<R>::T;
    PostOp bp{};
    bp.kind = 1;
    bp.uR = uR;
    bp.vR = vR;
    bp.quirks = quirks;
    bp.du = du;
    bp.dv = dv;
    bp.fmin = NAin / lambda;
    bp.fmax = NAout / lambda;
    bp.scale = 1.0 / ((unsigned)uR * (unsigned)vR);   // ScaleField(1.0/(R0*R1))
    int n1, n2;
    band_columns(bp, uR, &n1, &n2);
    const int Nc = n1 + n2;
    const int nrows = Nd * ss;
    // work layout: twiddles of uR and vR, columns, compact spectrum, rows
    const size_t o_tu = 0, o_tv = align256(sizeof(C) * uR);
    const size_t o_cols = o_tv + align256(sizeof(C) * vR);
    const size_t o_G = o_cols + align256(sizeof(int) * std::max(Nc, 1));
    const size_t o_I = o_G + align256(sizeof(C) * 2 * (size_t)std::max(Nc, 1) * vR);
    const size_t need = o_I + sizeof(R) * 2 * (size_t)nrows * uR;
    if (need > kNoAllocBytes)
        return set_error(ctx, BIMSIM_E_INVALID,
                         "the detector needs %zu bytes of work memory; the allocation-free "
                         "entry point has %zu",
                         need, kNoAllocBytes);
    C* tu = (C*)(work + o_tu);
    C* tv = (C*)(work + o_tv);
    int* cols = (int*)(work + o_cols);
    hipLaunchKernelGGL(twiddle_fill_kernel<C>, dim3((uR + 255) / 256), dim3(256), 0, ctx->stream,
                       tu, uR);
    hipLaunchKernelGGL(twiddle_fill_kernel<C>, dim3((vR + 255) / 256), dim3(256), 0, ctx->stream,
                       tv, vR);
    if (Nc > 0)
        hipLaunchKernelGGL(band_cols_kernel, dim3((Nc + 255) / 256), dim3(256), 0, ctx->stream,
                           cols, n1, n2, uR);
    BIMSIM_HIP_CHECK(ctx, hipGetLastError());
    return detector_band<R>(ctx, bp, const_cast<void*>(U), const_cast<void*>(Uf), uR, vR, padding,
                            Nd, ss, D, Di, Abs, 0, tu, tv, nullptr, cols, Nc, work + o_G,
                            work + o_I);
}

}  // namespace

namespace bimsim {

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kNoAllocDevices) return BIMSIM_E_INVALID;
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (g_noalloc_tried[device]) return g_noalloc_ptr[device] ? BIMSIM_OK : BIMSIM_E_HIP;
    g_noalloc_tried[device] = true;
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return BIMSIM_E_HIP;
    }
    g_noalloc_ptr[device] = ptr;
    return BIMSIM_OK;
}

}  // namespace bimsim

extern "C" int bimsim_detector_noalloc(bimsim_ctx* ctx, int prec, int quirks, const void* U,
                                       const void* Uf, int uR, int vR, double du, double dv,
                                       double NAin, double NAout, double lambda, int padding,
                                       int Nd, int ss, void* D, void* Di, void* Abs) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, U && Uf && D && Di, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && vR > 0 && Nd > 0 && ss > 0 && padding >= 0, "bad sizes");
    BIMSIM_CHECK_ARG(ctx, (long long)(2 * padding + 1) * Nd * ss <= std::min(uR, vR),
                     "detector crop exceeds the near-field grid");
    BIMSIM_CHECK_ARG(ctx, lambda > 0.0 && du > 0.0 && dv > 0.0, "lambda, du and dv must be positive");
    BIMSIM_CHECK_ARG(ctx, prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST, "unknown precision");
    const bool f32 = prec == BIMSIM_FP32_FAST;
    if (!band_path(ctx, f32 ? sizeof(float2) : sizeof(double2), uR, vR))
        return set_error(ctx, BIMSIM_E_UNSUPPORTED,
                         "the allocation-free detector takes FFT sizes of 64..4096 from "
                         "{2^k, 3 2^k} whose rows fit LDS (got %d x %d)",
                         uR, vR);
    void* work = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noalloc_mu);
        if (ctx->device >= 0 && ctx->device < kNoAllocDevices) work = g_noalloc_ptr[ctx->device];
    }
    if (!work)
        return set_error(ctx, BIMSIM_E_HIP, "the work memory of device %d was not set up",
                         ctx->device);
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return f32 ? detector_noalloc_impl<float>(ctx, (unsigned char*)work, quirks, U, Uf, uR, vR, du,
                                              dv, NAin, NAout, lambda, padding, Nd, ss, D, Di, Abs)
               : detector_noalloc_impl<double>(ctx, (unsigned char*)work, quirks, U, Uf, uR, vR,
                                               du, dv, NAin, NAout, lambda, padding, Nd, ss, D,
                                               Di, Abs);
}

extern "C" int bimsim_detector(bimsim_ctx* ctx, int prec, int quirks, void* U, void* Uf, int uR, int vR,
                               double du, double dv, double NAin, double NAout, double lambda,
                               int padding, int Nd, int ss, void* D, void* Di, void* Abs,
                               int write_filtered) {
    return detector_run(ctx, prec, quirks, U, Uf, uR, vR, du, dv, NAin, NAout, lambda, padding, Nd,
                        ss, D, Di, Abs, write_filtered, nullptr);
}

namespace bimsim {

int detector_zero_filled_run(bimsim_ctx* ctx, int prec, int quirks, const void* U, const void* Uf,
                             int uR, int vR, int fill, double du, double dv, double NAin,
                             double NAout, double lambda, void* D, void* Di, void* Abs) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, U && Uf && D && Di, "null pointer argument");
    BIMSIM_CHECK_ARG(ctx, uR > 0 && uR == vR, "the zero-filled detector needs a square field");
    BIMSIM_CHECK_ARG(ctx, fill >= 1 && fill <= 8 && (long long)fill * uR * fill * vR <= (1LL << 28),
                     "zero-fill factor must be in [1, 8] and the filled field at most 2^28 samples");
    BIMSIM_CHECK_ARG(ctx, lambda > 0.0 && du > 0.0 && dv > 0.0, "lambda, du and dv must be positive");
    BIMSIM_CHECK_ARG(ctx, prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST, "unknown precision");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    const int tmr = timer_begin(ctx, kStageDet);
    const int rc = prec == BIMSIM_FP32_FAST
                       ? detector_zero_filled_impl<float>(ctx, U, Uf, uR, vR, fill, du, dv, NAin,
                                                          NAout, lambda, D, Di, Abs, quirks)
                       : detector_zero_filled_impl<double>(ctx, U, Uf, uR, vR, fill, du, dv, NAin,
                                                           NAout, lambda, D, Di, Abs, quirks);
    if (rc != BIMSIM_OK) return rc;
    timer_end(ctx, tmr);
    return BIMSIM_OK;
}

}  // namespace bimsim

extern "C" int bimsim_detector_zero_filled(bimsim_ctx* ctx, int prec, int quirks, const void* U,
                                           const void* Uf, int uR, int vR, int fill, double du,
                                           double dv, double NAin, double NAout, double lambda,
                                           void* D, void* Di, void* Abs) {
    return detector_zero_filled_run(ctx, prec, quirks, U, Uf, uR, vR, fill, du, dv, NAin, NAout,
                                    lambda, D, Di, Abs);
}

extern "C" int bimsim_field_fft(bimsim_ctx* ctx, int prec, void* F, int uR, int vR, int inverse) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, F && uR > 0 && vR > 0, "bad FFT arguments");
    BIMSIM_CHECK_ARG(ctx, prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST, "unknown precision");
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (prec == BIMSIM_FP32_FAST) return field_fft_impl<float>(ctx, F, uR, vR, inverse);
    return field_fft_impl<double>(ctx, F, uR, vR, inverse);
}
