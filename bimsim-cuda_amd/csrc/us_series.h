// Scattered-field series in the FP32 form for one or many instances
// (slice, spheres, waves), by one of two algorithms.
// Direct form, two phases per chunk of instances:
//   1. a coefficient kernel writes, per (sphere, pixel), the wave-independent
//      products c'_l = s_l B_l h_l(kd) (or s_l A_l j_l(knd) inside) computed
//      in FP64 with the reference recurrences, and r-hat;
//   2. one persistent work-queue launch per order bucket runs the FP32
//      rescaled-Legendre series: work items are (instance, 256-pixel block,
//      wave split) and every workgroup takes items from a counter until the
//      queue is empty.
// Instances are grouped into chunks whose coefficient storage fits a memory
// budget; the same buffers serve every chunk.
// Harmonic form (addition theorem): per-wave Legendre values, azimuth
// phases and wave weights, then per (sphere, m, l) records of the wave sums
// (the wave tables, from the waves only), then one persistent launch per
// order bucket evaluates per pixel the order sums of the records and the
// FP64 radial factors.
#pragma once

#include <vector>

#include "nf_us.h"

namespace bimsim {

// One instance as the kernels see it (device table).
struct SeriesInst {
    QuadArg q;
    d3 f;               // focus
    double kmag;
    double sA;          // (1.0f/nk) * subA
    int uR, vR;
    int s_begin, ns;    // spheres [s_begin, s_begin + ns) of the sphere table
    int nk, nk_pad;     // waves, padded to a multiple of 2
    int wchunk, nsplit; // waves per split, number of splits
    int accumulate;     // U (+)= result
    int lh;             // harmonic form: orders of the instance's Legendre table
    long long wave_off; // first wave (rows of 3 doubles) of the instance
    long long u_off;    // element offset of the instance's field in U
    long long part_off; // element offset of its split partial sums
    long long pl_off;   // harmonic form: its normalised Legendre table
    long long e_off;    // harmonic form: its azimuth phase table
    long long perm_off; // harmonic form: its pixel order in the plan's permutation, or -1
};

struct SeriesSph {
    SphereDev s;
    int inst;           // owning instance
    int nb;             // blocks of 8 orders: ceil((Nl + 1) / 8)
    long long woff;     // offset of its WaveF records (and wave weights)
    long long coff;     // offset of c'[l][pix] in the chunk's coefficient buffer
    long long roff;     // offset of r-hat[pix] in the chunk's r-hat buffer
    long long toff;     // harmonic form: offset of its harmonic records
    int tl;             // harmonic form: record stride (its order bucket)
    int planar;         // harmonic form: the slice contains the centre and is normal
                        // to z, so every pixel has x = r-hat.z = 0
};

struct SeriesItem {
    int inst, pix0, split;
    int grp;            // harmonic form: its sphere group
};

// Harmonic form: the spheres of one instance in one order bucket, a
// consecutive range of the sphere table (the plan sorts each instance's
// spheres by bucket).  The first group of an instance in launch order writes
// (or, with the instance's accumulate, adds to) its field, later groups add.
struct SeriesGroup {
    int s_begin, ns;
    int accumulate;
    int L;
};

// Host description of a batch: the caller fills every SeriesInst field up
// to accumulate / wave_off / u_off, and SeriesSph.s / inst.
struct SeriesBatch {
    std::vector<SeriesInst> inst;
    std::vector<SeriesSph> sph;
};

struct UfInst;
struct SumInst;

// The other work of an FP32 batch's pre-pass, run in one launch with a
// harmonic plan's records (series_plan_launch_prepass): the Mie
// coefficients of n_mie_sph spheres (sphere i at mie_kmag[i]) in blocks of
// spb, Bessel arrays of wstride in LDS, and the complex64 incident fields
// of n_uf instances, uf_blocks blocks of 256 pixels each; with sum, also
// the clearing of each instance's detector images in D and Di (sum[i] for
// instance i), which the batch's sum kernel does otherwise.
struct PrepassArgs {
    const SphereDev* mie_sph;
    const double* mie_kmag;
    int n_mie_sph, spb, wstride, quirks;
    double2* A;
    double2* B;
    const UfInst* uf;
    int n_uf, uf_blocks;
    const double* uf_tabs;
    const double* uf_waves;   // waves of the Monte-Carlo incident fields
    float2* Uf;
    const SumInst* sum;
    float* D;
    float* Di;
};

// LDS of the pre-pass launch's coefficient blocks, which every block of it
// reserves; the launch is taken only up to kPrepassLds.
constexpr size_t kPrepassLds = 16384;
size_t prepass_lds(const PrepassArgs& a);

// Algorithms of the FP32 series.
enum SeriesAlgo { kSeriesAuto = 0, kSeriesDirect = 1, kSeriesHarmonic = 2 };

// One launch of the harmonic pixel kernel over a bucket's work items.
struct HarmLaunch {
    dim3 grid;
    hipStream_t stream;
    const SeriesInst* inst;
    const SeriesSph* sph;
    const SeriesGroup* grp;
    const SeriesItem* items;
    int n_items;
    const float4* rec;      // records times B_l
    const float4* rec_in;   // records times A_l
    const float* cst;
    const int* perm;
    float2* U;
    int* counter;
    int planar;   // every sphere of the bucket's groups is planar
};

// Register budgets of the harmonic kernel: variant v of the bucket of L
// orders runs harm_variant_waves(L, v) waves per SIMD (0: no such variant).
// Variants 0-2 hold the order sums of all L orders at once (2L VGPRs), so
// the larger budgets spill some values to scratch; more waves in flight
// hide the latency of the FP64 radial chains and of the record loads
// better.  Variants 3-8 exist for planar buckets only: they sum the orders
// in chunks of harm_variant_chunk(L, v) (each chunk's rows of records, then
// its radial factors, the recurrences carried from chunk to chunk), which
// holds 2C sums instead of 2L and adds the phase steps of the rows below
// each chunk; 3-5 form the sums with VALU FMAs, 6-8 on the matrix cores
// (harm_variant_mfma).  All give the same bits as variants 0-2.  The plan's
// tuner times the variants on the batch's own inputs.
constexpr int kHarmVariants = 9;
// untuned choice: all order sums at the middle budget (general buckets), or
// chunks of 16 orders at 8 waves per SIMD (planar buckets)
constexpr int harm_default_variant(bool planar) { return planar ? 4 : 1; }
constexpr bool harm_variant_mfma(int v) { return v >= 6; }
constexpr int harm_variant_chunk(int L, int v) {
    return v < 3 ? 0 : v == 5 ? 8 : v == 6 ? 16 : v == 7 ? 32 : v == 8 ? L / 2 : (L <= 16 ? 8 : 16);
}
constexpr int harm_variant_waves(int L, int v) {
    if (v >= 6) {
        // the matrix-core chunks: 16, 32 or L/2 orders (2C AGPRs; L/2 = 8 on
        // L = 16 ties the VALU chunks there, at 7 waves), none that
        // repeats another variant's chunk (all orders at once ran slowest on
        // the buckets above 32, at two or three waves per SIMD; 32-order
        // chunks at four waves ran 1 % faster on L = 32 but spilled, 0.5 GB
        // more HBM writes per cfg4 pass), and no chunk over 40
        // orders (48 and 56 spilled their sums, and the compiler then read
        // them right after the products, inside the wait the MFMA needs).
        // The budgets are the largest wave counts that spill nothing (the
        // 16-order chunks spill two registers on L >= 48).
        if (v == 7 && L < 32) return 0;
        if (v == 8 && (L < 16 || L == 32 || L == 64 || L > 80)) return 0;
        const int c = harm_variant_chunk(L, v);
        return c <= 8 ? 7 : c <= 12 ? 6 : c <= 16 ? 5 : c <= 24 ? 4 : 3;
    }
    if (v >= 3) return v == 3 ? 7 : 8;
    return L <= 16 ? 6 + v
           : L <= 24 ? 5 + v
           : L <= 40 ? 4 + v
           : L <= 48 ? 3 + v
           : L <= 64 ? 2 + v
           : L <= 80 ? (v < 2 ? 2 + v : 0)
                     : (v < 2 ? 1 + v : 0);
}
// Whether the tuner times variant v on a bucket: planar buckets take the
// chunked variants only (the variants holding all order sums ran 1.3 to 4
// times slower on every planar bucket measured), general buckets variants
// 0-2 (which are all that exist for them).
constexpr bool harm_variant_tuned(int L, int v, bool planar) {
    return planar ? v >= 3 && v < kHarmVariants && harm_variant_waves(L, v) > 0
                  : v >= 0 && v < 3 && harm_variant_waves(L, v) > 0;
}

// whether variant v runs on a bucket (chunked variants: planar buckets)
constexpr bool harm_variant_exists(int L, int v, bool planar) {
    return v >= 0 && v < kHarmVariants && harm_variant_waves(L, v) > 0 && (v < 3 || planar);
}

// Launches bucket L in variant v (defined per bucket in us_harm.hip, in
// three objects: the chunked variants are built with machine LICM off, the
// others with it); returns 0, or -1 for a variant that does not exist.
template <int L>
int harm_launch_full(const HarmLaunch& a, int variant);
template <int L>
int harm_launch_chunked(const HarmLaunch& a, int variant);
template <int L>
int harm_launch_mfma(const HarmLaunch& a, int variant);
template <int L>
inline int harm_launch(const HarmLaunch& a, int variant) {
    return variant < 3   ? harm_launch_full<L>(a, variant)
           : variant < 6 ? harm_launch_chunked<L>(a, variant)
                         : harm_launch_mfma<L>(a, variant);
}

struct SeriesPlan {
    int n_inst = 0, n_sph = 0;
    bool harmonic = false;    // addition-theorem form
    int split_mult = 1;       // waves split into up to this many items per block
    int groups_per_cu = 8;    // resident workgroups of the persistent launch
    int nk_max = 0, max_npix = 0;
    long long field_elems = 0;   // elements of U the plan writes
    long long n_wave_rec = 0, n_part = 0;
    long long coef_elems = 0, rhat_elems = 0;   // per chunk (maximum)
    struct Bucket {
        int L, item_begin, item_count, counter;
        int planar;    // harmonic form: every sphere of its instances is planar
        int variant;   // harmonic form: register budget (harm_variant_waves)
    };
    struct Chunk {
        int sph_begin, sph_count;
        std::vector<Bucket> buckets;
    };
    std::vector<Chunk> chunks;
    int n_counters = 0;
    std::vector<int> split_insts;
    SeriesInst* d_inst = nullptr;
    SeriesSph* d_sph = nullptr;
    SeriesItem* d_items = nullptr;
    int* d_split_insts = nullptr;
    int* d_counters = nullptr;
    WaveF* d_wt = nullptr;
    float2* d_part = nullptr;
    float2* d_coef = nullptr;
    float4* d_rhat = nullptr;
    // harmonic form
    long long n_pl = 0, n_e = 0, n_rec = 0;
    int max_lh = 0, max_tl = 0;
    int max_rec_combos = 0;   // harmonic records: the largest row's (order, 4 spheres) count
    int max_rec_nlv = 0;      // and its largest order count and sphere count
    int max_rec_ns = 0;
    double* d_pl = nullptr;       // normalised Legendre values per (m, wave, l)
    double2* d_e = nullptr;       // exp(i m phi_k) per (m, wave)
    double2* d_phi = nullptr;     // wave weights (1/nk) subA exp(i k k.c) per sphere
    float4* d_rec = nullptr;      // harmonic records per (sphere, m, l), then times B_l
    float4* d_rec_in = nullptr;   // the same records times A_l
    float* d_cst = nullptr;       // scaled Legendre recurrence constants per (m, l)
    int* d_perm = nullptr;        // pixel orders of the instances with perm_off >= 0
    SeriesGroup* d_grp = nullptr; // sphere groups of the harmonic form
    double* d_h64 = nullptr;      // FP64 Legendre constants and scales of the wave sums
};

// Chooses splits and chunks, builds and uploads the tables and allocates
// the plan's buffers (coefficient storage at most `coef_budget` bytes per
// chunk unless one instance needs more).  Synchronises the stream.
// split_mult = 0 picks the split from the queue length; groups_per_cu sets
// the persistent grid.
int series_plan_create(bimsim_ctx* ctx, SeriesBatch& batch, size_t coef_budget, SeriesPlan* plan,
                       int split_mult = 0, int groups_per_cu = 8);

// As series_plan_create, but when the queue is short (a few items per CU)
// the split and grid are chosen by timing the candidates on these inputs
// (into a temporary field); the choice is kept in the context per batch
// shape.  A, B and waves must be ready on the stream.
int series_plan_create_tuned(bimsim_ctx* ctx, SeriesBatch& batch, size_t coef_budget,
                             const double* waves, const double2* A, const double2* B,
                             SeriesPlan* plan);
void series_plan_destroy(SeriesPlan* plan);

// Launches the plan on the context stream; allocation- and
// synchronisation-free, so it can be captured into a graph.  The two halves
// can run apart: the wave tables (harmonic sums or wave records, from the
// waves only) and the per-pixel sums (which read A and B).
int series_plan_launch(bimsim_ctx* ctx, const SeriesPlan& plan, const double* waves,
                       const double2* A, const double2* B, float2* U);
int series_plan_launch_waves(bimsim_ctx* ctx, const SeriesPlan& plan, const double* waves);
// Legendre tables, then one launch of the records of P beside the
// coefficients and incident fields of a (instead of
// series_plan_launch_waves, the Mie launch and the incident-field launch).
// P must be a harmonic plan and prepass_lds(a) <= kPrepassLds.
int series_plan_launch_prepass(bimsim_ctx* ctx, const SeriesPlan& P, const double* waves,
                               const PrepassArgs& a);
// after_fold: an event the sums (not the records' coefficient fold before
// them) wait for.
int series_plan_launch_sums(bimsim_ctx* ctx, const SeriesPlan& plan, const double2* A,
                            const double2* B, float2* U, hipEvent_t after_fold = nullptr);

// Order capacity for Nl + 1 orders, or -1.
int series_capacity(int nl1);

// Plan for `batch`, reused while the batch tables are unchanged (the key is
// the tables' bytes).  The context keeps the last few plans, and every plan
// used during a stream capture until it ends; a plan missing during a
// capture is an error (BIMSIM_E_UNSUPPORTED).
int series_cached_plan(bimsim_ctx* ctx, SeriesBatch& batch, const double* waves, const double2* A,
                       const double2* B, const SeriesPlan** out);
void series_cache_release(bimsim_ctx* ctx);

}  // namespace bimsim
