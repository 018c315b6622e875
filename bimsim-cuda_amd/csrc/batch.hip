// Batches of independent absorbance images: the microscopeStruct::Simulate
// sequence (reference microscope.cu:849-857, nearfield.cpp:194-218) for many
// instances at once, optionally captured into one HIP graph.
//
//   Mie coefficients   one launch, three waves per block of spheres
//   incident fields    one launch over (pixel blocks, instances); for FP32
//                      each on a side stream of its own beside the wave
//                      tables and records, joined before the scattered
//                      fields (beside those persistent launches the fields
//                      ran starved of CUs, and slowed them)
//   FP32 pre-pass      with one harmonic plan and few spheres instead: the
//                      Legendre tables, then coefficients, records and
//                      incident fields in one launch (series_plan_launch_prepass);
//                      with band-limited detectors the sum is then formed by
//                      their first row pass and the images cleared here
//   scattered fields   FP32: coefficient kernel + one persistent work-queue
//                      launch per order bucket (us_series.hip); FP64 forms:
//                      the per-instance entry point
//   Uf + Us            one launch over (pixel blocks, instances)
//   detector           per instance (its own FFT size)
#include <map>
#include <vector>

#include "batch.h"
#include "internal.h"
#include "mie_device.h"
#include "nf_us.h"
#include "us_series.h"

using namespace bimsim;

struct bimsim_batch {
    bimsim_ctx* ctx = nullptr;
    int prec = 0, quirks = 1, n = 0;
    bool f32 = false;
    std::vector<bimsim_instance> inst;
    std::vector<std::vector<bimsim_sphere>> inst_spheres;   // per instance, packed coeff_off
    std::vector<double> kmag, subA, du, dv;
    const double* waves = nullptr;
    void *D = nullptr, *Di = nullptr, *Abs = nullptr;
    size_t image_elems = 0;
    int n_sph = 0, lmax = 0, max_npix = 0;
    // the incident-field launch's items per instance: the most of any
    // instance's work items and the image pixels it clears
    int uf_items = 1;
    long long ncoef = 0, field_elems = 0;
    std::vector<long long> field_off;
    // device
    SphereDev* d_sph = nullptr;
    double* d_kmag = nullptr;
    void* d_work = nullptr;
    double2 *d_A = nullptr, *d_B = nullptr;
    UfInst* d_uf = nullptr;
    double* d_uftab = nullptr;      // incident-field tables of complex64 batches
    SumInst* d_sum = nullptr;
    double4* d_sph4 = nullptr;
    void *d_Uf = nullptr, *d_U = nullptr;
    double* d_kdirs = nullptr;      // plane-wave directions, one row per instance
    // complex64 batches: the instances whose incident field is the
    // Monte-Carlo sum, evaluated by the tiled kernel after the per-pixel one
    int* d_mc_list = nullptr;
    int n_mc_uf = 0, mc_tiles_max = 0;
    // complex64 batches: instances with orders beyond the FP32 series
    // kernels, whose scattered field the FP64 form computes
    std::vector<int> wide;
    SeriesPlan plan_mc, plan_pw;    // FP32 series: Monte-Carlo / plane-wave instances
    bool have_mc = false, have_pw = false;
    hipStream_t capture_stream = nullptr;
    // FP32 passes compute the Mie coefficients and the incident fields on
    // streams of their own, beside the wave tables and records (joined
    // before the sums)
    hipStream_t side = nullptr, side_f = nullptr;
    hipEvent_t ev_fork = nullptr, ev_coef = nullptr, ev_fields = nullptr;
    // FP32 passes of one harmonic plan whose coefficient blocks fit
    // kPrepassLds: coefficients, records and incident fields in one launch
    // on the one stream instead
    bool prepass = false;
    PrepassArgs pre{};
    // and with every detector on the band-limited path, the Uf + Us sum
    // formed by the detectors' first row pass (the images cleared in the
    // pre-pass launch)
    bool fuse_sum = false;
    std::vector<SumInst> sum_host;
    // with fuse_sum: each instance's inside-a-sphere pixel bits (the
    // geometry's, formed once), at inside_off[i] words
    unsigned* d_inside = nullptr;
    std::vector<long long> inside_off;
    // detectors: one batched pass per FFT size, and the instances that
    // need their own (supersampling, zero fill, sizes without a fixed plan)
    std::vector<DetBatch> det;
    std::vector<int> det_single;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

template <class T>
int dev_upload(bimsim_ctx* ctx, T** dst, const std::vector<T>& v) {
    if (v.empty()) return BIMSIM_OK;
    if (hipMalloc((void**)dst, sizeof(T) * v.size()) != hipSuccess)
        return set_error(ctx, BIMSIM_E_NOMEM, "out of device memory for batch tables");
    BIMSIM_HIP_CHECK(ctx, hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return BIMSIM_OK;
}

int dev_alloc(bimsim_ctx* ctx, void** dst, size_t bytes) {
    if (bytes == 0) return BIMSIM_OK;
    if (hipMalloc(dst, bytes) != hipSuccess)
        return set_error(ctx, BIMSIM_E_NOMEM, "out of device memory for batch buffers (%zu bytes)",
                         bytes);
    return BIMSIM_OK;
}

void release(bimsim_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->capture_stream) (void)hipStreamDestroy(b->capture_stream);
    if (b->side) (void)hipStreamDestroy(b->side);
    if (b->side_f) (void)hipStreamDestroy(b->side_f);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_coef) (void)hipEventDestroy(b->ev_coef);
    if (b->ev_fields) (void)hipEventDestroy(b->ev_fields);
    series_plan_destroy(&b->plan_mc);
    series_plan_destroy(&b->plan_pw);
    for (DetBatch& d : b->det) det_batch_destroy(&d);
    void* ptrs[] = {b->d_sph, b->d_kmag, b->d_work, b->d_A, b->d_B, b->d_uf, b->d_uftab,
                    b->d_sum, b->d_sph4, b->d_Uf, b->d_U, b->d_kdirs, b->d_mc_list,
                    b->d_inside};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

// FP64 batches: coefficients, incident fields, then each instance's
// scattered field through the per-instance entry point.
int enqueue_fp64_fields(bimsim_batch* b) {
    bimsim_ctx* ctx = b->ctx;
    const size_t fb = sizeof(double2);
    BIMSIM_TRY(mie_batch_launch(ctx, b->d_sph, b->d_kmag, b->n_sph, b->lmax, b->quirks, b->d_work,
                                b->d_A, b->d_B));
    BIMSIM_TRY(uf_batch_launch(ctx, b->d_uf, b->n, b->uf_items, b->f32, b->d_uftab, b->waves,
                               b->d_Uf));
    for (int i = 0; i < b->n; ++i) {
        const bimsim_instance& I = b->inst[i];
        if (I.ns == 0) continue;
        const bool plane = I.uf_mode == BIMSIM_UF_PLANE;
        const double* w = plane ? b->d_kdirs + 3 * (size_t)i : b->waves + 3 * (size_t)I.wave_begin;
        const int nk = plane ? 1 : I.nk;
        BIMSIM_TRY(bimsim_nf_us(ctx, &I.pos, I.N, I.N, I.focus, b->kmag[i],
                                b->inst_spheres[i].data(), I.ns, b->d_A, b->d_B, w, nk, b->subA[i],
                                b->prec, b->quirks, (char*)b->d_U + fb * b->field_off[i], 0));
    }
    return BIMSIM_OK;
}

// Complex64 batches: the scattered fields of the instances beyond the FP32
// series orders (the FP64 form, rounded).
int enqueue_wide(bimsim_batch* b) {
    bimsim_ctx* ctx = b->ctx;
    for (int i : b->wide) {
        const bimsim_instance& I = b->inst[i];
        const bool plane = I.uf_mode == BIMSIM_UF_PLANE;
        const double* w = plane ? b->d_kdirs + 3 * (size_t)i : b->waves + 3 * (size_t)I.wave_begin;
        BIMSIM_TRY(us_wide_fp32(ctx, &I.pos, I.N, I.N, I.focus, b->kmag[i],
                                b->inst_spheres[i].data(), I.ns, b->d_A, b->d_B, w,
                                plane ? 1 : I.nk, b->subA[i],
                                (float2*)b->d_U + b->field_off[i], 0));
    }
    return BIMSIM_OK;
}

// One pass of the batch on ctx->stream.
int enqueue(bimsim_batch* b) {
    bimsim_ctx* ctx = b->ctx;
    const size_t fb = b->f32 ? sizeof(float2) : sizeof(double2);
    const size_t rb = b->f32 ? sizeof(float) : sizeof(double);
    if (b->f32 && b->prepass) {
        const bool mc = b->have_mc;
        BIMSIM_TRY(series_plan_launch_prepass(ctx, mc ? b->plan_mc : b->plan_pw,
                                              mc ? b->waves : b->d_kdirs, b->pre));
        BIMSIM_TRY(uf_mc_batch_launch(ctx, b->d_uf, b->d_mc_list, b->n_mc_uf, b->mc_tiles_max,
                                      b->waves, (float2*)b->d_Uf));
        BIMSIM_TRY(series_plan_launch_sums(ctx, mc ? b->plan_mc : b->plan_pw, b->d_A, b->d_B,
                                           (float2*)b->d_U));
    } else if (b->f32) {
        // fork: coefficients on one side stream, incident fields on the
        // other, wave tables and records here
        BIMSIM_HIP_CHECK(ctx, hipEventRecord(b->ev_fork, ctx->stream));
        BIMSIM_HIP_CHECK(ctx, hipStreamWaitEvent(b->side, b->ev_fork, 0));
        BIMSIM_HIP_CHECK(ctx, hipStreamWaitEvent(b->side_f, b->ev_fork, 0));
        const hipStream_t main = ctx->stream;
        ctx->stream = b->side;
        int rc = mie_batch_launch(ctx, b->d_sph, b->d_kmag, b->n_sph, b->lmax, b->quirks,
                                  b->d_work, b->d_A, b->d_B);
        if (rc == BIMSIM_OK && hipEventRecord(b->ev_coef, b->side) != hipSuccess)
            rc = set_error(ctx, BIMSIM_E_HIP, "hipEventRecord failed");
        ctx->stream = b->side_f;
        // (with the sum formed by the detectors, this launch clears the images)
        if (rc == BIMSIM_OK)
            rc = uf_batch_launch(ctx, b->d_uf, b->n, b->uf_items, b->f32, b->d_uftab, b->waves,
                                 b->d_Uf, b->fuse_sum ? b->d_sum : nullptr, b->D, b->Di);
        if (rc == BIMSIM_OK)
            rc = uf_mc_batch_launch(ctx, b->d_uf, b->d_mc_list, b->n_mc_uf, b->mc_tiles_max,
                                    b->waves, (float2*)b->d_Uf);
        if (rc == BIMSIM_OK && hipEventRecord(b->ev_fields, b->side_f) != hipSuccess)
            rc = set_error(ctx, BIMSIM_E_HIP, "hipEventRecord failed");
        ctx->stream = main;
        BIMSIM_TRY(rc);
        if (b->have_mc) BIMSIM_TRY(series_plan_launch_waves(ctx, b->plan_mc, b->waves));
        if (b->have_pw) BIMSIM_TRY(series_plan_launch_waves(ctx, b->plan_pw, b->d_kdirs));
        // the incident fields join before the series sums (beside their
        // persistent launches they would wait for free slots to the end),
        // after the coefficient fold
        BIMSIM_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, b->ev_coef, 0));
        if (b->have_mc)
            BIMSIM_TRY(series_plan_launch_sums(ctx, b->plan_mc, b->d_A, b->d_B, (float2*)b->d_U,
                                               b->ev_fields));
        if (b->have_pw)
            BIMSIM_TRY(series_plan_launch_sums(ctx, b->plan_pw, b->d_A, b->d_B, (float2*)b->d_U,
                                               b->ev_fields));
        if (!b->have_mc && !b->have_pw)
            BIMSIM_HIP_CHECK(ctx, hipStreamWaitEvent(ctx->stream, b->ev_fields, 0));
    } else {
        BIMSIM_TRY(enqueue_fp64_fields(b));
    }
    if (b->f32) BIMSIM_TRY(enqueue_wide(b));
    if (!b->fuse_sum)
        BIMSIM_TRY(sum_batch_launch(ctx, b->d_sum, b->n, b->max_npix, b->d_sph4, b->f32, b->d_Uf,
                                    b->d_U, b->D, b->Di));
    for (const DetBatch& d : b->det)
        BIMSIM_TRY(det_batch_launch(ctx, d, b->d_U, b->d_Uf, b->D, b->Di, b->Abs));
    for (int i : b->det_single) {
        const bimsim_instance& I = b->inst[i];
        void* Ab = b->Abs ? (char*)b->Abs + rb * I.image_off : nullptr;
        if (I.zero_fill >= 2) {
            BIMSIM_TRY(detector_zero_filled_run(
                ctx, b->prec, b->quirks, (char*)b->d_U + fb * b->field_off[i],
                (char*)b->d_Uf + fb * b->field_off[i], I.N, I.N, I.zero_fill, b->du[i], b->dv[i],
                I.objective[0], I.objective[1], I.lambda, (char*)b->D + rb * I.image_off,
                (char*)b->Di + rb * I.image_off, Ab));
            continue;
        }
        SumFuse fz{};
        if (b->fuse_sum) {
            const SumInst& S = b->sum_host[i];
            fz = SumFuse{S.q, S.uR, S.vR, S.ns, b->d_sph4 + S.s_begin,
                         b->d_inside + b->inside_off[i]};
        }
        BIMSIM_TRY(detector_run(ctx, b->prec, b->quirks, (char*)b->d_U + fb * b->field_off[i],
                                (char*)b->d_Uf + fb * b->field_off[i], I.N, I.N, b->du[i], b->dv[i],
                                I.objective[0], I.objective[1], I.lambda, I.padding, I.Nd, I.ss,
                                (char*)b->D + rb * I.image_off, (char*)b->Di + rb * I.image_off, Ab,
                                0, b->fuse_sum ? &fz : nullptr));
    }
    return BIMSIM_OK;
}

double len3(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

int build(bimsim_batch* b, const bimsim_instance* inst, int n, const bimsim_sphere* spheres,
          int n_spheres) {
    bimsim_ctx* ctx = b->ctx;
    b->n = n;
    b->inst.assign(inst, inst + n);
    b->f32 = b->prec == BIMSIM_FP32_FAST;
    std::vector<SphereDev> sph_dev;
    std::vector<double> sph_k;
    std::vector<double4> sph4;
    std::vector<UfInst> uf(n);
    std::vector<double> uftab;
    std::vector<SumInst> sum(n);
    std::vector<double> kdirs(3 * (size_t)n);
    std::vector<int> mc_list;
    SeriesBatch mc, pw;
    b->inst_spheres.resize(n);
    b->kmag.resize(n);
    b->subA.resize(n);
    b->du.resize(n);
    b->dv.resize(n);
    b->field_off.resize(n);
    long long coef = 0, field = 0;
    for (int i = 0; i < n; ++i) {
        const bimsim_instance& I = inst[i];
        BIMSIM_CHECK_ARG(ctx, I.N > 0 && (long long)I.N * I.N <= (1LL << 28), "instance N out of range");
        BIMSIM_CHECK_ARG(ctx, I.lambda > 0.0 && std::isfinite(I.lambda), "instance lambda must be positive");
        BIMSIM_CHECK_ARG(ctx, I.m >= 1 && I.m <= 100000, "instance m out of range");
        BIMSIM_CHECK_ARG(ctx, I.ns >= 0 && I.sphere_begin >= 0 && I.sphere_begin + I.ns <= n_spheres,
                         "instance sphere range outside the sphere array");
        BIMSIM_CHECK_ARG(ctx, I.uf_mode == BIMSIM_UF_SERIES || I.uf_mode == BIMSIM_UF_PLANE ||
                                  I.uf_mode == BIMSIM_UF_MC,
                         "instance incident field must be the series, a plane wave or the "
                         "Monte-Carlo sum");
        const bool plane = I.uf_mode == BIMSIM_UF_PLANE;
        BIMSIM_CHECK_ARG(ctx, plane || (I.nk >= 1 && I.wave_begin >= 0), "instance needs waves");
        BIMSIM_CHECK_ARG(ctx, I.zero_fill <= 1 || (I.zero_fill <= 8 && I.Nd == I.N &&
                                                    I.padding == 0 && I.ss == 1),
                         "a zero-filled detector takes Nd = N, padding 0, ss 1 and a factor "
                         "of at most 8");
        BIMSIM_CHECK_ARG(ctx, I.Nd > 0 && I.ss > 0 && I.padding >= 0 &&
                                  (long long)(2 * I.padding + 1) * I.Nd * I.ss <= I.N && I.image_off >= 0,
                         "instance detector does not fit its near field");
        const double lx = len3(I.pos.X), ly = len3(I.pos.Y);
        BIMSIM_CHECK_ARG(ctx, lx > 0.0 && ly > 0.0, "instance slice has a zero edge");
        b->kmag[i] = 2.0 * (double)kPIf / I.lambda;
        const double ca = std::cos(std::asin(I.condenser[0])), cb = std::cos(std::asin(I.condenser[1]));
        b->subA[i] = plane ? I.amp : 2.0 * (double)kPIf * I.amp * ((1.0 - cb) - (1.0 - ca));
        b->du[i] = 1.0 / lx;
        b->dv[i] = 1.0 / ly;
        b->image_elems = std::max(b->image_elems, (size_t)(I.image_off + (long long)I.Nd * I.Nd));
        const int npix = I.N * I.N;
        b->max_npix = std::max(b->max_npix, npix);
        b->field_off[i] = field;
        const QuadArg q = to_quad(I.pos);
        long long tab = -1;
        if (b->f32 && I.uf_mode == BIMSIM_UF_SERIES) {
            tab = (long long)uftab.size();
            uf_fast_table(ca, cb, I.m, &uftab, uf_flat(q, mk(I.focus), mk(I.kdir)));
        }
        uf[i] = UfInst{q, mk(I.focus), mk(I.kdir), b->kmag[i], I.amp, ca, cb, I.m, I.N, I.N,
                       I.uf_mode, field, tab, uf_flat(q, mk(I.focus), mk(I.kdir)) ? 1 : 0,
                       plane ? 1 : I.nk, plane ? 0 : (long long)I.wave_begin,
                       plane ? I.amp : (double)(1.0f / (float)I.nk) * b->subA[i]};
        uf_mirror(uf[i]);
        if (b->f32 && I.uf_mode == BIMSIM_UF_MC) {
            mc_list.push_back(i);
            b->mc_tiles_max = std::max(b->mc_tiles_max, mc_tiles(I.N, I.N));
        }
        // images to clear each pass: those of the instances whose detector
        // adds to them (the batched detector writes its images whole)
        const int nd2 = det_batchable(ctx, b->prec, I.N, I.ss, I.zero_fill) ? 0 : I.Nd * I.Nd;
        sum[i] = SumInst{q, I.N, I.N, (int)sph_dev.size(), I.ns, field, I.image_off, nd2};
        for (int c = 0; c < 3; ++c) kdirs[3 * (size_t)i + c] = I.kdir[c];
        // spheres of this instance, coefficients packed
        SeriesBatch& sb = plane ? pw : mc;
        SeriesInst si{};
        si.q = q;
        si.f = mk(I.focus);
        si.kmag = b->kmag[i];
        const int nk = plane ? 1 : I.nk;
        si.sA = (double)(1.0f / (float)nk) * b->subA[i];
        si.uR = si.vR = I.N;
        si.s_begin = (int)sb.sph.size();
        si.ns = I.ns;
        si.nk = nk;
        si.accumulate = 0;
        si.wave_off = plane ? i : I.wave_begin;
        si.u_off = field;
        for (int s = 0; s < I.ns; ++s) {
            bimsim_sphere S = spheres[I.sphere_begin + s];
            BIMSIM_CHECK_ARG(ctx, S.a > 0.0, "sphere radius must be positive");
            if (S.Nl <= 0) S.Nl = bimsim_sphere_order(S.a, I.lambda);
            BIMSIM_CHECK_ARG(ctx, S.Nl >= 2 && S.Nl <= 4096, "sphere order Nl must be in [2, 4096]");
            S.coeff_off = (int)coef;
            coef += S.Nl + 1;
            b->lmax = std::max(b->lmax, S.Nl);
            b->inst_spheres[i].push_back(S);
            SphereDev d;
            for (int c = 0; c < 3; ++c) d.p[c] = S.p[c];
            d.a = S.a;
            d.n[0] = S.n[0];
            d.n[1] = S.n[1];
            d.Nl = S.Nl;
            d.coeff_off = S.coeff_off;
            sph_dev.push_back(d);
            sph_k.push_back(b->kmag[i]);
            sph4.push_back(make_double4(S.p[0], S.p[1], S.p[2], S.a));
            SeriesSph ss{};
            ss.s = d;
            ss.inst = (int)sb.inst.size();
            sb.sph.push_back(ss);
        }
        // complex64 batches: orders beyond the FP32 series go to the FP64 form
        int nl1 = 1;
        for (const bimsim_sphere& S : b->inst_spheres[i]) nl1 = std::max(nl1, S.Nl + 1);
        if (b->f32 && I.ns > 0 && series_capacity(nl1) < 0) {
            b->wide.push_back(i);
            sb.sph.resize(sb.sph.size() - I.ns);
        } else if (I.ns > 0) {
            sb.inst.push_back(si);
        }
        field += npix;
    }
    b->n_sph = (int)sph_dev.size();
    b->ncoef = coef;
    b->field_elems = field;
    const size_t fb = b->f32 ? sizeof(float2) : sizeof(double2);
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    BIMSIM_TRY(dev_upload(ctx, &b->d_sph, sph_dev));
    BIMSIM_TRY(dev_upload(ctx, &b->d_kmag, sph_k));
    BIMSIM_TRY(dev_upload(ctx, &b->d_uf, uf));
    BIMSIM_TRY(dev_upload(ctx, &b->d_uftab, uftab));
    for (int i = 0; i < n; ++i)
        b->uf_items = std::max(b->uf_items, std::max(uf_work_items(uf[i]), sum[i].nd2));
    BIMSIM_TRY(dev_upload(ctx, &b->d_sum, sum));
    BIMSIM_TRY(dev_upload(ctx, &b->d_sph4, sph4));
    BIMSIM_TRY(dev_upload(ctx, &b->d_kdirs, kdirs));
    BIMSIM_TRY(dev_upload(ctx, &b->d_mc_list, mc_list));
    b->n_mc_uf = (int)mc_list.size();
    BIMSIM_TRY(dev_alloc(ctx, &b->d_work, sizeof(double2) * 8 * (size_t)(b->lmax + 2) * std::max(1, b->n_sph)));
    BIMSIM_TRY(dev_alloc(ctx, (void**)&b->d_A, sizeof(double2) * (size_t)std::max<long long>(1, coef)));
    BIMSIM_TRY(dev_alloc(ctx, (void**)&b->d_B, sizeof(double2) * (size_t)std::max<long long>(1, coef)));
    BIMSIM_TRY(dev_alloc(ctx, &b->d_Uf, fb * (size_t)field));
    BIMSIM_TRY(dev_alloc(ctx, &b->d_U, fb * (size_t)field));
    if (b->f32) {
        if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&b->side_f, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_coef, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&b->ev_fields, hipEventDisableTiming) != hipSuccess)
            return set_error(ctx, BIMSIM_E_HIP, "cannot create the batch's side streams");
        // the series plans are timed on real inputs: coefficients first
        const size_t budget = (size_t)8 << 30;
        BIMSIM_TRY(mie_batch_launch(ctx, b->d_sph, b->d_kmag, b->n_sph, b->lmax, b->quirks,
                                    b->d_work, b->d_A, b->d_B));
        if (!mc.inst.empty()) {
            BIMSIM_TRY(series_plan_create_tuned(ctx, mc, budget, b->waves, b->d_A, b->d_B, &b->plan_mc));
            b->have_mc = true;
        }
        if (!pw.inst.empty()) {
            BIMSIM_TRY(series_plan_create_tuned(ctx, pw, budget, b->d_kdirs, b->d_A, b->d_B, &b->plan_pw));
            b->have_pw = true;
        }
        const int wstride = b->lmax + 2;
        b->pre = PrepassArgs{b->d_sph, b->d_kmag, b->n_sph, mie::mie_spheres_per_block(wstride),
                             wstride, b->quirks, b->d_A, b->d_B, b->d_uf, b->n,
                             (b->uf_items + 255) / 256, b->d_uftab, b->waves, (float2*)b->d_Uf};
        const SeriesPlan& P = b->have_mc ? b->plan_mc : b->plan_pw;
        b->prepass = b->have_mc != b->have_pw && P.harmonic && b->n_sph > 0 &&
                     prepass_lds(b->pre) <= kPrepassLds;
        // the Uf + Us sum in the detectors' first row pass: the images are
        // then cleared by the pre-pass launch or the incident-field launch
        b->fuse_sum = true;
        for (int i = 0; i < n; ++i)
            b->fuse_sum = b->fuse_sum && inst[i].zero_fill <= 1 &&
                          detector_fusable(ctx, inst[i].N, inst[i].N);
        if (b->fuse_sum) {
            b->sum_host = sum;
            b->pre.sum = b->d_sum;
            b->pre.D = (float*)b->D;
            b->pre.Di = (float*)b->Di;
            // the pixels inside a sphere, once: the detectors' first row
            // pass then reads a bit per pixel instead of testing the spheres
            long long words = 0;
            b->inside_off.resize(n);
            for (int i = 0; i < n; ++i) {
                b->inside_off[i] = words;
                words += ((long long)sum[i].uR * sum[i].vR + 31) / 32;
            }
            BIMSIM_TRY(dev_alloc(ctx, (void**)&b->d_inside, sizeof(unsigned) * (size_t)words));
            for (int i = 0; i < n; ++i) {
                const SumInst& S = sum[i];
                BIMSIM_TRY(sum_mask_launch(ctx, SumFuse{S.q, S.uR, S.vR, S.ns, b->d_sph4 + S.s_begin},
                                           b->d_inside + b->inside_off[i]));
            }
        }
    }
    // detectors by FFT size
    std::map<int, std::vector<DetParams>> by_size;
    for (int i = 0; i < n; ++i) {
        const bimsim_instance& I = inst[i];
        if (!det_batchable(ctx, b->prec, I.N, I.ss, I.zero_fill)) {
            b->det_single.push_back(i);
            continue;
        }
        DetParams q{};
        q.uoff = b->field_off[i];
        q.image_off = I.image_off;
        q.padding = I.padding;
        q.Nd = I.Nd;
        q.du = b->du[i];
        q.dv = b->dv[i];
        q.NAin = I.objective[0];
        q.NAout = I.objective[1];
        q.lambda = I.lambda;
        if (b->fuse_sum) {
            const SumInst& S = sum[i];
            q.fz = SumFuse{S.q, S.uR, S.vR, S.ns, b->d_sph4 + S.s_begin,
                           b->d_inside + b->inside_off[i]};
        }
        by_size[I.N].push_back(q);
    }
    for (auto& kv : by_size) {
        b->det.emplace_back();
        BIMSIM_TRY(det_batch_create(ctx, b->prec, b->quirks, kv.first, kv.second, b->fuse_sum,
                                    &b->det.back()));
    }
    return BIMSIM_OK;
}

}  // namespace

extern "C" {

int bimsim_batch_create(bimsim_ctx* ctx, const bimsim_instance* inst, int n,
                        const bimsim_sphere* spheres, int n_spheres, const double* waves, int prec,
                        int quirks, void* D, void* Di, void* Abs, int use_graph, bimsim_batch** out) {
    if (!ctx || !out) return BIMSIM_E_INVALID;
    *out = nullptr;
    BIMSIM_CHECK_ARG(ctx, inst && n >= 1 && D && Di && n_spheres >= 0 && (n_spheres == 0 || spheres),
                     "bad batch arguments");
    BIMSIM_CHECK_ARG(ctx, prec >= BIMSIM_FP64_PARITY && prec <= BIMSIM_FP32_FAST, "unknown precision");
    if (!quirks)
        return set_error(ctx, BIMSIM_E_UNSUPPORTED, "batches run with the reference recurrences only");
    bool need_waves = false;
    for (int i = 0; i < n; ++i)
        need_waves |= inst[i].uf_mode == BIMSIM_UF_MC ||
                      (inst[i].uf_mode != BIMSIM_UF_PLANE && inst[i].ns > 0);
    BIMSIM_CHECK_ARG(ctx, waves || !need_waves, "waves must be given");
    bimsim_batch* b = new bimsim_batch();
    b->ctx = ctx;
    b->prec = prec;
    b->quirks = quirks;
    b->waves = waves;
    b->D = D;
    b->Di = Di;
    b->Abs = Abs;
    int rc = build(b, inst, n, spheres, n_spheres);
    // one eager pass: results, and every scratch buffer and table the pass
    // needs is in place before a capture
    if (rc == BIMSIM_OK) rc = enqueue(b);
    if (rc == BIMSIM_OK && hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = set_error(ctx, BIMSIM_E_HIP, "batch pass failed");
    if (rc == BIMSIM_OK && use_graph) {
        // capture on a stream of our own (the default stream cannot be captured)
        const hipStream_t saved = ctx->stream;
        const bool timing = ctx->timing;
        if (hipStreamCreateWithFlags(&b->capture_stream, hipStreamNonBlocking) != hipSuccess) {
            rc = set_error(ctx, BIMSIM_E_HIP, "cannot create the capture stream");
        } else {
            ctx->stream = b->capture_stream;
            ctx->timing = false;
            if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                rc = set_error(ctx, BIMSIM_E_HIP, "hipStreamBeginCapture failed");
            } else {
                rc = enqueue(b);
                hipGraph_t g = nullptr;
                const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
                if (rc == BIMSIM_OK && e != hipSuccess)
                    rc = set_error(ctx, BIMSIM_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
                b->graph = g;
                if (rc == BIMSIM_OK && hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess)
                    rc = set_error(ctx, BIMSIM_E_HIP, "hipGraphInstantiate failed");
            }
            ctx->stream = saved;
            ctx->timing = timing;
        }
    }
    if (rc != BIMSIM_OK) {
        (void)hipDeviceSynchronize();
        release(b);
        delete b;
        return rc;
    }
    *out = b;
    return BIMSIM_OK;
}

int bimsim_batch_launch(bimsim_batch* b) {
    if (!b) return BIMSIM_E_INVALID;
    bimsim_ctx* ctx = b->ctx;
    BIMSIM_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (b->exec) {
        BIMSIM_HIP_CHECK(ctx, hipGraphLaunch(b->exec, ctx->stream));
        return BIMSIM_OK;
    }
    return enqueue(b);
}

int bimsim_batch_destroy(bimsim_batch* b) {
    if (!b) return BIMSIM_E_INVALID;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    release(b);
    delete b;
    return BIMSIM_OK;
}

int bimsim_batch_run(bimsim_ctx* ctx, const bimsim_instance* inst, int n,
                     const bimsim_sphere* spheres, int n_spheres, const double* waves, int prec,
                     int quirks, void* D, void* Di, void* Abs) {
    bimsim_batch* b = nullptr;
    BIMSIM_TRY(bimsim_batch_create(ctx, inst, n, spheres, n_spheres, waves, prec, quirks, D, Di, Abs,
                                   0, &b));
    return bimsim_batch_destroy(b);
}

}  // extern "C"
