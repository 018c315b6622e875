// Context lifecycle, errors, scratch, timers, and the host-side helpers of the
// C ABI (sphere order, Monte-Carlo directions).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "internal.h"
#include "us_series.h"

namespace bimsim {

int set_error(bimsim_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

int scratch_get(bimsim_ctx* ctx, int slot, size_t bytes, void** out) {
    auto& s = ctx->scratch[slot];
    if (s.second < bytes) {
        if (s.first) {
            // kernels in flight and captured batch graphs may still use the
            // old buffer: it is freed with the context
            ctx->scratch_retired.push_back(s.first);
            s.first = nullptr;
            s.second = 0;
        }
        size_t want = bytes + bytes / 8 + 256;
        if (hipMalloc(&s.first, want) != hipSuccess) {
            s.first = nullptr;
            (void)hipGetLastError();
            return set_error(ctx, BIMSIM_E_NOMEM, "device scratch of %zu bytes unavailable", want);
        }
        s.second = want;
    }
    *out = s.first;
    return BIMSIM_OK;
}

namespace {

template <class T>
int table_get(bimsim_ctx* ctx, std::map<std::vector<T>, T*>& tables, const std::vector<T>& v,
              const T** out) {
    auto it = tables.find(v);
    if (it == tables.end()) {
        T* d = nullptr;
        if (hipMalloc(&d, sizeof(T) * std::max<size_t>(v.size(), 1)) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(ctx, BIMSIM_E_NOMEM, "table allocation failed");
        }
        if (!v.empty()) {
            BIMSIM_HIP_CHECK(ctx, hipMemcpyAsync(d, v.data(), sizeof(T) * v.size(),
                                                 hipMemcpyHostToDevice, ctx->stream));
            BIMSIM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        }
        it = tables.emplace(v, d).first;
    }
    *out = it->second;
    return BIMSIM_OK;
}

}  // namespace

int int_table_get(bimsim_ctx* ctx, const std::vector<int>& v, const int** out) {
    return table_get(ctx, ctx->int_tables, v, out);
}

int dbl_table_get(bimsim_ctx* ctx, const std::vector<double>& v, const double** out) {
    return table_get(ctx, ctx->dbl_tables, v, out);
}

int twiddles_get(bimsim_ctx* ctx, int n, const double** d, const float** f) {
    auto it = ctx->twiddles.find(n);
    if (it == ctx->twiddles.end()) {
        std::vector<double> hd(2 * (size_t)n);
        std::vector<float> hf(2 * (size_t)n);
        for (int t = 0; t < n; ++t) {
            // cos/sin of the angle of the exact integer fraction t/n, in double
            const double ang = -2.0 * M_PI * (double)t / (double)n;
            hd[2 * t] = std::cos(ang);
            hd[2 * t + 1] = std::sin(ang);
            hf[2 * t] = (float)hd[2 * t];
            hf[2 * t + 1] = (float)hd[2 * t + 1];
        }
        double* dd = nullptr;
        float* df = nullptr;
        if (hipMalloc(&dd, sizeof(double) * 2 * n) != hipSuccess ||
            hipMalloc(&df, sizeof(float) * 2 * n) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(ctx, BIMSIM_E_NOMEM, "twiddle table allocation failed");
        }
        BIMSIM_HIP_CHECK(ctx, hipMemcpyAsync(dd, hd.data(), sizeof(double) * 2 * n,
                                             hipMemcpyHostToDevice, ctx->stream));
        BIMSIM_HIP_CHECK(ctx, hipMemcpyAsync(df, hf.data(), sizeof(float) * 2 * n,
                                             hipMemcpyHostToDevice, ctx->stream));
        BIMSIM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        it = ctx->twiddles.emplace(n, std::make_pair(dd, df)).first;
    }
    *d = it->second.first;
    *f = it->second.second;
    return BIMSIM_OK;
}

static hipEvent_t pooled_event(bimsim_ctx* ctx) {
    if (!ctx->event_pool.empty()) {
        hipEvent_t e = ctx->event_pool.back();
        ctx->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

int timer_begin(bimsim_ctx* ctx, int stage) {
    if (!ctx->timing) return -1;
    bimsim_ctx::Span sp{stage, pooled_event(ctx), pooled_event(ctx)};
    if (!sp.a || !sp.b) return -1;
    (void)hipEventRecord(sp.a, ctx->stream);
    ctx->spans.push_back(sp);
    return (int)ctx->spans.size() - 1;
}

void timer_end(bimsim_ctx* ctx, int handle) {
    if (handle < 0 || handle >= (int)ctx->spans.size()) return;
    (void)hipEventRecord(ctx->spans[handle].b, ctx->stream);
}

}  // namespace bimsim

using namespace bimsim;

extern "C" {

int bimsim_abi_version(void) { return BIMSIM_ABI_VERSION; }

int bimsim_ctx_create(int device, void* hip_stream, bimsim_ctx** out) {
    if (!out) return BIMSIM_E_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        return BIMSIM_E_NODEVICE;
    }
    if (device < 0 || device >= ndev) return BIMSIM_E_INVALID;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return BIMSIM_E_HIP;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return BIMSIM_E_NODEVICE;
    if (hipSetDevice(device) != hipSuccess) return BIMSIM_E_HIP;
    // the allocation-free detector's work memory on this device (a failure
    // only disables that entry point, which then reports it)
    (void)noalloc_scratch_init(device);
    bimsim_ctx* ctx = new bimsim_ctx();
    ctx->device = device;
    ctx->cu_count = prop.multiProcessorCount;
    ctx->lds_max = prop.sharedMemPerBlock;
    // NULL selects the device's default stream: the stream the reference's
    // kernels run on and the one torch's default stream work is ordered with.
    ctx->stream = (hipStream_t)hip_stream;
    *out = ctx;
    return BIMSIM_OK;
}

int bimsim_ctx_destroy(bimsim_ctx* ctx) {
    if (!ctx) return BIMSIM_E_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    series_cache_release(ctx);
    for (auto& kv : ctx->scratch)
        if (kv.second.first) (void)hipFree(kv.second.first);
    for (void* p : ctx->scratch_retired) (void)hipFree(p);
    for (auto& kv : ctx->int_tables) (void)hipFree(kv.second);
    for (auto& kv : ctx->dbl_tables) (void)hipFree(kv.second);
    for (auto& kv : ctx->twiddles) {
        (void)hipFree(kv.second.first);
        (void)hipFree(kv.second.second);
    }
    for (auto& sp : ctx->spans) {
        (void)hipEventDestroy(sp.a);
        (void)hipEventDestroy(sp.b);
    }
    for (hipEvent_t e : ctx->event_pool) (void)hipEventDestroy(e);
    delete ctx;
    return BIMSIM_OK;
}

int bimsim_sync(bimsim_ctx* ctx) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return BIMSIM_OK;
}

const char* bimsim_last_error(const bimsim_ctx* ctx) {
    return ctx ? ctx->err.c_str() : "null context";
}

int bimsim_set_series_algorithm(bimsim_ctx* ctx, int algo) {
    if (!ctx) return BIMSIM_E_INVALID;
    BIMSIM_CHECK_ARG(ctx, algo >= BIMSIM_SERIES_AUTO && algo <= BIMSIM_SERIES_HARMONIC,
                     "unknown series algorithm");
    ctx->series_algo = algo;
    return BIMSIM_OK;
}

int bimsim_set_autotune(bimsim_ctx* ctx, int enable) {
    if (!ctx) return BIMSIM_E_INVALID;
    ctx->autotune = enable != 0;
    return BIMSIM_OK;
}

int bimsim_set_timing(bimsim_ctx* ctx, int enable) {
    if (!ctx) return BIMSIM_E_INVALID;
    ctx->timing = enable != 0;
    return BIMSIM_OK;
}

int bimsim_get_timings(bimsim_ctx* ctx, float* ms, int* counts) {
    if (!ctx || !ms) return BIMSIM_E_INVALID;
    for (int k = 0; k < kNumStages; ++k) {
        ms[k] = 0.f;
        if (counts) counts[k] = 0;
    }
    int rc = BIMSIM_OK;
    for (auto& sp : ctx->spans) {
        float t = 0.f;
        if (hipEventSynchronize(sp.b) != hipSuccess || hipEventElapsedTime(&t, sp.a, sp.b) != hipSuccess) {
            rc = set_error(ctx, BIMSIM_E_HIP, "stage timer events failed");
            t = 0.f;
        }
        if (sp.stage >= 0 && sp.stage < kNumStages) {
            ms[sp.stage] += t;
            if (counts) counts[sp.stage] += 1;
        }
        ctx->event_pool.push_back(sp.a);
        ctx->event_pool.push_back(sp.b);
    }
    ctx->spans.clear();
    return rc;
}

// sphere::calcNl, sphere.h:71-74 (float arithmetic, PI = 3.14159f)
int bimsim_sphere_order(double a, double lambda) {
    const float x = (2 * kPIf * (float)a) / (float)lambda;
    return (int)std::ceil(x + 4 * std::pow(x, (float)(1.0 / 3.0)) + 2);
}

}  // extern "C"
