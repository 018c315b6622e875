// Context and error plumbing shared by the C-ABI entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "common.h"

struct bimsim_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // stage timers (nearfieldStruct::t_Uf / t_Us): event pairs recorded on
    // the stream without waiting, resolved by bimsim_get_timings
    bool timing = false;
    struct Span {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Span> spans;
    std::vector<hipEvent_t> event_pool;
    // grow-only device scratch, split into named slots; a slot's outgrown
    // buffers are kept until the context ends, since a captured batch graph
    // may still hold them
    std::map<int, std::pair<void*, size_t>> scratch;
    std::vector<void*> scratch_retired;
    // FFT twiddle tables per length: e^{-2 pi i t/N}, t = 0..N-1
    std::map<int, std::pair<double*, float*>> twiddles;
    // device copies of small host tables, keyed by content (e.g. the
    // detector's pass-band column list)
    std::map<std::vector<int>, int*> int_tables;
    std::map<std::vector<double>, double*> dbl_tables;
    int cu_count = 256;
    size_t lds_max = 65536;
    // last plan of the FP32 series (opaque, owned by us_series.hip)
    void* series_cache = nullptr;
    // series launch shapes chosen by timing, per batch shape
    bool autotune = true;
    int series_algo = 0;   // FP32 series: 0 auto, 1 direct, 2 harmonic
    std::map<std::string, std::pair<int, int>> tune;
    // harmonic series: register-budget variant per bucket, per batch shape
    std::map<std::string, std::vector<int>> tune_harm;
};

namespace bimsim {

enum ScratchSlot {
    kSlotCoef = 0,    // per-pixel series coefficients
    kSlotRhat = 1,    // per-pixel unit vectors
    kSlotWaves = 2,   // per-sphere wave tables
    kSlotSph = 3,     // device sphere records
    kSlotPartial = 4, // split partial sums
    kSlotFftA = 5,    // detector work copies
    kSlotFftB = 6,
    kSlotDet = 7,
    kSlotMisc = 8,
    kSlotUfLut = 9,   // incident-field Bessel table
    kSlotWide = 10,   // complex128 scattered field of FP32 calls beyond the FP32 orders
};

int set_error(bimsim_ctx* ctx, int code, const char* fmt, ...);

// Device scratch of at least `bytes` in `slot` (grown with hipMalloc; an
// outgrown buffer stays allocated until the context ends).
int scratch_get(bimsim_ctx* ctx, int slot, size_t bytes, void** out);

// Device copy of an int table (cached by content; uploaded once).
int int_table_get(bimsim_ctx* ctx, const std::vector<int>& v, const int** out);

// Device copy of a double table (cached by content; uploaded once).
int dbl_table_get(bimsim_ctx* ctx, const std::vector<double>& v, const double** out);

// Twiddle table for length n on the device (double and float copies).
int twiddles_get(bimsim_ctx* ctx, int n, const double** d, const float** f);

// Takes the device address of bimsim_detector_noalloc's work memory on
// `device` (current) once per device; called by bimsim_ctx_create.
int noalloc_scratch_init(int device);

// Stages of the timers.
enum TimerStage { kStageUf = 0, kStageUs = 1, kStageDet = 2, kStageSeries = 3, kNumStages = 4 };

// Timer helpers: timer_begin records a start event for `stage` on the stream
// (when timing is on) and returns a handle for timer_end; nothing waits.
int timer_begin(bimsim_ctx* ctx, int stage);
void timer_end(bimsim_ctx* ctx, int handle);

}  // namespace bimsim

#define BIMSIM_HIP_CHECK(ctx, expr)                                                        \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::bimsim::set_error((ctx), BIMSIM_E_HIP, "%s failed: %s (%s:%d)", #expr, \
                                       hipGetErrorString(e_), __FILE__, __LINE__);         \
    } while (0)

#define BIMSIM_CHECK_ARG(ctx, cond, msg)                                           \
    do {                                                                           \
        if (!(cond)) return ::bimsim::set_error((ctx), BIMSIM_E_INVALID, "%s", msg); \
    } while (0)

#define BIMSIM_TRY(expr)          \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != BIMSIM_OK) return rc_; \
    } while (0)
