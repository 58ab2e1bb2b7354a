// Internal declarations shared by the HIP translation units of libprnu_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>

#include "../../include/prnu_api.h"

namespace prnu {

// ----------------------------------------------------------------- errors
void set_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

#define PRNU_CHECK_HIP(call)                                                            \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::prnu::set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),    \
                              __FILE__, __LINE__);                                      \
            return PRNU_EDEVICE;                                                        \
        }                                                                               \
    } while (0)

#define PRNU_CHECK_LAUNCH() PRNU_CHECK_HIP(hipGetLastError())

#define PRNU_TRY(expr)                                                                  \
    do {                                                                                \
        int s_ = (expr);                                                                \
        if (s_ != PRNU_OK) return s_;                                                   \
    } while (0)

#define PRNU_REQUIRE(cond, ...)                                                         \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::prnu::set_error(__VA_ARGS__);                                             \
            return PRNU_EINVAL;                                                         \
        }                                                                               \
    } while (0)

// --------------------------------------------------------------- workspace
// Device scratch owned by a context; grows on demand, never shrinks.
struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
    int ensure(size_t need);
    void release();
};

// ----------------------------------------------------------------- context
struct FftPlan;  // defined in fft.hip

struct Context {
    int h = 0, w = 0, mode = PRNU_MODE_NCC;
    int device = 0;
    hipStream_t own_stream = nullptr;
    Workspace ws_filter;    // extraction intermediates
    Workspace ws_gram;      // Gram partial tiles
    Workspace ws_small;     // reductions, pointer tables, scalars
    Workspace ws_pce;       // PCE correlation planes
    Workspace ws_fp8;       // FP8 copy of a pattern set and its scales
    FftPlan *fft_w = nullptr;       // length-w plan
    FftPlan *fft_h = nullptr;       // length-h plan
    prnu_gram_config gram_cfg[4] = {{0, 1, 0}, {0, 1, 0}, {0, 1, 0}, {0, 1, 0}};  // per precision; slabs 0 = automatic
    bool timing = false;            // record events around the dominant kernel
    bool timed = false;             // ev_begin/ev_end hold a measurement
    hipEvent_t ev_begin = nullptr, ev_end = nullptr;
    int64_t pixels() const { return (int64_t)h * w; }
    int64_t spectrum_elems() const { return (int64_t)h * (w / 2 + 1); }
};

// Event bracket around a dominant kernel when the context's timing is on.
inline void time_begin(Context *c, hipStream_t st) {
    if (c->timing) (void)hipEventRecord(c->ev_begin, st);
}
inline void time_end(Context *c, hipStream_t st) {
    if (c->timing) {
        (void)hipEventRecord(c->ev_end, st);
        c->timed = true;
    }
}

inline hipStream_t pick_stream(Context *c, prnu_stream s) {
    return s ? (hipStream_t)s : c->own_stream;
}

// Finish a call: synchronous when the caller passed no stream.
inline int finish(Context *c, prnu_stream s) {
    if (!s) PRNU_CHECK_HIP(hipStreamSynchronize(c->own_stream));
    return PRNU_OK;
}

// ------------------------------------------------------ kernel launchers
// NCC (ncc.hip)
int launch_pattern_stats(const float *const *d_ptrs, int count, int64_t n, double *d_sum,
                         double *d_sumsq, float *d_absmax, Workspace &ws, hipStream_t st);
int launch_pair_ncc(const float *x, const float *y, int64_t n, double *d_out, Workspace &ws,
                    hipStream_t st);
int launch_gram_ncc(Context *c, const float *const *d_ptrs, int count, int precision,
                    const double *d_sum, const double *d_sumsq, const float *d_absmax,
                    double *d_ncc, hipStream_t st);
// FP8 (fp8.hip)
int64_t fp8_row_stride(int64_t n);
int launch_quantize_fp8(const float *const *d_ptrs, int count, int64_t n, const float *d_absmax, uint8_t *d_q,
                        float *d_scale, hipStream_t st);
int launch_fp8_ncc(Workspace &ws, int a_rows, bool tri, const uint8_t *A, int64_t lda, int ma, const float *scale_a,
                   const double *sum_a, const double *sumsq_a, const uint8_t *B, int64_t ldb, int nb,
                   const float *scale_b, const double *sum_b, const double *sumsq_b, int64_t K, double *out,
                   int64_t ld_out, int slabs, hipStream_t st, Context *timing_ctx);
int gram_resolved_config(Context *c, int count, int precision, prnu_gram_config *out);
int autotune_gram(Context *c, const float *const *d_ptrs, int count, int precision, int reps,
                  prnu_gram_config *best, float *best_ms, hipStream_t st);

}  // namespace prnu
