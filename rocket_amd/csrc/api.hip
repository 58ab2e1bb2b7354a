// C ABI of libprnu_hip.so: error handling, memory/stream layer, context
// lifecycle and dispatch to the kernel launchers.
#include "prnu_internal.hpp"

#include <cstdio>
#include <cstring>
#include <new>

namespace prnu {

static thread_local char g_error[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

void clear_error() { g_error[0] = '\0'; }

int Workspace::ensure(size_t need) {
    if (need <= bytes) return PRNU_OK;
    if (ptr) {
        // the old buffer may still be read by queued work on any stream
        PRNU_CHECK_HIP(hipDeviceSynchronize());
        PRNU_CHECK_HIP(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
    }
    size_t alloc = (need + 4095) & ~size_t(4095);
    hipError_t e = hipMalloc(&ptr, alloc);
    if (e != hipSuccess) {
        ptr = nullptr;
        (void)hipGetLastError();
        set_error("device workspace of %zu bytes: %s", alloc, hipGetErrorString(e));
        return PRNU_ENOMEM;
    }
    bytes = alloc;
    return PRNU_OK;
}

void Workspace::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// FFT plans are created and destroyed in fft.hip.
void fft_plan_destroy(FftPlan *p);

}  // namespace prnu

using namespace prnu;

extern "C" {

const char *prnu_last_error(void) { return g_error; }
int prnu_api_version(void) { return PRNU_API_VERSION; }

// ------------------------------------------------------------------ memory
int prnu_device_count(int *count) {
    PRNU_REQUIRE(count, "device_count: null output");
    PRNU_CHECK_HIP(hipGetDeviceCount(count));
    return PRNU_OK;
}

int prnu_set_device(int device) {
    PRNU_CHECK_HIP(hipSetDevice(device));
    return PRNU_OK;
}

int prnu_mem_info(size_t *free_bytes, size_t *total_bytes) {
    PRNU_REQUIRE(free_bytes && total_bytes, "mem_info: null output");
    PRNU_CHECK_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return PRNU_OK;
}

int prnu_malloc_device(void **ptr, size_t bytes) {
    PRNU_REQUIRE(ptr, "malloc_device: null output");
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();
        set_error("hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return PRNU_ENOMEM;
    }
    return PRNU_OK;
}

int prnu_free_device(void *ptr) {
    if (ptr) PRNU_CHECK_HIP(hipFree(ptr));
    return PRNU_OK;
}

int prnu_malloc_pinned(void **ptr, size_t bytes) {
    PRNU_REQUIRE(ptr, "malloc_pinned: null output");
    hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        (void)hipGetLastError();
        set_error("hipHostMalloc(%zu): %s", bytes, hipGetErrorString(e));
        return PRNU_ENOMEM;
    }
    return PRNU_OK;
}

int prnu_free_pinned(void *ptr) {
    if (ptr) PRNU_CHECK_HIP(hipHostFree(ptr));
    return PRNU_OK;
}

static int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, prnu_stream stream) {
    if (bytes == 0) return PRNU_OK;
    PRNU_REQUIRE(dst && src, "memcpy: null pointer");
    if (stream) {
        PRNU_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, kind, (hipStream_t)stream));
    } else {
        PRNU_CHECK_HIP(hipMemcpy(dst, src, bytes, kind));
    }
    return PRNU_OK;
}

int prnu_memcpy_h2d(void *dst, const void *src, size_t bytes, prnu_stream stream) {
    return copy(dst, src, bytes, hipMemcpyHostToDevice, stream);
}
int prnu_memcpy_d2h(void *dst, const void *src, size_t bytes, prnu_stream stream) {
    return copy(dst, src, bytes, hipMemcpyDeviceToHost, stream);
}
int prnu_memcpy_d2d(void *dst, const void *src, size_t bytes, prnu_stream stream) {
    return copy(dst, src, bytes, hipMemcpyDeviceToDevice, stream);
}

int prnu_stream_create(prnu_stream *stream) {
    PRNU_REQUIRE(stream, "stream_create: null output");
    hipStream_t s = nullptr;
    PRNU_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (prnu_stream)s;
    return PRNU_OK;
}

int prnu_stream_destroy(prnu_stream stream) {
    if (stream) PRNU_CHECK_HIP(hipStreamDestroy((hipStream_t)stream));
    return PRNU_OK;
}

int prnu_stream_synchronize(prnu_stream stream) {
    PRNU_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
    return PRNU_OK;
}

// ----------------------------------------------------------------- context
int prnu_ctx_create(int height, int width, int mode, prnu_ctx **out) {
    PRNU_REQUIRE(out, "ctx_create: null output");
    *out = nullptr;
    PRNU_REQUIRE(height >= 2 && width >= 2, "ctx_create: image %dx%d must be at least 2x2", height, width);
    PRNU_REQUIRE((int64_t)height * width <= (int64_t(1) << 31) - 1, "ctx_create: image %dx%d too large",
                 height, width);
    PRNU_REQUIRE(mode == PRNU_MODE_NCC || mode == PRNU_MODE_PCE, "ctx_create: unknown mode %d", mode);
    Context *c = new (std::nothrow) Context();
    if (!c) {
        set_error("ctx_create: out of host memory");
        return PRNU_ENOMEM;
    }
    c->h = height;
    c->w = width;
    c->mode = mode;
    hipError_t e = hipGetDevice(&c->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_error("ctx_create: %s", hipGetErrorString(e));
        delete c;
        return PRNU_EDEVICE;
    }
    *out = reinterpret_cast<prnu_ctx *>(c);
    return PRNU_OK;
}

int prnu_ctx_destroy(prnu_ctx *ctx) {
    if (!ctx) return PRNU_OK;
    Context *c = reinterpret_cast<Context *>(ctx);
    if (c->own_stream) (void)hipStreamSynchronize(c->own_stream);
    (void)hipDeviceSynchronize();
    c->ws_filter.release();
    c->ws_gram.release();
    c->ws_small.release();
    c->ws_pce.release();
    c->ws_fp8.release();
    if (c->fft_w) fft_plan_destroy(c->fft_w);
    if (c->fft_h && c->fft_h != c->fft_w) fft_plan_destroy(c->fft_h);
    if (c->ev_begin) (void)hipEventDestroy(c->ev_begin);
    if (c->ev_end) (void)hipEventDestroy(c->ev_end);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return PRNU_OK;
}

int prnu_ctx_set_timing(prnu_ctx *ctx, int enable) {
    PRNU_REQUIRE(ctx, "ctx_set_timing: null context");
    Context *c = reinterpret_cast<Context *>(ctx);
    if (enable && !c->ev_begin) {
        PRNU_CHECK_HIP(hipEventCreate(&c->ev_begin));
        PRNU_CHECK_HIP(hipEventCreate(&c->ev_end));
    }
    c->timing = enable != 0;
    c->timed = false;
    return PRNU_OK;
}

int prnu_ctx_last_kernel_ms(prnu_ctx *ctx, float *ms) {
    PRNU_REQUIRE(ctx && ms, "ctx_last_kernel_ms: null pointer");
    Context *c = reinterpret_cast<Context *>(ctx);
    PRNU_REQUIRE(c->timing && c->timed, "ctx_last_kernel_ms: timing is off or nothing was timed yet");
    PRNU_CHECK_HIP(hipEventSynchronize(c->ev_end));
    PRNU_CHECK_HIP(hipEventElapsedTime(ms, c->ev_begin, c->ev_end));
    return PRNU_OK;
}

int64_t prnu_max_file_size(const prnu_ctx *ctx) {
    const Context *c = reinterpret_cast<const Context *>(ctx);
    return c ? 3 * c->pixels() : -1;
}

int64_t prnu_max_input_size(const prnu_ctx *ctx) {
    const Context *c = reinterpret_cast<const Context *>(ctx);
    if (!c) return -1;
    return c->mode == PRNU_MODE_PCE ? 8 * c->spectrum_elems() : 4 * c->pixels();
}

int64_t prnu_max_output_size(const prnu_ctx *ctx) { return ctx ? 8 : -1; }

// ------------------------------------------------------------- batched NCC
int prnu_pattern_stats(const float *const *d_patterns, int count, int64_t n, double *d_sum, double *d_sumsq,
                       float *d_absmax, prnu_stream stream) {
    PRNU_REQUIRE(d_patterns && d_sum && d_sumsq, "pattern_stats: null pointer");
    // stateless entry point: a per-thread workspace keeps it allocation-free after the first call
    static thread_local Workspace ws;
    hipStream_t st = (hipStream_t)stream;
    PRNU_TRY(launch_pattern_stats(d_patterns, count, n, d_sum, d_sumsq, d_absmax, ws, st));
    if (!stream) PRNU_CHECK_HIP(hipStreamSynchronize(st));
    return PRNU_OK;
}

int prnu_gram_ncc(prnu_ctx *ctx, const float *const *d_patterns, int count, int precision, const double *d_sum,
                  const double *d_sumsq, const float *d_absmax, double *d_ncc, prnu_stream stream) {
    PRNU_REQUIRE(ctx && d_patterns && d_sum && d_sumsq && d_ncc, "gram_ncc: null pointer");
    Context *c = reinterpret_cast<Context *>(ctx);
    PRNU_REQUIRE(c->mode == PRNU_MODE_NCC, "gram_ncc: context is not in NCC mode");
    PRNU_REQUIRE(precision != PRNU_PREC_FP8 || d_absmax, "gram_ncc: FP8 needs d_absmax");
    hipStream_t st = pick_stream(c, stream);
    PRNU_TRY(launch_gram_ncc(c, d_patterns, count, precision, d_sum, d_sumsq, d_absmax, d_ncc, st));
    return finish(c, stream);
}

int prnu_gram_set_config(prnu_ctx *ctx, int precision, const prnu_gram_config *cfg) {
    PRNU_REQUIRE(ctx, "gram_set_config: null context");
    PRNU_REQUIRE(precision >= 0 && precision <= 3, "gram_set_config: unknown precision %d", precision);
    Context *c = reinterpret_cast<Context *>(ctx);
    if (!cfg) {
        c->gram_cfg[precision] = prnu_gram_config{0, 1, 0};
        return PRNU_OK;
    }
    PRNU_REQUIRE(cfg->slabs >= 0 && (cfg->xcd_remap == 0 || cfg->xcd_remap == 1) && cfg->variant >= 0 &&
                     cfg->variant <= 3,
                 "gram_set_config: slabs must be >= 0, xcd_remap 0 or 1, variant 0 to 3");
    c->gram_cfg[precision] = *cfg;
    return PRNU_OK;
}

int prnu_gram_get_config(prnu_ctx *ctx, int precision, int count, prnu_gram_config *out) {
    PRNU_REQUIRE(ctx && out, "gram_get_config: null pointer");
    PRNU_REQUIRE(precision >= 0 && precision <= 2 && count >= 1, "gram_get_config: bad precision or count");
    return gram_resolved_config(reinterpret_cast<Context *>(ctx), count, precision, out);
}

int prnu_gram_autotune(prnu_ctx *ctx, const float *const *d_patterns, int count, int precision, int reps,
                       prnu_gram_config *best, float *best_ms, prnu_stream stream) {
    PRNU_REQUIRE(ctx && d_patterns, "gram_autotune: null pointer");
    Context *c = reinterpret_cast<Context *>(ctx);
    PRNU_REQUIRE(c->mode == PRNU_MODE_NCC, "gram_autotune: context is not in NCC mode");
    return autotune_gram(c, d_patterns, count, precision, reps, best, best_ms, pick_stream(c, stream));
}

}  // extern "C"
