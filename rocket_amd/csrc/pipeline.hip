// Per-item (preprocess), per-pair (correlate) and single-filter entry points.
#include "prnu_internal.hpp"

namespace prnu {
int run_preprocess(Context *c, int fmt, const void *in, size_t in_bytes, void *out, size_t out_cap,
                   size_t *out_bytes, hipStream_t st);
int run_preprocess_batch(Context *c, int fmt, const void *in, size_t in_bytes, int64_t in_stride, int count,
                         void *out, size_t out_cap, int64_t out_stride, size_t *out_bytes, hipStream_t st);
int run_filter_grayscale(Context *c, int fmt, const void *in, size_t in_bytes, float *out, hipStream_t st);
int run_filter_fastnoise(Context *c, float *img, hipStream_t st);
int run_filter_zeromean(Context *c, float *img, hipStream_t st);
int run_filter_wiener(Context *c, float *img, hipStream_t st);
int run_filter_spectral(Context *c, const float *img, float *out, size_t out_cap, hipStream_t st);
int run_pce(Context *c, int variant, const void *x, const void *y, double *out, hipStream_t st);
int run_pce_pairs(Context *c, int variant, const void *const *d_tab, const int *d_pairs, int npairs, double *out,
                  int batch, hipStream_t st);
}  // namespace prnu

using namespace prnu;

#define PRNU_CTX(ctx)                                                     \
    PRNU_REQUIRE(ctx, "null context");                                    \
    Context *c = reinterpret_cast<Context *>(ctx);                        \
    hipStream_t st = pick_stream(c, stream)

extern "C" {

int prnu_preprocess(prnu_ctx *ctx, int input_format, const void *d_in, size_t in_bytes, void *d_out,
                    size_t out_capacity, size_t *out_bytes, prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(d_in && d_out, "preprocess: null buffer");
    PRNU_TRY(run_preprocess(c, input_format, d_in, in_bytes, d_out, out_capacity, out_bytes, st));
    return finish(c, stream);
}

int prnu_preprocess_batch(prnu_ctx *ctx, int input_format, const void *d_in, size_t in_bytes, int64_t in_stride,
                          int count, void *d_out, size_t out_capacity, int64_t out_stride, prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(d_in && d_out, "preprocess_batch: null buffer");
    if (count == 0) return PRNU_OK;
    PRNU_REQUIRE(count > 0, "preprocess_batch: negative count %d", count);
    PRNU_TRY(run_preprocess_batch(c, input_format, d_in, in_bytes, in_stride, count, d_out, out_capacity,
                                  out_stride, nullptr, st));
    return finish(c, stream);
}

int prnu_correlate(prnu_ctx *ctx, const void *d_left, size_t left_bytes, const void *d_right,
                   size_t right_bytes, void *d_out, size_t out_capacity, size_t *out_bytes,
                   prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(d_left && d_right && d_out, "correlate: null buffer");
    const size_t need = (size_t)prnu_max_input_size(ctx);
    // NormalizedCrossCorrelation.java:83-87, PeakToCorrelationEnergy.java:203-207
    PRNU_REQUIRE(left_bytes >= need && right_bytes >= need && out_capacity >= 8,
                 "correlate: buffers smaller than the pattern (%zu bytes) or the output (8 bytes)", need);
    if (c->mode == PRNU_MODE_NCC) {
        PRNU_TRY(launch_pair_ncc(static_cast<const float *>(d_left), static_cast<const float *>(d_right),
                                 c->pixels(), static_cast<double *>(d_out), c->ws_small, st));
    } else {
        PRNU_TRY(run_pce(c, PRNU_PCE_REFERENCE, d_left, d_right, static_cast<double *>(d_out), st));
    }
    if (out_bytes) *out_bytes = 8;
    return finish(c, stream);
}

int prnu_correlate_pce(prnu_ctx *ctx, int variant, const void *d_left, const void *d_right, double *d_out,
                       prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(c->mode == PRNU_MODE_PCE, "correlate_pce: context is not in PCE mode");
    PRNU_REQUIRE(d_left && d_right && d_out, "correlate_pce: null buffer");
    PRNU_TRY(run_pce(c, variant, d_left, d_right, d_out, st));
    return finish(c, stream);
}

int prnu_pce_pairs(prnu_ctx *ctx, int variant, const void *const *d_spectra, const int32_t *d_pairs, int npairs,
                   double *d_out, prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(c->mode == PRNU_MODE_PCE, "pce_pairs: context is not in PCE mode");
    PRNU_REQUIRE(d_spectra && d_pairs && d_out, "pce_pairs: null buffer");
    PRNU_TRY(run_pce_pairs(c, variant, d_spectra, d_pairs, npairs, d_out, 0, st));
    return finish(c, stream);
}

int prnu_filter_grayscale(prnu_ctx *ctx, int input_format, const void *d_in, size_t in_bytes, float *d_out,
                          prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(d_in && d_out, "grayscale: null buffer");
    PRNU_TRY(run_filter_grayscale(c, input_format, d_in, in_bytes, d_out, st));
    return finish(c, stream);
}

int prnu_filter_fastnoise(prnu_ctx *ctx, float *d_image, prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(d_image, "fastnoise: null buffer");
    PRNU_TRY(run_filter_fastnoise(c, d_image, st));
    return finish(c, stream);
}

int prnu_filter_zeromean(prnu_ctx *ctx, float *d_image, prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(d_image, "zeromean: null buffer");
    PRNU_TRY(run_filter_zeromean(c, d_image, st));
    return finish(c, stream);
}

int prnu_filter_wiener(prnu_ctx *ctx, float *d_image, prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(d_image, "wiener: null buffer");
    PRNU_TRY(run_filter_wiener(c, d_image, st));
    return finish(c, stream);
}

int prnu_filter_spectral(prnu_ctx *ctx, const float *d_image, float *d_out, size_t out_capacity,
                         prnu_stream stream) {
    PRNU_CTX(ctx);
    PRNU_REQUIRE(d_image && d_out, "spectral: null buffer");
    PRNU_TRY(run_filter_spectral(c, d_image, d_out, out_capacity, st));
    return finish(c, stream);
}

}  // extern "C"
