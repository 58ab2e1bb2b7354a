/*
 * A C caller of libprnu_hip.so through include/prnu_api.h alone, the way a
 * JNI shim calls it: a context for 96 x 128 images in NCC mode, M synthetic
 * 8-bit RGB frames of two "cameras" uploaded with prnu_memcpy_h2d,
 * preprocessInputGPU per image (prnu_preprocess), correlateGPU per pair
 * (prnu_correlate), then the same pairs as one split-BF16 Gram
 * (prnu_pattern_stats + prnu_gram_ncc).  Checks, on the host:
 *   - every pattern's NCC with itself is 1 (up to FP32 rounding of x*x);
 *   - the Gram equals the per-pair scores within 1e-5;
 *   - the per-pair scores equal the FP64 NCC of the downloaded patterns
 *     (FP32 products, FP64 sums, no 1/n) within 1e-9;
 *   - an undersized output buffer gives PRNU_EINVAL and a message.
 * Exit status 0 when all hold.  Built and run by tests/test_abi_driver.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "prnu_api.h"

#define H 96
#define W 128
#define M 6

#define CHECK(call)                                                                   \
    do {                                                                              \
        int rc_ = (call);                                                             \
        if (rc_ != PRNU_OK) {                                                         \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, prnu_last_error());   \
            return 2;                                                                 \
        }                                                                             \
    } while (0)

static uint32_t rng_state = 12345u;
static double uniform(void) {
    rng_state = rng_state * 1664525u + 1013904223u;
    return (double)(rng_state >> 8) / 16777216.0;
}

static double ncc_fp64(const float *x, const float *y, int64_t n) {
    double sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0;
    for (int64_t i = 0; i < n; ++i) {
        sx += x[i];
        sy += y[i];
        sxx += (double)(x[i] * x[i]);
        syy += (double)(y[i] * y[i]);
        sxy += (double)(x[i] * y[i]);
    }
    return sxy / (sqrt(sxx - sx * sx) * sqrt(syy - sy * sy));
}

int main(void) {
    const int64_t n = (int64_t)H * W;
    prnu_ctx *ctx = NULL;
    CHECK(prnu_ctx_create(H, W, PRNU_MODE_NCC, &ctx));
    if (prnu_max_file_size(ctx) != 3 * n || prnu_max_input_size(ctx) != 4 * n || prnu_max_output_size(ctx) != 8) {
        fprintf(stderr, "size queries disagree with getMax*Size\n");
        return 1;
    }
    /* two camera fingerprints, each frame = smooth scene * (1 + 0.05 K) + noise */
    static float cam[2][H * W];
    for (int c = 0; c < 2; ++c)
        for (int64_t i = 0; i < n; ++i) cam[c][i] = (float)(uniform() - 0.5);
    uint8_t *rgb = (uint8_t *)malloc((size_t)(3 * n * M));
    for (int k = 0; k < M; ++k)
        for (int64_t i = 0; i < n; ++i) {
            const int y = (int)(i / W), x = (int)(i % W);
            const double scene = 120.0 + 60.0 * sin(0.05 * x + 0.3 * k) * cos(0.07 * y);
            for (int ch = 0; ch < 3; ++ch) {
                double v = scene * (1.0 + 0.05 * cam[k % 2][i]) + 4.0 * (uniform() - 0.5) + 5 * ch;
                rgb[3 * (n * k + i) + ch] = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
            }
        }
    void *d_rgb = NULL, *d_pat = NULL, *d_out = NULL, *d_ptrs = NULL, *d_sum = NULL, *d_sumsq = NULL;
    CHECK(prnu_malloc_device(&d_rgb, (size_t)(3 * n * M)));
    CHECK(prnu_malloc_device(&d_pat, sizeof(float) * (size_t)(n * M)));
    CHECK(prnu_malloc_device(&d_out, sizeof(double) * M * M));
    CHECK(prnu_malloc_device(&d_ptrs, sizeof(float *) * M));
    CHECK(prnu_malloc_device(&d_sum, sizeof(double) * M));
    CHECK(prnu_malloc_device(&d_sumsq, sizeof(double) * M));
    CHECK(prnu_memcpy_h2d(d_rgb, rgb, (size_t)(3 * n * M), NULL));
    for (int k = 0; k < M; ++k) {
        size_t out_bytes = 0;
        CHECK(prnu_preprocess(ctx, PRNU_INPUT_RGB8, (uint8_t *)d_rgb + 3 * n * k, (size_t)(3 * n),
                              (float *)d_pat + n * k, sizeof(float) * (size_t)n, &out_bytes, NULL));
        if (out_bytes != sizeof(float) * (size_t)n) {
            fprintf(stderr, "preprocess returned %zu bytes\n", out_bytes);
            return 1;
        }
    }
    double pair[M][M];
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            size_t out_bytes = 0;
            CHECK(prnu_correlate(ctx, (float *)d_pat + n * i, sizeof(float) * (size_t)n, (float *)d_pat + n * j,
                                 sizeof(float) * (size_t)n, d_out, 8, &out_bytes, NULL));
            CHECK(prnu_memcpy_d2h(&pair[i][j], d_out, 8, NULL));
        }
    float *h_ptrs[M];
    for (int k = 0; k < M; ++k) h_ptrs[k] = (float *)d_pat + n * k;
    CHECK(prnu_memcpy_h2d(d_ptrs, h_ptrs, sizeof(h_ptrs), NULL));
    CHECK(prnu_pattern_stats((const float *const *)d_ptrs, M, n, (double *)d_sum, (double *)d_sumsq, NULL, NULL));
    CHECK(prnu_gram_ncc(ctx, (const float *const *)d_ptrs, M, PRNU_PREC_SPLIT_BF16, (const double *)d_sum,
                        (const double *)d_sumsq, NULL, (double *)d_out, NULL));
    CHECK(prnu_stream_synchronize(NULL));
    double gram[M][M];
    CHECK(prnu_memcpy_d2h(gram, d_out, sizeof(gram), NULL));
    float *pats = (float *)malloc(sizeof(float) * (size_t)(n * M));
    CHECK(prnu_memcpy_d2h(pats, d_pat, sizeof(float) * (size_t)(n * M), NULL));
    double worst_gram = 0, worst_fp64 = 0, worst_self = 0, same_min = 1e9, other_max = -1e9;
    for (int i = 0; i < M; ++i)
        for (int j = 0; j < M; ++j) {
            worst_gram = fmax(worst_gram, fabs(gram[i][j] - pair[i][j]));
            worst_fp64 = fmax(worst_fp64, fabs(pair[i][j] - ncc_fp64(pats + n * i, pats + n * j, n)));
            if (i == j) worst_self = fmax(worst_self, fabs(pair[i][i] - 1.0));
            else if (i % 2 == j % 2) same_min = fmin(same_min, pair[i][j]);
            else other_max = fmax(other_max, pair[i][j]);
        }
    /* an undersized output buffer: the reference's IllegalArgumentException */
    size_t out_bytes = 0;
    const int rc = prnu_correlate(ctx, d_pat, sizeof(float) * (size_t)n, d_pat, sizeof(float) * (size_t)n, d_out, 4,
                                  &out_bytes, NULL);
    printf("abi driver: |gram - per pair| %.3e, |per pair - fp64| %.3e, |self - 1| %.3e, same camera >= %.4f, "
           "other <= %.4f, undersized output -> %d (%s)\n",
           worst_gram, worst_fp64, worst_self, same_min, other_max, rc, prnu_last_error());
    prnu_free_device(d_rgb);
    prnu_free_device(d_pat);
    prnu_free_device(d_out);
    prnu_free_device(d_ptrs);
    prnu_free_device(d_sum);
    prnu_free_device(d_sumsq);
    prnu_ctx_destroy(ctx);
    free(rgb);
    free(pats);
    const int ok = worst_gram <= 1e-5 && worst_fp64 <= 1e-9 && worst_self <= 1e-6 && same_min > other_max &&
                   rc == PRNU_EINVAL;
    return ok ? 0 : 1;
}
