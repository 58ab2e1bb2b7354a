/*
 * C restatement of the reference's CPU path for the PRNU hot path
 * (TEST INFRASTRUCTURE ONLY: used by tests/ and by the CPU-baseline leg of
 * bench.py, never by the product library).
 *
 * Reference paths are relative to
 * src/main/java/nl/esciencecenter/common_source_identification/.
 *
 * Built by oracle/Makefile into oracle/_build/libprnu_cpu_port.so.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* NormalizedCrossCorrelation.applyCPU (kernels/compare/NormalizedCrossCorrelation.java:127-133):
 * sum_xy accumulates the FP32 product x[i]*y[i] in FP64, and
 * NCC = sum_xy / sqrt(sumsq_x * sumsq_y). */
double port_ncc_cpu(const float *x, double sumsq_x, const float *y, double sumsq_y, int64_t n) {
    double sum_xy = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        float p = x[i] * y[i];
        sum_xy += (double)p;
    }
    return sum_xy / sqrt(sumsq_x * sumsq_y);
}

/* Sum of FP32 squares in FP64 (the sumsq argument the caller of applyCPU passes). */
double port_sumsq(const float *x, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        float p = x[i] * x[i];
        s += (double)p;
    }
    return s;
}

/* A list of pairs over a pattern table, parallel over pairs with `threads`
 * OpenMP threads (0 = all).  patterns[k] points at pattern k (n floats);
 * pairs holds (i, j) index pairs; sumsq[k] is port_sumsq(patterns[k]). */
void port_ncc_pairs(const float *const *patterns, const double *sumsq, const int32_t *pairs, int64_t n_pairs,
                    int64_t n, double *out, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t p = 0; p < n_pairs; ++p) {
        const int32_t i = pairs[2 * p], j = pairs[2 * p + 1];
        out[p] = port_ncc_cpu(patterns[i], sumsq[i], patterns[j], sumsq[j], n);
    }
}

void port_sumsq_many(const float *const *patterns, int64_t count, int64_t n, double *out, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int64_t k = 0; k < count; ++k) out[k] = port_sumsq(patterns[k], n);
}

int port_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
