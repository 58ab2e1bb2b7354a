// Peak-to-correlation energy on gfx950 (CSI/kernels/compare/
// peaktocorrelationenergy.cu and PeakToCorrelationEnergy.java:202-232).
//
// Pipeline per pair, inputs are two R2C half spectra h x (w/2+1):
//   1. pce_rows_kernel   builds row i of the full cross spectrum C and runs
//                        the inverse row FFT (length w) in LDS;
//   2. transpose         (h x w -> w x h);
//   3. pce_cols_kernel   inverse FFT of each column (length h) in LDS; keeps
//                        Re c, and per column the max |Re c| with its index
//                        and sum (Re c)^2 in FP64;
//   4. pce_final_kernel  argmax over columns, energy = total - window (FP64),
//                        peak value, PCE.
// Reference variant (quirk list in the header of prnu_api.h's correlate):
// right half of C uses row i instead of (h-i) mod h; window rows/cols strictly
// inside (p-5, p+5), clipped, divisor N-121; peak value = Re c[h-1][w-1],
// squared in FP32.  Textbook variant: true cross spectrum; window 11x11 around
// the argmax with cyclic wrap; peak value = c at the argmax; sign kept.
#include "prnu_internal.hpp"
#include "fft_device.hpp"
#include "fft_host.hpp"

#include <algorithm>

namespace prnu {
namespace {

constexpr int kRadius = 5;   // peaktocorrelationenergy.cu:218
constexpr int kSquare = 11;  // peaktocorrelationenergy.cu:217

struct PeakPart {
    double energy;  // sum (Re c)^2 over the column
    float absmax;   // max |Re c| over the column
    int index;      // linear index i*w + j of that max (smallest on ties)
};

__device__ inline double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline double block_sum_256(double v, double *sh) {
    v = wave_sum_d(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// (value, index) max with ties going to the smaller index.
__device__ __forceinline__ void argmax_merge(float &v, int &i, float v2, int i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

__device__ inline void block_argmax_256(float &v, int &idx, float *shv, int *shi) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float v2 = __shfl_xor(v, off, 64);
        const int i2 = __shfl_xor(idx, off, 64);
        argmax_merge(v, idx, v2, i2);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        shv[threadIdx.x >> 6] = v;
        shi[threadIdx.x >> 6] = idx;
    }
    __syncthreads();
    v = shv[0];
    idx = shi[0];
    for (int k = 1; k < 4; ++k) argmax_merge(v, idx, shv[k], shi[k]);
}

// Pair p of a batch: spectra from a device table and an (i, j) pair list,
// or the two pointers of a single pair.
struct PceBatch {
    const float2 *const *tab;  // spectra table (null: use x1, y1)
    const int *pairs;          // 2 ints per pair
    int base;                  // first pair of this batch
    const float2 *x1, *y1;
};

__device__ __forceinline__ void pce_operands(const PceBatch &b, int p, const float2 *&X, const float2 *&Y) {
    if (b.tab) {
        X = b.tab[b.pairs[2 * (b.base + p)]];
        Y = b.tab[b.pairs[2 * (b.base + p) + 1]];
    } else {
        X = b.x1;
        Y = b.y1;
    }
}

template <int VARIANT>
__global__ __launch_bounds__(kFftThreads) void pce_rows_kernel(FftDesc d, int h, PceBatch bt, float2 *__restrict__ Aall,
                                                               int64_t pstride) {
    extern __shared__ float2 buf[];
    const int i = blockIdx.x;
    const int w = d.n, wh = w / 2 + 1;
    const float2 *X, *Y;
    pce_operands(bt, blockIdx.y, X, Y);
    float2 *A = Aall + (int64_t)blockIdx.y * pstride;
    // rows of the half spectra feeding the right half of C
    const int mi = (VARIANT == PRNU_PCE_REFERENCE) ? i : (h - i) % h;
    for (int j = threadIdx.x; j < w; j += kFftThreads) {
        float2 c;
        if (j < wh) {
            const float2 x = X[(int64_t)i * wh + j], y = Y[(int64_t)i * wh + j];
            c = cmul(cconj(x), y);  // (xRe*yRe + xIm*yIm, xRe*yIm - xIm*yRe)
        } else {
            const float2 x = X[(int64_t)mi * wh + (w - j)], y = Y[(int64_t)mi * wh + (w - j)];
            c = cmul(x, cconj(y));
        }
        buf[j] = c;
    }
    __syncthreads();
    fft_lds<1>(buf, d);
    for (int j = threadIdx.x; j < w; j += kFftThreads) A[(int64_t)i * w + j] = buf[j];
}

// Row j of B is column j of the plane; inverse FFT over h.  grid (w, batch).
__global__ __launch_bounds__(kFftThreads) void pce_cols_kernel(FftDesc d, int w, const float2 *__restrict__ Ball,
                                                               float *__restrict__ Rall, PeakPart *__restrict__ partall,
                                                               int64_t pstride, int64_t rstride) {
    extern __shared__ float2 buf[];
    __shared__ double shd[4];
    __shared__ float shv[4];
    __shared__ int shi[4];
    const int j = blockIdx.x;
    const int h = d.n;
    const float2 *B = Ball + (int64_t)blockIdx.y * pstride;
    float *R = Rall + (int64_t)blockIdx.y * rstride;
    for (int i = threadIdx.x; i < h; i += kFftThreads) buf[i] = B[(int64_t)j * h + i];
    __syncthreads();
    fft_lds<1>(buf, d);
    double e = 0.0;
    float best = -1.0f;
    int bi = 0x7fffffff;
    for (int i = threadIdx.x; i < h; i += kFftThreads) {
        const float v = buf[i].x;
        R[(int64_t)j * h + i] = v;
        e += (double)v * (double)v;
        argmax_merge(best, bi, fabsf(v), i * w + j);
    }
    e = block_sum_256(e, shd);
    block_argmax_256(best, bi, shv, shi);
    if (threadIdx.x == 0) {
        PeakPart p;
        p.energy = e;
        p.absmax = best;
        p.index = bi;
        partall[(int64_t)blockIdx.y * w + j] = p;
    }
}

// R is the correlation plane transposed: c[i][j] = R[j*h + i].  grid (batch).
template <int VARIANT>
__global__ __launch_bounds__(256) void pce_final_kernel(const PeakPart *__restrict__ partall,
                                                        const float *__restrict__ Rall, int h, int w,
                                                        double *__restrict__ outall, int base, int64_t rstride) {
    __shared__ double shd[4];
    __shared__ float shv[4];
    __shared__ int shi[4];
    __shared__ int s_peak;
    const PeakPart *part = partall + (int64_t)blockIdx.x * w;
    const float *R = Rall + (int64_t)blockIdx.x * rstride;
    double e = 0.0;
    float best = -1.0f;
    int bi = 0x7fffffff;
    for (int j = threadIdx.x; j < w; j += 256) {
        const PeakPart p = part[j];
        e += p.energy;
        argmax_merge(best, bi, p.absmax, p.index);
    }
    const double total = block_sum_256(e, shd);
    block_argmax_256(best, bi, shv, shi);
    if (threadIdx.x == 0) s_peak = bi;
    __syncthreads();
    const int peak = s_peak;
    const int py = peak / w, px = peak - py * w;
    // window energy
    double win = 0.0;
    int excluded = 0;
    if (VARIANT == PRNU_PCE_REFERENCE) {
        const int r0 = max(0, py - kRadius + 1), r1 = min(h - 1, py + kRadius - 1);
        const int c0 = max(0, px - kRadius + 1), c1 = min(w - 1, px + kRadius - 1);
        const int nr = r1 - r0 + 1, nc = c1 - c0 + 1;
        for (int k = threadIdx.x; k < nr * nc; k += 256) {
            const int i = r0 + k / nc, jj = c0 + k % nc;
            const double v = R[(int64_t)jj * h + i];
            win += v * v;
        }
        excluded = kSquare * kSquare;  // the reference divides by N - 121
    } else {
        const int nr = h < kSquare ? h : kSquare, nc = w < kSquare ? w : kSquare;
        const int r0 = h < kSquare ? 0 : py - kRadius, c0 = w < kSquare ? 0 : px - kRadius;
        for (int k = threadIdx.x; k < nr * nc; k += 256) {
            const int i = ((r0 + k / nc) % h + h) % h, jj = ((c0 + k % nc) % w + w) % w;
            const double v = R[(int64_t)jj * h + i];
            win += v * v;
        }
        excluded = nr * nc;
    }
    win = block_sum_256(win, shd);
    if (threadIdx.x == 0) {
        const double energy = (total - win) / (double)((int64_t)h * w - excluded);
        double pce;
        if (VARIANT == PRNU_PCE_REFERENCE) {
            const float pk = R[(int64_t)(w - 1) * h + (h - 1)];  // last element c[h-1][w-1]
            const float sq = pk * pk;                            // squared in FP32
            pce = (double)sq / energy;
        } else {
            const double pk = R[(int64_t)px * h + py];
            pce = (pk < 0 ? -1.0 : 1.0) * pk * pk / energy;
        }
        outall[base + blockIdx.x] = pce;
    }
}

int pce_batch(Context *c, int variant, const PceBatch &bt, int nb, double *out, hipStream_t st) {
    const int h = c->h, w = c->w;
    const int64_t n = c->pixels();
    const size_t plane = ((sizeof(float2) * n) + 255) & ~size_t(255);
    const size_t rbytes = ((sizeof(float) * n) + 255) & ~size_t(255);
    const size_t pbytes = ((sizeof(PeakPart) * w) + 255) & ~size_t(255);
    const int64_t pstride = (int64_t)(plane / sizeof(float2)), rstride = (int64_t)(rbytes / sizeof(float));
    PRNU_TRY(c->ws_pce.ensure((2 * plane + rbytes + pbytes) * nb));
    char *base = static_cast<char *>(c->ws_pce.ptr);
    float2 *A = reinterpret_cast<float2 *>(base);
    float2 *B = reinterpret_cast<float2 *>(base + plane * nb);
    float *R = reinterpret_cast<float *>(base + 2 * plane * nb);
    PeakPart *part = reinterpret_cast<PeakPart *>(base + (2 * plane + rbytes) * nb);
    const FftDesc &dw = fft_desc(c->fft_w);
    const FftDesc &dh = fft_desc(c->fft_h);
    if (variant == PRNU_PCE_REFERENCE)
        hipLaunchKernelGGL(pce_rows_kernel<PRNU_PCE_REFERENCE>, dim3(h, nb), dim3(kFftThreads), sizeof(float2) * w,
                           st, dw, h, bt, A, pstride);
    else
        hipLaunchKernelGGL(pce_rows_kernel<PRNU_PCE_TEXTBOOK>, dim3(h, nb), dim3(kFftThreads), sizeof(float2) * w,
                           st, dw, h, bt, A, pstride);
    PRNU_CHECK_LAUNCH();
    PRNU_TRY(launch_transpose_c(A, B, h, w, nb, pstride, st));
    hipLaunchKernelGGL(pce_cols_kernel, dim3(w, nb), dim3(kFftThreads), sizeof(float2) * h, st, dh, w, B, R, part,
                       pstride, rstride);
    PRNU_CHECK_LAUNCH();
    if (variant == PRNU_PCE_REFERENCE)
        hipLaunchKernelGGL(pce_final_kernel<PRNU_PCE_REFERENCE>, dim3(nb), dim3(256), 0, st, part, R, h, w, out,
                           bt.base, rstride);
    else
        hipLaunchKernelGGL(pce_final_kernel<PRNU_PCE_TEXTBOOK>, dim3(nb), dim3(256), 0, st, part, R, h, w, out,
                           bt.base, rstride);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

}  // namespace

int run_pce(Context *c, int variant, const void *x, const void *y, double *out, hipStream_t st) {
    PRNU_REQUIRE(variant == PRNU_PCE_REFERENCE || variant == PRNU_PCE_TEXTBOOK, "pce: unknown variant %d", variant);
    PRNU_TRY(ensure_fft_plans(c));
    PceBatch bt;
    bt.tab = nullptr;
    bt.pairs = nullptr;
    bt.base = 0;
    bt.x1 = static_cast<const float2 *>(x);
    bt.y1 = static_cast<const float2 *>(y);
    time_begin(c, st);
    PRNU_TRY(pce_batch(c, variant, bt, 1, out, st));
    time_end(c, st);
    return PRNU_OK;
}

// PCE of `npairs` (i, j) pairs over a device table of spectra, in batches
// whose planes fit `batch` pairs at a time.
int run_pce_pairs(Context *c, int variant, const void *const *d_tab, const int *d_pairs, int npairs, double *out,
                  int batch, hipStream_t st) {
    PRNU_REQUIRE(variant == PRNU_PCE_REFERENCE || variant == PRNU_PCE_TEXTBOOK, "pce: unknown variant %d", variant);
    PRNU_REQUIRE(npairs >= 0, "pce_pairs: negative pair count");
    PRNU_TRY(ensure_fft_plans(c));
    if (batch <= 0) {
        // about 1 GiB of planes
        const int64_t per = 20 * c->pixels();
        batch = (int)std::max<int64_t>(1, std::min<int64_t>(256, (int64_t(1) << 30) / per));
    }
    time_begin(c, st);
    for (int b0 = 0; b0 < npairs; b0 += batch) {
        PceBatch bt;
        bt.tab = reinterpret_cast<const float2 *const *>(d_tab);
        bt.pairs = d_pairs;
        bt.base = b0;
        bt.x1 = bt.y1 = nullptr;
        PRNU_TRY(pce_batch(c, variant, bt, std::min(batch, npairs - b0), out, st));
    }
    time_end(c, st);
    return PRNU_OK;
}

}  // namespace prnu
