// Peak-to-correlation energy on gfx950 (CSI/kernels/compare/
// peaktocorrelationenergy.cu and PeakToCorrelationEnergy.java:202-232).
//
// Only Re c = Re IFFT2(C) enters the PCE, and Re IFFT2(C) = IFFT2(H) with
// H(k) = (C(k) + conj(C(-k)))/2 Hermitian, so IFFT2 of Z = H_a + i H_b gives
// Re c_a + i Re c_b: two pairs share one complex transform.  Pipeline per
// couple of pairs, inputs are R2C half spectra h x (w/2+1):
//   1. pce_rows2_kernel  builds row i of Z and runs the inverse row FFT
//                        (length w) in LDS;
//   2. pce_cols2_kernel  loads a strip of columns (coalesced rows of the
//                        strip, no transpose), inverse FFT of each column
//                        (length h) in LDS; writes Re c of both pairs and per
//                        column the max |Re c| with its index and sum (Re c)^2
//                        in FP64;
//   3. pce_final_kernel  per pair: argmax over columns, energy = total -
//                        window (FP64), peak value, PCE.
// A pair's result depends on its couple partner only through FP32 rounding
// (a fixed pair list always forms the same couples).
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

// Cross spectrum of a pair at (i, j) of the full h x w grid from the half
// spectra (wh = w/2 + 1 columns).  Reference variant
// (peaktocorrelationenergy.cu:102-134): conj(X)Y on the left half, X conj(Y)
// mirrored from column w - j of the SAME row i on the right.  Textbook: the
// true cross spectrum (right half mirrored from row (h - i) mod h).
template <int VARIANT>
__device__ __forceinline__ float2 cross_at(const float2 *X, const float2 *Y, int h, int w, int i, int j) {
    const int wh = w / 2 + 1;
    if (j < wh) {
        const float2 x = X[(int64_t)i * wh + j], y = Y[(int64_t)i * wh + j];
        return cmul(cconj(x), y);
    }
    const int mi = (VARIANT == PRNU_PCE_REFERENCE) ? i : (h - i) % h;
    const float2 x = X[(int64_t)mi * wh + (w - j)], y = Y[(int64_t)mi * wh + (w - j)];
    return cmul(x, cconj(y));
}

// Hermitian part H(i, j) = (C(i, j) + conj(C(-i, -j)))/2; the textbook cross
// spectrum of real patterns is Hermitian already.
template <int VARIANT>
__device__ __forceinline__ float2 herm_at(const float2 *X, const float2 *Y, int h, int w, int i, int j) {
    const float2 c = cross_at<VARIANT>(X, Y, h, w, i, j);
    if (VARIANT != PRNU_PCE_REFERENCE) return c;
    const float2 m = cross_at<VARIANT>(X, Y, h, w, (h - i) % h, (w - j) % w);
    return make_float2(0.5f * (c.x + m.x), 0.5f * (c.y - m.y));
}

// Row i of Z = H_a + i H_b for couple blockIdx.y (pairs 2q, 2q+1 of the
// batch; the second may be absent), inverse row FFT: the first pass reads
// the Hermitian parts straight from the half spectra, the last pass writes
// the row of A.  grid (h, couples), LDS w complex.
template <int VARIANT, int MAXN>
__global__ __launch_bounds__(kFftThreads) void pce_rows2_kernel(FftDesc d, int h, int nb, PceBatch bt,
                                                                float2 *__restrict__ Aall, int64_t pstride) {
    extern __shared__ float2 buf[];
    const int q = blockIdx.y, i = blockIdx.x;
    const int w = d.n;
    const float2 *Xa, *Ya, *Xb = nullptr, *Yb = nullptr;
    pce_operands(bt, 2 * q, Xa, Ya);
    const bool two = 2 * q + 1 < nb;
    if (two) pce_operands(bt, 2 * q + 1, Xb, Yb);
    float2 *A = Aall + (int64_t)q * pstride + (int64_t)i * w;
    fft_row<1, MAXN>(
        buf, d,
        [&](int j) {
            const float2 ha = herm_at<VARIANT>(Xa, Ya, h, w, i, j);
            if (!two) return ha;
            const float2 hb = herm_at<VARIANT>(Xb, Yb, h, w, i, j);
            return make_float2(ha.x - hb.y, ha.y + hb.x);
        },
        [&](int j, float2 f) { A[j] = f; });
}

constexpr int kMaxStrip = 16;
// Values of one column-pass strip (64 KiB of LDS, two workgroups per CU): a
// strip of 1024-row columns is 8 wide, so its rows are read in 64-byte
// pieces; the columns are transformed kFftMultiMax values at a time.
constexpr int kPceStripMax = 8192;

// Columns [j0, j0 + sw) of couple blockIdx.y: rows of the strip are read
// contiguously into column-major LDS, the columns are inverse-transformed
// together (length h), and Re / Im go to the two pairs' correlation planes
// (column-major, R[j*h + i]) with per-column energy and |value| argmax.
// grid (ceil(w/sw), couples), LDS sw*h complex, sw <= kMaxStrip.
template <bool BIG>
__global__ __launch_bounds__(kFftThreads, BIG ? 1 : 2) void pce_cols2_kernel(FftDesc d, int w, int sw, int nb,
                                                                             const float2 *__restrict__ Aall,
                                                                             int64_t pstride, float *__restrict__ Rall,
                                                                             int64_t rstride,
                                                                             PeakPart *__restrict__ partall) {
    extern __shared__ float2 buf[];
    __shared__ double sh_e[2 * kMaxStrip][4];
    __shared__ float sh_v[2 * kMaxStrip][4];
    __shared__ int sh_i[2 * kMaxStrip][4];
    const int h = d.n, q = blockIdx.y;
    const int j0 = blockIdx.x * sw;
    const int ncol = min(sw, w - j0);
    const float2 *A = Aall + (int64_t)q * pstride;
    for (int idx = threadIdx.x; idx < h * ncol; idx += kFftThreads) {
        const int i = idx / ncol, c = idx - (idx / ncol) * ncol;
        buf[c * h + i] = A[(int64_t)i * w + j0 + c];
    }
    __syncthreads();
    if constexpr (BIG)
        fft_lds<1>(buf, d);  // one column longer than kFftMultiMax (sw == 1)
    else {
        const int per = max(1, kFftMultiMax / h);  // columns per multi-transform call
        for (int c0 = 0; c0 < ncol; c0 += per) fft_lds_multi<1>(buf + c0 * h, d, min(per, ncol - c0));
    }
    const int halves = 2 * q + 1 < nb ? 2 : 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < ncol; ++c) {
        const int j = j0 + c;
        for (int hf = 0; hf < halves; ++hf) {
            float *R = Rall + (int64_t)(2 * q + hf) * rstride;
            double e = 0.0;
            float best = -1.0f;
            int bi = 0x7fffffff;
            for (int i = threadIdx.x; i < h; i += kFftThreads) {
                const float2 z = buf[c * h + i];
                const float v = hf ? z.y : z.x;
                R[(int64_t)j * h + i] = v;
                e += (double)v * (double)v;
                argmax_merge(best, bi, fabsf(v), i * w + j);
            }
            e = wave_sum_d(e);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const float v2 = __shfl_xor(best, off, 64);
                const int i2 = __shfl_xor(bi, off, 64);
                argmax_merge(best, bi, v2, i2);
            }
            if (lane == 0) {
                sh_e[2 * c + hf][wave] = e;
                sh_v[2 * c + hf][wave] = best;
                sh_i[2 * c + hf][wave] = bi;
            }
        }
    }
    __syncthreads();
    // one thread per (column, pair): fixed-order merge of the 4 waves
    const int t = threadIdx.x;
    if (t < 2 * ncol && (t & 1) < halves) {
        const int c = t >> 1, hf = t & 1;
        PeakPart p;
        p.energy = (sh_e[t][0] + sh_e[t][1]) + (sh_e[t][2] + sh_e[t][3]);
        float best = sh_v[t][0];
        int bi = sh_i[t][0];
        for (int k = 1; k < 4; ++k) argmax_merge(best, bi, sh_v[t][k], sh_i[t][k]);
        p.absmax = best;
        p.index = bi;
        partall[(int64_t)(2 * q + hf) * w + j0 + c] = p;
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
    // no argmax at all (every value NaN): index 0 keeps the window and peak
    // reads inside the plane; the energy, hence the PCE, is NaN anyway
    if (threadIdx.x == 0) s_peak = (bi >= 0 && (int64_t)bi < (int64_t)h * w) ? bi : 0;
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

// Columns per strip of the column pass: a power of two with the strip's
// values (<= kPceStripMax, 64 KiB) in one workgroup's LDS, transformed
// kFftMultiMax values at a time; a column longer than kFftMultiMax runs alone
// through fft_lds (the BIG kernel, up to kFftMaxN).  The row pass runs one row per workgroup through fft_row (rows
// grouped per workgroup measured slower: 184 VGPRs, 2 waves per SIMD).
int pce_strip(int h) {
    int sw = kMaxStrip;
    if (h > kFftMultiMax) return 1;
    while (sw > 1 && (size_t)sw * h > (size_t)kPceStripMax) sw >>= 1;
    return sw;
}

int pce_batch(Context *c, int variant, const PceBatch &bt, int nb, double *out, hipStream_t st) {
    const int h = c->h, w = c->w;
    const int64_t n = c->pixels();
    const int couples = (nb + 1) / 2;
    const size_t plane = ((sizeof(float2) * n) + 255) & ~size_t(255);
    const size_t rbytes = ((sizeof(float) * n) + 255) & ~size_t(255);
    const size_t pbytes = ((sizeof(PeakPart) * w) + 255) & ~size_t(255);
    const int64_t pstride = (int64_t)(plane / sizeof(float2)), rstride = (int64_t)(rbytes / sizeof(float));
    PRNU_TRY(c->ws_pce.ensure(plane * couples + (rbytes + pbytes) * nb));
    char *base = static_cast<char *>(c->ws_pce.ptr);
    float2 *A = reinterpret_cast<float2 *>(base);
    float *R = reinterpret_cast<float *>(base + plane * couples);
    PeakPart *part = reinterpret_cast<PeakPart *>(base + plane * couples + rbytes * nb);
    const FftDesc &dw = fft_desc(c->fft_w);
    const FftDesc &dh = fft_desc(c->fft_h);
    const bool big_h = h > kFftMultiMax;
    const dim3 rgrid(h, couples);
    const size_t rsm = fft_row_smem(w);
#define PRNU_PCE_ROWS(V, M) \
    hipLaunchKernelGGL((pce_rows2_kernel<V, M>), rgrid, dim3(kFftThreads), rsm, st, dw, h, nb, bt, A, pstride)
    if (variant == PRNU_PCE_REFERENCE) {
        if (w <= kFftSmallN) PRNU_PCE_ROWS(PRNU_PCE_REFERENCE, kFftSmallN);
        else if (w <= kFftMidN) PRNU_PCE_ROWS(PRNU_PCE_REFERENCE, kFftMidN);
        else PRNU_PCE_ROWS(PRNU_PCE_REFERENCE, kFftMaxN);
    } else {
        if (w <= kFftSmallN) PRNU_PCE_ROWS(PRNU_PCE_TEXTBOOK, kFftSmallN);
        else if (w <= kFftMidN) PRNU_PCE_ROWS(PRNU_PCE_TEXTBOOK, kFftMidN);
        else PRNU_PCE_ROWS(PRNU_PCE_TEXTBOOK, kFftMaxN);
    }
#undef PRNU_PCE_ROWS
    PRNU_CHECK_LAUNCH();
    const int sw = pce_strip(h);
    const dim3 cgrid((w + sw - 1) / sw, couples);
    const size_t csm = sizeof(float2) * sw * h;  // <= 64 KiB (one column of up to kFftMaxN)
    if (big_h)
        hipLaunchKernelGGL((pce_cols2_kernel<true>), cgrid, dim3(kFftThreads), csm, st, dh, w, sw, nb, A, pstride, R,
                           rstride, part);
    else
        hipLaunchKernelGGL((pce_cols2_kernel<false>), cgrid, dim3(kFftThreads), csm, st, dh, w, sw, nb, A, pstride, R,
                           rstride, part);
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
        const int64_t per = 12 * c->pixels();  // half a complex plane + a real plane per pair
        batch = (int)std::max<int64_t>(2, std::min<int64_t>(512, (int64_t(1) << 30) / per)) & ~1;
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
