// Peak-to-correlation energy on gfx950 (CSI/kernels/compare/
// peaktocorrelationenergy.cu and PeakToCorrelationEnergy.java:202-232).
//
// Only Re c = Re IFFT2(C) enters the PCE, and Re IFFT2(C) = IFFT2(H) with
// H(k) = (C(k) + conj(C(-k)))/2 Hermitian, so IFFT2 of Z = H_a + i H_b gives
// Re c_a + i Re c_b: two pairs share one complex transform.  Pipeline per
// couple of pairs, inputs are R2C half spectra h x (w/2+1):
//   1. pce_rows2_kernel  builds row i of Z and runs the inverse row FFT
//                        (length w) in LDS;
//   2. transpose         the row-pass output, so that a column is a
//                        contiguous run of h values;
//      pce_cols_kernel   one column per workgroup: inverse FFT (length h)
//                        through fft_row, whose first pass loads the column
//                        and whose last pass hands every output to the
//                        per-column sums: the max |Re c| with its index and
//                        sum (Re c)^2 in FP64 for each pair of the couple
//                        (the correlation planes themselves are never
//                        written);
//   3. pce_window_kernel per pair and window row: argmax over columns; the
//                        few values the PCE needs beyond the sums -- the
//                        exclusion window around the peak and the peak value
//                        -- are recomputed as direct column DFTs of the
//                        row-pass output (FP64 sums, about a hundred values of
//                        h terms each);
//      pce_score_kernel  energy = total - window (FP64), PCE.
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

// H(u, j) and, for 1 <= j < w/2, H(u, w - j) -- herm_at's values bit for bit
// -- from the four spectrum values X, Y at (u, j) and (um, j), um = (h - u)
// mod h: every cross-spectrum term either needs reads at (u, j) or (um, j).
// herm_at reads each of them twice over the row (once for j, once for its
// mirror column).
template <int VARIANT>
__device__ __forceinline__ void herm_pair(const float2 *X, const float2 *Y, int w, int u, int um, int j, float2 &hj,
                                          float2 &hm) {
    const int wh = w / 2 + 1;
    const float2 x0 = X[(int64_t)u * wh + j], y0 = Y[(int64_t)u * wh + j];
    const float2 x1 = X[(int64_t)um * wh + j], y1 = Y[(int64_t)um * wh + j];
    if (VARIANT != PRNU_PCE_REFERENCE) {
        hj = cmul(cconj(x0), y0);  // cross_at(u, j), j < wh
        hm = cmul(x1, cconj(y1));  // cross_at(u, w - j): row um, column j
        return;
    }
    // reference: c = cross_at(u, j); m = cross_at(um, (w - j) mod w), which is
    // row um's column j (mirrored from w - j) unless j is 0 or w/2
    const float2 c = cmul(cconj(x0), y0);
    const float2 m = (j == 0 || 2 * j == w) ? cmul(cconj(x1), y1) : cmul(x1, cconj(y1));
    hj = make_float2(0.5f * (c.x + m.x), 0.5f * (c.y - m.y));
    // the mirror column w - j: c' = cross_at(u, w - j) = x0 conj(y0) (same row
    // in the reference), m' = cross_at(um, j) = conj(x1) y1
    const float2 c2 = cmul(x0, cconj(y0));
    const float2 m2 = cmul(cconj(x1), y1);
    hm = make_float2(0.5f * (c2.x + m2.x), 0.5f * (c2.y - m2.y));
}

// Row i of Z = H_a + i H_b for couple blockIdx.y (pairs 2q, 2q+1 of the
// batch; the second may be absent), inverse row FFT: the first pass reads
// the Hermitian parts straight from the half spectra, the last pass writes
// the row of A.  grid (16 ceil((h/2 + 1) / 8), couples), LDS w complex.
template <int VARIANT, int MAXN>
__global__ __launch_bounds__(kFftThreads) void pce_rows2_kernel(FftDesc d, int h, int nb, PceBatch bt,
                                                                float2 *__restrict__ Aall, int64_t pstride) {
    extern __shared__ float2 buf[];
    // Row i reads spectrum rows i and (h - i) mod h (the mirrored half of the
    // cross spectrum), and so does row (h - i) mod h: workgroups x and x + 8
    // of each group of 16 -- the same XCD, as workgroup ids go to the 8 XCDs
    // round-robin -- take the two rows of a mirror pair u, h - u, so the
    // second reads them from that XCD's L2.  Rows 0 and h/2 are their own
    // mirrors (the second workgroup of the pair leaves); ids past the last
    // pair leave too (grid rows_grid(h) >= h).
    const int x = blockIdx.x;
    const int u = (x >> 4) * 8 + (x & 7), side = (x >> 3) & 1;
    if (u > h / 2) return;
    const int i = side ? (h - u) % h : u;
    if (side && i == u) return;
    const int q = blockIdx.y;
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

// Both rows of a mirror pair u, h - u (one when u is its own mirror) of
// couple blockIdx.y.  The Hermitian parts of row u, H(u, j) for every j,
// are formed once into LDS; row -u needs no loads of its own, since
// H(-u, j) = conj(H(u, -j)) (H is the Hermitian part of the cross spectrum
// in both variants), so its transform input is
// Z(-u, j) = conj(H_a(u, -j)) + i conj(H_b(u, -j)).  That halves the
// spectrum reads of pce_rows2_kernel, whose two workgroups per pair each
// form their own row's Hermitian parts.  grid (h/2 + 1, couples), LDS 3w
// complex (H_a, H_b of row u, the transform buffer).
template <int VARIANT, int MAXN>
__global__ __launch_bounds__(kFftThreads) void pce_rows_mirror_kernel(FftDesc d, int h, int nb, PceBatch bt,
                                                                      float2 *__restrict__ Aall, int64_t pstride) {
    extern __shared__ float2 smem[];
    const int w = d.n;
    float2 *ha = smem, *hb = smem + w, *buf = smem + 2 * w;
    const int u = blockIdx.x;
    if (u > h / 2) return;
    const int um = (h - u) % h;
    const int q = blockIdx.y;
    const float2 *Xa, *Ya, *Xb = nullptr, *Yb = nullptr;
    pce_operands(bt, 2 * q, Xa, Ya);
    const bool two = 2 * q + 1 < nb;
    if (two) pce_operands(bt, 2 * q + 1, Xb, Yb);
    for (int j = threadIdx.x; j <= w / 2; j += kFftThreads) {  // each spectrum value read once
        float2 a, am, b = make_float2(0.f, 0.f), bm = make_float2(0.f, 0.f);
        herm_pair<VARIANT>(Xa, Ya, w, u, um, j, a, am);
        if (two) herm_pair<VARIANT>(Xb, Yb, w, u, um, j, b, bm);
        ha[j] = a;
        hb[j] = b;
        if (j > 0 && 2 * j < w) {
            ha[w - j] = am;
            hb[w - j] = bm;
        }
    }
    __syncthreads();
    float2 *A = Aall + (int64_t)q * pstride;
    fft_row<1, MAXN>(
        buf, d,
        [&](int j) {
            const float2 a = ha[j], b = hb[j];
            return make_float2(a.x - b.y, a.y + b.x);
        },
        [&](int j, float2 f) { A[(int64_t)u * w + j] = f; });
    if (um == u) return;
    fft_row<1, MAXN>(
        buf, d,
        [&](int j) {
            const int mj = j == 0 ? 0 : w - j;
            const float2 a = ha[mj], b = hb[mj];
            return make_float2(a.x + b.y, b.x - a.y);
        },
        [&](int j, float2 f) { A[(int64_t)um * w + j] = f; });
}

// Column j of couple blockIdx.y from the transposed row-pass output
// (B[j*h + i] = A[i][j]): inverse FFT of length h; every output value goes
// straight into this thread's sums of Re (pair 2q) and Im (pair 2q + 1) --
// FP64 energy and the max |value| with its linear index i*w + j, smallest
// index on ties -- which a fixed-order block reduction turns into the
// column's partials.  grid (w, couples), LDS h complex.
template <int MAXN>
__global__ __launch_bounds__(kFftThreads) void pce_cols_kernel(FftDesc d, int w, int nb, const float2 *__restrict__ Ball,
                                                               int64_t pstride, PeakPart *__restrict__ partall) {
    extern __shared__ float2 buf[];
    __shared__ double sh_e[2][4];
    __shared__ float sh_v[2][4];
    __shared__ int sh_i[2][4];
    const int h = d.n, q = blockIdx.y, j = blockIdx.x;
    const float2 *col = Ball + (int64_t)q * pstride + (int64_t)j * h;
    double e0 = 0.0, e1 = 0.0;
    float b0 = -1.0f, b1 = -1.0f;
    int i0 = 0x7fffffff, i1 = 0x7fffffff;
    fft_row<1, MAXN>(
        buf, d, [&](int i) { return col[i]; },
        [&](int i, float2 z) {
            const int idx = i * w + j;
            e0 += (double)z.x * (double)z.x;
            e1 += (double)z.y * (double)z.y;
            argmax_merge(b0, i0, fabsf(z.x), idx);
            argmax_merge(b1, i1, fabsf(z.y), idx);
        });
    const int halves = 2 * q + 1 < nb ? 2 : 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    e0 = wave_sum_d(e0);
    e1 = wave_sum_d(e1);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float v2 = __shfl_xor(b0, off, 64);
        int k2 = __shfl_xor(i0, off, 64);
        argmax_merge(b0, i0, v2, k2);
        v2 = __shfl_xor(b1, off, 64);
        k2 = __shfl_xor(i1, off, 64);
        argmax_merge(b1, i1, v2, k2);
    }
    if (lane == 0) {
        sh_e[0][wave] = e0;
        sh_e[1][wave] = e1;
        sh_v[0][wave] = b0;
        sh_v[1][wave] = b1;
        sh_i[0][wave] = i0;
        sh_i[1][wave] = i1;
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < halves) {  // fixed-order merge of the 4 waves
        PeakPart p;
        p.energy = (sh_e[t][0] + sh_e[t][1]) + (sh_e[t][2] + sh_e[t][3]);
        float best = sh_v[t][0];
        int bi = sh_i[t][0];
        for (int k = 1; k < 4; ++k) argmax_merge(best, bi, sh_v[t][k], sh_i[t][k]);
        p.absmax = best;
        p.index = bi;
        partall[(int64_t)(2 * q + t) * w + j] = p;
    }
}

// pce_rows_mirror_kernel for w = 1024 with one wave per row: the two waves
// form H_a, H_b of row u into LDS together, then wave 0 transforms row u and
// wave 1 row -u at the same time (fft_wave_1024), their transform buffers
// laid over H_a, H_b once the first pass has read them.  When u is its own
// mirror, wave 1 transforms and stores nothing of its own (it still takes
// part in every barrier).  grid (h/2 + 1, couples), 128 threads.
template <int VARIANT>
__global__ __launch_bounds__(128) void pce_rows_mirror_wave_kernel(const float2 *__restrict__ tw, int h, int nb,
                                                                   PceBatch bt, float2 *__restrict__ Aall,
                                                                   int64_t pstride) {
    constexpr int w = 1024;
    __shared__ float2 smem[2 * fft_wave_lds(w)];
    float2 *ha = smem, *hb = smem + w;
    const int u = blockIdx.x;
    const int um = (h - u) % h;
    const int q = blockIdx.y;
    const float2 *Xa, *Ya, *Xb = nullptr, *Yb = nullptr;
    pce_operands(bt, 2 * q, Xa, Ya);
    const bool two = 2 * q + 1 < nb;
    if (two) pce_operands(bt, 2 * q + 1, Xb, Yb);
    for (int j = threadIdx.x; j <= w / 2; j += 128) {  // each spectrum value read once
        float2 a, am, b = make_float2(0.f, 0.f), bm = make_float2(0.f, 0.f);
        herm_pair<VARIANT>(Xa, Ya, w, u, um, j, a, am);
        if (two) herm_pair<VARIANT>(Xb, Yb, w, u, um, j, b, bm);
        ha[j] = a;
        hb[j] = b;
        if (j > 0 && 2 * j < w) {
            ha[w - j] = am;
            hb[w - j] = bm;
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6;
    float2 *buf = smem + wave * fft_wave_lds(w);
    float2 *A = Aall + (int64_t)q * pstride;
    const int row = wave == 0 ? u : um;
    const bool store = wave == 0 || um != u;
    auto get = [&](int j) {
        if (wave == 0) {
            const float2 a = ha[j], b = hb[j];
            return make_float2(a.x - b.y, a.y + b.x);
        }
        const int mj = j == 0 ? 0 : w - j;
        const float2 a = ha[mj], b = hb[mj];
        return make_float2(a.x + b.y, b.x - a.y);
    };
    auto put = [&](int j, float2 f) {
        if (store) A[(int64_t)row * w + j] = f;
    };
    // fft_wave_1024's passes; the first with workgroup barriers (both waves
    // read H_a, H_b, which their buffers overlay), the others each in its
    // own buffer with wave-level ordering
    fft_wave_pass<16, 1, true, false, 1024, 1, true>(buf, tw, get, put);
    fft_wave_pass<16, 1, false, false, 1024, 16, false, 4, 64, true>(buf, tw, get, put);
    fft_wave_pass<4, 1, false, true, 1024, 256, false, 4, 64, true>(buf, tw, get, put);
}

// pce_cols_kernel for h = 1024 with one wave per column (fft_wave_1024: 16
// values per lane, radix 16, 16, 4): every lane works in every pass and the
// barriers between passes are one wave's.  Same sums; the FP64 energy is
// added in another order (one wave's lanes instead of four waves' partials).
__global__ __launch_bounds__(64) void pce_cols_wave_kernel(const float2 *__restrict__ tw, int w, int nb,
                                                           const float2 *__restrict__ Ball, int64_t pstride,
                                                           PeakPart *__restrict__ partall) {
    __shared__ float2 buf[fft_wave_lds(1024)];
    const int q = blockIdx.y, j = blockIdx.x;
    const float2 *col = Ball + (int64_t)q * pstride + (int64_t)j * 1024;
    double e0 = 0.0, e1 = 0.0;
    float b0 = -1.0f, b1 = -1.0f;
    int i0 = 0x7fffffff, i1 = 0x7fffffff;
    fft_wave_1024<1>(
        buf, tw, [&](int i) { return col[i]; },
        [&](int i, float2 z) {
            const int idx = i * w + j;
            e0 += (double)z.x * (double)z.x;
            e1 += (double)z.y * (double)z.y;
            argmax_merge(b0, i0, fabsf(z.x), idx);
            argmax_merge(b1, i1, fabsf(z.y), idx);
        });
    e0 = wave_sum_d(e0);
    e1 = wave_sum_d(e1);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float v2 = __shfl_xor(b0, off, 64);
        int k2 = __shfl_xor(i0, off, 64);
        argmax_merge(b0, i0, v2, k2);
        v2 = __shfl_xor(b1, off, 64);
        k2 = __shfl_xor(i1, off, 64);
        argmax_merge(b1, i1, v2, k2);
    }
    const int t = threadIdx.x;
    const int halves = 2 * q + 1 < nb ? 2 : 1;
    if (t < halves) {
        PeakPart p;
        p.energy = t == 0 ? e0 : e1;
        p.absmax = t == 0 ? b0 : b1;
        p.index = t == 0 ? i0 : i1;
        partall[(int64_t)(2 * q + t) * w + j] = p;
    }
}

// pce_cols_wave_kernel for h = 1024 reading the row-major row-pass output A
// itself, so no transpose precedes it: a workgroup of 8 waves stages a strip
// of 8 adjacent columns into LDS (every load a 64-byte piece of a row, each
// column in its own padded buffer), then wave c transforms column c there.
// Workgroup g takes strip (g mod 8) * per + g / 8, so that neighbouring strips
// (the two halves of a 128-byte line) run on one XCD.  grid (8 per, couples).
constexpr int kStripBuf = 1092;  // fft_wave_lds(1024) = 1088, +4 spreads the staging writes over banks
__global__ __launch_bounds__(512) void pce_cols_strip_kernel(const float2 *__restrict__ tw, int w, int nb,
                                                             const float2 *__restrict__ Aall, int64_t pstride,
                                                             PeakPart *__restrict__ partall, int per) {
    __shared__ __attribute__((aligned(16))) float2 lds[8 * kStripBuf];
    const int g = blockIdx.x;
    const int strip = (g & 7) * per + (g >> 3);
    if (strip >= w / 8) return;
    const int q = blockIdx.y;
    const int j0 = strip * 8;
    const float2 *A = Aall + (int64_t)q * pstride;
    const int t = (int)threadIdx.x;
    {
        // two adjacent columns per thread, one 16-byte load (w and j0 even)
        const int c2 = (t & 3) * 2, r = t >> 2;
#pragma unroll 4
        for (int k = 0; k < 1024 / 128; ++k) {
            const int i = r + 128 * k;
            const float4 v = *reinterpret_cast<const float4 *>(A + (int64_t)i * w + j0 + c2);
            lds[c2 * kStripBuf + fft_wave_pad<4>(i)] = make_float2(v.x, v.y);
            lds[(c2 + 1) * kStripBuf + fft_wave_pad<4>(i)] = make_float2(v.z, v.w);
        }
    }
    __syncthreads();
    const int wave = t >> 6;
    float2 *buf = lds + wave * kStripBuf;
    const int j = j0 + wave;
    double e0 = 0.0, e1 = 0.0;
    float b0 = -1.0f, b1 = -1.0f;
    int i0 = 0x7fffffff, i1 = 0x7fffffff;
    fft_wave_1024<1, true, true>(  // after the staging barrier each wave works in its own buffer
        buf, tw, [&](int i) { return buf[fft_wave_pad<4>(i)]; },
        [&](int i, float2 z) {
            const int idx = i * w + j;
            e0 += (double)z.x * (double)z.x;
            e1 += (double)z.y * (double)z.y;
            argmax_merge(b0, i0, fabsf(z.x), idx);
            argmax_merge(b1, i1, fabsf(z.y), idx);
        });
    e0 = wave_sum_d(e0);
    e1 = wave_sum_d(e1);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float v2 = __shfl_xor(b0, off, 64);
        int k2 = __shfl_xor(i0, off, 64);
        argmax_merge(b0, i0, v2, k2);
        v2 = __shfl_xor(b1, off, 64);
        k2 = __shfl_xor(i1, off, 64);
        argmax_merge(b1, i1, v2, k2);
    }
    const int lane = t & 63;
    const int halves = 2 * q + 1 < nb ? 2 : 1;
    if (lane < halves) {
        PeakPart p;
        p.energy = lane == 0 ? e0 : e1;
        p.absmax = lane == 0 ? b0 : b1;
        p.index = lane == 0 ? i0 : i1;
        partall[(int64_t)(2 * q + lane) * w + j] = p;
    }
}

// Value c[i][j] of pair hf (0: Re, 1: Im of the couple's complex plane) as
// the direct inverse DFT of column j of the row-pass output A (h x w,
// row-major): sum_k A[k][j] exp(+2 pi i i k / h), summed in FP64 per thread
// over k = t, t + 256, ... and then over the block (fixed order).  `tw` is
// the length-h table exp(-2 pi i m / h); ncells cells (i, j[c]) of one row i
// at once, which share the twiddle of each k: its index (i k) mod h is
// stepped by (256 i) mod h from one k to the next.
constexpr int kMaxCells = 11;

__device__ void dft_cells(const float2 *__restrict__ A, int h, int w, const float2 *__restrict__ tw, int hf,
                          int crow, const int *cj, int ncells, double *out, double (*sh)[4]) {
    double acc[kMaxCells];
#pragma unroll
    for (int c = 0; c < kMaxCells; ++c) acc[c] = 0.0;
    const int step = (int)(((int64_t)crow * 256) % h);
    int ti = (int)(((int64_t)crow * threadIdx.x) % h);
    for (int k = threadIdx.x; k < h; k += 256) {
        const float2 *row = A + (int64_t)k * w;
        const float2 t = tw[ti];  // exp(-i theta); the inverse uses exp(+i theta)
        ti += step;
        if (ti >= h) ti -= h;
#pragma unroll
        for (int c = 0; c < kMaxCells; ++c) {
            if (c < ncells) {
                const float2 a = row[cj[c]];
                // (a.x + i a.y)(t.x - i t.y): Re = a.x t.x + a.y t.y, Im = a.y t.x - a.x t.y
                acc[c] += hf ? ((double)a.y * t.x - (double)a.x * t.y) : ((double)a.x * t.x + (double)a.y * t.y);
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < kMaxCells; ++c) {
        if (c < ncells) {
            const double v = wave_sum_d(acc[c]);
            if (lane == 0) sh[c][wave] = v;
        }
    }
    __syncthreads();
    if (threadIdx.x < ncells) {
        const int c = threadIdx.x;
        out[c] = (sh[c][0] + sh[c][1]) + (sh[c][2] + sh[c][3]);
    }
    __syncthreads();
}

// Per pair p of the batch, 16 doubles of window work: [0, 11) the sums of
// squares of window rows, [11] the peak value, [12] the total energy, [13]
// the excluded cell count of the energy divisor.
constexpr int kWinSlots = 16;

// Argmax and total energy of pair p from its column partials (every block of
// the pair computes them the same way, in a fixed order).
__device__ void pair_peak(const PeakPart *part, int w, int h, double *shd, float *shv, int *shi, int *s_peak,
                          double *total) {
    double e = 0.0;
    float best = -1.0f;
    int bi = 0x7fffffff;
    for (int j = threadIdx.x; j < w; j += 256) {
        const PeakPart q = part[j];
        e += q.energy;
        argmax_merge(best, bi, q.absmax, q.index);
    }
    *total = block_sum_256(e, shd);
    block_argmax_256(best, bi, shv, shi);
    // no argmax at all (every value NaN): index 0 keeps the window and peak
    // reads inside the plane; the energy, hence the PCE, is NaN anyway
    if (threadIdx.x == 0) *s_peak = (bi >= 0 && (int64_t)bi < (int64_t)h * w) ? bi : 0;
    __syncthreads();
}

// Block (p, y): pair p's window row y (y < 11), or its peak value (y == 11:
// the last element c[h-1][w-1] in the reference variant; the textbook
// variant's peak value comes with the window row that holds it).  grid
// (batch, 12).
// Per pair, once: the total energy and the argmax of the column partials
// (pair_peak) into slots 12 and 14 (the index as an exact double), read by
// every window row's block.  grid (pairs), 256 threads.
__global__ __launch_bounds__(256) void pce_peak_kernel(const PeakPart *__restrict__ partall, int h, int w,
                                                       double *__restrict__ wbuf) {
    __shared__ double shd[4];
    __shared__ float shv[4];
    __shared__ int shi[4];
    __shared__ int s_peak;
    const int p = blockIdx.x;
    double total;
    pair_peak(partall + (int64_t)p * w, w, h, shd, shv, shi, &s_peak, &total);
    if (threadIdx.x == 0) {
        wbuf[(int64_t)p * kWinSlots + 12] = total;
        wbuf[(int64_t)p * kWinSlots + 14] = (double)s_peak;
    }
}

template <int VARIANT>
__global__ __launch_bounds__(256) void pce_window_kernel(const PeakPart *__restrict__ partall,
                                                         const float2 *__restrict__ Aall, int64_t pstride,
                                                         const float2 *__restrict__ tw, int h, int w,
                                                         double *__restrict__ wbuf) {
    __shared__ double sh_cell[kMaxCells][4];
    __shared__ double cells[kMaxCells];
    __shared__ int ci[kMaxCells], cj[kMaxCells];
    const int p = blockIdx.x, y = blockIdx.y, hf = p & 1;
    const float2 *A = Aall + (int64_t)(p >> 1) * pstride;
    double *slots = wbuf + (int64_t)p * kWinSlots;
    const int peak = (int)slots[14];  // pce_peak_kernel's
    const int py = peak / w, px = peak - py * w;
    int r0, nr, c0, nc;
    bool wrap;
    if (VARIANT == PRNU_PCE_REFERENCE) {  // strictly inside (p - 5, p + 5), clipped
        r0 = max(0, py - kRadius + 1);
        nr = min(h - 1, py + kRadius - 1) - r0 + 1;
        c0 = max(0, px - kRadius + 1);
        nc = min(w - 1, px + kRadius - 1) - c0 + 1;
        wrap = false;
    } else {  // 11 x 11 around the peak, cyclic
        nr = h < kSquare ? h : kSquare;
        nc = w < kSquare ? w : kSquare;
        r0 = h < kSquare ? 0 : py - kRadius;
        c0 = w < kSquare ? 0 : px - kRadius;
        wrap = true;
    }
    if (y == 0 && threadIdx.x == 0)  // the reference divides by N - 121
        slots[13] = VARIANT == PRNU_PCE_REFERENCE ? (double)(kSquare * kSquare) : (double)(nr * nc);
    const int t = (int)threadIdx.x;  // signed: c0 may be negative before the wrap
    if (y < kSquare) {
        if (y >= nr) {
            if (t == 0) slots[y] = 0.0;
            return;
        }
        if (t < nc) {
            ci[t] = wrap ? ((r0 + y) % h + h) % h : r0 + y;
            cj[t] = wrap ? ((c0 + t) % w + w) % w : c0 + t;
        }
        __syncthreads();
        dft_cells(A, h, w, tw, hf, ci[0], cj, nc, cells, sh_cell);  // one window row: every ci[c] is ci[0]
        if (t == 0) {
            double win = 0.0;
            for (int c = 0; c < nc; ++c) {
                win += cells[c] * cells[c];
                if (VARIANT != PRNU_PCE_REFERENCE && ci[c] == py && cj[c] == px) slots[11] = cells[c];
            }
            slots[y] = win;
        }
    } else if (VARIANT == PRNU_PCE_REFERENCE) {  // the peak value is the last element c[h-1][w-1]
        if (t == 0) {
            ci[0] = h - 1;
            cj[0] = w - 1;
        }
        __syncthreads();
        dft_cells(A, h, w, tw, hf, ci[0], cj, 1, cells, sh_cell);
        if (t == 0) slots[11] = cells[0];
    }
}

// PCE of each pair of the batch from its window work (window rows summed in
// order).  grid (ceil(batch / 64)), 64 threads.
template <int VARIANT>
__global__ __launch_bounds__(64) void pce_score_kernel(const double *__restrict__ wbuf, int nb, int h, int w,
                                                       double *__restrict__ outall, int base) {
    const int p = blockIdx.x * 64 + threadIdx.x;
    if (p >= nb) return;
    const double *slots = wbuf + (int64_t)p * kWinSlots;
    double win = 0.0;
    for (int y = 0; y < kSquare; ++y) win += slots[y];
    const double energy = (slots[12] - win) / ((double)h * (double)w - slots[13]);
    double pce;
    if (VARIANT == PRNU_PCE_REFERENCE) {
        const float pk = (float)slots[11];
        const float sq = pk * pk;  // squared in FP32
        pce = (double)sq / energy;
    } else {
        const double pk = slots[11];
        pce = (pk < 0 ? -1.0 : 1.0) * pk * pk / energy;
    }
    outall[base + p] = pce;
}

int pce_batch(Context *c, int variant, const PceBatch &bt, int nb, double *out, hipStream_t st) {
    const int h = c->h, w = c->w;
    const int64_t n = c->pixels();
    const int couples = (nb + 1) / 2;
    const size_t plane = ((sizeof(float2) * n) + 255) & ~size_t(255);
    const size_t pbytes = ((sizeof(PeakPart) * w) + 255) & ~size_t(255);
    const int64_t pstride = (int64_t)(plane / sizeof(float2));
    const size_t wbytes = ((sizeof(double) * kWinSlots * nb) + 255) & ~size_t(255);
    PRNU_TRY(c->ws_pce.ensure(2 * plane * couples + pbytes * nb + wbytes));
    char *base = static_cast<char *>(c->ws_pce.ptr);
    float2 *A = reinterpret_cast<float2 *>(base);                  // row-pass output, h x w per couple
    float2 *B = reinterpret_cast<float2 *>(base + plane * couples);  // its transpose, w x h per couple
    PeakPart *part = reinterpret_cast<PeakPart *>(base + 2 * plane * couples);
    double *wbuf = reinterpret_cast<double *>(base + 2 * plane * couples + pbytes * nb);
    const FftDesc &dw = fft_desc(c->fft_w);
    const FftDesc &dh = fft_desc(c->fft_h);
    const dim3 rgrid((unsigned)((h / 2 + 1 + 7) / 8 * 16), couples);  // mirror pairs of rows (pce_rows2_kernel)
    const size_t rsm = fft_row_smem(w);
    // rows up to 1024: both rows of a mirror pair in one workgroup (24 KiB of LDS)
    const bool mirror = w <= 1024;
    const dim3 mgrid((unsigned)(h / 2 + 1), couples);
    const size_t msm = 3 * sizeof(float2) * (size_t)w;
#define PRNU_PCE_ROWS(V, M)                                                                                          \
    do {                                                                                                             \
        if (w == 1024)                                                                                               \
            hipLaunchKernelGGL((pce_rows_mirror_wave_kernel<V>), mgrid, dim3(128), 0, st, dw.tw, h, nb, bt, A,       \
                               pstride);                                                                             \
        else if (mirror)                                                                                             \
            hipLaunchKernelGGL((pce_rows_mirror_kernel<V, M>), mgrid, dim3(kFftThreads), msm, st, dw, h, nb, bt, A, \
                               pstride);                                                                             \
        else                                                                                                         \
            hipLaunchKernelGGL((pce_rows2_kernel<V, M>), rgrid, dim3(kFftThreads), rsm, st, dw, h, nb, bt, A,       \
                               pstride);                                                                             \
    } while (0)
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
    const dim3 cgrid(w, couples);
    const size_t csm = fft_row_smem(h);
    const bool strips = h == 1024 && w % 8 == 0;
    if (!strips) PRNU_TRY(launch_transpose_c(A, B, h, w, couples, pstride, st));
    if (strips) {
        const int per = (w / 8 + 7) / 8;
        hipLaunchKernelGGL(pce_cols_strip_kernel, dim3(8 * per, couples), dim3(512), 0, st, dh.tw, w, nb, A, pstride,
                           part, per);
    } else if (h == 1024)
        hipLaunchKernelGGL(pce_cols_wave_kernel, cgrid, dim3(64), 0, st, dh.tw, w, nb, B, pstride, part);
    else if (h <= kFftSmallN)
        hipLaunchKernelGGL((pce_cols_kernel<kFftSmallN>), cgrid, dim3(kFftThreads), csm, st, dh, w, nb, B, pstride,
                           part);
    else if (h <= kFftMidN)
        hipLaunchKernelGGL((pce_cols_kernel<kFftMidN>), cgrid, dim3(kFftThreads), csm, st, dh, w, nb, B, pstride,
                           part);
    else
        hipLaunchKernelGGL((pce_cols_kernel<kFftMaxN>), cgrid, dim3(kFftThreads), csm, st, dh, w, nb, B, pstride,
                           part);
    PRNU_CHECK_LAUNCH();
    // window rows (and, for the reference variant, the last element's value)
    const dim3 wgrid_ref(nb, kSquare + 1), wgrid_txt(nb, kSquare), sgrid((nb + 63) / 64);
    hipLaunchKernelGGL(pce_peak_kernel, dim3(nb), dim3(256), 0, st, part, h, w, wbuf);
    PRNU_CHECK_LAUNCH();
    if (variant == PRNU_PCE_REFERENCE) {
        hipLaunchKernelGGL(pce_window_kernel<PRNU_PCE_REFERENCE>, wgrid_ref, dim3(256), 0, st, part, A, pstride, dh.tw,
                           h, w, wbuf);
        PRNU_CHECK_LAUNCH();
        hipLaunchKernelGGL(pce_score_kernel<PRNU_PCE_REFERENCE>, sgrid, dim3(64), 0, st, wbuf, nb, h, w, out,
                           bt.base);
    } else {
        hipLaunchKernelGGL(pce_window_kernel<PRNU_PCE_TEXTBOOK>, wgrid_txt, dim3(256), 0, st, part, A, pstride, dh.tw,
                           h, w, wbuf);
        PRNU_CHECK_LAUNCH();
        hipLaunchKernelGGL(pce_score_kernel<PRNU_PCE_TEXTBOOK>, sgrid, dim3(64), 0, st, wbuf, nb, h, w, out, bt.base);
    }
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
        // about 2 GiB of planes: 256 pairs of 1024x1024 (a sweep of 8 to 2048
        // pairs there peaks at 256; smaller batches pay each launch's ramp and
        // tail more often, larger ones lose the planes' L2 and cache reuse)
        const int64_t per = 8 * c->pixels();  // half of a couple's complex plane and of its transpose per pair
        batch = (int)std::max<int64_t>(2, std::min<int64_t>(512, (int64_t(2) << 30) / per)) & ~1;
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
