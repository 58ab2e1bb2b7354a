// NCC on gfx950: per-pattern FP64 statistics, the per-pair comparator of the
// reference (CSI/kernels/compare/normalizedcrosscorrelation.cu:85-178), and
// the all-pairs Gram on MFMA (FP32, split-BF16) with an FP64 epilogue.
//
// Gram design.  G = P P^T over K = h*w is tiny-M, huge-K, so the triangle of
// 256x256 tiles is split along K into S slabs.  One 512-thread workgroup
// computes one (tile, slab) partial in FP32 MFMA accumulators and writes it
// out; a second kernel sums the S partials of each element in FP64 in a fixed
// order (deterministic) and applies the NCC epilogue.  Workgroup ids are
// remapped so that the T tiles of one K slab land on the same XCD (workgroups
// are dealt round-robin over the 8 XCDs), which lets those T workgroups share
// the slab through that XCD's L2 instead of each pulling it from HBM.
#include "prnu_internal.hpp"

#include <algorithm>
#include <mutex>

namespace prnu {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kThreads = 256;

// ------------------------------------------------------------ reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline float wave_max(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
    return v;
}

// Deterministic block sum for blockDim.x == 256 (4 waves).
__device__ inline double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    return ((sh[0] + sh[1]) + (sh[2] + sh[3]));
}

__device__ inline float block_max(float v, float *sh) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    return fmaxf(fmaxf(sh[0], sh[1]), fmaxf(sh[2], sh[3]));
}

// ------------------------------------------------------- pattern stats
// grid (chunks, count), block 256.  Partial (sum, sumsq, absmax) per chunk.
struct StatPart {
    double s, ss;
    float am;
    float pad;
};

__global__ __launch_bounds__(kThreads) void stats_partial_kernel(const float *const *ptrs, int64_t n,
                                                                 int64_t chunk, StatPart *part) {
    __shared__ double shd[4];
    __shared__ float shf[4];
    const int p = blockIdx.y;
    const float *x = ptrs[p];
    const int64_t beg = (int64_t)blockIdx.x * chunk;
    const int64_t end = min(n, beg + chunk);
    double s = 0.0, ss = 0.0;
    float am = 0.0f;
    if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
        const int64_t vb = beg >> 2, ve = end >> 2;  // beg is a multiple of 4
        const float4 *x4 = reinterpret_cast<const float4 *>(x);
        for (int64_t i = vb + threadIdx.x; i < ve; i += kThreads) {
            float4 v = x4[i];
            s += ((double)v.x + (double)v.y) + ((double)v.z + (double)v.w);
            // products rounded to FP32, then accumulated in FP64 (reference semantics)
            float px = v.x * v.x, py = v.y * v.y, pz = v.z * v.z, pw = v.w * v.w;
            ss += ((double)px + (double)py) + ((double)pz + (double)pw);
            am = fmaxf(am, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
        }
        for (int64_t i = (ve << 2) + threadIdx.x; i < end; i += kThreads) {
            float v = x[i];
            float pv = v * v;
            s += v;
            ss += pv;
            am = fmaxf(am, fabsf(v));
        }
    } else {
        for (int64_t i = beg + threadIdx.x; i < end; i += kThreads) {
            float v = x[i];
            float pv = v * v;
            s += v;
            ss += pv;
            am = fmaxf(am, fabsf(v));
        }
    }
    s = block_sum(s, shd);
    ss = block_sum(ss, shd);
    am = block_max(am, shf);
    if (threadIdx.x == 0) {
        StatPart o;
        o.s = s;
        o.ss = ss;
        o.am = am;
        o.pad = 0.f;
        part[(int64_t)p * gridDim.x + blockIdx.x] = o;
    }
}

__global__ void stats_final_kernel(const StatPart *part, int chunks, int count, double *sum,
                                   double *sumsq, float *absmax) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    double s = 0.0, ss = 0.0;
    float am = 0.0f;
    for (int c = 0; c < chunks; ++c) {
        const StatPart q = part[(int64_t)p * chunks + c];
        s += q.s;
        ss += q.ss;
        am = fmaxf(am, q.am);
    }
    sum[p] = s;
    sumsq[p] = ss;
    if (absmax) absmax[p] = am;
}

// ------------------------------------------------------ per-pair NCC
// normalizedcrosscorrelation.cu:130-178 semantics: FP32 products, FP64 sums,
// deterministic two-level reduction.
constexpr int kPairBlocks = 1024;

__global__ __launch_bounds__(kThreads) void pair_sums_kernel(const float *__restrict__ x,
                                                             const float *__restrict__ y, int64_t n,
                                                             double *__restrict__ part) {
    __shared__ double sh[4];
    double sx = 0, sxx = 0, sy = 0, syy = 0, sxy = 0;
    const int64_t stride = (int64_t)gridDim.x * kThreads;
    const bool vec = ((reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(y)) & 15) == 0;
    int64_t tail_begin = 0;
    if (vec) {
        const int64_t n4 = n >> 2;
        const float4 *x4 = reinterpret_cast<const float4 *>(x);
        const float4 *y4 = reinterpret_cast<const float4 *>(y);
        for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n4; i += stride) {
            const float4 a = x4[i], b = y4[i];
            const float ax[4] = {a.x, a.y, a.z, a.w};
            const float bx[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float v = ax[e], w = bx[e];
                const float vv = v * v, ww = w * w, vw = v * w;
                sx += v;
                sxx += vv;
                sy += w;
                syy += ww;
                sxy += vw;
            }
        }
        tail_begin = n4 << 2;
    }
    for (int64_t i = tail_begin + (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        const float v = x[i], w = y[i];
        const float vv = v * v, ww = w * w, vw = v * w;
        sx += v;
        sxx += vv;
        sy += w;
        syy += ww;
        sxy += vw;
    }
    sx = block_sum(sx, sh);
    sxx = block_sum(sxx, sh);
    sy = block_sum(sy, sh);
    syy = block_sum(syy, sh);
    sxy = block_sum(sxy, sh);
    if (threadIdx.x == 0) {
        double *p = part + 5 * blockIdx.x;
        p[0] = sx;
        p[1] = sxx;
        p[2] = sy;
        p[3] = syy;
        p[4] = sxy;
    }
}

__global__ __launch_bounds__(kThreads) void pair_ncc_final_kernel(const double *part, int blocks,
                                                                  double *out) {
    __shared__ double sh[4];
    double a[5] = {0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < blocks; b += kThreads)
        for (int q = 0; q < 5; ++q) a[q] += part[5 * b + q];
    for (int q = 0; q < 5; ++q) a[q] = block_sum(a[q], sh);
    if (threadIdx.x == 0) {
        const double sdx = sqrt(a[1] - a[0] * a[0]);
        const double sdy = sqrt(a[3] - a[2] * a[2]);
        out[0] = a[4] / (sdx * sdy);
    }
}

// --------------------------------------------------------------- Gram
constexpr int kTile = 256;         // patterns per tile side
constexpr int kKT = 32;            // K per LDS stage
constexpr int kGramThreads = 512;  // 8 waves: 2 (rows) x 4 (cols), wave tile 128 x 64
constexpr int kRows = 2 * kTile;   // A rows then B rows in LDS

struct GramArgs {
    const float *const *ptrs;
    float *partials;
    int64_t K;
    int64_t kchunk;
    int M;
    int nt;  // tiles per side
    int T;   // tiles in the upper triangle, nt*(nt+1)/2
    int S;   // K slabs
    int xcd_remap;  // 1: the T tiles of a slab run on one XCD
};

// Linear index over the upper triangle (row-major, I <= J) -> (I, J).
__device__ inline int2 tri_tile(int t, int nt) {
    int I = 0;
    while (t >= nt - I) {
        t -= nt - I;
        ++I;
    }
    return make_int2(I, I + t);
}

__device__ inline uint16_t bf16_bits(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
__device__ inline float bf16_val(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

// LDS tiles are unpadded; 16-byte chunks are XOR-swizzled by row so that the
// 8 lanes an LDS cycle serves for ds_read_b128 (8 consecutive rows, same
// chunk) land on 8 different 16-byte bank groups.
//   bf16: row = 32 values = 64 B = 4 chunks, chunk' = chunk ^ ((row >> 1) & 3)
//   fp32: row = 32 values = 128 B = 8 chunks, chunk' = chunk ^ (row & 7)
__device__ __forceinline__ int bf_off(int row, int chunk) { return row * 32 + ((chunk ^ ((row >> 1) & 3)) << 3); }
__device__ __forceinline__ int f_off(int row, int chunk) { return row * 32 + ((chunk ^ (row & 7)) << 2); }

// Per-thread view of the rows it loads: row r_lo + 64*it, it = 0..7.
struct RowSet {
    const float *p[8];  // always a readable pattern row (absent rows alias a present one)
    bool valid[8];
};

// Main K loop of one (tile, slab) workgroup.  Double-buffered LDS, one
// barrier per stage: loads for stage s+1 are issued before the MFMAs of
// stage s and written to the other buffer after them.
// FAST: every row 16-byte aligned and K % 4 == 0, so a float4 at a multiple
// of 4 is wholly inside or wholly outside [k_begin, k_end); its address is
// clamped into the row and out-of-slab values are zeroed when stored.
template <int PREC, bool FAST>
__device__ __forceinline__ void gram_mainloop(uint32_t *lds, const RowSet &rs, int nrows, int bbase, int wrow0,
                                              int wcol0, int64_t k_begin, int64_t k_end, int64_t K,
                                              floatx4 (&acc)[8][4]) {
    constexpr bool kSplit = PREC == PRNU_PREC_SPLIT_BF16;
    constexpr int kBuf = kRows * 32;  // uint32 per stage buffer (64 KiB)
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int r_lo = tid >> 3;
    const int c4 = (tid & 7) * 4;
    float4 st[8];

    auto load = [&](int64_t k0) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            if (FAST) {
                const int64_t kk = k0 + c4;
                const int64_t kc = kk < K ? kk : K - 4;
                // global (not flat) loads: flat loads also count on lgkmcnt and
                // would make every LDS wait of the compute phase wait for them
                typedef const __attribute__((address_space(1))) floatx4 *gf4;
                const floatx4 v = *(gf4)(rs.p[it] + kc);
                st[it] = make_float4(v[0], v[1], v[2], v[3]);
            } else {
                float e[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int64_t kk = k0 + c4 + q;
                    e[q] = (rs.valid[it] && kk < k_end) ? rs.p[it][kk] : 0.f;
                }
                st[it] = make_float4(e[0], e[1], e[2], e[3]);
            }
        }
    };

    auto store = [&](uint32_t *buf, int64_t k0) {
        const bool kin = k0 + c4 < k_end;
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int r = r_lo + 64 * it;
            if (r < nrows) {
                float4 v = st[it];
                if (FAST && !(rs.valid[it] && kin)) v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (kSplit) {
                    const uint16_t hx = bf16_bits(v.x), hy = bf16_bits(v.y);
                    const uint16_t hz = bf16_bits(v.z), hw = bf16_bits(v.w);
                    const uint16_t lx = bf16_bits(v.x - bf16_val(hx));
                    const uint16_t ly = bf16_bits(v.y - bf16_val(hy));
                    const uint16_t lz = bf16_bits(v.z - bf16_val(hz));
                    const uint16_t lw = bf16_bits(v.w - bf16_val(hw));
                    uint16_t *hi = reinterpret_cast<uint16_t *>(buf);
                    uint16_t *lo = hi + kRows * 32;
                    const int off = bf_off(r, c4 >> 3) + (c4 & 4);
                    *reinterpret_cast<uint2 *>(hi + off) =
                        make_uint2((uint32_t)hx | ((uint32_t)hy << 16), (uint32_t)hz | ((uint32_t)hw << 16));
                    *reinterpret_cast<uint2 *>(lo + off) =
                        make_uint2((uint32_t)lx | ((uint32_t)ly << 16), (uint32_t)lz | ((uint32_t)lw << 16));
                } else {
                    *reinterpret_cast<float4 *>(reinterpret_cast<float *>(buf) + f_off(r, c4 >> 2)) = v;
                }
            }
        }
    };

    auto compute = [&](const uint32_t *buf) {
        const int frow = lane & 15;
        if (kSplit) {
            const uint16_t *hi = reinterpret_cast<const uint16_t *>(buf);
            const uint16_t *lo = hi + kRows * 32;
            const int chunk = lane >> 4;
            bf16x8 bh[4], bl[4];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const int off = bf_off(bbase + wcol0 + cb * 16 + frow, chunk);
                bh[cb] = *reinterpret_cast<const bf16x8 *>(hi + off);
                bl[cb] = *reinterpret_cast<const bf16x8 *>(lo + off);
            }
#pragma unroll
            for (int rb = 0; rb < 8; ++rb) {
                const int off = bf_off(wrow0 + rb * 16 + frow, chunk);
                const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(hi + off);
                const bf16x8 al = *reinterpret_cast<const bf16x8 *>(lo + off);
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[cb], acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[cb], acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[cb], acc[rb][cb], 0, 0, 0);
                }
            }
        } else {
            const float *f = reinterpret_cast<const float *>(buf);
#pragma unroll
            for (int kg = 0; kg < 2; ++kg) {
                const int chunk = kg * 4 + (lane >> 4);
                float4 bf[4];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    bf[cb] = *reinterpret_cast<const float4 *>(f + f_off(bbase + wcol0 + cb * 16 + frow, chunk));
#pragma unroll
                for (int rb = 0; rb < 8; ++rb) {
                    const float4 af = *reinterpret_cast<const float4 *>(f + f_off(wrow0 + rb * 16 + frow, chunk));
#pragma unroll
                    for (int cb = 0; cb < 4; ++cb) {
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af.x, bf[cb].x, acc[rb][cb], 0, 0, 0);
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af.y, bf[cb].y, acc[rb][cb], 0, 0, 0);
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af.z, bf[cb].z, acc[rb][cb], 0, 0, 0);
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(af.w, bf[cb].w, acc[rb][cb], 0, 0, 0);
                    }
                }
            }
        }
    };

    const int64_t nst = (k_end - k_begin + kKT - 1) / kKT;
    load(k_begin);
    store(lds, k_begin);
    __syncthreads();
    for (int64_t s = 0; s < nst; ++s) {
        const bool next = s + 1 < nst;
        const int64_t kn = k_begin + (s + 1) * kKT;
        if (next) load(kn);
        compute(lds + (s & 1) * kBuf);
        if (next) store(lds + ((s + 1) & 1) * kBuf, kn);
        __syncthreads();
    }
}

template <int PREC>
__global__ __launch_bounds__(kGramThreads, 1) void gram_tile_kernel(GramArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t lds[2 * kRows * 32];  // 2 x 64 KiB stage buffers

    // ---- workgroup -> (tile, slab).  Workgroups are dealt round-robin to the
    // 8 XCDs (bid mod 8); with xcd_remap the T tiles of slab s all have
    // bid = s mod 8 (mod 8) and run back to back on one XCD.
    const int bid = blockIdx.x;
    int t, s;
    if (a.xcd_remap) {
        const int local = bid >> 3;
        t = local % a.T;
        s = (local / a.T) * 8 + (bid & 7);
    } else {
        t = bid % a.T;
        s = bid / a.T;
    }
    if (s >= a.S) return;
    const int2 tile = tri_tile(t, a.nt);
    const bool diag = tile.x == tile.y;
    const int nrows = diag ? kTile : kRows;
    const int64_t k_begin = (int64_t)s * a.kchunk;
    const int64_t k_end = min(a.K, k_begin + a.kchunk);

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wrow0 = (wave >> 2) * 128;  // A rows of this wave
    const int wcol0 = (wave & 3) * 64;    // B rows of this wave
    const int bbase = diag ? 0 : kTile;   // B rows start in LDS

    RowSet rs;
    const float *any_row = a.ptrs[min(tile.x * kTile, a.M - 1)];
    int aligned = (a.K & 3) == 0;
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int r = (tid >> 3) + 64 * it;
        const int g = (r < kTile) ? tile.x * kTile + r : tile.y * kTile + (r - kTile);
        const bool v = r < nrows && g < a.M;
        rs.valid[it] = v;
        rs.p[it] = v ? a.ptrs[g] : any_row;
        if (reinterpret_cast<uintptr_t>(rs.p[it]) & 15) aligned = 0;
    }
    aligned = __syncthreads_and(aligned);

    floatx4 acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    if (aligned)
        gram_mainloop<PREC, true>(lds, rs, nrows, bbase, wrow0, wcol0, k_begin, k_end, a.K, acc);
    else
        gram_mainloop<PREC, false>(lds, rs, nrows, bbase, wrow0, wcol0, k_begin, k_end, a.K, acc);

    // ---- epilogue: FP32 partial tile, row = A index, col = B index
    float *out = a.partials + ((int64_t)t * a.S + s) * (kTile * kTile);
    const int ocol = lane & 15;
    const int orow = (lane >> 4) * 4;
#pragma unroll
    for (int rb = 0; rb < 8; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(wrow0 + rb * 16 + orow + r) * kTile + wcol0 + cb * 16 + ocol] = acc[rb][cb][r];
}

// Variant 1 of the split-BF16 tile kernel: 4 waves (one per SIMD, so each
// has the whole 512-entry register file), wave tile 128 x 128 with the
// accumulators in AGPRs, and two K stages of global loads in flight (the
// latency of a K stage that misses L2 is longer than one stage of MFMA work).
// Row pointers live in LDS.  Same LDS layout and math as gram_tile_kernel.
constexpr int kG4Threads = 256;

template <bool FAST>
__device__ __forceinline__ void g4_load(float4 (&st)[16], const float *const *rowp, const uint16_t *rowv, int tid,
                                        int64_t k0, int64_t k_end, int64_t K) {
    const int r_lo = tid >> 3, c4 = (tid & 7) * 4;
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const float *p = rowp[r_lo + 32 * it];
        if (FAST) {
            const int64_t kk = k0 + c4;
            const int64_t kc = kk < K ? kk : K - 4;
            typedef const __attribute__((address_space(1))) floatx4 *gf4;
            const floatx4 v = *(gf4)(p + kc);
            st[it] = make_float4(v[0], v[1], v[2], v[3]);
        } else {
            const bool ok = rowv[r_lo + 32 * it] != 0;
            float e[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t kk = k0 + c4 + q;
                e[q] = (ok && kk < k_end) ? p[kk] : 0.f;
            }
            st[it] = make_float4(e[0], e[1], e[2], e[3]);
        }
    }
}

template <bool FAST>
__device__ __forceinline__ void g4_store(const float4 (&st)[16], uint32_t *buf, const uint16_t *rowv, int nrows,
                                         int tid, int64_t k0, int64_t k_end) {
    const int r_lo = tid >> 3, c4 = (tid & 7) * 4;
    const bool kin = k0 + c4 < k_end;
    uint16_t *hi = reinterpret_cast<uint16_t *>(buf);
    uint16_t *lo = hi + kRows * 32;
#pragma unroll
    for (int it = 0; it < 16; ++it) {
        const int r = r_lo + 32 * it;
        if (r < nrows) {
            float4 v = st[it];
            if (FAST && !(rowv[r] && kin)) v = make_float4(0.f, 0.f, 0.f, 0.f);
            const uint16_t hx = bf16_bits(v.x), hy = bf16_bits(v.y);
            const uint16_t hz = bf16_bits(v.z), hw = bf16_bits(v.w);
            const uint16_t lx = bf16_bits(v.x - bf16_val(hx));
            const uint16_t ly = bf16_bits(v.y - bf16_val(hy));
            const uint16_t lz = bf16_bits(v.z - bf16_val(hz));
            const uint16_t lw = bf16_bits(v.w - bf16_val(hw));
            const int off = bf_off(r, c4 >> 3) + (c4 & 4);
            *reinterpret_cast<uint2 *>(hi + off) =
                make_uint2((uint32_t)hx | ((uint32_t)hy << 16), (uint32_t)hz | ((uint32_t)hw << 16));
            *reinterpret_cast<uint2 *>(lo + off) =
                make_uint2((uint32_t)lx | ((uint32_t)ly << 16), (uint32_t)lz | ((uint32_t)lw << 16));
        }
    }
}

__device__ __forceinline__ void g4_compute(const uint32_t *buf, int bbase, int wrow0, int wcol0, int lane,
                                           floatx4 (&acc)[8][8]) {
    const uint16_t *hi = reinterpret_cast<const uint16_t *>(buf);
    const uint16_t *lo = hi + kRows * 32;
    const int frow = lane & 15, chunk = lane >> 4;
    bf16x8 bh[8], bl[8];
#pragma unroll
    for (int cb = 0; cb < 8; ++cb) {
        const int off = bf_off(bbase + wcol0 + cb * 16 + frow, chunk);
        bh[cb] = *reinterpret_cast<const bf16x8 *>(hi + off);
        bl[cb] = *reinterpret_cast<const bf16x8 *>(lo + off);
    }
#pragma unroll
    for (int rb = 0; rb < 8; ++rb) {
        const int off = bf_off(wrow0 + rb * 16 + frow, chunk);
        const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(hi + off);
        const bf16x8 al = *reinterpret_cast<const bf16x8 *>(lo + off);
#pragma unroll
        for (int cb = 0; cb < 8; ++cb) {
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[cb], acc[rb][cb], 0, 0, 0);
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[cb], acc[rb][cb], 0, 0, 0);
            acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[cb], acc[rb][cb], 0, 0, 0);
        }
    }
}

template <bool FAST>
__device__ __forceinline__ void g4_mainloop(uint32_t *lds, const float *const *rowp, const uint16_t *rowv, int nrows,
                                            int bbase, int wrow0, int wcol0, int64_t k_begin, int64_t k_end,
                                            int64_t K, floatx4 (&acc)[8][8]) {
    constexpr int kBuf = kRows * 32;
    const int tid = threadIdx.x, lane = tid & 63;
    float4 s0[16], s1[16];
    const int64_t nst = (k_end - k_begin + kKT - 1) / kKT;
    // prologue: stage 0 into LDS buffer 0, stage 1 in flight in s1
    g4_load<FAST>(s0, rowp, rowv, tid, k_begin, k_end, K);
    if (nst > 1) g4_load<FAST>(s1, rowp, rowv, tid, k_begin + kKT, k_end, K);
    g4_store<FAST>(s0, lds, rowv, nrows, tid, k_begin, k_end);
    __syncthreads();
    // iteration s: LDS buf[s&1] holds stage s, stage s+1 is in flight in the
    // other register set; issue stage s+2, compute s, then land s+1.
    for (int64_t s = 0; s < nst; s += 2) {
        {
            const int64_t k2 = k_begin + (s + 2) * kKT;
            if (s + 2 < nst) g4_load<FAST>(s0, rowp, rowv, tid, k2, k_end, K);
            g4_compute(lds, bbase, wrow0, wcol0, lane, acc);
            if (s + 1 < nst) g4_store<FAST>(s1, lds + kBuf, rowv, nrows, tid, k_begin + (s + 1) * kKT, k_end);
            __syncthreads();
        }
        if (s + 1 < nst) {
            const int64_t k3 = k_begin + (s + 3) * kKT;
            if (s + 3 < nst) g4_load<FAST>(s1, rowp, rowv, tid, k3, k_end, K);
            g4_compute(lds + kBuf, bbase, wrow0, wcol0, lane, acc);
            if (s + 2 < nst) g4_store<FAST>(s0, lds, rowv, nrows, tid, k_begin + (s + 2) * kKT, k_end);
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(kG4Threads, 1) void gram_tile4_kernel(GramArgs a) {
    __shared__ __attribute__((aligned(16))) uint32_t lds[2 * kRows * 32];  // 2 x 64 KiB stage buffers
    __shared__ const float *rowp[kRows];                                     // row pointers of this tile
    __shared__ uint16_t rowv[kRows];                                         // row present
    const int bid = blockIdx.x;
    int t, s;
    if (a.xcd_remap) {
        const int local = bid >> 3;
        t = local % a.T;
        s = (local / a.T) * 8 + (bid & 7);
    } else {
        t = bid % a.T;
        s = bid / a.T;
    }
    if (s >= a.S) return;
    const int2 tile = tri_tile(t, a.nt);
    const bool diag = tile.x == tile.y;
    const int nrows = diag ? kTile : kRows;
    const int64_t k_begin = (int64_t)s * a.kchunk;
    const int64_t k_end = min(a.K, k_begin + a.kchunk);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wrow0 = (wave >> 1) * 128, wcol0 = (wave & 1) * 128;
    const int bbase = diag ? 0 : kTile;

    const float *any_row = a.ptrs[min(tile.x * kTile, a.M - 1)];
    int aligned = (a.K & 3) == 0;
    for (int r = tid; r < kRows; r += kG4Threads) {
        const int g = (r < kTile) ? tile.x * kTile + r : tile.y * kTile + (r - kTile);
        const bool v = r < nrows && g < a.M;
        const float *p = v ? a.ptrs[g] : any_row;
        rowp[r] = p;
        rowv[r] = v;
        if (reinterpret_cast<uintptr_t>(p) & 15) aligned = 0;
    }
    aligned = __syncthreads_and(aligned);

    floatx4 acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (aligned)
        g4_mainloop<true>(lds, rowp, rowv, nrows, bbase, wrow0, wcol0, k_begin, k_end, a.K, acc);
    else
        g4_mainloop<false>(lds, rowp, rowv, nrows, bbase, wrow0, wcol0, k_begin, k_end, a.K, acc);

    float *out = a.partials + ((int64_t)t * a.S + s) * (kTile * kTile);
    const int ocol = lane & 15, orow = (lane >> 4) * 4;
#pragma unroll
    for (int rb = 0; rb < 8; ++rb)
#pragma unroll
        for (int cb = 0; cb < 8; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(wrow0 + rb * 16 + orow + r) * kTile + wcol0 + cb * 16 + ocol] = acc[rb][cb][r];
}

// Sum the S partials of each element in FP64 (fixed order) and apply the NCC
// epilogue of normalizedcrosscorrelation.cu:121-123.  grid (256, T), block 256.
__global__ __launch_bounds__(kThreads) void gram_reduce_kernel(const float *__restrict__ partials, int S,
                                                               int nt, int M,
                                                               const double *__restrict__ sum,
                                                               const double *__restrict__ sumsq,
                                                               double *__restrict__ out) {
    const int t = blockIdx.y;
    const int e = blockIdx.x * kThreads + threadIdx.x;
    const int2 tile = tri_tile(t, nt);
    const int i = tile.x * kTile + (e >> 8);
    const int j = tile.y * kTile + (e & 255);
    if (i >= M || j >= M) return;
    if (tile.x == tile.y && i > j) return;
    const float *p = partials + (int64_t)t * S * (kTile * kTile) + e;
    double g = 0.0;
    for (int q = 0; q < S; ++q) g += (double)p[(int64_t)q * (kTile * kTile)];
    const double di = sqrt(sumsq[i] - sum[i] * sum[i]);
    const double dj = sqrt(sumsq[j] - sum[j] * sum[j]);
    const double v = g / (di * dj);
    out[(int64_t)i * M + j] = v;
    out[(int64_t)j * M + i] = v;
}

}  // namespace

// ================================================================ launchers
// K split of the statistics: about 64K values per block (a multiple of 4),
// at most 4096 blocks per pattern.
static void stats_split(int64_t n, int64_t *chunk_out, int64_t *chunks_out) {
    int64_t chunk = 65536;
    int64_t chunks = (n + chunk - 1) / chunk;
    if (chunks > 4096) {
        chunks = 4096;
        chunk = ((n + chunks - 1) / chunks + 3) & ~int64_t(3);
        chunks = (n + chunk - 1) / chunk;
    }
    *chunk_out = chunk;
    *chunks_out = chunks;
}

// Both kernels of the statistics over a partial-sum buffer of chunks * count
// entries (every entry is written by stats_partial_kernel before the final
// kernel reads it, so the buffer needs no clearing).
static int launch_stats_kernels(const float *const *d_ptrs, int count, int64_t n, int64_t chunk, int64_t chunks,
                                StatPart *part, double *d_sum, double *d_sumsq, float *d_absmax, hipStream_t st) {
    hipLaunchKernelGGL(stats_partial_kernel, dim3((unsigned)chunks, (unsigned)count), dim3(kThreads), 0, st,
                       d_ptrs, n, chunk, part);
    PRNU_CHECK_LAUNCH();
    hipLaunchKernelGGL(stats_final_kernel, dim3((count + 255) / 256), dim3(256), 0, st, part, (int)chunks, count,
                       d_sum, d_sumsq, d_absmax);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

int launch_pattern_stats(const float *const *d_ptrs, int count, int64_t n, double *d_sum,
                         double *d_sumsq, float *d_absmax, Workspace &ws, hipStream_t st) {
    PRNU_REQUIRE(count > 0 && n > 0, "pattern_stats: count and n must be positive");
    PRNU_REQUIRE(count <= 65535, "pattern_stats: count %d exceeds 65535", count);
    int64_t chunk, chunks;
    stats_split(n, &chunk, &chunks);
    PRNU_TRY(ws.ensure(sizeof(StatPart) * (size_t)chunks * count));
    return launch_stats_kernels(d_ptrs, count, n, chunk, chunks, static_cast<StatPart *>(ws.ptr), d_sum, d_sumsq,
                                d_absmax, st);
}

// ------------------------------------------ statistics without allocation
// The partial sums of prnu_pattern_stats_noalloc live in a fixed device
// array (each device holds its own copy, at its own address, until the
// process ends).  Host code never passes `stats_scratch` itself to a kernel
// or a copy: the array's name is a host-side stand-in there, so its device
// address is looked up once per device with hipGetSymbolAddress.
constexpr int64_t kStatsScratchParts = 131072;  // 3 MiB of 24-byte (sum, sumsq, absmax) partials
static __device__ __align__(256) StatPart stats_scratch[kStatsScratchParts];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static StatPart *g_scratch_ptr[kMaxDevices] = {};

static int stats_scratch_ptr(StatPart **out) {
    int dev = 0;
    PRNU_CHECK_HIP(hipGetDevice(&dev));
    PRNU_REQUIRE(dev >= 0 && dev < kMaxDevices, "pattern_stats_noalloc: device %d beyond %d devices", dev,
                 kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        PRNU_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(stats_scratch)));
        PRNU_REQUIRE(p, "pattern_stats_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<StatPart *>(p);
    }
    *out = g_scratch_ptr[dev];
    return PRNU_OK;
}

int launch_pair_ncc(const float *x, const float *y, int64_t n, double *d_out, Workspace &ws, hipStream_t st) {
    PRNU_TRY(ws.ensure(sizeof(double) * 5 * kPairBlocks));
    double *part = static_cast<double *>(ws.ptr);
    hipLaunchKernelGGL(pair_sums_kernel, dim3(kPairBlocks), dim3(kThreads), 0, st, x, y, n, part);
    PRNU_CHECK_LAUNCH();
    hipLaunchKernelGGL(pair_ncc_final_kernel, dim3(1), dim3(kThreads), 0, st, part, kPairBlocks, d_out);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

namespace {

struct GramPlan {
    int nt, T;
    int64_t S, kchunk;
    int xcd_remap;
    int variant;
};

// K split of the Gram: `slabs` per tile (0: about 5 workgroups per CU, >= 16
// stages per slab), partials capped at 2 GiB.
GramPlan plan_gram(const Context *c, int count, int precision) {
    GramPlan p;
    const int64_t K = c->pixels();
    p.nt = (count + kTile - 1) / kTile;
    p.T = p.nt * (p.nt + 1) / 2;
    const int64_t stages = (K + kKT - 1) / kKT;
    int64_t S = c->gram_cfg[precision].slabs;
    if (S <= 0) {
        S = std::max<int64_t>(1, (256 * 5 + p.T - 1) / p.T);
        S = std::min<int64_t>(S, std::max<int64_t>(1, stages / 16));
    }
    S = std::min<int64_t>(S, stages);
    const int64_t tile_bytes = (int64_t)kTile * kTile * sizeof(float);
    S = std::min<int64_t>(S, std::max<int64_t>(1, (int64_t(2) << 30) / (tile_bytes * p.T)));
    p.kchunk = ((stages + S - 1) / S) * kKT;
    p.S = (K + p.kchunk - 1) / p.kchunk;
    p.xcd_remap = c->gram_cfg[precision].xcd_remap;
    p.variant = precision == PRNU_PREC_SPLIT_BF16 ? c->gram_cfg[precision].variant : 0;
    return p;
}

int launch_gram_tiles(Context *c, const GramPlan &p, const float *const *d_ptrs, int count, int precision,
                      float *partials, hipStream_t st) {
    GramArgs a;
    a.ptrs = d_ptrs;
    a.partials = partials;
    a.K = c->pixels();
    a.kchunk = p.kchunk;
    a.M = count;
    a.nt = p.nt;
    a.T = p.T;
    a.S = (int)p.S;
    a.xcd_remap = p.xcd_remap;
    const int64_t S8 = p.xcd_remap ? (p.S + 7) / 8 * 8 : p.S;
    const dim3 grid((unsigned)(p.T * S8));
    if (precision == PRNU_PREC_SPLIT_BF16 && p.variant == 1)
        hipLaunchKernelGGL(gram_tile4_kernel, grid, dim3(kG4Threads), 0, st, a);
    else if (precision == PRNU_PREC_SPLIT_BF16)
        hipLaunchKernelGGL(gram_tile_kernel<PRNU_PREC_SPLIT_BF16>, grid, dim3(kGramThreads), 0, st, a);
    else
        hipLaunchKernelGGL(gram_tile_kernel<PRNU_PREC_FP32>, grid, dim3(kGramThreads), 0, st, a);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

}  // namespace

int launch_gram_ncc(Context *c, const float *const *d_ptrs, int count, int precision, const double *d_sum,
                    const double *d_sumsq, const float *d_absmax, double *d_ncc, hipStream_t st) {
    PRNU_REQUIRE(count >= 1, "gram: count must be >= 1");
    PRNU_REQUIRE(precision == PRNU_PREC_FP32 || precision == PRNU_PREC_SPLIT_BF16 || precision == PRNU_PREC_FP8,
                 "gram: unknown precision %d", precision);
    if (precision == PRNU_PREC_FP8) {
        // e4m3 copy of the set (rows padded to 128 bytes), then the FP8 triangle
        PRNU_REQUIRE(count <= 65535, "gram: FP8 supports at most 65535 patterns");
        const int64_t stride = fp8_row_stride(c->pixels());
        const size_t qbytes = (size_t)stride * count;
        PRNU_TRY(c->ws_fp8.ensure(qbytes + 256 + sizeof(float) * count));
        uint8_t *q = static_cast<uint8_t *>(c->ws_fp8.ptr);
        float *scale = reinterpret_cast<float *>(q + ((qbytes + 255) & ~size_t(255)));
        PRNU_TRY(launch_quantize_fp8(d_ptrs, count, c->pixels(), d_absmax, q, scale, st));
        return launch_fp8_ncc(c->ws_gram, 256, true, q, stride, count, scale, d_sum, d_sumsq, nullptr, 0, 0,
                              nullptr, nullptr, nullptr, stride, d_ncc, count, c->gram_cfg[precision].slabs, st, c);
    }
    const GramPlan p = plan_gram(c, count, precision);
    const size_t part_bytes = (size_t)kTile * kTile * sizeof(float) * p.T * p.S;
    PRNU_TRY(c->ws_gram.ensure(part_bytes));
    float *partials = static_cast<float *>(c->ws_gram.ptr);
    time_begin(c, st);
    PRNU_TRY(launch_gram_tiles(c, p, d_ptrs, count, precision, partials, st));
    time_end(c, st);
    hipLaunchKernelGGL(gram_reduce_kernel, dim3(kTile * kTile / kThreads, p.T), dim3(kThreads), 0, st, partials,
                       (int)p.S, p.nt, count, d_sum, d_sumsq, d_ncc);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

int gram_resolved_config(Context *c, int count, int precision, prnu_gram_config *out) {
    const GramPlan p = plan_gram(c, count, precision);
    out->slabs = (int32_t)p.S;
    out->xcd_remap = p.xcd_remap;
    out->variant = p.variant;
    return PRNU_OK;
}

// Time the tile kernel over candidate K splits and XCD mappings on the
// caller's data; keep the fastest in the context.
int autotune_gram(Context *c, const float *const *d_ptrs, int count, int precision, int reps,
                  prnu_gram_config *best, float *best_ms, hipStream_t st) {
    PRNU_REQUIRE(precision == PRNU_PREC_FP32 || precision == PRNU_PREC_SPLIT_BF16,
                 "gram_autotune: precision %d not supported", precision);
    PRNU_REQUIRE(count >= 1 && reps >= 1, "gram_autotune: count and reps must be positive");
    const prnu_gram_config saved = c->gram_cfg[precision];
    c->gram_cfg[precision] = prnu_gram_config{0, 1, 0};
    const int64_t s0 = plan_gram(c, count, precision).S;
    const int64_t cand_s[] = {s0, std::max<int64_t>(8, s0 / 2), s0 * 2, s0 * 4};
    hipEvent_t e0, e1;
    PRNU_CHECK_HIP(hipEventCreate(&e0));
    PRNU_CHECK_HIP(hipEventCreate(&e1));
    float best_t = 3.4e38f;
    prnu_gram_config best_cfg = saved;
    int status = PRNU_OK;
    const int nvar = precision == PRNU_PREC_SPLIT_BF16 ? 2 : 1;
    for (int var = 0; var < nvar; ++var)
    for (int64_t S : cand_s) {
        for (int xcd = 1; xcd >= 0 && status == PRNU_OK; --xcd) {
            c->gram_cfg[precision] = prnu_gram_config{(int32_t)S, xcd, var};
            const GramPlan p = plan_gram(c, count, precision);
            const size_t part_bytes = (size_t)kTile * kTile * sizeof(float) * p.T * p.S;
            status = c->ws_gram.ensure(part_bytes);
            if (status != PRNU_OK) break;
            float *partials = static_cast<float *>(c->ws_gram.ptr);
            status = launch_gram_tiles(c, p, d_ptrs, count, precision, partials, st);  // warm-up
            float tmin = 3.4e38f;
            for (int r = 0; r < reps && status == PRNU_OK; ++r) {
                (void)hipEventRecord(e0, st);
                status = launch_gram_tiles(c, p, d_ptrs, count, precision, partials, st);
                (void)hipEventRecord(e1, st);
                if (hipEventSynchronize(e1) != hipSuccess) {
                    set_error("gram_autotune: event wait failed");
                    status = PRNU_EDEVICE;
                    break;
                }
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                tmin = std::min(tmin, ms);
            }
            if (status == PRNU_OK && tmin < best_t) {
                best_t = tmin;
                best_cfg = prnu_gram_config{(int32_t)p.S, xcd, var};
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->gram_cfg[precision] = (status == PRNU_OK) ? best_cfg : saved;
    if (status != PRNU_OK) return status;
    if (best) *best = best_cfg;
    if (best_ms) *best_ms = best_t;
    return PRNU_OK;
}

}  // namespace prnu

using namespace prnu;

extern "C" {

int prnu_pattern_stats_noalloc(const float *const *d_patterns, int count, int64_t n, double *d_sum, double *d_sumsq,
                               float *d_absmax, prnu_stream stream) {
    PRNU_REQUIRE(d_patterns && d_sum && d_sumsq, "pattern_stats_noalloc: null pointer");
    PRNU_REQUIRE(count > 0 && n > 0, "pattern_stats_noalloc: count and n must be positive");
    PRNU_REQUIRE(count <= 65535, "pattern_stats_noalloc: count %d exceeds 65535", count);
    int64_t chunk, chunks;
    stats_split(n, &chunk, &chunks);
    PRNU_REQUIRE(chunks * count <= kStatsScratchParts,
                 "pattern_stats_noalloc: %d patterns of %lld values need %lld partial sums, the fixed scratch holds "
                 "%lld (at most %lld patterns of this length per call)",
                 count, (long long)n, (long long)(chunks * count), (long long)kStatsScratchParts,
                 (long long)(kStatsScratchParts / chunks));
    StatPart *part = nullptr;
    PRNU_TRY(stats_scratch_ptr(&part));
    hipStream_t st = (hipStream_t)stream;
    PRNU_TRY(launch_stats_kernels(d_patterns, count, n, chunk, chunks, part, d_sum, d_sumsq, d_absmax, st));
    if (!stream) PRNU_CHECK_HIP(hipStreamSynchronize(st));
    return PRNU_OK;
}

}  // extern "C"
