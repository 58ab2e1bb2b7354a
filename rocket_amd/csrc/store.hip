// Device-resident pattern store in a K-blocked layout, and the all-pairs Gram
// over it.
//
// Layout.  A pattern of n values is cut into nkb = ceil(n/32) blocks of 32.
// Block b of every slot is contiguous: byte offset (b * cap_pad + slot) * 128.
// A block is 128 bytes: split-BF16 stores hi = rne(x) (32 bf16) then
// lo = rne(x - hi) (32 bf16); FP32 stores the 32 floats.  The FP8 format
// instead cuts patterns into blocks of 128 values and stores e4m3(x / scale),
// scale = max|x| / 448 per pattern; the FP16 format into blocks of 64 values
// of rne(x / scale), scale = the power of two with max|x| / scale in
// [2^14, 2^15).  Values past n are 0.
//
// Why.  One K stage of a 256 x 256 Gram tile reads 32 values of 512 patterns.
// With row-major patterns those are 512 reads 4*n bytes apart, one page each;
// the per-CU TLB cannot hold them and half of all translations miss.  In this
// layout the same stage reads 2 (or 4) contiguous 32 KiB runs, the hi/lo
// split is done once at insert time instead of in every tile that reads the
// value, and the per-pattern FP64 statistics are taken in the same pass.
#include "prnu_internal.hpp"

#include <cstring>

#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

namespace prnu {

struct Store {
    int64_t n = 0, nkb = 0;
    int capacity = 0, cap_pad = 0;
    int format = PRNU_PREC_SPLIT_BF16;
    int device = 0;
    uint8_t *blocks = nullptr;
    double *sum = nullptr, *sumsq = nullptr;
    float *absmax = nullptr;
    float *scale = nullptr;  // FP8: max|x| / 448, FP16: a power of two, per slot
    Workspace ws_part;  // statistics partials
};

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int intx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

constexpr int kBlk = 32;        // values per block (split-BF16, FP32)
constexpr int kBlk8 = 128;      // values per block (FP8)
constexpr int kBlk16 = 64;      // values per block (FP16)
constexpr float kE4M3Max = 448.0f;
constexpr int kBlkBytes = 128;  // bytes per block
constexpr int kTileS = 256;
constexpr int kRowsS = 2 * kTileS;
constexpr int kThrS = 512;
constexpr int kMaxPlanGroups = 4096;
constexpr int kMaxStoreSlots = (1 << 24) - kTileS;  // a block column (cap_pad x 128 B) stays below 2^31 bytes

__device__ inline uint16_t bf16_bits_s(float v) { return __builtin_bit_cast(uint16_t, (__bf16)v); }
__device__ inline float bf16_val_s(uint16_t b) { return __builtin_bit_cast(float, (uint32_t)b << 16); }

// ------------------------------------------------------------------- insert
// block 256 = 8 rows x 32 blocks; grid (ceil(nkb/32), ceil(count/8)).
// Each thread packs one 32-value block and keeps its (sum, sumsq, absmax);
// the 32 threads of a row reduce them to one partial per (slot, grid.x).
struct StatPartS {
    double s, ss;
    float am;
    float pad;
};

template <int FMT>
__global__ __launch_bounds__(256) void store_put_kernel(const float *const *ptrs, int count, int64_t n, int64_t nkb,
                                                        uint8_t *blocks, int cap_pad, int first, StatPartS *part) {
    const int r = threadIdx.x >> 5, bl = threadIdx.x & 31;
    const int p = blockIdx.y * 8 + r;
    const int64_t b = (int64_t)blockIdx.x * 32 + bl;
    double s = 0.0, ss = 0.0;
    float am = 0.0f;
    if (p < count && b < nkb) {
        const float *x = ptrs[p];
        float v[kBlk];
        const int64_t k0 = b * kBlk;
        if (k0 + kBlk <= n && ((reinterpret_cast<uintptr_t>(x + k0) & 15) == 0)) {
#pragma unroll
            for (int q = 0; q < kBlk / 4; ++q) {
                const float4 f = reinterpret_cast<const float4 *>(x + k0)[q];
                v[4 * q] = f.x;
                v[4 * q + 1] = f.y;
                v[4 * q + 2] = f.z;
                v[4 * q + 3] = f.w;
            }
        } else {
#pragma unroll
            for (int e = 0; e < kBlk; ++e) v[e] = (k0 + e < n) ? x[k0 + e] : 0.f;
        }
#pragma unroll
        for (int e = 0; e < kBlk; ++e) {
            const float sq = v[e] * v[e];  // FP32 product, FP64 sum (reference semantics)
            s += v[e];
            ss += sq;
            am = fmaxf(am, fabsf(v[e]));
        }
        uint4 *dst = reinterpret_cast<uint4 *>(blocks + ((size_t)b * cap_pad + first + p) * kBlkBytes);
        if (FMT == PRNU_PREC_SPLIT_BF16) {
            uint32_t hi[16], lo[16];
#pragma unroll
            for (int e = 0; e < kBlk; e += 2) {
                const uint16_t h0 = bf16_bits_s(v[e]), h1 = bf16_bits_s(v[e + 1]);
                const uint16_t l0 = bf16_bits_s(v[e] - bf16_val_s(h0)), l1 = bf16_bits_s(v[e + 1] - bf16_val_s(h1));
                hi[e / 2] = (uint32_t)h0 | ((uint32_t)h1 << 16);
                lo[e / 2] = (uint32_t)l0 | ((uint32_t)l1 << 16);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) dst[q] = make_uint4(hi[4 * q], hi[4 * q + 1], hi[4 * q + 2], hi[4 * q + 3]);
#pragma unroll
            for (int q = 0; q < 4; ++q) dst[4 + q] = make_uint4(lo[4 * q], lo[4 * q + 1], lo[4 * q + 2], lo[4 * q + 3]);
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q)
                dst[q] = make_uint4(__float_as_uint(v[4 * q]), __float_as_uint(v[4 * q + 1]),
                                    __float_as_uint(v[4 * q + 2]), __float_as_uint(v[4 * q + 3]));
        }
    }
    // reduce over the 32 lanes of this row (lanes bl = 0..31 of one half-wave)
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) {
        s += __shfl_xor(s, off, 64);
        ss += __shfl_xor(ss, off, 64);
        am = fmaxf(am, __shfl_xor(am, off, 64));
    }
    if (bl == 0 && p < count) {
        StatPartS o;
        o.s = s;
        o.ss = ss;
        o.am = am;
        o.pad = 0.f;
        part[(int64_t)p * gridDim.x + blockIdx.x] = o;
    }
}

// FP8 insert: thread -> 32 values (a quarter of a 128-value block) of one
// pattern, quantised with the pattern's scale; grid (ceil(4*nkb/256), count).
__global__ __launch_bounds__(256) void store_put_fp8_kernel(const float *const *ptrs, int count, int64_t n,
                                                            int64_t nkb, uint8_t *blocks, int cap_pad, int first,
                                                            const float *absmax, float *scale) {
    const int p = blockIdx.y;
    const int64_t sub = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const float am = absmax[first + p];
    const float inv = am > 0.f ? kE4M3Max / am : 1.0f;
    if (sub == 0) scale[first + p] = am > 0.f ? am / kE4M3Max : 1.0f;
    if (sub >= 4 * nkb) return;
    const float *x = ptrs[p];
    const int64_t k0 = sub * 32;
    float v[32];
    if (k0 + 32 <= n && ((reinterpret_cast<uintptr_t>(x + k0) & 15) == 0)) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 f = reinterpret_cast<const float4 *>(x + k0)[q];
            v[4 * q] = f.x;
            v[4 * q + 1] = f.y;
            v[4 * q + 2] = f.z;
            v[4 * q + 3] = f.w;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 32; ++e) v[e] = (k0 + e < n) ? x[k0 + e] : 0.f;
    }
    int w[8];
#pragma unroll
    for (int g = 0; g < 8; ++g) {
        int t = 0;
        t = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * g] * inv, v[4 * g + 1] * inv, t, false);
        t = __builtin_amdgcn_cvt_pk_fp8_f32(v[4 * g + 2] * inv, v[4 * g + 3] * inv, t, true);
        w[g] = t;
    }
    uint4 *dst = reinterpret_cast<uint4 *>(blocks + ((size_t)(sub >> 2) * cap_pad + first + p) * kBlkBytes +
                                           (sub & 3) * 32);
    dst[0] = make_uint4(w[0], w[1], w[2], w[3]);
    dst[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// FP16 insert: thread -> 32 values (half a 64-value block) of one pattern,
// multiplied by the power of two 2^(15 - e), max|x| < 2^e, and rounded to
// nearest; grid (ceil(2*nkb/256), count).  The exponent is clamped from
// below so the multiplier of a tiny pattern stays a finite float.
__global__ __launch_bounds__(256) void store_put_fp16_kernel(const float *const *ptrs, int count, int64_t n,
                                                             int64_t nkb, uint8_t *blocks, int cap_pad, int first,
                                                             const float *absmax, float *scale) {
    const int p = blockIdx.y;
    const int64_t sub = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const float am = absmax[first + p];
    int e = 0;
    (void)frexpf(am, &e);  // am = m * 2^e, m in [0.5, 1): am < 2^e
    e = am > 0.f ? max(-100, e) : 15;
    const float mult = ldexpf(1.0f, 15 - e);
    if (sub == 0) scale[first + p] = ldexpf(1.0f, e - 15);
    if (sub >= 2 * nkb) return;
    const float *x = ptrs[p];
    const int64_t k0 = sub * 32;
    float v[32];
    if (k0 + 32 <= n && ((reinterpret_cast<uintptr_t>(x + k0) & 15) == 0)) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const float4 f = reinterpret_cast<const float4 *>(x + k0)[q];
            v[4 * q] = f.x;
            v[4 * q + 1] = f.y;
            v[4 * q + 2] = f.z;
            v[4 * q + 3] = f.w;
        }
    } else {
#pragma unroll
        for (int q = 0; q < 32; ++q) v[q] = (k0 + q < n) ? x[k0 + q] : 0.f;
    }
    uint32_t w[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const uint16_t h0 = __builtin_bit_cast(uint16_t, (_Float16)(v[2 * q] * mult));
        const uint16_t h1 = __builtin_bit_cast(uint16_t, (_Float16)(v[2 * q + 1] * mult));
        w[q] = (uint32_t)h0 | ((uint32_t)h1 << 16);
    }
    uint4 *dst = reinterpret_cast<uint4 *>(blocks + ((size_t)(sub >> 1) * cap_pad + first + p) * kBlkBytes +
                                           (sub & 1) * 64);
#pragma unroll
    for (int q = 0; q < 4; ++q) dst[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

// One 256-thread workgroup per pattern: strided partial sums, then a fixed
// tree over the block (deterministic).
__global__ __launch_bounds__(256) void store_stats_final_kernel(const StatPartS *part, int chunks, int count,
                                                                int first, double *sum, double *sumsq,
                                                                float *absmax) {
    __shared__ double sh_s[256], sh_ss[256];
    __shared__ float sh_am[256];
    const int p = blockIdx.x, tid = threadIdx.x;
    double s = 0.0, ss = 0.0;
    float am = 0.0f;
    for (int c = tid; c < chunks; c += 256) {
        const StatPartS q = part[(int64_t)p * chunks + c];
        s += q.s;
        ss += q.ss;
        am = fmaxf(am, q.am);
    }
    sh_s[tid] = s;
    sh_ss[tid] = ss;
    sh_am[tid] = am;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            sh_s[tid] += sh_s[tid + w];
            sh_ss[tid] += sh_ss[tid + w];
            sh_am[tid] = fmaxf(sh_am[tid], sh_am[tid + w]);
        }
        __syncthreads();
    }
    if (tid == 0) {
        sum[first + p] = sh_s[0];
        sumsq[first + p] = sh_ss[0];
        absmax[first + p] = sh_am[0];
    }
}

// -------------------------------------------------------------------- Gram
struct StoreGramArgs {
    const uint8_t *blocks;
    float *partials;
    int64_t cap_bytes;   // cap_pad * 128: distance between consecutive blocks of one slot
    int64_t nkb;
    int64_t kcb;         // blocks per slab
    int first, count;
    int nt, T, S;        // T: work items per slab
    int xcd_remap;
};

__device__ inline int2 tri_tile_s(int t, int nt) {
    int I = 0;
    while (t >= nt - I) {
        t -= nt - I;
        ++I;
    }
    return make_int2(I, I + t);
}

// 16-byte chunk c of LDS row r (128-byte rows), XOR-swizzled by row.
__device__ __forceinline__ int s_off(int row, int chunk) { return row * kBlkBytes + ((chunk ^ (row & 7)) << 4); }

constexpr int kBufS = kRowsS * kBlkBytes;  // 64 KiB per LDS stage

// Work items.  The LDS stage holds four quarters of 128 pattern rows; a
// wave multiplies the 128 rows of one quarter by 64 rows of another (a
// 128 x 64 block of the output), and waves 2k, 2k+1 together cover the
// 128 x 128 sub-tile of one quarter pair (unit k of the item).  The triangle
// of Q = ceil(count / 128) quarters has Q (Q + 1) / 2 units (quarter pairs
// qa <= qb); an item holds at most 4 units over at most 4 quarters, so a
// slab needs at least ceil(Q (Q + 1) / 8) items.  With m = floor(Q / 2)
// full 256-row tiles, per slab:
//  - m (m - 1) / 2 off-diagonal tiles (I, J): quarters 2I, 2I+1, 2J, 2J+1,
//    all four units of the pair;
//  - Q even: ceil(6 m / 8) diagonal packs.  A diagonal tile needs 3 units,
//    the 6 wave blocks (wr, wc) = (0,0) (0,1) (0,2) (0,3) (1,2) (1,3); the
//    6m such blocks are numbered e = 6 d + l and pack p gives block 8p + w
//    to wave w.  Eight consecutive blocks span at most two diagonal tiles
//    d0 = 8p / 6 and d0 + 1, whose quarters fill the LDS.
//  - Q odd (last quarter L = 2m alone in tile m): m "edge" items d with
//    quarters 2d, 2d+1, L and units (2d,2d) (2d,2d+1) (2d,L) (2d+1,L), then
//    ceil((m + 1) / 4) items of the remaining one-quarter units
//    (2j+1, 2j+1), j < m, and (L, L), four per item.
// Both reach the bound: m(m-1)/2 + ceil(6m/8) = ceil(Q(Q+1)/8) for even Q,
// m(m-1)/2 + m + ceil((m+1)/4) = ceil(Q(Q+1)/8) for odd Q.  The output of
// unit (qa, qb) is the sub-tile at rows 128 (qa & 1), columns 128 (qb & 1)
// of partial tile (qa / 2, qb / 2).
struct ItemWaves {
    int q[4];             // first pattern (relative to the range) of LDS rows [128 k, 128 k + 128)
    int nq;               // quarters loaded: LDS rows [0, 128 nq)
    int a_off, b_off;     // this wave's first A row / first B row in LDS
    int out_tile;         // triangle index of the tile this wave writes, -1: idle wave
    int out_off;          // offset of the wave's 128 x 64 block inside that tile
};

__device__ __forceinline__ int tri_index(int I, int J, int nt) { return I * nt - I * (I - 1) / 2 + (J - I); }

__device__ __forceinline__ void item_tiles(ItemWaves &r, int x, int y) {
    r.q[0] = x * kTileS;
    r.q[1] = x * kTileS + 128;
    r.q[2] = y * kTileS;
    r.q[3] = y * kTileS + 128;
}

// Item u of a slab of the triangle of `count` patterns (nt = ceil(count /
// 256) partial tiles), for wave `wave`.
__device__ __forceinline__ ItemWaves item_waves(int u, int count, int wave) {
    ItemWaves r;
    const int Q = (count + 127) >> 7, m = Q >> 1, nt = (count + kTileS - 1) / kTileS;
    const int wr = wave >> 2, wc = wave & 3;
    const int toff = m * (m - 1) / 2;
    if (u < toff) {
        int I = 0, t = u;
        while (t >= m - 1 - I) {
            t -= m - 1 - I;
            ++I;
        }
        const int J = I + 1 + t;
        item_tiles(r, I, J);
        r.nq = 4;
        r.a_off = wr * 128;
        r.b_off = kTileS + wc * 64;
        r.out_tile = tri_index(I, J, nt);
        r.out_off = wr * 128 * kTileS + wc * 64;
        return r;
    }
    const int p = u - toff;
    r.out_tile = -1;
    r.a_off = r.b_off = r.out_off = 0;
    if (!(Q & 1)) {
        const int d0 = (8 * p) / 6;
        const int dl = min(m - 1, (8 * p + 7) / 6);
        item_tiles(r, d0, dl > d0 ? d0 + 1 : d0);
        r.nq = dl > d0 ? 4 : 2;
        const int e = 8 * p + wave;
        if (e >= 6 * m) return r;
        const int d = e / 6, l = e - 6 * (e / 6);
        const int bwr = l < 4 ? 0 : 1;
        const int bwc = l < 4 ? l : l - 2;
        const int base = (d - d0) * kTileS;
        r.a_off = base + bwr * 128;
        r.b_off = base + bwc * 64;
        r.out_tile = tri_index(d, d, nt);
        r.out_off = bwr * 128 * kTileS + bwc * 64;
        return r;
    }
    const int k = wave >> 1, h = wave & 1;
    if (p < m) {  // edge item of tile d = p: quarters 2d, 2d+1, L
        const int d = p;
        r.q[0] = d * kTileS;
        r.q[1] = d * kTileS + 128;
        r.q[2] = r.q[3] = m * kTileS;
        r.nq = 3;
        const int sa = k == 3 ? 1 : 0;           // A quarter: 2d for units 0..2, 2d+1 for unit 3
        const int sb = k == 0 ? 0 : k == 1 ? 1 : 2;
        r.a_off = sa * 128;
        r.b_off = sb * 128 + h * 64;
        r.out_tile = tri_index(d, k < 2 ? d : m, nt);
        r.out_off = sa * 128 * kTileS + (k == 1 ? 128 : 0) + h * 64;
        return r;
    }
    // one-quarter units 4j' .. 4j'+3 of the list (2j+1, 2j+1) j < m, then (L, L)
    const int j0 = 4 * (p - m);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int j = min(j0 + i, m);
        r.q[i] = j * kTileS + (j < m ? 128 : 0);
    }
    r.nq = min(4, m + 1 - j0);
    const int j = j0 + k;
    if (j > m) return r;
    r.a_off = k * 128;
    r.b_off = k * 128 + h * 64;
    r.out_tile = tri_index(j, j, nt);
    r.out_off = (j < m ? 128 * kTileS + 128 : 0) + h * 64;
    return r;
}

// Items per slab of the triangle of `count` patterns (item_waves).
__host__ __device__ inline int tri_items_per_slab(int count) {
    const int Q = (count + 127) >> 7, m = Q >> 1;
    return m * (m - 1) / 2 + ((Q & 1) ? m + (m + 4) / 4 : (6 * m + 7) / 8);
}

// One work item over the K blocks [b_begin, b_end): LDS rows
// [128 k, 128 k + 128) hold the patterns w.q[k] + [0, 128) of the slots
// [first, first+count) for k = 0, 1 and of [first_b, first_b+count_b) for
// k = 2, 3 (the same range for a triangle), for k < w.nq; each wave
// multiplies its 128 A rows by its 64 B rows and writes its FP32 128 x 64
// partial block to `out` (row stride 256; null for an idle wave).  Ends
// with a barrier, so a caller may start the next item on the same LDS.
// VAR: 0 = direct-to-LDS loads, 1 = loads staged through registers,
// 2 = direct-to-LDS loads and raised wave priority over the MFMA block.
template <int FMT, int VAR>
__device__ __forceinline__ void gram_item(const uint8_t *__restrict__ blocks, int64_t cap_bytes, int first, int count,
                                          int first_b, int count_b, const ItemWaves &w, int64_t b_begin,
                                          int64_t b_end, float *__restrict__ out, uint8_t *lds,
                                          int tid = threadIdx.x) {
    constexpr bool kSplit = FMT == PRNU_PREC_SPLIT_BF16;
    constexpr bool kFp8 = FMT == PRNU_PREC_FP8;
    constexpr bool kF16 = FMT == PRNU_PREC_FP16;
    constexpr bool DMA = VAR != 1;
    constexpr bool kPrio = VAR == 2;
    const bool half = w.nq <= 2;  // an item of three quarters loads a fourth (it is never read)
    const int nrows = half ? kTileS : kRowsS;
    const bool active = out != nullptr;

    const int lane = tid & 63;
    const int wrow0 = w.a_off, wcol0 = w.b_off;

    // thread -> (row = (tid >> 3) + 64*it, chunk = tid & 7); rows past count
    // read a clamped slot and are zeroed on the LDS store
    const int r0 = tid >> 3, chunk = tid & 7;
    uint32_t valid = 0;
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int r = r0 + 64 * it;  // in quarter it >> 1
        const int loc = w.q[it >> 1] + r0 + 64 * (it & 1);
        valid |= (uint32_t)(r < nrows && loc < (it < 4 ? count : count_b)) << it;
    }
    // per-row byte offsets inside one block column (rows clamped into their
    // range: LDS rows [0,256) come from slots [first, first+count), rows
    // [256,512) from [first_b, first_b+count_b))
    auto row_off = [&](int it) -> int64_t {
        const int loc = w.q[it >> 1] + r0 + 64 * (it & 1);
        if (it < 4) return (int64_t)(first + min(loc, count - 1)) * kBlkBytes;
        return (int64_t)(first_b + min(loc, count_b - 1)) * kBlkBytes;
    };

    floatx4 acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    // Register-staged loads (DMA = false): global -> VGPR st[] -> LDS after
    // the MFMAs, rows past count zeroed.  Direct-to-LDS loads (DMA = true,
    // buffer_load_dwordx4 ... lds): each wave instruction fills 1 KiB of LDS,
    // lane i at byte 16*i, i.e. rows r0 + 64*it of this thread's wave, LDS
    // slot (tid & 7); the lane therefore fetches global chunk
    // (tid & 7) ^ (r & 7), which is the XOR swizzle s_off reads.  Rows past
    // count hold a clamped duplicate: they only feed output rows/columns
    // >= count, which the reduction never reads.  The buffer form (a
    // resource per block column, 32-bit row offsets; the column is below
    // 2 GiB) is tracked by the compiler under vmcnt alone, so the fragment
    // reads of the stage in LDS keep in-order lgkmcnt(k) waits; the global
    // form would force lgkmcnt(0) before every fragment read while a fill is
    // in flight.
    uint4 st[DMA ? 1 : 8];
    const int dchunk = chunk ^ (r0 & 7);
    uint32_t voff[DMA ? 8 : 1];
    if constexpr (DMA) {
#pragma unroll
        for (int it = 0; it < 8; ++it) voff[it] = (uint32_t)row_off(it) + (uint32_t)(dchunk << 4);
    }
    auto load_to = [&](int64_t b, uint8_t *buf) {
        if constexpr (DMA) {
            const __amdgpu_buffer_rsrc_t rsrc =
                __builtin_amdgcn_make_buffer_rsrc((void *)(blocks + b * cap_bytes), 0, (int)cap_bytes, 0x00020000);
            const int wbase = (tid >> 6) * 8;  // first row of this wave within a 64-row group
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                if (it < 4 || !half) {
                    typedef __attribute__((address_space(3))) void *lptr;
                    __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lptr)(buf + (64 * it + wbase) * kBlkBytes), 16,
                                                             voff[it], 0, 0, 0);
                }
            }
        } else {
            (void)buf;
            const uint8_t *col = blocks + b * cap_bytes + (chunk << 4);
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                if (it < 4 || !half) {
                    typedef const __attribute__((address_space(1))) unsigned int __attribute__((ext_vector_type(4))) *gu4;
                    const auto v = *(gu4)(col + row_off(it));
                    st[it] = make_uint4(v[0], v[1], v[2], v[3]);
                }
            }
        }
    };
    auto store = [&](uint8_t *buf) {
        if constexpr (DMA) {
            (void)buf;
            __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this thread's LDS fills have landed
        } else {
#pragma unroll
            for (int it = 0; it < 8; ++it) {
                const int r = r0 + 64 * it;
                if (r < nrows) {
                    uint4 v = st[it];
                    if (!((valid >> it) & 1)) v = make_uint4(0u, 0u, 0u, 0u);
                    *reinterpret_cast<uint4 *>(buf + s_off(r, chunk)) = v;
                }
            }
        }
    };
    auto compute = [&](const uint8_t *buf) {
        const int frow = lane & 15;
        if constexpr (kFp8) {
            // one 128-value block per stage: lane (row frow, k-group lane>>4)
            // holds bytes [32*(lane>>4), +32) = 16-byte chunks ch, ch+1
            const int ch = (lane >> 4) * 2;
            intx8 bq[4];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const int r = wcol0 + cb * 16 + frow;
                const int4 lo = *reinterpret_cast<const int4 *>(buf + s_off(r, ch));
                const int4 hi = *reinterpret_cast<const int4 *>(buf + s_off(r, ch + 1));
                bq[cb] = intx8{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            }
#pragma unroll
            for (int rb = 0; rb < 8; ++rb) {
                const int r = wrow0 + rb * 16 + frow;
                const int4 lo = *reinterpret_cast<const int4 *>(buf + s_off(r, ch));
                const int4 hi = *reinterpret_cast<const int4 *>(buf + s_off(r, ch + 1));
                const intx8 aq = intx8{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb][cb] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(aq, bq[cb], acc[rb][cb], 0, 0, 0,
                                                                                   127, 0, 127);
            }
        } else if (kF16) {
            // one 64-value block per stage as two K = 32 halves: chunks
            // 0..3 and 4..7, lane k-group ch = lane >> 4 in each
            const int ch = lane >> 4;
            halfx8 b0[4], b1[4];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const int r = wcol0 + cb * 16 + frow;
                b0[cb] = *reinterpret_cast<const halfx8 *>(buf + s_off(r, ch));
                b1[cb] = *reinterpret_cast<const halfx8 *>(buf + s_off(r, ch + 4));
            }
#pragma unroll
            for (int rb = 0; rb < 8; ++rb) {
                const int r = wrow0 + rb * 16 + frow;
                const halfx8 a0 = *reinterpret_cast<const halfx8 *>(buf + s_off(r, ch));
                const halfx8 a1 = *reinterpret_cast<const halfx8 *>(buf + s_off(r, ch + 4));
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b0[cb], acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, b1[cb], acc[rb][cb], 0, 0, 0);
                }
            }
        } else if (kSplit) {
            // Fragments are read in the order the MFMAs use them: first what
            // the first MFMAs need (A lo of row blocks 0 and 1, B hi), then
            // the rest of pair 0 of row blocks; the A fragments of pair p + 1
            // are issued after the first eight MFMAs of pair p.  LDS returns
            // reads in order, so each MFMA waits (lgkmcnt(k), k > 0) only for
            // the reads it uses, not for a burst of twelve; at most 15 reads
            // are ever outstanding (the counter's range).  Per accumulator
            // the products keep the order lo*hi, hi*lo, hi*hi; the two row
            // blocks of a pair alternate, so dependent MFMAs are eight apart.
            const int ch = lane >> 4;  // hi chunks 0..3, lo chunks 4..7
            auto rd = [&](int r, int c) { return *reinterpret_cast<const bf16x8 *>(buf + s_off(r, c)); };
            bf16x8 bh[4], bl[4], ah[4], al[4];  // A fragments: two pairs of row blocks in flight
            al[0] = rd(wrow0 + frow, ch + 4);
            al[1] = rd(wrow0 + 16 + frow, ch + 4);
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) bh[cb] = rd(wcol0 + cb * 16 + frow, ch);
            __builtin_amdgcn_sched_barrier(0);
            ah[0] = rd(wrow0 + frow, ch);
            ah[1] = rd(wrow0 + 16 + frow, ch);
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) bl[cb] = rd(wcol0 + cb * 16 + frow, ch + 4);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int pr = 0; pr < 4; ++pr) {
                const int c0 = (2 * pr) & 3, c1 = (2 * pr + 1) & 3;  // this pair's A slots
                const int rb0 = 2 * pr, rb1 = 2 * pr + 1;
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[c0], bh[cb], acc[rb0][cb], 0, 0, 0);
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[c1], bh[cb], acc[rb1][cb], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (pr + 1 < 4) {
                    const int n0 = (2 * pr + 2) & 3, n1 = (2 * pr + 3) & 3;
                    const int r = wrow0 + (2 * pr + 2) * 16 + frow;
                    al[n0] = rd(r, ch + 4);
                    al[n1] = rd(r + 16, ch + 4);
                    ah[n0] = rd(r, ch);
                    ah[n1] = rd(r + 16, ch);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c0], bl[cb], acc[rb0][cb], 0, 0, 0);
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c1], bl[cb], acc[rb1][cb], 0, 0, 0);
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb0][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c0], bh[cb], acc[rb0][cb], 0, 0, 0);
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    acc[rb1][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c1], bh[cb], acc[rb1][cb], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int kg = 0; kg < 2; ++kg) {
                const int ch = kg * 4 + (lane >> 4);
                float4 bf[4];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    bf[cb] = *reinterpret_cast<const float4 *>(buf + s_off(wcol0 + cb * 16 + frow, ch));
#pragma unroll
                for (int rb = 0; rb < 8; ++rb) {
                    const float4 af = *reinterpret_cast<const float4 *>(buf + s_off(wrow0 + rb * 16 + frow, ch));
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

    const int64_t nst = b_end - b_begin;
    load_to(b_begin, lds);
    store(lds);
    __syncthreads();
    for (int64_t q = 0; q < nst; ++q) {
        const bool next = q + 1 < nst;
        if (next) load_to(b_begin + q + 1, lds + ((q + 1) & 1) * kBufS);
        if (active) {
            if constexpr (kPrio) __builtin_amdgcn_s_setprio(1);
            compute(lds + (q & 1) * kBufS);
            if constexpr (kPrio) __builtin_amdgcn_s_setprio(0);
        }
        if (next) store(lds + ((q + 1) & 1) * kBufS);
        __syncthreads();
    }

    if (!active) return;
    const int ocol = lane & 15, orow = (lane >> 4) * 4;
#pragma unroll
    for (int rb = 0; rb < 8; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(rb * 16 + orow + r) * kTileS + cb * 16 + ocol] = acc[rb][cb][r];
}

// Run item u of slab s: partial tile t of slab s lives at partials + (t*S + s) * 256*256.
// ntb > 0: a rectangle of nt x ntb tiles between slots [first, first+count)
// (rows) and [first_b, first_b+count_b) (columns), item u = tile
// (u / ntb, u % ntb); ntb == 0: the triangle of [first, first+count).
__device__ __forceinline__ ItemWaves rect_waves(int u, int ntb, int wave) {
    ItemWaves r;
    const int wr = wave >> 2, wc = wave & 3;
    item_tiles(r, u / ntb, u - (u / ntb) * ntb);
    r.nq = 4;
    r.a_off = wr * 128;
    r.b_off = kTileS + wc * 64;
    r.out_tile = u;
    r.out_off = wr * 128 * kTileS + wc * 64;
    return r;
}

template <int FMT, int VAR>
__device__ __forceinline__ void run_item(const uint8_t *blocks, int64_t cap_bytes, int first, int count, int first_b,
                                         int count_b, int ntb, int S, int64_t nkb, int64_t kcb,
                                         float *partials, int u, int s, uint8_t *lds, int tid = threadIdx.x) {
    // wave-uniform: keep it in scalar registers
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    ItemWaves w = ntb > 0 ? rect_waves(u, ntb, wave) : item_waves(u, count, wave);
#pragma unroll
    for (int k = 0; k < 4; ++k) w.q[k] = __builtin_amdgcn_readfirstlane(w.q[k]);
    w.nq = __builtin_amdgcn_readfirstlane(w.nq);
    w.a_off = __builtin_amdgcn_readfirstlane(w.a_off >> 6) << 6;  // multiples of 64 rows: the swizzle
    w.b_off = __builtin_amdgcn_readfirstlane(w.b_off >> 6) << 6;  // term (row & 7) stays per lane
    w.out_tile = __builtin_amdgcn_readfirstlane(w.out_tile);
    w.out_off = __builtin_amdgcn_readfirstlane(w.out_off);
    const int64_t b_begin = (int64_t)s * kcb;
    const int64_t b_end = min(nkb, b_begin + kcb);
    float *out = w.out_tile < 0 ? nullptr
                                : partials + ((int64_t)w.out_tile * S + s) * (kTileS * kTileS) + w.out_off;
    gram_item<FMT, VAR>(blocks, cap_bytes, first, count, first_b, count_b, w, b_begin, b_end, out, lds, tid);
}

// ------------------------------------------------ 4-wave split-BF16 item
// Same tile and LDS stage as gram_item, run by 4 waves (256 threads) of
// 128 x 128 each: a wave holds 8 x 8 accumulator tiles (256 registers, in
// AGPRs) with the full 512-register file of one wave per SIMD, and reads
// 32 fragments per stage for 192 MFMAs (the 8-wave layout reads 24 for 96).
// Diagonal tiles need 3 of their 4 wave blocks, (0,0) (0,1) (1,1); pack p
// gives blocks 4p .. 4p+3 of the 3*nt to its waves (at most two diagonal
// tiles per pack).
constexpr int kThr4 = 256;

struct ItemWaves4 {
    int rows_x, rows_y;   // tiles whose 256 rows go to LDS rows [0,256) / [256,512); rows_y < 0: none
    int a_off, b_off;     // this wave's first A row / first B row in LDS
    int out_tile;         // triangle index of the tile this wave writes, -1: idle wave
    int out_off;          // offset of the wave's 128 x 128 block inside that tile
};

__device__ __forceinline__ ItemWaves4 item_waves4(int u, int nt, int wave) {
    ItemWaves4 r;
    const int wr = wave >> 1, wc = wave & 1;
    const int toff = nt * (nt - 1) / 2;
    if (u < toff) {
        int I = 0, t = u;
        while (t >= nt - 1 - I) {
            t -= nt - 1 - I;
            ++I;
        }
        const int J = I + 1 + t;
        r.rows_x = I;
        r.rows_y = J;
        r.a_off = wr * 128;
        r.b_off = kTileS + wc * 128;
        r.out_tile = tri_index(I, J, nt);
        r.out_off = wr * 128 * kTileS + wc * 128;
        return r;
    }
    const int p = u - toff;
    const int d0 = (4 * p) / 3;
    const int dl = min(nt - 1, (4 * p + 3) / 3);
    r.rows_x = d0;
    r.rows_y = dl > d0 ? d0 + 1 : -1;
    const int e = 4 * p + wave;
    if (e >= 3 * nt) {
        r.a_off = r.b_off = 0;
        r.out_tile = -1;
        r.out_off = 0;
        return r;
    }
    const int d = e / 3, l = e - 3 * (e / 3);
    const int bwr = l == 2 ? 1 : 0, bwc = l == 0 ? 0 : 1;
    const int base = (d - d0) * kTileS;
    r.a_off = base + bwr * 128;
    r.b_off = base + bwc * 128;
    r.out_tile = tri_index(d, d, nt);
    r.out_off = bwr * 128 * kTileS + bwc * 128;
    return r;
}

template <int FMT>
__global__ __launch_bounds__(kThr4, 1) void store_gram4_kernel(StoreGramArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[2 * kBufS];
    const int bid = blockIdx.x;
    int u, sl;
    if (a.xcd_remap) {
        const int local = bid >> 3;
        u = local % a.T;
        sl = (local / a.T) * 8 + (bid & 7);
    } else {
        u = bid % a.T;
        sl = bid / a.T;
    }
    if (sl >= a.S) return;
    ItemWaves4 w = item_waves4(u, a.nt, __builtin_amdgcn_readfirstlane(threadIdx.x >> 6));
    const int rows_x = __builtin_amdgcn_readfirstlane(w.rows_x);
    const int rows_y = __builtin_amdgcn_readfirstlane(w.rows_y);
    const int a_off = __builtin_amdgcn_readfirstlane(w.a_off >> 6) << 6;
    const int b_off = __builtin_amdgcn_readfirstlane(w.b_off >> 6) << 6;
    const int out_tile = __builtin_amdgcn_readfirstlane(w.out_tile);
    const int out_off = __builtin_amdgcn_readfirstlane(w.out_off);
    const int64_t b_begin = (int64_t)sl * a.kcb;
    const int64_t b_end = min(a.nkb, b_begin + a.kcb);
    const bool half = rows_y < 0;
    const int ty = half ? rows_x : rows_y;
    const int tid = threadIdx.x, lane = tid & 63;

    // direct-to-LDS loads: thread -> rows (tid >> 3) + 32*it, it < 16; one
    // wave instruction fills 8 consecutive rows (1 KiB)
    const int r0 = tid >> 3, chunk = tid & 7;
    const int dchunk = chunk ^ (r0 & 7);
    // byte offsets inside one block column fit 32 bits (capacity * 128 < 2^31)
    auto row_off = [&](int it) -> uint32_t {
        const int r = r0 + 32 * it;
        const int loc = r < kTileS ? rows_x * kTileS + r : ty * kTileS + (r - kTileS);
        return (uint32_t)(a.first + min(loc, a.count - 1)) * (uint32_t)kBlkBytes;
    };
    // buffer loads with 32-bit per-row offsets: the resource points at the
    // block column of the stage (scalar), the offsets stay fixed
    uint32_t voff[16];
#pragma unroll
    for (int it = 0; it < 16; ++it) voff[it] = row_off(it) + (uint32_t)(dchunk << 4);
    const int wbase = (tid >> 6) * 8;
    auto load_to = [&](int64_t b, uint8_t *buf) {
        __amdgpu_buffer_rsrc_t rsrc =
            __builtin_amdgcn_make_buffer_rsrc((void *)(a.blocks + b * a.cap_bytes), (short)0,
                                              (int)(a.cap_bytes), 0x00020000);
#pragma unroll
        for (int it = 0; it < 16; ++it) {
            if (it < 8 || !half) {
                typedef __attribute__((address_space(3))) void *lptr;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lptr)(buf + (32 * it + wbase) * kBlkBytes), 16,
                                                         voff[it], 0, 0, 0);
            }
        }
    };

    floatx4 acc[8][8];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 8; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const bool active = out_tile >= 0;
    const int frow = lane & 15, ch = lane >> 4;
    auto compute = [&](const uint8_t *buf) {
        if constexpr (FMT == PRNU_PREC_FP8) {
            // one 128-value block per stage; lane (row frow, k-group lane>>4)
            // holds 32 bytes = 16-byte chunks 2*(lane>>4), +1
            const int c8 = (lane >> 4) * 2;
            intx8 bq[8];
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) {
                const int r = b_off + cb * 16 + frow;
                const int4 lo = *reinterpret_cast<const int4 *>(buf + s_off(r, c8));
                const int4 hi = *reinterpret_cast<const int4 *>(buf + s_off(r, c8 + 1));
                bq[cb] = intx8{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
            }
#pragma unroll
            for (int rb = 0; rb < 8; ++rb) {
                const int r = a_off + rb * 16 + frow;
                const int4 lo = *reinterpret_cast<const int4 *>(buf + s_off(r, c8));
                const int4 hi = *reinterpret_cast<const int4 *>(buf + s_off(r, c8 + 1));
                const intx8 aq = intx8{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
                for (int cb = 0; cb < 8; ++cb)
                    acc[rb][cb] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(aq, bq[cb], acc[rb][cb], 0, 0, 0,
                                                                                   127, 0, 127);
            }
        } else if constexpr (FMT == PRNU_PREC_FP16) {
            // one 64-value block per stage as two K = 32 halves (chunks ch, ch + 4)
            halfx8 b0[8], b1[8];
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) {
                const int r = b_off + cb * 16 + frow;
                b0[cb] = *reinterpret_cast<const halfx8 *>(buf + s_off(r, ch));
                b1[cb] = *reinterpret_cast<const halfx8 *>(buf + s_off(r, ch + 4));
            }
#pragma unroll
            for (int rb = 0; rb < 8; ++rb) {
                const int r = a_off + rb * 16 + frow;
                const halfx8 a0 = *reinterpret_cast<const halfx8 *>(buf + s_off(r, ch));
                const halfx8 a1 = *reinterpret_cast<const halfx8 *>(buf + s_off(r, ch + 4));
#pragma unroll
                for (int cb = 0; cb < 8; ++cb) {
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0, b0[cb], acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1, b1[cb], acc[rb][cb], 0, 0, 0);
                }
            }
        } else {
            // Reads in use order: A lo of row block 0 and B hi first (the
            // first eight MFMAs), then A hi and B lo, then each next row
            // block's A one row block (24 MFMAs) ahead; in-order LDS returns
            // give each MFMA an lgkmcnt(k) wait for just its operands.  Per
            // accumulator the products stay lo*hi, hi*lo, hi*hi, eight MFMAs
            // apart.
            auto rd = [&](int r, int c) { return *reinterpret_cast<const bf16x8 *>(buf + s_off(r, c)); };
            bf16x8 bh[8], bl[8], ah[2], al[2];
            al[0] = rd(a_off + frow, ch + 4);
#pragma unroll
            for (int cb = 0; cb < 8; ++cb) bh[cb] = rd(b_off + cb * 16 + frow, ch);
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int rb = 0; rb < 8; ++rb) {
                const int c = rb & 1;
#pragma unroll
                for (int cb = 0; cb < 8; ++cb)
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[c], bh[cb], acc[rb][cb], 0, 0, 0);
                __builtin_amdgcn_sched_barrier(0);
                if (rb == 0) {
                    ah[0] = rd(a_off + frow, ch);
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int cb = 0; cb < 8; ++cb) {
                        bl[cb] = rd(b_off + cb * 16 + frow, ch + 4);
                        __builtin_amdgcn_sched_barrier(0);  // in use order
                    }
                }
                if (rb + 1 < 8) {
                    const int r = a_off + (rb + 1) * 16 + frow;
                    al[c ^ 1] = rd(r, ch + 4);
                    ah[c ^ 1] = rd(r, ch);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int cb = 0; cb < 8; ++cb)
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c], bl[cb], acc[rb][cb], 0, 0, 0);
#pragma unroll
                for (int cb = 0; cb < 8; ++cb)
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c], bh[cb], acc[rb][cb], 0, 0, 0);
            }
        }
    };

    const int64_t nst = b_end - b_begin;
    load_to(b_begin, lds);
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0)
    __syncthreads();
    for (int64_t q = 0; q < nst; ++q) {
        const bool next = q + 1 < nst;
        if (next) load_to(b_begin + q + 1, lds + ((q + 1) & 1) * kBufS);
        // idle waves (a partial diagonal pack) compute too and write nothing:
        // a branch around the MFMAs makes the compiler copy the AGPR
        // accumulators through VGPRs every stage
        compute(lds + (q & 1) * kBufS);
        if (next) __builtin_amdgcn_s_waitcnt(0x0f70);
        __syncthreads();
    }
    if (!active) return;
    float *out = a.partials + ((int64_t)out_tile * a.S + sl) * (kTileS * kTileS) + out_off;
    const int ocol = lane & 15, orow = (lane >> 4) * 4;
#pragma unroll
    for (int rb = 0; rb < 8; ++rb)
#pragma unroll
        for (int cb = 0; cb < 8; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r) out[(rb * 16 + orow + r) * kTileS + cb * 16 + ocol] = acc[rb][cb][r];
}

// One launch = one work item per workgroup.  With xcd_remap, the items of
// a slab run on one XCD (workgroup ids are dealt to the 8 XCDs round-robin).
template <int FMT, int VAR>
__global__ __launch_bounds__(kThrS, 1) void store_gram_kernel(StoreGramArgs a) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[2 * kBufS];
    const int bid = blockIdx.x;
    int u, s;
    if (a.xcd_remap) {
        const int local = bid >> 3;
        u = local % a.T;
        s = (local / a.T) * 8 + (bid & 7);
    } else {
        u = bid % a.T;
        s = bid / a.T;
    }
    if (s >= a.S) return;
    run_item<FMT, VAR>(a.blocks, a.cap_bytes, a.first, a.count, a.first, a.count, 0, a.S, a.nkb, a.kcb,
                       a.partials, u, s, lds);
}

// ------------------------------------------------------------ grouped plan
// Several pattern groups (each all-pairs within itself) in one launch over
// their concatenated work-item lists.  Group g (T items per slab) owns the
// virtual ids [v_begin, v_begin + T*S8), with S8 = S rounded up to 8 and
// v_begin a multiple of 8; id v maps to item ((v - v_begin) >> 3) % T and slab
// ((v - v_begin) >> 3) / T * 8 + (v & 7).  Every slab has the same number of
// K blocks in every group, so all items cost the same; the dispatcher hands
// workgroups to CUs in id order as they free up, which is the work queue.
// v & 7 is the XCD, so the T items of a slab share that XCD's L2.
struct GroupDev {
    const uint8_t *blocks;
    float *partials;
    int64_t cap_bytes, nkb, kcb, v_begin;
    int first, count, nt, T, S;
    int first_b, count_b, ntb;  // ntb > 0: rectangle rows x [first_b, first_b+count_b)
};

template <int FMT, int VAR>
__global__ __launch_bounds__(kThrS, 1) void store_gram_groups_kernel(const GroupDev *__restrict__ groups,
                                                                    int ngroups) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[2 * kBufS];
    const int64_t v = blockIdx.x;
    int lo = 0, hi = ngroups - 1;  // last group with v_begin <= v
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (groups[mid].v_begin <= v) lo = mid;
        else hi = mid - 1;
    }
    const GroupDev &g = groups[lo];
    const int64_t vl = v - g.v_begin;
    const int local = (int)(vl >> 3);
    const int u = local % g.T;
    const int s = (local / g.T) * 8 + (int)(vl & 7);
    if (s >= g.S) return;  // padding item
    run_item<FMT, VAR>(g.blocks, g.cap_bytes, g.first, g.count, g.first_b, g.count_b, g.ntb, g.S, g.nkb,
                       g.kcb, g.partials, u, s, lds);
}

// Persistent form: a grid of one workgroup per CU walks the item list.
// Items v with v & 7 == x form queue x, taken in order through the counter
// queue[16 * x] by the workgroups with blockIdx.x & 7 == x (the XCD the
// dispatcher gives them), so the items of one slab stay on one XCD's L2 as
// in the one-item-per-workgroup launch, and a workgroup that finishes early
// takes the next item instead of waiting for the launch's tail.  Every
// workgroup leaves once its queue passes v_total; the counters must be zero
// at launch (store_queue_reset_kernel).
// With items of equal cost this form gains nothing over one item per
// workgroup and measures 1-4 % slower: the hardware dispatcher starts the
// items of a slab together (in order, gated by the busiest XCD), so they
// read each K block in step and the L2 serves it once, while persistent
// workgroups drift apart item by item (more HBM reads, a lower clock under
// the power cap).  The plan autotuner times both.
// The item number taken by thread 0 reaches the other waves through the
// first word of the stage buffers (free between items): with a second LDS
// variable the compiler can no longer tell the stage being read from the one
// the global-to-LDS loads are filling, and waits for those loads (vmcnt(0))
// before every stage's fragment reads, which serialises loads and MFMAs
// (items ran 21 % longer).
template <int FMT>
__global__ __launch_bounds__(kThrS, 1) void store_gram_groups_persistent_kernel(const GroupDev *__restrict__ groups,
                                                                               int ngroups, int64_t v_total,
                                                                               int *__restrict__ queue) {
    __shared__ __attribute__((aligned(16))) uint8_t lds[2 * kBufS];
    int *taken = reinterpret_cast<int *>(lds);
    const int x = blockIdx.x & 7;
    for (;;) {
        if (threadIdx.x == 0) *taken = atomicAdd(queue + 16 * x, 1);
        __syncthreads();
        const int64_t v = (int64_t)__builtin_amdgcn_readfirstlane(*taken) * 8 + x;
        __syncthreads();  // every thread has read the word before the item's loads overwrite it
        if (v >= v_total) break;
        int lo = 0, hi = ngroups - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (groups[mid].v_begin <= v) lo = mid;
            else hi = mid - 1;
        }
        const GroupDev &g = groups[lo];
        const int64_t vl = v - g.v_begin;
        const int local = (int)(vl >> 3);
        const int u = local % g.T;
        const int s = (local / g.T) * 8 + (int)(vl & 7);
        if (s >= g.S) continue;  // padding item (uniform across the workgroup)
        // the thread id is opaque per item, so its address terms are not
        // hoisted out of this loop (they would stay live across every item)
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        run_item<FMT, 0>(g.blocks, g.cap_bytes, g.first, g.count, g.first_b, g.count_b, g.ntb, g.S, g.nkb,
                         g.kcb, g.partials, u, s, lds, tid);
    }
}

__global__ void store_queue_reset_kernel(int *queue) { queue[threadIdx.x] = 0; }

// FP64 sum of the S partials + NCC epilogue (normalizedcrosscorrelation.cu:121-123).
// Four consecutive elements of a partial tile per thread: one 16-byte load
// per slab, four FP64 sums, each in slab order (the same order as one
// element per thread, so the results do not depend on this grouping).
// grid (kTileS * kTileS / 1024, tiles), 256 threads.
__device__ __forceinline__ void sum_slabs4(const float *__restrict__ p, int S, double g[4]) {
    g[0] = g[1] = g[2] = g[3] = 0.0;
    const float4 *q4 = reinterpret_cast<const float4 *>(p);
    constexpr int kStride = kTileS * kTileS / 4;
#pragma unroll 8
    for (int q = 0; q < S; ++q) {
        const float4 v = q4[(int64_t)q * kStride];
        g[0] += (double)v.x;
        g[1] += (double)v.y;
        g[2] += (double)v.z;
        g[3] += (double)v.w;
    }
}

__global__ __launch_bounds__(256) void store_reduce_kernel(const float *__restrict__ partials, int S, int nt,
                                                           int count, const double *__restrict__ sum,
                                                           const double *__restrict__ sumsq,
                                                           const float *__restrict__ scale,
                                                           double *__restrict__ out, int64_t ld) {
    const int t = blockIdx.y;
    const int e = (blockIdx.x * 256 + threadIdx.x) * 4;
    const int2 tile = tri_tile_s(t, nt);
    const int i = tile.x * kTileS + (e >> 8);
    const int j0 = tile.y * kTileS + (e & 255);
    if (i >= count || j0 >= count) return;
    if (tile.x == tile.y && i > j0 + 3) return;
    double g[4];
    sum_slabs4(partials + (int64_t)t * S * (kTileS * kTileS) + e, S, g);
    const double di = sqrt(sumsq[i] - sum[i] * sum[i]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int j = j0 + c;
        if (j >= count || (tile.x == tile.y && i > j)) continue;
        double v = g[c];
        if (scale) v *= (double)scale[i] * (double)scale[j];  // FP8 / FP16: undo the per-pattern scales
        const double dj = sqrt(sumsq[j] - sum[j] * sum[j]);
        v = v / (di * dj);
        out[(int64_t)i * ld + j] = v;
        out[(int64_t)j * ld + i] = v;
    }
}

// Rectangle: element (i, j) of tile t = (I, J) is rows i of [first_a..) and
// column j of [first_b..); no mirroring; out[i * ld + j].
__global__ __launch_bounds__(256) void store_reduce_rect_kernel(const float *__restrict__ partials, int S, int ntb,
                                                                int count_a, int count_b,
                                                                const double *__restrict__ sum_a,
                                                                const double *__restrict__ sumsq_a,
                                                                const double *__restrict__ sum_b,
                                                                const double *__restrict__ sumsq_b,
                                                                const float *__restrict__ scale_a,
                                                                const float *__restrict__ scale_b,
                                                                double *__restrict__ out, int64_t ld) {
    const int t = blockIdx.y;
    const int e = (blockIdx.x * 256 + threadIdx.x) * 4;
    const int I = t / ntb, J = t - (t / ntb) * ntb;
    const int i = I * kTileS + (e >> 8);
    const int j0 = J * kTileS + (e & 255);
    if (i >= count_a || j0 >= count_b) return;
    double g[4];
    sum_slabs4(partials + (int64_t)t * S * (kTileS * kTileS) + e, S, g);
    const double di = sqrt(sumsq_a[i] - sum_a[i] * sum_a[i]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int j = j0 + c;
        if (j >= count_b) continue;
        double v = g[c];
        if (scale_a) v *= (double)scale_a[i] * (double)scale_b[j];
        const double dj = sqrt(sumsq_b[j] - sum_b[j] * sum_b[j]);
        out[(int64_t)i * ld + j] = v / (di * dj);
    }
}

}  // namespace

// Values per 128-byte block of a store format.
int block_values(int format) {
    return format == PRNU_PREC_FP8 ? kBlk8 : format == PRNU_PREC_FP16 ? kBlk16 : kBlk;
}

// Work items per K slab of the 4-wave kernel: off-diagonal tiles plus
// diagonal packs of 4 blocks (item_waves4).
int items_per_slab4(int nt) { return nt * (nt - 1) / 2 + (3 * nt + 3) / 4; }

int device_cus(int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 256;
    return cus;
}

// Blocks per slab for groups of (items per slab, K blocks): all work items
// cost about kcb block-stages and one CU runs one item at a time, so the
// launch takes about ceil(items / cus) * kcb stages, plus a per-slab cost
// (the partial tile each item writes and the reduction reads, ~2 stages).
// Searches kcb and returns the cheapest whose partial tiles fit `max_bytes`.
int64_t choose_kcb(const int *items, const int *tiles, const int64_t *nkb, int ngroups, int cus, size_t max_bytes) {
    int64_t max_nkb = 1;
    for (int g = 0; g < ngroups; ++g) max_nkb = std::max(max_nkb, nkb[g]);
    double best = 1e300;
    int64_t best_kcb = max_nkb;
    for (int64_t kcb = 16; kcb <= max_nkb; kcb += (kcb < 256 ? 8 : kcb < 2048 ? 32 : 256)) {
        double total = 0.0, part = 0.0;
        for (int g = 0; g < ngroups; ++g) {
            const int64_t S = (nkb[g] + kcb - 1) / kcb;
            total += (double)items[g] * S;
            part += (double)tiles[g] * S * kTileS * kTileS * sizeof(float);
        }
        if (part > (double)max_bytes) continue;
        const double rounds = std::ceil(total / cus);
        const double cost = rounds * (double)(kcb + 2);
        if (cost < best * (1.0 - 1e-9)) {
            best = cost;
            best_kcb = kcb;
        }
    }
    return best_kcb;
}

int store_create(int64_t n, int capacity, int format, Store **out) {
    *out = nullptr;
    PRNU_REQUIRE(n >= 1 && capacity >= 1, "store_create: n and capacity must be positive");
    // one block column (capacity x 128 bytes) is addressed by a buffer
    // resource with 32-bit offsets in the Gram kernels
    PRNU_REQUIRE(capacity <= kMaxStoreSlots, "store_create: capacity %d exceeds %d slots", capacity,
                 kMaxStoreSlots);
    PRNU_REQUIRE(format == PRNU_PREC_SPLIT_BF16 || format == PRNU_PREC_FP32 || format == PRNU_PREC_FP8 ||
                     format == PRNU_PREC_FP16,
                 "store_create: format must be PRNU_PREC_SPLIT_BF16, PRNU_PREC_FP32, PRNU_PREC_FP8 or "
                 "PRNU_PREC_FP16");
    Store *s = new (std::nothrow) Store();
    if (!s) {
        set_error("store_create: out of host memory");
        return PRNU_ENOMEM;
    }
    s->n = n;
    const int vals = block_values(format);
    s->nkb = (n + vals - 1) / vals;
    s->capacity = capacity;
    s->cap_pad = (capacity + kTileS - 1) / kTileS * kTileS;
    s->format = format;
    (void)hipGetDevice(&s->device);
    const size_t bytes = (size_t)s->nkb * s->cap_pad * kBlkBytes;
    hipError_t e = hipMalloc(&s->blocks, bytes);
    if (e == hipSuccess) e = hipMalloc(&s->sum, sizeof(double) * s->cap_pad);
    if (e == hipSuccess) e = hipMalloc(&s->sumsq, sizeof(double) * s->cap_pad);
    if (e == hipSuccess) e = hipMalloc(&s->absmax, sizeof(float) * s->cap_pad);
    if (e == hipSuccess && (format == PRNU_PREC_FP8 || format == PRNU_PREC_FP16))
        e = hipMalloc(&s->scale, sizeof(float) * s->cap_pad);
    // empty slots read as zeros (the Gram masks them anyway)
    if (e == hipSuccess) e = hipMemset(s->blocks, 0, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("store_create: %zu bytes of device memory: %s", bytes, hipGetErrorString(e));
        if (s->blocks) (void)hipFree(s->blocks);
        if (s->sum) (void)hipFree(s->sum);
        if (s->sumsq) (void)hipFree(s->sumsq);
        if (s->absmax) (void)hipFree(s->absmax);
        if (s->scale) (void)hipFree(s->scale);
        delete s;
        return PRNU_ENOMEM;
    }
    *out = s;
    return PRNU_OK;
}

void store_destroy(Store *s) {
    if (!s) return;
    (void)hipDeviceSynchronize();
    if (s->blocks) (void)hipFree(s->blocks);
    if (s->sum) (void)hipFree(s->sum);
    if (s->sumsq) (void)hipFree(s->sumsq);
    if (s->absmax) (void)hipFree(s->absmax);
    if (s->scale) (void)hipFree(s->scale);
    s->ws_part.release();
    delete s;
}

int store_put(Store *s, const float *const *d_ptrs, int count, int first, hipStream_t st) {
    PRNU_REQUIRE(count >= 1 && first >= 0 && first + count <= s->capacity,
                 "store_put: slots [%d, %d) outside the capacity %d", first, first + count, s->capacity);
    if (s->format == PRNU_PREC_FP8) {
        // statistics (and max|x|, which sets the scale) first, then quantise
        PRNU_REQUIRE(count <= 65535, "store_put: at most 65535 FP8 patterns per call");
        PRNU_TRY(launch_pattern_stats(d_ptrs, count, s->n, s->sum + first, s->sumsq + first, s->absmax + first,
                                      s->ws_part, st));
        const dim3 grid8((unsigned)((4 * s->nkb + 255) / 256), (unsigned)count);
        hipLaunchKernelGGL(store_put_fp8_kernel, grid8, dim3(256), 0, st, d_ptrs, count, s->n, s->nkb, s->blocks,
                           s->cap_pad, first, s->absmax, s->scale);
        PRNU_CHECK_LAUNCH();
        return PRNU_OK;
    }
    if (s->format == PRNU_PREC_FP16) {
        // statistics (and max|x|, which sets the power-of-two scale) first
        PRNU_REQUIRE(count <= 65535, "store_put: at most 65535 FP16 patterns per call");
        PRNU_TRY(launch_pattern_stats(d_ptrs, count, s->n, s->sum + first, s->sumsq + first, s->absmax + first,
                                      s->ws_part, st));
        const dim3 grid16((unsigned)((2 * s->nkb + 255) / 256), (unsigned)count);
        hipLaunchKernelGGL(store_put_fp16_kernel, grid16, dim3(256), 0, st, d_ptrs, count, s->n, s->nkb, s->blocks,
                           s->cap_pad, first, s->absmax, s->scale);
        PRNU_CHECK_LAUNCH();
        return PRNU_OK;
    }
    const unsigned gx = (unsigned)((s->nkb + 31) / 32);
    PRNU_REQUIRE((count + 7) / 8 <= 65535, "store_put: at most 524280 patterns per call");
    PRNU_TRY(s->ws_part.ensure(sizeof(StatPartS) * (size_t)gx * count));
    StatPartS *part = static_cast<StatPartS *>(s->ws_part.ptr);
    const dim3 grid(gx, (unsigned)((count + 7) / 8));
    if (s->format == PRNU_PREC_SPLIT_BF16)
        hipLaunchKernelGGL(store_put_kernel<PRNU_PREC_SPLIT_BF16>, grid, dim3(256), 0, st, d_ptrs, count, s->n,
                           s->nkb, s->blocks, s->cap_pad, first, part);
    else
        hipLaunchKernelGGL(store_put_kernel<PRNU_PREC_FP32>, grid, dim3(256), 0, st, d_ptrs, count, s->n, s->nkb,
                           s->blocks, s->cap_pad, first, part);
    PRNU_CHECK_LAUNCH();
    hipLaunchKernelGGL(store_stats_final_kernel, dim3(count), dim3(256), 0, st, part, (int)gx, count, first, s->sum,
                       s->sumsq, s->absmax);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

int store_gram_ncc(Context *c, Store *s, int first, int count, double *d_ncc, hipStream_t st) {
    PRNU_REQUIRE(count >= 1 && first >= 0 && first + count <= s->capacity,
                 "store_gram: slots [%d, %d) outside the capacity %d", first, first + count, s->capacity);
    StoreGramArgs a;
    a.blocks = s->blocks;
    a.cap_bytes = (int64_t)s->cap_pad * kBlkBytes;
    a.nkb = s->nkb;
    a.first = first;
    a.count = count;
    a.nt = (count + kTileS - 1) / kTileS;
    const prnu_gram_config cfg = c->gram_cfg[s->format];
    const int var = cfg.variant;  // 0 direct-to-LDS, 1 register-staged, 2 + priority, 3 four waves (not FP32)
    const bool four = var == 3 && s->format != PRNU_PREC_FP32;
    a.T = four ? items_per_slab4(a.nt) : tri_items_per_slab(count);
    const int tiles = a.nt * (a.nt + 1) / 2;  // partial tiles per slab (triangle incl. diagonal)
    const int64_t tile_bytes = (int64_t)kTileS * kTileS * sizeof(float);
    if (cfg.slabs > 0) {
        int64_t S = std::min<int64_t>(cfg.slabs, s->nkb);
        S = std::min<int64_t>(S, std::max<int64_t>(1, (int64_t(2) << 30) / (tile_bytes * tiles)));
        a.kcb = (s->nkb + S - 1) / S;
    } else {
        const int64_t nkb = s->nkb;
        a.kcb = choose_kcb(&a.T, &tiles, &nkb, 1, device_cus(c->device), size_t(2) << 30);
    }
    a.S = (int)((s->nkb + a.kcb - 1) / a.kcb);
    a.xcd_remap = cfg.xcd_remap;
    PRNU_TRY(c->ws_gram.ensure((size_t)tile_bytes * tiles * a.S));
    a.partials = static_cast<float *>(c->ws_gram.ptr);
    const int64_t S8 = a.xcd_remap ? (a.S + 7) / 8 * 8 : a.S;
    const dim3 grid((unsigned)(a.T * S8));
    time_begin(c, st);
    if (four) {
        if (s->format == PRNU_PREC_FP8)
            hipLaunchKernelGGL(store_gram4_kernel<PRNU_PREC_FP8>, grid, dim3(kThr4), 0, st, a);
        else if (s->format == PRNU_PREC_FP16)
            hipLaunchKernelGGL(store_gram4_kernel<PRNU_PREC_FP16>, grid, dim3(kThr4), 0, st, a);
        else
            hipLaunchKernelGGL(store_gram4_kernel<PRNU_PREC_SPLIT_BF16>, grid, dim3(kThr4), 0, st, a);
        PRNU_CHECK_LAUNCH();
        time_end(c, st);
        hipLaunchKernelGGL(store_reduce_kernel, dim3(kTileS * kTileS / 1024, tiles), dim3(256), 0, st, a.partials,
                           a.S, a.nt, count, s->sum + first, s->sumsq + first, s->scale ? s->scale + first : nullptr,
                           d_ncc, (int64_t)count);
        PRNU_CHECK_LAUNCH();
        return PRNU_OK;
    }
#define PRNU_STORE_GRAM(F)                                                                        \
    do {                                                                                          \
        if (var == 1)                                                                             \
            hipLaunchKernelGGL((store_gram_kernel<F, 1>), grid, dim3(kThrS), 0, st, a);           \
        else if (var == 2)                                                                        \
            hipLaunchKernelGGL((store_gram_kernel<F, 2>), grid, dim3(kThrS), 0, st, a);           \
        else                                                                                      \
            hipLaunchKernelGGL((store_gram_kernel<F, 0>), grid, dim3(kThrS), 0, st, a);           \
    } while (0)
    if (s->format == PRNU_PREC_SPLIT_BF16)
        PRNU_STORE_GRAM(PRNU_PREC_SPLIT_BF16);
    else if (s->format == PRNU_PREC_FP8)
        PRNU_STORE_GRAM(PRNU_PREC_FP8);
    else if (s->format == PRNU_PREC_FP16)
        PRNU_STORE_GRAM(PRNU_PREC_FP16);
    else
        PRNU_STORE_GRAM(PRNU_PREC_FP32);
#undef PRNU_STORE_GRAM
    PRNU_CHECK_LAUNCH();
    time_end(c, st);
    hipLaunchKernelGGL(store_reduce_kernel, dim3(kTileS * kTileS / 1024, tiles), dim3(256), 0, st, a.partials, a.S,
                       a.nt, count, s->sum + first, s->sumsq + first, s->scale ? s->scale + first : nullptr, d_ncc,
                       (int64_t)count);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

// ------------------------------------------------------------ grouped plan
// One block of a plan: the triangle of a slot range (tri) or the rectangle
// rows x cols of two slot ranges of one store.
struct PlanBlock {
    const Store *store;
    int row_first, row_count, col_first, col_count;
    double *out;
    int64_t ld;
    bool tri;
    int nta() const { return (row_count + kTileS - 1) / kTileS; }
    int ntb() const { return (col_count + kTileS - 1) / kTileS; }
    int items() const { return tri ? tri_items_per_slab(row_count) : nta() * ntb(); }
    int tiles() const { return tri ? nta() * (nta() + 1) / 2 : nta() * ntb(); }
};

struct GramPlan {
    Context *c = nullptr;
    int format = PRNU_PREC_SPLIT_BF16;
    std::vector<PlanBlock> groups;
    std::vector<GroupDev> gd;        // host copy of the device descriptors
    GroupDev *d_groups = nullptr;
    float *partials = nullptr;
    size_t partial_bytes = 0;
    int64_t kcb = 0;                 // K blocks per slab
    int64_t v_total = 0, items = 0;  // workgroups launched, work items (padding excluded)
    int cus = 0;
    int var = 0;                     // load variant (the context's config of the format)
    int persistent = 0;              // 1: one workgroup per CU takes items from per-XCD queues
    int *queue = nullptr;            // 8 counters, 64 bytes apart
    double flop = 0.0;               // one product over every block (triangles incl. diagonal)
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

constexpr size_t kPlanPartialCap = size_t(8) << 30;

void plan_drop_graph(GramPlan *p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
}

// Work list for slabs of `kcb` blocks: descriptors, partial-tile buffer.
int plan_layout(GramPlan *p, int64_t kcb) {
    // a run already queued on any stream (e.g. an asynchronous run on a
    // non-blocking stream) still reads the descriptors and partial tiles that
    // are rewritten or freed below: let it finish first
    if (p->d_groups || p->partials) PRNU_CHECK_HIP(hipDeviceSynchronize());
    plan_drop_graph(p);
    const int64_t tile_elems = (int64_t)kTileS * kTileS;
    std::vector<GroupDev> gd(p->groups.size());
    int64_t v = 0, items = 0, part = 0;
    for (size_t i = 0; i < p->groups.size(); ++i) {
        const PlanBlock &g = p->groups[i];
        const Store *s = g.store;
        GroupDev &d = gd[i];
        d.blocks = s->blocks;
        d.cap_bytes = (int64_t)s->cap_pad * kBlkBytes;
        d.nkb = s->nkb;
        d.kcb = std::min<int64_t>(kcb, s->nkb);
        d.first = g.row_first;
        d.count = g.row_count;
        d.nt = g.nta();
        d.first_b = g.tri ? g.row_first : g.col_first;
        d.count_b = g.tri ? g.row_count : g.col_count;
        d.ntb = g.tri ? 0 : g.ntb();
        d.T = g.items();
        d.S = (int)((s->nkb + d.kcb - 1) / d.kcb);
        d.v_begin = v;
        d.partials = reinterpret_cast<float *>((uintptr_t)part);  // offset for now
        v += (int64_t)d.T * ((d.S + 7) / 8 * 8);
        items += (int64_t)d.T * d.S;
        part += tile_elems * sizeof(float) * g.tiles() * d.S;
    }
    PRNU_REQUIRE(v < (int64_t(1) << 31), "gram_plan: %lld work items; use longer slabs", (long long)v);
    PRNU_REQUIRE((size_t)part <= kPlanPartialCap,
                 "gram_plan: %lld bytes of partial tiles exceed the %zu-byte limit; use longer slabs",
                 (long long)part, kPlanPartialCap);
    if ((size_t)part > p->partial_bytes) {
        if (p->partials) PRNU_CHECK_HIP(hipFree(p->partials));
        p->partials = nullptr;
        p->partial_bytes = 0;
        hipError_t e = hipMalloc(&p->partials, (size_t)part);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            set_error("gram_plan: %lld bytes of partial tiles: %s", (long long)part, hipGetErrorString(e));
            return PRNU_ENOMEM;
        }
        p->partial_bytes = (size_t)part;
    }
    for (GroupDev &d : gd)
        d.partials = reinterpret_cast<float *>(reinterpret_cast<uint8_t *>(p->partials) + (uintptr_t)d.partials);
    if (!p->d_groups) PRNU_CHECK_HIP(hipMalloc(&p->d_groups, sizeof(GroupDev) * kMaxPlanGroups));
    if (!p->queue) PRNU_CHECK_HIP(hipMalloc(&p->queue, 128 * sizeof(int)));
    PRNU_CHECK_HIP(hipMemcpy(p->d_groups, gd.data(), sizeof(GroupDev) * gd.size(), hipMemcpyHostToDevice));
    p->gd = gd;
    p->kcb = kcb;
    p->v_total = v;
    p->items = items;
    return PRNU_OK;
}

// Default slab: the cheapest by the rounds model of choose_kcb.
int64_t plan_default_kcb(const GramPlan *p) {
    std::vector<int> items, tiles;
    std::vector<int64_t> nkb;
    for (const PlanBlock &g : p->groups) {
        items.push_back(g.items());
        tiles.push_back(g.tiles());
        nkb.push_back(g.store->nkb);
    }
    return choose_kcb(items.data(), tiles.data(), nkb.data(), (int)items.size(), p->cus, kPlanPartialCap);
}

// Persistent launch: one workgroup per CU, a multiple of 8 so blockIdx.x & 7
// is the XCD.
int plan_persistent_grid(const GramPlan *p) {
    return (int)std::max<int64_t>(8, std::min<int64_t>(p->v_total, p->cus / 8 * 8));
}

int plan_launch(GramPlan *p, hipStream_t st) {
    const int ng = (int)p->gd.size();
    const dim3 grid((unsigned)p->v_total);
    const dim3 pgrid((unsigned)plan_persistent_grid(p));
    if (p->persistent) {
        hipLaunchKernelGGL(store_queue_reset_kernel, dim3(1), dim3(128), 0, st, p->queue);
        PRNU_CHECK_LAUNCH();
    }
#define PRNU_GROUPS_GRAM(F)                                                                              \
    do {                                                                                                 \
        if (p->persistent)                                                                               \
            hipLaunchKernelGGL((store_gram_groups_persistent_kernel<F>), pgrid, dim3(kThrS), 0, st,         \
                               p->d_groups, ng, p->v_total, p->queue);                                   \
        else if (p->var == 1)                                                                            \
            hipLaunchKernelGGL((store_gram_groups_kernel<F, 1>), grid, dim3(kThrS), 0, st, p->d_groups, ng); \
        else if (p->var == 2)                                                                            \
            hipLaunchKernelGGL((store_gram_groups_kernel<F, 2>), grid, dim3(kThrS), 0, st, p->d_groups, ng); \
        else                                                                                             \
            hipLaunchKernelGGL((store_gram_groups_kernel<F, 0>), grid, dim3(kThrS), 0, st, p->d_groups, ng); \
    } while (0)
    if (p->format == PRNU_PREC_SPLIT_BF16)
        PRNU_GROUPS_GRAM(PRNU_PREC_SPLIT_BF16);
    else if (p->format == PRNU_PREC_FP8)
        PRNU_GROUPS_GRAM(PRNU_PREC_FP8);
    else if (p->format == PRNU_PREC_FP16)
        PRNU_GROUPS_GRAM(PRNU_PREC_FP16);
    else
        PRNU_GROUPS_GRAM(PRNU_PREC_FP32);
#undef PRNU_GROUPS_GRAM
    PRNU_CHECK_LAUNCH();
    for (int i = 0; i < ng; ++i) {
        const GroupDev &d = p->gd[i];
        const PlanBlock &g = p->groups[i];
        const Store *s = g.store;
        if (g.tri) {
            hipLaunchKernelGGL(store_reduce_kernel, dim3(kTileS * kTileS / 1024, g.tiles()), dim3(256), 0, st,
                               d.partials, d.S, d.nt, d.count, s->sum + d.first, s->sumsq + d.first,
                               s->scale ? s->scale + d.first : nullptr, g.out, g.ld);
        } else {
            hipLaunchKernelGGL(store_reduce_rect_kernel, dim3(kTileS * kTileS / 1024, g.tiles()), dim3(256), 0, st,
                               d.partials, d.S, d.ntb, d.count, d.count_b, s->sum + d.first, s->sumsq + d.first,
                               s->sum + d.first_b, s->sumsq + d.first_b, s->scale ? s->scale + d.first : nullptr,
                               s->scale ? s->scale + d.first_b : nullptr, g.out, g.ld);
        }
        PRNU_CHECK_LAUNCH();
    }
    return PRNU_OK;
}

int plan_capture(GramPlan *p) {
    if (!p->cap_stream) PRNU_CHECK_HIP(hipStreamCreateWithFlags(&p->cap_stream, hipStreamNonBlocking));
    PRNU_CHECK_HIP(hipStreamBeginCapture(p->cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = plan_launch(p, p->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(p->cap_stream, &g);
    if (rc != PRNU_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    PRNU_CHECK_HIP(e);
    p->graph = g;
    PRNU_CHECK_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    return PRNU_OK;
}

int plan_run(GramPlan *p, int use_graph, hipStream_t st) {
    Context *c = p->c;
    if (use_graph && !p->exec) PRNU_TRY(plan_capture(p));
    time_begin(c, st);
    if (use_graph) {
        PRNU_CHECK_HIP(hipGraphLaunch(p->exec, st));
    } else {
        PRNU_TRY(plan_launch(p, st));
    }
    time_end(c, st);
    return PRNU_OK;
}

void plan_free(GramPlan *p) {
    if (!p) return;
    (void)hipDeviceSynchronize();
    plan_drop_graph(p);
    if (p->cap_stream) (void)hipStreamDestroy(p->cap_stream);
    if (p->partials) (void)hipFree(p->partials);
    if (p->d_groups) (void)hipFree(p->d_groups);
    if (p->queue) (void)hipFree(p->queue);
    delete p;
}

}  // namespace

int gram_plan_create_blocks(Context *c, const std::vector<PlanBlock> &blocks, int64_t slab_values, GramPlan **out) {
    *out = nullptr;
    const int nblocks = (int)blocks.size();
    PRNU_REQUIRE(nblocks >= 1 && nblocks <= kMaxPlanGroups, "gram_plan: 1..%d blocks, got %d", kMaxPlanGroups,
                 nblocks);
    PRNU_REQUIRE(slab_values >= 0, "gram_plan: slab_values must be non-negative");
    int format = -1;
    for (int i = 0; i < nblocks; ++i) {
        const PlanBlock &g = blocks[i];
        PRNU_REQUIRE(g.store && g.out, "gram_plan: block %d has a null store or output", i);
        const Store *s = g.store;
        PRNU_REQUIRE(g.row_count >= 1 && g.row_first >= 0 && g.row_first + g.row_count <= s->capacity &&
                         g.col_count >= 1 && g.col_first >= 0 && g.col_first + g.col_count <= s->capacity,
                     "gram_plan: block %d slots rows [%d, %d) x cols [%d, %d) outside the capacity %d", i,
                     g.row_first, g.row_first + g.row_count, g.col_first, g.col_first + g.col_count, s->capacity);
        PRNU_REQUIRE(g.ld >= (g.tri ? g.row_count : g.col_count), "gram_plan: block %d output stride too small", i);
        if (format < 0) format = s->format;
        PRNU_REQUIRE(s->format == format, "gram_plan: every block's store must have the same format");
    }
    GramPlan *p = new (std::nothrow) GramPlan();
    if (!p) {
        set_error("gram_plan: out of host memory");
        return PRNU_ENOMEM;
    }
    p->c = c;
    p->format = format;
    p->var = c->gram_cfg[format].variant;
    p->groups = blocks;
    p->cus = device_cus(c->device);  // one 512-thread, 128 KiB-LDS workgroup per CU at a time
    for (const PlanBlock &g : p->groups)
        p->flop += 2.0 * (double)g.store->n *
                   (g.tri ? (double)g.row_count * (g.row_count + 1) / 2.0 : (double)g.row_count * g.col_count);
    const int vals = block_values(format);
    if (slab_values % vals != 0) {
        set_error("gram_plan: slab_values must be a multiple of %d for this store format", vals);
        plan_free(p);
        return PRNU_EINVAL;
    }
    const int64_t kcb = slab_values ? slab_values / vals : plan_default_kcb(p);
    const int rc = plan_layout(p, kcb);
    if (rc != PRNU_OK) {
        plan_free(p);
        return rc;
    }
    *out = p;
    return PRNU_OK;
}

int gram_plan_create(Context *c, const prnu_gram_group *groups, int ngroups, int64_t slab_values, GramPlan **out) {
    *out = nullptr;
    PRNU_REQUIRE(groups && ngroups >= 1 && ngroups <= kMaxPlanGroups, "gram_plan: 1..%d groups, got %d",
                 kMaxPlanGroups, ngroups);
    std::vector<PlanBlock> blocks;
    for (int i = 0; i < ngroups; ++i) {
        const prnu_gram_group &g = groups[i];
        PRNU_REQUIRE(g.store, "gram_plan: group %d has a null store", i);
        blocks.push_back(PlanBlock{reinterpret_cast<const Store *>(g.store), g.first_slot, g.count, g.first_slot,
                                   g.count, g.d_ncc, (int64_t)g.count, true});
    }
    return gram_plan_create_blocks(c, blocks, slab_values, out);
}

// Times every candidate slab length (eager launches, best of `reps`) and keeps
// the fastest.
int gram_plan_autotune(GramPlan *p, hipStream_t st, int64_t *best_values, float *best_ms) {
    const int64_t base = plan_default_kcb(p);
    const int64_t cands[] = {base / 4, base / 2, base, base * 2, base * 4, 256, 512, 768, 1024};
    hipEvent_t e0, e1;
    PRNU_CHECK_HIP(hipEventCreate(&e0));
    PRNU_CHECK_HIP(hipEventCreate(&e1));
    float best = 3.4e38f;
    int64_t best_kcb = p->kcb;
    int best_launch = p->persistent;
    int rc = PRNU_OK;
    for (int64_t kcb : cands) {
        if (kcb < 16) continue;
        if (plan_layout(p, kcb) != PRNU_OK) continue;  // e.g. partials over the limit
        for (int launch = 0; launch < 2 && rc == PRNU_OK; ++launch) {
            p->persistent = launch;
            rc = plan_launch(p, st);  // warm-up
            float ms_min = 3.4e38f;
            for (int r = 0; r < 3 && rc == PRNU_OK; ++r) {
                (void)hipEventRecord(e0, st);
                rc = plan_launch(p, st);
                (void)hipEventRecord(e1, st);
                if (hipEventSynchronize(e1) != hipSuccess) rc = PRNU_EDEVICE;
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, e0, e1);
                ms_min = std::min(ms_min, ms);
            }
            if (rc == PRNU_OK && ms_min < best) {
                best = ms_min;
                best_kcb = kcb;
                best_launch = launch;
            }
        }
        if (rc != PRNU_OK) break;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    p->persistent = best_launch;
    if (rc != PRNU_OK) return rc;
    PRNU_TRY(plan_layout(p, best_kcb));
    if (best_values) *best_values = best_kcb * block_values(p->format);
    if (best_ms) *best_ms = best;
    return PRNU_OK;
}

}  // namespace prnu

using namespace prnu;

extern "C" {

int prnu_store_create(int64_t n, int capacity, int format, prnu_store **out) {
    PRNU_REQUIRE(out, "store_create: null output");
    Store *s = nullptr;
    PRNU_TRY(store_create(n, capacity, format, &s));
    *out = reinterpret_cast<prnu_store *>(s);
    return PRNU_OK;
}

int prnu_store_destroy(prnu_store *store) {
    store_destroy(reinterpret_cast<Store *>(store));
    return PRNU_OK;
}

int prnu_store_info(const prnu_store *store, int64_t *n, int *capacity, int *format, size_t *bytes) {
    PRNU_REQUIRE(store, "store_info: null store");
    const Store *s = reinterpret_cast<const Store *>(store);
    if (n) *n = s->n;
    if (capacity) *capacity = s->capacity;
    if (format) *format = s->format;
    if (bytes) *bytes = (size_t)s->nkb * s->cap_pad * kBlkBytes;
    return PRNU_OK;
}

int prnu_store_put(prnu_store *store, const float *const *d_patterns, int count, int first_slot,
                   prnu_stream stream) {
    PRNU_REQUIRE(store && d_patterns, "store_put: null pointer");
    Store *s = reinterpret_cast<Store *>(store);
    PRNU_TRY(store_put(s, d_patterns, count, first_slot, (hipStream_t)stream));
    if (!stream) PRNU_CHECK_HIP(hipStreamSynchronize(nullptr));
    return PRNU_OK;
}

// Slot images (the spill tier): a header, then the slots' K blocks, block by
// block (count x 128 bytes each), then per slot the FP64 sum and sum of
// squares and the FP32 max|x| and scale (0 for formats without a scale).
struct SlotImageHeader {
    char magic[8];  // "PRNUSLOT"
    int64_t n;      // values per pattern
    int32_t format, count;
    int64_t nkb;    // K blocks per pattern
    int64_t reserved[4];
};
static_assert(sizeof(SlotImageHeader) == 64, "slot image header is 64 bytes");

static int64_t slot_image_bytes(const Store *s, int count) {
    return (int64_t)sizeof(SlotImageHeader) +
           (int64_t)count * (s->nkb * kBlkBytes + 2 * (int64_t)sizeof(double) + 2 * (int64_t)sizeof(float));
}

int64_t prnu_store_image_bytes(const prnu_store *store, int count) {
    if (!store || count < 0) return -1;
    return slot_image_bytes(reinterpret_cast<const Store *>(store), count);
}

int prnu_store_save(const prnu_store *store, int first_slot, int count, void *h_image, prnu_stream stream) {
    PRNU_REQUIRE(store && h_image, "store_save: null pointer");
    const Store *s = reinterpret_cast<const Store *>(store);
    PRNU_REQUIRE(count >= 1 && first_slot >= 0 && first_slot + count <= s->capacity,
                 "store_save: slots [%d, %d) outside the capacity %d", first_slot, first_slot + count, s->capacity);
    hipStream_t st = (hipStream_t)stream;
    uint8_t *h = static_cast<uint8_t *>(h_image);
    SlotImageHeader hd;
    std::memset(&hd, 0, sizeof(hd));
    std::memcpy(hd.magic, "PRNUSLOT", 8);
    hd.n = s->n;
    hd.format = s->format;
    hd.count = count;
    hd.nkb = s->nkb;
    std::memcpy(h, &hd, sizeof(hd));
    uint8_t *data = h + sizeof(hd);
    const size_t row = (size_t)count * kBlkBytes;
    PRNU_CHECK_HIP(hipMemcpy2DAsync(data, row, s->blocks + (size_t)first_slot * kBlkBytes,
                                    (size_t)s->cap_pad * kBlkBytes, row, (size_t)s->nkb, hipMemcpyDeviceToHost, st));
    uint8_t *stats = data + row * (size_t)s->nkb;
    PRNU_CHECK_HIP(hipMemcpyAsync(stats, s->sum + first_slot, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    stats += sizeof(double) * count;
    PRNU_CHECK_HIP(hipMemcpyAsync(stats, s->sumsq + first_slot, sizeof(double) * count, hipMemcpyDeviceToHost, st));
    stats += sizeof(double) * count;
    PRNU_CHECK_HIP(hipMemcpyAsync(stats, s->absmax + first_slot, sizeof(float) * count, hipMemcpyDeviceToHost, st));
    stats += sizeof(float) * count;
    if (s->scale)
        PRNU_CHECK_HIP(hipMemcpyAsync(stats, s->scale + first_slot, sizeof(float) * count, hipMemcpyDeviceToHost, st));
    else
        std::memset(stats, 0, sizeof(float) * count);
    if (!stream) PRNU_CHECK_HIP(hipStreamSynchronize(nullptr));
    return PRNU_OK;
}

int prnu_store_load(prnu_store *store, int first_slot, const void *h_image, size_t image_bytes,
                    prnu_stream stream) {
    PRNU_REQUIRE(store && h_image, "store_load: null pointer");
    Store *s = reinterpret_cast<Store *>(store);
    const uint8_t *h = static_cast<const uint8_t *>(h_image);
    PRNU_REQUIRE(image_bytes >= sizeof(SlotImageHeader), "store_load: %zu bytes hold no slot image header",
                 image_bytes);
    SlotImageHeader hd;
    std::memcpy(&hd, h, sizeof(hd));
    PRNU_REQUIRE(std::memcmp(hd.magic, "PRNUSLOT", 8) == 0, "store_load: not a slot image");
    PRNU_REQUIRE(hd.n == s->n && hd.format == s->format && hd.nkb == s->nkb,
                 "store_load: image of %lld values in format %d, store of %lld in format %d", (long long)hd.n,
                 hd.format, (long long)s->n, s->format);
    const int count = hd.count;
    PRNU_REQUIRE(count >= 1 && first_slot >= 0 && first_slot + count <= s->capacity,
                 "store_load: slots [%d, %d) outside the capacity %d", first_slot, first_slot + count, s->capacity);
    // every byte the copies below read lies inside the caller's buffer
    PRNU_REQUIRE((int64_t)image_bytes >= slot_image_bytes(s, count),
                 "store_load: image of %d slots needs %lld bytes, the buffer holds %zu", count,
                 (long long)slot_image_bytes(s, count), image_bytes);
    hipStream_t st = (hipStream_t)stream;
    const uint8_t *data = h + sizeof(hd);
    const size_t row = (size_t)count * kBlkBytes;
    PRNU_CHECK_HIP(hipMemcpy2DAsync(s->blocks + (size_t)first_slot * kBlkBytes, (size_t)s->cap_pad * kBlkBytes, data,
                                    row, row, (size_t)s->nkb, hipMemcpyHostToDevice, st));
    const uint8_t *stats = data + row * (size_t)s->nkb;
    PRNU_CHECK_HIP(hipMemcpyAsync(s->sum + first_slot, stats, sizeof(double) * count, hipMemcpyHostToDevice, st));
    stats += sizeof(double) * count;
    PRNU_CHECK_HIP(hipMemcpyAsync(s->sumsq + first_slot, stats, sizeof(double) * count, hipMemcpyHostToDevice, st));
    stats += sizeof(double) * count;
    PRNU_CHECK_HIP(hipMemcpyAsync(s->absmax + first_slot, stats, sizeof(float) * count, hipMemcpyHostToDevice, st));
    stats += sizeof(float) * count;
    if (s->scale)
        PRNU_CHECK_HIP(hipMemcpyAsync(s->scale + first_slot, stats, sizeof(float) * count, hipMemcpyHostToDevice, st));
    if (!stream) PRNU_CHECK_HIP(hipStreamSynchronize(nullptr));
    return PRNU_OK;
}

int prnu_store_stats(const prnu_store *store, const double **d_sum, const double **d_sumsq, const float **d_absmax) {
    PRNU_REQUIRE(store, "store_stats: null store");
    const Store *s = reinterpret_cast<const Store *>(store);
    if (d_sum) *d_sum = s->sum;
    if (d_sumsq) *d_sumsq = s->sumsq;
    if (d_absmax) *d_absmax = s->absmax;
    return PRNU_OK;
}

int prnu_store_gram_ncc(prnu_ctx *ctx, prnu_store *store, int first_slot, int count, double *d_ncc,
                        prnu_stream stream) {
    PRNU_REQUIRE(ctx && store && d_ncc, "store_gram_ncc: null pointer");
    Context *c = reinterpret_cast<Context *>(ctx);
    Store *s = reinterpret_cast<Store *>(store);
    hipStream_t st = pick_stream(c, stream);
    PRNU_TRY(store_gram_ncc(c, s, first_slot, count, d_ncc, st));
    return finish(c, stream);
}

int prnu_gram_plan_create(prnu_ctx *ctx, const prnu_gram_group *groups, int ngroups, int64_t slab_values,
                          prnu_gram_plan **out) {
    PRNU_REQUIRE(ctx && out, "gram_plan_create: null pointer");
    GramPlan *p = nullptr;
    PRNU_TRY(gram_plan_create(reinterpret_cast<Context *>(ctx), groups, ngroups, slab_values, &p));
    *out = reinterpret_cast<prnu_gram_plan *>(p);
    return PRNU_OK;
}

int prnu_gram_plan_create_blocks(prnu_ctx *ctx, const prnu_gram_block *blocks, int nblocks, int64_t slab_values,
                                 prnu_gram_plan **out) {
    PRNU_REQUIRE(ctx && out && blocks, "gram_plan_create_blocks: null pointer");
    PRNU_REQUIRE(nblocks >= 1 && nblocks <= kMaxPlanGroups, "gram_plan: 1..%d blocks, got %d", kMaxPlanGroups,
                 nblocks);
    std::vector<PlanBlock> v;
    for (int i = 0; i < nblocks; ++i) {
        const prnu_gram_block &b = blocks[i];
        PRNU_REQUIRE(b.store, "gram_plan: block %d has a null store", i);
        const bool tri = b.row_first == b.col_first && b.row_count == b.col_count;
        v.push_back(PlanBlock{reinterpret_cast<const Store *>(b.store), b.row_first, b.row_count, b.col_first,
                              b.col_count, b.d_out, b.ld_out, tri});
    }
    GramPlan *p = nullptr;
    PRNU_TRY(gram_plan_create_blocks(reinterpret_cast<Context *>(ctx), v, slab_values, &p));
    *out = reinterpret_cast<prnu_gram_plan *>(p);
    return PRNU_OK;
}

int prnu_gram_plan_destroy(prnu_gram_plan *plan) {
    plan_free(reinterpret_cast<GramPlan *>(plan));
    return PRNU_OK;
}

int prnu_gram_plan_run(prnu_gram_plan *plan, int use_graph, prnu_stream stream) {
    PRNU_REQUIRE(plan, "gram_plan_run: null plan");
    GramPlan *p = reinterpret_cast<GramPlan *>(plan);
    hipStream_t st = pick_stream(p->c, stream);
    PRNU_TRY(plan_run(p, use_graph, st));
    return finish(p->c, stream);
}

int prnu_gram_plan_autotune(prnu_gram_plan *plan, prnu_stream stream, int64_t *best_slab_values, float *best_ms) {
    PRNU_REQUIRE(plan, "gram_plan_autotune: null plan");
    GramPlan *p = reinterpret_cast<GramPlan *>(plan);
    hipStream_t st = pick_stream(p->c, stream);
    PRNU_TRY(gram_plan_autotune(p, st, best_slab_values, best_ms));
    return finish(p->c, stream);
}

int prnu_gram_plan_set_launch(prnu_gram_plan *plan, int launch) {
    PRNU_REQUIRE(plan, "gram_plan_set_launch: null plan");
    PRNU_REQUIRE(launch == PRNU_PLAN_PER_ITEM || launch == PRNU_PLAN_PERSISTENT,
                 "gram_plan_set_launch: launch must be 0 (per item) or 1 (persistent), got %d", launch);
    GramPlan *p = reinterpret_cast<GramPlan *>(plan);
    if (p->persistent != launch) plan_drop_graph(p);
    p->persistent = launch;
    return PRNU_OK;
}

int prnu_gram_plan_get_launch(const prnu_gram_plan *plan, int *launch) {
    PRNU_REQUIRE(plan && launch, "gram_plan_get_launch: null pointer");
    *launch = reinterpret_cast<const GramPlan *>(plan)->persistent;
    return PRNU_OK;
}

int prnu_gram_plan_info(const prnu_gram_plan *plan, int64_t *work_items, int *workgroups, int64_t *slab_values,
                        size_t *partial_bytes, double *flop) {
    PRNU_REQUIRE(plan, "gram_plan_info: null plan");
    const GramPlan *p = reinterpret_cast<const GramPlan *>(plan);
    if (work_items) *work_items = p->items;
    if (workgroups) *workgroups = p->persistent ? plan_persistent_grid(p) : (int)p->v_total;
    if (slab_values) *slab_values = p->kcb * block_values(p->format);
    if (partial_bytes) *partial_bytes = p->partial_bytes;
    if (flop) *flop = p->flop;
    return PRNU_OK;
}

}  // extern "C"
