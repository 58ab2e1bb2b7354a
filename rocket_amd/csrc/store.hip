// Device-resident pattern store in a K-blocked layout, and the all-pairs Gram
// over it.
//
// Layout.  A pattern of n values is cut into nkb = ceil(n/32) blocks of 32.
// Block b of every slot is contiguous: byte offset (b * cap_pad + slot) * 128.
// A block is 128 bytes: split-BF16 stores hi = rne(x) (32 bf16) then
// lo = rne(x - hi) (32 bf16); FP32 stores the 32 floats.  Values past n are 0.
//
// Why.  One K stage of a 256 x 256 Gram tile reads 32 values of 512 patterns.
// With row-major patterns those are 512 reads 4*n bytes apart, one page each;
// the per-CU TLB cannot hold them and half of all translations miss.  In this
// layout the same stage reads 2 (or 4) contiguous 32 KiB runs, the hi/lo
// split is done once at insert time instead of in every tile that reads the
// value, and the per-pattern FP64 statistics are taken in the same pass.
#include "prnu_internal.hpp"

#include <algorithm>
#include <new>

namespace prnu {

struct Store {
    int64_t n = 0, nkb = 0;
    int capacity = 0, cap_pad = 0;
    int format = PRNU_PREC_SPLIT_BF16;
    int device = 0;
    uint8_t *blocks = nullptr;
    double *sum = nullptr, *sumsq = nullptr;
    float *absmax = nullptr;
    Workspace ws_part;  // statistics partials
};

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kBlk = 32;        // values per block
constexpr int kBlkBytes = 128;  // bytes per block
constexpr int kTileS = 256;
constexpr int kRowsS = 2 * kTileS;
constexpr int kThrS = 512;

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

__global__ void store_stats_final_kernel(const StatPartS *part, int chunks, int count, int first, double *sum,
                                         double *sumsq, float *absmax) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= count) return;
    double s = 0.0, ss = 0.0;
    float am = 0.0f;
    for (int c = 0; c < chunks; ++c) {
        const StatPartS q = part[(int64_t)p * chunks + c];
        s += q.s;
        ss += q.ss;
        am = fmaxf(am, q.am);
    }
    sum[first + p] = s;
    sumsq[first + p] = ss;
    absmax[first + p] = am;
}

// -------------------------------------------------------------------- Gram
struct StoreGramArgs {
    const uint8_t *blocks;
    float *partials;
    int64_t cap_bytes;   // cap_pad * 128: distance between consecutive blocks of one slot
    int64_t nkb;
    int64_t kcb;         // blocks per slab
    int first, count;
    int nt, T, S;
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

// 16-byte chunk c of LDS row r, XOR-swizzled by row.
__device__ __forceinline__ int s_off(int row, int chunk) { return row * kBlkBytes + ((chunk ^ (row & 7)) << 4); }

template <int FMT>
__global__ __launch_bounds__(kThrS, 1) void store_gram_kernel(StoreGramArgs a) {
    constexpr bool kSplit = FMT == PRNU_PREC_SPLIT_BF16;
    constexpr int kBuf = kRowsS * kBlkBytes;  // 64 KiB per stage
    __shared__ __attribute__((aligned(16))) uint8_t lds[2 * kBuf];

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
    const int2 tile = tri_tile_s(t, a.nt);
    const bool diag = tile.x == tile.y;
    const int nrows = diag ? kTileS : kRowsS;
    const int bbase = diag ? 0 : kTileS;
    const int64_t b_begin = (int64_t)s * a.kcb;
    const int64_t b_end = min(a.nkb, b_begin + a.kcb);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wrow0 = (wave >> 2) * 128, wcol0 = (wave & 3) * 64;

    // thread -> (row = (tid >> 3) + 64*it, chunk = tid & 7); rows past count
    // read a clamped slot and are zeroed on the LDS store
    const int r0 = tid >> 3, chunk = tid & 7;
    const int la = min(tile.x * kTileS + r0, a.count - 1);
    const int lb = min(tile.y * kTileS + r0, a.count - 1);
    uint32_t valid = 0;
#pragma unroll
    for (int it = 0; it < 8; ++it) {
        const int r = r0 + 64 * it;
        const int loc = (r < kTileS) ? tile.x * kTileS + r : tile.y * kTileS + (r - kTileS);
        valid |= (uint32_t)(r < nrows && loc < a.count) << it;
    }
    // per-row byte offsets inside one block column (rows clamped into the set)
    auto row_off = [&](int it) -> int64_t {
        if (it < 4) return (int64_t)(a.first + min(tile.x * kTileS + r0 + 64 * it, a.count - 1)) * kBlkBytes;
        return (int64_t)(a.first + min(tile.y * kTileS + r0 + 64 * (it - 4), a.count - 1)) * kBlkBytes;
    };
    (void)la;
    (void)lb;

    floatx4 acc[8][4];
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    uint4 st[8];
    auto load = [&](int64_t b) {
        const uint8_t *col = a.blocks + b * a.cap_bytes + (chunk << 4);
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            if (it < 4 || !diag) {
                typedef const __attribute__((address_space(1))) unsigned int __attribute__((ext_vector_type(4))) *gu4;
                const auto v = *(gu4)(col + row_off(it));
                st[it] = make_uint4(v[0], v[1], v[2], v[3]);
            }
        }
    };
    auto store = [&](uint8_t *buf) {
#pragma unroll
        for (int it = 0; it < 8; ++it) {
            const int r = r0 + 64 * it;
            if (r < nrows) {
                uint4 v = st[it];
                if (!((valid >> it) & 1)) v = make_uint4(0u, 0u, 0u, 0u);
                *reinterpret_cast<uint4 *>(buf + s_off(r, chunk)) = v;
            }
        }
    };
    auto compute = [&](const uint8_t *buf) {
        const int frow = lane & 15;
        if (kSplit) {
            const int ch = lane >> 4;  // hi chunks 0..3, lo chunks 4..7
            bf16x8 bh[4], bl[4];
#pragma unroll
            for (int cb = 0; cb < 4; ++cb) {
                const int r = bbase + wcol0 + cb * 16 + frow;
                bh[cb] = *reinterpret_cast<const bf16x8 *>(buf + s_off(r, ch));
                bl[cb] = *reinterpret_cast<const bf16x8 *>(buf + s_off(r, ch + 4));
            }
#pragma unroll
            for (int rb = 0; rb < 8; ++rb) {
                const int r = wrow0 + rb * 16 + frow;
                const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(buf + s_off(r, ch));
                const bf16x8 al = *reinterpret_cast<const bf16x8 *>(buf + s_off(r, ch + 4));
#pragma unroll
                for (int cb = 0; cb < 4; ++cb) {
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[cb], acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[cb], acc[rb][cb], 0, 0, 0);
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[cb], acc[rb][cb], 0, 0, 0);
                }
            }
        } else {
#pragma unroll
            for (int kg = 0; kg < 2; ++kg) {
                const int ch = kg * 4 + (lane >> 4);
                float4 bf[4];
#pragma unroll
                for (int cb = 0; cb < 4; ++cb)
                    bf[cb] = *reinterpret_cast<const float4 *>(buf + s_off(bbase + wcol0 + cb * 16 + frow, ch));
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
    load(b_begin);
    store(lds);
    __syncthreads();
    for (int64_t q = 0; q < nst; ++q) {
        const bool next = q + 1 < nst;
        if (next) load(b_begin + q + 1);
        compute(lds + (q & 1) * kBuf);
        if (next) store(lds + ((q + 1) & 1) * kBuf);
        __syncthreads();
    }

    float *out = a.partials + ((int64_t)t * a.S + s) * (kTileS * kTileS);
    const int ocol = lane & 15, orow = (lane >> 4) * 4;
#pragma unroll
    for (int rb = 0; rb < 8; ++rb)
#pragma unroll
        for (int cb = 0; cb < 4; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(wrow0 + rb * 16 + orow + r) * kTileS + wcol0 + cb * 16 + ocol] = acc[rb][cb][r];
}

// FP64 sum of the S partials + NCC epilogue (normalizedcrosscorrelation.cu:121-123).
__global__ __launch_bounds__(256) void store_reduce_kernel(const float *__restrict__ partials, int S, int nt,
                                                           int count, const double *__restrict__ sum,
                                                           const double *__restrict__ sumsq,
                                                           double *__restrict__ out) {
    const int t = blockIdx.y;
    const int e = blockIdx.x * 256 + threadIdx.x;
    const int2 tile = tri_tile_s(t, nt);
    const int i = tile.x * kTileS + (e >> 8);
    const int j = tile.y * kTileS + (e & 255);
    if (i >= count || j >= count) return;
    if (tile.x == tile.y && i > j) return;
    const float *p = partials + (int64_t)t * S * (kTileS * kTileS) + e;
    double g = 0.0;
    for (int q = 0; q < S; ++q) g += (double)p[(int64_t)q * (kTileS * kTileS)];
    const double di = sqrt(sumsq[i] - sum[i] * sum[i]);
    const double dj = sqrt(sumsq[j] - sum[j] * sum[j]);
    const double v = g / (di * dj);
    out[(int64_t)i * count + j] = v;
    out[(int64_t)j * count + i] = v;
}

}  // namespace

int store_create(int64_t n, int capacity, int format, Store **out) {
    *out = nullptr;
    PRNU_REQUIRE(n >= 1 && capacity >= 1, "store_create: n and capacity must be positive");
    PRNU_REQUIRE(format == PRNU_PREC_SPLIT_BF16 || format == PRNU_PREC_FP32,
                 "store_create: format must be PRNU_PREC_SPLIT_BF16 or PRNU_PREC_FP32");
    Store *s = new (std::nothrow) Store();
    if (!s) {
        set_error("store_create: out of host memory");
        return PRNU_ENOMEM;
    }
    s->n = n;
    s->nkb = (n + kBlk - 1) / kBlk;
    s->capacity = capacity;
    s->cap_pad = (capacity + kTileS - 1) / kTileS * kTileS;
    s->format = format;
    (void)hipGetDevice(&s->device);
    const size_t bytes = (size_t)s->nkb * s->cap_pad * kBlkBytes;
    hipError_t e = hipMalloc(&s->blocks, bytes);
    if (e == hipSuccess) e = hipMalloc(&s->sum, sizeof(double) * s->cap_pad);
    if (e == hipSuccess) e = hipMalloc(&s->sumsq, sizeof(double) * s->cap_pad);
    if (e == hipSuccess) e = hipMalloc(&s->absmax, sizeof(float) * s->cap_pad);
    // empty slots read as zeros (the Gram masks them anyway)
    if (e == hipSuccess) e = hipMemset(s->blocks, 0, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("store_create: %zu bytes of device memory: %s", bytes, hipGetErrorString(e));
        if (s->blocks) (void)hipFree(s->blocks);
        if (s->sum) (void)hipFree(s->sum);
        if (s->sumsq) (void)hipFree(s->sumsq);
        if (s->absmax) (void)hipFree(s->absmax);
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
    s->ws_part.release();
    delete s;
}

int store_put(Store *s, const float *const *d_ptrs, int count, int first, hipStream_t st) {
    PRNU_REQUIRE(count >= 1 && first >= 0 && first + count <= s->capacity,
                 "store_put: slots [%d, %d) outside the capacity %d", first, first + count, s->capacity);
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
    hipLaunchKernelGGL(store_stats_final_kernel, dim3((count + 255) / 256), dim3(256), 0, st, part, (int)gx, count,
                       first, s->sum, s->sumsq, s->absmax);
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
    a.T = a.nt * (a.nt + 1) / 2;
    const prnu_gram_config cfg = c->gram_cfg[s->format];
    int64_t S = cfg.slabs;
    if (S <= 0) {
        S = std::max<int64_t>(1, (256 * 5 + a.T - 1) / a.T);
        S = std::min<int64_t>(S, std::max<int64_t>(1, s->nkb / 16));
    }
    S = std::min<int64_t>(S, s->nkb);
    const int64_t tile_bytes = (int64_t)kTileS * kTileS * sizeof(float);
    S = std::min<int64_t>(S, std::max<int64_t>(1, (int64_t(2) << 30) / (tile_bytes * a.T)));
    a.kcb = (s->nkb + S - 1) / S;
    a.S = (int)((s->nkb + a.kcb - 1) / a.kcb);
    a.xcd_remap = cfg.xcd_remap;
    PRNU_TRY(c->ws_gram.ensure((size_t)tile_bytes * a.T * a.S));
    a.partials = static_cast<float *>(c->ws_gram.ptr);
    const int64_t S8 = a.xcd_remap ? (a.S + 7) / 8 * 8 : a.S;
    const dim3 grid((unsigned)(a.T * S8));
    time_begin(c, st);
    if (s->format == PRNU_PREC_SPLIT_BF16)
        hipLaunchKernelGGL(store_gram_kernel<PRNU_PREC_SPLIT_BF16>, grid, dim3(kThrS), 0, st, a);
    else
        hipLaunchKernelGGL(store_gram_kernel<PRNU_PREC_FP32>, grid, dim3(kThrS), 0, st, a);
    PRNU_CHECK_LAUNCH();
    time_end(c, st);
    hipLaunchKernelGGL(store_reduce_kernel, dim3(kTileS * kTileS / 256, a.T), dim3(256), 0, st, a.partials, a.S, a.nt,
                       count, s->sum + first, s->sumsq + first, d_ncc);
    PRNU_CHECK_LAUNCH();
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

}  // extern "C"
