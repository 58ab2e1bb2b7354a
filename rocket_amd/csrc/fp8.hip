// FP8-e4m3 NCC on gfx950: per-pattern scaled quantisation and an MFMA
// (v_mfma_scale_f32_16x16x128_f8f6f4) tile kernel for C = A B^T over K, in
// two shapes: the upper triangle of one set (all-pairs Gram) and the full
// rectangle of a query block against a store (serving scan).
//
// Quantisation: scale_i = max|x_i| / 448 (e4m3 max normal), q = rne(x/scale).
// G_ij = scale_i * scale_j * sum_k q_ik q_jk, accumulated in FP32 per K slab
// and summed over slabs in FP64.  Representation error is <= 2^-4 relative per
// element; the FP32 re-score of the serving path removes it for the top hits.
#include "prnu_internal.hpp"

#include <climits>

#include <algorithm>

namespace prnu {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef int intx8 __attribute__((ext_vector_type(8)));

constexpr float kE4M3Max = 448.0f;
constexpr int kKT8 = 128;          // K bytes per stage = one MFMA K
constexpr int kThreads8 = 512;     // 8 waves
constexpr int kBTile = 256;        // B rows per tile

// ------------------------------------------------------------ quantisation
// grid (ceil(stride/4096), count), block 256; each thread packs 16 values.
__global__ __launch_bounds__(256) void quantize_fp8_kernel(const float *const *ptrs, int64_t n, int64_t stride,
                                                           const float *absmax, uint8_t *q, float *scale) {
    const int p = blockIdx.y;
    const float am = absmax[p];
    const float inv = am > 0.f ? kE4M3Max / am : 1.0f;
    if (blockIdx.x == 0 && threadIdx.x == 0) scale[p] = am > 0.f ? am / kE4M3Max : 1.0f;
    const float *x = ptrs[p];
    const int64_t k0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 16;
    if (k0 >= stride) return;
    int packed[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int64_t k = k0 + 4 * g + e;
            v[e] = k < n ? x[k] * inv : 0.f;
        }
        int w = 0;
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[0], v[1], w, false);
        w = __builtin_amdgcn_cvt_pk_fp8_f32(v[2], v[3], w, true);
        packed[g] = w;
    }
    *reinterpret_cast<int4 *>(q + (int64_t)p * stride + k0) = make_int4(packed[0], packed[1], packed[2], packed[3]);
}

// ------------------------------------------------------------- tile kernel
struct Fp8Args {
    const uint8_t *A;
    const uint8_t *B;
    int64_t lda, ldb;  // row strides in bytes (multiples of 128)
    int ma, nb;        // valid rows
    int64_t K;         // padded K (multiple of 128)
    int64_t kchunk;    // K per slab (multiple of 128)
    int tri;           // 1: upper triangle of A with itself (B == A)
    int nta, ntb, T, S;
    int xcd_remap;
    float *partials;
};

__device__ inline int2 tri_tile8(int t, int nt) {
    int I = 0;
    while (t >= nt - I) {
        t -= nt - I;
        ++I;
    }
    return make_int2(I, I + t);
}

// 16-byte chunks of a 128-byte row, XOR-swizzled by row (conflict-free
// ds_read_b128 for 8 consecutive rows).
__device__ __forceinline__ int q_off(int row, int chunk) { return row * 128 + ((chunk ^ (row & 7)) << 4); }

template <int AR>
__global__ __launch_bounds__(kThreads8, 1) void fp8_tile_kernel(Fp8Args a) {
    constexpr int kRowsL = AR + kBTile;
    constexpr int kBuf = kRowsL * kKT8;                  // bytes per stage buffer
    constexpr int WM = AR == 256 ? 2 : 1, WN = 8 / WM;   // wave grid
    constexpr int RB = AR / WM / 16, CB = kBTile / WN / 16;
    constexpr int kLoads = kRowsL * 8 / kThreads8;       // 16-byte chunks per thread per stage
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
    int I, J;
    if (a.tri) {
        const int2 ij = tri_tile8(t, a.nta);
        I = ij.x;
        J = ij.y;
    } else {
        I = t / a.ntb;
        J = t % a.ntb;
    }
    const bool diag = a.tri && I == J;
    const int nrows = diag ? AR : kRowsL;
    const int bbase = diag ? 0 : AR;
    const int64_t k_begin = (int64_t)s * a.kchunk;
    const int64_t k_end = min(a.K, k_begin + a.kchunk);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wrow0 = (wave / WN) * (AR / WM);
    const int wcol0 = (wave % WN) * (kBTile / WN);

    // loader: chunk c = tid + 512*it -> row c >> 3, 16-byte column c & 7.
    // Row addresses are recomputed per load (row clamped into the set) to
    // keep registers for the accumulators; absent rows are zeroed on store.
    const int r0 = tid >> 3;
    const int col = (tid & 7) << 4;
    uint32_t valid = 0;
#pragma unroll
    for (int it = 0; it < kLoads; ++it) {
        const int r = r0 + 64 * it;
        const bool v = (r < AR) ? (I * AR + r < a.ma) : (r < nrows && J * kBTile + (r - AR) < a.nb);
        valid |= (uint32_t)v << it;
    }
    auto row_ptr = [&](int it) -> const uint8_t * {
        const int r = r0 + 64 * it;
        if (r < AR) {
            const int g = min(I * AR + r, a.ma - 1);
            return a.A + (int64_t)g * a.lda + col;
        }
        const int g = min(J * kBTile + (r - AR), a.nb - 1);
        return a.B + (int64_t)g * a.ldb + col;
    };

    floatx4 acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    int4 st[kLoads];
    auto load = [&](int64_t k0) {
#pragma unroll
        for (int it = 0; it < kLoads; ++it) {
            typedef const __attribute__((address_space(1))) int __attribute__((ext_vector_type(4))) *gi4;
            const auto v = *(gi4)(row_ptr(it) + k0);
            st[it] = make_int4(v[0], v[1], v[2], v[3]);
        }
    };
    auto store = [&](uint8_t *buf) {
#pragma unroll
        for (int it = 0; it < kLoads; ++it) {
            const int c = tid + kThreads8 * it;
            const int r = c >> 3;
            if (r < nrows) {
                int4 v = st[it];
                if (!((valid >> it) & 1)) v = make_int4(0, 0, 0, 0);
                *reinterpret_cast<int4 *>(buf + q_off(r, c & 7)) = v;
            }
        }
    };
    auto compute = [&](const uint8_t *buf) {
        const int frow = lane & 15, ch = (lane >> 4) * 2;
        intx8 bf[CB];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
            const int r = bbase + wcol0 + cb * 16 + frow;
            const int4 lo = *reinterpret_cast<const int4 *>(buf + q_off(r, ch));
            const int4 hi = *reinterpret_cast<const int4 *>(buf + q_off(r, ch + 1));
            bf[cb] = intx8{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        }
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const int r = wrow0 + rb * 16 + frow;
            const int4 lo = *reinterpret_cast<const int4 *>(buf + q_off(r, ch));
            const int4 hi = *reinterpret_cast<const int4 *>(buf + q_off(r, ch + 1));
            const intx8 af = intx8{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
                acc[rb][cb] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(af, bf[cb], acc[rb][cb], 0, 0, 0, 127,
                                                                               0, 127);
        }
    };

    const int64_t nst = (k_end - k_begin) / kKT8;
    load(k_begin);
    store(lds);
    __syncthreads();
    for (int64_t q = 0; q < nst; ++q) {
        const bool next = q + 1 < nst;
        if (next) load(k_begin + (q + 1) * kKT8);
        compute(lds + (q & 1) * kBuf);
        if (next) store(lds + ((q + 1) & 1) * kBuf);
        __syncthreads();
    }

    float *out = a.partials + ((int64_t)t * a.S + s) * (AR * kBTile);
    const int ocol = lane & 15, orow = (lane >> 4) * 4;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(wrow0 + rb * 16 + orow + r) * kBTile + wcol0 + cb * 16 + ocol] = acc[rb][cb][r];
}

// Query scan: 64 query rows against 256 store rows per tile, the store
// streamed once per call.  The work per stage is a few MFMAs against 32 KiB
// of store rows, so the kernel is bound by how many bytes it keeps in flight:
// a ring of kScanRing LDS stages filled by direct global-to-LDS loads
// (global_load_lds_dwordx4), kScanRing - 1 of them in flight while one is
// multiplied.  Each stage starts with an explicit wait for this wave's loads
// of that stage and an s_barrier in inline asm (a memory clobber, no fence):
// a workgroup fence would wait for every load in flight, which is what the
// ring avoids.  Rows past ma / nb
// hold clamped duplicates; they only feed partial rows and columns the
// reduction skips.  Same partial layout as fp8_tile_kernel<64>.
constexpr int kScanRing = 4;
constexpr int kScanRows = 64 + kBTile;          // LDS rows per stage
constexpr int kScanBuf = kScanRows * kKT8;      // 40 KiB per stage
constexpr int kScanLoads = kScanRows * 8 / kThreads8;  // 16-byte loads per thread per stage

__global__ __launch_bounds__(kThreads8, 1) void fp8_scan_kernel(Fp8Args a) {
    constexpr int RB = 4, CB = 2;  // wave tile: 64 rows x 32 columns (8 waves across the 256 columns)
    __shared__ __attribute__((aligned(16))) uint8_t lds[kScanRing * kScanBuf];

    const int bid = blockIdx.x;
    const int local = bid >> 3;
    const int t = local % a.T;
    const int s = (local / a.T) * 8 + (bid & 7);
    if (s >= a.S) return;
    const int I = t / a.ntb, J = t % a.ntb;
    const int64_t k_begin = (int64_t)s * a.kchunk;
    const int64_t k_end = min(a.K, k_begin + a.kchunk);
    const int nst = (int)((k_end - k_begin) / kKT8);

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wcol0 = wave * 32;

    // load `it` of a stage fills LDS rows 64*it + 8*wave .. +8, lane l at
    // byte 16*l: row 64*it + (tid >> 3), 16-byte slot tid & 7, which holds
    // chunk (tid & 7) ^ (row & 7) of the row (the swizzle q_off reads)
    const int r0 = tid >> 3;
    const int col = ((tid & 7) ^ (r0 & 7)) << 4;
    const uint8_t *src[kScanLoads];
#pragma unroll
    for (int it = 0; it < kScanLoads; ++it) {
        const int r = r0 + 64 * it;
        if (r < 64) {
            const int g = min(I * 64 + r, a.ma - 1);
            src[it] = a.A + (int64_t)g * a.lda + col + k_begin;
        } else {
            const int g = min(J * kBTile + (r - 64), a.nb - 1);
            src[it] = a.B + (int64_t)g * a.ldb + col + k_begin;
        }
    }
    auto issue = [&](int q) {
        uint8_t *buf = lds + (q % kScanRing) * kScanBuf;
        const int64_t k0 = (int64_t)q * kKT8;
#pragma unroll
        for (int it = 0; it < kScanLoads; ++it) {
            typedef const __attribute__((address_space(1))) void *gptr;
            typedef __attribute__((address_space(3))) void *lptr;
            __builtin_amdgcn_global_load_lds((gptr)(src[it] + k0), (lptr)(buf + (64 * it + 8 * wave) * kKT8), 16, 0,
                                             0);
        }
    };

    floatx4 acc[RB][CB];
#pragma unroll
    for (int i = 0; i < RB; ++i)
#pragma unroll
        for (int j = 0; j < CB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

#pragma unroll
    for (int d = 0; d < kScanRing - 1; ++d)
        if (d < nst) issue(d);
    const int frow = lane & 15, ch = (lane >> 4) * 2;
    const uint32_t lds_base = (uint32_t)(uintptr_t)(__attribute__((address_space(3))) uint8_t *)lds;
    const uint32_t b_off_lo = (uint32_t)q_off(64 + wcol0 + frow, ch), b_off_hi = (uint32_t)q_off(64 + wcol0 + frow, ch + 1);
    const uint32_t a_off_lo = (uint32_t)q_off(frow, ch), a_off_hi = (uint32_t)q_off(frow, ch + 1);
    for (int q = 0; q < nst; ++q) {
        // stages issued past q: min(nst - 1 - q, kScanRing - 2); wait until
        // only those are outstanding, then every wave's stage q has landed
        // and every wave is done reading the buffer stage q + 3 will fill
        const int ahead = min(nst - 1 - q, kScanRing - 2);
        static_assert(kScanRing == 4 && kScanLoads == 5, "the waits below count 5 loads per stage");
        // s_waitcnt immediates: vmcnt in bits 3:0, expcnt 6:4 (7 = no wait),
        // lgkmcnt 11:8 (0 = this wave's LDS reads of the last stage are done)
        if (ahead >= 2)
            __builtin_amdgcn_s_waitcnt(0x007A);  // vmcnt(10) lgkmcnt(0)
        else if (ahead == 1)
            __builtin_amdgcn_s_waitcnt(0x0075);  // vmcnt(5) lgkmcnt(0)
        else
            __builtin_amdgcn_s_waitcnt(0x0070);  // vmcnt(0) lgkmcnt(0)
        asm volatile("s_barrier" ::: "memory");
        if (q + kScanRing - 1 < nst) issue(q + kScanRing - 1);
        // this wave's fragments of the stage: B rows 64 + wcol0 + {0, 16} + frow
        // and A rows {0, 16, 32, 48} + frow, chunks ch and ch + 1.  Rows 16
        // apart are 2048 bytes apart under the swizzle (16 is a multiple of
        // 8), so four addresses and immediate offsets cover the 12 reads.
        // They are one asm block ending in lgkmcnt(0): as plain LDS loads the
        // compiler would wait for every LDS fill in flight before them.
        const uint32_t sb = lds_base + (uint32_t)((q % kScanRing) * kScanBuf);
        const uint32_t b_lo = sb + b_off_lo, b_hi = sb + b_off_hi, a_lo = sb + a_off_lo, a_hi = sb + a_off_hi;
        int4 f[12];
        asm volatile(
            "ds_read_b128 %0, %12\n\t"
            "ds_read_b128 %1, %13\n\t"
            "ds_read_b128 %2, %12 offset:2048\n\t"
            "ds_read_b128 %3, %13 offset:2048\n\t"
            "ds_read_b128 %4, %14\n\t"
            "ds_read_b128 %5, %15\n\t"
            "ds_read_b128 %6, %14 offset:2048\n\t"
            "ds_read_b128 %7, %15 offset:2048\n\t"
            "ds_read_b128 %8, %14 offset:4096\n\t"
            "ds_read_b128 %9, %15 offset:4096\n\t"
            "ds_read_b128 %10, %14 offset:6144\n\t"
            "ds_read_b128 %11, %15 offset:6144\n\t"
            "s_waitcnt lgkmcnt(0)"
            : "=&v"(f[0]), "=&v"(f[1]), "=&v"(f[2]), "=&v"(f[3]), "=&v"(f[4]), "=&v"(f[5]), "=&v"(f[6]),
              "=&v"(f[7]), "=&v"(f[8]), "=&v"(f[9]), "=&v"(f[10]), "=&v"(f[11])
            : "v"(b_lo), "v"(b_hi), "v"(a_lo), "v"(a_hi)
            : "memory");
        intx8 bf[CB];
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
            const int4 lo = f[2 * cb], hi = f[2 * cb + 1];
            bf[cb] = intx8{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
        }
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
            const int4 lo = f[4 + 2 * rb], hi = f[5 + 2 * rb];
            const intx8 af = intx8{lo.x, lo.y, lo.z, lo.w, hi.x, hi.y, hi.z, hi.w};
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
                acc[rb][cb] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(af, bf[cb], acc[rb][cb], 0, 0, 0, 127,
                                                                               0, 127);
        }
    }

    float *out = a.partials + ((int64_t)t * a.S + s) * (64 * kBTile);
    const int ocol = lane & 15, orow = (lane >> 4) * 4;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
        for (int cb = 0; cb < CB; ++cb)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                out[(rb * 16 + orow + r) * kBTile + wcol0 + cb * 16 + ocol] = acc[rb][cb][r];
}

// Sum the slabs in FP64, undo the scales and apply the NCC epilogue
// G / (sqrt(Sxx_i - Sx_i^2) sqrt(Sxx_j - Sx_j^2)).  Four consecutive elements
// per thread (one 16-byte load per slab, each element summed in slab order).
// grid (AR*256/1024, T).
template <int AR>
__global__ __launch_bounds__(256) void fp8_reduce_kernel(Fp8Args a, const float *scale_a, const float *scale_b,
                                                         const double *sum_a, const double *sumsq_a,
                                                         const double *sum_b, const double *sumsq_b, double *out,
                                                         int64_t ld_out) {
    const int t = blockIdx.y;
    const int e = (blockIdx.x * 256 + threadIdx.x) * 4;
    int I, J;
    if (a.tri) {
        const int2 ij = tri_tile8(t, a.nta);
        I = ij.x;
        J = ij.y;
    } else {
        I = t / a.ntb;
        J = t % a.ntb;
    }
    const int i = I * AR + e / kBTile;
    const int j0 = J * kBTile + e % kBTile;
    if (i >= a.ma || j0 >= a.nb) return;
    if (a.tri && I == J && i > j0 + 3) return;
    const float4 *p = reinterpret_cast<const float4 *>(a.partials + (int64_t)t * a.S * (AR * kBTile) + e);
    double g[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
    for (int q = 0; q < a.S; ++q) {
        const float4 v = p[(int64_t)q * (AR * kBTile / 4)];
        g[0] += (double)v.x;
        g[1] += (double)v.y;
        g[2] += (double)v.z;
        g[3] += (double)v.w;
    }
    const double di = sqrt(sumsq_a[i] - sum_a[i] * sum_a[i]);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int j = j0 + c;
        if (j >= a.nb || (a.tri && I == J && i > j)) continue;
        const double dj = sqrt(sumsq_b[j] - sum_b[j] * sum_b[j]);
        const double v = g[c] * ((double)scale_a[i] * (double)scale_b[j]) / (di * dj);
        out[(int64_t)i * ld_out + j] = v;
        if (a.tri) out[(int64_t)j * ld_out + i] = v;
    }
}

}  // namespace

int64_t fp8_row_stride(int64_t n) { return (n + kKT8 - 1) / kKT8 * kKT8; }

int launch_quantize_fp8(const float *const *d_ptrs, int count, int64_t n, const float *d_absmax, uint8_t *d_q,
                        float *d_scale, hipStream_t st) {
    PRNU_REQUIRE(count >= 1 && count <= 65535 && n >= 1, "quantize_fp8: bad count or length");
    const int64_t stride = fp8_row_stride(n);
    dim3 grid((unsigned)((stride + 4095) / 4096), (unsigned)count);
    hipLaunchKernelGGL(quantize_fp8_kernel, grid, dim3(256), 0, st, d_ptrs, n, stride, d_absmax, d_q, d_scale);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

// Workgroups of the a_rows kernel (fp8_tile_kernel<256> or fp8_scan_kernel)
// the device runs at once (per device, queried once).
int64_t fp8_slots(int a_rows) {
    static int64_t cached[16][2] = {};
    int dev = 0;
    (void)hipGetDevice(&dev);
    const int v = a_rows == 256 ? 0 : 1;
    if (dev < 0 || dev >= 16) dev = 0;
    if (!cached[dev][v]) {
        int per_cu = 0, cus = 0;
        if (a_rows == 256)
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fp8_tile_kernel<256>, kThreads8, 0);
        else
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fp8_scan_kernel, kThreads8, 0);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipGetLastError();
        cached[dev][v] = (int64_t)std::max(1, per_cu) * std::max(1, cus);
    }
    return cached[dev][v];
}

// C = A B^T in FP8 with NCC epilogue.  tri: B is A and only I <= J tiles run.
// a_rows: 256 (Gram) or 64 (query blocks).
int launch_fp8_ncc(Workspace &ws, int a_rows, bool tri, const uint8_t *A, int64_t lda, int ma, const float *scale_a,
                   const double *sum_a, const double *sumsq_a, const uint8_t *B, int64_t ldb, int nb,
                   const float *scale_b, const double *sum_b, const double *sumsq_b, int64_t K, double *out,
                   int64_t ld_out, int slabs, hipStream_t st, Context *timing_ctx) {
    PRNU_REQUIRE(a_rows == 256 || a_rows == 64, "fp8: a_rows must be 256 or 64");
    PRNU_REQUIRE(K % kKT8 == 0 && lda % kKT8 == 0 && ldb % kKT8 == 0, "fp8: K and strides must be multiples of 128");
    Fp8Args a;
    a.A = A;
    a.B = tri ? A : B;
    a.lda = lda;
    a.ldb = tri ? lda : ldb;
    a.ma = ma;
    a.nb = tri ? ma : nb;
    a.K = K;
    a.tri = tri ? 1 : 0;
    a.nta = (ma + a_rows - 1) / a_rows;
    a.ntb = (a.nb + kBTile - 1) / kBTile;
    PRNU_REQUIRE(!tri || a_rows == kBTile, "fp8: the triangle needs square tiles");
    a.T = tri ? a.nta * (a.nta + 1) / 2 : a.nta * a.ntb;
    const int64_t stages = K / kKT8;
    const int64_t tile_bytes = (int64_t)a_rows * kBTile * sizeof(float);
    const int64_t s_max = std::max<int64_t>(
        1, std::min<int64_t>(std::min<int64_t>(stages / 8, 1024), (int64_t(2) << 30) / (tile_bytes * a.T)));
    int64_t S;
    if (slabs > 0) {
        S = std::min<int64_t>(slabs, s_max);
    } else {
        // the slab count minimising dispatch rounds x (slab length + a fixed
        // cost per slab: filling the load pipeline, writing the partial tile
        // and reading it back, about kSlabOverhead stages), with the
        // workgroups the device runs at once (a function of the tile count,
        // K and the device only: a query's scores do not depend on its batch)
        constexpr int64_t kSlabOverhead = 8;
        const int64_t slots = fp8_slots(a_rows);
        int64_t best = INT64_MAX;
        S = 1;
        for (int64_t c = 1; c <= s_max; ++c) {
            const int64_t cost = ((a.T * c + slots - 1) / slots) * ((stages + c - 1) / c + kSlabOverhead);
            if (cost < best) {
                best = cost;
                S = c;
            }
        }
    }
    a.kchunk = ((stages + S - 1) / S) * kKT8;
    a.S = (int)((K + a.kchunk - 1) / a.kchunk);
    a.xcd_remap = 1;
    PRNU_TRY(ws.ensure((size_t)tile_bytes * a.T * a.S));
    a.partials = static_cast<float *>(ws.ptr);
    const dim3 grid((unsigned)(a.T * ((a.S + 7) / 8 * 8)));
    if (timing_ctx) time_begin(timing_ctx, st);
    if (a_rows == 256)
        hipLaunchKernelGGL(fp8_tile_kernel<256>, grid, dim3(kThreads8), 0, st, a);
    else
        hipLaunchKernelGGL(fp8_scan_kernel, grid, dim3(kThreads8), 0, st, a);
    PRNU_CHECK_LAUNCH();
    if (timing_ctx) time_end(timing_ctx, st);
    const dim3 rgrid((unsigned)(a_rows * kBTile / 1024), (unsigned)a.T);
    if (a_rows == 256)
        hipLaunchKernelGGL(fp8_reduce_kernel<256>, rgrid, dim3(256), 0, st, a, scale_a, tri ? scale_a : scale_b,
                           sum_a, sumsq_a, tri ? sum_a : sum_b, tri ? sumsq_a : sumsq_b, out, ld_out);
    else
        hipLaunchKernelGGL(fp8_reduce_kernel<64>, rgrid, dim3(256), 0, st, a, scale_a, scale_b, sum_a, sumsq_a,
                           sum_b, sumsq_b, out, ld_out);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

}  // namespace prnu
