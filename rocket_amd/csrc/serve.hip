// Camera-identification serving on gfx950: a device-resident fingerprint
// store (FP8-e4m3 copy for scanning, FP32 originals for exact re-scoring,
// FP64 statistics) and the query path
//
//   query patterns -> FP64 stats -> FP8 quantisation -> FP8 MFMA scan of the
//   whole store (NCC estimate per (query, fingerprint)) -> per-query top-k
//   -> exact NCC of the k candidates (FP32 products, FP64 sums, reference
//   formula of normalizedcrosscorrelation.cu:121-123) -> sort.
//
// Every per-query result depends only on that query: the scan runs the
// queries as independent rows of 64-row tiles with a K split fixed by the
// store size, top-k and re-score are per query, so a query scores the same
// alone and in any batch.
#include "prnu_internal.hpp"

#include <algorithm>
#include <new>

namespace prnu {

struct FpStore {
    int64_t n = 0, stride = 0;
    int capacity = 0, size = 0;  // size: slots filled (1 + highest slot put)
    uint8_t *q = nullptr;        // capacity x stride, e4m3
    float *scale = nullptr;      // capacity
    float *f32 = nullptr;        // capacity x n, exact copy
    double *sum = nullptr, *sumsq = nullptr;
    float *absmax = nullptr;
    Workspace ws_stats;
};

namespace {

constexpr int kMaxK = 16;
constexpr int kQueryBlock = 64;  // queries per scan tile (fp8_tile_kernel<64>)

// out[i] row copy: dst + (first + i) * n <- ptrs[i]
__global__ __launch_bounds__(256) void copy_rows_kernel(const float *const *ptrs, int64_t n, float *dst, int first) {
    const float *src = ptrs[blockIdx.y];
    float *d = dst + (int64_t)(first + blockIdx.y) * n;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) d[i] = src[i];
}

// Top-k of a range of one row of `scores`, k <= kMaxK: one 256-thread block
// per range.  Each thread keeps the best kMaxK of its strided share, sorted,
// so the union of the heads contains the range's top k; then k rounds of a
// block arg-max over the list heads.  Ties go to the smaller index
// (deterministic).
__device__ __forceinline__ bool better(double a, int ia, double b, int ib) { return a > b || (a == b && ia < ib); }

// grid (rows, kTopkChunks): block (q, c) takes the top `k` of columns
// [c * chunk_len, (c + 1) * chunk_len) of row q into out_idx[q * out_ld + c * k
// ...] (-1 past the chunk's end).  With k = kMaxK per chunk the chunk lists
// hold every member of the row's top k, which topk_merge_kernel picks out.
constexpr int kTopkChunks = 8;

__global__ __launch_bounds__(256) void topk_kernel(const double *scores, int64_t ld, int N, int k, int *out_idx,
                                                   int chunk_len, int64_t out_ld) {
    __shared__ double sv[256 * kMaxK];
    __shared__ int si[256 * kMaxK];
    __shared__ double wv[4];
    __shared__ int wi[4], wt[4];
    const double *row = scores + (int64_t)blockIdx.x * ld;
    const int j0 = blockIdx.y * chunk_len, j1 = min(N, j0 + chunk_len);
    double v[kMaxK];
    int ix[kMaxK];
#pragma unroll
    for (int q = 0; q < kMaxK; ++q) {
        v[q] = -1e300;
        ix[q] = 0x7fffffff;
    }
    for (int j = j0 + threadIdx.x; j < j1; j += 256) {
        double cv = row[j];
        int ci = j;
        if (better(cv, ci, v[kMaxK - 1], ix[kMaxK - 1])) {
#pragma unroll
            for (int q = 0; q < kMaxK; ++q) {
                if (better(cv, ci, v[q], ix[q])) {
                    const double tv = v[q];
                    const int ti = ix[q];
                    v[q] = cv;
                    ix[q] = ci;
                    cv = tv;
                    ci = ti;
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kMaxK; ++q) {
        sv[threadIdx.x * kMaxK + q] = v[q];
        si[threadIdx.x * kMaxK + q] = ix[q];
    }
    int head = 0;
    __syncthreads();
    for (int r = 0; r < k; ++r) {
        double bv = head < kMaxK ? sv[threadIdx.x * kMaxK + head] : -1e300;
        int bi = head < kMaxK ? si[threadIdx.x * kMaxK + head] : 0x7fffffff;
        int bt = threadIdx.x;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(bv, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            const int ot = __shfl_xor(bt, off, 64);
            if (better(ov, oi, bv, bi)) {
                bv = ov;
                bi = oi;
                bt = ot;
            }
        }
        if ((threadIdx.x & 63) == 0) {
            wv[threadIdx.x >> 6] = bv;
            wi[threadIdx.x >> 6] = bi;
            wt[threadIdx.x >> 6] = bt;
        }
        __syncthreads();
        bv = wv[0];
        bi = wi[0];
        bt = wt[0];
        for (int q = 1; q < 4; ++q)
            if (better(wv[q], wi[q], bv, bi)) {
                bv = wv[q];
                bi = wi[q];
                bt = wt[q];
            }
        if (threadIdx.x == 0) out_idx[(int64_t)blockIdx.x * out_ld + (int64_t)blockIdx.y * k + r] = bi < N ? bi : -1;
        if (threadIdx.x == bt) ++head;
        __syncthreads();
    }
}

// grid rows, block 64: the top k of row q among its kTopkChunks * kMaxK
// chunk candidates cand[q * kTopkChunks * kMaxK ...] (-1: none), ties to the
// smaller index, into out_idx[q * k ...].
__global__ __launch_bounds__(64) void topk_merge_kernel(const double *scores, int64_t ld, const int *cand, int k,
                                                        int *out_idx) {
    constexpr int kC = kTopkChunks * kMaxK / 64;  // candidates per lane
    const int q = blockIdx.x, lane = threadIdx.x;
    const double *row = scores + (int64_t)q * ld;
    double v[kC];
    int ix[kC];
#pragma unroll
    for (int e = 0; e < kC; ++e) {
        const int c = cand[(int64_t)q * kTopkChunks * kMaxK + e * 64 + lane];
        v[e] = c >= 0 ? row[c] : -1e300;
        ix[e] = c >= 0 ? c : 0x7fffffff;
    }
    for (int r = 0; r < k; ++r) {
        double bv = v[0];
        int bi = ix[0], be = 0;
#pragma unroll
        for (int e = 1; e < kC; ++e)
            if (better(v[e], ix[e], bv, bi)) {
                bv = v[e];
                bi = ix[e];
                be = e;
            }
        double wv = bv;
        int wi = bi;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const double ov = __shfl_xor(wv, off, 64);
            const int oi = __shfl_xor(wi, off, 64);
            if (better(ov, oi, wv, wi)) {
                wv = ov;
                wi = oi;
            }
        }
        if (lane == 0) out_idx[(int64_t)q * k + r] = wi == 0x7fffffff ? -1 : wi;
        if (bi == wi && wi != 0x7fffffff) {  // indices are distinct: one lane holds the winner
#pragma unroll
            for (int e = 0; e < kC; ++e)
                if (e == be) {
                    v[e] = -1e300;
                    ix[e] = 0x7fffffff;
                }
        }
    }
}

// Exact NCC of each query with its k candidates, in two passes over a fixed
// split of the n values into slices (a function of n only, so a query's
// scores do not depend on its batch).  Pass 1, grid (slices, nq), block 256:
// a block reads its slice of the query once and the same slice of every
// candidate, FP32 products summed in FP64 (16-byte loads where every row is
// 16-byte aligned), and writes one partial per candidate.  Pass 2, grid nq:
// the partials of each candidate summed in slice order, then the NCC.
constexpr int64_t kRescoreSliceMin = 16384;  // values per slice, at least

int64_t rescore_slices(int64_t n) { return std::min<int64_t>(256, std::max<int64_t>(1, n / kRescoreSliceMin)); }

int64_t rescore_slice_len(int64_t n) {
    const int64_t sl = rescore_slices(n);
    return ((n + sl - 1) / sl + 3) / 4 * 4;
}

__global__ __launch_bounds__(256) void rescore_slice_kernel(const float *const *qptrs, const float *f32, int64_t n,
                                                            int64_t slice_len, const int *idx, int k,
                                                            double *part) {
    __shared__ double sh[4][kMaxK];
    const int sl = blockIdx.x, slices = gridDim.x, q = blockIdx.y;
    const int64_t lo = (int64_t)sl * slice_len, hi = min(n, lo + slice_len);
    const float *x = qptrs[q];
    const bool vec = ((uintptr_t)x & 15) == 0 && (n & 3) == 0 && ((uintptr_t)f32 & 15) == 0;
    // All kMaxK slots run, branch-free: a slot past k or an absent candidate
    // (index < 0) re-reads the query row (cache hits) and is not written;
    // the final pass writes -1e300 for an absent one.  Global-address-space
    // pointers, so the loads are global_load (the tables hold generic ones).
    typedef float fv4 __attribute__((ext_vector_type(4)));
    typedef const __attribute__((address_space(1))) float *gf;
    typedef const __attribute__((address_space(1))) fv4 *gf4;
    const gf xg = (gf)x;
    gf y[kMaxK];
    double s[kMaxK];
#pragma unroll
    for (int r = 0; r < kMaxK; ++r) {
        const int c = r < k ? idx[q * k + r] : -1;
        y[r] = c >= 0 ? (gf)(f32 + (int64_t)c * n) : xg;
        s[r] = 0.0;
    }
    if (vec) {
        for (int64_t i = lo + 4 * threadIdx.x; i < hi; i += 1024) {
            const fv4 xv = *(gf4)(xg + i);
            fv4 yv[kMaxK];
#pragma unroll
            for (int r = 0; r < kMaxK; ++r) yv[r] = *(gf4)(y[r] + i);
#pragma unroll
            for (int r = 0; r < kMaxK; ++r) {
                const float p0 = xv.x * yv[r].x, p1 = xv.y * yv[r].y, p2 = xv.z * yv[r].z, p3 = xv.w * yv[r].w;
                s[r] += (double)p0;
                s[r] += (double)p1;
                s[r] += (double)p2;
                s[r] += (double)p3;
            }
        }
    } else {
        for (int64_t i = lo + threadIdx.x; i < hi; i += 256) {
            const float xv = xg[i];
            float yv[kMaxK];
#pragma unroll
            for (int r = 0; r < kMaxK; ++r) yv[r] = y[r][i];
#pragma unroll
            for (int r = 0; r < kMaxK; ++r) {
                const float p = xv * yv[r];
                s[r] += (double)p;
            }
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int r = 0; r < kMaxK; ++r) {
        double v = s[r];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) sh[wave][r] = v;
    }
    __syncthreads();
    if (threadIdx.x < k)
        part[((int64_t)q * k + threadIdx.x) * slices + sl] =
            (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

// grid nq, block 64: thread r < k finishes candidate r of query blockIdx.x.
__global__ __launch_bounds__(64) void rescore_final_kernel(const double *part, int slices, const double *qsum,
                                                           const double *qsumsq, const double *fsum,
                                                           const double *fsumsq, const int *idx, int k,
                                                           double *score) {
    const int q = blockIdx.x, r = threadIdx.x;
    if (r >= k) return;
    const int c = idx[q * k + r];
    if (c < 0) {
        score[q * k + r] = -1e300;
        return;
    }
    const double *p = part + ((int64_t)q * k + r) * slices;
    double sxy = 0.0;
    for (int i = 0; i < slices; ++i) sxy += p[i];
    const double dx = sqrt(qsumsq[q] - qsum[q] * qsum[q]);
    const double dy = sqrt(fsumsq[c] - fsum[c] * fsum[c]);
    score[q * k + r] = sxy / (dx * dy);
}

// Sort each query's k (score, index) by descending exact score (ties: index).
__global__ void sort_k_kernel(double *score, int *idx, int nq, int k) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    double *s = score + (int64_t)q * k;
    int *ix = idx + (int64_t)q * k;
    for (int a = 1; a < k; ++a) {
        const double v = s[a];
        const int iv = ix[a];
        int b = a - 1;
        while (b >= 0 && better(v, iv, s[b], ix[b])) {
            s[b + 1] = s[b];
            ix[b + 1] = ix[b];
            --b;
        }
        s[b + 1] = v;
        ix[b + 1] = iv;
    }
}

size_t a256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

int fpstore_create(int64_t n, int capacity, FpStore **out) {
    *out = nullptr;
    PRNU_REQUIRE(n >= 1 && capacity >= 1, "fpstore_create: n and capacity must be positive");
    FpStore *s = new (std::nothrow) FpStore();
    if (!s) {
        set_error("fpstore_create: out of host memory");
        return PRNU_ENOMEM;
    }
    s->n = n;
    s->stride = fp8_row_stride(n);
    s->capacity = capacity;
    hipError_t e = hipMalloc(&s->q, (size_t)s->stride * capacity);
    if (e == hipSuccess) e = hipMalloc(&s->f32, sizeof(float) * (size_t)n * capacity);
    if (e == hipSuccess) e = hipMalloc(&s->scale, sizeof(float) * capacity);
    if (e == hipSuccess) e = hipMalloc(&s->sum, sizeof(double) * capacity);
    if (e == hipSuccess) e = hipMalloc(&s->sumsq, sizeof(double) * capacity);
    if (e == hipSuccess) e = hipMalloc(&s->absmax, sizeof(float) * capacity);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("fpstore_create: device memory for %d fingerprints of %lld values: %s", capacity, (long long)n,
                  hipGetErrorString(e));
        for (void *p : {(void *)s->q, (void *)s->f32, (void *)s->scale, (void *)s->sum, (void *)s->sumsq,
                        (void *)s->absmax})
            if (p) (void)hipFree(p);
        delete s;
        return PRNU_ENOMEM;
    }
    *out = s;
    return PRNU_OK;
}

void fpstore_destroy(FpStore *s) {
    if (!s) return;
    (void)hipDeviceSynchronize();
    for (void *p : {(void *)s->q, (void *)s->f32, (void *)s->scale, (void *)s->sum, (void *)s->sumsq,
                    (void *)s->absmax})
        if (p) (void)hipFree(p);
    s->ws_stats.release();
    delete s;
}

int fpstore_put(FpStore *s, const float *const *d_ptrs, int count, int first, hipStream_t st) {
    PRNU_REQUIRE(count >= 1 && first >= 0 && first + count <= s->capacity,
                 "fpstore_put: slots [%d, %d) outside the capacity %d", first, first + count, s->capacity);
    // the filled slots stay the prefix [0, size) that every query scans: a put
    // may overwrite filled slots or extend the prefix, never leave a gap of
    // slots whose memory was never written
    PRNU_REQUIRE(first <= s->size, "fpstore_put: slots [%d, %d) would leave slots [%d, %d) unfilled", first,
                 first + count, s->size, first);
    PRNU_TRY(launch_pattern_stats(d_ptrs, count, s->n, s->sum + first, s->sumsq + first, s->absmax + first,
                                  s->ws_stats, st));
    PRNU_TRY(launch_quantize_fp8(d_ptrs, count, s->n, s->absmax + first, s->q + (size_t)first * s->stride,
                                 s->scale + first, st));
    const unsigned gx = (unsigned)std::min<int64_t>(1024, (s->n + 255) / 256);
    hipLaunchKernelGGL(copy_rows_kernel, dim3(gx, count), dim3(256), 0, st, d_ptrs, s->n, s->f32, first);
    PRNU_CHECK_LAUNCH();
    s->size = std::max(s->size, first + count);
    return PRNU_OK;
}

// nq query patterns (device pointer table) against slots [0, size).
int fpstore_query(Context *c, FpStore *s, const float *const *d_qptrs, int nq, int k, int *d_idx, double *d_score,
                  double *d_scan, hipStream_t st) {
    PRNU_REQUIRE(nq >= 1 && nq <= 65535, "query: 1..65535 queries per call");
    PRNU_REQUIRE(k >= 1 && k <= kMaxK, "query: k must be in [1, %d]", kMaxK);
    PRNU_REQUIRE(s->size >= 1, "query: the store is empty");
    PRNU_REQUIRE(c->pixels() == s->n, "query: context size %lld does not match the store (%lld)",
                 (long long)c->pixels(), (long long)s->n);
    const int N = s->size;
    PRNU_REQUIRE(k <= N, "query: k=%d exceeds the %d stored fingerprints", k, N);
    // workspace: query stats, FP8 queries + scales, scan scores
    const size_t b_stats = a256(sizeof(double) * nq) * 2 + a256(sizeof(float) * nq);
    const size_t b_q = a256((size_t)s->stride * nq) + a256(sizeof(float) * nq);
    const size_t b_scan = d_scan ? 0 : a256(sizeof(double) * (size_t)nq * N);
    const int64_t slices = rescore_slices(s->n);
    const size_t b_part = a256(sizeof(double) * (size_t)nq * k * slices);
    const size_t b_cand = a256(sizeof(int) * (size_t)nq * kTopkChunks * kMaxK);
    PRNU_TRY(c->ws_fp8.ensure(b_stats + b_q + b_scan + b_part + b_cand));
    char *p = static_cast<char *>(c->ws_fp8.ptr);
    double *qsum = reinterpret_cast<double *>(p);
    double *qsumsq = reinterpret_cast<double *>(p + a256(sizeof(double) * nq));
    float *qabs = reinterpret_cast<float *>(p + 2 * a256(sizeof(double) * nq));
    uint8_t *qq = reinterpret_cast<uint8_t *>(p + b_stats);
    float *qscale = reinterpret_cast<float *>(p + b_stats + a256((size_t)s->stride * nq));
    double *scan = d_scan ? d_scan : reinterpret_cast<double *>(p + b_stats + b_q);
    double *part = reinterpret_cast<double *>(p + b_stats + b_q + b_scan);
    int *cand = reinterpret_cast<int *>(p + b_stats + b_q + b_scan + b_part);

    PRNU_TRY(launch_pattern_stats(d_qptrs, nq, s->n, qsum, qsumsq, qabs, c->ws_small, st));
    PRNU_TRY(launch_quantize_fp8(d_qptrs, nq, s->n, qabs, qq, qscale, st));
    time_begin(c, st);
    for (int q0 = 0; q0 < nq; q0 += kQueryBlock) {
        const int m = std::min(kQueryBlock, nq - q0);
        PRNU_TRY(launch_fp8_ncc(c->ws_gram, kQueryBlock, false, qq + (size_t)q0 * s->stride, s->stride, m,
                                qscale + q0, qsum + q0, qsumsq + q0, s->q, s->stride, N, s->scale, s->sum, s->sumsq,
                                s->stride, scan + (size_t)q0 * N, N, c->gram_cfg[PRNU_PREC_FP8].slabs, st, nullptr));
    }
    time_end(c, st);
    const int chunk_len = (N + kTopkChunks - 1) / kTopkChunks;
    hipLaunchKernelGGL(topk_kernel, dim3(nq, kTopkChunks), dim3(256), 0, st, scan, (int64_t)N, N, kMaxK, cand,
                       chunk_len, (int64_t)kTopkChunks * kMaxK);
    PRNU_CHECK_LAUNCH();
    hipLaunchKernelGGL(topk_merge_kernel, dim3(nq), dim3(64), 0, st, scan, (int64_t)N, cand, k, d_idx);
    PRNU_CHECK_LAUNCH();
    hipLaunchKernelGGL(rescore_slice_kernel, dim3((unsigned)slices, nq), dim3(256), 0, st, d_qptrs, s->f32, s->n,
                       rescore_slice_len(s->n), d_idx, k, part);
    PRNU_CHECK_LAUNCH();
    hipLaunchKernelGGL(rescore_final_kernel, dim3(nq), dim3(64), 0, st, part, (int)slices, qsum, qsumsq, s->sum,
                       s->sumsq, d_idx, k, d_score);
    PRNU_CHECK_LAUNCH();
    hipLaunchKernelGGL(sort_k_kernel, dim3((nq + 63) / 64), dim3(64), 0, st, d_score, d_idx, nq, k);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

}  // namespace prnu

using namespace prnu;

extern "C" {

int prnu_fpstore_create(int64_t n, int capacity, prnu_fpstore **out) {
    PRNU_REQUIRE(out, "fpstore_create: null output");
    FpStore *s = nullptr;
    PRNU_TRY(fpstore_create(n, capacity, &s));
    *out = reinterpret_cast<prnu_fpstore *>(s);
    return PRNU_OK;
}

int prnu_fpstore_destroy(prnu_fpstore *store) {
    fpstore_destroy(reinterpret_cast<FpStore *>(store));
    return PRNU_OK;
}

int prnu_fpstore_info(const prnu_fpstore *store, int64_t *n, int *capacity, int *size, size_t *bytes) {
    PRNU_REQUIRE(store, "fpstore_info: null store");
    const FpStore *s = reinterpret_cast<const FpStore *>(store);
    if (n) *n = s->n;
    if (capacity) *capacity = s->capacity;
    if (size) *size = s->size;
    if (bytes) *bytes = (size_t)s->capacity * (s->stride + 4 * s->n + 28);
    return PRNU_OK;
}

int prnu_fpstore_put(prnu_fpstore *store, const float *const *d_patterns, int count, int first_slot,
                     prnu_stream stream) {
    PRNU_REQUIRE(store && d_patterns, "fpstore_put: null pointer");
    FpStore *s = reinterpret_cast<FpStore *>(store);
    PRNU_TRY(fpstore_put(s, d_patterns, count, first_slot, (hipStream_t)stream));
    if (!stream) PRNU_CHECK_HIP(hipStreamSynchronize(nullptr));
    return PRNU_OK;
}

int prnu_fpstore_query(prnu_ctx *ctx, prnu_fpstore *store, const float *const *d_queries, int nq, int k,
                       int32_t *d_idx, double *d_score, double *d_scan, prnu_stream stream) {
    PRNU_REQUIRE(ctx && store && d_queries && d_idx && d_score, "fpstore_query: null pointer");
    Context *c = reinterpret_cast<Context *>(ctx);
    FpStore *s = reinterpret_cast<FpStore *>(store);
    hipStream_t st = pick_stream(c, stream);
    PRNU_TRY(fpstore_query(c, s, d_queries, nq, k, d_idx, d_score, d_scan, st));
    return finish(c, stream);
}

}  // extern "C"
