/*
 * prnu_api.h -- C ABI of libprnu_hip.so, the MI355X (gfx950) implementation of
 * Rocket's PRNU common-source-identification hot path.
 *
 * Reference paths are relative to src/main/java/nl/esciencecenter/ of
 * JungleComputing/rocket:  CSI/ = common_source_identification/,
 * R/ = rocket/.
 *
 * Conventions
 *  - Every call returns an int status (prnu_status).  On failure a message is
 *    kept per thread and returned by prnu_last_error().  No C++ exception
 *    crosses this boundary.  PRNU_EINVAL is what the reference raises as
 *    IllegalArgumentException (undersized buffer, bad size); the other codes
 *    map to its CudaException.
 *  - Device pointers are plain addresses in the current HIP device's memory.
 *  - `stream` is a hipStream_t.  NULL selects the context's private stream and
 *    makes the call synchronous, as every reference GPU method is
 *    (CSI/kernels/compare/NormalizedCrossCorrelation.java:111-113,
 *    CSI/kernels/filter/PRNUFilter.java:135-137).  A non-NULL stream makes the
 *    call asynchronous on that stream.
 *  - A context is not re-entrant: calls on one context must be serialised by
 *    the caller (the reference calls them from one device thread,
 *    R/scheduler/DeviceWorker.java:94-95).  Separate contexts may be used from
 *    separate threads.
 *  - The context owns its workspaces and twiddle tables; buffers passed in are
 *    borrowed for the duration of the call and never freed by the library.
 */
#ifndef PRNU_API_H
#define PRNU_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PRNU_API_VERSION 1

#if defined(__GNUC__) || defined(__clang__)
#define PRNU_EXPORT __attribute__((visibility("default")))
#else
#define PRNU_EXPORT
#endif

typedef struct prnu_ctx prnu_ctx;
typedef void *prnu_stream; /* hipStream_t */

typedef enum {
    PRNU_OK = 0,
    PRNU_EINVAL = 1,   /* bad argument / undersized buffer */
    PRNU_ENOMEM = 2,   /* device or host allocation failed */
    PRNU_EDEVICE = 3,  /* HIP runtime error */
    PRNU_ENOTSUP = 4   /* valid request this build does not implement */
} prnu_status;

typedef enum { PRNU_MODE_NCC = 0, PRNU_MODE_PCE = 1 } prnu_mode;

typedef enum {
    PRNU_INPUT_RGB8 = 0,   /* packed 8-bit, 3 bytes per pixel, libjpeg order */
    PRNU_INPUT_GRAY8 = 1,  /* 8-bit gray, treated as R = G = B */
    PRNU_INPUT_FLOAT32 = 2 /* already-gray FP32 image */
} prnu_input_format;

typedef enum {
    PRNU_PREC_FP32 = 0,       /* FP32 MFMA (v_mfma_f32_16x16x4_f32) */
    PRNU_PREC_SPLIT_BF16 = 1, /* hi+lo BF16, three MFMA products, FP32 accumulate */
    PRNU_PREC_FP8 = 2,        /* e4m3 with per-pattern scale, FP32 accumulate */
    PRNU_PREC_FP16 = 3        /* FP16 with a per-pattern power-of-two scale, one MFMA product, FP32
                                 accumulate; pattern stores only */
} prnu_precision;

typedef enum {
    PRNU_PCE_REFERENCE = 0, /* the reference's semantics, quirks included */
    PRNU_PCE_TEXTBOOK = 1   /* true cross spectrum, |c| peak, 11x11 cyclic window */
} prnu_pce_variant;

/* ------------------------------------------------------------------ errors */
/* Message of the last failed call on this thread ("" if none). */
PRNU_EXPORT const char *prnu_last_error(void);
PRNU_EXPORT int prnu_api_version(void);

/* --------------------------------------------- memory and streams
 * These replace the JCuda wrapper the reference's host reaches HIP through:
 * R/cubaapi/CudaContext.java (allocBytes/allocFloats, getMemInfo),
 * R/cubaapi/CudaMem.java (copyFromHostAsync/copyToHostAsync/copyToDeviceAsync),
 * R/cubaapi/CudaPinned.java, R/cubaapi/CudaStream.java:28-31. */
PRNU_EXPORT int prnu_device_count(int *count);
PRNU_EXPORT int prnu_set_device(int device);
PRNU_EXPORT int prnu_mem_info(size_t *free_bytes, size_t *total_bytes);
PRNU_EXPORT int prnu_malloc_device(void **ptr, size_t bytes);
PRNU_EXPORT int prnu_free_device(void *ptr);
PRNU_EXPORT int prnu_malloc_pinned(void **ptr, size_t bytes);
PRNU_EXPORT int prnu_free_pinned(void *ptr);
PRNU_EXPORT int prnu_memcpy_h2d(void *dst, const void *src, size_t bytes, prnu_stream stream);
PRNU_EXPORT int prnu_memcpy_d2h(void *dst, const void *src, size_t bytes, prnu_stream stream);
PRNU_EXPORT int prnu_memcpy_d2d(void *dst, const void *src, size_t bytes, prnu_stream stream);
PRNU_EXPORT int prnu_stream_create(prnu_stream *stream);
PRNU_EXPORT int prnu_stream_destroy(prnu_stream stream);
PRNU_EXPORT int prnu_stream_synchronize(prnu_stream stream);

/* ------------------------------------------------------------------ context
 * CommonSourceIdentificationContext(CudaContext, Dimension, ComparisonStrategy)
 * (CSI/CommonSourceIdentificationContext.java:32-61): a PRNU filter for one
 * image size plus the comparator of `mode` (prnu_mode).  h, w >= 2. */
PRNU_EXPORT int prnu_ctx_create(int height, int width, int mode, prnu_ctx **out);
PRNU_EXPORT int prnu_ctx_destroy(prnu_ctx *ctx);
/* getMaxFileSize = 3*h*w (CommonSourceIdentificationContext.java:64-66) */
PRNU_EXPORT int64_t prnu_max_file_size(const prnu_ctx *ctx);
/* getMaxInputSize = pattern bytes: 4*h*w (NCC) or 8*h*(w/2+1) (PCE)
 * (CommonSourceIdentificationContext.java:69-71, PRNUFilter.java:103-112) */
PRNU_EXPORT int64_t prnu_max_input_size(const prnu_ctx *ctx);
/* getMaxOutputSize = 8 (CommonSourceIdentificationContext.java:74-76) */
PRNU_EXPORT int64_t prnu_max_output_size(const prnu_ctx *ctx);

/* ------------------------------------------------- drop-in per item / pair
 * preprocessInputGPU(K, CudaMem input, CudaMem output)
 * (R/scheduler/ApplicationContext.java:66, CommonSourceIdentificationContext.java:105-108,
 * PRNUFilter.applyGPU at CSI/kernels/filter/PRNUFilter.java:124-138):
 * image -> PRNU pattern (NCC mode: FP32 h*w) or its R2C spectrum (PCE mode:
 * complex FP32 h*(w/2+1)).  `input_format` is a prnu_input_format; in_bytes
 * must be at least h*w*{3,1,4}.  *out_bytes receives the pattern size. */
PRNU_EXPORT int prnu_preprocess(prnu_ctx *ctx, int input_format, const void *d_in, size_t in_bytes,
                    void *d_out, size_t out_capacity, size_t *out_bytes, prnu_stream stream);

/* Batched preprocessInputGPU: `count` images, in_stride BYTES apart, into
 * `count` patterns, out_stride BYTES apart, with one launch per filter stage
 * for the whole batch (the reference calls preprocessInputGPU once per file,
 * R/scheduler/ApplicationContext.java:66; this is the same computation for
 * many files).  Each output equals prnu_preprocess of that image alone.
 * in_bytes / out_capacity bound the whole strided range. */
PRNU_EXPORT int prnu_preprocess_batch(prnu_ctx *ctx, int input_format, const void *d_in, size_t in_bytes,
                                      int64_t in_stride, int count, void *d_out, size_t out_capacity,
                                      int64_t out_stride, prnu_stream stream);

/* correlateGPU(K, CudaMem left, K, CudaMem right, CudaMem output)
 * (R/scheduler/ApplicationContext.java:81, CommonSourceIdentificationContext.java:111-115,
 * PatternComparator.applyGPU at CSI/kernels/compare/PatternComparator.java:45):
 * one FP64 score into d_out.  NCC mode follows
 * CSI/kernels/compare/normalizedcrosscorrelation.cu:85-178 (FP32 products,
 * FP64 sums); PCE mode follows CSI/kernels/compare/peaktocorrelationenergy.cu
 * (reference variant).  Inputs are read only. */
PRNU_EXPORT int prnu_correlate(prnu_ctx *ctx, const void *d_left, size_t left_bytes,
                   const void *d_right, size_t right_bytes,
                   void *d_out, size_t out_capacity, size_t *out_bytes, prnu_stream stream);

/* PCE of one pair with a chosen variant (prnu_pce_variant); PCE contexts only.
 * PeakToCorrelationEnergy.compare (CSI/kernels/compare/PeakToCorrelationEnergy.java:202-232). */
PRNU_EXPORT int prnu_correlate_pce(prnu_ctx *ctx, int variant, const void *d_left, const void *d_right,
                       double *d_out, prnu_stream stream);

/* PCE of `npairs` pairs in batches (no reference counterpart: replaces
 * `npairs` correlateGPU calls in PCE mode).  d_spectra is a DEVICE array of
 * pointers to R2C half spectra (as prnu_preprocess writes them in PCE mode),
 * d_pairs a DEVICE array of 2*npairs int32 indices into it, d_out npairs
 * doubles.  PCE contexts only. */
PRNU_EXPORT int prnu_pce_pairs(prnu_ctx *ctx, int variant, const void *const *d_spectra, const int32_t *d_pairs,
                               int npairs, double *d_out, prnu_stream stream);

/* ------------------------------------------------------- single filters
 * The individual filter objects of PRNUFilter (PRNUFilter.java:157-175), for
 * testing each stage.  All work in place on an FP32 h*w image except
 * grayscale (in -> FP32) and spectral (FP32 -> complex h*(w/2+1)). */
/* GrayscaleFilter.applyGPU, CSI/kernels/filter/GrayscaleFilter.java:72-85 */
PRNU_EXPORT int prnu_filter_grayscale(prnu_ctx *ctx, int input_format, const void *d_in, size_t in_bytes,
                          float *d_out, prnu_stream stream);
/* FastNoiseFilter.applyGPU, CSI/kernels/filter/FastNoiseFilter.java:98-120 */
PRNU_EXPORT int prnu_filter_fastnoise(prnu_ctx *ctx, float *d_image, prnu_stream stream);
/* ZeroMeanTotalFilter.applyGPU, CSI/kernels/filter/ZeroMeanTotalFilter.java:77-99 */
PRNU_EXPORT int prnu_filter_zeromean(prnu_ctx *ctx, float *d_image, prnu_stream stream);
/* WienerFilter.applyGPU, CSI/kernels/filter/WienerFilter.java:187-240 */
PRNU_EXPORT int prnu_filter_wiener(prnu_ctx *ctx, float *d_image, prnu_stream stream);
/* SpectralFilter.applyGPU, CSI/kernels/filter/SpectralFilter.java:44-54 */
PRNU_EXPORT int prnu_filter_spectral(prnu_ctx *ctx, const float *d_image, float *d_out, size_t out_capacity,
                         prnu_stream stream);

/* ------------------------------------------------------ batched extensions
 * The reference has no batched entry points: it calls correlateGPU once per
 * pair.  These replace M(M-1)/2 such calls.
 *
 * Per-pattern statistics: d_sum[i] = sum x, d_sumsq[i] = sum fl32(x*x) in
 * FP64 (the per-pattern half of normalizedcrosscorrelation.cu:153-161), and
 * optionally d_absmax[i] = max |x| (FP32, for FP8 scaling).
 * d_patterns is a DEVICE array of `count` device pointers, each to n floats. */
PRNU_EXPORT int prnu_pattern_stats(const float *const *d_patterns, int count, int64_t n,
                       double *d_sum, double *d_sumsq, float *d_absmax, prnu_stream stream);

/* prnu_pattern_stats without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  The
 * same kernels and the same results, bit for bit; the partial sums go to a
 * fixed 3 MiB device array per GPU (131072 partials of 24 bytes) instead of a workspace,
 * so a call may cover at most 131072 / ceil(n / 65536) patterns (e.g. 2730 of
 * 2048x1536) and returns PRNU_EINVAL, with that limit in prnu_last_error(),
 * beyond it.  The array is shared by every call on a GPU: calls that overlap
 * on one GPU (on different streams or threads) race and must be serialised
 * by the caller.  Calls no hipMalloc/hipFree; with stream 0 it returns when
 * done. */
PRNU_EXPORT int prnu_pattern_stats_noalloc(const float *const *d_patterns, int count, int64_t n, double *d_sum,
                                           double *d_sumsq, float *d_absmax, prnu_stream stream);

/* All-pairs NCC of `count` patterns of the context's size as a triangular
 * Gram on MFMA (`precision` is a prnu_precision), with the FP64 epilogue
 * NCC = G_ij / (sqrt(Sxx_i - Sx_i^2) sqrt(Sxx_j - Sx_j^2)) of
 * normalizedcrosscorrelation.cu:121-123.  d_sum/d_sumsq (and d_absmax for
 * FP8) come from prnu_pattern_stats.  d_ncc is count*count doubles,
 * row-major; both triangles and the diagonal are written.
 * Replaces count*(count-1)/2 calls of correlateGPU in NCC mode. */
PRNU_EXPORT int prnu_gram_ncc(prnu_ctx *ctx, const float *const *d_patterns, int count, int precision,
                  const double *d_sum, const double *d_sumsq, const float *d_absmax,
                  double *d_ncc, prnu_stream stream);

/* ----------------------------------------------------------- pattern store
 * Device-resident pattern set for the batched path, the counterpart of the
 * device cache slots Rocket keeps preprocessed patterns in
 * (R/cache/DeviceCache.java:19-44), laid out for the Gram: each pattern is cut
 * into 128-byte blocks and block b of all slots is contiguous, stored as
 * split-BF16 (hi, lo; 32 values per block), FP32 (32 values), FP8 e4m3
 * (128 values of x / scale, scale = max|x| / 448 per pattern) or FP16 (64
 * values of x / scale, scale the power of two with max|x| / scale in
 * [2^14, 2^15)) for `format` = PRNU_PREC_SPLIT_BF16, PRNU_PREC_FP32,
 * PRNU_PREC_FP8 or PRNU_PREC_FP16.  Inserting computes
 * the per-pattern FP64 statistics (FP64 sums of FP32 products, as the
 * reference's NCC), so the NCC epilogue is exact whatever the format.
 * capacity <= 16,776,960 slots (one block column of the store, capacity
 * rounded up to 256 times 128 bytes, stays below 2^31 bytes). */
typedef struct prnu_store prnu_store;

PRNU_EXPORT int prnu_store_create(int64_t n, int capacity, int format, prnu_store **out);
PRNU_EXPORT int prnu_store_destroy(prnu_store *store);
PRNU_EXPORT int prnu_store_info(const prnu_store *store, int64_t *n, int *capacity, int *format, size_t *bytes);
/* Copy `count` FP32 patterns (DEVICE array of pointers, n floats each) into
 * slots [first_slot, first_slot + count) and record sum x, sum fl32(x*x)
 * (FP64) and max|x| of each. */
PRNU_EXPORT int prnu_store_put(prnu_store *store, const float *const *d_patterns, int count, int first_slot,
                               prnu_stream stream);
/* Device arrays (capacity entries) of the statistics recorded by put. */
PRNU_EXPORT int prnu_store_stats(const prnu_store *store, const double **d_sum, const double **d_sumsq,
                                 const float **d_absmax);
/* Spill tier (R/cache/HostCache.java:12-67, R/cache/FileCache.java:18-130:
 * the reference keeps what does not fit on the device in pinned host memory
 * and files).  A "slot image" is `count` slots of a store in the store's own
 * layout: a 64-byte header (magic "PRNUSLOT", values per pattern, format,
 * count, K blocks), the K blocks, and the per-slot statistics, so that a
 * pattern set larger than HBM can wait in host memory or a file and come back
 * without extraction or conversion.  prnu_store_image_bytes: the size of an
 * image of `count` slots (-1 for a null store).  prnu_store_save writes slots
 * [first_slot, first_slot + count) into h_image; prnu_store_load copies an
 * image of `image_bytes` bytes into slots [first_slot, first_slot + its
 * count) of a store of the same length and format (PRNU_EINVAL otherwise,
 * and when image_bytes is shorter than an image of the header's count, so a
 * truncated file is refused before anything is read past its end).  Both are
 * asynchronous on
 * `stream` (pinned host memory lets them overlap kernels; with stream 0 they
 * return when done). */
PRNU_EXPORT int64_t prnu_store_image_bytes(const prnu_store *store, int count);
PRNU_EXPORT int prnu_store_save(const prnu_store *store, int first_slot, int count, void *h_image, prnu_stream stream);
PRNU_EXPORT int prnu_store_load(prnu_store *store, int first_slot, const void *h_image, size_t image_bytes,
                                prnu_stream stream);
/* All-pairs NCC of slots [first_slot, first_slot + count) in the store's
 * format, same epilogue and output as prnu_gram_ncc.  `ctx` supplies the
 * workspace, the launch shape (its prnu_gram_config of the store's format)
 * and kernel timing. */
PRNU_EXPORT int prnu_store_gram_ncc(prnu_ctx *ctx, prnu_store *store, int first_slot, int count, double *d_ncc,
                                    prnu_stream stream);

/* ------------------------------------------------- grouped all-pairs plan
 * Image sets of several sizes (cameras of different resolutions): the
 * reference compares only patterns of one size per context
 * (CSI/CommonSourceIdentificationContext.java:97-99) and enumerates the pairs
 * of one set with an index space (R/indexspace/HilbertCurveIndexSpace.java);
 * a job over G sets is G such runs.  A plan computes the all-pairs NCC of
 * every group in ONE launch over the concatenated triangular tile
 * lists of all groups (equal-cost work items: every group is cut into K slabs
 * of the same length), then one FP64 reduction per group; each group's output
 * is what prnu_store_gram_ncc writes for it.  Groups may share a store (slot
 * ranges) or use different stores; all stores must have the same format. */
typedef struct prnu_gram_group {
    prnu_store *store;
    int32_t first_slot;
    int32_t count;
    double *d_ncc;          /* count x count FP64, row-major, device */
} prnu_gram_group;
typedef struct prnu_gram_plan prnu_gram_plan;

/* slab_values: K values per work item (a multiple of 32), 0 = automatic
 * (about 8 work items per CU).  The plan owns its partial-tile
 * buffer (at most 8 GiB) and keeps `ctx` for timing; destroy it before ctx. */
PRNU_EXPORT int prnu_gram_plan_create(prnu_ctx *ctx, const prnu_gram_group *groups, int ngroups,
                                      int64_t slab_values, prnu_gram_plan **out);
PRNU_EXPORT int prnu_gram_plan_destroy(prnu_gram_plan *plan);
/* use_graph != 0: the first run captures the launches into a HIP graph,
 * later runs replay it (one hipGraphLaunch); 0: plain launches. */
PRNU_EXPORT int prnu_gram_plan_run(prnu_gram_plan *plan, int use_graph, prnu_stream stream);
/* Times candidate slab lengths with both launch forms on the device (writes
 * the outputs) and keeps the fastest pair; drops a captured graph. */
PRNU_EXPORT int prnu_gram_plan_autotune(prnu_gram_plan *plan, prnu_stream stream, int64_t *best_slab_values,
                                        float *best_ms);
/* Launch form of the plan's product: PRNU_PLAN_PER_ITEM (default) launches
 * one workgroup per work item; PRNU_PLAN_PERSISTENT launches one workgroup
 * per CU that takes work items from per-XCD queues until the tile list is
 * done (direct-to-LDS loads whatever the load variant).  Results are
 * bitwise the same.  Drops a captured graph. */
typedef enum { PRNU_PLAN_PER_ITEM = 0, PRNU_PLAN_PERSISTENT = 1 } prnu_plan_launch;
PRNU_EXPORT int prnu_gram_plan_set_launch(prnu_gram_plan *plan, int launch);
PRNU_EXPORT int prnu_gram_plan_get_launch(const prnu_gram_plan *plan, int *launch);
/* A block of a plan: NCC of slots [row_first, row_first + row_count) against
 * [col_first, col_first + col_count) of one store into d_out (row-major,
 * leading dimension ld_out).  Identical ranges make a triangle block: the
 * symmetric count x count matrix, both triangles written, as
 * prnu_store_gram_ncc writes it.  Other blocks are rectangles (no mirror
 * written), e.g. the leaves of an index-space
 * division (R/indexspace/HilbertCurveIndexSpace.java:92-123) of the pair
 * triangle, so a host scheduler can hand a whole division to one launch. */
typedef struct prnu_gram_block {
    prnu_store *store;
    int32_t row_first, row_count;
    int32_t col_first, col_count;
    double *d_out;
    int64_t ld_out;
} prnu_gram_block;
/* Plan over up to 4096 blocks of stores of one format; otherwise as
 * prnu_gram_plan_create. */
PRNU_EXPORT int prnu_gram_plan_create_blocks(prnu_ctx *ctx, const prnu_gram_block *blocks, int nblocks,
                                             int64_t slab_values, prnu_gram_plan **out);
/* Work items, workgroups launched (padding items included for the per-item
 * form), slab length, partial bytes and the
 * algorithmic FLOP of one product over every group's triangle (diagonal
 * included). */
PRNU_EXPORT int prnu_gram_plan_info(const prnu_gram_plan *plan, int64_t *work_items, int *workgroups,
                                    int64_t *slab_values, size_t *partial_bytes, double *flop);

/* ------------------------------------------------------ fingerprint serving
 * No reference counterpart (the reference has no serving mode).  A device-
 * resident store of camera fingerprints: an FP8-e4m3 copy with a per-slot
 * FP32 scale for scanning, the FP32 originals for exact re-scoring, and FP64
 * sum / sum of squares per slot. */
typedef struct prnu_fpstore prnu_fpstore;

PRNU_EXPORT int prnu_fpstore_create(int64_t n, int capacity, prnu_fpstore **out);
PRNU_EXPORT int prnu_fpstore_destroy(prnu_fpstore *store);
/* size = the number of filled slots, which are always slots [0, size);
 * bytes = device memory held */
PRNU_EXPORT int prnu_fpstore_info(const prnu_fpstore *store, int64_t *n, int *capacity, int *size, size_t *bytes);
/* Quantise and copy `count` FP32 fingerprints (DEVICE pointer table) into
 * slots [first_slot, first_slot + count).  first_slot may not exceed the
 * current size (PRNU_EINVAL): a put overwrites filled slots or appends, so a
 * query never scans a slot that was never written. */
PRNU_EXPORT int prnu_fpstore_put(prnu_fpstore *store, const float *const *d_patterns, int count, int first_slot,
                                 prnu_stream stream);
/* Top-k (1 <= k <= 16) fingerprints of each of `nq` query patterns (DEVICE
 * pointer table, the context's h*w floats each): FP8 MFMA scan of every
 * filled slot, then the k best re-scored exactly (FP32 products, FP64 sums,
 * NCC of normalizedcrosscorrelation.cu:121-123) and sorted by that score.
 * d_idx / d_score: nq*k int32 / doubles.  d_scan (optional, nq*size doubles)
 * receives the FP8 NCC estimates.  A query's result does not depend on the
 * other queries of the call. */
PRNU_EXPORT int prnu_fpstore_query(prnu_ctx *ctx, prnu_fpstore *store, const float *const *d_queries, int nq, int k,
                                   int32_t *d_idx, double *d_score, double *d_scan, prnu_stream stream);

/* Measurement (no reference counterpart): the dense MFMA rate this device
 * sustains at the clock it holds under a full matrix load.  Every wave issues
 * independent MFMAs on register operands only (2 waves per SIMD on every CU),
 * `iterations` rounds of 32 per wave; precision selects the instruction:
 * PRNU_PREC_SPLIT_BF16 -> 16x16x32 BF16 (one product), PRNU_PREC_FP16 ->
 * 16x16x32 FP16, PRNU_PREC_FP8 -> 16x16x128 e4m3.  random_operands = 1 fills
 * the operands with hashed bits (noise-like data: the power drawn, hence the
 * clock the chip holds, depends on how many operand bits toggle), 0 with
 * small constants.  Best of 3 timed launches after a warm-up; *tflops =
 * FLOP / time, *ms (optional) = that time. */
PRNU_EXPORT int prnu_measure_mfma_peak(int precision, int random_operands, int iterations, double *tflops,
                                       float *ms);

/* Launch shape of the Gram tile kernel (no reference counterpart).
 * slabs: K slabs per tile (0 = automatic); xcd_remap: 1 keeps the tiles of one
 * K slab on one XCD so they share it through that XCD's L2; variant: for the
 * split-BF16 Gram over pointer tables, 1 = 4 waves with the full register
 * file each, two K stages in flight, otherwise 8 waves, one K stage in
 * flight; for the store Gram and grouped plans, 0 = direct global-to-LDS
 * loads, 1 = loads staged through registers, 2 = direct loads with the wave
 * priority raised over the MFMA block, 3 (store Gram of split-BF16 and FP8
 * stores) = 4 waves of 128 x 128 with accumulators in AGPRs.  Every variant
 * gives bitwise the same results. */
typedef struct {
    int32_t slabs;
    int32_t xcd_remap;
    int32_t variant;
} prnu_gram_config;

/* Set (cfg != NULL) or reset to automatic (cfg == NULL) the Gram launch shape
 * of `precision` for this context. */
PRNU_EXPORT int prnu_gram_set_config(prnu_ctx *ctx, int precision, const prnu_gram_config *cfg);
/* The shape prnu_gram_ncc would use for `count` patterns. */
PRNU_EXPORT int prnu_gram_get_config(prnu_ctx *ctx, int precision, int count, prnu_gram_config *out);
/* Autotuner: times the tile kernel on the caller's patterns for candidate K
 * splits and XCD mappings (`reps` timed launches each, minimum kept), keeps
 * the fastest in the context and reports it.  Synchronous on `stream`. */
PRNU_EXPORT int prnu_gram_autotune(prnu_ctx *ctx, const float *const *d_patterns, int count, int precision,
                                   int reps, prnu_gram_config *best, float *best_ms, prnu_stream stream);

/* --------------------------------------------------------- pair scheduling
 * Host-only (no device call).  Upper-triangle index spaces of
 * R/indexspace/HilbertCurveIndexSpace.java:31-123 and
 * R/indexspace/TilingIndexSpace.java:27-91 and the pair count of
 * R/util/Util.java:51-77.  A node covers rows (left) [left_offset,
 * left_offset+left_length) x columns (right) [right_offset, ...); pair (i, j)
 * with i < j (or i <= j with include_diagonal) belongs to it. */
typedef enum { PRNU_INDEX_HILBERT = 0, PRNU_INDEX_TILING = 1 } prnu_index_kind;

typedef struct {
    int32_t left_offset, left_length;
    int32_t right_offset, right_length;
    int32_t rule; /* Hilbert rule 0..3 (A..D); 0 for tiling */
} prnu_index_node;

/* Util.calculateTriangleSize(rowBegin, rowEnd, colBegin, colEnd, includeDiagonal) */
PRNU_EXPORT int64_t prnu_triangle_size(int row_begin, int row_end, int col_begin, int col_end,
                                       int include_diagonal);
/* IndexSpace.divide(IndexSpaceDivision) of one node: `param` is minSplitSize
 * (Hilbert) or tileSize (tiling).  Children to spawn go to `children`, entries
 * the step emits directly go to `pairs` as (i, j) int32 pairs. */
PRNU_EXPORT int prnu_index_divide(int kind, const prnu_index_node *node, int param, int include_diagonal,
                                  prnu_index_node *children, int children_cap, int *n_children,
                                  int32_t *pairs, int64_t pairs_cap, int64_t *n_pairs);
/* Every pair of a `count`-key space in depth-first (curve) order.  With
 * pairs == NULL only *n_pairs is computed. */
PRNU_EXPORT int prnu_index_pairs(int kind, int count, int param, int include_diagonal, int32_t *pairs,
                                 int64_t pairs_cap, int64_t *n_pairs);

/* ------------------------------------------------------ device-cache policy
 * Host-only (no device call).  Which keys (pattern blocks) occupy which of
 * `nslots` slots of a pattern store: the policy of R/cache/AbstractCache.java:
 * 19-409 over the buffer stack of R/cache/DeviceCache.java:19-62.  Entries
 * are reference counted; status WRITING (one holder fills it), READONLY
 * (filled, shared) or INVALID (a writer released it unfilled).  Transactions
 * are served in submission order; one that does not fit waits at the head,
 * holding what it has, while unreferenced entries are evicted least recently
 * used first (a released readable entry is most recently used, a new or
 * invalid one least).  Thread-safe: one lock per cache. */
typedef struct prnu_slotcache prnu_slotcache;
typedef enum { PRNU_SLOT_WRITING = 0, PRNU_SLOT_READONLY = 1, PRNU_SLOT_INVALID = 2 } prnu_slot_status;

/* Slots for entries of `entry_bytes`: R/RocketLauncher.java:371,391-398 (the
 * request if positive, else free memory plus that negative margin, at most
 * 5000 entries) with DeviceCache.java:27-31 (the largest n with
 * n * entry_bytes below the total). */
PRNU_EXPORT int prnu_slotcache_slots_for(size_t entry_bytes, int64_t requested, size_t free_bytes,
                                         int64_t *nslots);
/* new DeviceCache(ctx, total, entry) without the allocation: slots
 * 0..nslots-1 (the caller's store), nslots >= 2 (DeviceWorker.java:87-89). */
PRNU_EXPORT int prnu_slotcache_create(int64_t nslots, prnu_slotcache **cache);
PRNU_EXPORT int prnu_slotcache_destroy(prnu_slotcache *cache);
/* AbstractCache.acquireAllAsync(keys...): duplicates removed; *txn numbers
 * the transaction (1, 2, ... per cache). */
PRNU_EXPORT int prnu_slotcache_acquire(prnu_slotcache *cache, const int64_t *keys, int n, int64_t *txn);
/* The transaction's future: *done = 0 while it waits (*n = its key count);
 * once complete, keys / slots / writer (1: the caller fills the slot and
 * calls prnu_slotcache_ready, a miss -- DeviceWorker.java:208-226) for its
 * *n distinct keys, and the transaction is retired (polling it again is
 * PRNU_EINVAL).  Hits and misses are counted here.  With keys, slots and
 * writer all NULL only *done and *n are reported (nothing is retired). */
PRNU_EXPORT int prnu_slotcache_poll(prnu_slotcache *cache, int64_t txn, int *done, int64_t *keys, int64_t *slots,
                                    int32_t *writer, int cap, int *n);
/* Entry.downgradeToReader(meta): PRNU_EINVAL unless the entry is WRITING. */
PRNU_EXPORT int prnu_slotcache_ready(prnu_slotcache *cache, int64_t key, int64_t meta);
/* Entry.release(): one reference dropped; PRNU_EINVAL at refcount 0. */
PRNU_EXPORT int prnu_slotcache_release(prnu_slotcache *cache, int64_t key);
/* tryAcquireImmediately(key): *slot of a READONLY entry, now referenced, or -1. */
PRNU_EXPORT int prnu_slotcache_try_acquire(prnu_slotcache *cache, int64_t key, int64_t *slot);
/* Entry.size(): the meta of a READONLY entry, else PRNU_EINVAL. */
PRNU_EXPORT int prnu_slotcache_meta(prnu_slotcache *cache, int64_t key, int64_t *meta);
/* Inspection: slot, prnu_slot_status (-1: not resident) and refcount. */
PRNU_EXPORT int prnu_slotcache_entry(prnu_slotcache *cache, int64_t key, int64_t *slot, int *status,
                                     int64_t *refcount);
/* Unreferenced keys in eviction order (first evicted first); keys == NULL
 * only counts. */
PRNU_EXPORT int prnu_slotcache_lru(prnu_slotcache *cache, int64_t *keys, int64_t cap, int64_t *n);
/* stats[8]: hits, misses, evictions, resident entries, free slots, waiting
 * transactions, slots, unreferenced entries. */
PRNU_EXPORT int prnu_slotcache_stats(prnu_slotcache *cache, int64_t *stats);
/* AbstractCache.cleanup(): evict everything; PRNU_EINVAL if any entry is
 * still referenced. */
PRNU_EXPORT int prnu_slotcache_cleanup(prnu_slotcache *cache);

/* ------------------------------------------------------------ JPEG input
 * ReadJPEG.readJPEG / readJPEG (CSI/ReadJPEG.java:25-45, readjpeg/src/main/
 * cpp/native.cpp:59-134): libjpeg's default decode of a 3-component JPEG to
 * packed RGB8, here for a batch of files on the GPU, bit for bit what
 * libjpeg returns (ISLOW inverse DCT, libjpeg's choice of fancy or
 * replicating chroma upsampling, its YCbCr tables).  The reference rejects
 * anything that does not decode to RGB (native.cpp:99-103); so does this,
 * and it also rejects progressive, lossless, arithmetic-coded and 12-bit
 * files and 1x2 chroma subsampling (PRNU_ENOTSUP for that file). */
typedef struct prnu_jpeg_decoder prnu_jpeg_decoder;
/* per-file status beyond prnu_status: decoded, but the entropy-coded data
 * held an invalid code (libjpeg warns and decodes on; readJPEG succeeds) */
#define PRNU_JPEG_CORRUPT_DATA 16
PRNU_EXPORT int prnu_jpeg_decoder_create(prnu_jpeg_decoder **out);
PRNU_EXPORT int prnu_jpeg_decoder_destroy(prnu_jpeg_decoder *dec);
/* Width and height from the headers of a JPEG file in host memory. */
PRNU_EXPORT int prnu_jpeg_info(const void *h_file, int64_t size, int32_t *width, int32_t *height);
/* Decode `count` files (host memory: h_files[k], h_sizes[k] bytes) into
 * d_rgb (device), file k at byte k * rgb_stride, rows of width * 3 bytes.
 * h_width / h_height / h_status (optional, `count` each) receive each
 * file's size and status (0, PRNU_JPEG_CORRUPT_DATA, or the prnu_status it
 * was rejected with: EINVAL for a malformed file or one that does not fit
 * rgb_stride, ENOTSUP for an unsupported kind).  Every accepted file is
 * decoded; the call returns PRNU_OK when all were accepted, else the status
 * of the first rejected one with its reason in prnu_last_error().  The call
 * returns after the decode (synchronous on `stream`). */
PRNU_EXPORT int prnu_jpeg_decode_batch(prnu_jpeg_decoder *dec, const void *const *h_files, const int64_t *h_sizes,
                                       int count, uint8_t *d_rgb, int64_t rgb_stride, int64_t rgb_capacity,
                                       int32_t *h_width, int32_t *h_height, int32_t *h_status, prnu_stream stream);
/* readJPEG itself (readjpeg/src/main/cpp/native.cpp:59-134, bound by JNA in
 * CSI/ReadJPEG.java:12-23): the same name, arguments and result, so the
 * binding can load this library in place of libreadjpeg.  One file from host
 * memory into the host buffer `output` (packed RGB8 rows); returns 0, or 1
 * with the reason on stderr (not 3-component YCbCr, unsupported kind,
 * malformed, or output_capacity < 3 * width * height).  Decoded on the
 * current device; calls are serialised. */
PRNU_EXPORT int readJPEG(const char *input, int32_t input_size, int32_t *output_width, int32_t *output_height,
                         char *output, int32_t output_capacity);

/* --------------------------------------------------------------- timing
 * No reference counterpart (the reference times whole tasks on the host,
 * R/RocketLauncher.java:115-203).  When enabled, the context records HIP
 * events around the dominant kernel of each batched call, on the stream that
 * kernel runs on; prnu_ctx_last_kernel_ms waits for the end event and returns
 * the elapsed milliseconds of the last such call. */
PRNU_EXPORT int prnu_ctx_set_timing(prnu_ctx *ctx, int enable);
PRNU_EXPORT int prnu_ctx_last_kernel_ms(prnu_ctx *ctx, float *ms);

#ifdef __cplusplus
}
#endif
#endif /* PRNU_API_H */
