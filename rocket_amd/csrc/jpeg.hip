// Baseline JPEG decoding into packed RGB8 frames on the GPU: the reference's
// readJPEG (readjpeg/src/main/cpp/native.cpp:59-134) for a batch of files.
// The reference calls libjpeg with its default parameters; the results here
// are those of libjpeg's default decode, bit for bit:
//   * sequential Huffman decoding (ITU T.81 F.2.2), DC prediction per
//     component, reset at restart markers;
//   * the ISLOW integer inverse DCT (13-bit fixed-point Loeffler-Ligtenberg-
//     Moschytz factorisation, 2 extra bits between the passes, output
//     level-shifted and clamped through the range table indexed by the low
//     10 bits);
//   * chroma upsampling as libjpeg chooses it: the triangle ("fancy") filter
//     for 2x1 and 2x2 subsampling of components more than 2 samples wide,
//     replication for every other integral factor;
//   * YCbCr -> RGB with libjpeg's 16-bit fixed-point tables.
// Files that do not decode to 3-component YCbCr (gray, RGB-coded, CMYK) are
// rejected, as readJPEG rejects non-RGB output (native.cpp:99-103); so are
// progressive, lossless, arithmetic-coded and 12-bit files and 1x2 chroma
// subsampling (which libjpeg versions upsample differently).
//
// Work split.  The host parses the markers (tables, frame, scans), finds the
// restart-interval segments of each scan and copies their entropy-coded
// bytes, unstuffed, to the device.  Huffman decoding is parallel inside a
// segment (a whole scan when the file has no restart markers): segments are
// cut into chunks of kChunkBytes, and every chunk is decoded from a guessed
// state (bit position, block of the MCU, coefficient index); the exit state
// of chunk t becomes the entry state of chunk t+1, and chunks whose entry
// changed are decoded again until no entry changes.  Huffman codes
// resynchronise within a few symbols, so a guess that is wrong at the start
// of a chunk gives the right state at its end and two rounds usually settle
// every chunk; the result does not depend on that speed, only on the exact
// fixed point.  Then a prefix sum over each segment's chunks gives each
// chunk its first block, a write pass stores the coefficients (DC as
// differences), and a per-component scan turns the differences into DC
// values (the prediction chain).  jpeg_idct_kernel runs one block per
// thread, jpeg_color_kernel one output pixel per thread.
#include "prnu_internal.hpp"

#include <algorithm>
#include <climits>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace prnu {
namespace {

constexpr int kLook = 9;  // Huffman lookahead bits

struct HuffTab {
    uint16_t look[1 << kLook];  // (length << 8) | symbol for codes of <= kLook bits, 0 otherwise
    int32_t maxcode[18];        // largest code of each length (-1: none); [17] = INT_MAX
    int32_t valoff[18];         // symbol index of a code of length l = valoff[l] + code
    uint8_t vals[256];
};

struct JComp {
    int h, v;        // sampling factors (blocks per MCU across / down)
    int hf, vf;      // upsampling factors (max sampling factor / this component's)
    int bw, bh;      // block columns / rows of the whole-MCU grid
    int dw, dh;      // samples of the component
    int tq;          // global quantisation table index
    int pitch;       // plane row bytes (bw * 8)
    int64_t coef;    // first coefficient block (global)
    int64_t plane;   // first plane byte (global)
};

struct JImage {
    int W, H;        // 0 for a rejected file: every kernel skips it
    JComp c[3];
    int64_t rgb;     // output offset (bytes)
};

constexpr int kMaxBlocksPerMcu = 10;  // T.81 B.2.3
constexpr int kChunkBytes = 256;      // unstuffed entropy-coded bytes per chunk

struct JSeg {
    int image, ns, nbu;         // components in the scan, blocks per MCU
    int comp[4], dc[4], ac[4];  // per scan slot: component index, global DC / AC table
    int uslot[kMaxBlocksPerMcu];  // per block of the MCU: scan slot, block row and column offsets
    int uv[kMaxBlocksPerMcu], uh[kMaxBlocksPerMcu];
    int mcu0, mcus;             // first MCU (block, for ns == 1) and count
    int mcux;                   // MCU (block, for ns == 1) columns of the scan
    int chunk0, nchunk;         // this segment's chunks
    int64_t begin, end;         // unstuffed entropy-coded bytes in the batch buffer
};

// zigzag position -> natural (row-major) index; positions past 63 (corrupt
// data) land on 63, as libjpeg's extended table does
__constant__ uint8_t kZigzag[80] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,
    6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31,
    39, 46, 53, 60, 61, 54, 47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

// ----------------------------------------------------------- bit reading
// MSB-first bits of an unstuffed segment; past its end the stream continues
// with zero bits (libjpeg does the same at a premature marker).  pos() is
// the absolute bit index of the next unread bit.
struct BitReader {
    const uint8_t *base, *p, *end;
    uint64_t acc;
    int n;
    __device__ void init(const uint8_t *data, int64_t bit, const uint8_t *seg_end) {
        base = data;
        p = data + (bit >> 3);
        end = seg_end;
        acc = 0;
        n = 0;
        fill();
        const int sk = (int)(bit & 7);
        acc <<= sk;
        n -= sk;
    }
    __device__ int64_t pos() const { return (int64_t)(p - base) * 8 - n; }
    __device__ void fill() {
        while (n <= 56) {
            const uint64_t b = p < end ? *p : 0;
            ++p;
            acc |= b << (56 - n);
            n += 8;
        }
    }
    __device__ int get(int k) {  // k in [0, 16]
        if (k == 0) return 0;
        fill();
        const int v = (int)(acc >> (64 - k));
        acc <<= k;
        n -= k;
        return v;
    }
    __device__ int huff(const HuffTab &t, int &bad) {
        fill();
        const int e = t.look[acc >> (64 - kLook)];
        if (e) {
            const int len = e >> 8;
            acc <<= len;
            n -= len;
            return e & 0xFF;
        }
        for (int l = kLook + 1; l <= 16; ++l) {
            const int code = (int)(acc >> (64 - l));
            if (code <= t.maxcode[l]) {
                acc <<= l;
                n -= l;
                return t.vals[(t.valoff[l] + code) & 0xFF];
            }
        }
        bad = 1;  // no such code: libjpeg warns and takes symbol 0
        acc <<= 16;
        n -= 16;
        return 0;
    }
};

__device__ __forceinline__ int extend(int v, int s) { return (s && v < (1 << (s - 1))) ? v - (1 << s) + 1 : v; }

// Decoder state between two symbols: bit position, block of the MCU, next
// coefficient index (0: a DC symbol is next).
__device__ __forceinline__ int64_t pack_state(int64_t pos, int u, int k) { return (pos << 16) | (u << 8) | k; }

// Coefficient block of the segment's block number b (decode order).
__device__ __forceinline__ int16_t *block_at(const JSeg &g, const JImage &im, int16_t *coef, int b) {
    if (g.ns == 1) {
        const JComp &c = im.c[g.comp[0]];
        const int q = g.mcu0 + b;
        return coef + (c.coef + (int64_t)(q / g.mcux) * c.bw + q % g.mcux) * 64;
    }
    const int m = g.mcu0 + b / g.nbu, u = b % g.nbu;
    const JComp &c = im.c[g.comp[g.uslot[u]]];
    return coef + (c.coef + (int64_t)((m / g.mcux) * c.v + g.uv[u]) * c.bw + (m % g.mcux) * c.h + g.uh[u]) * 64;
}

// Decode symbols from `state` while the position is before `stop`.  WRITE:
// store coefficients (DC as the difference) of blocks [0, total) into
// `coef`; the block in progress at the entry is number first - 1, the next
// one `first`.  Returns the exit state; *started = DC symbols decoded.
template <bool WRITE>
__device__ int64_t run_chunk(const JSeg &g, const JImage &im, const HuffTab *__restrict__ tabs,
                             const uint8_t *__restrict__ data, int64_t state, int64_t stop, int *started, int first,
                             int16_t *__restrict__ coef, int &bad) {
    BitReader br;
    br.init(data, state >> 16, data + g.end);
    int u = (int)((state >> 8) & 0xFF), k = (int)(state & 0xFF);
    const int total = g.mcus * g.nbu;
    int cur = first - 1;  // the block in progress (k > 0) or the one before the next DC symbol
    int n = 0;
    int16_t *blk = (WRITE && cur >= 0 && cur < total) ? block_at(g, im, coef, cur) : nullptr;
    while (br.pos() < stop) {
        const int slot = g.uslot[u];
        if (k == 0) {
            const int s = br.huff(tabs[g.dc[slot]], bad) & 15;
            const int diff = extend(br.get(s), s);
            ++cur;
            ++n;
            if (WRITE) {
                blk = cur < total ? block_at(g, im, coef, cur) : nullptr;
                if (blk) blk[0] = (int16_t)diff;
            }
            k = 1;
        } else {
            const int rs = br.huff(tabs[g.ac[slot]], bad);
            const int r = rs >> 4, sz = rs & 15;
            if (sz) {
                k += r;
                const int v = extend(br.get(sz), sz);
                if (WRITE && blk) blk[kZigzag[min(k, 79)]] = (int16_t)v;
                ++k;
            } else {
                k = (r == 15) ? k + 16 : 64;  // ZRL or end of block
            }
        }
        if (k >= 64) {
            k = 0;
            u = (u + 1 == g.nbu) ? 0 : u + 1;
            // the segment's last block is complete: what follows is the
            // 1-bit padding before the marker, never decoded (by libjpeg
            // either), so it cannot raise the corrupt-data flag
            if (WRITE && cur + 1 >= total) break;
        }
    }
    *started = n;
    return pack_state(br.pos(), u, k);
}

__device__ __forceinline__ int64_t chunk_start_bit(const JSeg &g, int ch) {
    return (g.begin + (int64_t)(ch - g.chunk0) * kChunkBytes) * 8;
}
__device__ __forceinline__ int64_t chunk_stop_bit(const JSeg &g, int ch) {
    return min(g.end, g.begin + (int64_t)(ch - g.chunk0 + 1) * kChunkBytes) * 8;
}

// Every chunk starts from its own first bit, at the start of an MCU (exact
// for each segment's first chunk, a guess for the others).
__global__ __launch_bounds__(256) void jpeg_chunk_init_kernel(const JSeg *__restrict__ segs,
                                                              const int *__restrict__ chunk_seg, int nchunk,
                                                              int64_t *__restrict__ entry, int *__restrict__ dirty) {
    const int ch = blockIdx.x * 256 + threadIdx.x;
    if (ch >= nchunk) return;
    entry[ch] = pack_state(chunk_start_bit(segs[chunk_seg[ch]], ch), 0, 0);
    dirty[ch] = 1;
}

// Decode the chunks whose entry state changed; record exit states.
__global__ __launch_bounds__(64) void jpeg_chunk_sync_kernel(const JSeg *__restrict__ segs,
                                                             const int *__restrict__ chunk_seg, int nchunk,
                                                             const JImage *__restrict__ imgs,
                                                             const HuffTab *__restrict__ tabs,
                                                             const uint8_t *__restrict__ data,
                                                             const int64_t *__restrict__ entry,
                                                             int64_t *__restrict__ exit_state,
                                                             int *__restrict__ started, int *__restrict__ dirty) {
    const int ch = blockIdx.x * 64 + threadIdx.x;
    if (ch >= nchunk || !dirty[ch]) return;
    const JSeg &g = segs[chunk_seg[ch]];
    int bad = 0, n = 0;
    exit_state[ch] = run_chunk<false>(g, imgs[g.image], tabs, data, entry[ch], chunk_stop_bit(g, ch), &n, 0,
                                      nullptr, bad);
    started[ch] = n;
    dirty[ch] = 0;
}

// entry[t] = exit[t-1] inside a segment; flags what changed.
__global__ __launch_bounds__(256) void jpeg_chunk_link_kernel(const JSeg *__restrict__ segs,
                                                              const int *__restrict__ chunk_seg, int nchunk,
                                                              const int64_t *__restrict__ exit_state,
                                                              int64_t *__restrict__ entry, int *__restrict__ dirty,
                                                              int *__restrict__ changed) {
    const int ch = blockIdx.x * 256 + threadIdx.x;
    if (ch >= nchunk || ch == segs[chunk_seg[ch]].chunk0) return;
    const int64_t e = exit_state[ch - 1];
    if (e != entry[ch]) {
        entry[ch] = e;
        dirty[ch] = 1;
        *changed = 1;
    }
}

// First block of each chunk: prefix sum of the blocks started before it in
// its segment.  One thread per segment.
__global__ __launch_bounds__(64) void jpeg_chunk_scan_kernel(const JSeg *__restrict__ segs, int nseg,
                                                             const int *__restrict__ started,
                                                             int *__restrict__ first) {
    const int si = blockIdx.x * 64 + threadIdx.x;
    if (si >= nseg) return;
    const JSeg &g = segs[si];
    int run = 0;
    for (int ch = g.chunk0; ch < g.chunk0 + g.nchunk; ++ch) {
        first[ch] = run;
        run += started[ch];
    }
}

__global__ __launch_bounds__(64) void jpeg_chunk_write_kernel(const JSeg *__restrict__ segs,
                                                              const int *__restrict__ chunk_seg, int nchunk,
                                                              const JImage *__restrict__ imgs,
                                                              const HuffTab *__restrict__ tabs,
                                                              const uint8_t *__restrict__ data,
                                                              const int64_t *__restrict__ entry,
                                                              const int *__restrict__ first,
                                                              int16_t *__restrict__ coef, int *__restrict__ bad_out) {
    const int ch = blockIdx.x * 64 + threadIdx.x;
    if (ch >= nchunk) return;
    const JSeg &g = segs[chunk_seg[ch]];
    int bad = 0, n = 0;
    (void)run_chunk<true>(g, imgs[g.image], tabs, data, entry[ch], chunk_stop_bit(g, ch), &n, first[ch], coef, bad);
    if (bad) bad_out[g.image] = 1;
}

// DC prediction: per (segment, scan slot), the component's blocks in decode
// order hold differences; replace them by running sums (reset at each
// segment, i.e. restart interval).  grid (segments, 4), 256 threads: each
// thread sums a contiguous run, then a block-wide exclusive scan.
__global__ __launch_bounds__(256) void jpeg_dc_scan_kernel(const JSeg *__restrict__ segs,
                                                           const JImage *__restrict__ imgs,
                                                           int16_t *__restrict__ coef) {
    __shared__ int part[256];
    const JSeg &g = segs[blockIdx.x];
    const int slot = blockIdx.y;
    if (slot >= g.ns) return;
    const JImage &im = imgs[g.image];
    const JComp &c = im.c[g.comp[slot]];
    const int per_mcu = g.ns == 1 ? 1 : c.v * c.h;
    const int count = g.mcus * per_mcu;
    const int span = (count + 255) / 256;
    const int i0 = min(count, (int)threadIdx.x * span), i1 = min(count, i0 + span);
    auto at = [&](int i) -> int16_t * {
        if (g.ns == 1) {
            const int q = g.mcu0 + i;
            return coef + (c.coef + (int64_t)(q / g.mcux) * c.bw + q % g.mcux) * 64;
        }
        const int m = g.mcu0 + i / per_mcu, r = i % per_mcu;
        return coef + (c.coef + (int64_t)((m / g.mcux) * c.v + r / c.h) * c.bw + (m % g.mcux) * c.h + r % c.h) * 64;
    };
    int sum = 0;
    for (int i = i0; i < i1; ++i) sum += at(i)[0];
    part[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int t = 0; t < 256; ++t) {
            const int v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    int run = part[threadIdx.x];
    for (int i = i0; i < i1; ++i) {
        int16_t *b = at(i);
        run += b[0];
        b[0] = (int16_t)run;
    }
}

// ------------------------------------------------------------------- IDCT
constexpr int kConstBits = 13, kPass1Bits = 2;
// |sqrt(2) x combinations of cos(k pi / 16)| in 13-bit fixed point, named by
// their values to 9 decimals
constexpr int F0_298631336 = 2446, F0_390180644 = 3196, F0_541196100 = 4433, F0_765366865 = 6270,
              F0_899976223 = 7373, F1_175875602 = 9633, F1_501321110 = 12299, F1_847759065 = 15137,
              F1_961570560 = 16069, F2_053119869 = 16819, F2_562915447 = 20995, F3_072711026 = 25172;

// One 8-point pass; outputs before descaling.
__device__ __forceinline__ void idct8(const int d[8], int o[8]) {
    int z2 = d[2], z3 = d[6];
    int z1 = (z2 + z3) * F0_541196100;
    const int tmp2 = z1 - z3 * F1_847759065;
    const int tmp3 = z1 + z2 * F0_765366865;
    const int tmp0 = (d[0] + d[4]) * (1 << kConstBits);
    const int tmp1 = (d[0] - d[4]) * (1 << kConstBits);
    const int tmp10 = tmp0 + tmp3, tmp13 = tmp0 - tmp3, tmp11 = tmp1 + tmp2, tmp12 = tmp1 - tmp2;
    int t0 = d[7], t1 = d[5], t2 = d[3], t3 = d[1];
    z1 = t0 + t3;
    z2 = t1 + t2;
    z3 = t0 + t2;
    int z4 = t1 + t3;
    const int z5 = (z3 + z4) * F1_175875602;
    t0 *= F0_298631336;
    t1 *= F2_053119869;
    t2 *= F3_072711026;
    t3 *= F1_501321110;
    z1 *= -F0_899976223;
    z2 *= -F2_562915447;
    z3 = z3 * -F1_961570560 + z5;
    z4 = z4 * -F0_390180644 + z5;
    t0 += z1 + z3;
    t1 += z2 + z4;
    t2 += z2 + z3;
    t3 += z1 + z4;
    o[0] = tmp10 + t3;
    o[7] = tmp10 - t3;
    o[1] = tmp11 + t2;
    o[6] = tmp11 - t2;
    o[2] = tmp12 + t1;
    o[5] = tmp12 - t1;
    o[3] = tmp13 + t0;
    o[4] = tmp13 - t0;
}

__device__ __forceinline__ int descale(int x, int n) { return (x + (1 << (n - 1))) >> n; }

// libjpeg's post-IDCT range table by its index v & 1023: v + 128 clamped to
// [0, 255] for v in [-512, 511], wrapping outside
__device__ __forceinline__ uint32_t range_limit(int v) {
    const int k = v & 1023;
    return (uint32_t)(k < 128 ? k + 128 : k < 512 ? 255 : k < 896 ? 0 : k - 896);
}

// One 8x8 block per thread: dequantize, ISLOW, write 8 rows of 8 samples.
// grid (ceil(max blocks / 256), 3 components, images)
__global__ __launch_bounds__(256) void jpeg_idct_kernel(const JImage *__restrict__ imgs,
                                                        const uint16_t *__restrict__ qtab,
                                                        const int16_t *__restrict__ coef,
                                                        uint8_t *__restrict__ planes) {
    const JImage &im = imgs[blockIdx.z];
    if (im.W == 0) return;
    const JComp &c = im.c[blockIdx.y];
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= c.bw * c.bh) return;
    const uint4 *src = reinterpret_cast<const uint4 *>(coef + (c.coef + b) * 64);
    const uint16_t *q = qtab + c.tq * 64;
    int x[64];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const uint4 u = src[i];
        const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            x[8 * i + 2 * j] = (int)(int16_t)(w[j] & 0xFFFF) * (int)q[8 * i + 2 * j];
            x[8 * i + 2 * j + 1] = (int)(int16_t)(w[j] >> 16) * (int)q[8 * i + 2 * j + 1];
        }
    }
    int ws[64];
#pragma unroll
    for (int col = 0; col < 8; ++col) {  // pass 1: columns
        int d[8], o[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) d[r] = x[8 * r + col];
        idct8(d, o);
#pragma unroll
        for (int r = 0; r < 8; ++r) ws[8 * r + col] = descale(o[r], kConstBits - kPass1Bits);
    }
    const int by = b / c.bw, bx = b - (b / c.bw) * c.bw;
    uint8_t *dst = planes + c.plane + (int64_t)by * 8 * c.pitch + bx * 8;
#pragma unroll
    for (int r = 0; r < 8; ++r) {  // pass 2: rows
        int d[8], o[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) d[k] = ws[8 * r + k];
        idct8(d, o);
        uint32_t lo = 0, hi = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            lo |= range_limit(descale(o[k], kConstBits + kPass1Bits + 3)) << (8 * k);
            hi |= range_limit(descale(o[k + 4], kConstBits + kPass1Bits + 3)) << (8 * k);
        }
        *reinterpret_cast<uint2 *>(dst + (int64_t)r * c.pitch) = make_uint2(lo, hi);
    }
}

// ------------------------------------------------------ upsampling, color
// Sample of component c at output pixel (x, y), upsampled as libjpeg does.
__device__ __forceinline__ int sample_at(const JComp &c, const uint8_t *__restrict__ planes, int x, int y) {
    const uint8_t *P = planes + c.plane;
    if (c.hf == 1 && c.vf == 1) return P[(int64_t)y * c.pitch + x];
    const int cx = x / c.hf, cy = y / c.vf;
    const bool fancy = c.dw > 2 && c.hf == 2 && c.vf <= 2;
    if (!fancy) return P[(int64_t)cy * c.pitch + cx];  // replication
    const uint8_t *row = P + (int64_t)cy * c.pitch;
    if (c.vf == 1) {  // 2x1 triangle filter
        const int p = row[cx];
        if ((x & 1) == 0) return cx == 0 ? p : (3 * p + row[cx - 1] + 1) >> 2;
        return cx == c.dw - 1 ? p : (3 * p + row[cx + 1] + 2) >> 2;
    }
    // 2x2: the nearer row is cy, the further one above (even y) or below
    // (odd y), edge rows replicated
    const int r2 = (y & 1) == 0 ? max(cy - 1, 0) : min(cy + 1, c.dh - 1);
    const uint8_t *row2 = P + (int64_t)r2 * c.pitch;
    const int cs = 3 * row[cx] + row2[cx];
    if ((x & 1) == 0) return cx == 0 ? (4 * cs + 8) >> 4 : (3 * cs + 3 * row[cx - 1] + row2[cx - 1] + 8) >> 4;
    return cx == c.dw - 1 ? (4 * cs + 7) >> 4 : (3 * cs + 3 * row[cx + 1] + row2[cx + 1] + 7) >> 4;
}

__device__ __forceinline__ uint8_t clamp255(int v) { return (uint8_t)min(255, max(0, v)); }

// grid (ceil(max W / 256), max H, images)
__global__ __launch_bounds__(256) void jpeg_color_kernel(const JImage *__restrict__ imgs,
                                                         const uint8_t *__restrict__ planes,
                                                         uint8_t *__restrict__ rgb) {
    const JImage &im = imgs[blockIdx.z];
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= im.W || y >= im.H) return;
    const int Y = sample_at(im.c[0], planes, x, y);
    const int cb = sample_at(im.c[1], planes, x, y) - 128;
    const int cr = sample_at(im.c[2], planes, x, y) - 128;
    // libjpeg's tables: FIX(x) = x * 2^16 rounded; R and B rounded by 2^15
    const int r = Y + ((91881 * cr + 32768) >> 16);
    const int g = Y + ((-22554 * cb + 32768 - 46802 * cr) >> 16);
    const int b = Y + ((116130 * cb + 32768) >> 16);
    uint8_t *o = rgb + im.rgb + ((int64_t)y * im.W + x) * 3;
    o[0] = clamp255(r);
    o[1] = clamp255(g);
    o[2] = clamp255(b);
}

// ------------------------------------------------------------------- host
struct RawHuff {
    uint8_t counts[16];
    uint8_t vals[256];
    bool set = false;
};

struct ScanInfo {
    int ns = 0;
    int comp[4] = {0, 0, 0, 0};
    HuffTab dc[4], ac[4];
    int64_t begin = 0, end = 0;                        // entropy-coded bytes in the file
    std::vector<std::pair<int64_t, int64_t>> segs;     // restart segments [begin, end), stuffed
    std::vector<int64_t> seg_len;                      // their unstuffed lengths
};

// Entropy-coded bytes without the 0x00 that follows each 0xFF data byte
// (T.81 F.1.2.3); returns the length, writes to dst when it is not null.
int64_t unstuff(const uint8_t *src, int64_t n, uint8_t *dst) {
    int64_t i = 0, o = 0;
    while (i < n) {
        const void *ff = std::memchr(src + i, 0xFF, (size_t)(n - i));
        const int64_t j = ff ? (const uint8_t *)ff - src : n;
        if (dst) std::memcpy(dst + o, src + i, (size_t)(j - i));
        o += j - i;
        if (j >= n) break;
        if (dst) dst[o] = 0xFF;
        ++o;
        i = j + 1;
        if (i < n && src[i] == 0x00) ++i;
    }
    return o;
}

struct FileInfo {
    int status = 0;  // 0 ok, PRNU_EINVAL / PRNU_ENOTSUP: rejected
    std::string err;
    int W = 0, H = 0, ncomp = 0;
    int id[4] = {0, 0, 0, 0}, h[4] = {0, 0, 0, 0}, v[4] = {0, 0, 0, 0}, tq[4] = {0, 0, 0, 0};
    uint16_t q[4][64];       // latched per component at its first scan
    bool latched[4] = {false, false, false, false};
    int ri = 0;
    bool jfif = false, adobe = false;
    int adobe_t = -1;
    std::vector<ScanInfo> scans;
};

const int kZz[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                     41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                     30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

bool build_huff(const RawHuff &r, HuffTab &t) {
    std::memset(&t, 0, sizeof(t));
    int total = 0;
    for (int l = 0; l < 16; ++l) total += r.counts[l];
    if (total > 256) return false;
    std::memcpy(t.vals, r.vals, (size_t)total);
    int code = 0, k = 0;
    for (int l = 1; l <= 16; ++l) {
        t.valoff[l] = k - code;
        const int cnt = r.counts[l - 1];
        for (int i = 0; i < cnt; ++i) {
            if (l <= kLook) {
                const int lo = code << (kLook - l), n = 1 << (kLook - l);
                for (int j = 0; j < n; ++j) t.look[lo + j] = (uint16_t)((l << 8) | r.vals[k]);
            }
            ++code;
            ++k;
        }
        if (code > (1 << l)) return false;  // over-subscribed
        t.maxcode[l] = cnt ? code - 1 : -1;
        code <<= 1;
    }
    t.maxcode[17] = INT_MAX;
    return true;
}

int fail(FileInfo &f, int status, const std::string &msg) {
    f.status = status;
    f.err = msg;
    return status;
}

// Markers of one file (T.81 B.2); entropy-coded segments located with memchr.
int parse_file(const uint8_t *d, int64_t n, FileInfo &f) {
    if (n < 4 || d[0] != 0xFF || d[1] != 0xD8) return fail(f, PRNU_EINVAL, "not a JPEG file (no SOI marker)");
    uint16_t qt[4][64];
    bool have_q[4] = {false, false, false, false};
    RawHuff dc[4], ac[4];
    bool frame = false, eoi = false;
    int64_t i = 2;
    while (i < n) {
        if (d[i] != 0xFF) return fail(f, PRNU_EINVAL, "marker expected");
        while (i < n && d[i] == 0xFF) ++i;
        if (i >= n) break;
        const int m = d[i++];
        if (m == 0xD9) {
            eoi = true;
            break;
        }
        if (m == 0x01 || (m >= 0xD0 && m <= 0xD7)) continue;
        if (i + 2 > n) return fail(f, PRNU_EINVAL, "truncated marker segment");
        const int len = (d[i] << 8) | d[i + 1];
        if (len < 2 || i + len > n) return fail(f, PRNU_EINVAL, "truncated marker segment");
        const uint8_t *s = d + i + 2;
        const int sl = len - 2;
        i += len;
        if (m == 0xDB) {
            for (int j = 0; j < sl;) {
                const int pq = s[j] >> 4, tq = s[j] & 15;
                if (tq > 3 || j + 1 + (pq ? 128 : 64) > sl) return fail(f, PRNU_EINVAL, "bad DQT segment");
                for (int k = 0; k < 64; ++k)
                    qt[tq][kZz[k]] = pq ? (uint16_t)((s[j + 1 + 2 * k] << 8) | s[j + 2 + 2 * k]) : s[j + 1 + k];
                have_q[tq] = true;
                j += 1 + (pq ? 128 : 64);
            }
        } else if (m == 0xC4) {
            for (int j = 0; j < sl;) {
                if (j + 17 > sl) return fail(f, PRNU_EINVAL, "bad DHT segment");
                const int tc = s[j] >> 4, th = s[j] & 15;
                if (tc > 1 || th > 3) return fail(f, PRNU_EINVAL, "bad DHT segment");
                RawHuff &r = tc ? ac[th] : dc[th];
                int total = 0;
                for (int l = 0; l < 16; ++l) total += (r.counts[l] = s[j + 1 + l]);
                if (total > 256 || j + 17 + total > sl) return fail(f, PRNU_EINVAL, "bad DHT segment");
                std::memcpy(r.vals, s + j + 17, (size_t)total);
                r.set = true;
                j += 17 + total;
            }
        } else if (m == 0xC0 || m == 0xC1) {
            if (sl < 6 || s[0] != 8) return fail(f, PRNU_ENOTSUP, "only 8-bit JPEG samples are supported");
            f.H = (s[1] << 8) | s[2];
            f.W = (s[3] << 8) | s[4];
            f.ncomp = s[5];
            if (f.ncomp < 1 || f.ncomp > 4 || sl < 6 + 3 * f.ncomp) return fail(f, PRNU_EINVAL, "bad SOF segment");
            for (int k = 0; k < f.ncomp; ++k) {
                f.id[k] = s[6 + 3 * k];
                f.h[k] = s[7 + 3 * k] >> 4;
                f.v[k] = s[7 + 3 * k] & 15;
                f.tq[k] = s[8 + 3 * k] & 3;
                if (f.h[k] < 1 || f.h[k] > 4 || f.v[k] < 1 || f.v[k] > 4)
                    return fail(f, PRNU_EINVAL, "bad sampling factors");
            }
            frame = true;
        } else if ((m >= 0xC2 && m <= 0xC3) || (m >= 0xC5 && m <= 0xC7) || (m >= 0xC9 && m <= 0xCB) ||
                   (m >= 0xCD && m <= 0xCF)) {
            return fail(f, PRNU_ENOTSUP, "progressive, lossless or arithmetic-coded JPEG is not supported");
        } else if (m == 0xDD) {
            if (sl < 2) return fail(f, PRNU_EINVAL, "bad DRI segment");
            f.ri = (s[0] << 8) | s[1];
        } else if (m == 0xE0) {
            if (sl >= 5 && std::memcmp(s, "JFIF\0", 5) == 0) f.jfif = true;
        } else if (m == 0xEE) {
            if (sl >= 12 && std::memcmp(s, "Adobe", 5) == 0) {
                f.adobe = true;
                f.adobe_t = s[11];
            }
        } else if (m == 0xDA) {
            if (!frame) return fail(f, PRNU_EINVAL, "scan before the frame header");
            ScanInfo sc;
            sc.ns = sl >= 1 ? s[0] : 0;
            if (sc.ns < 1 || sc.ns > 4 || sl < 4 + 2 * sc.ns) return fail(f, PRNU_EINVAL, "bad SOS segment");
            for (int k = 0; k < sc.ns; ++k) {
                const int cid = s[1 + 2 * k];
                int ci = -1;
                for (int c = 0; c < f.ncomp; ++c)
                    if (f.id[c] == cid) ci = c;
                const int td = s[2 + 2 * k] >> 4, ta = s[2 + 2 * k] & 15;
                if (ci < 0 || td > 3 || ta > 3 || !dc[td].set || !ac[ta].set)
                    return fail(f, PRNU_EINVAL, "scan references an unknown component or table");
                if (!build_huff(dc[td], sc.dc[k]) || !build_huff(ac[ta], sc.ac[k]))
                    return fail(f, PRNU_EINVAL, "invalid Huffman table");
                sc.comp[k] = ci;
                if (!f.latched[ci]) {
                    if (!have_q[f.tq[ci]]) return fail(f, PRNU_EINVAL, "missing quantization table");
                    std::memcpy(f.q[ci], qt[f.tq[ci]], sizeof(f.q[ci]));
                    f.latched[ci] = true;
                }
            }
            const int ss = s[1 + 2 * sc.ns], se = s[2 + 2 * sc.ns], ahl = s[3 + 2 * sc.ns];
            if (ss != 0 || se != 63 || ahl != 0) return fail(f, PRNU_ENOTSUP, "not a sequential scan");
            // entropy-coded data: up to the first marker that is not stuffing or RSTn
            sc.begin = i;
            int64_t seg0 = i, j = i;
            while (true) {
                const void *ff = j < n ? std::memchr(d + j, 0xFF, (size_t)(n - j)) : nullptr;
                if (!ff || (const uint8_t *)ff - d + 1 >= n) {
                    j = n;
                    break;
                }
                j = (const uint8_t *)ff - d;
                const int nx = d[j + 1];
                if (nx == 0x00 || nx == 0xFF) {
                    j += (nx == 0x00) ? 2 : 1;
                    continue;
                }
                if (nx >= 0xD0 && nx <= 0xD7) {
                    sc.segs.push_back({seg0, j});
                    j += 2;
                    seg0 = j;
                    continue;
                }
                break;
            }
            sc.end = j;
            sc.segs.push_back({seg0, j});
            for (const auto &sg : sc.segs) sc.seg_len.push_back(unstuff(d + sg.first, sg.second - sg.first, nullptr));
            f.scans.push_back(sc);
            i = j;
        }
        // APPn, COM and the rest: skipped
    }
    if (!frame) return fail(f, PRNU_EINVAL, "no baseline or extended sequential frame (SOF0/SOF1)");
    // readJPEG's source fails when libjpeg asks for bytes past the end
    // (native.cpp:26-29), which it does for a file without EOI
    if (!eoi) return fail(f, PRNU_EINVAL, "truncated JPEG file (no EOI marker)");
    if (f.ncomp != 3) return fail(f, PRNU_ENOTSUP, "image not in RGB color space (not 3 components)");
    // libjpeg's color space for 3 components: YCbCr unless (no JFIF marker
    // and) an Adobe marker says transform 0 or the ids are 'R','G','B'
    const bool rgb_coded = !f.jfif && ((f.adobe && f.adobe_t == 0) ||
                                        (!f.adobe && f.id[0] == 82 && f.id[1] == 71 && f.id[2] == 66));
    if (rgb_coded) return fail(f, PRNU_ENOTSUP, "RGB-coded JPEG (no YCbCr transform) is not supported");
    if (f.W < 1 || f.H < 1) return fail(f, PRNU_EINVAL, "zero image dimension");
    int hmax = 1, vmax = 1;
    for (int c = 0; c < 3; ++c) {
        hmax = std::max(hmax, f.h[c]);
        vmax = std::max(vmax, f.v[c]);
    }
    for (int c = 0; c < 3; ++c) {
        if (hmax % f.h[c] || vmax % f.v[c]) return fail(f, PRNU_ENOTSUP, "non-integral chroma subsampling");
        if (hmax / f.h[c] == 1 && vmax / f.v[c] == 2)
            return fail(f, PRNU_ENOTSUP, "1x2 chroma subsampling is not supported");
        if (!f.latched[c]) return fail(f, PRNU_EINVAL, "a component is in no scan");
    }
    return PRNU_OK;
}

struct Pinned {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need) {
        if (need <= bytes) return PRNU_OK;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        if (hipHostMalloc(&p, need, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            set_error("jpeg: %zu bytes of pinned host memory", need);
            return PRNU_ENOMEM;
        }
        bytes = need;
        return PRNU_OK;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
    }
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

struct JpegDecoder {
    Pinned staging;        // files, descriptors, tables (one H2D copy)
    Workspace dev;         // the same on the device, then coefficients and planes
    Workspace bad;         // per-image corrupt-data flags
    hipStream_t own = nullptr;
};

namespace {

int decode_batch(JpegDecoder *dec, const uint8_t *const *files, const int64_t *sizes, int count, uint8_t *d_rgb,
                 int64_t rgb_stride, int64_t rgb_cap, int32_t *widths, int32_t *heights, int32_t *status,
                 hipStream_t st) {
    std::vector<FileInfo> info((size_t)count);
    {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int nt = std::max(1, std::min(count, (int)std::min(16u, hw)));
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t)
            pool.emplace_back([&, t]() {
                for (int k = t; k < count; k += nt) {
                    if (!files[k] || sizes[k] <= 0) {
                        info[k].status = PRNU_EINVAL;
                        info[k].err = "empty input";
                        continue;
                    }
                    (void)parse_file(files[k], sizes[k], info[k]);
                }
            });
        for (auto &th : pool) th.join();
    }
    // layout: descriptors, tables, files in one staging block; coefficients
    // and planes after it on the device
    std::vector<JImage> imgs((size_t)count);
    std::vector<JSeg> segs;
    std::vector<int> chunk_seg;
    std::vector<HuffTab> tabs;
    std::vector<uint16_t> qtab;
    std::vector<int64_t> file_data((size_t)count, -1);  // first unstuffed byte of each accepted file
    int64_t data_bytes = 0, coef_blocks = 0, plane_bytes = 0;
    int max_blocks = 0, max_w = 0, max_h = 0, max_chunks = 1;
    for (int k = 0; k < count; ++k) {
        FileInfo &f = info[k];
        JImage &im = imgs[k];
        std::memset(&im, 0, sizeof(im));
        if (f.status == PRNU_OK && (int64_t)f.W * f.H * 3 > rgb_stride) {
            f.status = PRNU_EINVAL;
            f.err = "insufficient buffer space provided for image";
        }
        if (f.status == PRNU_OK && (int64_t)k * rgb_stride + (int64_t)f.W * f.H * 3 > rgb_cap) {
            f.status = PRNU_EINVAL;
            f.err = "insufficient buffer space provided for image";
        }
        if (f.status != PRNU_OK) continue;
        int hmax = 1, vmax = 1;
        for (int c = 0; c < 3; ++c) {
            hmax = std::max(hmax, f.h[c]);
            vmax = std::max(vmax, f.v[c]);
        }
        const int mcux = (f.W + 8 * hmax - 1) / (8 * hmax), mcuy = (f.H + 8 * vmax - 1) / (8 * vmax);
        im.W = f.W;
        im.H = f.H;
        im.rgb = (int64_t)k * rgb_stride;
        for (int c = 0; c < 3; ++c) {
            JComp &jc = im.c[c];
            jc.h = f.h[c];
            jc.v = f.v[c];
            jc.hf = hmax / f.h[c];
            jc.vf = vmax / f.v[c];
            jc.bw = mcux * f.h[c];
            jc.bh = mcuy * f.v[c];
            jc.dw = (int)(((int64_t)f.W * f.h[c] + hmax - 1) / hmax);
            jc.dh = (int)(((int64_t)f.H * f.v[c] + vmax - 1) / vmax);
            jc.pitch = jc.bw * 8;
            jc.tq = (int)(qtab.size() / 64);
            qtab.insert(qtab.end(), f.q[c], f.q[c] + 64);
            jc.coef = coef_blocks;
            coef_blocks += (int64_t)jc.bw * jc.bh;
            jc.plane = plane_bytes;
            plane_bytes += (int64_t)jc.bw * 8 * jc.bh * 8;
            max_blocks = std::max(max_blocks, jc.bw * jc.bh);
        }
        max_w = std::max(max_w, f.W);
        max_h = std::max(max_h, f.H);
        // segments, their chunks and their place in the unstuffed data
        const size_t seg_first = segs.size(), chunk_first = chunk_seg.size();
        const int64_t data_first = data_bytes;
        for (const ScanInfo &sc : f.scans) {
            JSeg base;
            std::memset(&base, 0, sizeof(base));
            base.image = k;
            base.ns = sc.ns;
            int total;
            if (sc.ns == 1) {
                const JComp &jc = im.c[sc.comp[0]];
                base.mcux = (jc.dw + 7) / 8;
                total = base.mcux * ((jc.dh + 7) / 8);
                base.nbu = 1;
            } else {
                base.mcux = mcux;
                total = mcux * mcuy;
                int u = 0;
                for (int q = 0; q < sc.ns && u <= kMaxBlocksPerMcu; ++q) {
                    const JComp &jc = im.c[sc.comp[q]];
                    for (int vv = 0; vv < jc.v; ++vv)
                        for (int hh = 0; hh < jc.h; ++hh) {
                            if (u < kMaxBlocksPerMcu) {
                                base.uslot[u] = q;
                                base.uv[u] = vv;
                                base.uh[u] = hh;
                            }
                            ++u;
                        }
                }
                base.nbu = u;
            }
            if (base.nbu > kMaxBlocksPerMcu) {
                f.status = PRNU_EINVAL;
                f.err = "more than 10 blocks per MCU";
                break;
            }
            for (int q = 0; q < sc.ns; ++q) {
                base.comp[q] = sc.comp[q];
                base.dc[q] = (int)tabs.size();
                tabs.push_back(sc.dc[q]);
                base.ac[q] = (int)tabs.size();
                tabs.push_back(sc.ac[q]);
            }
            const int ri = f.ri > 0 ? f.ri : total;
            const int nseg = (total + ri - 1) / ri;
            if ((int)sc.segs.size() != nseg) {
                f.status = PRNU_EINVAL;
                f.err = "restart markers do not match the restart interval";
                break;
            }
            for (int q = 0; q < nseg; ++q) {
                JSeg g = base;
                g.mcu0 = q * ri;
                g.mcus = std::min(ri, total - q * ri);
                g.begin = data_bytes;
                g.end = data_bytes + sc.seg_len[q];
                data_bytes = g.end;
                g.chunk0 = (int)chunk_seg.size();
                g.nchunk = (int)std::max<int64_t>(1, (sc.seg_len[q] + kChunkBytes - 1) / kChunkBytes);
                max_chunks = std::max(max_chunks, g.nchunk);
                for (int t = 0; t < g.nchunk; ++t) chunk_seg.push_back((int)segs.size());
                segs.push_back(g);
            }
        }
        if (f.status != PRNU_OK) {
            segs.resize(seg_first);
            chunk_seg.resize(chunk_first);
            data_bytes = data_first;
            im.W = im.H = 0;
            continue;
        }
        file_data[k] = data_first;
    }
    const int nseg = (int)segs.size(), nchunk = (int)chunk_seg.size();
    const size_t o_img = 0, o_seg = align256(sizeof(JImage) * count);
    const size_t o_tab = o_seg + align256(sizeof(JSeg) * std::max(1, nseg));
    const size_t o_q = o_tab + align256(sizeof(HuffTab) * std::max<size_t>(1, tabs.size()));
    const size_t o_cseg = o_q + align256(sizeof(uint16_t) * std::max<size_t>(64, qtab.size()));
    const size_t o_data = o_cseg + align256(sizeof(int) * std::max(1, nchunk));
    const size_t staged = o_data + align256((size_t)data_bytes + 8);
    // device only: chunk states, coefficients, planes
    const size_t o_entry = staged;
    const size_t o_exit = o_entry + align256(sizeof(int64_t) * std::max(1, nchunk));
    const size_t o_started = o_exit + align256(sizeof(int64_t) * std::max(1, nchunk));
    const size_t o_first = o_started + align256(sizeof(int) * std::max(1, nchunk));
    const size_t o_dirty = o_first + align256(sizeof(int) * std::max(1, nchunk));
    const size_t o_coef = o_dirty + align256(sizeof(int) * std::max(1, nchunk));
    const size_t o_plane = o_coef + align256((size_t)coef_blocks * 128);
    const size_t total = o_plane + align256((size_t)plane_bytes);
    PRNU_TRY(dec->staging.ensure(staged));
    PRNU_TRY(dec->dev.ensure(total));
    PRNU_TRY(dec->bad.ensure(sizeof(int) * ((size_t)count + 1)));
    uint8_t *hs = static_cast<uint8_t *>(dec->staging.p);
    std::memcpy(hs + o_img, imgs.data(), sizeof(JImage) * count);
    if (nseg) std::memcpy(hs + o_seg, segs.data(), sizeof(JSeg) * nseg);
    if (!tabs.empty()) std::memcpy(hs + o_tab, tabs.data(), sizeof(HuffTab) * tabs.size());
    if (!qtab.empty()) std::memcpy(hs + o_q, qtab.data(), sizeof(uint16_t) * qtab.size());
    if (nchunk) std::memcpy(hs + o_cseg, chunk_seg.data(), sizeof(int) * nchunk);
    {
        // unstuffed entropy-coded data, file by file in parallel
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int nt = std::max(1, std::min(count, (int)std::min(16u, hw)));
        std::vector<std::thread> pool;
        for (int t = 0; t < nt; ++t)
            pool.emplace_back([&, t]() {
                for (int k = t; k < count; k += nt) {
                    if (file_data[k] < 0) continue;
                    int64_t off = file_data[k];
                    for (const ScanInfo &sc : info[k].scans)
                        for (size_t q = 0; q < sc.segs.size(); ++q)
                            off += unstuff(files[k] + sc.segs[q].first, sc.segs[q].second - sc.segs[q].first,
                                           hs + o_data + off);
                }
            });
        for (auto &th : pool) th.join();
    }
    uint8_t *dv = static_cast<uint8_t *>(dec->dev.ptr);
    int *d_bad = static_cast<int *>(dec->bad.ptr);
    int *d_changed = d_bad + count;
    PRNU_CHECK_HIP(hipMemcpyAsync(dv, hs, staged, hipMemcpyHostToDevice, st));
    PRNU_CHECK_HIP(hipMemsetAsync(d_bad, 0, sizeof(int) * (size_t)count, st));
    const JImage *d_img = reinterpret_cast<const JImage *>(dv + o_img);
    const JSeg *d_seg = reinterpret_cast<const JSeg *>(dv + o_seg);
    const HuffTab *d_tab = reinterpret_cast<const HuffTab *>(dv + o_tab);
    const int *d_cseg = reinterpret_cast<const int *>(dv + o_cseg);
    const uint8_t *d_data = dv + o_data;
    int64_t *d_entry = reinterpret_cast<int64_t *>(dv + o_entry);
    int64_t *d_exit = reinterpret_cast<int64_t *>(dv + o_exit);
    int *d_started = reinterpret_cast<int *>(dv + o_started);
    int *d_first = reinterpret_cast<int *>(dv + o_first);
    int *d_dirty = reinterpret_cast<int *>(dv + o_dirty);
    int16_t *d_coef = reinterpret_cast<int16_t *>(dv + o_coef);
    uint8_t *d_plane = dv + o_plane;
    if (nseg) {
        PRNU_CHECK_HIP(hipMemsetAsync(d_coef, 0, (size_t)coef_blocks * 128, st));
        const dim3 g256((unsigned)((nchunk + 255) / 256)), g64((unsigned)((nchunk + 63) / 64));
        hipLaunchKernelGGL(jpeg_chunk_init_kernel, g256, dim3(256), 0, st, d_seg, d_cseg, nchunk, d_entry, d_dirty);
        PRNU_CHECK_LAUNCH();
        // entry states settle: after round r the first r + 1 chunks of every
        // segment are exact, so max_chunks + 1 rounds always suffice
        for (int round = 0; round <= max_chunks; ++round) {
            hipLaunchKernelGGL(jpeg_chunk_sync_kernel, g64, dim3(64), 0, st, d_seg, d_cseg, nchunk, d_img, d_tab,
                               d_data, d_entry, d_exit, d_started, d_dirty);
            PRNU_CHECK_LAUNCH();
            PRNU_CHECK_HIP(hipMemsetAsync(d_changed, 0, sizeof(int), st));
            hipLaunchKernelGGL(jpeg_chunk_link_kernel, g256, dim3(256), 0, st, d_seg, d_cseg, nchunk, d_exit,
                               d_entry, d_dirty, d_changed);
            PRNU_CHECK_LAUNCH();
            int changed = 0;
            PRNU_CHECK_HIP(hipMemcpyAsync(&changed, d_changed, sizeof(int), hipMemcpyDeviceToHost, st));
            PRNU_CHECK_HIP(hipStreamSynchronize(st));
            if (!changed) break;
        }
        hipLaunchKernelGGL(jpeg_chunk_scan_kernel, dim3((unsigned)((nseg + 63) / 64)), dim3(64), 0, st, d_seg, nseg,
                           d_started, d_first);
        PRNU_CHECK_LAUNCH();
        hipLaunchKernelGGL(jpeg_chunk_write_kernel, g64, dim3(64), 0, st, d_seg, d_cseg, nchunk, d_img, d_tab, d_data,
                           d_entry, d_first, d_coef, d_bad);
        PRNU_CHECK_LAUNCH();
        hipLaunchKernelGGL(jpeg_dc_scan_kernel, dim3((unsigned)nseg, 4), dim3(256), 0, st, d_seg, d_img, d_coef);
        PRNU_CHECK_LAUNCH();
        hipLaunchKernelGGL(jpeg_idct_kernel, dim3((unsigned)((max_blocks + 255) / 256), 3, (unsigned)count),
                           dim3(256), 0, st, d_img, reinterpret_cast<const uint16_t *>(dv + o_q), d_coef, d_plane);
        PRNU_CHECK_LAUNCH();
        hipLaunchKernelGGL(jpeg_color_kernel, dim3((unsigned)((max_w + 255) / 256), (unsigned)max_h, (unsigned)count),
                           dim3(256), 0, st, d_img, d_plane, d_rgb);
        PRNU_CHECK_LAUNCH();
    }
    std::vector<int> bad((size_t)count, 0);
    PRNU_CHECK_HIP(hipMemcpyAsync(bad.data(), d_bad, sizeof(int) * (size_t)count, hipMemcpyDeviceToHost, st));
    PRNU_CHECK_HIP(hipStreamSynchronize(st));
    int first_err = -1;
    for (int k = 0; k < count; ++k) {
        const FileInfo &f = info[k];
        if (widths) widths[k] = f.status == PRNU_OK ? f.W : 0;
        if (heights) heights[k] = f.status == PRNU_OK ? f.H : 0;
        int s = f.status;
        if (s == PRNU_OK && bad[k]) s = PRNU_JPEG_CORRUPT_DATA;
        if (status) status[k] = s;
        if (f.status != PRNU_OK && first_err < 0) first_err = k;
    }
    if (first_err >= 0) {
        set_error("jpeg: file %d: %s", first_err, info[first_err].err.c_str());
        return info[first_err].status;
    }
    return PRNU_OK;
}

}  // namespace

int jpeg_info(const uint8_t *d, int64_t n, int *w, int *h) {
    FileInfo f;
    const int rc = parse_file(d, n, f);
    if (rc != PRNU_OK) {
        set_error("jpeg: %s", f.err.c_str());
        return rc;
    }
    *w = f.W;
    *h = f.H;
    return PRNU_OK;
}

}  // namespace prnu

using namespace prnu;

extern "C" {

int prnu_jpeg_decoder_create(prnu_jpeg_decoder **out) {
    PRNU_REQUIRE(out, "jpeg_decoder_create: null output");
    JpegDecoder *d = new (std::nothrow) JpegDecoder();
    if (!d) {
        set_error("jpeg_decoder_create: out of host memory");
        return PRNU_ENOMEM;
    }
    *out = reinterpret_cast<prnu_jpeg_decoder *>(d);
    return PRNU_OK;
}

int prnu_jpeg_decoder_destroy(prnu_jpeg_decoder *dec) {
    if (!dec) return PRNU_OK;
    JpegDecoder *d = reinterpret_cast<JpegDecoder *>(dec);
    (void)hipDeviceSynchronize();
    d->staging.release();
    d->dev.release();
    d->bad.release();
    delete d;
    return PRNU_OK;
}

int prnu_jpeg_info(const void *h_file, int64_t size, int32_t *width, int32_t *height) {
    PRNU_REQUIRE(h_file && width && height && size > 0, "jpeg_info: null pointer or empty file");
    int w = 0, h = 0;
    PRNU_TRY(jpeg_info(static_cast<const uint8_t *>(h_file), size, &w, &h));
    *width = w;
    *height = h;
    return PRNU_OK;
}

int prnu_jpeg_decode_batch(prnu_jpeg_decoder *dec, const void *const *h_files, const int64_t *h_sizes, int count,
                           uint8_t *d_rgb, int64_t rgb_stride, int64_t rgb_capacity, int32_t *h_width,
                           int32_t *h_height, int32_t *h_status, prnu_stream stream) {
    PRNU_REQUIRE(dec && h_files && h_sizes && d_rgb, "jpeg_decode_batch: null pointer");
    PRNU_REQUIRE(count >= 1 && count <= 65535, "jpeg_decode_batch: 1..65535 files, got %d", count);
    PRNU_REQUIRE(rgb_stride > 0 && rgb_capacity >= rgb_stride, "jpeg_decode_batch: bad output stride or capacity");
    return decode_batch(reinterpret_cast<JpegDecoder *>(dec), reinterpret_cast<const uint8_t *const *>(h_files),
                        h_sizes, count, d_rgb, rgb_stride, rgb_capacity, h_width, h_height, h_status,
                        (hipStream_t)stream);
}

}  // extern "C"
