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
// cut into chunks of kChunkBytes, one lane each, and every chunk is first
// decoded from a guessed state (its first bit, block 0 of an MCU, a DC symbol
// next).  Linking rounds then give chunk t the exit state (bit position,
// block of the MCU, coefficient index) of chunk t-1 and decode again the
// chunks whose entry changed, until a round changes nothing.  Huffman codes
// resynchronise within some hundreds of bits on photographic data, so a
// guessed chunk usually ends in the exact state and a few rounds settle a
// batch; the result does not depend on that speed, only on the exact fixed
// point (rounds are bounded by the chunks per segment).  The rounds also
// count the blocks each chunk starts and sum its DC differences per
// component; a prefix over each segment's chunks then gives every chunk its
// first block and DC predictors, and a write pass stores the coefficients
// with final DC values.  jpeg_idct_kernel runs one block per thread,
// jpeg_color_kernel four output pixels per thread.
#include "prnu_internal.hpp"

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstring>
#include <mutex>
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
    int tab_off, tab_len;       // the scan's two-level tables (uint16 entries) in the batch table array
    int trel[3];                // per scan slot: DC table offset | AC table offset << 16, in the scan's tables
    int64_t begin, end;         // unstuffed entropy-coded bytes in the batch buffer
};

// zigzag position -> natural (row-major) index; positions past 63 (corrupt
// data) land on 63, as libjpeg's extended table does
__constant__ uint8_t kZigzag[80] = {
    0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13,
    6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51, 58, 59, 52, 45, 38, 31,
    39, 46, 53, 60, 61, 54, 47, 55, 62, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63, 63};

// ---------------------------------------------------------- chunk decoding
// One lane decodes one chunk.  A workgroup is one wave of kLanes chunk
// decoders of one scan (the host pads every scan's chunk list to a multiple
// of kLanes).  Before decoding it stages in LDS
//   * each working lane's window: the chunk's bytes from the 4-byte boundary
//     at or before its first byte, as big-endian dwords, zero from the
//     segment's end on (past the end of a segment the stream continues with
//     zero bits, as libjpeg reads a premature marker);
//   * the scan's two-level decoding tables (build_lut2).
// A lane's reads stay inside its window: it enters at most 30 bits past its
// chunk's first bit (the exit of the chunk before it, whose last symbol
// started before that chunk's end), stops before the first symbol that
// starts at or after its own end, and a symbol (code and extra bits) spans at
// most 31 bits; with 64 bits in the accumulator and the next word preloaded
// that is at most 3 + kChunkBytes + 12 bytes of the 4 * kWinWords.
constexpr int kLanes = 64;
constexpr int kWinWords = ((kChunkBytes + 24) / 4) | 1;  // odd, which spreads the lanes' windows over the LDS banks
constexpr int kScanTabWords = 8192;  // a scan's tables: first levels (512 entries each) and subtables

struct ChunkLds {
    uint32_t win[kLanes * kWinWords];
    uint32_t tab[kScanTabWords / 2];  // uint16 entries
    int64_t wbase[kLanes];            // per compact working slot: first window byte
    int64_t wend[kLanes];             // and the end of its segment (bytes)
};

// MSB-first bits of one lane's window: acc holds the next n bits, nxt the
// window word after them.
struct WinBits {
    const uint32_t *w;
    uint64_t acc;
    uint32_t nxt;
    int n, wi;  // valid bits in acc; window index of nxt
    __device__ uint32_t word(int i) const { return i < kWinWords ? w[i] : 0u; }
    __device__ void init(const uint32_t *win, int rel) {  // rel: first bit, relative to the window
        w = win;
        const int i = rel >> 5, sh = rel & 31;
        acc = (((uint64_t)word(i) << 32) | word(i + 1)) << sh;
        n = 64 - sh;
        wi = i + 2;
        nxt = word(wi);
    }
    __device__ void refill() {  // n >= 33 afterwards
        if (n <= 32) {
            acc |= (uint64_t)nxt << (32 - n);
            n += 32;
            nxt = word(++wi);
        }
    }
    __device__ int pos() const { return wi * 32 - n; }  // relative to the window
    __device__ int take(int s) {                         // s in [0, 15]
        if (s == 0) return 0;
        const int v = (int)(acc >> (64 - s));
        acc <<= s;
        n -= s;
        return v;
    }
};

__device__ __forceinline__ int extend(int v, int s) { return (s && v < (1 << (s - 1))) ? v - (1 << s) + 1 : v; }

// Entry of the code at the head of acc in a two-level table T: (length << 8)
// | symbol, 0 for no such code, -1 when the canonical walk must decide.
__device__ __forceinline__ int lut_entry(const uint16_t *T, uint64_t acc) {
    int e = T[(int)(acc >> 55)];
    if (e & 0x8000) {
        const int off = e & 0xFFF, nb = (e >> 12) & 7;
        e = off == 0xFFF ? -1 : (int)T[512 + off + ((int)(acc >> (55 - nb)) & ((1 << nb) - 1))];
    }
    return e;
}

// Canonical decoding (T.81 F.2.2.3) of the code at the head of acc.
__device__ int huff_walk(const HuffTab &t, uint64_t acc) {
    for (int l = 1; l <= 16; ++l) {
        const int code = (int)(acc >> (64 - l));
        if (code <= t.maxcode[l]) return (l << 8) | t.vals[(t.valoff[l] + code) & 0xFF];
    }
    return 0;
}

// Decoder state between two symbols: bit position, block of the MCU, next
// coefficient index (0: a DC symbol is next).
__device__ __forceinline__ int64_t pack_state(int64_t pos, int u, int k) { return (pos << 16) | (u << 8) | k; }

template <class T>
__device__ __forceinline__ T sel3(int s, T a, T b, T c) {
    return s == 0 ? a : s == 1 ? b : c;
}

// What a lane knows of its segment: per block of the MCU (6 bits each) the
// scan slot, block row and block column inside the MCU; per scan slot the
// offsets of its DC and AC tables in the scan's two-level tables (low / high
// half) and, for the canonical walk, the global index of slot 0's DC table
// (DC and AC alternate per slot).
struct Lane {
    uint64_t umap;
    int nbu, gt0, tr0, tr1, tr2;
};

__device__ __forceinline__ Lane lane_of(const JSeg &g) {
    Lane ln;
    ln.nbu = g.nbu;
    ln.gt0 = g.dc[0];
    ln.tr0 = g.trel[0];
    ln.tr1 = g.trel[g.ns > 1 ? 1 : 0];
    ln.tr2 = g.trel[g.ns > 2 ? 2 : 0];
    uint64_t m = 0;
    for (int u = 0; u < g.nbu && u < kMaxBlocksPerMcu; ++u)
        m |= (uint64_t)(g.uslot[u] | (g.uv[u] << 2) | (g.uh[u] << 4)) << (6 * u);
    ln.umap = m;
    return ln;
}

// Where the write pass puts coefficient blocks: per scan slot the
// component's first block, block columns and blocks per MCU across / down
// (1 x 1 in a single-component scan); the MCU of the current block.
struct Cursor {
    int16_t *coef, *blk;
    int64_t cb0, cb1, cb2;
    int bw0, bw1, bw2, h0, h1, h2, v0, v1, v2;
    int cur, total, mcux, mrow, mcol;
    __device__ int16_t *addr(uint64_t umap, int u) const {
        if (cur < 0 || cur >= total) return nullptr;
        const int f = (int)(umap >> (6 * u)) & 63, s = f & 3, uv = (f >> 2) & 3, uh = f >> 4;
        const int64_t b = sel3(s, cb0, cb1, cb2) + (int64_t)(mrow * sel3(s, v0, v1, v2) + uv) * sel3(s, bw0, bw1, bw2) +
                          mcol * sel3(s, h0, h1, h2) + uh;
        return coef + b * 64;
    }
};

// Decode symbols from (bit position, block u, coefficient k) while the
// position is before `stop` (relative to the lane's window).  d0..d2: per
// scan slot, running sums of the DC differences (the DC predictors in the
// write pass, which also stores the coefficients and stops after the
// segment's last block).  A code that does not exist: libjpeg warns, takes
// symbol 0 after 17 bits; flagged in `bad`.  Returns the exit position; u
// and k are updated.
template <bool WRITE>
__device__ int decode_chunk(WinBits &br, int stop, int &u, int &k, const Lane &ln, const uint16_t *T,
                            const HuffTab *__restrict__ tabs, int &started, int &d0, int &d1, int &d2, Cursor &cz,
                            int &bad) {
    int slot = (int)(ln.umap >> (6 * u)) & 3;
    int trel = sel3(slot, ln.tr0, ln.tr1, ln.tr2);
    int n = 0;
    while (br.pos() < stop) {
        br.refill();
        int e = lut_entry(T + (k == 0 ? (trel & 0xFFFF) : (trel >> 16)), br.acc);
        if (e < 0) e = huff_walk(tabs[ln.gt0 + 2 * slot + (k == 0 ? 0 : 1)], br.acc);
        int len = e >> 8;
        if (e == 0) {
            bad = 1;
            len = 17;
        }
        br.acc <<= len;
        br.n -= len;
        const int sym = e & 0xFF, s = sym & 15, r = sym >> 4;
        const int v = extend(br.take(s), s);
        if (k == 0) {  // DC: the difference category s (<= 15, checked on the host)
            ++n;
            const int dv = sel3(slot, d0, d1, d2) + v;
            d0 = slot == 0 ? dv : d0;
            d1 = slot == 1 ? dv : d1;
            d2 = slot == 2 ? dv : d2;
            if (WRITE) {
                ++cz.cur;
                cz.blk = cz.addr(ln.umap, u);
                if (cz.blk) cz.blk[0] = (int16_t)dv;
            }
            k = 1;
        } else if (s) {
            k += r;
            if (WRITE && cz.blk) cz.blk[kZigzag[min(k, 79)]] = (int16_t)v;
            ++k;
        } else {
            k = (r == 15) ? k + 16 : 64;  // ZRL or end of block
        }
        if (k >= 64) {
            k = 0;
            if (++u == ln.nbu) {
                u = 0;
                if (WRITE && ++cz.mcol == cz.mcux) {
                    cz.mcol = 0;
                    ++cz.mrow;
                }
            }
            slot = (int)(ln.umap >> (6 * u)) & 3;
            trel = sel3(slot, ln.tr0, ln.tr1, ln.tr2);
            // the segment's last block is complete: what follows is the
            // 1-bit padding before the marker, never decoded (by libjpeg
            // either), so it cannot raise the corrupt-data flag
            if (WRITE && cz.cur + 1 >= cz.total) break;
        }
    }
    started = n;
    return br.pos();
}

__device__ __forceinline__ int64_t chunk_start_byte(const JSeg &g, int ch) {
    return g.begin + (int64_t)(ch - g.chunk0) * kChunkBytes;
}

// Stage the working lanes' windows and the scan's tables (lane 0's, the
// same for every lane of the workgroup).  Returns this lane's compact slot
// (-1 when it does not work).
__device__ int stage_chunks(ChunkLds &S, const uint8_t *__restrict__ data, const uint16_t *__restrict__ lut,
                            bool work, int64_t start_byte, int64_t seg_end, int tab_off, int tab_len) {
    const int t = threadIdx.x;
    const uint64_t mask = __ballot(work);
    const int nwork = __popcll(mask);
    const int slot = work ? __popcll(mask & ((1ull << t) - 1)) : -1;
    if (work) {
        S.wbase[slot] = start_byte & ~(int64_t)3;
        S.wend[slot] = seg_end;
    }
    const int toff = __shfl(tab_off, 0), twords = __shfl(tab_len, 0) / 2;
    __syncthreads();
    for (int L = t; L < nwork * kWinWords; L += kLanes) {
        const int s = L / kWinWords, w = L - s * kWinWords;
        const int64_t byte = S.wbase[s] + 4 * w, end = S.wend[s];
        uint32_t v = 0;
        if (byte < end) {
            v = __builtin_bswap32(*reinterpret_cast<const uint32_t *>(data + byte));
            if (byte + 4 > end) v &= ~0u << (8 * (int)(byte + 4 - end));
        }
        S.win[s * kWinWords + w] = v;
    }
    const uint32_t *src = reinterpret_cast<const uint32_t *>(lut + toff);
    for (int L = t; L < twords; L += kLanes) S.tab[L] = src[L];
    __syncthreads();
    return slot;
}

// One linking round.  Round 0 decodes every chunk from its own first byte
// at the start of an MCU (exact for a segment's first chunk, a guess for the
// others).  Round r > 0 gives chunk t the exit state of chunk t-1 from round
// r-1 and decodes it again if that differs from its entry; exits go to
// exit_out (exit_in is the previous round's).  After round r the first r + 1
// chunks of every segment are exact; changed[r] is set when any chunk was
// decoded again, and a round after one that changed nothing returns at once.
// Per chunk: blocks started (DC symbols) and per scan slot the sum of the DC
// differences.  chunk_seg[ch] < 0: padding.
__global__ __launch_bounds__(kLanes) void jpeg_chunk_round_kernel(
    const JSeg *__restrict__ segs, const int *__restrict__ chunk_seg, int nchunk, const HuffTab *__restrict__ tabs,
    const uint16_t *__restrict__ lut, const uint8_t *__restrict__ data, int round, int64_t *__restrict__ entry,
    const int64_t *__restrict__ exit_in, int64_t *__restrict__ exit_out, int *__restrict__ started,
    int *__restrict__ dcsum, int *__restrict__ changed) {
    __shared__ ChunkLds S;
    if (round > 0 && changed[round - 1] == 0) return;
    const int ch = blockIdx.x * kLanes + threadIdx.x;
    const int sg = ch < nchunk ? chunk_seg[ch] : -1;
    const bool valid = sg >= 0;
    const JSeg &g = segs[valid ? sg : chunk_seg[blockIdx.x * kLanes]];
    const int64_t start = chunk_start_byte(g, valid ? ch : g.chunk0);
    bool work = false;
    int64_t e = 0;
    if (valid) {
        if (round == 0) {
            e = pack_state(start * 8, 0, 0);
            entry[ch] = e;
            work = true;
        } else if (ch == g.chunk0) {
            exit_out[ch] = exit_in[ch];
        } else {
            e = exit_in[ch - 1];
            if (e != entry[ch]) {
                entry[ch] = e;
                work = true;
            } else {
                exit_out[ch] = exit_in[ch];
            }
        }
    }
    if (__ballot(work) == 0) return;
    const int slot = stage_chunks(S, data, lut, work, start, g.end, g.tab_off, g.tab_len);
    if (!work) return;
    const Lane ln = lane_of(g);
    WinBits br;
    const int64_t wbit = (start & ~(int64_t)3) * 8;
    br.init(S.win + slot * kWinWords, (int)((e >> 16) - wbit));
    const int stop = (int)(min(g.end, start + kChunkBytes) * 8 - wbit);
    int u = (int)((e >> 8) & 0xFF), k = (int)(e & 0xFF), n = 0, d0 = 0, d1 = 0, d2 = 0, bad = 0;
    Cursor cz;
    const int p = decode_chunk<false>(br, stop, u, k, ln, reinterpret_cast<const uint16_t *>(S.tab), tabs, n, d0, d1,
                                      d2, cz, bad);
    exit_out[ch] = pack_state(wbit + p, u, k);
    started[ch] = n;
    dcsum[3 * ch] = d0;
    dcsum[3 * ch + 1] = d1;
    dcsum[3 * ch + 2] = d2;
    changed[round] = 1;
}

// Per segment (one workgroup each): exclusive prefix over its chunks of the
// blocks started and the per-slot DC difference sums, i.e. each chunk's
// first block and its DC predictors (prediction restarts at each segment).
__global__ __launch_bounds__(256) void jpeg_chunk_scan_kernel(const JSeg *__restrict__ segs,
                                                              const int *__restrict__ started,
                                                              const int *__restrict__ dcsum, int *__restrict__ first,
                                                              int *__restrict__ pred) {
    __shared__ int4 part[256];
    const JSeg &g = segs[blockIdx.x];
    const int t = threadIdx.x, span = (g.nchunk + 255) / 256;
    const int c0 = g.chunk0 + min(g.nchunk, t * span), c1 = g.chunk0 + min(g.nchunk, (t + 1) * span);
    int4 s = make_int4(0, 0, 0, 0);
    for (int ch = c0; ch < c1; ++ch) {
        s.x += started[ch];
        s.y += dcsum[3 * ch];
        s.z += dcsum[3 * ch + 1];
        s.w += dcsum[3 * ch + 2];
    }
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {  // inclusive Hillis-Steele scan
        const int4 v = t >= off ? part[t - off] : make_int4(0, 0, 0, 0);
        __syncthreads();
        part[t] = make_int4(part[t].x + v.x, part[t].y + v.y, part[t].z + v.z, part[t].w + v.w);
        __syncthreads();
    }
    int4 run = t > 0 ? part[t - 1] : make_int4(0, 0, 0, 0);
    for (int ch = c0; ch < c1; ++ch) {
        first[ch] = run.x;
        pred[3 * ch] = run.y;
        pred[3 * ch + 1] = run.z;
        pred[3 * ch + 2] = run.w;
        run.x += started[ch];
        run.y += dcsum[3 * ch];
        run.z += dcsum[3 * ch + 1];
        run.w += dcsum[3 * ch + 2];
    }
}

// Every chunk from its exact entry state: coefficients into the zeroed
// coefficient blocks, DC as the predicted value.
__global__ __launch_bounds__(kLanes) void jpeg_chunk_write_kernel(
    const JSeg *__restrict__ segs, const int *__restrict__ chunk_seg, int nchunk, const JImage *__restrict__ imgs,
    const HuffTab *__restrict__ tabs, const uint16_t *__restrict__ lut, const uint8_t *__restrict__ data,
    const int64_t *__restrict__ entry, const int *__restrict__ first, const int *__restrict__ pred,
    int16_t *__restrict__ coef, int *__restrict__ bad_out) {
    __shared__ ChunkLds S;
    const int ch = blockIdx.x * kLanes + threadIdx.x;
    const int sg = ch < nchunk ? chunk_seg[ch] : -1;
    const bool valid = sg >= 0;
    const JSeg &g = segs[valid ? sg : chunk_seg[blockIdx.x * kLanes]];
    const int64_t start = chunk_start_byte(g, valid ? ch : g.chunk0);
    const int slot = stage_chunks(S, data, lut, valid, start, g.end, g.tab_off, g.tab_len);
    if (!valid) return;
    const Lane ln = lane_of(g);
    const int64_t e = entry[ch];
    WinBits br;
    const int64_t wbit = (start & ~(int64_t)3) * 8;
    br.init(S.win + slot * kWinWords, (int)((e >> 16) - wbit));
    const int stop = (int)(min(g.end, start + kChunkBytes) * 8 - wbit);
    int u = (int)((e >> 8) & 0xFF), k = (int)(e & 0xFF), n = 0, bad = 0;
    int d0 = pred[3 * ch], d1 = pred[3 * ch + 1], d2 = pred[3 * ch + 2];
    const JImage &im = imgs[g.image];
    Cursor cz;
    cz.coef = coef;
    const bool single = g.ns == 1;
    {
        const JComp &a = im.c[g.comp[0]];
        const JComp &b = im.c[g.comp[g.ns > 1 ? 1 : 0]];
        const JComp &c = im.c[g.comp[g.ns > 2 ? 2 : 0]];
        cz.cb0 = a.coef;
        cz.cb1 = b.coef;
        cz.cb2 = c.coef;
        cz.bw0 = a.bw;
        cz.bw1 = b.bw;
        cz.bw2 = c.bw;
        cz.h0 = single ? 1 : a.h;
        cz.h1 = b.h;
        cz.h2 = c.h;
        cz.v0 = single ? 1 : a.v;
        cz.v1 = b.v;
        cz.v2 = c.v;
    }
    const int f = first[ch];
    cz.total = g.mcus * g.nbu;
    cz.cur = f - 1;  // the block in progress (k > 0) or the last one completed
    const int bpos = max(0, k > 0 ? f - 1 : f);
    const int m = g.mcu0 + bpos / g.nbu;
    cz.mcux = g.mcux;
    cz.mrow = m / g.mcux;
    cz.mcol = m - cz.mrow * g.mcux;
    cz.blk = k > 0 ? cz.addr(ln.umap, u) : nullptr;
    (void)decode_chunk<true>(br, stop, u, k, ln, reinterpret_cast<const uint16_t *>(S.tab), tabs, n, d0, d1, d2, cz,
                             bad);
    if (bad) bad_out[g.image] = 1;
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

// Chroma of four output pixels x0..x0+3 (x0 a multiple of 4) of row y from
// a plane upsampled by the triangle filter, 2x1 (VF 1) or 2x2 (VF 2): the
// four samples around columns x0/2 and x0/2 + 1, clamped to the component's
// width, which is libjpeg's edge rule.
template <int VF>
__device__ __forceinline__ void chroma_fancy4(const JComp &c, const uint8_t *__restrict__ planes, int x0, int y,
                                              int out[4]) {
    const uint8_t *P = planes + c.plane;
    const int cx = x0 >> 1;
    int col[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) col[j] = min(max(cx - 1 + j, 0), c.dw - 1);
    if (VF == 1) {
        const uint8_t *row = P + (int64_t)y * c.pitch;
        int p[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) p[j] = row[col[j]];
        out[0] = (3 * p[1] + p[0] + 1) >> 2;
        out[1] = (3 * p[1] + p[2] + 2) >> 2;
        out[2] = (3 * p[2] + p[1] + 1) >> 2;
        out[3] = (3 * p[2] + p[3] + 2) >> 2;
    } else {
        const int cy = y >> 1;
        const int r2 = (y & 1) == 0 ? max(cy - 1, 0) : min(cy + 1, c.dh - 1);
        const uint8_t *row = P + (int64_t)cy * c.pitch, *row2 = P + (int64_t)r2 * c.pitch;
        int cs[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) cs[j] = 3 * row[col[j]] + row2[col[j]];
        out[0] = (3 * cs[1] + cs[0] + 8) >> 4;
        out[1] = (3 * cs[1] + cs[2] + 7) >> 4;
        out[2] = (3 * cs[2] + cs[1] + 8) >> 4;
        out[3] = (3 * cs[2] + cs[3] + 7) >> 4;
    }
}

// Four output pixels per thread (12 bytes, three dword stores when aligned).
// Full-resolution luma with both chroma planes at 1x1, or both triangle-
// filtered 2x1 or 2x2 (the common files), take the direct path; every other
// combination goes through sample_at.
// grid (ceil(max W / 1024), max H, images)
__global__ __launch_bounds__(256) void jpeg_color_kernel(const JImage *__restrict__ imgs,
                                                         const uint8_t *__restrict__ planes,
                                                         uint8_t *__restrict__ rgb) {
    const JImage &im = imgs[blockIdx.z];
    const int x0 = (blockIdx.x * 256 + threadIdx.x) * 4, y = blockIdx.y;
    if (x0 >= im.W || y >= im.H) return;
    const int np = min(4, im.W - x0);
    const JComp &c0 = im.c[0], &c1 = im.c[1], &c2 = im.c[2];
    int Y[4], Cb[4], Cr[4];
    const bool full_y = c0.hf == 1 && c0.vf == 1;
    const bool same_c = c1.hf == c2.hf && c1.vf == c2.vf && c1.dw == c2.dw && c1.dh == c2.dh;
    const bool fancy = same_c && c1.hf == 2 && c1.vf <= 2 && c1.dw > 2;
    if (full_y && (fancy || (same_c && c1.hf == 1 && c1.vf == 1))) {
        const uint32_t yw = *reinterpret_cast<const uint32_t *>(planes + c0.plane + (int64_t)y * c0.pitch + x0);
#pragma unroll
        for (int i = 0; i < 4; ++i) Y[i] = (int)((yw >> (8 * i)) & 0xFF);
        if (!fancy) {
            const uint32_t bw = *reinterpret_cast<const uint32_t *>(planes + c1.plane + (int64_t)y * c1.pitch + x0);
            const uint32_t rw = *reinterpret_cast<const uint32_t *>(planes + c2.plane + (int64_t)y * c2.pitch + x0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                Cb[i] = (int)((bw >> (8 * i)) & 0xFF);
                Cr[i] = (int)((rw >> (8 * i)) & 0xFF);
            }
        } else if (c1.vf == 1) {
            chroma_fancy4<1>(c1, planes, x0, y, Cb);
            chroma_fancy4<1>(c2, planes, x0, y, Cr);
        } else {
            chroma_fancy4<2>(c1, planes, x0, y, Cb);
            chroma_fancy4<2>(c2, planes, x0, y, Cr);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int x = min(x0 + i, im.W - 1);
            Y[i] = sample_at(c0, planes, x, y);
            Cb[i] = sample_at(c1, planes, x, y);
            Cr[i] = sample_at(c2, planes, x, y);
        }
    }
    uint32_t px[3] = {0, 0, 0};  // 12 bytes, little-endian
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int cb = Cb[i] - 128, cr = Cr[i] - 128;
        // libjpeg's tables: FIX(x) = x * 2^16 rounded; R and B rounded by 2^15
        const uint32_t r = clamp255(Y[i] + ((91881 * cr + 32768) >> 16));
        const uint32_t g = clamp255(Y[i] + ((-22554 * cb + 32768 - 46802 * cr) >> 16));
        const uint32_t b = clamp255(Y[i] + ((116130 * cb + 32768) >> 16));
        const int o = 3 * i;  // byte offsets o, o + 1, o + 2 of the 12
        px[o >> 2] |= r << (8 * (o & 3));
        px[(o + 1) >> 2] |= g << (8 * ((o + 1) & 3));
        px[(o + 2) >> 2] |= b << (8 * ((o + 2) & 3));
    }
    uint8_t *dst = rgb + im.rgb + ((int64_t)y * im.W + x0) * 3;
    if (np == 4 && (reinterpret_cast<uintptr_t>(dst) & 3) == 0) {
        uint32_t *d = reinterpret_cast<uint32_t *>(dst);
        d[0] = px[0];
        d[1] = px[1];
        d[2] = px[2];
    } else {
        for (int j = 0; j < 3 * np; ++j) dst[j] = (uint8_t)(px[j >> 2] >> (8 * (j & 3)));
    }
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
    int td[4] = {0, 0, 0, 0}, ta[4] = {0, 0, 0, 0};  // DC / AC table ids per slot
    HuffTab dc[4], ac[4];
    RawHuff rdc[4], rac[4];
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

// Two-level decoding table of one Huffman table, appended to `out` (uint16
// entries): kLook2 first-level entries indexed by the next 9 bits, then the
// second-level subtables.  An entry is (length << 8) | symbol for a code of
// at most 9 bits (first level) or of any length (second level), 0 for no
// code, or, in the first level, 0x8000 | (nb << 12) | off for the 9-bit
// prefix of longer codes: their entry is at second-level index off + the nb
// bits after the prefix (nb = longest such code - 9).  off == 0xFFF: the
// subtable did not fit in `budget` entries; decoded by the canonical walk.
// The table has passed build_huff (no over-subscribed length).
void build_lut2(const RawHuff &r, std::vector<uint16_t> &out, int budget) {
    std::vector<uint16_t> L(512, 0), sub;
    int maxlen[512] = {0};
    struct LongCode {
        int l, code, sym;
    };
    std::vector<LongCode> longc;
    int code = 0, k = 0;
    for (int l = 1; l <= 16; ++l) {
        for (int i = 0; i < r.counts[l - 1]; ++i, ++code, ++k) {
            if (l <= 9) {
                for (int j = 0; j < (1 << (9 - l)); ++j) L[(code << (9 - l)) + j] = (uint16_t)((l << 8) | r.vals[k]);
            } else {
                const int p = code >> (l - 9);
                maxlen[p] = std::max(maxlen[p], l);
                longc.push_back({l, code, r.vals[k]});
            }
        }
        code <<= 1;
    }
    for (int p = 0; p < 512; ++p) {
        if (!maxlen[p]) continue;
        const int nb = maxlen[p] - 9, size = 1 << nb;
        if ((int)sub.size() + size <= budget && (int)sub.size() < 0xFFF) {
            L[p] = (uint16_t)(0x8000 | (nb << 12) | (int)sub.size());
            sub.resize(sub.size() + size, 0);
        } else {
            L[p] = (uint16_t)(0x8000 | 0xFFF);
        }
    }
    for (const LongCode &c : longc) {
        const int p = c.code >> (c.l - 9);
        if ((L[p] & 0xFFF) == 0xFFF) continue;
        const int nb = (L[p] >> 12) & 7, off = L[p] & 0xFFF, extra = c.l - 9;
        const int rest = c.code & ((1 << extra) - 1), span = 1 << (nb - extra);
        for (int j = 0; j < span; ++j) sub[off + (rest << (nb - extra)) + j] = (uint16_t)((c.l << 8) | c.sym);
    }
    out.insert(out.end(), L.begin(), L.end());
    out.insert(out.end(), sub.begin(), sub.end());
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
                for (int j = 0; j < k; ++j)
                    if (sc.comp[j] == ci) return fail(f, PRNU_EINVAL, "a component appears twice in a scan");
                if (!build_huff(dc[td], sc.dc[k]) || !build_huff(ac[ta], sc.ac[k]))
                    return fail(f, PRNU_EINVAL, "invalid Huffman table");
                // DC symbols are difference categories (libjpeg rejects > 15)
                int ndc = 0;
                for (int l = 0; l < 16; ++l) ndc += dc[td].counts[l];
                for (int j = 0; j < ndc; ++j)
                    if (dc[td].vals[j] > 15) return fail(f, PRNU_EINVAL, "invalid Huffman table");
                sc.td[k] = td;
                sc.ta[k] = ta;
                sc.rdc[k] = dc[td];
                sc.rac[k] = ac[ta];
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
    std::vector<uint16_t> lut;  // two-level tables, scan by scan
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
        const size_t seg_first = segs.size(), chunk_first = chunk_seg.size(), tab_first = tabs.size(),
                     lut_first = lut.size();
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
            // the scan's distinct tables, two-level, within kScanTabWords
            {
                if (lut.size() & 1) lut.push_back(0);  // blocks start and end on 4-byte boundaries
                base.tab_off = (int)lut.size();
                int key[8], rel[8], nd = 0;
                for (int q = 0; q < sc.ns; ++q) {
                    int pr[2];
                    for (int cls = 0; cls < 2; ++cls) {
                        const int id = cls * 4 + (cls ? sc.ta[q] : sc.td[q]);
                        int j = 0;
                        while (j < nd && key[j] != id) ++j;
                        if (j == nd) {
                            key[nd] = id;
                            rel[nd] = (int)lut.size() - base.tab_off;
                            ++nd;
                            // level-2 room: what is left after the first levels still to come
                            const int left = kScanTabWords - rel[j] - 512 * (2 * sc.ns - (2 * q + cls));
                            build_lut2(cls ? sc.rac[q] : sc.rdc[q], lut, std::max(0, left));
                        }
                        pr[cls] = rel[j];
                    }
                    base.trel[q] = pr[0] | (pr[1] << 16);
                }
                if (lut.size() & 1) lut.push_back(0);
                base.tab_len = (int)lut.size() - base.tab_off;
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
            // a workgroup decodes chunks of one scan only (its tables in LDS)
            while (chunk_seg.size() % kLanes) chunk_seg.push_back(-1);
        }
        if (f.status != PRNU_OK) {
            segs.resize(seg_first);
            chunk_seg.resize(chunk_first);
            tabs.resize(tab_first);
            lut.resize(lut_first);
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
    const size_t o_lut = o_cseg + align256(sizeof(int) * std::max(1, nchunk));
    const size_t o_data = o_lut + align256(sizeof(uint16_t) * std::max<size_t>(2, lut.size()));
    const size_t staged = o_data + align256((size_t)data_bytes + 8);
    // device only: chunk states, per-chunk sums and prefixes, round flags,
    // coefficients, planes
    const size_t n8 = align256(sizeof(int64_t) * std::max(1, nchunk)), n4 = align256(sizeof(int) * std::max(1, nchunk));
    const size_t o_entry = staged, o_exit0 = o_entry + n8, o_exit1 = o_exit0 + n8;
    const size_t o_started = o_exit1 + n8, o_dcsum = o_started + n4, o_first = o_dcsum + 3 * n4;
    const size_t o_pred = o_first + n4, o_changed = o_pred + 3 * n4;
    const size_t o_coef = o_changed + align256(sizeof(int) * ((size_t)max_chunks + 2));
    const size_t o_plane = o_coef + align256((size_t)coef_blocks * 128);
    const size_t total = o_plane + align256((size_t)plane_bytes);
    PRNU_TRY(dec->staging.ensure(staged));
    PRNU_TRY(dec->dev.ensure(total));
    PRNU_TRY(dec->bad.ensure(sizeof(int) * (size_t)count));
    uint8_t *hs = static_cast<uint8_t *>(dec->staging.p);
    std::memcpy(hs + o_img, imgs.data(), sizeof(JImage) * count);
    if (nseg) std::memcpy(hs + o_seg, segs.data(), sizeof(JSeg) * nseg);
    if (!tabs.empty()) std::memcpy(hs + o_tab, tabs.data(), sizeof(HuffTab) * tabs.size());
    if (!qtab.empty()) std::memcpy(hs + o_q, qtab.data(), sizeof(uint16_t) * qtab.size());
    if (nchunk) std::memcpy(hs + o_cseg, chunk_seg.data(), sizeof(int) * nchunk);
    if (!lut.empty()) std::memcpy(hs + o_lut, lut.data(), sizeof(uint16_t) * lut.size());
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
    PRNU_CHECK_HIP(hipMemcpyAsync(dv, hs, staged, hipMemcpyHostToDevice, st));
    PRNU_CHECK_HIP(hipMemsetAsync(d_bad, 0, sizeof(int) * (size_t)count, st));
    const JImage *d_img = reinterpret_cast<const JImage *>(dv + o_img);
    const JSeg *d_seg = reinterpret_cast<const JSeg *>(dv + o_seg);
    const HuffTab *d_tab = reinterpret_cast<const HuffTab *>(dv + o_tab);
    const int *d_cseg = reinterpret_cast<const int *>(dv + o_cseg);
    const uint16_t *d_lut = reinterpret_cast<const uint16_t *>(dv + o_lut);
    const uint8_t *d_data = dv + o_data;
    int64_t *d_entry = reinterpret_cast<int64_t *>(dv + o_entry);
    int64_t *d_exit[2] = {reinterpret_cast<int64_t *>(dv + o_exit0), reinterpret_cast<int64_t *>(dv + o_exit1)};
    int *d_started = reinterpret_cast<int *>(dv + o_started);
    int *d_dcsum = reinterpret_cast<int *>(dv + o_dcsum);
    int *d_first = reinterpret_cast<int *>(dv + o_first);
    int *d_pred = reinterpret_cast<int *>(dv + o_pred);
    int *d_changed = reinterpret_cast<int *>(dv + o_changed);
    int16_t *d_coef = reinterpret_cast<int16_t *>(dv + o_coef);
    uint8_t *d_plane = dv + o_plane;
    if (nseg) {
        PRNU_CHECK_HIP(hipMemsetAsync(d_coef, 0, (size_t)coef_blocks * 128, st));
        PRNU_CHECK_HIP(hipMemsetAsync(d_changed, 0, sizeof(int) * ((size_t)max_chunks + 2), st));
        const dim3 gch((unsigned)((nchunk + kLanes - 1) / kLanes));
        // linking rounds until one changes nothing; after round r the first
        // r + 1 chunks of every segment are exact, so rounds 0..max_chunks
        // always suffice.  The host looks at the flag once per group of
        // rounds; a round after the fixed point returns at once.
        constexpr int kRoundsPerCheck = 4;
        for (int r = 0; r <= max_chunks;) {
            const int r_end = std::min(max_chunks + 1, r + kRoundsPerCheck);
            for (; r < r_end; ++r) {
                hipLaunchKernelGGL(jpeg_chunk_round_kernel, gch, dim3(kLanes), 0, st, d_seg, d_cseg, nchunk, d_tab,
                                   d_lut, d_data, r, d_entry, d_exit[(r + 1) & 1], d_exit[r & 1], d_started, d_dcsum,
                                   d_changed);
                PRNU_CHECK_LAUNCH();
            }
            int changed = 0;
            PRNU_CHECK_HIP(hipMemcpyAsync(&changed, d_changed + (r - 1), sizeof(int), hipMemcpyDeviceToHost, st));
            PRNU_CHECK_HIP(hipStreamSynchronize(st));
            if (!changed) break;
        }
        hipLaunchKernelGGL(jpeg_chunk_scan_kernel, dim3((unsigned)nseg), dim3(256), 0, st, d_seg, d_started, d_dcsum,
                           d_first, d_pred);
        PRNU_CHECK_LAUNCH();
        hipLaunchKernelGGL(jpeg_chunk_write_kernel, gch, dim3(kLanes), 0, st, d_seg, d_cseg, nchunk, d_img, d_tab,
                           d_lut, d_data, d_entry, d_first, d_pred, d_coef, d_bad);
        PRNU_CHECK_LAUNCH();
        hipLaunchKernelGGL(jpeg_idct_kernel, dim3((unsigned)((max_blocks + 255) / 256), 3, (unsigned)count),
                           dim3(256), 0, st, d_img, reinterpret_cast<const uint16_t *>(dv + o_q), d_coef, d_plane);
        PRNU_CHECK_LAUNCH();
        hipLaunchKernelGGL(jpeg_color_kernel, dim3((unsigned)((max_w + 1023) / 1024), (unsigned)max_h, (unsigned)count),
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

// readJPEG (readjpeg/src/main/cpp/native.cpp:59-134) under its own name and
// signature, so that ReadJPEG.Bindings (CSI/ReadJPEG.java:12-23) can load
// this library in place of libreadjpeg: one file into a host buffer, 0 on
// success, 1 on failure with the reason on stderr, width and height set only
// on success.  It decodes on the calling thread's current device through one
// process-wide decoder and device buffer (calls are serialised) and copies
// the frame back to `output`.
int readJPEG(const char *input, int32_t input_size, int32_t *output_width, int32_t *output_height, char *output,
             int32_t output_capacity) {
    static std::mutex mu;
    static JpegDecoder *dec = nullptr;
    static Workspace frame;
    std::lock_guard<std::mutex> lock(mu);
    if (!input || input_size <= 0 || !output_width || !output_height || !output) {
        std::fprintf(stderr, "error: readJPEG: null pointer or empty input\n");
        return 1;
    }
    FileInfo f;
    if (parse_file(reinterpret_cast<const uint8_t *>(input), input_size, f) != PRNU_OK) {
        std::fprintf(stderr, "error: %s\n", f.err.c_str());
        return 1;
    }
    const int64_t bytes = (int64_t)f.W * f.H * 3;
    if (bytes > output_capacity) {
        std::fprintf(stderr, "error: insufficient buffer space provided for image\n");
        return 1;
    }
    if (!dec) dec = new (std::nothrow) JpegDecoder();
    if (!dec || frame.ensure((size_t)bytes) != PRNU_OK) {
        std::fprintf(stderr, "error: readJPEG: out of memory\n");
        return 1;
    }
    const uint8_t *file = reinterpret_cast<const uint8_t *>(input);
    const int64_t size = input_size;
    int32_t w = 0, h = 0, st = 0;
    if (decode_batch(dec, &file, &size, 1, static_cast<uint8_t *>(frame.ptr), bytes, bytes, &w, &h, &st, nullptr) !=
            PRNU_OK ||
        hipMemcpy(output, frame.ptr, (size_t)bytes, hipMemcpyDeviceToHost) != hipSuccess) {
        (void)hipGetLastError();
        std::fprintf(stderr, "error: %s\n", prnu_last_error());
        return 1;
    }
    *output_width = w;
    *output_height = h;
    return 0;
}

}  // extern "C"
