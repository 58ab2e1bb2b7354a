// PRNU extraction on gfx950: grayscale + FastNoise (one LDS halo-tile
// kernel), zero-mean total, frequency-domain Wiener filter and the R2C
// spectrum of the PCE strategy.  Every kernel takes a batch of images on the
// last grid axis (per-image strides), so a batch of B images is one launch
// per stage.
//
// Reference semantics (paths under CSI/kernels/filter/):
//   grayscalefilter.cu:47-66      out = 0.299*c.z + 0.587*c.y + 0.114*c.x
//   fastnoisefilter.cu:150-204    out = hgrad(dx*s) + vgrad(dy*s), s = 1/(1+|grad|)
//   zeromeantotalfilter.cu        per column then per row: subtract even/odd means
//   wienerfilter.cu               F = FFT2(x); varest = min_f box_f(|F|^2);
//                                 F *= var/max(var, varest); x = Re(IFFT2(F))/n
//   SpectralFilter.java:44-54     R2C FFT2, h x (w/2+1)
//
// Layout in HBM: images and patterns are row-major h x w FP32.  The Wiener
// filter runs its row FFTs in LDS (one row per workgroup), then its column
// FFTs: for columns up to 1024 long, in LDS over strips of >= 4 adjacent
// columns read row by row (no transpose, variance stencil in the natural
// layout); for longer columns through a transpose (w x h) so they are LDS row
// FFTs, the stencil in that transposed domain (the zero-padded box-min is
// transpose invariant), and a transpose back.  It ends with the inverse row
// FFTs.
#include "prnu_internal.hpp"
#include "fft_device.hpp"
#include "fft_host.hpp"

#include <algorithm>

namespace prnu {
namespace {

constexpr float kWR = 0.299f, kWG = 0.587f, kWB = 0.114f;

// ---------------------------------------------------------------- reductions
__device__ inline double wave_sum_d(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
// Deterministic sum over a 256-thread block.
__device__ inline double block_sum_256(double v, double *sh) {
    v = wave_sum_d(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// ---------------------------------------------------------------- grayscale
template <int FMT>
__device__ __forceinline__ float gray_at(const uint8_t *in, int64_t idx) {
    if (FMT == PRNU_INPUT_RGB8) {
        const uint8_t *p = in + 3 * idx;
        // byte 2 (c.z) carries 0.299, byte 0 (c.x) 0.114
        return kWR * (float)p[2] + kWG * (float)p[1] + kWB * (float)p[0];
    } else if (FMT == PRNU_INPUT_GRAY8) {
        const float g = (float)in[idx];
        return kWR * g + kWG * g + kWB * g;
    } else {
        return reinterpret_cast<const float *>(in)[idx];
    }
}

template <int FMT>
__global__ __launch_bounds__(256) void gray_kernel(const uint8_t *in, float *out, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        out[i] = gray_at<FMT>(in, i);
}

// ---------------------------------------------------------------- FastNoise
constexpr int kFnTX = 64, kFnTY = 32;  // output tile

// One-sided at the borders, central inside (fastnoisefilter.cu:150-175).
__device__ __forceinline__ float grad1(float left, float center, float right, int pos, int len) {
    if (pos == 0) return right - center;
    if (pos == len - 1) return center - left;
    return 0.5f * (right - left);
}

// Partial sums of the zero-mean pass for one 64 x 32 output tile held in O:
// per column the sums of even and odd rows, per row the sums of even and odd
// columns (FP64).
__device__ __forceinline__ void zm_tile_sums(const float (*O)[kFnTX + 1], int x0, int y0, int h, int w,
                                             double *colpart, double *rowpart) {
    const int tid = threadIdx.x;
    if (tid < kFnTX) {
        const int gx = x0 + tid;
        double se = 0.0, so = 0.0;
        for (int ly = 0; ly < kFnTY; ++ly) {
            if (((y0 + ly) & 1) == 0) se += O[ly][tid];
            else so += O[ly][tid];
        }
        if (gx < w) {
            colpart[((int64_t)blockIdx.y * w + gx) * 2 + 0] = se;
            colpart[((int64_t)blockIdx.y * w + gx) * 2 + 1] = so;
        }
    } else if (tid >= 64 && tid < 64 + kFnTY) {
        const int ly = tid - 64, gy = y0 + ly;
        double se = 0.0, so = 0.0;
        for (int lx = 0; lx < kFnTX; ++lx) {
            if (((x0 + lx) & 1) == 0) se += O[ly][lx];
            else so += O[ly][lx];
        }
        if (gy < h) {
            rowpart[((int64_t)blockIdx.x * h + gy) * 2 + 0] = se;
            rowpart[((int64_t)blockIdx.x * h + gy) * 2 + 1] = so;
        }
    }
}

// Gray conversion + FastNoise for a 64 x 32 tile with a 2-pixel halo, and
// optionally the zero-mean partial sums of the output tile.  Thread = one
// output column (lane) and the rows wave + 4k, k < 8, all of one parity:
// its column partial is a register sum, and the even- and odd-column sums of
// a row are one 5-step shuffle reduction of the wave (offsets 2..32 keep the
// lane parity apart).  Tiles are numbered image by image, row by row; XCD x
// (workgroup ids go to the XCDs round-robin) takes tiles [x * per, (x+1) *
// per) in order, so neighbouring tiles share halo lines in its L2.
// grid (8 * per)
template <int FMT>
__global__ __launch_bounds__(256) void fastnoise_kernel(const uint8_t *in_all, int64_t in_stride, float *out_all,
                                                        int64_t out_stride, int h, int w, double *colpart_all,
                                                        int64_t cp_stride, double *rowpart_all, int64_t rp_stride,
                                                        int ntx, int nty, int ntiles) {
    __shared__ float G[kFnTY + 4][kFnTX + 4];
    __shared__ float U[kFnTY + 2][kFnTX + 2];
    __shared__ float V[kFnTY + 2][kFnTX + 2];
    __shared__ double csum[4][kFnTX];
    const int per = (ntiles + 7) / 8;
    const int t = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (t >= ntiles) return;
    const int bz = t / (ntx * nty), rem = t - bz * (ntx * nty);
    const int tyi = rem / ntx, txi = rem - (rem / ntx) * ntx;
    const uint8_t *in = in_all + (int64_t)bz * in_stride;
    float *out = out_all + (int64_t)bz * out_stride;
    const int x0 = txi * kFnTX, y0 = tyi * kFnTY;
    const int tid = threadIdx.x;

    for (int idx = tid; idx < (kFnTY + 4) * (kFnTX + 4); idx += 256) {
        const int ly = idx / (kFnTX + 4), lx = idx % (kFnTX + 4);
        const int gy = y0 + ly - 2, gx = x0 + lx - 2;
        float v = 0.f;
        if (gy >= 0 && gy < h && gx >= 0 && gx < w) v = gray_at<FMT>(in, (int64_t)gy * w + gx);
        G[ly][lx] = v;
    }
    __syncthreads();
    for (int idx = tid; idx < (kFnTY + 2) * (kFnTX + 2); idx += 256) {
        const int ly = idx / (kFnTX + 2), lx = idx % (kFnTX + 2);
        const int gy = y0 + ly - 1, gx = x0 + lx - 1;
        float u = 0.f, v = 0.f;
        if (gy >= 0 && gy < h && gx >= 0 && gx < w) {
            const int cy = ly + 1, cx = lx + 1;  // position in G
            const float dx = grad1(G[cy][cx - 1], G[cy][cx], G[cy][cx + 1], gx, w);
            const float dy = grad1(G[cy - 1][cx], G[cy][cx], G[cy + 1][cx], gy, h);
            // hardware square root and reciprocal (about 1 ulp each) in place
            // of the correctly rounded library sequences
            const float norm = __builtin_amdgcn_sqrtf(dx * dx + dy * dy);
            const float scale = __builtin_amdgcn_rcpf(1.0f + norm);
            u = scale * dx;
            v = scale * dy;
        }
        U[ly][lx] = u;
        V[ly][lx] = v;
    }
    __syncthreads();
    // G is free now: it holds the output tile O (32 x 64) for the row sums
    float(*O)[kFnTX + 4] = G;
    const int lane = tid & 63, wave = tid >> 6;
    const int gx = x0 + lane;
    const bool zm = colpart_all != nullptr;
    double col = 0.0;
#pragma unroll
    for (int k = 0; k < kFnTY / 4; ++k) {
        const int ly = wave + 4 * k, gy = y0 + ly;
        float o = 0.f;
        if (gy < h && gx < w) {
            const int cy = ly + 1, cx = lane + 1;  // position in U/V
            const float a = grad1(U[cy][cx - 1], U[cy][cx], U[cy][cx + 1], gx, w);
            const float b = grad1(V[cy - 1][cx], V[cy][cx], V[cy + 1][cx], gy, h);
            o = a + b;
            out[(int64_t)gy * w + gx] = o;
        }
        col += (double)o;
        O[ly][lane] = o;
    }
    if (!zm) return;
    csum[wave][lane] = col;
    __syncthreads();
    if (tid < kFnTX && gx < w) {
        // y0 is even, so waves 0 and 2 hold the even rows
        double *cp = colpart_all + (int64_t)bz * cp_stride + ((int64_t)tyi * w + gx) * 2;
        cp[0] = csum[0][tid] + csum[2][tid];
        cp[1] = csum[1][tid] + csum[3][tid];
    }
    // row r, column parity p: four threads (quad c = 0..3) sum 8 columns each
    const int r = tid >> 3, p = (tid >> 2) & 1, c = tid & 3;
    double rs = 0.0;
#pragma unroll
    for (int i = 0; i < 8; ++i) rs += (double)O[r][p + 16 * c + 2 * i];
    rs += __shfl_xor(rs, 1, 64);
    rs += __shfl_xor(rs, 2, 64);
    const int gy = y0 + r;
    if (c == 0 && gy < h) rowpart_all[(int64_t)bz * rp_stride + ((int64_t)txi * h + gy) * 2 + p] = rs;
}

// Row-walking FastNoise for widths that are multiples of 4 (and 4-byte
// aligned inputs): lane l of a wave holds the 4 pixels of columns
// x0 + 4 (l - 1) + [0, 4) of one row, x0 = 248 * segment; lanes 0 and 63 are
// the halo (their pixels feed lanes 1 and 62 and are not written).  A wave
// walks down a band of 64 rows keeping three gray rows, one U row and three
// V rows in registers: at step t it loads gray row t (one 12-byte load per
// lane for RGB8), forms U and V of row t - 1 and the output of row t - 2.
// Horizontal neighbours come from the adjacent lanes through DPP moves, so
// nothing goes through LDS.  The same arithmetic per pixel as
// fastnoise_kernel (fastnoisefilter.cu:150-204).  Zero-mean partial sums:
// per column and row parity over the band (FP64, bands of 64 rows), per row
// and column parity over each 64-column group of 16 lanes (FP32 sums of the
// group's 32 pixels of a parity, stored as FP64).
// grid (8 * per) x 256 threads: a workgroup = 4 waves = 4 bands of one
// segment; units (image, segment, band group) are dealt in order to the
// XCDs as in fastnoise_kernel.
constexpr int kFnSeg = 248, kFnBand = 64;

__device__ __forceinline__ float dpp_from_left(float v) {   // lane i <- lane i - 1 (0 for lane 0)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
}
__device__ __forceinline__ float dpp_from_right(float v) {  // lane i <- lane i + 1 (0 for lane 63)
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
}
// sum over the 16 lanes of each DPP row, in lane 15 of the row
__device__ __forceinline__ float dpp_row_sum16(float v) {
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x111, 0xf, 0xf, true));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x112, 0xf, 0xf, true));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x114, 0xf, 0xf, true));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x118, 0xf, 0xf, true));
    return v;
}

// Gray values of 4 consecutive pixels starting at pixel `px` (a multiple of
// 4) of an image at `in`.  8-bit images may start at any byte: the pixels'
// bytes are read as the aligned dwords that hold them (each holds at least
// one of the wanted bytes, so no read leaves the pages of the image) and
// realigned with v_alignbyte; `off` = (address of `in`) & 3, the same for
// every group since 3 * px and px are multiples of 4.
template <int FMT>
__device__ __forceinline__ void gray4(const uint8_t *in, int64_t px, int off, float g[4]) {
    if constexpr (FMT == PRNU_INPUT_RGB8) {
        const uint32_t *q = reinterpret_cast<const uint32_t *>(in + 3 * px - off);
        uint32_t a = q[0], b = q[1], c = q[2];
        if (off) {
            const uint32_t d = q[3];
            a = __builtin_amdgcn_alignbyte(b, a, off);
            b = __builtin_amdgcn_alignbyte(c, b, off);
            c = __builtin_amdgcn_alignbyte(d, c, off);
        }
        const uint32_t by[12] = {a & 0xff, (a >> 8) & 0xff, (a >> 16) & 0xff, a >> 24,
                                 b & 0xff, (b >> 8) & 0xff, (b >> 16) & 0xff, b >> 24,
                                 c & 0xff, (c >> 8) & 0xff, (c >> 16) & 0xff, c >> 24};
#pragma unroll
        for (int k = 0; k < 4; ++k)
            g[k] = kWR * (float)by[3 * k + 2] + kWG * (float)by[3 * k + 1] + kWB * (float)by[3 * k];
    } else if constexpr (FMT == PRNU_INPUT_GRAY8) {
        const uint32_t *q = reinterpret_cast<const uint32_t *>(in + px - off);
        uint32_t a = q[0];
        if (off) a = __builtin_amdgcn_alignbyte(q[1], a, off);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float v = (float)((a >> (8 * k)) & 0xff);
            g[k] = kWR * v + kWG * v + kWB * v;
        }
    } else {
        // FP32 images are 4-byte aligned; a 16-byte load needs no more
        typedef float f4a4 __attribute__((ext_vector_type(4), aligned(4)));
        const f4a4 v = *reinterpret_cast<const f4a4 *>(reinterpret_cast<const float *>(in) + px);
        g[0] = v.x;
        g[1] = v.y;
        g[2] = v.z;
        g[3] = v.w;
    }
}

template <int FMT>
__global__ __launch_bounds__(256) void fastnoise_rows_kernel(const uint8_t *in_all, int64_t in_stride, float *out_all,
                                                             int64_t out_stride, int h, int w, double *colpart_all,
                                                             int64_t cp_stride, double *rowpart_all, int64_t rp_stride,
                                                             int nseg, int ngrp, int nunits) {
    const int per = (nunits + 7) / 8;
    const int t = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (t >= nunits) return;
    const int bz = t / (nseg * ngrp), rem = t - bz * (nseg * ngrp);
    const int seg = rem / ngrp, grp = rem - (rem / ngrp) * ngrp;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int band = grp * 4 + wave;
    const int y0 = band * kFnBand;
    if (y0 >= h) return;  // no barriers in this kernel: a wave may leave alone
    const int y1 = min(h, y0 + kFnBand);
    const uint8_t *in = in_all + (int64_t)bz * in_stride;
    const int off = FMT == PRNU_INPUT_FLOAT32 ? 0 : (int)(reinterpret_cast<uintptr_t>(in) & 3);
    float *out = out_all + (int64_t)bz * out_stride;
    const int xl = seg * kFnSeg + 4 * (lane - 1);  // first column of this lane
    const bool inx = xl >= 0 && xl < w;             // all 4 pixels (w is a multiple of 4)
    const bool own = inx && lane >= 1 && lane <= 62;
    const bool first_col = xl == 0, last_col = xl + 4 == w;
    const bool zm = colpart_all != nullptr;

    // gray rows t-2, t-1, t; U row t-2 (output row) and t-1; V rows t-3 .. t-1
    float G0[4] = {}, G1[4] = {}, G2[4] = {};
    float Ua[4] = {}, Ub[4] = {};
    float V0[4] = {}, V1[4] = {}, V2[4] = {};
    double ce[4] = {}, co[4] = {};  // column sums of even / odd rows

    auto dx4 = [&](const float g[4], float d[4]) {
        const float left = dpp_from_left(g[3]), right = dpp_from_right(g[0]);
        d[0] = first_col ? g[1] - g[0] : 0.5f * (g[1] - left);
        d[1] = 0.5f * (g[2] - g[0]);
        d[2] = 0.5f * (g[3] - g[1]);
        d[3] = last_col ? g[3] - g[2] : 0.5f * (right - g[2]);
    };
    for (int r = y0 - 2; r <= y1 + 1; ++r) {
        // shift the windows by one row
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            G0[k] = G1[k];
            G1[k] = G2[k];
            V0[k] = V1[k];
            V1[k] = V2[k];
            Ua[k] = Ub[k];
        }
        // gray row r
        if (inx && r >= 0 && r < h) {
            gray4<FMT>(in, (int64_t)r * w + xl, off, G2);
        } else {
#pragma unroll
            for (int k = 0; k < 4; ++k) G2[k] = 0.f;
        }
        // U, V of row r - 1 from gray rows r - 2, r - 1, r
        // (every lane runs the DPP moves; lanes outside the image keep zeros)
        const int ru = r - 1;
        if (ru >= y0 - 1) {
            const bool valid = inx && ru >= 0 && ru < h;
            float dx[4];
            dx4(G1, dx);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float dy = ru == 0 ? G2[k] - G1[k] : ru == h - 1 ? G1[k] - G0[k] : 0.5f * (G2[k] - G0[k]);
                const float norm = __builtin_amdgcn_sqrtf(dx[k] * dx[k] + dy * dy);
                const float scale = __builtin_amdgcn_rcpf(1.0f + norm);
                Ub[k] = valid ? scale * dx[k] : 0.f;
                V2[k] = valid ? scale * dy : 0.f;
            }
        }
        // output row r - 2 from U row r - 2 and V rows r - 3 .. r - 1
        const int ro = r - 2;
        if (ro >= y0) {
            float a[4];
            dx4(Ua, a);  // the same one-sided / central rule on U
            float o[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const float b = ro == 0 ? V2[k] - V1[k] : ro == h - 1 ? V1[k] - V0[k] : 0.5f * (V2[k] - V0[k]);
                o[k] = own ? a[k] + b : 0.f;
            }
            if (own) *reinterpret_cast<float4 *>(out + (int64_t)ro * w + xl) = make_float4(o[0], o[1], o[2], o[3]);
            if (zm) {
                if (ro & 1) {
#pragma unroll
                    for (int k = 0; k < 4; ++k) co[k] += (double)o[k];
                } else {
#pragma unroll
                    for (int k = 0; k < 4; ++k) ce[k] += (double)o[k];
                }
                // columns xl + k have parity k & 1 (xl is a multiple of 4)
                const float se = dpp_row_sum16(o[0] + o[2]);
                const float so = dpp_row_sum16(o[1] + o[3]);
                if ((lane & 15) == 15) {
                    double *rp = rowpart_all + (int64_t)bz * rp_stride +
                                 ((int64_t)(seg * 4 + (lane >> 4)) * h + ro) * 2;
                    rp[0] = (double)se;
                    rp[1] = (double)so;
                }
            }
        }
    }
    if (zm && own) {
        double *cp = colpart_all + (int64_t)bz * cp_stride + ((int64_t)band * w + xl) * 2;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cp[2 * k + 0] = ce[k];
            cp[2 * k + 1] = co[k];
        }
    }
}

// Zero-mean partial-sum tiles of the FastNoise kernels: colpart holds one
// row of column sums per tile row, rowpart one column of row sums per tile
// column.
constexpr int kZmTileRows = kFnTY, kZmTileCols = kFnTX;             // fastnoise_kernel, zm_partials_kernel
int zm_tiles_y_rows(int h) { return (h + kFnBand - 1) / kFnBand; }  // fastnoise_rows_kernel
int zm_tiles_x_rows(int w) { return 4 * ((w + kFnSeg - 1) / kFnSeg); }

// -------------------------------------------------------------- zero-mean
// Partial sums of an arbitrary image in the fastnoise tiling.  grid (tx, ty, B)
__global__ __launch_bounds__(256) void zm_partials_kernel(const float *img_all, int64_t img_stride, int h, int w,
                                                          double *colpart_all, int64_t cp_stride,
                                                          double *rowpart_all, int64_t rp_stride) {
    __shared__ float O[kFnTY][kFnTX + 1];
    const float *img = img_all + (int64_t)blockIdx.z * img_stride;
    const int x0 = blockIdx.x * kFnTX, y0 = blockIdx.y * kFnTY;
    for (int idx = threadIdx.x; idx < kFnTY * kFnTX; idx += 256) {
        const int ly = idx / kFnTX, lx = idx % kFnTX;
        const int gy = y0 + ly, gx = x0 + lx;
        O[ly][lx] = (gy < h && gx < w) ? img[(int64_t)gy * w + gx] : 0.f;
    }
    __syncthreads();
    zm_tile_sums(O, x0, y0, h, w, colpart_all + (int64_t)blockIdx.z * cp_stride,
                 rowpart_all + (int64_t)blockIdx.z * rp_stride);
}

// colsum[p][x] = sum over tile rows; rowmean[p][y] = row sums / column counts.
// grid (ceil(max(h,w)/256), B)
__global__ __launch_bounds__(256) void zm_reduce_kernel(const double *colpart_all, int64_t cp_stride,
                                                        const double *rowpart_all, int64_t rp_stride, int h, int w,
                                                        int tiles_y, int tiles_x, double *colsum_all,
                                                        double *rowmean_all, float *rz_all) {
    const int b = blockIdx.y;
    const double *colpart = colpart_all + (int64_t)b * cp_stride;
    const double *rowpart = rowpart_all + (int64_t)b * rp_stride;
    double *colsum = colsum_all + (int64_t)b * 2 * w;
    double *rowmean = rowmean_all + (int64_t)b * 2 * h;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < w) {
        double se = 0.0, so = 0.0;
        for (int t = 0; t < tiles_y; ++t) {
            se += colpart[((int64_t)t * w + i) * 2 + 0];
            so += colpart[((int64_t)t * w + i) * 2 + 1];
        }
        colsum[i] = se;
        colsum[w + i] = so;
    }
    if (i < h) {
        double se = 0.0, so = 0.0;
        for (int t = 0; t < tiles_x; ++t) {
            se += rowpart[((int64_t)t * h + i) * 2 + 0];
            so += rowpart[((int64_t)t * h + i) * 2 + 1];
        }
        rowmean[i] = se / (double)((w + 1) >> 1);
        rowmean[h + i] = so / (double)(w >> 1);
        if (rz_all) {
            float *rz = rz_all + (int64_t)b * 2 * h;
            rz[i] = (float)rowmean[i];
            rz[h + i] = (float)rowmean[h + i];
        }
    }
}

// g[pi*2+pj] = mean of the (row parity pi, column parity pj) sub-lattice; also
// turns colsum into column means in place, and (cz_all non-null) writes the
// FP32 column terms cz[pi][x] = colmean[pi][x] - g[pi][x & 1] of the
// Wiener row kernel.  grid (B), block 256.
__global__ __launch_bounds__(256) void zm_lattice_kernel(double *colsum_all, int h, int w, double *g_all,
                                                         float *cz_all) {
    __shared__ double sh[4];
    double *colsum = colsum_all + (int64_t)blockIdx.x * 2 * w;
    double *g = g_all + (int64_t)blockIdx.x * 4;
    double s[4] = {0, 0, 0, 0};
    for (int x = threadIdx.x; x < w; x += 256) {
        s[0 + (x & 1)] += colsum[x];
        s[2 + (x & 1)] += colsum[w + x];
    }
    double tot[4];
    for (int q = 0; q < 4; ++q) tot[q] = block_sum_256(s[q], sh);
    const double rows_e = (double)((h + 1) >> 1), rows_o = (double)(h >> 1);
    const double cols_e = (double)((w + 1) >> 1), cols_o = (double)(w >> 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        g[0] = tot[0] / (rows_e * cols_e);
        g[1] = tot[1] / (rows_e * cols_o);
        g[2] = tot[2] / (rows_o * cols_e);
        g[3] = tot[3] / (rows_o * cols_o);
    }
    const double gl[4] = {tot[0] / (rows_e * cols_e), tot[1] / (rows_e * cols_o), tot[2] / (rows_o * cols_e),
                          tot[3] / (rows_o * cols_o)};
    float *cz = cz_all ? cz_all + (int64_t)blockIdx.x * 2 * w : nullptr;
    for (int x = threadIdx.x; x < w; x += 256) {
        colsum[x] /= rows_e;
        colsum[w + x] /= rows_o;
        if (cz) {
            cz[x] = (float)(colsum[x] - gl[x & 1]);
            cz[w + x] = (float)(colsum[w + x] - gl[2 + (x & 1)]);
        }
    }
}

// out = x - colmean[pi][x] - rowmean[pj][y] + g[pi][pj]: columns-then-rows
// zero-mean of zeromeantotalfilter.cu in closed form.
__device__ __forceinline__ float zm_value(float x, int y, int xx, int h, int w, const double *colmean,
                                          const double *rowmean, const double *g) {
    const int pi = y & 1, pj = xx & 1;
    return (float)((double)x - colmean[pi * w + xx] - rowmean[pj * h + y] + g[pi * 2 + pj]);
}

// grid (blocks, B)
__global__ __launch_bounds__(256) void zm_apply_kernel(float *img_all, int64_t img_stride, int h, int w,
                                                       const double *colmean_all, const double *rowmean_all,
                                                       const double *g_all) {
    const int b = blockIdx.y;
    float *img = img_all + (int64_t)b * img_stride;
    const double *colmean = colmean_all + (int64_t)b * 2 * w;
    const double *rowmean = rowmean_all + (int64_t)b * 2 * h;
    const double *g = g_all + (int64_t)b * 4;
    const int64_t n = (int64_t)h * w;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const int y = (int)(i / w), x = (int)(i - (int64_t)y * w);
        img[i] = zm_value(img[i], y, x, h, w, colmean, rowmean, g);
    }
}

// ------------------------------------------------------------------ Wiener
// The input is real, so its 2-D spectrum F is Hermitian and the filter runs
// on half spectra (h x (w/2+1)): Re IFFT2(F S) = IFFT2(F S'), with
// S'(k) = (S(k) + S(-k)) / 2, and F S' is Hermitian again, so one complex
// inverse transform over the w/2+1 columns and a real inverse over the rows
// give the reference's real part.  Rows go two at a time through one complex
// transform z = a + i b of length w: A_k = (Z_k + conj Z_{w-k}) / 2 and
// B_k = (Z_k - conj Z_{w-k}) / 2i.

// Rows 2r and 2r+1 of image b: (optional zero-mean) -> sums of squares
// (FP64) -> forward FFT (w) -> half spectra into A (h x (w/2+1)).
// grid (ceil(h/2), B).
// Zero-mean (cz non-null): x - cz[row parity][j] - rz[j & 1][row] in FP32,
// the column terms cz = colmean - g and row means rz of zm_lattice_kernel /
// zm_reduce_kernel (within an ulp of the FP64 closed form zm_value).
// PAD2048: rows of 2048 by fft_block_2048_r8 (LDS kFft2048PadLds values).
template <int MAXN, bool PAD2048 = false>
__global__ __launch_bounds__(kFftThreads) void wiener_rows_fwd_kernel(FftDesc d, const float *img_all,
                                                                      int64_t img_stride, int h, float2 *A_all,
                                                                      int64_t plane, double *rowsq_all,
                                                                      const float *cz_all, const float *rz_all,
                                                                      int lda) {
    extern __shared__ float2 buf[];
    __shared__ double sh[8];
    const int i0 = 2 * blockIdx.x, i1 = i0 + 1, b = blockIdx.y;
    const bool two = i1 < h;
    const int w = d.n, wh = w / 2 + 1;
    const float *img = img_all + (int64_t)b * img_stride;
    const bool zm = cz_all != nullptr;
    // rows i0 (even) and i1 (odd): column terms of parities 0 and 1
    const float *cz0 = zm ? cz_all + (int64_t)b * 2 * w : nullptr;
    const float *cz1 = zm ? cz0 + w : nullptr;
    const float *rz = zm ? rz_all + (int64_t)b * 2 * h : nullptr;
    const float ra0 = zm ? rz[i0] : 0.f, ra1 = zm ? rz[h + i0] : 0.f;   // row i0, even / odd columns
    const float rb0 = zm && two ? rz[i1] : 0.f, rb1 = zm && two ? rz[h + i1] : 0.f;
    double sq0 = 0.0, sq1 = 0.0;
    const float *r0 = img + (int64_t)i0 * w;
    const float *r1 = img + (int64_t)(two ? i1 : i0) * w;
    auto get = [&](int j) {
            float va = r0[j], vb = two ? r1[j] : 0.f;
            if (zm) {
                const bool odd = j & 1;
                va = (va - cz0[j]) - (odd ? ra1 : ra0);
                vb = two ? (vb - cz1[j]) - (odd ? rb1 : rb0) : 0.f;
            }
            sq0 += (double)va * (double)va;
            sq1 += (double)vb * (double)vb;
            return make_float2(va, vb);
        };
    auto put = [&](int j, float2 f) { buf[j] = f; };  // the last pass has read buf before it writes
    if constexpr (PAD2048)
        fft_block_2048_r8<-1>(buf, d.tw, get, put);
    else
        fft_row<-1, MAXN>(buf, d, get, put);
    float2 *a0 = A_all + (int64_t)b * plane + (int64_t)i0 * lda;  // rows lda values apart
    for (int k = threadIdx.x; k < wh; k += kFftThreads) {
        const float2 zk = buf[k], zm_ = buf[k == 0 ? 0 : w - k];
        a0[k] = make_float2(0.5f * (zk.x + zm_.x), 0.5f * (zk.y - zm_.y));
        if (two) a0[lda + k] = make_float2(0.5f * (zk.y + zm_.y), -0.5f * (zk.x - zm_.x));
    }
    // both row sums in one fixed-order block reduction
    sq0 = wave_sum_d(sq0);
    sq1 = wave_sum_d(sq1);
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6] = sq0;
        sh[4 + (threadIdx.x >> 6)] = sq1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        rowsq_all[(int64_t)b * h + i0] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
        if (two) rowsq_all[(int64_t)b * h + i1] = (sh[4] + sh[5]) + (sh[6] + sh[7]);
    }
}

// var = n/(n-1) * sum(x^2) over the row partials, FP64, fixed order.  grid (B)
__global__ __launch_bounds__(256) void wiener_var_kernel(const double *rowsq_all, int h, int64_t n, double *var_all) {
    __shared__ double sh[4];
    const double *rowsq = rowsq_all + (int64_t)blockIdx.x * h;
    double s = 0.0;
    for (int i = threadIdx.x; i < h; i += 256) s += rowsq[i];
    s = block_sum_256(s, sh);
    if (threadIdx.x == 0) var_all[blockIdx.x] = s * (double)n / (double)(n - 1);
}

// Variance estimate: min over f in {3,5,7,9} of the zero-padded f x f mean of
// P (wienerfilter.cu:164-247), then S = var / max(var, varest).
// A workgroup takes a 64-column x 128-row output tile with its 4-pixel halo
// in LDS (136 x 72 floats); wave k owns output rows [32 k, 32 k + 32) and a
// lane one column.  The lane walks its 40 LDS rows once (fully unrolled):
// per row it forms the horizontal 3/5/7/9 sums around its column (9 LDS
// reads), and every output row sums the horizontal sums of the rows its box
// covers, so a row's horizontal sums are formed once for 32 outputs instead
// of once per 8.  All sums add nonnegative terms directly (no running
// differences, which would carry the rounding error of a large bin past its
// window).  Tiles are numbered image by image, row by row, and dealt to the
// XCDs in ranges as in fastnoise_kernel.  grid (8 * per)
// P holds half of the power spectrum, which is symmetric, P(-k) = P(k):
// MIRROR_ROWS, rows [0, rows/2] of length cols (full row r > rows/2 is row
// rows - r reversed: P(r, c) = P(rows - r, (cols - c) % cols)); otherwise
// all rows with columns [0, cols/2] (P(r, c) = P((rows - r) % rows, cols - c)
// for c > cols/2).  S is written on the full rows x cols grid.
constexpr int kVX = 64, kVY = 128, kVB = 4, kVRows = 32;
template <bool MIRROR_ROWS>
__device__ __forceinline__ float power_full(const float *P, int r, int c, int rows, int cols) {
    if (MIRROR_ROWS) {
        if (r <= rows / 2) return P[(int64_t)r * cols + c];
        return P[(int64_t)(rows - r) * cols + (c == 0 ? 0 : cols - c)];
    }
    const int ch = cols / 2 + 1;
    if (c <= cols / 2) return P[(int64_t)r * ch + c];
    return P[(int64_t)(r == 0 ? 0 : rows - r) * ch + (cols - c)];
}

template <bool MIRROR_ROWS>
__global__ __launch_bounds__(256) void varest_scale_kernel(const float *P_all, float *S_all, int rows, int cols,
                                                           const double *var_all, int ntx, int nty, int ntiles) {
    constexpr int LW = kVX + 2 * kVB, LH = kVY + 2 * kVB;
    __shared__ __align__(16) float T[LH][LW];
    const int per = (ntiles + 7) / 8;
    const int t = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
    if (t >= ntiles) return;
    const int b = t / (ntx * nty), rem = t - b * (ntx * nty);
    const int r0 = (rem / ntx) * kVY, c0 = (rem - (rem / ntx) * ntx) * kVX;
    const int64_t plane = (int64_t)rows * cols;
    const float *P = P_all + (int64_t)b * plane;
    float *S = S_all + (int64_t)b * plane;
    // thread -> (LDS row, group of 4 columns): 18 groups per row.  In the
    // transposed layout (MIRROR_ROWS) a group inside the grid is 4
    // consecutive values of one source row -- forward (one 16-byte load when
    // cols is a multiple of 4: gx0 is) or backward in the mirrored half --
    // so only groups at the grid's edges take the per-value path.
    const bool vec4 = MIRROR_ROWS && (cols & 3) == 0;
    for (int q = threadIdx.x; q < LH * (LW / 4); q += 256) {
        const int ly = q / (LW / 4), lx = (q - ly * (LW / 4)) * 4;
        const int gy = r0 + ly - kVB;
        const int gx0 = c0 + lx - kVB;
        const bool row_in = gy >= 0 && gy < rows;
        float4 val = make_float4(0.f, 0.f, 0.f, 0.f);
        if (vec4 && row_in && gx0 >= 0 && gx0 + 3 < cols) {
            if (gy <= rows / 2) {
                val = *reinterpret_cast<const float4 *>(P + (int64_t)gy * cols + gx0);
            } else {
                const float *pr = P + (int64_t)(rows - gy) * cols;
                val.x = pr[gx0 == 0 ? 0 : cols - gx0];
                val.y = pr[cols - gx0 - 1];
                val.z = pr[cols - gx0 - 2];
                val.w = pr[cols - gx0 - 3];
            }
        } else if (row_in) {
            float e[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int gx = gx0 + k;
                e[k] = (gx >= 0 && gx < cols) ? power_full<MIRROR_ROWS>(P, gy, gx, rows, cols) : 0.f;
            }
            val = make_float4(e[0], e[1], e[2], e[3]);
        }
        *reinterpret_cast<float4 *>(&T[ly][lx]) = val;
    }
    __syncthreads();
    const int c = threadIdx.x & 63, oy0 = (threadIdx.x >> 6) * kVRows;
    const float v = (float)var_all[b];
    const int gx = c0 + c;
    // LDS rows oy0 .. oy0 + 39; output row j sits at LDS row oy0 + j + 4
    float h3[kVRows + 8], h5[kVRows + 8], h7[kVRows + 8], h9[kVRows + 8];
#pragma unroll
    for (int i = 0; i < kVRows + 8; ++i) {
        const float *row = &T[oy0 + i][c + kVB];
        h3[i] = row[-1] + row[0] + row[1];
        h5[i] = h3[i] + row[-2] + row[2];
        h7[i] = h5[i] + row[-3] + row[3];
        h9[i] = h7[i] + row[-4] + row[4];
        const int j = i - 8;  // the last row output j needs
        if (j >= 0) {
            // each sum starts from its first term: the same value as 0 + first
            // term, since the terms are sums of powers (never -0), one add fewer
            float s3 = h3[j + 3], s5 = h5[j + 2], s7 = h7[j + 1], s9 = h9[j];
#pragma unroll
            for (int d = 0; d <= 1; ++d) s3 += h3[j + 4 + d];
#pragma unroll
            for (int d = -1; d <= 2; ++d) s5 += h5[j + 4 + d];
#pragma unroll
            for (int d = -2; d <= 3; ++d) s7 += h7[j + 4 + d];
#pragma unroll
            for (int d = -3; d <= 4; ++d) s9 += h9[j + 4 + d];
            // means by the rounded reciprocals of f^2, the quotient by the
            // hardware reciprocal (each within an ulp or so)
            float best = s3 * (1.f / 9.f);
            best = fminf(best, s5 * (1.f / 25.f));
            best = fminf(best, s7 * (1.f / 49.f));
            best = fminf(best, s9 * (1.f / 81.f));
            const int gy = r0 + oy0 + j;
            if (gy < rows && gx < cols) S[(int64_t)gy * cols + gx] = v * __builtin_amdgcn_rcpf(fmaxf(v, best));
        }
    }
}

int launch_varest(const float *P, float *S, int rows, int cols, bool mirror_rows, int nb, const double *var,
                  hipStream_t st) {
    const int ntx = (cols + kVX - 1) / kVX, nty = (rows + kVY - 1) / kVY;
    const int64_t ntiles = (int64_t)ntx * nty * nb;
    PRNU_REQUIRE(ntiles < (int64_t(1) << 30), "wiener: %lld variance tiles", (long long)ntiles);
    const int per = (int)((ntiles + 7) / 8);
    if (mirror_rows)
        hipLaunchKernelGGL(varest_scale_kernel<true>, dim3((unsigned)(8 * per)), dim3(256), 0, st, P, S, rows, cols,
                           var, ntx, nty, (int)ntiles);
    else
        hipLaunchKernelGGL(varest_scale_kernel<false>, dim3((unsigned)(8 * per)), dim3(256), 0, st, P, S, rows, cols,
                           var, ntx, nty, (int)ntiles);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

// Rows 2r and 2r+1 from their half spectra in A (h x (w/2+1)): one inverse
// transform of Z_k = A_k + i B_k (upper half from A_{w-k} = conj A_k; the
// imaginary parts of bins 0 and w/2 dropped, as a real transform does);
// out = z / (w*h), row 2r the real, row 2r+1 the imaginary part
// (normalizeToReal).  grid (ceil(h/2), B)
// BLOCKN = 2048: rows of 2048 by the two-wave transform (fft_block_2048, 128
// threads, LDS fft_wave_lds<2048>() values); 0: fft_row on kFftThreads.
template <int MAXN, int BLOCKN = 0>
__global__ __launch_bounds__(BLOCKN ? 128 : kFftThreads) void rows_c2r_kernel(FftDesc d, const float2 *A,
                                                                              int64_t plane, int h, float *out_all,
                                                                              int64_t out_stride, float inv_n,
                                                                              int lda) {
    extern __shared__ float2 buf[];
    const int i0 = 2 * blockIdx.x, i1 = i0 + 1;
    const bool two = i1 < h;
    const int w = d.n;
    const float2 *a0 = A + (int64_t)blockIdx.y * plane + (int64_t)i0 * lda;  // rows lda values apart
    float *out = out_all + (int64_t)blockIdx.y * out_stride + (int64_t)i0 * w;
    auto get = [&](int k) {
        const bool up = k > w / 2;
        const int kk = up ? w - k : k;
        float2 x = a0[kk], y = two ? a0[lda + kk] : make_float2(0.f, 0.f);
        if (kk == 0 || 2 * kk == w) {
            x.y = 0.f;
            y.y = 0.f;
        }
        if (up) {
            x.y = -x.y;
            y.y = -y.y;
        }
        return make_float2(x.x - y.y, x.y + y.x);
    };
    auto put = [&](int j, float2 z) {
        out[j] = z.x * inv_n;
        if (two) out[w + j] = z.y * inv_n;
    };
    if constexpr (BLOCKN == 2048)
        fft_block_2048<1>(buf, d.tw, get, put);
    else
        fft_row<1, MAXN>(buf, d, get, put);
}

// ---------------------------------------------------------------- spectrum
// Forward FFTs (length w) of real rows 2r and 2r+1 as one complex transform
// (z = a + i b, split as in wiener_rows_fwd_kernel); the w/2+1 bins of each
// row go to A (h x wh, rows wh apart).  grid (ceil(h/2), B)
// PAD2048 as wiener_rows_fwd_kernel's.
template <int MAXN, bool PAD2048 = false>
__global__ __launch_bounds__(kFftThreads) void rows_fft_half_kernel(FftDesc d, const float *img_all,
                                                                    int64_t img_stride, float2 *A, int64_t plane,
                                                                    int h, int lda) {
    extern __shared__ float2 buf[];
    const int i0 = 2 * blockIdx.x, i1 = i0 + 1;
    const bool two = i1 < h;
    const int w = d.n, wh = w / 2 + 1;
    const float *r0 = img_all + (int64_t)blockIdx.y * img_stride + (int64_t)i0 * w;
    const float *r1 = two ? r0 + w : r0;
    auto get = [&](int j) { return make_float2(r0[j], two ? r1[j] : 0.f); };
    auto put = [&](int j, float2 f) { buf[j] = f; };  // the last pass has read buf before it writes
    if constexpr (PAD2048)
        fft_block_2048_r8<-1>(buf, d.tw, get, put);
    else
        fft_row<-1, MAXN>(buf, d, get, put);
    float2 *a0 = A + (int64_t)blockIdx.y * plane + (int64_t)i0 * lda;  // rows lda values apart
    for (int k = threadIdx.x; k < wh; k += kFftThreads) {
        const float2 zk = buf[k], zm_ = buf[k == 0 ? 0 : w - k];
        a0[k] = make_float2(0.5f * (zk.x + zm_.x), 0.5f * (zk.y - zm_.y));
        if (two) a0[lda + k] = make_float2(0.5f * (zk.y + zm_.y), -0.5f * (zk.x - zm_.x));
    }
}

// ------------------------------------------- transposed-layout row kernels
// Used when a column strip would be narrower than 4 columns (long columns):
// the plane is transposed and columns become LDS row transforms.
// In-place forward FFT of the rows of B (length d.n) and |F|^2 into P.
// grid (rows, batch); planes `plane` elements apart.
template <int MAXN>
__global__ __launch_bounds__(kFftThreads) void rows_fft_power_kernel(FftDesc d, float2 *B, float *P, int64_t plane) {
    extern __shared__ float2 buf[];
    const int r = blockIdx.x;
    const int n = d.n;
    float2 *row = B + (int64_t)blockIdx.y * plane + (int64_t)r * n;
    float *prow = P + (int64_t)blockIdx.y * plane + (int64_t)r * n;
    fft_row<-1, MAXN>(
        buf, d, [&](int j) { return row[j]; },
        [&](int j, float2 f) {
            row[j] = f;
            prow[j] = f.x * f.x + f.y * f.y;
        });
}

// Row r of B (r <= R/2, length d.n = C) times S'(r, c) = (S(r, c) +
// S((R - r) % R, (C - c) % C)) / 2 from the full R x C scale S, then the
// inverse FFT in place.  grid (R/2 + 1, B); planes `plane` apart.
template <int MAXN>
__global__ __launch_bounds__(kFftThreads) void rows_ifft_scaled_kernel(FftDesc d, float2 *B, const float *S,
                                                                       int64_t plane, int R) {
    extern __shared__ float2 buf[];
    const int r = blockIdx.x;
    const int n = d.n;
    float2 *row = B + (int64_t)blockIdx.y * plane + (int64_t)r * n;
    const float *srow = S + (int64_t)blockIdx.y * plane + (int64_t)r * n;
    const float *mrow = S + (int64_t)blockIdx.y * plane + (int64_t)(r == 0 ? 0 : R - r) * n;
    fft_row<1, MAXN>(
        buf, d, [&](int j) { return cscale(row[j], 0.5f * (srow[j] + mrow[j == 0 ? 0 : n - j])); },
        [&](int j, float2 f) { row[j] = f; });
}

// grid (rows, B)
// rows_fft_power_kernel for columns of N = 1024 or 1536 points, reading the
// row-major half spectra A (h x wh per frame) itself, so no transpose
// precedes it: each column by one wave (fft_wave: every lane works in every
// pass).  A workgroup of 8 waves
// stages a strip of 8 adjacent columns into LDS (64-byte row pieces, each
// column in its own padded buffer), then wave c transforms column c there and
// writes B, P in the transposed layout as before.  Workgroup g takes strip
// (g mod 8) * per + g / 8 (neighbouring strips on one XCD); the columns of
// the last strip past wh are neither loaded nor stored.  grid (8 per, frames).
template <int N>
__global__ __launch_bounds__(512) void cols_fft_power_strip_kernel(const float2 *__restrict__ tw,
                                                                  const float2 *__restrict__ A, int wh, float2 *B,
                                                                  float *P, int64_t plane, int per, int lda) {
    constexpr int SH = FftWavePlan<N>::kPadShift;
    constexpr int SB = fft_wave_lds<N>() + 4;  // +4 spreads the staging writes of the 8 buffers over banks
    __shared__ __attribute__((aligned(16))) float2 lds[8 * SB];
    const int g = blockIdx.x;
    const int strip = (g & 7) * per + (g >> 3);
    if (strip >= (wh + 7) / 8) return;
    const int j0 = strip * 8;
    const float2 *Af = A + (int64_t)blockIdx.y * plane;
    const int t = (int)threadIdx.x;
    {
        // two adjacent columns per thread, one 16-byte load when lda is even
        // (j0 is, and the plane is 16-byte aligned); the second column of a
        // strip's last pair may lie past wh, in the row's pitch or the next
        // row, inside the plane, and is never stored
        const int c2 = (t & 3) * 2, r = t >> 2;
        if (j0 + c2 < wh) {
#pragma unroll 4
            for (int k = 0; k < N / 128; ++k) {
                const int i = r + 128 * k;
                const float2 *src = Af + (int64_t)i * lda + j0 + c2;
                float4 v;
                if ((lda & 1) == 0) {  // 16-byte aligned pairs
                    v = *reinterpret_cast<const float4 *>(src);
                } else {  // an odd (unpadded) pitch: two 8-byte loads
                    const float2 a = src[0], b = src[1];
                    v = make_float4(a.x, a.y, b.x, b.y);
                }
                lds[c2 * SB + fft_wave_pad<SH>(i)] = make_float2(v.x, v.y);
                lds[(c2 + 1) * SB + fft_wave_pad<SH>(i)] = make_float2(v.z, v.w);
            }
        }
    }
    __syncthreads();
    const int wave = t >> 6, col = j0 + wave;
    float2 *buf = lds + wave * SB;
    const bool store = col < wh;
    float2 *row = B + (int64_t)blockIdx.y * plane + (int64_t)col * N;
    float *prow = P + (int64_t)blockIdx.y * plane + (int64_t)col * N;
    fft_wave<-1, N, true, true>(  // after the staging barrier each wave works in its own buffer
        buf, tw, [&](int i) { return buf[fft_wave_pad<SH>(i)]; },
        [&](int j, float2 f) {
            if (store) {
                row[j] = f;
                prow[j] = f.x * f.x + f.y * f.y;
            }
        });
}

// The spectrum's column FFTs in the row-major layout, without transposes:
// strips of 8 columns of A (h x wh per frame) staged into LDS as in
// cols_fft_power_strip_kernel, transformed there in place by one wave per
// column, and written to out (h x wh per frame, frames out_stride apart) as
// 64-byte row pieces.  grid (8 per, frames).
template <int N>
__global__ __launch_bounds__(512) void spectrum_cols_strip_kernel(const float2 *__restrict__ tw,
                                                            const float2 *__restrict__ A, int64_t plane, int wh,
                                                            float2 *__restrict__ out, int64_t out_stride, int per,
                                                            int lda) {
    constexpr int SH = FftWavePlan<N>::kPadShift;
    constexpr int SB = fft_wave_lds<N>() + 4;
    __shared__ __attribute__((aligned(16))) float2 lds[8 * SB];
    const int g = blockIdx.x;
    const int strip = (g & 7) * per + (g >> 3);
    if (strip >= (wh + 7) / 8) return;
    const int j0 = strip * 8;
    const float2 *Af = A + (int64_t)blockIdx.y * plane;
    const int t = (int)threadIdx.x, c = t & 7, r = t >> 3;
    const bool cin = j0 + c < wh;
    {
        // two adjacent columns per thread, one 16-byte load (as in
        // cols_fft_power_strip_kernel)
        const int c2 = (t & 3) * 2, r2 = t >> 2;
        if (j0 + c2 < wh) {
#pragma unroll 4
            for (int k = 0; k < N / 128; ++k) {
                const int i = r2 + 128 * k;
                const float2 *src = Af + (int64_t)i * lda + j0 + c2;
                float4 v;
                if ((lda & 1) == 0) {  // 16-byte aligned pairs
                    v = *reinterpret_cast<const float4 *>(src);
                } else {  // an odd (unpadded) pitch: two 8-byte loads
                    const float2 a = src[0], b = src[1];
                    v = make_float4(a.x, a.y, b.x, b.y);
                }
                lds[c2 * SB + fft_wave_pad<SH>(i)] = make_float2(v.x, v.y);
                lds[(c2 + 1) * SB + fft_wave_pad<SH>(i)] = make_float2(v.z, v.w);
            }
        }
    }
    __syncthreads();
    float2 *buf = lds + (t >> 6) * SB;
    // the last pass has read buf before it writes (a barrier between); each
    // wave works in its own buffer until the store of the whole strip
    fft_wave<-1, N, true, true>(
        buf, tw, [&](int i) { return buf[fft_wave_pad<SH>(i)]; },
        [&](int j, float2 f) { buf[fft_wave_pad<SH>(j)] = f; });
    __syncthreads();
    float2 *of = out + (int64_t)blockIdx.y * out_stride;
    if (cin) {
#pragma unroll 4
        for (int k = 0; k < N / 64; ++k) {
            const int i = r + 64 * k;
            of[(int64_t)i * wh + j0 + c] = lds[c * SB + fft_wave_pad<SH>(i)];
        }
    }
}

// rows_ifft_scaled_kernel for N = 1024 or 1536 writing the row-major plane A
// (h x wh per frame) itself, so no transpose follows it: wave c of a
// workgroup transforms transposed row col = j0 + c of B (scaled as there) by
// one-wave transforms into its LDS
// buffer, then the workgroup writes the strip's 8 columns as 64-byte row
// pieces.  Strip mapping as cols_fft_power_strip_kernel; R as there.
template <int N>
__global__ __launch_bounds__(512) void cols_ifft_scaled_strip_kernel(const float2 *__restrict__ tw,
                                                                    const float2 *__restrict__ B,
                                                                    const float *__restrict__ S, int64_t plane,
                                                                    int R, int wh, float2 *__restrict__ A, int per,
                                                                    int lda) {
    constexpr int SH = FftWavePlan<N>::kPadShift;
    constexpr int SB = fft_wave_lds<N>() + 4;
    __shared__ __attribute__((aligned(16))) float2 lds[8 * SB];
    const int g = blockIdx.x;
    const int strip = (g & 7) * per + (g >> 3);
    if (strip >= (wh + 7) / 8) return;
    const int j0 = strip * 8;
    const int t = (int)threadIdx.x, wave = t >> 6;
    const int col = min(j0 + wave, wh - 1);  // a column past wh transforms the last one again and stores nothing
    float2 *buf = lds + wave * SB;
    const float2 *row = B + (int64_t)blockIdx.y * plane + (int64_t)col * N;
    const float *srow = S + (int64_t)blockIdx.y * plane + (int64_t)col * N;
    const float *mrow = S + (int64_t)blockIdx.y * plane + (int64_t)(col == 0 ? 0 : R - col) * N;
    fft_wave<1, N, false, true>(  // each wave in its own buffer until the store of the whole strip
        buf, tw, [&](int j) { return cscale(row[j], 0.5f * (srow[j] + mrow[j == 0 ? 0 : N - j])); },
        [&](int j, float2 f) { buf[fft_wave_pad<SH>(j)] = f; });  // the last pass has read buf before it writes
    __syncthreads();
    const int c = t & 7, r = t >> 3;
    if (j0 + c < wh) {
        float2 *Af = A + (int64_t)blockIdx.y * plane;
#pragma unroll 4
        for (int k = 0; k < N / 64; ++k) {
            const int i = r + 64 * k;
            Af[(int64_t)i * lda + j0 + c] = lds[c * SB + fft_wave_pad<SH>(i)];
        }
    }
}

template <int MAXN>
__global__ __launch_bounds__(kFftThreads) void rows_fft_fwd_inplace_kernel(FftDesc d, float2 *B, int64_t plane) {
    extern __shared__ float2 buf[];
    const int n = d.n;
    float2 *row = B + (int64_t)blockIdx.y * plane + (int64_t)blockIdx.x * n;
    fft_row<-1, MAXN>(buf, d, [&](int j) { return row[j]; }, [&](int j, float2 f) { row[j] = f; });
}

// ------------------------------------------------------ column transforms
// Column FFTs (length h = d.n) of a row-major plane, strip by strip: a
// workgroup reads `sw` adjacent columns row by row (contiguous pieces of
// each row, no transpose), transforms them together in LDS (column-major
// there) and writes them back in place.  MODE 0: plain; MODE 1: also write
// |F|^2 to P (laid out as A); MODE 2: multiply element (i, j) by
// (S(i, j) + S((h - i) % h, (scols - j) % scols)) / 2 on the way in, S the
// full h x scols scale.  BIG: one column longer
// than kFftMultiMax through fft_lds.  grid (ceil(ncols/sw), images); planes
// `plane` elements apart, rows `ld` elements apart.
template <int DIR, int MODE, bool BIG>
__global__ __launch_bounds__(kFftThreads, BIG ? 1 : 2) void cols_fft_strip_kernel(FftDesc d, float2 *A, int64_t plane,
                                                                                   int ld, int ncols, int sw,
                                                                                   const float *S, float *P,
                                                                                   int scols) {
    extern __shared__ float2 buf[];
    const int h = d.n;
    const int j0 = blockIdx.x * sw;
    const int ncol = min(sw, ncols - j0);
    float2 *a = A + (int64_t)blockIdx.y * plane;
    for (int idx = threadIdx.x; idx < h * ncol; idx += kFftThreads) {
        const int i = idx / ncol, c = idx - (idx / ncol) * ncol;
        const int64_t at = (int64_t)i * ld + j0 + c;
        float2 v = a[at];
        if (MODE == 2) {
            const float *Sb = S + (int64_t)blockIdx.y * plane;
            const int j = j0 + c;
            const float s1 = Sb[(int64_t)i * scols + j];
            const float s2 = Sb[(int64_t)(i == 0 ? 0 : h - i) * scols + (j == 0 ? 0 : scols - j)];
            v = cscale(v, 0.5f * (s1 + s2));
        }
        buf[c * h + i] = v;
    }
    __syncthreads();
    if constexpr (BIG)
        fft_lds<DIR>(buf, d);
    else
        fft_lds_multi<DIR>(buf, d, ncol);
    for (int idx = threadIdx.x; idx < h * ncol; idx += kFftThreads) {
        const int i = idx / ncol, c = idx - (idx / ncol) * ncol;
        const int64_t at = (int64_t)i * ld + j0 + c;
        const float2 f = buf[c * h + i];
        a[at] = f;
        if (MODE == 1) P[(int64_t)blockIdx.y * plane + at] = f.x * f.x + f.y * f.y;
    }
}

// ------------------------------------------------------------- workspace
// Columns per strip: a power of two <= 16 with sw * h <= kFftMultiMax.
int col_strip(int h) {
    int sw = 16;
    while (sw > 1 && (size_t)sw * h > (size_t)kFftMultiMax) sw >>= 1;
    return sw;
}
// Column transforms by strips up to 128 rows, through transposes above
// (batches of 32 measured both ways, NCC / PCE images/s): 1024 x 1024
// 26.1k / 20.8k by 4-column strips (32-byte row pieces) -> 37.5k / 29.2k;
// 512 x 512 99k / 78k by 8-column strips -> 123k / 98k.
bool strips_for(int h) { return h <= 128; }

// Per-image buffers, each an array over the batch with the given stride.
struct ExtractWs {
    int nb;                 // images the workspace holds
    int64_t n;              // plane stride (elements) of A, P, S
    float2 *A, *B;          // complex planes of the Wiener filter (B: transposed path only)
    float *P, *S;           // power spectrum / scale (S also holds the FastNoise output)
    double *rowsq, *var, *colpart, *rowpart, *colmean, *rowmean, *g;
    float *cz, *rz;         // FP32 zero-mean column terms (2w) and row means (2h) per image
    int64_t cp_stride, rp_stride;
    int tiles_x, tiles_y;
};

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Zero-mean partial-sum tiles the workspace holds: enough for either FastNoise kernel.
int zm_tiles_x_max(int w) { return std::max((w + kZmTileCols - 1) / kZmTileCols, zm_tiles_x_rows(w)); }
int zm_tiles_y_max(int h) { return std::max((h + kZmTileRows - 1) / kZmTileRows, zm_tiles_y_rows(h)); }

size_t per_image_bytes(int h, int w) {
    const int64_t n = (int64_t)h * w;
    const int tx = zm_tiles_x_max(w), ty = zm_tiles_y_max(h);
    return (strips_for(h) ? 16 : 24) * (size_t)n + 8 * (size_t)(h + 1 + 2 * w * ty + 2 * h * tx + 2 * w + 2 * h + 4) +
           4 * (size_t)(2 * w + 2 * h) + 13 * 256;
}

int get_ws(Context *c, int nb, ExtractWs &e) {
    const int h = c->h, w = c->w;
    const int64_t n = c->pixels();
    e.nb = nb;
    e.n = n;
    // the tiles of the zm_partials_kernel layout; launch_fastnoise reports its own
    e.tiles_x = (w + kZmTileCols - 1) / kZmTileCols;
    e.tiles_y = (h + kZmTileRows - 1) / kZmTileRows;
    e.cp_stride = 2 * (int64_t)w * zm_tiles_y_max(h);
    e.rp_stride = 2 * (int64_t)h * zm_tiles_x_max(w);
    const size_t sz[] = {align256(sizeof(float2) * n * nb),
                         strips_for(h) ? 0 : align256(sizeof(float2) * n * nb),
                         align256(sizeof(float) * n * nb),    align256(sizeof(float) * n * nb),
                         align256(sizeof(double) * h * nb),   align256(sizeof(double) * nb),
                         align256(sizeof(double) * e.cp_stride * nb), align256(sizeof(double) * e.rp_stride * nb),
                         align256(sizeof(double) * 2 * w * nb), align256(sizeof(double) * 2 * h * nb),
                         align256(sizeof(double) * 4 * nb),   align256(sizeof(float) * 2 * w * nb),
                         align256(sizeof(float) * 2 * h * nb)};
    size_t total = 0;
    for (size_t s : sz) total += s;
    PRNU_TRY(c->ws_filter.ensure(total));
    char *p = static_cast<char *>(c->ws_filter.ptr);
    void **slots[] = {(void **)&e.A,       (void **)&e.B,       (void **)&e.P,       (void **)&e.S,
                      (void **)&e.rowsq,   (void **)&e.var,     (void **)&e.colpart, (void **)&e.rowpart,
                      (void **)&e.colmean, (void **)&e.rowmean, (void **)&e.g,       (void **)&e.cz,
                      (void **)&e.rz};
    for (int k = 0; k < 13; ++k) {
        *slots[k] = p;
        p += sz[k];
    }
    return PRNU_OK;
}

size_t fft_smem(const FftDesc &d) { return sizeof(float2) * d.n; }

// Row-FFT kernel K<MAXN> over rows of length n: the register budget of the
// smallest of kFftSmallN, kFftMidN, kFftMaxN that holds the row.
#define PRNU_LAUNCH_ROW(K, n, grid, st, ...)                                                                    \
    do {                                                                                                        \
        if ((n) <= kFftSmallN)                                                                                  \
            hipLaunchKernelGGL((K<kFftSmallN>), grid, dim3(kFftThreads), fft_row_smem(n), st, __VA_ARGS__);     \
        else if ((n) <= kFftMidN)                                                                               \
            hipLaunchKernelGGL((K<kFftMidN>), grid, dim3(kFftThreads), fft_row_smem(n), st, __VA_ARGS__);       \
        else                                                                                                    \
            hipLaunchKernelGGL((K<kFftMaxN>), grid, dim3(kFftThreads), fft_row_smem(n), st, __VA_ARGS__);       \
    } while (0)

// The row-walking kernel needs widths that are multiples of 4 (>= 8), two
// rows, and 16-byte aligned output rows (the chain's workspace planes).  The
// choice depends on the shape alone for the chain, so a batch and its
// images one by one take the same kernel.
bool fastnoise_rows_ok(const float *out, int64_t out_stride, int h, int w, int nb) {
    if (w % 4 != 0 || w < 8 || h < 2) return false;
    if (reinterpret_cast<uintptr_t>(out) % 16 != 0) return false;
    return nb == 1 || out_stride % 4 == 0;
}

template <int FMT>
int launch_fastnoise_fmt(const void *in, int64_t in_stride, float *out, int64_t out_stride, int h, int w, int nb,
                         double *colpart, int64_t cp_stride, double *rowpart, int64_t rp_stride, int *tiles_x,
                         int *tiles_y, hipStream_t st) {
    if (fastnoise_rows_ok(out, out_stride, h, w, nb)) {
        const int nseg = (w + kFnSeg - 1) / kFnSeg, ngrp = (zm_tiles_y_rows(h) + 3) / 4;
        const int64_t nunits = (int64_t)nseg * ngrp * nb;
        PRNU_REQUIRE(nunits < (int64_t(1) << 30), "fastnoise: %lld units", (long long)nunits);
        const int per = (int)((nunits + 7) / 8);
        hipLaunchKernelGGL(fastnoise_rows_kernel<FMT>, dim3((unsigned)(8 * per)), dim3(256), 0, st,
                           static_cast<const uint8_t *>(in), in_stride, out, out_stride, h, w, colpart, cp_stride,
                           rowpart, rp_stride, nseg, ngrp, (int)nunits);
        PRNU_CHECK_LAUNCH();
        *tiles_x = zm_tiles_x_rows(w);
        *tiles_y = zm_tiles_y_rows(h);
        return PRNU_OK;
    }
    const int ntx = (w + kFnTX - 1) / kFnTX, nty = (h + kFnTY - 1) / kFnTY;
    const int64_t ntiles = (int64_t)ntx * nty * nb;
    PRNU_REQUIRE(ntiles < (int64_t(1) << 30), "fastnoise: %lld tiles", (long long)ntiles);
    const int per = (int)((ntiles + 7) / 8);
    hipLaunchKernelGGL(fastnoise_kernel<FMT>, dim3((unsigned)(8 * per)), dim3(256), 0, st,
                       static_cast<const uint8_t *>(in), in_stride, out, out_stride, h, w, colpart, cp_stride, rowpart,
                       rp_stride, ntx, nty, (int)ntiles);
    PRNU_CHECK_LAUNCH();
    *tiles_x = ntx;
    *tiles_y = nty;
    return PRNU_OK;
}

// FastNoise of nb images; *tiles_x / *tiles_y: the zero-mean partial-sum
// tiles the kernel wrote (for zm_reduce_kernel).
int launch_fastnoise(int fmt, const void *in, int64_t in_stride, float *out, int64_t out_stride, int h, int w, int nb,
                     double *colpart, int64_t cp_stride, double *rowpart, int64_t rp_stride, int *tiles_x,
                     int *tiles_y, hipStream_t st) {
    switch (fmt) {
        case PRNU_INPUT_RGB8:
            return launch_fastnoise_fmt<PRNU_INPUT_RGB8>(in, in_stride, out, out_stride, h, w, nb, colpart, cp_stride,
                                                         rowpart, rp_stride, tiles_x, tiles_y, st);
        case PRNU_INPUT_GRAY8:
            return launch_fastnoise_fmt<PRNU_INPUT_GRAY8>(in, in_stride, out, out_stride, h, w, nb, colpart,
                                                          cp_stride, rowpart, rp_stride, tiles_x, tiles_y, st);
        default:
            return launch_fastnoise_fmt<PRNU_INPUT_FLOAT32>(in, in_stride, out, out_stride, h, w, nb, colpart,
                                                            cp_stride, rowpart, rp_stride, tiles_x, tiles_y, st);
    }
}

int launch_zm_means(Context *c, ExtractWs &e, int nb, hipStream_t st) {
    const int h = c->h, w = c->w;
    const int m = h > w ? h : w;
    hipLaunchKernelGGL(zm_reduce_kernel, dim3((m + 255) / 256, nb), dim3(256), 0, st, e.colpart, e.cp_stride,
                       e.rowpart, e.rp_stride, h, w, e.tiles_y, e.tiles_x, e.colmean, e.rowmean, e.rz);
    PRNU_CHECK_LAUNCH();
    hipLaunchKernelGGL(zm_lattice_kernel, dim3(nb), dim3(256), 0, st, e.colmean, h, w, e.g, e.cz);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

template <int DIR, int MODE>
int launch_cols_fft(const FftDesc &dh, float2 *A, int64_t plane, int ld, int ncols, int nb, const float *S, float *P,
                    hipStream_t st, int scols = 0) {
    const int h = dh.n;
    const bool big = h > kFftMultiMax;
    const int sw = big ? 1 : col_strip(h);
    const dim3 grid((ncols + sw - 1) / sw, nb);
    const size_t smem = sizeof(float2) * (size_t)sw * h;
    if (big)
        hipLaunchKernelGGL((cols_fft_strip_kernel<DIR, MODE, true>), grid, dim3(kFftThreads), smem, st, dh, A, plane,
                           ld, ncols, sw, S, P, scols);
    else
        hipLaunchKernelGGL((cols_fft_strip_kernel<DIR, MODE, false>), grid, dim3(kFftThreads), smem, st, dh, A, plane,
                           ld, ncols, sw, S, P, scols);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

// Row pitch of the chain's half-spectrum plane A: 128-byte rows (a multiple
// of 16 complex values) where the strip kernels read or write A in 64-byte
// row pieces (1024 and 1536 rows), so that no piece straddles two lines;
// w/2 + 1 elsewhere (transposes and short-column strips take A packed).  A
// plane holds h x w complex values, so the padded pitch is used only when it
// fits in w.
int half_plane_pitch(int h, int w) {
    const int wh = w / 2 + 1, padded = (wh + 15) & ~15;
    return (h == 1024 || h == 1536) && padded <= w ? padded : wh;
}

// Wiener filter of nb images (img, img_stride floats apart) into out, on
// half spectra: row FFTs two rows at a time, column FFTs (+ |F|^2) over the
// w/2+1 columns, the variance stencil on the full grid (mirrored reads of
// the half power spectrum), the symmetrized scale + inverse column FFTs,
// real inverse row FFTs two rows at a time; columns by strips or through
// transposes (strips_for).
int launch_wiener(Context *c, ExtractWs &e, int nb, const float *img, int64_t img_stride, float *out,
                  int64_t out_stride, bool zm, hipStream_t st) {
    const int h = c->h, w = c->w, wh = w / 2 + 1;
    const int64_t n = c->pixels();
    const FftDesc &dw = fft_desc(c->fft_w);
    const FftDesc &dh = fft_desc(c->fft_h);
    const int pairs = (h + 1) / 2;
    const int lda = half_plane_pitch(h, w);
    // (the two-wave 2048-point form measured slower here, 331 -> 365 us per
    // 32 frames of 2048x1536: four loads per value and FP64 sums at 4 waves
    // per SIMD; it pays in rows_c2r_kernel)
    if (w == 2048)
        hipLaunchKernelGGL((wiener_rows_fwd_kernel<kFftSmallN, true>), dim3(pairs, nb), dim3(kFftThreads),
                           sizeof(float2) * kFft2048PadLds, st, dw, img, img_stride, h, e.A, n, e.rowsq,
                           zm ? e.cz : nullptr, zm ? e.rz : nullptr, lda);
    else
        PRNU_LAUNCH_ROW(wiener_rows_fwd_kernel, w, dim3(pairs, nb), st, dw, img, img_stride, h, e.A, n, e.rowsq,
                        zm ? e.cz : nullptr, zm ? e.rz : nullptr, lda);
    PRNU_CHECK_LAUNCH();
    hipLaunchKernelGGL(wiener_var_kernel, dim3(nb), dim3(256), 0, st, e.rowsq, h, n, e.var);
    PRNU_CHECK_LAUNCH();
    if (strips_for(h)) {
        PRNU_TRY((launch_cols_fft<-1, 1>(dh, e.A, n, wh, wh, nb, nullptr, e.P, st)));
        PRNU_TRY(launch_varest(e.P, e.S, h, w, false, nb, e.var, st));
        PRNU_TRY((launch_cols_fft<1, 2>(dh, e.A, n, wh, wh, nb, e.S, nullptr, st, w)));
    } else {
        // transposed layout: B = A^T (wh x h), column FFTs as rows of B; the
        // zero-padded box-min stencil is transpose invariant
        const int per = ((wh + 7) / 8 + 7) / 8;
        if (h == 1024) {
            hipLaunchKernelGGL(cols_fft_power_strip_kernel<1024>, dim3(8 * per, nb), dim3(512), 0, st, dh.tw, e.A,
                               wh, e.B, e.P, n, per, lda);
        } else if (h == 1536) {
            hipLaunchKernelGGL(cols_fft_power_strip_kernel<1536>, dim3(8 * per, nb), dim3(512), 0, st, dh.tw, e.A,
                               wh, e.B, e.P, n, per, lda);
        } else {
            PRNU_TRY(launch_transpose_c(e.A, e.B, h, wh, nb, n, st));
            PRNU_LAUNCH_ROW(rows_fft_power_kernel, h, dim3(wh, nb), st, dh, e.B, e.P, n);
        }
        PRNU_CHECK_LAUNCH();
        PRNU_TRY(launch_varest(e.P, e.S, w, h, true, nb, e.var, st));
        if (h == 1024) {
            hipLaunchKernelGGL(cols_ifft_scaled_strip_kernel<1024>, dim3(8 * per, nb), dim3(512), 0, st, dh.tw, e.B,
                               e.S, n, w, wh, e.A, per, lda);
        } else if (h == 1536) {
            hipLaunchKernelGGL(cols_ifft_scaled_strip_kernel<1536>, dim3(8 * per, nb), dim3(512), 0, st, dh.tw, e.B,
                               e.S, n, w, wh, e.A, per, lda);
        } else {
            PRNU_LAUNCH_ROW(rows_ifft_scaled_kernel, h, dim3(wh, nb), st, dh, e.B, e.S, n, w);
            PRNU_CHECK_LAUNCH();
            PRNU_TRY(launch_transpose_c(e.B, e.A, wh, h, nb, n, st));
        }
        PRNU_CHECK_LAUNCH();
    }
    if (w == 2048)
        hipLaunchKernelGGL((rows_c2r_kernel<kFftSmallN, 2048>), dim3(pairs, nb), dim3(128),
                           sizeof(float2) * fft_wave_lds<2048>(), st, dw, e.A, n, h, out, out_stride,
                           1.0f / (float)((double)w * h), lda);
    else
        PRNU_LAUNCH_ROW(rows_c2r_kernel, w, dim3(pairs, nb), st, dw, e.A, n, h, out, out_stride,
                        1.0f / (float)((double)w * h), lda);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

// R2C spectrum of nb images into out (h x (w/2+1) complex each, out_stride
// float2 apart): half-row FFTs, then column FFTs (strips in place in out, or
// through transposes).
int launch_spectrum(Context *c, ExtractWs &e, int nb, const float *img, int64_t img_stride, float2 *out,
                    int64_t out_stride, hipStream_t st) {
    const int h = c->h, w = c->w, wh = w / 2 + 1;
    const FftDesc &dw = fft_desc(c->fft_w);
    const FftDesc &dh = fft_desc(c->fft_h);
    if (strips_for(h)) {
        PRNU_LAUNCH_ROW(rows_fft_half_kernel, w, dim3((h + 1) / 2, nb), st, dw, img, img_stride, out, out_stride, h,
                        wh);
        PRNU_CHECK_LAUNCH();
        PRNU_TRY((launch_cols_fft<-1, 0>(dh, out, out_stride, wh, wh, nb, nullptr, nullptr, st)));
        return PRNU_OK;
    }
    const int64_t n = c->pixels();
    const int lda = half_plane_pitch(h, w);
    if (w == 2048)
        hipLaunchKernelGGL((rows_fft_half_kernel<kFftSmallN, true>), dim3((h + 1) / 2, nb), dim3(kFftThreads),
                           sizeof(float2) * kFft2048PadLds, st, dw, img, img_stride, e.A, n, h, lda);
    else
        PRNU_LAUNCH_ROW(rows_fft_half_kernel, w, dim3((h + 1) / 2, nb), st, dw, img, img_stride, e.A, n, h, lda);
    PRNU_CHECK_LAUNCH();
    if (h == 1024 || h == 1536) {  // no transposes: strips of 8 columns through LDS
        const int per = ((wh + 7) / 8 + 7) / 8;
        if (h == 1024)
            hipLaunchKernelGGL(spectrum_cols_strip_kernel<1024>, dim3(8 * per, nb), dim3(512), 0, st, dh.tw, e.A, n,
                               wh, out, out_stride, per, lda);
        else
            hipLaunchKernelGGL(spectrum_cols_strip_kernel<1536>, dim3(8 * per, nb), dim3(512), 0, st, dh.tw, e.A, n,
                               wh, out, out_stride, per, lda);
        PRNU_CHECK_LAUNCH();
        return PRNU_OK;
    }
    PRNU_TRY(launch_transpose_c(e.A, e.B, h, wh, nb, n, st));
    PRNU_LAUNCH_ROW(rows_fft_fwd_inplace_kernel, h, dim3(wh, nb), st, dh, e.B, n);
    PRNU_CHECK_LAUNCH();
    PRNU_TRY(launch_transpose_c(e.B, out, wh, h, nb, n, st, out_stride));
    return PRNU_OK;
}

int input_bytes_per_pixel(int fmt) {
    return fmt == PRNU_INPUT_RGB8 ? 3 : fmt == PRNU_INPUT_GRAY8 ? 1 : fmt == PRNU_INPUT_FLOAT32 ? 4 : 0;
}

}  // namespace

// Whole PRNUFilter chain (PRNUFilter.java:124-138) for `count` images
// in_stride bytes apart into patterns out_stride bytes apart: NCC mode writes
// h*w FP32, PCE mode the h*(w/2+1) complex spectrum.
int run_preprocess_batch(Context *c, int fmt, const void *in, size_t in_bytes, int64_t in_stride, int count,
                         void *out, size_t out_cap, int64_t out_stride, size_t *out_bytes, hipStream_t st) {
    const int bpp = input_bytes_per_pixel(fmt);
    PRNU_REQUIRE(bpp > 0, "preprocess: unknown input format %d", fmt);
    PRNU_REQUIRE(count >= 1, "preprocess: count must be >= 1");
    PRNU_REQUIRE(fmt != PRNU_INPUT_FLOAT32 || reinterpret_cast<uintptr_t>(in) % 4 == 0,
                 "preprocess: FP32 input must be 4-byte aligned");
    const size_t img_bytes = (size_t)bpp * c->pixels();
    const size_t need = c->mode == PRNU_MODE_PCE ? 8 * (size_t)c->spectrum_elems() : 4 * (size_t)c->pixels();
    if (count > 1) {
        PRNU_REQUIRE(in_stride >= (int64_t)img_bytes && out_stride >= (int64_t)need,
                     "preprocess: strides (%lld, %lld) smaller than an image (%zu) or a pattern (%zu)",
                     (long long)in_stride, (long long)out_stride, img_bytes, need);
        // 8-bit images may start at any byte; FP32 images and patterns need their alignment
        PRNU_REQUIRE((fmt != PRNU_INPUT_FLOAT32 || in_stride % 4 == 0) &&
                         out_stride % (c->mode == PRNU_MODE_PCE ? 8 : 4) == 0,
                     "preprocess: FP32 strides must be multiples of the element size");
    } else {
        in_stride = (int64_t)img_bytes;
        out_stride = (int64_t)need;
    }
    PRNU_REQUIRE(in_bytes >= (size_t)in_stride * (count - 1) + img_bytes,
                 "preprocess: input has %zu bytes, needs %zu", in_bytes, (size_t)in_stride * (count - 1) + img_bytes);
    PRNU_REQUIRE(out_cap >= (size_t)out_stride * (count - 1) + need, "preprocess: output has %zu bytes, needs %zu",
                 out_cap, (size_t)out_stride * (count - 1) + need);
    PRNU_TRY(ensure_fft_plans(c));
    // images per pass: the whole batch if its workspace stays under 4 GiB
    const size_t per = per_image_bytes(c->h, c->w);
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>((size_t)count, (size_t(4) << 30) / per));
    ExtractWs e;
    PRNU_TRY(get_ws(c, chunk, e));
    const int64_t n = c->pixels();
    time_begin(c, st);
    for (int b0 = 0; b0 < count; b0 += chunk) {
        const int nb = std::min(chunk, count - b0);
        const uint8_t *src = static_cast<const uint8_t *>(in) + (size_t)b0 * in_stride;
        uint8_t *dst = static_cast<uint8_t *>(out) + (size_t)b0 * out_stride;
        // gray + FastNoise into S (FP32 images n apart), zero-mean sums on the side
        float *noise = e.S;
        PRNU_TRY(launch_fastnoise(fmt, src, in_stride, noise, n, c->h, c->w, nb, e.colpart, e.cp_stride, e.rowpart,
                                  e.rp_stride, &e.tiles_x, &e.tiles_y, st));
        PRNU_TRY(launch_zm_means(c, e, nb, st));
        if (c->mode == PRNU_MODE_NCC) {
            PRNU_TRY(launch_wiener(c, e, nb, noise, n, reinterpret_cast<float *>(dst), out_stride / 4, true, st));
        } else {
            // Wiener output into P (FP32 images n apart), then its spectrum
            PRNU_TRY(launch_wiener(c, e, nb, noise, n, e.P, n, true, st));
            PRNU_TRY(launch_spectrum(c, e, nb, e.P, n, reinterpret_cast<float2 *>(dst), out_stride / 8, st));
        }
    }
    time_end(c, st);
    if (out_bytes) *out_bytes = need;
    return PRNU_OK;
}

int run_preprocess(Context *c, int fmt, const void *in, size_t in_bytes, void *out, size_t out_cap,
                   size_t *out_bytes, hipStream_t st) {
    return run_preprocess_batch(c, fmt, in, in_bytes, 0, 1, out, out_cap, 0, out_bytes, st);
}

int run_filter_grayscale(Context *c, int fmt, const void *in, size_t in_bytes, float *out, hipStream_t st) {
    const int bpp = input_bytes_per_pixel(fmt);
    PRNU_REQUIRE(bpp > 0, "grayscale: unknown input format %d", fmt);
    // GrayscaleFilter.java:72-74
    PRNU_REQUIRE(in_bytes >= (size_t)bpp * c->pixels(), "grayscale: input has %zu bytes, needs %lld", in_bytes,
                 (long long)(bpp * c->pixels()));
    const int64_t n = c->pixels();
    const unsigned blocks = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    const uint8_t *src = static_cast<const uint8_t *>(in);
    if (fmt == PRNU_INPUT_RGB8)
        hipLaunchKernelGGL(gray_kernel<PRNU_INPUT_RGB8>, dim3(blocks), dim3(256), 0, st, src, out, n);
    else if (fmt == PRNU_INPUT_GRAY8)
        hipLaunchKernelGGL(gray_kernel<PRNU_INPUT_GRAY8>, dim3(blocks), dim3(256), 0, st, src, out, n);
    else
        hipLaunchKernelGGL(gray_kernel<PRNU_INPUT_FLOAT32>, dim3(blocks), dim3(256), 0, st, src, out, n);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

int run_filter_fastnoise(Context *c, float *img, hipStream_t st) {
    // the kernel reads a halo around each tile, so it cannot write in place
    ExtractWs e;
    PRNU_TRY(get_ws(c, 1, e));
    int tx = 0, ty = 0;
    PRNU_TRY(launch_fastnoise(PRNU_INPUT_FLOAT32, img, 4 * c->pixels(), e.P, c->pixels(), c->h, c->w, 1, nullptr, 0,
                              nullptr, 0, &tx, &ty, st));
    PRNU_CHECK_HIP(hipMemcpyAsync(img, e.P, 4 * (size_t)c->pixels(), hipMemcpyDeviceToDevice, st));
    return PRNU_OK;
}

int run_filter_zeromean(Context *c, float *img, hipStream_t st) {
    ExtractWs e;
    PRNU_TRY(get_ws(c, 1, e));
    dim3 grid(e.tiles_x, e.tiles_y, 1);
    hipLaunchKernelGGL(zm_partials_kernel, grid, dim3(256), 0, st, img, c->pixels(), c->h, c->w, e.colpart,
                       e.cp_stride, e.rowpart, e.rp_stride);
    PRNU_CHECK_LAUNCH();
    PRNU_TRY(launch_zm_means(c, e, 1, st));
    const int64_t n = c->pixels();
    const unsigned blocks = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    hipLaunchKernelGGL(zm_apply_kernel, dim3(blocks, 1), dim3(256), 0, st, img, n, c->h, c->w, e.colmean, e.rowmean,
                       e.g);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

int run_filter_wiener(Context *c, float *img, hipStream_t st) {
    PRNU_TRY(ensure_fft_plans(c));
    ExtractWs e;
    PRNU_TRY(get_ws(c, 1, e));
    return launch_wiener(c, e, 1, img, c->pixels(), img, c->pixels(), false, st);
}

int run_filter_spectral(Context *c, const float *img, float *out, size_t out_cap, hipStream_t st) {
    // SpectralFilter.java:45-47
    PRNU_REQUIRE(out_cap >= 8 * (size_t)c->spectrum_elems(), "spectral: output has %zu bytes, needs %lld", out_cap,
                 (long long)(8 * c->spectrum_elems()));
    PRNU_TRY(ensure_fft_plans(c));
    ExtractWs e;
    PRNU_TRY(get_ws(c, 1, e));
    return launch_spectrum(c, e, 1, img, c->pixels(), reinterpret_cast<float2 *>(out), c->spectrum_elems(), st);
}

}  // namespace prnu
