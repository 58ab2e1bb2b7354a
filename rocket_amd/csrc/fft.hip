// FFT plans (twiddle tables) and the row/transposition kernels shared by the
// Wiener filter, the spectral filter and PCE.
#include "prnu_internal.hpp"
#include "fft_device.hpp"
#include "fft_host.hpp"

#include <cmath>
#include <new>
#include <vector>

namespace prnu {

struct FftPlan {
    FftDesc desc;
    float2 *d_tw = nullptr;
};

// Radices 8, 4, 2 first (the twiddle exponent's j mod Ns is then a mask),
// then 3 and 5; a power-of-two part 8^k * 2 becomes 8^(k-1) * 4 * 4 (the
// same number of passes, no radix-2 pass).
bool fft_factor(int n, int *radix, int *passes) {
    int p = 0, e2 = 0;
    while (n % 2 == 0) { ++e2; n /= 2; }
    int n8 = e2 / 3, n4 = 0, n2 = 0;
    if (e2 % 3 == 2) n4 = 1;
    else if (e2 % 3 == 1) {
        if (n8 > 0) { --n8; n4 = 2; } else n2 = 1;
    }
    for (int i = 0; i < n8 && p < kFftMaxPasses; ++i) radix[p++] = 8;
    for (int i = 0; i < n4 && p < kFftMaxPasses; ++i) radix[p++] = 4;
    for (int i = 0; i < n2 && p < kFftMaxPasses; ++i) radix[p++] = 2;
    if (p < n8 + n4 + n2) return false;
    while (n % 3 == 0 && p < kFftMaxPasses) { radix[p++] = 3; n /= 3; }
    while (n % 5 == 0 && p < kFftMaxPasses) { radix[p++] = 5; n /= 5; }
    *passes = p;
    return n == 1;
}

int fft_plan_create(int n, FftPlan **out) {
    *out = nullptr;
    PRNU_REQUIRE(n >= 2 && n <= kFftMaxN, "fft: length %d outside [2, %d]", n, kFftMaxN);
    FftPlan *p = new (std::nothrow) FftPlan();
    if (!p) {
        set_error("fft: out of host memory");
        return PRNU_ENOMEM;
    }
    p->desc.n = n;
    if (!fft_factor(n, p->desc.radix, &p->desc.passes)) {
        delete p;
        set_error("fft: length %d has a prime factor other than 2, 3, 5", n);
        return PRNU_ENOTSUP;
    }
    std::vector<float2> tw(n);
    const double two_pi = 6.283185307179586476925286766559;
    for (int t = 0; t < n; ++t) {
        const double a = -two_pi * (double)t / (double)n;
        tw[t] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    hipError_t e = hipMalloc(&p->d_tw, sizeof(float2) * n);
    if (e == hipSuccess) e = hipMemcpy(p->d_tw, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (p->d_tw) (void)hipFree(p->d_tw);
        delete p;
        (void)hipGetLastError();
        set_error("fft: twiddle table: %s", hipGetErrorString(e));
        return PRNU_EDEVICE;
    }
    p->desc.tw = p->d_tw;
    *out = p;
    return PRNU_OK;
}

void fft_plan_destroy(FftPlan *p) {
    if (!p) return;
    if (p->d_tw) (void)hipFree(p->d_tw);
    delete p;
}

const FftDesc &fft_desc(const FftPlan *p) { return p->desc; }

int ensure_fft_plans(Context *c) {
    if (!c->fft_w) PRNU_TRY(fft_plan_create(c->w, &c->fft_w));
    if (!c->fft_h) {
        if (c->h == c->w) {
            c->fft_h = c->fft_w;
        } else {
            PRNU_TRY(fft_plan_create(c->h, &c->fft_h));
        }
    }
    return PRNU_OK;
}

// ------------------------------------------------------------ transposition
// out[b][c][r] = in[b][r][c] for an (rows x cols) complex matrix per batch item.
namespace {
constexpr int kTT = 32;

__global__ __launch_bounds__(256) void transpose_c_kernel(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                          int rows, int cols, int64_t in_stride, int64_t out_stride) {
    __shared__ float2 tile[kTT][kTT + 1];
    const int64_t off = (int64_t)blockIdx.z * in_stride, ooff = (int64_t)blockIdx.z * out_stride;
    const int c0 = blockIdx.x * kTT, r0 = blockIdx.y * kTT;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
#pragma unroll
    for (int i = 0; i < kTT; i += 8) {
        const int r = r0 + ty + i, c = c0 + tx;
        if (r < rows && c < cols) tile[ty + i][tx] = in[off + (int64_t)r * cols + c];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kTT; i += 8) {
        const int c = c0 + ty + i, r = r0 + tx;
        if (r < rows && c < cols) out[ooff + (int64_t)c * rows + r] = tile[tx][ty + i];
    }
}
}  // namespace

int launch_transpose_c(const float2 *in, float2 *out, int rows, int cols, int batch, int64_t batch_stride,
                       hipStream_t st, int64_t out_batch_stride) {
    dim3 grid((cols + kTT - 1) / kTT, (rows + kTT - 1) / kTT, batch);
    hipLaunchKernelGGL(transpose_c_kernel, grid, dim3(256), 0, st, in, out, rows, cols, batch_stride,
                       out_batch_stride < 0 ? batch_stride : out_batch_stride);
    PRNU_CHECK_LAUNCH();
    return PRNU_OK;
}

}  // namespace prnu
