// MFMA throughput the device sustains at the clock it holds under a full
// matrix load (a measurement of the roofline's ceiling; no reference
// counterpart).  Every wave issues independent MFMAs on register operands
// only -- no memory traffic in the loop -- with 2 waves per SIMD, the
// occupancy of the Gram kernels, so the rate is bounded by the MFMA pipes and
// by the shader clock the chip's power management allows under that load.
#include "prnu_internal.hpp"

namespace prnu {
namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef int intx8 __attribute__((ext_vector_type(8)));

constexpr int kChains = 32;  // independent accumulators per wave (8 x 4, as a Gram wave's 128 x 64 block)

__device__ __forceinline__ uint32_t mix32(uint32_t x) {  // an integer hash (bit avalanche)
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// FMT: PRNU_PREC_SPLIT_BF16 -> v_mfma_f32_16x16x32_bf16, PRNU_PREC_FP16 ->
// v_mfma_f32_16x16x32_f16, PRNU_PREC_FP8 -> v_mfma_scale_f32_16x16x128_f8f6f4
// (e4m3, unit scales).  Chain j multiplies A operand (j >> 2) & 3 by B operand
// j & 3.
// RANDOM: the operands hold hashed bits (values in (-2, 2), every mantissa bit
// toggling between operands, as with noise-like patterns); else small
// constants (few toggling bits: the chip holds a higher clock on them).
// `magic` is never equal to the result; the store behind it keeps the loop
// from being removed.
template <int FMT, bool RANDOM>
__global__ __launch_bounds__(512, 1) void mfma_peak_kernel(int iters, float magic, float *sink) {
    const int t = threadIdx.x;
    floatx4 acc[kChains];
#pragma unroll
    for (int j = 0; j < kChains; ++j) {  // distinct chains: none can be folded into another
        const float v = (float)(j * 64 + (t & 63));
        acc[j] = floatx4{v, v + 1.f, v + 2.f, v + 3.f};
    }
    const uint32_t seed = (uint32_t)(blockIdx.x * 512 + t) * 64u;
    auto val = [&](int k, int e) -> float {
        if (!RANDOM) return k < 4 ? 1.0f + 0.001f * (t & 7) : 0.5f - 0.001f * (t & 3);
        return (float)(int)(mix32(seed + (uint32_t)(k * 8 + e)) & 0xffffu) * (1.0f / 16384.0f) - 2.0f;
    };
    if constexpr (FMT == PRNU_PREC_FP8) {
        intx8 a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                // e4m3 bytes with the top exponent bit clear: no NaN, |x| < 2
                const uint32_t ra = RANDOM ? mix32(seed + (uint32_t)(k * 8 + e)) : 0x38383838u ^ (uint32_t)(t & 1);
                const uint32_t rb = RANDOM ? mix32(seed + 4096u + (uint32_t)(k * 8 + e)) : 0x30303030u;
                a[k][e] = (int)(ra & 0xBFBFBFBFu);
                b[k][e] = (int)(rb & 0xBFBFBFBFu);
            }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int j = 0; j < kChains; ++j)
                acc[j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a[(j >> 2) & 3], b[j & 3], acc[j], 0, 0, 0, 127, 0,
                                                                          127);
        }
    } else if constexpr (FMT == PRNU_PREC_FP16) {
        halfx8 a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                a[k][e] = (_Float16)val(k, e);
                b[k][e] = (_Float16)val(k + 4, e);
            }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int j = 0; j < kChains; ++j)
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[(j >> 2) & 3], b[j & 3], acc[j], 0, 0, 0);
        }
    } else {
        bf16x8 a[4], b[4];
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                a[k][e] = (__bf16)val(k, e);
                b[k][e] = (__bf16)val(k + 4, e);
            }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int j = 0; j < kChains; ++j)
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[(j >> 2) & 3], b[j & 3], acc[j], 0, 0, 0);
        }
    }
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < kChains; ++j) s += (acc[j][0] + acc[j][1]) + (acc[j][2] + acc[j][3]);
    if (s == magic) sink[0] = s;
}

}  // namespace
}  // namespace prnu

using namespace prnu;

extern "C" {

int prnu_measure_mfma_peak(int precision, int random_operands, int iterations, double *tflops, float *ms) {
    PRNU_REQUIRE(tflops, "measure_mfma_peak: null output");
    PRNU_REQUIRE(precision == PRNU_PREC_SPLIT_BF16 || precision == PRNU_PREC_FP16 || precision == PRNU_PREC_FP8,
                 "measure_mfma_peak: precision %d has no dense MFMA probe (bf16 = split-BF16, fp16, fp8)",
                 precision);
    PRNU_REQUIRE(iterations >= 1 && iterations <= (1 << 24), "measure_mfma_peak: iterations %d outside [1, 2^24]",
                 iterations);
    int dev = 0;
    PRNU_CHECK_HIP(hipGetDevice(&dev));
    int cus = 0;
    PRNU_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned grid = (unsigned)cus;  // one workgroup of 8 waves per CU: 2 waves per SIMD
    float *sink = nullptr;
    PRNU_CHECK_HIP(hipMalloc(&sink, 256));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto launch = [&](int it) {
#define PRNU_PEAK(F, R) hipLaunchKernelGGL((mfma_peak_kernel<F, R>), dim3(grid), dim3(512), 0, nullptr, it, -1.0f, sink)
        if (random_operands) {
            if (precision == PRNU_PREC_FP8) PRNU_PEAK(PRNU_PREC_FP8, true);
            else if (precision == PRNU_PREC_FP16) PRNU_PEAK(PRNU_PREC_FP16, true);
            else PRNU_PEAK(PRNU_PREC_SPLIT_BF16, true);
        } else {
            if (precision == PRNU_PREC_FP8) PRNU_PEAK(PRNU_PREC_FP8, false);
            else if (precision == PRNU_PREC_FP16) PRNU_PEAK(PRNU_PREC_FP16, false);
            else PRNU_PEAK(PRNU_PREC_SPLIT_BF16, false);
        }
#undef PRNU_PEAK
    };
    float best = 0.f;
    if (e == hipSuccess) {
        launch(iterations);  // warm-up: clocks settle under the load
        for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
            e = hipEventRecord(e0, nullptr);
            launch(iterations);
            if (e == hipSuccess) e = hipGetLastError();
            if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (e == hipSuccess && (rep == 0 || t < best)) best = t;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(sink);
    if (e != hipSuccess) {
        set_error("measure_mfma_peak: %s", hipGetErrorString(e));
        return PRNU_EDEVICE;
    }
    const double flop_per_mfma = precision == PRNU_PREC_FP8 ? 2.0 * 16 * 16 * 128 : 2.0 * 16 * 16 * 32;
    const double flop = flop_per_mfma * kChains * (double)iterations * 8.0 * grid;
    *tflops = flop / (best * 1e-3) / 1e12;
    if (ms) *ms = best;
    return PRNU_OK;
}

}  // extern "C"
