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

constexpr int kChains = 16;  // independent accumulators per wave

// FMT: PRNU_PREC_SPLIT_BF16 -> v_mfma_f32_16x16x32_bf16, PRNU_PREC_FP16 ->
// v_mfma_f32_16x16x32_f16, PRNU_PREC_FP8 -> v_mfma_scale_f32_16x16x128_f8f6f4
// (e4m3, unit scales).  `magic` is never equal to the result; the store
// behind it keeps the loop from being removed.
template <int FMT>
__global__ __launch_bounds__(256) void mfma_peak_kernel(int iters, float magic, float *sink) {
    const int t = threadIdx.x;
    floatx4 acc[kChains];
#pragma unroll
    for (int j = 0; j < kChains; ++j) {  // distinct chains: none can be folded into another
        const float v = (float)(j * 64 + (t & 63));
        acc[j] = floatx4{v, v + 1.f, v + 2.f, v + 3.f};
    }
    if constexpr (FMT == PRNU_PREC_FP8) {
        intx8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            a[e] = 0x38383838 ^ (t & 1);
            b[e] = 0x30303030 ^ ((t >> 1) & 1);
        }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int j = 0; j < kChains; ++j)
                acc[j] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, acc[j], 0, 0, 0, 127, 0, 127);
        }
    } else if constexpr (FMT == PRNU_PREC_FP16) {
        halfx8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            a[e] = (_Float16)(1.0f + 0.001f * (t & 7));
            b[e] = (_Float16)(0.5f - 0.001f * (t & 3));
        }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int j = 0; j < kChains; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[j], 0, 0, 0);
        }
    } else {
        bf16x8 a, b;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            a[e] = (__bf16)(1.0f + 0.001f * (t & 7));
            b[e] = (__bf16)(0.5f - 0.001f * (t & 3));
        }
        for (int i = 0; i < iters; ++i) {
#pragma unroll
            for (int j = 0; j < kChains; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[j], 0, 0, 0);
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

int prnu_measure_mfma_peak(int precision, int iterations, double *tflops, float *ms) {
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
    const unsigned grid = (unsigned)cus * 2;  // 2 workgroups of 4 waves per CU: 2 waves per SIMD
    float *sink = nullptr;
    PRNU_CHECK_HIP(hipMalloc(&sink, 256));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    auto launch = [&](int it) {
        if (precision == PRNU_PREC_FP8)
            hipLaunchKernelGGL(mfma_peak_kernel<PRNU_PREC_FP8>, dim3(grid), dim3(256), 0, nullptr, it, -1.0f, sink);
        else if (precision == PRNU_PREC_FP16)
            hipLaunchKernelGGL(mfma_peak_kernel<PRNU_PREC_FP16>, dim3(grid), dim3(256), 0, nullptr, it, -1.0f, sink);
        else
            hipLaunchKernelGGL(mfma_peak_kernel<PRNU_PREC_SPLIT_BF16>, dim3(grid), dim3(256), 0, nullptr, it, -1.0f,
                               sink);
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
    const double flop = flop_per_mfma * kChains * (double)iterations * 4.0 * grid;
    *tflops = flop / (best * 1e-3) / 1e12;
    if (ms) *ms = best;
    return PRNU_OK;
}

}  // extern "C"
