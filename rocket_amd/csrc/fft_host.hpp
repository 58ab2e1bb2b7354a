// Host-side FFT plan helpers shared by the extraction and PCE translation units.
#pragma once

#include "fft_device.hpp"
#include "prnu_internal.hpp"

namespace prnu {

bool fft_factor(int n, int *radix, int *passes);
int fft_plan_create(int n, FftPlan **out);
void fft_plan_destroy(FftPlan *p);
const FftDesc &fft_desc(const FftPlan *p);
// Create the context's length-w and length-h plans on first use.
int ensure_fft_plans(Context *c);
// out[b][c][r] = in[b][r][c], complex; input items batch_stride elements
// apart, output items out_batch_stride apart (negative: batch_stride).
int launch_transpose_c(const float2 *in, float2 *out, int rows, int cols, int batch, int64_t batch_stride,
                       hipStream_t st, int64_t out_batch_stride = -1);

}  // namespace prnu
