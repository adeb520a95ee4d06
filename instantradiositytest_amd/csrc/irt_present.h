// Presentation — present.wgsl:1-8 drawn into the Rgba8UnormSrgb capture
// target (renderer.rs:1300-1306, presented at renderer.rs:1961): a
// full-screen triangle whose fragment at pixel (x, y) loads the denoised
// history at (x, y), so the display is the history's rgb with alpha 1,
// clamped to [0, 1], sRGB-encoded and stored in 8 bits.
//
// gfx950 has no sRGB render targets, so the encode is explicit: a channel's
// code is the number of decision thresholds at or below it, where threshold
// k is the linear value at which the exact sRGB encode (IEC 61966-2-1)
// times 255 reaches k + 0.5, rounded up to a float.  That is the correctly
// rounded 8-bit code of the exact curve, with no transcendental function on
// the device; the thresholds are computed once on the host in double.
#pragma once

#include "irt_device.h"

namespace irt {

constexpr int SRGB_THRESHOLDS = 255;

__device__ __forceinline__ uint32_t srgb_code(float v, const float* __restrict__ t) {
  v = fminf(fmaxf(v, 0.0f), 1.0f);  // UNORM clamp; NaN -> 0
  uint32_t k = 0;
#pragma unroll
  for (uint32_t step = 128; step; step >>= 1)
    if (k + step <= (uint32_t)SRGB_THRESHOLDS && t[k + step - 1] <= v) k += step;
  return k;
}

template <class P>
__global__ __launch_bounds__(256) void k_present(const typename P::R4* __restrict__ history, uint32_t n,
                                                 const float* __restrict__ thresholds, uchar4* __restrict__ out) {
  __shared__ float t[SRGB_THRESHOLDS];
  for (uint32_t i = threadIdx.x; i < (uint32_t)SRGB_THRESHOLDS; i += blockDim.x) t[i] = thresholds[i];
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float4 v = ld4(history, i);
  out[i] = make_uchar4((unsigned char)srgb_code(v.x, t), (unsigned char)srgb_code(v.y, t),
                       (unsigned char)srgb_code(v.z, t), (unsigned char)255);
}

}  // namespace irt
