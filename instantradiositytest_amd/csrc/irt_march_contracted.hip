// The mixed profile's march kernels, compiled with multiply-add contraction
// inside each expression and the hardware's 1-ulp division and square root
// (build flags of this file: -ffp-contract=on
// -fno-hip-fp32-correctly-rounded-divide-sqrt).  The kernels' source is
// irt_march.h; only the arithmetic contract differs from the FP32 profile's.
#define IRT_MARCH_MATH IRT_MARCH_CONTRACTED
#include "irt_march.h"

namespace irt {

hipError_t launch_march_contracted(bool counters, dim3 grid, hipStream_t stream, const SceneView& sc,
                                   const irt_frame& F, uint32_t n_meshes, const RayReq* reqs, uint32_t n_rays,
                                   uint32_t* queue, float* vis, unsigned long long* counter_out) {
  if (counters)
    hipLaunchKernelGGL((k_march<true, IRT_MARCH_CONTRACTED>), grid, dim3(256), 0, stream, sc, F, n_meshes, reqs,
                       n_rays, queue, vis, counter_out);
  else
    hipLaunchKernelGGL((k_march<false, IRT_MARCH_CONTRACTED>), grid, dim3(256), 0, stream, sc, F, n_meshes, reqs,
                       n_rays, queue, vis, counter_out);
  return hipGetLastError();
}

hipError_t launch_march_views_contracted(dim3 grid, hipStream_t stream, const MarchView* views, uint32_t n_views) {
  hipLaunchKernelGGL((k_march_views<IRT_MARCH_CONTRACTED>), grid, dim3(256), 0, stream, views, n_views);
  return hipGetLastError();
}

}  // namespace irt
