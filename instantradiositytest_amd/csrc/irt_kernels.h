// Kernels of the lighting + denoising path.  Two schedules share the same
// per-pixel bodies (irt_passes.h):
//   passes: one 8x8-workgroup kernel per reference entry point, the
//           reference's buffer bindings and copies (renderer.rs:1889-1941,
//           svgf.rs:228-260);
//   fused : reservoirs A (candidates + temporal, both domains), reservoirs B
//           (spatial + both resolves, direct lighting kept in registers),
//           SVGF temporal, variance, a-trous with the feedback / final copies
//           folded into the a-trous stores.
#pragma once

#include "irt_passes.h"

namespace irt {

constexpr int TILE = 8;  // 8x8 = one 64-lane wavefront per workgroup

struct Px {
  int32_t x, y;
  uint32_t id;
  bool active;
};
__device__ __forceinline__ Px pixel(uint32_t W, uint32_t H) {
  Px p;
  p.x = (int32_t)(blockIdx.x * TILE + threadIdx.x);
  p.y = (int32_t)(blockIdx.y * TILE + threadIdx.y);
  p.active = (uint32_t)p.x < W && (uint32_t)p.y < H;
  p.id = p.active ? (uint32_t)p.y * W + (uint32_t)p.x : 0u;
  return p;
}

template <class P>
__device__ __forceinline__ bool valid_surface(const LightArgs<P>& a, const Px& p, Surface* s) {
  bool valid = p.active && (a.guide ? a.guide[p.id].valid != 0u : a.depth[p.id] < 1.0f);
  if (valid) {
    *s = load_surface(a, p.x, p.y);
  } else {
    *s = Surface{v3{0.f, 0.f, 0.f}, v3{0.f, 0.f, 0.f}, v3{0.f, 0.f, 0.f}, 0.f, 0.f};
  }
  return valid;
}

// ------------------------------------------------- per-pass (reference) schedule
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_candidates_direct(LightArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) a.direct_out[p.id] = candidates_direct(a, p.id, valid, s, cnt);
  cnt.flush(a.counters);
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_candidates_indirect(LightArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) a.indirect_out[p.id] = candidates_indirect(a, p.id, valid, s, cnt);
  cnt.flush(a.counters);
}
template <class P>
__device__ __forceinline__ bool temporal_gate(const LightArgs<P>& a, const Px& p, bool valid, const Surface& s,
                                              uint32_t* q) {
  int32_t qx, qy;
  reproject(a, p.x, p.y, &qx, &qy);
  bool hv = valid && history_valid(a, s, qx, qy);
  *q = hv ? (uint32_t)qy * a.F.size[0] + (uint32_t)qx : 0u;
  return hv;
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_temporal_direct(LightArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) {
    uint32_t q;
    bool hv = temporal_gate(a, p, valid, s, &q);
    a.direct_out[p.id] = temporal_direct(a, a.direct_in[p.id], p.id, hv, q, s, cnt);
  }
  cnt.flush(a.counters);
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_temporal_indirect(LightArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) {
    uint32_t q;
    bool hv = temporal_gate(a, p, valid, s, &q);
    a.indirect_out[p.id] = temporal_indirect(a, a.indirect_in[p.id], p.id, hv, q, s, cnt);
  }
  cnt.flush(a.counters);
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_spatial_direct(LightArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) a.direct_out[p.id] = spatial_direct(a, a.direct_in[p.id], p.x, p.y, valid, s, cnt);
  cnt.flush(a.counters);
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_spatial_indirect(LightArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) a.indirect_out[p.id] = spatial_indirect(a, a.indirect_in[p.id], p.x, p.y, valid, s, cnt);
  cnt.flush(a.counters);
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_resolve_direct(LightArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) {
    v3 r = resolve_direct(a, a.direct_in[p.id], p.id, valid, s, cnt);
    st4(a.direct_lighting, p.id, make_float4(r.x, r.y, r.z, 1.0f));
  }
  cnt.flush(a.counters);
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_resolve_indirect(LightArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) {
    v3 base = xyz(ld4(a.direct_lighting, p.id));
    v3 r = resolve_indirect(a, a.indirect_in[p.id], base, valid, s, cnt);
    st_r16(a.lighting, p.id, make_float4(r.x, r.y, r.z, 1.0f));
  }
  cnt.flush(a.counters);
}

// ---------------------------------------------------------- fused schedule
// Reservoirs A: candidates (both domains) then temporal reuse (both domains)
// on the in-register candidate reservoirs.  direct0/indirect0 keep the
// candidate reservoirs observable when non-null.
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_reservoirs_a(LightArgs<P> a, irt_reservoir* direct0,
                                                     irt_reservoir* indirect0) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) {
    irt_reservoir rd = candidates_direct(a, p.id, valid, s, cnt);
    irt_reservoir ri = candidates_indirect(a, p.id, valid, s, cnt);
    if (direct0) direct0[p.id] = rd;
    if (indirect0) indirect0[p.id] = ri;
    uint32_t q;
    bool hv = temporal_gate(a, p, valid, s, &q);
    a.direct_out[p.id] = temporal_direct(a, rd, p.id, hv, q, s, cnt);
    a.indirect_out[p.id] = temporal_indirect(a, ri, p.id, hv, q, s, cnt);
  }
  cnt.flush(a.counters);
}

// Reservoirs B: spatial reuse (both domains) and both resolves.  The direct
// radiance stays in registers (FP32) between the two resolves; it is stored
// to direct_lighting only when that buffer is requested.
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_reservoirs_b(LightArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (p.active) {
    irt_reservoir rd = spatial_direct(a, a.direct_in[p.id], p.x, p.y, valid, s, cnt);
    irt_reservoir ri = spatial_indirect(a, a.indirect_in[p.id], p.x, p.y, valid, s, cnt);
    a.direct_out[p.id] = rd;
    a.indirect_out[p.id] = ri;
    v3 dl = resolve_direct(a, rd, p.id, valid, s, cnt);
    if (a.direct_lighting) st4(a.direct_lighting, p.id, make_float4(dl.x, dl.y, dl.z, 1.0f));
    v3 r = resolve_indirect(a, ri, dl, valid, s, cnt);
    st_r16(a.lighting, p.id, make_float4(r.x, r.y, r.z, 1.0f));
  }
  cnt.flush(a.counters);
}

// ------------------------------------------------------------------- SVGF
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_svgf_temporal(SvgfArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  if (p.active) {
    float4 rad, mom;
    svgf_temporal(a, p.x, p.y, &rad, &mom, cnt);
    st_r16(a.out, p.id, rad);
    a.moments_out[p.id] = mom;
  }
  cnt.flush(a.counters);
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_svgf_variance(SvgfArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  if (p.active) st_r16(a.out, p.id, svgf_variance(a, p.x, p.y, cnt));
  cnt.flush(a.counters);
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_svgf_atrous(SvgfArgs<P> a) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  if (p.active) {
    float4 r = svgf_atrous(a, p.x, p.y, cnt);
    st_r16(a.out, p.id, r);
    if (a.out2) st_r16(a.out2, p.id, r);
  }
  cnt.flush(a.counters);
}

// Guide records of one frame: world position (world(), common.wgsl:46) and
// the binary16 channels of the G-buffer.
template <class P>
__global__ __launch_bounds__(256) void k_guide(irt_frame F, const float* depth, const typename P::N4* normal,
                                               const typename P::N4* color, const typename P::P2* params,
                                               GuidePx* out) {
  const uint32_t W = F.size[0], H = F.size[1];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  const int32_t x = (int32_t)(i % W), y = (int32_t)(i / W);
  const float d = depth[i];
  const v3 p = world(pixel_u(x, W), pixel_u(y, H), d, F.inv_vp);
  const float4 n = ld4(normal, i), c = ld4(color, i);
  const float2 m = ld2(params, i);
  GuidePx g;
  g.px = p.x;
  g.py = p.y;
  g.pz = p.z;
  g.valid = d < 1.0f ? 1u : 0u;
  g.nx = (_Float16)n.x;
  g.ny = (_Float16)n.y;
  g.nz = (_Float16)n.z;
  g.cr = (_Float16)c.x;
  g.cg = (_Float16)c.y;
  g.cb = (_Float16)c.z;
  g.m0 = (_Float16)m.x;
  g.m1 = (_Float16)m.y;
  out[i] = g;
}

// ------------------------------------------------------------- format I/O
// Reference formats (Depth32Float + binary16 channels) -> internal G-buffer.
template <class P>
__global__ __launch_bounds__(256) void k_ingest_gbuffer(uint32_t n, const float* depth, const uint16_t* normal,
                                                        const uint16_t* color, const uint16_t* params,
                                                        const uint16_t* motion, float* o_depth,
                                                        typename P::N4* o_normal, typename P::N4* o_color,
                                                        typename P::P2* o_params, typename P::P2* o_motion) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto h = [](uint16_t b) { return (float)__builtin_bit_cast(_Float16, b); };
  o_depth[i] = depth[i];
  const ushort4 nn = reinterpret_cast<const ushort4*>(normal)[i];
  const ushort4 cc = reinterpret_cast<const ushort4*>(color)[i];
  const ushort2 pp = reinterpret_cast<const ushort2*>(params)[i];
  const ushort2 mm = reinterpret_cast<const ushort2*>(motion)[i];
  st4(o_normal, i, make_float4(h(nn.x), h(nn.y), h(nn.z), h(nn.w)));
  st4(o_color, i, make_float4(h(cc.x), h(cc.y), h(cc.z), h(cc.w)));
  float2 pv = make_float2(h(pp.x), h(pp.y));
  float2 mv = make_float2(h(mm.x), h(mm.y));
  if constexpr (sizeof(typename P::P2) == sizeof(float2)) {
    reinterpret_cast<float2*>(o_params)[i] = pv;
    reinterpret_cast<float2*>(o_motion)[i] = mv;
  } else {
    reinterpret_cast<ushort2*>(o_params)[i] = pp;
    reinterpret_cast<ushort2*>(o_motion)[i] = mm;
  }
}

}  // namespace irt
