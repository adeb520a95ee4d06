// Per-pixel bodies of the eight ReSTIR/resolve passes and the three SVGF
// passes.  Kernels (irt_kernels.hip) call these one per launch (the
// reference's pass schedule) or several per launch (fused schedule).
#pragma once

#include <type_traits>

#include "irt_device.h"

namespace irt {

// Everything a lighting pass reads or writes.  The frame record travels in
// the kernel-argument segment, so its fields are scalar loads.
template <class P>
struct LightArgs {
  irt_frame F;
  SceneView sc;
  const float* depth;
  const typename P::N4* normal;
  const typename P::N4* color;
  const typename P::P2* params;
  const typename P::P2* motion;
  const float* prev_depth;
  const typename P::N4* prev_normal;
  const irt_reservoir* direct_in;
  const irt_reservoir* indirect_in;
  irt_reservoir* direct_out;
  irt_reservoir* indirect_out;
  const irt_reservoir* direct_history;
  const irt_reservoir* indirect_history;
  const irt_vpl* vpl_ring;  // (restir_max_history + 1) pools of vpl_count records
  typename P::L4* direct_lighting;
  typename P::R4* lighting;
  unsigned long long* counters;
  const GuidePx* guide;  // when set, surface data comes from the per-pixel guide records
};

template <class P>
__device__ __forceinline__ Surface load_surface(const LightArgs<P>& a, int32_t x, int32_t y) {
  const uint32_t W = a.F.size[0], H = a.F.size[1];
  const uint32_t id = (uint32_t)y * W + (uint32_t)x;
  if (a.guide) {
    const GuidePx g = ld_guide(a.guide, id);
    Surface s;
    s.p = v3{g.px, g.py, g.pz};
    s.n = normalize(v3{(float)g.nx, (float)g.ny, (float)g.nz});
    s.color = v3{(float)g.cr, (float)g.cg, (float)g.cb};
    s.rough = (float)g.m0;
    s.metal = (float)g.m1;
    surface_view(s, ld3(a.F.camera));
    return s;
  }
  float2 rm = ld2(a.params, id);
  Surface s;
  s.p = world(pixel_u(x, W), pixel_u(y, H), a.depth[id], a.F.inv_vp);
  s.n = normalize(xyz(ld4(a.normal, id)));
  s.color = xyz(ld4(a.color, id));
  s.rough = rm.x;
  s.metal = rm.y;
  surface_view(s, ld3(a.F.camera));
  return s;
}

// vpl_sample_valid — lighting.wgsl:22-24.
__device__ __forceinline__ bool vpl_sample_valid(const irt_frame& F, const irt_reservoir& r) {
  return r.generation == F.size[3] && r.selected_target > 0.0f && r.id < F.counts[1] &&
         F.size[2] - r.source <= f2u_sat(F.rr[2]);
}
// selected_vpl — lighting.wgsl:25-29.  The current and previous pools are
// the ring slots of this and the previous frame, so one ring lookup serves
// all three cases of the reference.
__device__ __forceinline__ const irt_vpl& selected_vpl(const irt_frame& F, const irt_vpl* ring,
                                                      const irt_reservoir& r) {
  const uint32_t slots = f2u_sat(F.rr[2]) + 1u;
  return ring[(r.source % slots) * F.counts[1] + r.id];
}

// compatible — lighting.wgsl:36-40.
template <class P>
__device__ __forceinline__ bool compatible(const LightArgs<P>& a, const Surface& s, int32_t qx, int32_t qy) {
  const uint32_t W = a.F.size[0], H = a.F.size[1];
  const uint32_t q = (uint32_t)qy * W + (uint32_t)qx;
  v3 p, n;
  if (a.guide) {
    const GuidePx g = ld_guide(a.guide, q);
    if (!g.valid) return false;
    p = v3{g.px, g.py, g.pz};
    n = normalize(v3{(float)g.nx, (float)g.ny, (float)g.nz});
  } else {
    float d = a.depth[q];
    if (d >= 1.0f) return false;
    p = world(pixel_u(qx, W), pixel_u(qy, H), d, a.F.inv_vp);
    n = normalize(xyz(ld4(a.normal, q)));
  }
  v3 cam = ld3(a.F.camera);
  return dot(s.n, n) > a.F.temporal[2] && length(s.p - p) < a.F.temporal[1] * fmaxf(length(s.p - cam), 1.0f);
}

// reproject — lighting.wgsl:41-44.
template <class P>
__device__ __forceinline__ void reproject(const LightArgs<P>& a, int32_t x, int32_t y, int32_t* qx, int32_t* qy) {
  const uint32_t W = a.F.size[0], H = a.F.size[1];
  float2 m = ld2(a.motion, (uint32_t)y * W + (uint32_t)x);
  float u = pixel_u(x, W) - m.x, v = pixel_u(y, H) - m.y;
  *qx = f2i_sat(floorf(u * (float)W));
  *qy = f2i_sat(floorf(v * (float)H));
}

// history_valid — lighting.wgsl:45-50.
template <class P>
__device__ __forceinline__ bool history_valid(const LightArgs<P>& a, const Surface& s, int32_t qx, int32_t qy) {
  const irt_frame& F = a.F;
  const int32_t W = (int32_t)F.size[0], H = (int32_t)F.size[1];
  if (F.flags[0] == 0u || F.flags[1] != F.size[3] || qx < 0 || qy < 0 || qx >= W || qy >= H) return false;
  const uint32_t q = (uint32_t)qy * (uint32_t)W + (uint32_t)qx;
  float d = a.prev_depth[q];
  if (d >= 1.0f) return false;
  v3 oldp = world(pixel_u(qx, W), pixel_u(qy, H), d, F.previous_inv_vp);
  v3 cam = ld3(F.camera);
  return dot(s.n, xyz(ld4(a.prev_normal, q))) > F.temporal[2] &&
         length(s.p - oldp) < F.temporal[1] * fmaxf(length(s.p - cam), 1.0f);
}

// candidates_direct — lighting.wgsl:52-58.
template <class P, class C>
__device__ __forceinline__ irt_reservoir candidates_direct(const LightArgs<P>& a, uint32_t id, bool valid,
                                                            const Surface& s, C& cnt) {
  const irt_frame& F = a.F;
  irt_reservoir r = empty_reservoir(F.size[2], F.size[3]);
  uint32_t rng = seed(id, 11u, F.size[2]);
  if (valid && F.counts[0] > 0u) {
    const float pdf = 1.0f / (float)F.counts[0];
    const float ypdf = rcp_refined(pdf);
    for (uint32_t i = 0; i < F.restir[0]; ++i) {
      uint32_t light_id = min(f2u_sat(rnd(rng) * (float)F.counts[0]), F.counts[0] - 1u);
      float target = luminance(direct(s, a.sc.lights[light_id], cnt));
      add(r, light_id, F.size[2], target, pdf, div_weight(target, pdf, ypdf), 1u, 0u, rng);
    }
  }
  return r;
}

// candidates_indirect — lighting.wgsl:59-65.
template <class P, class C>
__device__ __forceinline__ irt_reservoir candidates_indirect(const LightArgs<P>& a, uint32_t id, bool valid,
                                                              const Surface& s, C& cnt) {
  const irt_frame& F = a.F;
  irt_reservoir r = empty_reservoir(F.size[2], F.size[3]);
  uint32_t rng = seed(id, 13u, F.size[2]);
  if (valid) {
    const float pdf = 1.0f / (float)F.counts[1];
    const float ypdf = rcp_refined(pdf);
    const uint32_t slots = f2u_sat(F.rr[2]) + 1u;
    const irt_vpl* current = a.vpl_ring + (size_t)(F.size[2] % slots) * F.counts[1];
    for (uint32_t i = 0; i < F.restir[1]; ++i) {
      uint32_t vpl_id = min(f2u_sat(rnd(rng) * (float)F.counts[1]), F.counts[1] - 1u);
      float target = luminance(indirect(s, current[vpl_id], cnt));
      add(r, vpl_id, F.size[2], target, pdf, div_weight(target, pdf, ypdf), 1u, 0u, rng);
    }
  }
  return r;
}

// temporal_direct — lighting.wgsl:66-74.  `hv`: valid && history_valid, `q`: reprojected id.
template <class P, class C>
__device__ __forceinline__ irt_reservoir temporal_direct(const LightArgs<P>& a, irt_reservoir r, uint32_t id,
                                                          bool hv, uint32_t q, const Surface& s, C& cnt) {
  const irt_frame& F = a.F;
  uint32_t rng = seed(id, 17u, F.size[2]);
  if (hv) {
    const uint32_t hist = f2u_sat(F.rr[2]);
    irt_reservoir prev = cap(a.direct_history[q], hist * F.restir[0]);
    if (prev.generation == F.size[3] && prev.selected_target > 0.0f && prev.id < F.counts[0] && prev.age < hist) {
      float target = luminance(direct(s, a.sc.lights[prev.id], cnt));
      add(r, prev.id, prev.source, target, prev.pdf, target * normalization(prev) * (float)prev.count, prev.count,
          prev.age + 1u, rng);
    }
  }
  return r;
}

// temporal_indirect — lighting.wgsl:75-85.
template <class P, class C>
__device__ __forceinline__ irt_reservoir temporal_indirect(const LightArgs<P>& a, irt_reservoir r, uint32_t id,
                                                            bool hv, uint32_t q, const Surface& s, C& cnt) {
  const irt_frame& F = a.F;
  uint32_t rng = seed(id, 19u, F.size[2]);
  if (hv) {
    irt_reservoir prev = cap(a.indirect_history[q], f2u_sat(F.rr[2]) * F.restir[1]);
    if (vpl_sample_valid(F, prev)) {
      const irt_vpl& v = selected_vpl(F, a.vpl_ring, prev);
      float target = luminance(indirect(s, v, cnt));
      add(r, prev.id, prev.source, target, prev.pdf, target * normalization(prev) * (float)prev.count, prev.count,
          prev.age + 1u, rng);
    }
  }
  return r;
}

// neighbor — lighting.wgsl:86-89 (round() is half-to-even: rintf; cos and sin
// shared with the oracle, so the rounding lands on the same pixel).
__device__ __forceinline__ void neighbor(const irt_frame& F, int32_t x, int32_t y, uint32_t& rng, int32_t* qx,
                                         int32_t* qy) {
  float angle = (2.0f * IRT_PI) * rnd(rng);
  float radius = sqrtf(rnd(rng)) * (float)F.restir[3];
  float sn, cs;
  shared_sincos(angle, &sn, &cs);
  *qx = x + f2i_sat(rintf(cs * radius));
  *qy = y + f2i_sat(rintf(sn * radius));
}

// spatial_direct — lighting.wgsl:90-97.  `r` is direct_in[id].
template <class P, class C>
__device__ __forceinline__ irt_reservoir spatial_direct(const LightArgs<P>& a, irt_reservoir r, int32_t x, int32_t y,
                                                         bool valid, const Surface& s, C& cnt) {
  const irt_frame& F = a.F;
  const int32_t W = (int32_t)F.size[0], H = (int32_t)F.size[1];
  const uint32_t id = (uint32_t)y * (uint32_t)W + (uint32_t)x;
  uint32_t rng = seed(id, 23u, F.size[2]);
  if (valid) {
    for (uint32_t i = 0; i < F.restir[2]; ++i) {
      int32_t qx, qy;
      neighbor(F, x, y, rng, &qx, &qy);
      if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
      if (!compatible(a, s, qx, qy)) continue;
      irt_reservoir prev = a.direct_in[(uint32_t)qy * (uint32_t)W + (uint32_t)qx];
      // The reference indexes lights[prev.id] unchecked; ids written this frame
      // are always < counts.x, and an out-of-range id (injected data) is skipped.
      if (prev.selected_target > 0.0f && prev.id < F.counts[0]) {
        float target = luminance(direct(s, a.sc.lights[prev.id], cnt));
        add(r, prev.id, prev.source, target, prev.pdf, target * normalization(prev) * (float)prev.count, prev.count,
            prev.age, rng);
      }
    }
  }
  return cap(r, f2u_sat(F.rr[2]) * F.restir[0] * 8u);
}

// spatial_indirect — lighting.wgsl:98-105.
template <class P, class C>
__device__ __forceinline__ irt_reservoir spatial_indirect(const LightArgs<P>& a, irt_reservoir r, int32_t x,
                                                           int32_t y, bool valid, const Surface& s, C& cnt) {
  const irt_frame& F = a.F;
  const int32_t W = (int32_t)F.size[0], H = (int32_t)F.size[1];
  const uint32_t id = (uint32_t)y * (uint32_t)W + (uint32_t)x;
  uint32_t rng = seed(id, 29u, F.size[2]);
  if (valid) {
    for (uint32_t i = 0; i < F.restir[2]; ++i) {
      int32_t qx, qy;
      neighbor(F, x, y, rng, &qx, &qy);
      if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
      if (!compatible(a, s, qx, qy)) continue;
      irt_reservoir prev = a.indirect_in[(uint32_t)qy * (uint32_t)W + (uint32_t)qx];
      if (vpl_sample_valid(F, prev)) {
        float target = luminance(indirect(s, selected_vpl(F, a.vpl_ring, prev), cnt));
        add(r, prev.id, prev.source, target, prev.pdf, target * normalization(prev) * (float)prev.count, prev.count,
            prev.age, rng);
      }
    }
  }
  return cap(r, f2u_sat(F.rr[2]) * F.restir[1] * 8u);
}

// spatial_direct and spatial_indirect of one pixel, neighbour i of both
// domains per iteration (the fused set-up).  Each domain keeps its own random
// stream and its own order of neighbour, compatibility test, reservoir load
// and add, so the reservoirs equal those of the two loops run one after the
// other; interleaving only lets the two domains' dependent loads overlap.
// The compatibility tests read the guide records (a.guide is set).
template <class P, class C>
__device__ __forceinline__ void spatial_both(const LightArgs<P>& a, irt_reservoir& rd, irt_reservoir& ri, int32_t x,
                                             int32_t y, bool valid, const Surface& s, C& cnt) {
  const irt_frame& F = a.F;
  const int32_t W = (int32_t)F.size[0], H = (int32_t)F.size[1];
  const uint32_t id = (uint32_t)y * (uint32_t)W + (uint32_t)x;
  uint32_t rng_d = seed(id, 23u, F.size[2]), rng_i = seed(id, 29u, F.size[2]);
  if (valid) {
    for (uint32_t i = 0; i < F.restir[2]; ++i) {
      int32_t dx, dy, ix, iy;
      neighbor(F, x, y, rng_d, &dx, &dy);
      neighbor(F, x, y, rng_i, &ix, &iy);
      const bool in_d = !(dx < 0 || dy < 0 || dx >= W || dy >= H);
      const bool in_i = !(ix < 0 || iy < 0 || ix >= W || iy >= H);
      // both guide records are loaded before either is tested (clamped
      // coordinates keep an outside neighbour's load inside the image)
      const bool ok_d = in_d && compatible(a, s, min(max(dx, 0), W - 1), min(max(dy, 0), H - 1));
      const bool ok_i = in_i && compatible(a, s, min(max(ix, 0), W - 1), min(max(iy, 0), H - 1));
      if (ok_d) {
        irt_reservoir prev = a.direct_in[(uint32_t)dy * (uint32_t)W + (uint32_t)dx];
        if (prev.selected_target > 0.0f && prev.id < F.counts[0]) {
          float target = luminance(direct(s, a.sc.lights[prev.id], cnt));
          add(rd, prev.id, prev.source, target, prev.pdf, target * normalization(prev) * (float)prev.count,
              prev.count, prev.age, rng_d);
        }
      }
      if (ok_i) {
        irt_reservoir prev = a.indirect_in[(uint32_t)iy * (uint32_t)W + (uint32_t)ix];
        if (vpl_sample_valid(F, prev)) {
          float target = luminance(indirect(s, selected_vpl(F, a.vpl_ring, prev), cnt));
          add(ri, prev.id, prev.source, target, prev.pdf, target * normalization(prev) * (float)prev.count,
              prev.count, prev.age, rng_i);
        }
      }
    }
  }
  rd = cap(rd, f2u_sat(F.rr[2]) * F.restir[0] * 8u);
  ri = cap(ri, f2u_sat(F.rr[2]) * F.restir[1] * 8u);
}

// resolve_direct — lighting.wgsl:106-122.  `d` is the final direct reservoir.
template <class P, class C>
__device__ __forceinline__ v3 resolve_direct(const LightArgs<P>& a, const irt_reservoir& d, uint32_t id, bool valid,
                                             const Surface& s, C& cnt) {
  const irt_frame& F = a.F;
  v3 result = v3{0.f, 0.f, 0.f};
  if (valid) {
    const uint32_t mode = F.flags[2];
    if (d.selected_target > 0.0f && d.id < F.counts[0] && mode != 2u && mode < 3u) {
      const irt_light& l = a.sc.lights[d.id];
      v3 value = direct(s, l, cnt);
      if (luminance(value) > 0.0f) {
        float w = normalization(d);
        float vis = visibility(a.sc, F, s.p, s.n, ld3(l.position_strength), v3{0.f, 0.f, 0.f}, cnt);
        result = result + value * w * vis;
      }
    }
    if (mode == 3u) result = s.color;
    if (mode == 4u) result = s.n * 0.5f + v3{0.5f, 0.5f, 0.5f};
    if (mode == 5u) {
      float2 m = ld2(a.motion, id);
      result = v3{clampf(0.5f + m.x * 50.0f, 0.0f, 1.0f), clampf(0.5f + m.y * 50.0f, 0.0f, 1.0f), 0.5f};
    }
  }
  return result;
}

// resolve_indirect — lighting.wgsl:123-134.  `base` is direct_lighting[id].rgb.
template <class P, class C>
__device__ __forceinline__ v3 resolve_indirect(const LightArgs<P>& a, const irt_reservoir& v, v3 base, bool valid,
                                               const Surface& s, C& cnt) {
  const irt_frame& F = a.F;
  v3 result = base;
  if (valid && F.flags[2] != 1u && F.flags[2] < 3u) {
    if (vpl_sample_valid(F, v)) {
      const irt_vpl& smp = selected_vpl(F, a.vpl_ring, v);
      v3 value = indirect(s, smp, cnt);
      if (luminance(value) > 0.0f) {
        float w = normalization(v);
        float vis = visibility(a.sc, F, s.p, s.n, ld3(smp.position), ld3(smp.normal), cnt);
        result = result + value * w * vis;
      }
    }
  }
  return vmin(result, v3{65000.f, 65000.f, 65000.f});
}

// ----------------------------------------------------------------- SVGF
template <class P>
struct SvgfArgs {
  irt_frame F;
  const float* depth;
  const typename P::N4* normal;
  const typename P::N4* color;
  const typename P::P2* params;
  const typename P::P2* motion;
  const float* old_depth;
  const typename P::N4* old_normal;
  const typename P::R4* noisy;
  const typename P::R4* old_radiance;
  const typename P::M4* old_moments;
  const typename P::R4* input;
  const typename P::M4* moments;
  typename P::R4* out;
  typename P::R4* out2;  // optional second destination (feedback / final copy)
  typename P::M4* moments_out;
  uint32_t stride;
  unsigned long long* counters;
  const GuidePx* guide;  // when set, positions and G-buffer channels come from the guide records
};

// inside, position, history_ok, guide, geometry_weight — svgf.wgsl:18-44.
struct Guide {
  v3 position, normal, color;
  float m0, m1, pixel_scale;
};

template <class P>
__device__ __forceinline__ bool inside(const SvgfArgs<P>& a, int32_t x, int32_t y) {
  return x >= 0 && y >= 0 && x < (int32_t)a.F.size[0] && y < (int32_t)a.F.size[1];
}
template <class P>
__device__ __forceinline__ v3 position(const SvgfArgs<P>& a, int32_t x, int32_t y) {
  const uint32_t W = a.F.size[0], H = a.F.size[1];
  const uint32_t id = (uint32_t)y * W + (uint32_t)x;
  if (a.guide) {
    const GuidePx g = ld_guide(a.guide, id);
    return v3{g.px, g.py, g.pz};
  }
  return world(pixel_u(x, W), pixel_u(y, H), a.depth[id], a.F.inv_vp);
}
template <class P>
__device__ __forceinline__ bool valid_px(const SvgfArgs<P>& a, uint32_t id) {
  return a.guide ? a.guide[id].valid != 0u : a.depth[id] < 1.0f;
}
template <class P>
__device__ __forceinline__ v3 normal_px(const SvgfArgs<P>& a, uint32_t id) {
  if (a.guide) {
    const GuidePx g = ld_guide(a.guide, id);
    return v3{(float)g.nx, (float)g.ny, (float)g.nz};
  }
  return xyz(ld4(a.normal, id));
}
template <class P>
__device__ __forceinline__ bool history_ok(const SvgfArgs<P>& a, int32_t px, int32_t py, int32_t qx, int32_t qy) {
  const irt_frame& F = a.F;
  if (F.flags[0] == 0u || F.flags[1] != F.size[3] || !inside(a, qx, qy)) return false;
  const uint32_t W = F.size[0], H = F.size[1];
  const uint32_t q = (uint32_t)qy * W + (uint32_t)qx;
  float d = a.old_depth[q];
  if (d >= 1.0f) return false;
  v3 pa = position(a, px, py);
  v3 pb = world(pixel_u(qx, W), pixel_u(qy, H), d, F.previous_inv_vp);
  v3 cam = ld3(F.camera);
  return length(pa - pb) < F.temporal[1] * fmaxf(length(pa - cam), 1.0f) &&
         dot(normal_px(a, (uint32_t)py * W + (uint32_t)px), xyz(ld4(a.old_normal, q))) > F.temporal[2];
}
// Guide of a pixel from its guide record.
__device__ __forceinline__ Guide guide_of(const irt_frame& F, const GuidePx& r) {
  Guide g;
  g.position = v3{r.px, r.py, r.pz};
  g.normal = v3{(float)r.nx, (float)r.ny, (float)r.nz};
  g.color = v3{(float)r.cr, (float)r.cg, (float)r.cb};
  g.m0 = (float)r.m0;
  g.m1 = (float)r.m1;
  g.pixel_scale = length(g.position - ld3(F.camera)) / (float)F.size[1];
  return g;
}
template <class P>
__device__ __forceinline__ Guide guide(const SvgfArgs<P>& a, int32_t x, int32_t y) {
  const uint32_t W = a.F.size[0];
  const uint32_t id = (uint32_t)y * W + (uint32_t)x;
  if (a.guide) return guide_of(a.F, ld_guide(a.guide, id));
  Guide g;
  g.position = position(a, x, y);
  g.normal = xyz(ld4(a.normal, id));
  g.color = xyz(ld4(a.color, id));
  float2 rm = ld2(a.params, id);
  g.m0 = rm.x;
  g.m1 = rm.y;
  g.pixel_scale = length(g.position - ld3(a.F.camera)) / (float)a.F.size[1];
  return g;
}
// FAST (fused schedule): per-pixel reciprocals of the tolerance and of phi
// replace the per-tap divisions and the hardware square root replaces the
// correctly rounded one (<= 1 ulp in FP32 before the binary16 store).
template <bool FAST>
__device__ __forceinline__ float wsqrt(float x) {
  if constexpr (FAST) return __builtin_amdgcn_sqrtf(x);
  return sqrtf(x);
}
__device__ __forceinline__ float tap_tolerance(const Guide& g, float stride) {
  return fmaxf(0.0001f, g.pixel_scale * stride * 0.5f);
}

// The weight of geometry_weight for a tap of normal nq, position wq,
// colour cq and material rm.
template <bool FAST>
__device__ __forceinline__ float tap_weight(const Guide& g, v3 nq, v3 wq, v3 cq, float2 rm, float stride,
                                            float inv_tol) {
  float wn = pow64(fmaxf(dot(g.normal, nq), 0.0f));
  float plane = fabsf(dot(wq - g.position, g.normal));
  float wz;
  if constexpr (FAST) {
    wz = __expf(-plane * inv_tol);
  } else {
    wz = __expf(-plane / tap_tolerance(g, stride));
  }
  float dm0 = g.m0 - rm.x, dm1 = g.m1 - rm.y;
  v3 dc = g.color - cq;
  float wc = __expf(-wsqrt<FAST>(dot(dc, dc)) * 8.0f - wsqrt<FAST>(dm0 * dm0 + dm1 * dm1) * 8.0f);
  return wn * wz * wc;
}
// Tap weight from a guide record (records of pixels without geometry weigh 0).
template <bool FAST>
__device__ __forceinline__ float guide_tap_weight(const Guide& g, const GuidePx& r, float stride, float inv_tol) {
  if (!r.valid) return 0.0f;
  return tap_weight<FAST>(g, v3{(float)r.nx, (float)r.ny, (float)r.nz}, v3{r.px, r.py, r.pz},
                          v3{(float)r.cr, (float)r.cg, (float)r.cb}, make_float2((float)r.m0, (float)r.m1), stride,
                          inv_tol);
}
// The mixed profile's fused SVGF taps.  Its radiance and moments are stored
// in 16 bits, so the tap arithmetic fuses each product and sum into one
// multiply-add (one rounding), lets the records' binary16 channels enter the
// multiply-adds as they are, and evaluates the product of the tap's
// exponential weights (svgf.wgsl:30-44 plane and colour terms, and the
// a-trous luminance term) as one exponential of the summed exponents.  A tap
// whose weight is zero adds zero instead of being skipped: staged radiance
// and moments are finite.  The FP32 profile keeps the reference's
// evaluation order (tap_weight).
template <class P>
constexpr bool fused_svgf() {
  return std::is_same<P, ProfileMixed>::value;
}
struct TapTerms {
  float wn;   // normal weight, 0 for records without geometry
  float arg;  // exponent of the plane and colour weights
};
__device__ __forceinline__ float lum_fused(float r, float g, float b) {
  return fmaf(b, 0.0722f, fmaf(g, 0.7152f, r * 0.2126f));
}
__device__ __forceinline__ TapTerms tap_terms_fused(const Guide& g, const GuidePx& r, float inv_tol) {
  const float nd = fmaf(g.normal.z, (float)r.nz, fmaf(g.normal.y, (float)r.ny, g.normal.x * (float)r.nx));
  const float plane = fabsf(fmaf(r.pz - g.position.z, g.normal.z,
                                 fmaf(r.py - g.position.y, g.normal.y, (r.px - g.position.x) * g.normal.x)));
  const float dr = fmaf(-1.0f, (float)r.cr, g.color.x), dg = fmaf(-1.0f, (float)r.cg, g.color.y);
  const float db = fmaf(-1.0f, (float)r.cb, g.color.z);
  const float d0 = fmaf(-1.0f, (float)r.m0, g.m0), d1 = fmaf(-1.0f, (float)r.m1, g.m1);
  const float lc = __builtin_amdgcn_sqrtf(fmaf(db, db, fmaf(dg, dg, dr * dr)));
  const float lm = __builtin_amdgcn_sqrtf(fmaf(d1, d1, d0 * d0));
  return TapTerms{r.valid ? pow64(fmaxf(nd, 0.0f)) : 0.0f, fmaf(-plane, inv_tol, -8.0f * (lc + lm))};
}
// wn * exp(arg), 0 for records without geometry (whose terms may not be finite).
__device__ __forceinline__ float tap_weight_fused(const TapTerms& t, bool valid, float extra_arg) {
  const float w = t.wn * __expf(t.arg + extra_arg);
  return valid ? w : 0.0f;
}

template <class P, class C, bool FAST = false>
__device__ __forceinline__ float geometry_weight(const SvgfArgs<P>& a, const Guide& g, int32_t qx, int32_t qy,
                                                 float stride, float inv_tol, C& cnt) {
  cnt.add(C_TAP, 1);
  if (!inside(a, qx, qy)) return 0.0f;
  const uint32_t W = a.F.size[0];
  const uint32_t q = (uint32_t)qy * W + (uint32_t)qx;
  if (a.guide) return guide_tap_weight<FAST>(g, ld_guide(a.guide, q), stride, inv_tol);
  float d = a.depth[q];
  if (d >= 1.0f) return 0.0f;
  return tap_weight<FAST>(g, xyz(ld4(a.normal, q)), world(pixel_u(qx, W), pixel_u(qy, a.F.size[1]), d, a.F.inv_vp),
                          xyz(ld4(a.color, q)), ld2(a.params, q), stride, inv_tol);
}

// svgf_temporal — svgf.wgsl:45-76.
template <class P, class C, bool FAST = false>
__device__ __forceinline__ void svgf_temporal(const SvgfArgs<P>& a, int32_t x, int32_t y, float4* rad_out,
                                              float4* mom_out, C& cnt) {
  const irt_frame& F = a.F;
  const uint32_t W = F.size[0], H = F.size[1];
  const uint32_t id = (uint32_t)y * W + (uint32_t)x;
  v3 c = xyz(ld4(a.noisy, id));
  float l = luminance(c);
  v3 history = v3{0.f, 0.f, 0.f};
  float hm0 = 0.f, hm1 = 0.f, length_history = 0.f, total = 0.f;
  if (valid_px(a, id)) {
    float2 m = ld2(a.motion, id);
    float px = (pixel_u(x, W) - m.x) * (float)W - 0.5f;
    float py = (pixel_u(y, H) - m.y) * (float)H - 0.5f;
    float flx = floorf(px), fly = floorf(py);
    int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
    float fx = px - flx, fy = py - fly;
    for (int32_t j = 0; j < 2; ++j) {
      for (int32_t i = 0; i < 2; ++i) {
        int32_t qx = bx + i, qy = by + j;
        if (!history_ok(a, x, y, qx, qy)) continue;
        const uint32_t q = (uint32_t)qy * W + (uint32_t)qx;
        float4 mo = ld_mom(a.old_moments, q);
        if (mo.z <= 0.0f) continue;
        float wx = i == 1 ? fx : 1.0f - fx;
        float wy = j == 1 ? fy : 1.0f - fy;
        float w = wx * wy;
        history = history + xyz(ld4(a.old_radiance, q)) * w;
        hm0 += mo.x * w;
        hm1 += mo.y * w;
        length_history += mo.z * w;
        total += w;
      }
    }
  }
  float age = 1.0f;
  v3 result = c;
  float m0 = l, m1 = l * l;
  if (total > 0.001f) {
    history = history / total;
    hm0 /= total;
    hm1 /= total;
    age = fminf(length_history / total + 1.0f, F.temporal[3]);
    float lo = 1e20f, hi = 0.0f;
    Guide g = guide(a, x, y);
    const float inv_tol = FAST ? 1.0f / tap_tolerance(g, 1.0f) : 0.0f;
    for (int32_t j = -1; j <= 1; ++j) {
      for (int32_t i = -1; i <= 1; ++i) {
        int32_t qx = min(max(x + i, 0), (int32_t)W - 1);
        int32_t qy = min(max(y + j, 0), (int32_t)H - 1);
        if (geometry_weight<P, C, FAST>(a, g, qx, qy, 1.0f, inv_tol, cnt) > 0.01f) {
          float v = luminance(xyz(ld4(a.noisy, (uint32_t)qy * W + (uint32_t)qx)));
          lo = fminf(lo, v);
          hi = fmaxf(hi, v);
        }
      }
    }
    float lh = luminance(history);
    history = history * (clampf(lh, lo * 0.5f, hi * 1.5f + 0.001f) / fmaxf(lh, 0.000001f));
    float alpha = fmaxf(1.0f - F.temporal[0], 1.0f / age);
    result = history * (1.0f - alpha) + c * alpha;
    float beta = fmaxf(0.2f, 1.0f / age);
    m0 = hm0 * (1.0f - beta) + m0 * beta;
    m1 = hm1 * (1.0f - beta) + m1 * beta;
  }
  float variance = fmaxf(m1 - m0 * m0, 0.0f);
  *rad_out = make_float4(result.x, result.y, result.z, fminf(sqrtf(variance), 65000.0f));
  *mom_out = make_float4(m0, m1, age, 0.0f);
}

// svgf_variance — svgf.wgsl:77-89.
template <class P, class C, bool FAST = false>
__device__ __forceinline__ float4 svgf_variance(const SvgfArgs<P>& a, int32_t x, int32_t y, C& cnt) {
  const uint32_t W = a.F.size[0];
  const uint32_t id = (uint32_t)y * W + (uint32_t)x;
  float4 c = ld4(a.input, id);
  float age = ld_mom(a.moments, id).z;
  float sigma = c.w;
  if (age < 4.0f && valid_px(a, id)) {
    float s0 = 0.f, s1 = 0.f, total = 0.f;
    Guide g = guide(a, x, y);
    const float inv_tol = FAST ? 1.0f / tap_tolerance(g, 1.0f) : 0.0f;
    for (int32_t j = -3; j <= 3; ++j) {
      for (int32_t i = -3; i <= 3; ++i) {
        int32_t qx = x + i, qy = y + j;
        float w = geometry_weight<P, C, FAST>(a, g, qx, qy, 1.0f, inv_tol, cnt);
        if (w <= 0.0f) continue;
        float4 mo = ld_mom(a.moments, (uint32_t)qy * W + (uint32_t)qx);
        s0 += mo.x * w;
        s1 += mo.y * w;
        total += w;
      }
    }
    float t = fmaxf(total, 0.00001f);
    s0 /= t;
    s1 /= t;
    sigma = sqrtf(fmaxf(s1 - s0 * s0, 0.0f) * 4.0f / fmaxf(age, 1.0f));
  }
  return make_float4(c.x, c.y, c.z, fminf(sigma, 65000.0f));
}

// history_ok with the current pixel's position and normal given (its guide
// record): the same tests in the same order.
template <class P>
__device__ __forceinline__ bool history_ok_at(const SvgfArgs<P>& a, v3 pa, v3 na, int32_t qx, int32_t qy) {
  const irt_frame& F = a.F;
  if (F.flags[0] == 0u || F.flags[1] != F.size[3] || !inside(a, qx, qy)) return false;
  const uint32_t W = F.size[0], H = F.size[1];
  const uint32_t q = (uint32_t)qy * W + (uint32_t)qx;
  float d = a.old_depth[q];
  if (d >= 1.0f) return false;
  v3 pb = world(pixel_u(qx, W), pixel_u(qy, H), d, F.previous_inv_vp);
  v3 cam = ld3(F.camera);
  return length(pa - pb) < F.temporal[1] * fmaxf(length(pa - cam), 1.0f) && dot(na, xyz(ld4(a.old_normal, q))) > F.temporal[2];
}

// svgf_temporal (FAST) for pixel (x, y) whose 3 x 3 clamp window is staged:
// guide records (two 16-byte halves) and noisy radiance of an n-wide patch
// with the pixel at (li, lj).  The reprojected history reads the previous
// frame's buffers.  Same arithmetic and order as svgf_temporal.
template <class P, class C>
__device__ __forceinline__ void svgf_temporal_staged(const SvgfArgs<P>& a, int32_t x, int32_t y, const uint4 (*sg)[2],
                                                     const typename P::R4* sr, int32_t li, int32_t lj, int32_t n,
                                                     float4* rad_out, float4* mom_out, C& cnt) {
  const irt_frame& F = a.F;
  const uint32_t W = F.size[0], H = F.size[1];
  const uint32_t id = (uint32_t)y * W + (uint32_t)x;
  const int32_t e0 = lj * n + li;
  v3 c = xyz(ld4(sr, (uint32_t)e0));
  float l = luminance(c);
  const GuidePx self = unpack_guide(sg[e0][0], sg[e0][1]);
  v3 history = v3{0.f, 0.f, 0.f};
  float hm0 = 0.f, hm1 = 0.f, length_history = 0.f, total = 0.f;
  if (self.valid) {
    const v3 pa = v3{self.px, self.py, self.pz};
    const v3 na = v3{(float)self.nx, (float)self.ny, (float)self.nz};
    float2 m = ld2(a.motion, id);
    float px = (pixel_u(x, W) - m.x) * (float)W - 0.5f;
    float py = (pixel_u(y, H) - m.y) * (float)H - 0.5f;
    float flx = floorf(px), fly = floorf(py);
    int32_t bx = f2i_sat(flx), by = f2i_sat(fly);
    float fx = px - flx, fy = py - fly;
    for (int32_t j = 0; j < 2; ++j) {
      for (int32_t i = 0; i < 2; ++i) {
        int32_t qx = bx + i, qy = by + j;
        if (!history_ok_at(a, pa, na, qx, qy)) continue;
        const uint32_t q = (uint32_t)qy * W + (uint32_t)qx;
        float4 mo = ld_mom(a.old_moments, q);
        if (mo.z <= 0.0f) continue;
        float wx = i == 1 ? fx : 1.0f - fx;
        float wy = j == 1 ? fy : 1.0f - fy;
        float w = wx * wy;
        history = history + xyz(ld4(a.old_radiance, q)) * w;
        hm0 += mo.x * w;
        hm1 += mo.y * w;
        length_history += mo.z * w;
        total += w;
      }
    }
  }
  float age = 1.0f;
  v3 result = c;
  float m0 = l, m1 = l * l;
  if (total > 0.001f) {
    history = history / total;
    hm0 /= total;
    hm1 /= total;
    age = fminf(length_history / total + 1.0f, F.temporal[3]);
    float lo = 1e20f, hi = 0.0f;
    const Guide g = guide_of(F, self);
    const float inv_tol = 1.0f / tap_tolerance(g, 1.0f);
    if constexpr (fused_svgf<P>()) {
      // one tap in flight at a time (see svgf_atrous_staged)
#pragma unroll 1
      for (int32_t j = -1; j <= 1; ++j) {
#pragma unroll 1
        for (int32_t i = -1; i <= 1; ++i) {
          int32_t qx = min(max(x + i, 0), (int32_t)W - 1);
          int32_t qy = min(max(y + j, 0), (int32_t)H - 1);
          cnt.add(C_TAP, 1);
          const int32_t e = (lj + qy - y) * n + li + qx - x;
          const GuidePx r = unpack_guide(sg[e][0], sg[e][1]);
          const float4 smp = ld4(sr, (uint32_t)e);
          if (tap_weight_fused(tap_terms_fused(g, r, inv_tol), r.valid != 0u, 0.0f) > 0.01f) {
            const float v = lum_fused(smp.x, smp.y, smp.z);
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
          }
        }
      }
    } else {
      for (int32_t j = -1; j <= 1; ++j) {
        for (int32_t i = -1; i <= 1; ++i) {
          int32_t qx = min(max(x + i, 0), (int32_t)W - 1);
          int32_t qy = min(max(y + j, 0), (int32_t)H - 1);
          cnt.add(C_TAP, 1);
          const int32_t e = (lj + qy - y) * n + li + qx - x;
          if (guide_tap_weight<true>(g, unpack_guide(sg[e][0], sg[e][1]), 1.0f, inv_tol) > 0.01f) {
            float v = luminance(xyz(ld4(sr, (uint32_t)e)));
            lo = fminf(lo, v);
            hi = fmaxf(hi, v);
          }
        }
      }
    }
    float lh = luminance(history);
    history = history * (clampf(lh, lo * 0.5f, hi * 1.5f + 0.001f) / fmaxf(lh, 0.000001f));
    float alpha = fmaxf(1.0f - F.temporal[0], 1.0f / age);
    result = history * (1.0f - alpha) + c * alpha;
    float beta = fmaxf(0.2f, 1.0f / age);
    m0 = hm0 * (1.0f - beta) + m0 * beta;
    m1 = hm1 * (1.0f - beta) + m1 * beta;
  }
  float variance = fmaxf(m1 - m0 * m0, 0.0f);
  *rad_out = make_float4(result.x, result.y, result.z, fminf(sqrtf(variance), 65000.0f));
  *mom_out = make_float4(m0, m1, age, 0.0f);
}

// svgf_variance for pixel (x, y) whose 7 x 7 neighbourhood is staged: the
// guide records (two 16-byte halves; valid = 0 outside the image) and the
// first two moments of an n-wide patch with the pixel at (li, lj).  Same
// arithmetic and tap order as svgf_variance with guide records.
template <class P, class C, bool FAST>
__device__ __forceinline__ float4 svgf_variance_staged(const SvgfArgs<P>& a, int32_t x, int32_t y,
                                                       const uint4 (*sg)[2], const float2* sm, int32_t li, int32_t lj,
                                                       int32_t n, C& cnt) {
  const uint32_t W = a.F.size[0];
  const uint32_t id = (uint32_t)y * W + (uint32_t)x;
  float4 c = ld4(a.input, id);
  float age = ld_mom(a.moments, id).z;
  float sigma = c.w;
  const int32_t e0 = lj * n + li;
  const GuidePx self = unpack_guide(sg[e0][0], sg[e0][1]);
  if (age < 4.0f && self.valid) {
    float s0 = 0.f, s1 = 0.f, total = 0.f;
    const Guide g = guide_of(a.F, self);
    const float inv_tol = FAST ? 1.0f / tap_tolerance(g, 1.0f) : 0.0f;
    if constexpr (fused_svgf<P>() && FAST) {
      // one tap in flight at a time (see svgf_atrous_staged)
#pragma unroll 1
      for (int32_t j = -3; j <= 3; ++j) {
#pragma unroll 1
        for (int32_t i = -3; i <= 3; ++i) {
          cnt.add(C_TAP, 1);
          const int32_t e = (lj + j) * n + li + i;
          const GuidePx r = unpack_guide(sg[e][0], sg[e][1]);
          const float w = tap_weight_fused(tap_terms_fused(g, r, inv_tol), r.valid != 0u, 0.0f);
          const float2 mo = sm[e];
          s0 = fmaf(mo.x, w, s0);
          s1 = fmaf(mo.y, w, s1);
          total += w;
        }
      }
    } else {
      for (int32_t j = -3; j <= 3; ++j) {
        for (int32_t i = -3; i <= 3; ++i) {
          cnt.add(C_TAP, 1);
          const int32_t e = (lj + j) * n + li + i;
          float w = guide_tap_weight<FAST>(g, unpack_guide(sg[e][0], sg[e][1]), 1.0f, inv_tol);
          if (w <= 0.0f) continue;
          const float2 mo = sm[e];
          s0 += mo.x * w;
          s1 += mo.y * w;
          total += w;
        }
      }
    }
    float t = fmaxf(total, 0.00001f);
    s0 /= t;
    s1 /= t;
    sigma = sqrtf(fmaxf(s1 - s0 * s0, 0.0f) * 4.0f / fmaxf(age, 1.0f));
  }
  return make_float4(c.x, c.y, c.z, fminf(sigma, 65000.0f));
}

// kernel, svgf_atrous — svgf.wgsl:90-110.
__device__ __forceinline__ float bspline(int32_t x) { return x == 0 ? 1.0f : ((x == 1 || x == -1) ? 2.0f / 3.0f : 1.0f / 6.0f); }

template <class P, class C, bool FAST = false>
__device__ __forceinline__ float4 svgf_atrous(const SvgfArgs<P>& a, int32_t x, int32_t y, C& cnt) {
  const uint32_t W = a.F.size[0], H = a.F.size[1];
  const uint32_t id = (uint32_t)y * W + (uint32_t)x;
  float4 c = ld4(a.input, id);
  if (!valid_px(a, id)) return c;
  float variance = 0.f, guide_weight = 0.f;
  Guide g = guide(a, x, y);
  for (int32_t j = -1; j <= 1; ++j) {
    for (int32_t i = -1; i <= 1; ++i) {
      int32_t qx = min(max(x + i, 0), (int32_t)W - 1);
      int32_t qy = min(max(y + j, 0), (int32_t)H - 1);
      float w = (i == 0 ? 2.0f : 1.0f) * (j == 0 ? 2.0f : 1.0f);
      float s = ld4(a.input, (uint32_t)qy * W + (uint32_t)qx).w;
      variance += s * s * w;
      guide_weight += w;
    }
  }
  float phi = 4.0f * sqrtf(variance / fmaxf(guide_weight, 0.00001f)) + 0.0001f;
  float lc = luminance(xyz(c));
  v3 sum = v3{0.f, 0.f, 0.f};
  float total = 0.f, v = 0.f;
  const int32_t step = (int32_t)a.stride;
  const float inv_tol = FAST ? 1.0f / tap_tolerance(g, (float)step) : 0.0f;
  const float inv_phi = FAST ? 1.0f / phi : 0.0f;
  for (int32_t j = -2; j <= 2; ++j) {
    for (int32_t i = -2; i <= 2; ++i) {
      int32_t qx = x + i * step, qy = y + j * step;
      float geom = geometry_weight<P, C, FAST>(a, g, qx, qy, (float)step, inv_tol, cnt);
      if (geom <= 0.0f) continue;
      float4 smp = ld4(a.input, (uint32_t)qy * W + (uint32_t)qx);
      float dl = fabsf(luminance(xyz(smp)) - lc);
      float w = bspline(i) * bspline(j) * geom * __expf(FAST ? -dl * inv_phi : -dl / phi);
      sum = sum + xyz(smp) * w;
      v += smp.w * smp.w * w * w;
      total += w;
    }
  }
  float t = fmaxf(total, 0.00001f);
  return make_float4(sum.x / t, sum.y / t, sum.z / t, fminf(sqrtf(v) / t, 65000.0f));
}

// svgf_atrous (FAST form) for a pixel whose tap neighbourhood is staged:
// `sg`/`sr` hold the guide records (two 16-byte halves; valid = 0 outside the
// image) and input radiance of an n x n patch of the stride-`step` lattice
// around the pixel, which sits at patch coordinates (li, lj).  `var_from_patch`
// (unit stride) reads the 3x3 variance neighbourhood from the patch as well;
// otherwise it is loaded from `a.input`.  Same arithmetic, same order.
template <class P, class C>
__device__ __forceinline__ float4 svgf_atrous_staged(const SvgfArgs<P>& a, int32_t x, int32_t y, const uint4 (*sg)[2],
                                                     const typename P::R4* sr, int32_t li, int32_t lj, int32_t n,
                                                     bool var_from_patch, C& cnt) {
  const uint32_t W = a.F.size[0], H = a.F.size[1];
  const int32_t e0 = lj * n + li;
  float4 c = ld4(sr, (uint32_t)e0);
  const GuidePx self = unpack_guide(sg[e0][0], sg[e0][1]);
  if (!self.valid) return c;
  float variance = 0.f, guide_weight = 0.f;
  const Guide g = guide_of(a.F, self);
  for (int32_t j = -1; j <= 1; ++j) {
    for (int32_t i = -1; i <= 1; ++i) {
      int32_t qx = min(max(x + i, 0), (int32_t)W - 1);
      int32_t qy = min(max(y + j, 0), (int32_t)H - 1);
      float w = (i == 0 ? 2.0f : 1.0f) * (j == 0 ? 2.0f : 1.0f);
      float sv = var_from_patch ? ld4(sr, (uint32_t)((lj + qy - y) * n + li + qx - x)).w
                                : ld4(a.input, (uint32_t)qy * W + (uint32_t)qx).w;
      variance += sv * sv * w;
      guide_weight += w;
    }
  }
  float phi = 4.0f * sqrtf(variance / fmaxf(guide_weight, 0.00001f)) + 0.0001f;
  float lc = luminance(xyz(c));
  v3 sum = v3{0.f, 0.f, 0.f};
  float total = 0.f, v = 0.f;
  const int32_t step = (int32_t)a.stride;
  const float inv_tol = 1.0f / tap_tolerance(g, (float)step);
  const float inv_phi = 1.0f / phi;
  if constexpr (fused_svgf<P>()) {
    const float lcf = lum_fused(c.x, c.y, c.z);
    // one tap in flight at a time: unrolled, the branch-free taps would be
    // scheduled with every tap's loads at once (registers for about 250
    // values, one wave per SIMD)
#pragma unroll 1
    for (int32_t j = -2; j <= 2; ++j) {
#pragma unroll 1
      for (int32_t i = -2; i <= 2; ++i) {
        cnt.add(C_TAP, 1);
        const int32_t e = (lj + j) * n + li + i;
        const GuidePx r = unpack_guide(sg[e][0], sg[e][1]);
        const float4 smp = ld4(sr, (uint32_t)e);
        const float dl = fabsf(lum_fused(smp.x, smp.y, smp.z) - lcf);
        const float w = (bspline(i) * bspline(j)) *
                        tap_weight_fused(tap_terms_fused(g, r, inv_tol), r.valid != 0u, -dl * inv_phi);
        sum = v3{fmaf(smp.x, w, sum.x), fmaf(smp.y, w, sum.y), fmaf(smp.z, w, sum.z)};
        v = fmaf(smp.w * smp.w, w * w, v);
        total += w;
      }
    }
    const float t = fmaxf(total, 0.00001f);
    return make_float4(sum.x / t, sum.y / t, sum.z / t, fminf(sqrtf(v) / t, 65000.0f));
  }
  for (int32_t j = -2; j <= 2; ++j) {
    for (int32_t i = -2; i <= 2; ++i) {
      cnt.add(C_TAP, 1);
      const int32_t e = (lj + j) * n + li + i;
      float geom = guide_tap_weight<true>(g, unpack_guide(sg[e][0], sg[e][1]), (float)step, inv_tol);
      if (geom <= 0.0f) continue;
      float4 smp = ld4(sr, (uint32_t)e);
      float dl = fabsf(luminance(xyz(smp)) - lc);
      float w = bspline(i) * bspline(j) * geom * __expf(-dl * inv_phi);
      sum = sum + xyz(smp) * w;
      v += smp.w * smp.w * w * w;
      total += w;
    }
  }
  float t = fmaxf(total, 0.00001f);
  return make_float4(sum.x / t, sum.y / t, sum.z / t, fminf(sqrtf(v) / t, 65000.0f));
}

}  // namespace irt
