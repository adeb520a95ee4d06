// material_at — common.wgsl:50-66 — with the material textures of
// renderer.rs:951-1046: one RGBA8 2-D array of square layers, read through a
// linear view or, for sRGB base colours, an sRGB view, with the sampler of
// renderer.rs:1024-1031 (repeat addressing, linear filtering, level 0).
//
// Sampling restates textureSampleLevel: texel coordinates uv * size - 0.5,
// the four texels around them with repeat wrapping, each decoded to float
// (unorm: c / 255; sRGB view: the IEC 61966-2-1 decode, a 256-entry table
// computed in double), then blended bilinearly in float.  Hardware filtering
// quantises the blend weights to a few bits; this blend does not, so filtered
// values can differ from a GPU's by that quantisation.
#pragma once

#include "irt_device.h"

namespace irt {

struct TextureView {
  const uint32_t* texels;  // layer-major, rows top to bottom, RGBA8 packed little-endian
  const float* srgb_lut;   // 256 decoded sRGB values
  uint32_t size, layers;
};

__device__ __forceinline__ float4 texel(const TextureView& tv, uint32_t layer, uint32_t x, uint32_t y, bool srgb) {
  const uint32_t p = tv.texels[((size_t)layer * tv.size + y) * tv.size + x];
  const uint32_t r = p & 0xffu, g = (p >> 8) & 0xffu, b = (p >> 16) & 0xffu, a = p >> 24;
  if (srgb) return make_float4(tv.srgb_lut[r], tv.srgb_lut[g], tv.srgb_lut[b], (float)a / 255.0f);
  return make_float4((float)r / 255.0f, (float)g / 255.0f, (float)b / 255.0f, (float)a / 255.0f);
}

__device__ __forceinline__ uint32_t wrap(int32_t i, uint32_t n) {
  const int32_t m = i % (int32_t)n;
  return (uint32_t)(m < 0 ? m + (int32_t)n : m);
}

// textureSampleLevel(t, repeat-linear sampler, uv, layer, 0).
__device__ __forceinline__ float4 sample_layer(const TextureView& tv, uint32_t layer, float u, float v, bool srgb) {
  const float x = u * (float)tv.size - 0.5f, y = v * (float)tv.size - 0.5f;
  const float fx0 = floorf(x), fy0 = floorf(y);
  const float fx = x - fx0, fy = y - fy0;
  // floor of a finite coordinate fits int32 for any uv a mesh carries; clamp the rest
  const int32_t ix = (int32_t)fminf(fmaxf(fx0, -1073741824.0f), 1073741824.0f);
  const int32_t iy = (int32_t)fminf(fmaxf(fy0, -1073741824.0f), 1073741824.0f);
  const uint32_t x0 = wrap(ix, tv.size), x1 = wrap(ix + 1, tv.size);
  const uint32_t y0 = wrap(iy, tv.size), y1 = wrap(iy + 1, tv.size);
  const float4 a = texel(tv, layer, x0, y0, srgb), b = texel(tv, layer, x1, y0, srgb);
  const float4 c = texel(tv, layer, x0, y1, srgb), d = texel(tv, layer, x1, y1, srgb);
  const float gx = 1.0f - fx, gy = 1.0f - fy;
  return make_float4((a.x * gx + b.x * fx) * gy + (c.x * gx + d.x * fx) * fy,
                     (a.y * gx + b.y * fx) * gy + (c.y * gx + d.y * fx) * fy,
                     (a.z * gx + b.z * fx) * gy + (c.z * gx + d.z * fx) * fy,
                     (a.w * gx + b.w * fx) * gy + (c.w * gx + d.w * fx) * fy);
}

__device__ __forceinline__ float channel(float4 v, uint32_t k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

__device__ __forceinline__ bool textured(const irt_material& m) {
  return m.textures[0] >= 0 || m.textures[1] >= 0 || m.textures[2] >= 0 || m.textures[3] >= 0;
}

// material_at: colour, roughness, metalness and shading normal of a surface
// with interpolated normal n, uv and tangent (xyz, w = handedness sign).
__device__ __forceinline__ void material_at(const irt_material& m, const TextureView& tv, float u_in, float v_in,
                                            v3 n, float4 tangent, Surface& s) {
  const float tu = u_in, tv_ = 1.0f - v_in;  // texture_uv = (uv.x, 1 - uv.y)
  s.color = v3{m.color[0], m.color[1], m.color[2]};
  s.rough = m.params[0];
  s.metal = m.params[1];
  s.n = normalize(n);
  if (m.textures[0] >= 0) {
    const float4 c = sample_layer(tv, (uint32_t)m.textures[0], tu, tv_, m.params[3] > 0.5f);
    s.color = s.color * v3{c.x, c.y, c.z};
  }
  if (m.textures[1] >= 0) s.rough = s.rough * channel(sample_layer(tv, (uint32_t)m.textures[1], tu, tv_, false), m.channels[1]);
  if (m.textures[2] >= 0) s.metal = s.metal * channel(sample_layer(tv, (uint32_t)m.textures[2], tu, tv_, false), m.channels[2]);
  if (m.textures[3] >= 0) {
    const float4 t4 = sample_layer(tv, (uint32_t)m.textures[3], tu, tv_, false);
    v3 nm = v3{t4.x * 2.0f - 1.0f, t4.y * 2.0f - 1.0f, t4.z * 2.0f - 1.0f};
    nm = normalize(v3{nm.x * m.params[2], nm.y * m.params[2], nm.z});
    const v3 tg = v3{tangent.x, tangent.y, tangent.z};
    const v3 t = normalize(tg - s.n * dot(s.n, tg));
    const v3 c = s.n;
    const v3 b = v3{c.y * t.z - c.z * t.y, c.z * t.x - c.x * t.z, c.x * t.y - c.y * t.x} * tangent.w;
    s.n = normalize(t * nm.x + b * nm.y + s.n * nm.z);
  }
  s.rough = clampf(s.rough, 0.045f, 1.0f);
  s.metal = clampf(s.metal, 0.0f, 1.0f);
}

}  // namespace irt
