// SDF baking — tools/export_scene.py:121-158 (build_sdf, inside) — and the
// field's Lipschitz bound (renderer.rs:330-358) on the GPU.
//
// Grid: the vertex bounds padded on every axis by |hi - lo| / (res - 1) * 2,
// samples on the nodes, x fastest; sample = sign * distance - extent * 2e-6
// with extent = |hi - lo| of the padded box.
//   distance: to the nearest triangle (closest point on a triangle: Ericson
//     2005, "Real-Time Collision Detection", 5.1.5), found by a stackless walk
//     of the mesh's triangle hierarchy that skips boxes farther than the best
//     distance so far;
//   sign: inside (negative) when at least two of three rays along the
//     exporter's fixed directions cross the surface an odd number of times;
//     each ray counts its closest hits within extent * 4 of the current start
//     and restarts max(extent * 1e-6, 1e-7) beyond each hit; a ray whose 256
//     casts all hit makes the bake fail, as the exporter raises.
// One lane per sample.  Positions and the final value are computed in double
// and rounded to float, as the exporter does in Python; distances and ray
// hits in FP32.
#pragma once

#include "irt_vpl.h"

namespace irt {

constexpr uint32_t BAKE_MAX_CROSSINGS = 256;

struct BakeGrid {
  double lo[3], hi[3];
  double extent;
  uint32_t res;
  uint32_t n_nodes;
};

// Squared distance from p to the box (0 inside).
__device__ __forceinline__ float box_distance2(v3 p, const float* lo, const float* hi) {
  const float dx = fmaxf(fmaxf(lo[0] - p.x, p.x - hi[0]), 0.0f);
  const float dy = fmaxf(fmaxf(lo[1] - p.y, p.y - hi[1]), 0.0f);
  const float dz = fmaxf(fmaxf(lo[2] - p.z, p.z - hi[2]), 0.0f);
  return dx * dx + dy * dy + dz * dz;
}

// Closest point on triangle abc to p (Ericson 2005, 5.1.5: Voronoi regions
// of the vertices, then of the edges, then the face).
__device__ __forceinline__ v3 closest_on_triangle(v3 p, v3 a, v3 b, v3 c) {
  const v3 ab = b - a, ac = c - a, ap = p - a;
  const float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  const v3 bp = p - b;
  const float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return a + ab * (d1 / (d1 - d3));
  const v3 cp = p - c;
  const float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return a + ac * (d2 / (d2 - d6));
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
    return b + (c - b) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
  const float denom = 1.0f / (va + vb + vc);
  const float v = vb * denom, w = vc * denom;
  return a + ab * v + ac * w;
}

__device__ __forceinline__ float nearest_distance2(const irt_tri_node* __restrict__ nodes, uint32_t n_nodes,
                                                   const irt_tri* __restrict__ tris, v3 p) {
  float best = INFINITY;
  uint32_t j = 0;
  while (j < n_nodes) {
    const irt_tri_node& nd = nodes[j];
    if (box_distance2(p, nd.lo, nd.hi) > best) {
      j = nd.count ? j + 1u : nd.link;
      continue;
    }
    for (uint32_t k = 0; k < nd.count; ++k) {
      const irt_tri& t = tris[nd.link + k];
      const v3 q = closest_on_triangle(p, ld3(t.a), ld3(t.b), ld3(t.c));
      const v3 d = p - q;
      best = fminf(best, dot(d, d));
    }
    ++j;
  }
  return best;
}

// Closest hit along o + t d, t in [0, tmax] (both faces; the tie rule of
// ray_triangle).  Returns t, or -1 for no hit.
__device__ __forceinline__ float cast_ray(const irt_tri_node* __restrict__ nodes, uint32_t n_nodes,
                                          const irt_tri* __restrict__ tris, v3 o, v3 d, float tmax) {
  RayHit h{tmax, 0.0f, 0.0f, 0xffffffffu, 0xffffffffu};
  const v3 inv = v3{slab_rcp(d.x), slab_rcp(d.y), slab_rcp(d.z)};
  uint32_t j = 0;
  while (j < n_nodes) {
    const irt_tri_node& nd = nodes[j];
    if (!ray_box(o, inv, nd.lo, nd.hi, 0.0f, h.t)) {
      j = nd.count ? j + 1u : nd.link;
      continue;
    }
    for (uint32_t k = 0; k < nd.count; ++k) {
      const irt_tri& tr = tris[nd.link + k];
      // ray_triangle with tmin 0
      const v3 a = ld3(tr.a);
      const v3 e1 = ld3(tr.b) - a, e2 = ld3(tr.c) - a;
      const v3 pv = cross(d, e2);
      const float det = dot(e1, pv);
      if (det == 0.0f) continue;
      const float inv_det = 1.0f / det;
      const v3 s = o - a;
      const float u = dot(s, pv) * inv_det;
      if (!(u >= 0.0f && u <= 1.0f)) continue;
      const v3 q = cross(s, e1);
      const float v = dot(d, q) * inv_det;
      if (!(v >= 0.0f && u + v <= 1.0f)) continue;
      const float t = dot(e2, q) * inv_det;
      if (!(t >= 0.0f) || !closer(t, 0u, tr.prim, h)) continue;
      h = RayHit{t, u, v, 0u, tr.prim};
    }
    ++j;
  }
  return h.object == 0xffffffffu ? -1.0f : h.t;
}

// The exporter's three parity directions (export_scene.py:124), normalised.
__device__ __forceinline__ v3 parity_direction(int k) {
  const v3 d = k == 0 ? v3{1.0f, 0.371f, 0.129f} : (k == 1 ? v3{0.173f, 1.0f, 0.419f} : v3{0.319f, 0.217f, 1.0f});
  return normalize(d);
}

__global__ __launch_bounds__(256) void k_bake_sdf(const irt_tri_node* __restrict__ nodes,
                                                  const irt_tri* __restrict__ tris, BakeGrid g,
                                                  float* __restrict__ out, uint32_t* __restrict__ overflow) {
  const uint32_t r = g.res;
  const uint64_t n = (uint64_t)r * r * r;
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t x = (uint32_t)(i % r), y = (uint32_t)((i / r) % r), z = (uint32_t)(i / ((uint64_t)r * r));
  const uint32_t c[3] = {x, y, z};
  float pf[3];
  for (int a = 0; a < 3; ++a) pf[a] = (float)(g.lo[a] + (g.hi[a] - g.lo[a]) * (double)c[a] / (double)(r - 1));
  const v3 p = v3{pf[0], pf[1], pf[2]};
  const float dist = sqrtf(nearest_distance2(nodes, g.n_nodes, tris, p));
  const float tmax = (float)(g.extent * 4.0);
  const float step = fmaxf((float)(g.extent * 1e-6), 1e-7f);
  int votes = 0;
  for (int k = 0; k < 3; ++k) {
    const v3 d = parity_direction(k);
    v3 start = p;
    uint32_t hits = 0;
    while (true) {
      const float t = cast_ray(nodes, g.n_nodes, tris, start, d, tmax);
      if (t < 0.0f) break;
      if (++hits >= BAKE_MAX_CROSSINGS) {  // the exporter's 256 casts all hit: it refuses the mesh
        atomicOr(overflow, 1u);
        break;
      }
      start = (start + d * t) + d * step;
    }
    votes += (int)(hits & 1u);
  }
  const double sign = votes >= 2 ? -1.0 : 1.0;
  out[i] = (float)(sign * (double)dist - g.extent * 2e-6);
}

// field_lipschitz_bound (renderer.rs:330-358) over the baked samples: the
// largest corner-gradient norm over cells with a sample above
// |spacing| * (0.5 + 1/128), at least 1, times 1.0001.  Grid-stride over the
// cells; per-lane, wave and workgroup maxima of the squared norms, then one
// atomicMax per workgroup on the float bits (non-negative floats order as
// integers).
__global__ __launch_bounds__(256) void k_lipschitz(const float* __restrict__ s, uint32_t r, float sx, float sy,
                                                   float sz, float threshold, uint32_t* __restrict__ max_bits) {
  __shared__ float wave_max[4];
  const uint64_t n = r, m = r - 1;
  const uint64_t cells = m * m * m;
  float best = 0.0f;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < cells; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = i % m, y = (i / m) % m, z = i / (m * m);
    float v[8];
    bool any = false;
    for (int k = 0; k < 8; ++k) {
      v[k] = s[(x + (k & 1)) + n * ((y + ((k >> 1) & 1)) + n * (z + (k >> 2)))];
      any |= !(v[k] <= threshold);
    }
    if (any)
      for (int k = 0; k < 8; ++k) {
        const float dx = (v[(k & 6) | 1] - v[k & 6]) / sx;
        const float dy = (v[(k & 5) | 2] - v[k & 5]) / sy;
        const float dz = (v[(k & 3) | 4] - v[k & 3]) / sz;
        best = fmaxf(best, dx * dx + dy * dy + dz * dz);
      }
  }
  for (int off = 32; off > 0; off >>= 1) best = fmaxf(best, __shfl_xor(best, off, 64));
  if (__lane_id() == 0) wave_max[threadIdx.x / 64] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = wave_max[0];
    for (uint32_t w = 1; w < blockDim.x / 64; ++w) b = fmaxf(b, wave_max[w]);
    if (b > 0.0f) atomicMax(max_bits, __float_as_uint(b));
  }
}

}  // namespace irt
