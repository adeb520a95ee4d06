// The VPL pass — vpl.wgsl:2-47 — for gfx950, which has no ray-tracing units.
//
// The reference traces each light path with hardware ray queries against a
// TLAS of per-mesh triangle BLASes (renderer.rs:1061-1106, 1786-1797).  Here
// the instance level is the objects' world boxes (the leaf boxes of the
// object hierarchy, which contain each object's triangles), tested side by
// side by the lanes of a group, and the mesh level is a stackless triangle
// hierarchy built on the host (irt_host::mesh_hierarchy).  A ray is moved
// into an object's space by its inverse model matrix, as a ray query moves it
// into instance space, so the hit distance is the world-space ray parameter.
//
// Closest hit: triangles are both-sided (opaque geometry, no cull flags),
// t in [0.0005, 10000] (the RayDesc of vpl.wgsl:18), Moller-Trumbore
// intersection (Moller & Trumbore 1997, "Fast, minimum storage ray/triangle
// intersection").  Equal distances go to the lower (object, triangle) pair,
// so the answer does not depend on the walk order: it equals a scan over
// every triangle of every object.
//
// A group of lanes per light path: the path's bounces are sequential, its
// rays' object tests run side by side across the group.
#pragma once

#include "irt_material.h"

namespace irt {

struct TriScene {
  const irt_object* objects;
  const irt_mesh* meshes;
  const uint2* mesh_nodes;  // per mesh: [first, end) of its triangle hierarchy
  const irt_tri_node* tri_nodes;
  const irt_tri* tris;
  const irt_vertex* vertices;
  const uint32_t* indices;
  const irt_material* materials;
  const irt_light* lights;
  TextureView tex;
};

struct RayHit {
  float t, u, v;
  uint32_t object, prim;
};

constexpr float RAY_TMIN = 0.0005f;
constexpr float RAY_TMAX = 10000.0f;

__device__ __forceinline__ v3 cross(v3 a, v3 b) {
  return v3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// 1/d per axis; a zero component gets a large finite reciprocal so that the
// slab products stay defined (0 * 1e30 = 0 where the origin lies on a slab).
__device__ __forceinline__ float slab_rcp(float d) { return d == 0.0f ? 1e30f : 1.0f / d; }

// Does the ray meet the box within [tmin, tmax]?
__device__ __forceinline__ bool ray_box(v3 o, v3 inv, const float* lo, const float* hi, float tmin, float tmax) {
  const float ax = (lo[0] - o.x) * inv.x, bx = (hi[0] - o.x) * inv.x;
  const float ay = (lo[1] - o.y) * inv.y, by = (hi[1] - o.y) * inv.y;
  const float az = (lo[2] - o.z) * inv.z, bz = (hi[2] - o.z) * inv.z;
  const float tn = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), fmaxf(fminf(az, bz), tmin));
  const float tf = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), fminf(fmaxf(az, bz), tmax));
  return tn <= tf;
}

__device__ __forceinline__ bool closer(float t, uint32_t object, uint32_t prim, const RayHit& h) {
  return t < h.t || (t == h.t && (object < h.object || (object == h.object && prim < h.prim)));
}

// Moller-Trumbore, both faces.  u, v weight the triangle's second and third
// corners (the barycentrics a ray query reports).
__device__ __forceinline__ void ray_triangle(v3 o, v3 d, const irt_tri& tr, uint32_t object, RayHit& h) {
  const v3 a = ld3(tr.a);
  const v3 e1 = ld3(tr.b) - a, e2 = ld3(tr.c) - a;
  const v3 p = cross(d, e2);
  const float det = dot(e1, p);
  if (det == 0.0f) return;
  const float inv = 1.0f / det;
  const v3 s = o - a;
  const float u = dot(s, p) * inv;
  if (!(u >= 0.0f && u <= 1.0f)) return;
  const v3 q = cross(s, e1);
  const float v = dot(d, q) * inv;
  if (!(v >= 0.0f && u + v <= 1.0f)) return;
  const float t = dot(e2, q) * inv;
  if (!(t >= RAY_TMIN) || !closer(t, object, tr.prim, h)) return;
  h = RayHit{t, u, v, object, tr.prim};
}

// Closest hit of the ray in object `obj`'s triangles, folded into h: the
// ray moves into object space by the inverse model matrix (a ray query's
// instance transform) and walks the mesh's hierarchy without a stack.
__device__ __forceinline__ void trace_object(const TriScene& sc, uint32_t obj, v3 o, v3 d, RayHit& h) {
  const irt_object& ob = sc.objects[obj];
  const float4 lo4 = mat_mul(ob.inverse, o.x, o.y, o.z, 1.0f);
  const float4 ld4 = mat_mul(ob.inverse, d.x, d.y, d.z, 0.0f);
  const v3 lo = v3{lo4.x, lo4.y, lo4.z}, ld = v3{ld4.x, ld4.y, ld4.z};
  const v3 linv = v3{slab_rcp(ld.x), slab_rcp(ld.y), slab_rcp(ld.z)};
  const uint2 range = sc.mesh_nodes[ob.info[0]];
  uint32_t j = range.x;
  while (j < range.y) {
    const irt_tri_node& tn = sc.tri_nodes[j];
    if (!ray_box(lo, linv, tn.lo, tn.hi, RAY_TMIN, h.t)) {
      j = tn.count ? j + 1u : tn.link;
      continue;
    }
    for (uint32_t k = 0; k < tn.count; ++k) ray_triangle(lo, ld, sc.tris[tn.link + k], obj, h);
    ++j;
  }
}

// rayQueryInitialize/Proceed/GetCommittedIntersection of vpl.wgsl:17-20 for
// a group of G lanes that share one ray: lane k takes objects k, k+G, ...
// (each one's box is its leaf box in the object hierarchy), walks their
// triangles, and the group keeps the closest hit by the tie rule above, so
// every lane returns the same hit (object == 0xffffffff: none).
template <int G>
__device__ __forceinline__ RayHit trace_group(const TriScene& sc, const float4* boxes, uint32_t n_objects, v3 o,
                                              v3 d, uint32_t lane) {
  RayHit h{RAY_TMAX, 0.0f, 0.0f, 0xffffffffu, 0xffffffffu};
  const v3 inv = v3{slab_rcp(d.x), slab_rcp(d.y), slab_rcp(d.z)};
  for (uint32_t obj = lane; obj < n_objects; obj += G) {
    const float4 blo = boxes[2 * obj], bhi = boxes[2 * obj + 1];
    const float lo[3] = {blo.x, blo.y, blo.z}, hi[3] = {bhi.x, bhi.y, bhi.z};
    if (ray_box(o, inv, lo, hi, RAY_TMIN, h.t)) trace_object(sc, obj, o, d, h);
  }
#pragma unroll
  for (int off = 1; off < G; off <<= 1) {
    RayHit q;
    q.t = __shfl_xor(h.t, off, G);
    q.u = __shfl_xor(h.u, off, G);
    q.v = __shfl_xor(h.v, off, G);
    q.object = (uint32_t)__shfl_xor((int)h.object, off, G);
    q.prim = (uint32_t)__shfl_xor((int)h.prim, off, G);
    if (closer(q.t, q.object, q.prim, h)) h = q;
  }
  return h;
}


// basis — common.wgsl:84: mat3x3(t, cross(n, t), n) * v.
__device__ __forceinline__ v3 basis_mul(v3 n, v3 v) {
  const v3 axis = fabsf(n.z) > 0.9f ? v3{1.0f, 0.0f, 0.0f} : v3{0.0f, 0.0f, 1.0f};
  const v3 t = normalize(cross(axis, n));
  const v3 b = cross(n, t);
  return v3{t.x * v.x + b.x * v.y + n.x * v.z, t.y * v.x + b.y * v.y + n.y * v.z, t.z * v.x + b.z * v.y + n.z * v.z};
}

// cosine_direction — common.wgsl:85.
__device__ __forceinline__ v3 cosine_direction(v3 n, uint32_t& rng) {
  const float u = rnd(rng);
  const float phi = 2.0f * IRT_PI * rnd(rng);
  float sn, cs;
  shared_sincos(phi, &sn, &cs);
  return basis_mul(n, v3{sqrtf(u) * cs, sqrtf(u) * sn, sqrtf(1.0f - u)});
}

struct Scatter {
  v3 direction, weight;
};

// sample_bsdf — common.wgsl:86-97, for a surface whose wo terms
// surface_toward() filled.  pow(rough, 4) is the (rough^2)^2 of those terms.
__device__ __forceinline__ Scatter sample_bsdf(const Surface& s, uint32_t& rng) {
  const v3 wo = s.wo;
  const float spec_probability = 0.25f * (1.0f - s.metal) + 0.9f * s.metal;  // mix(.25, .9, metal)
  v3 wi;
  if (rnd(rng) < spec_probability) {
    const float a2 = s.a2;
    const float u = rnd(rng);
    const float cos_theta = sqrtf((1.0f - u) / (1.0f + (a2 - 1.0f) * u));
    const float phi = 2.0f * IRT_PI * rnd(rng);
    float sn, cs;
    shared_sincos(phi, &sn, &cs);
    const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
    const v3 h = basis_mul(s.n, v3{sin_theta * cs, sin_theta * sn, cos_theta});
    const v3 e = -wo;  // reflect(e, h) = e - 2 dot(h, e) h
    const float k = 2.0f * dot(h, e);
    wi = e - h * k;
  } else {
    wi = cosine_direction(s.n, rng);
  }
  const float nl = fmaxf(dot(s.n, wi), 0.0f);
  if (nl <= 0.0f) return Scatter{wi, v3{0.0f, 0.0f, 0.0f}};
  const v3 h = normalize(wo + wi);
  const float nh = fmaxf(dot(s.n, h), 0.0f);
  const float a2 = s.a2;
  const float q = nh * nh * (a2 - 1.0f) + 1.0f;
  const float diffuse_pdf = nl / IRT_PI;
  const float spec_pdf = a2 * nh / (IRT_PI * q * q * fmaxf(4.0f * fabsf(dot(wo, h)), 0.00001f));
  const float pdf = diffuse_pdf * (1.0f - spec_probability) + spec_pdf * spec_probability;
  const v3 b = brdf(s, wi);
  const float den = fmaxf(pdf, 0.00001f);
  return Scatter{wi, v3{b.x * nl / den, b.y * nl / den, b.z * nl / den}};
}

// generate — vpl.wgsl:2-47.  G lanes per path id < frame.counts.y: every
// lane of a group evaluates the path's (identical) sampling and shading,
// the group shares each ray's object tests, and the group's first lane
// writes the record.  Whole groups leave together.
template <int G>
__global__ __launch_bounds__(256) void k_generate_vpls(TriScene sc, irt_frame F, const float4* __restrict__ boxes,
                                                       uint32_t n_objects, irt_vpl* __restrict__ pool) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t id = t / G, lane = t % G;
  const uint32_t V = F.counts[1];
  if (id >= V) return;
  uint32_t rng = seed(id, 7u, F.size[2]);
  irt_vpl v;
  for (int k = 0; k < 4; ++k) v.position[k] = v.normal[k] = v.flux[k] = 0.0f;
  const uint32_t L = F.counts[0];
  if (L == 0u) {
    if (lane == 0) pool[id] = v;
    return;
  }
  const uint32_t light_id = min(f2u_sat(rnd(rng) * (float)L), L - 1u);
  const irt_light& l = sc.lights[light_id];
  v3 origin = ld3(l.position_strength);
  v3 direction;
  float omega = 4.0f * IRT_PI;
  if (l.color_type[3] > 0.5f) {
    const float ch = shared_cos(l.direction_angle[3] * 0.5f);
    omega = 2.0f * IRT_PI * (1.0f - ch);
    const float r = rnd(rng);
    const float z = ch * (1.0f - r) + 1.0f * r;  // mix(cos(angle/2), 1, r)
    const float phi = 2.0f * IRT_PI * rnd(rng);
    float sn, cs;
    shared_sincos(phi, &sn, &cs);
    const float sz = sqrtf(1.0f - z * z);
    direction = basis_mul(normalize(ld3(l.direction_angle)), v3{sz * cs, sz * sn, z});
  } else {
    const float z = 2.0f * rnd(rng) - 1.0f;
    const float phi = 2.0f * IRT_PI * rnd(rng);
    float sn, cs;
    shared_sincos(phi, &sn, &cs);
    const float sz = sqrtf(1.0f - z * z);
    direction = v3{sz * cs, sz * sn, z};
  }
  // source power and path-budget normalisation stay out of the throughput
  const float Lf = (float)L;
  const v3 source_flux = v3{l.color_type[0] * l.position_strength[3] * omega * Lf,
                            l.color_type[1] * l.position_strength[3] * omega * Lf,
                            l.color_type[2] * l.position_strength[3] * omega * Lf};
  v3 throughput = v3{1.0f, 1.0f, 1.0f};
  const float Vf = (float)V;
  for (uint32_t bounce = 0; bounce < F.counts[2]; ++bounce) {
    const RayHit hit = trace_group<G>(sc, boxes, n_objects, origin, direction, lane);
    if (hit.object == 0xffffffffu) {
      v.flux[0] = v.flux[1] = v.flux[2] = v.flux[3] = 0.0f;
      break;
    }
    const irt_object& ob = sc.objects[hit.object];
    const irt_mesh& me = sc.meshes[ob.info[0]];
    const uint32_t base = me.geometry[2] + hit.prim * 3u;
    const irt_vertex& a = sc.vertices[me.geometry[0] + sc.indices[base]];
    const irt_vertex& b = sc.vertices[me.geometry[0] + sc.indices[base + 1u]];
    const irt_vertex& c = sc.vertices[me.geometry[0] + sc.indices[base + 2u]];
    const float bx = 1.0f - hit.u - hit.v, by = hit.u, bz = hit.v;
    const v3 ln = v3{a.normal[0] * bx + b.normal[0] * by + c.normal[0] * bz,
                     a.normal[1] * bx + b.normal[1] * by + c.normal[1] * bz,
                     a.normal[2] * bx + b.normal[2] * by + c.normal[2] * bz};
    const float4 wn = mat_mul(ob.model, ln.x, ln.y, ln.z, 0.0f);
    const v3 n = normalize(v3{wn.x, wn.y, wn.z});
    const irt_material& m = sc.materials[ob.info[1]];
    Surface s;
    float tu = 0.0f, tv = 0.0f;
    float4 tangent = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    if (textured(m)) {  // uv and tangent only feed the textures (vpl.wgsl:27-29)
      tu = a.uv[0] * bx + b.uv[0] * by + c.uv[0] * bz;
      tv = a.uv[1] * bx + b.uv[1] * by + c.uv[1] * bz;
      const float4 t4 = make_float4(a.tangent[0] * bx + b.tangent[0] * by + c.tangent[0] * bz,
                                    a.tangent[1] * bx + b.tangent[1] * by + c.tangent[1] * bz,
                                    a.tangent[2] * bx + b.tangent[2] * by + c.tangent[2] * bz,
                                    a.tangent[3] * bx + b.tangent[3] * by + c.tangent[3] * bz);
      const float4 wt = mat_mul(ob.model, t4.x, t4.y, t4.z, 0.0f);
      tangent = make_float4(wt.x, wt.y, wt.z, t4.w >= 0.0f ? 1.0f : -1.0f);
    }
    material_at(m, sc.tex, tu, tv, n, tangent, s);
    s.p = origin + direction * hit.t;
    const v3 wo = -direction;
    if (dot(s.n, wo) <= 0.0f) {
      v.flux[0] = v.flux[1] = v.flux[2] = v.flux[3] = 0.0f;
      break;
    }
    v.position[0] = s.p.x;
    v.position[1] = s.p.y;
    v.position[2] = s.p.z;
    v.position[3] = 1.0f;
    v.normal[0] = s.n.x;
    v.normal[1] = s.n.y;
    v.normal[2] = s.n.z;
    v.normal[3] = 0.0f;
    surface_toward(s, wo);
    const Scatter sc_ = sample_bsdf(s, rng);
    const v3 reflected = throughput * sc_.weight;
    v.flux[0] = source_flux.x * reflected.x / Vf;
    v.flux[1] = source_flux.y * reflected.y / Vf;
    v.flux[2] = source_flux.z * reflected.z / Vf;
    v.flux[3] = 0.0f;
    if (bounce + 1u >= F.counts[2]) break;
    const float survive = clampf(fmaxf(reflected.x, fmaxf(reflected.y, reflected.z)), F.rr[0], F.rr[1]);
    if (rnd(rng) > survive) break;
    if (luminance(sc_.weight) <= 0.0f) break;
    throughput = reflected / survive;
    origin = s.p + s.n * 0.001f;
    direction = sc_.direction;
  }
  if (lane == 0) pool[id] = v;
}

}  // namespace irt
