// CPU restatement of the reference WGSL (test infrastructure; see oracle.h).
//
// Every function follows one WGSL function of the reference line by line;
// the file:line is given beside it.  Scalar code, no vector intrinsics, built
// with -ffp-contract=off so FP32 arithmetic is one IEEE rounding per WGSL
// operation in source order.
#include "oracle.h"

#include <omp.h>

#include <cmath>
#include <cstring>
#include <vector>

namespace {

void shared_sincos(float x, float* sn, float* cs);  // defined with the VPL pass below

template <class R>
struct Vec {
  R x, y, z;
};
template <class R>
Vec<R> V(R x, R y, R z) {
  return Vec<R>{x, y, z};
}
template <class R>
Vec<R> add3(Vec<R> a, Vec<R> b) { return V<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R>
Vec<R> sub3(Vec<R> a, Vec<R> b) { return V<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R>
Vec<R> mul3(Vec<R> a, Vec<R> b) { return V<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R>
Vec<R> scale3(Vec<R> a, R s) { return V<R>(a.x * s, a.y * s, a.z * s); }
template <class R>
Vec<R> div3(Vec<R> a, R s) { return V<R>(a.x / s, a.y / s, a.z / s); }
template <class R>
R dot3(Vec<R> a, Vec<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R>
R length3(Vec<R> a) { return std::sqrt(dot3(a, a)); }
template <class R>
Vec<R> normalize3(Vec<R> a) { return div3(a, length3(a)); }  // WGSL: e / length(e)
template <class R>
R maxr(R a, R b) { return a > b ? a : b; }  // operands never NaN where used
// pow(x, 5) / pow(x, 64) for x >= 0: the FP32 instance evaluates the fixed
// multiply chains (x^2)^2*x and six squarings; the FP64 instance uses pow().
inline float pow5(float x) {
  const float x2 = x * x;
  return x2 * x2 * x;
}
inline double pow5(double x) { return std::pow(x, 5.0); }
inline float pow64(float x) {
  float y = x * x;
  y = y * y;
  y = y * y;
  y = y * y;
  y = y * y;
  return y * y;
}
inline double pow64(double x) { return std::pow(x, 64.0); }
template <class R>
R minr(R a, R b) { return a < b ? a : b; }
template <class R>
Vec<R> max3(Vec<R> a, Vec<R> b) { return V<R>(maxr(a.x, b.x), maxr(a.y, b.y), maxr(a.z, b.z)); }
template <class R>
Vec<R> min3(Vec<R> a, Vec<R> b) { return V<R>(minr(a.x, b.x), minr(a.y, b.y), minr(a.z, b.z)); }
template <class R>
R clampr(R e, R lo, R hi) { return minr(maxr(e, lo), hi); }  // WGSL clamp = min(max(e, lo), hi)
template <class R>
Vec<R> F3(const float* p) { return V<R>((R)p[0], (R)p[1], (R)p[2]); }
template <class R>
Vec<R> D3(const double* p) { return V<R>((R)p[0], (R)p[1], (R)p[2]); }

// WGSL saturating conversions.
template <class R>
uint32_t to_u32(R x) {
  if (!(x > R(0))) return 0u;
  if (x >= R(4294967295.0)) return 4294967295u;
  return (uint32_t)x;
}
template <class R>
int32_t to_i32(R x) {
  if (x != x) return 0;
  if (x >= R(2147483647.0)) return 2147483647;
  if (x <= R(-2147483648.0)) return (int32_t)(-2147483647 - 1);
  return (int32_t)x;
}

uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7fffffffu;
  if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u);
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);
  if (ax < 0x38800000u) {
    if (ax < 0x33000000u) return (uint16_t)sign;
    const uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    const int shift = 126 - (int)(ax >> 23);
    uint32_t q = mant >> shift;
    const uint32_t rem = mant & ((1u << shift) - 1u), half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return (uint16_t)(sign | q);
  }
  uint32_t q = (((ax >> 23) - 112u) << 10) | ((ax & 0x7fffffu) >> 13);
  const uint32_t rem = ax & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
  return (uint16_t)(sign | q);
}
float f16_to_f32(uint16_t h) {
  const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  float v;
  if (e == 0)
    v = std::ldexp((float)m, -24);
  else if (e == 31)
    v = m ? NAN : INFINITY;
  else
    v = std::ldexp((float)(m | 0x400u), (int)e - 25);
  return (h & 0x8000u) ? -v : v;
}

// Storage rounding of an Rgba16F target: binary16 for FP32, none for FP64.
inline float store16(float v) { return f16_to_f32(f32_to_f16(v)); }
inline double store16(double v) { return v; }

// random — common.wgsl:39-43 (the FP32 value, also used by the FP64 instance).
float random_f32(uint32_t* s) {
  *s = *s * 747796405u + 2891336453u;
  const uint32_t w = ((*s >> ((*s >> 28u) + 4u)) ^ *s) * 277803737u;
  return (float)((w >> 22u) ^ w) * (1.0f / 4294967296.0f);
}
// seed — common.wgsl:44.
uint32_t seed_of(uint32_t i, uint32_t salt, uint32_t frame) { return i * 1973u + frame * 9277u + salt * 26699u + 89173u; }

template <class R>
struct Res {
  uint32_t id, source, count, age;
  R sum, selected_target, pdf;
  uint32_t generation;
};
template <class R>
Res<R> from_io(const oracle_reservoir& o) {
  return Res<R>{o.id, o.source, o.count, o.age, (R)o.sum, (R)o.selected_target, (R)o.pdf, o.generation};
}
template <class R>
oracle_reservoir to_io(const Res<R>& r) {
  oracle_reservoir o;
  o.id = r.id;
  o.source = r.source;
  o.count = r.count;
  o.age = r.age;
  o.sum = (double)r.sum;
  o.selected_target = (double)r.selected_target;
  o.pdf = (double)r.pdf;
  o.generation = r.generation;
  o.pad = 0;
  return o;
}

template <class R>
struct Surface {
  Vec<R> p, n, color;
  R rough, metal;
};

template <class R>
struct Oracle {
  const irt_frame& F;
  const oracle_scene* sc;
  const oracle_gbuffer* gb;
  oracle_io* io;
  uint64_t cnt[7] = {0, 0, 0, 0, 0, 0, 0};
  const R PI = R(3.14159265359f);  // const PI: f32 (common.wgsl:2)

  Oracle(const irt_frame& f, const oracle_scene* s, const oracle_gbuffer* g, oracle_io* i) : F(f), sc(s), gb(g), io(i) {}

  uint32_t W() const { return F.size[0]; }
  uint32_t H() const { return F.size[1]; }
  R random(uint32_t* s) { return (R)random_f32(s); }

  // mat4 * vec4, column-major, row by row in column order.
  void mat_vec(const float* m, R x, R y, R z, R w, R out[4]) const {
    for (int r = 0; r < 4; ++r) out[r] = (R)m[r] * x + (R)m[4 + r] * y + (R)m[8 + r] * z + (R)m[12 + r] * w;
  }
  R luminance(Vec<R> c) const { return c.x * R(0.2126f) + c.y * R(0.7152f) + c.z * R(0.0722f); }  // common.wgsl:45
  Vec<R> world(R u, R v, R depth, const float* m) const {  // common.wgsl:46
    R h[4];
    mat_vec(m, u * R(2) - R(1), R(1) - v * R(2), depth, R(1), h);
    return V<R>(h[0] / h[3], h[1] / h[3], h[2] / h[3]);
  }
  R pixel_u(int64_t x) const { return ((R)x + R(0.5f)) / (R)W(); }  // common.wgsl:47
  R pixel_v(int64_t y) const { return ((R)y + R(0.5f)) / (R)H(); }
  Vec<R> camera() const { return F3<R>(F.camera); }

  // brdf — common.wgsl:67-74.
  Vec<R> brdf(const Surface<R>& s, Vec<R> wo, Vec<R> wi) {
    cnt[0]++;
    const R nv = maxr(dot3(s.n, wo), R(0.00001f));
    const R nl = maxr(dot3(s.n, wi), R(0));
    if (nl <= R(0)) return V<R>(0, 0, 0);
    const Vec<R> h = normalize3(add3(wo, wi));
    const R nh = maxr(dot3(s.n, h), R(0));
    const R vh = maxr(dot3(wo, h), R(0));
    const R a = s.rough * s.rough;
    const R a2 = a * a;
    const R q = nh * nh * (a2 - R(1)) + R(1);
    const R d = a2 / (PI * q * q);
    const R g1v = R(2) * nv / (nv + std::sqrt(a2 + (R(1) - a2) * nv * nv));
    const R g1l = R(2) * nl / (nl + std::sqrt(a2 + (R(1) - a2) * nl * nl));
    // mix(vec3(.04), color, metal) = .04*(1-metal) + color*metal
    const R c04 = R(0.04f);
    const Vec<R> f0 = V<R>(c04 * (R(1) - s.metal) + s.color.x * s.metal, c04 * (R(1) - s.metal) + s.color.y * s.metal,
                           c04 * (R(1) - s.metal) + s.color.z * s.metal);
    const R p5 = pow5(R(1) - vh);
    const Vec<R> f = V<R>(f0.x + (R(1) - f0.x) * p5, f0.y + (R(1) - f0.y) * p5, f0.z + (R(1) - f0.z) * p5);
    const R den = maxr(R(4) * nv * nl, R(0.00001f));
    Vec<R> out;
    const R fs[3] = {f.x, f.y, f.z}, cs[3] = {s.color.x, s.color.y, s.color.z};
    R o[3];
    for (int k = 0; k < 3; ++k) o[k] = (R(1) - fs[k]) * (R(1) - s.metal) * cs[k] / PI + fs[k] * d * g1v * g1l / den;
    out = V<R>(o[0], o[1], o[2]);
    return out;
  }
  // direct — common.wgsl:75-79.
  Vec<R> direct(const Surface<R>& s, uint32_t id) {
    const irt_light& l = sc->lights[id];
    const Vec<R> delta = sub3(F3<R>(l.position_strength), s.p);
    const R r2 = maxr(dot3(delta, delta), R(0.0001f));
    const Vec<R> wi = scale3(delta, R(1) / std::sqrt(r2));
    if ((R)l.color_type[3] > R(0.5f) &&
        dot3(scale3(wi, R(-1)), normalize3(F3<R>(l.direction_angle))) < std::cos((R)l.direction_angle[3] * R(0.5f)))
      return V<R>(0, 0, 0);
    const Vec<R> b = brdf(s, normalize3(sub3(camera(), s.p)), wi);
    const R k = maxr(dot3(s.n, wi), R(0));
    const R c[3] = {(R)l.color_type[0], (R)l.color_type[1], (R)l.color_type[2]};
    const R bb[3] = {b.x, b.y, b.z};
    R o[3];
    for (int i = 0; i < 3; ++i) o[i] = bb[i] * k * c[i] * (R)l.position_strength[3] / r2;
    return V<R>(o[0], o[1], o[2]);
  }
  // indirect — common.wgsl:80-83.
  Vec<R> indirect(const Surface<R>& s, const irt_vpl& v) {
    const Vec<R> delta = sub3(F3<R>(v.position), s.p);
    const R r2 = maxr(dot3(delta, delta), R(0.0025f));
    const Vec<R> wi = scale3(delta, R(1) / std::sqrt(r2));
    const Vec<R> b = brdf(s, normalize3(sub3(camera(), s.p)), wi);
    const R k1 = maxr(dot3(s.n, wi), R(0));
    const R k2 = maxr(dot3(F3<R>(v.normal), scale3(wi, R(-1))), R(0));
    const R den = R(4) * PI * r2;
    const R bb[3] = {b.x, b.y, b.z}, fl[3] = {(R)v.flux[0], (R)v.flux[1], (R)v.flux[2]};
    R o[3];
    for (int i = 0; i < 3; ++i) o[i] = bb[i] * k1 * k2 * fl[i] / den;
    return V<R>(o[0], o[1], o[2]);
  }

  // Reservoir helpers — common.wgsl:98-106, lighting.wgsl:51.
  Res<R> empty_reservoir() const { return Res<R>{0u, F.size[2], 0u, 0u, R(0), R(0), R(0), F.size[3]}; }
  static R normalization(const Res<R>& r) {
    if (r.count == 0u || r.selected_target <= R(0)) return R(0);
    return r.sum / ((R)(float)r.count * r.selected_target);
  }
  bool add(Res<R>* r, uint32_t id, uint32_t source, R target, R pdf, R weight, uint32_t count, uint32_t age,
           uint32_t* rng) {
    r->count += count;
    if (weight <= R(0) || weight > R(1e30f)) return false;  // NaN passes, as in WGSL
    r->sum += weight;
    if (random(rng) * r->sum < weight) {
      r->id = id;
      r->source = source;
      r->selected_target = target;
      r->pdf = pdf;
      r->age = age;
      return true;
    }
    return false;
  }
  static Res<R> cap(Res<R> a, uint32_t limit) {
    if (a.count > limit) {
      a.sum *= (R)(float)limit / (R)(float)a.count;
      a.count = limit;
    }
    return a;
  }

  // ------------------------------------------------------------- SDF
  static R signed_box_distance(Vec<R> p, Vec<R> lo, Vec<R> hi) {  // sdf.wgsl:1-4
    const Vec<R> delta = max3(sub3(lo, p), sub3(p, hi));
    return length3(max3(delta, V<R>(0, 0, 0))) + minr(maxr(delta.x, maxr(delta.y, delta.z)), R(0));
  }
  // sample_field — common.wgsl:31-38 (manual 8-tap variant).
  R sample_field(const uint32_t origin[3], Vec<R> q, uint32_t res) const {
    const uint32_t cell[3] = {minr(to_u32(std::floor(q.x)), res - 2u), minr(to_u32(std::floor(q.y)), res - 2u),
                              minr(to_u32(std::floor(q.z)), res - 2u)};
    const R f[3] = {q.x - (R)(float)cell[0], q.y - (R)(float)cell[1], q.z - (R)(float)cell[2]};
    R value = R(0);
    for (uint32_t z = 0; z < 2; ++z)
      for (uint32_t y = 0; y < 2; ++y)
        for (uint32_t x = 0; x < 2; ++x) {
          const R wx = x == 1 ? f[0] : R(1) - f[0];
          const R wy = y == 1 ? f[1] : R(1) - f[1];
          const R wz = z == 1 ? f[2] : R(1) - f[2];
          const size_t tx = origin[0] + cell[0] + x, ty = origin[1] + cell[1] + y, tz = origin[2] + cell[2] + z;
          const R t = (R)sc->atlas[tx + (size_t)sc->atlas_w * (ty + (size_t)sc->atlas_h * tz)];
          value += t * wx * wy * wz;
        }
    return value;
  }
  // mesh_distance — sdf.wgsl:5-33.
  R mesh_distance(Vec<R> p, const irt_mesh& mesh) {
    const R geometry_bound = signed_box_distance(p, F3<R>(mesh.geometry_lo), F3<R>(mesh.geometry_hi));
    if ((R)mesh.geometry_lo[3] > R(0.5f)) {
      cnt[2]++;
      return geometry_bound;
    }
    cnt[3]++;
    const uint32_t res = mesh.field[3];
    const Vec<R> lo = F3<R>(mesh.lo), hi = F3<R>(mesh.hi);
    const Vec<R> clamped = min3(max3(p, lo), hi);  // clamp(p, lo, hi)
    const Vec<R> q = mul3(sub3(clamped, lo), F3<R>(mesh.grid));
    const R value = sample_field(mesh.field, q, res);
    R bound = value - (R)mesh.grid[3] * (R(0.5f) + R(1.0f / 128.0f));
    const Vec<R> outside = max3(max3(sub3(lo, p), sub3(p, hi)), V<R>(0, 0, 0));
    const R outside_squared = dot3(outside, outside);
    if (outside_squared > R(0)) {
      const R positive = maxr(bound, R(0));
      bound = std::sqrt(positive * positive + outside_squared);
    }
    if (bound > R(0)) bound *= (R)mesh.geometry_hi[3];
    const R sphere_bound = length3(sub3(p, F3<R>(mesh.enclosing_sphere))) - (R)mesh.enclosing_sphere[3];
    return maxr(bound, maxr(geometry_bound, sphere_bound));
  }
  // object_distance — sdf.wgsl:35-41.
  void object_distance(Vec<R> p, uint32_t id, bool need_spacing, R* d_out, R* voxel_out) {
    const irt_object& obj = sc->objects[id];
    const irt_mesh& mesh = sc->meshes[obj.info[0]];
    R l[4];
    mat_vec(obj.inverse, p.x, p.y, p.z, R(1), l);
    float scale, voxel;
    std::memcpy(&scale, &obj.info[2], 4);
    std::memcpy(&voxel, &obj.info[3], 4);
    R d = mesh_distance(V<R>(l[0], l[1], l[2]), mesh) * (R)scale;
    if (d > R(0)) d *= R(0.9999f);
    *d_out = d;
    *voxel_out = need_spacing ? (R)voxel : (R)F.sdf[1];
  }
  // scene_distance — sdf.wgsl:42-60.
  void scene_distance(Vec<R> p, bool need_spacing, uint32_t hint, R* best_out, R* voxel_out, uint32_t* closest_out) {
    R best = R(1e20f), voxel = (R)F.sdf[1];
    uint32_t closest = hint, cursor = 0u;
    const uint32_t n = F.counts[3];
    if (hint < n) object_distance(p, hint, need_spacing, &best, &voxel);
    while (cursor < n * 2u - 1u) {
      const irt_node& node = sc->nodes[cursor];
      cnt[1]++;
      const Vec<R> delta = max3(max3(sub3(F3<R>(node.lo), p), sub3(p, F3<R>(node.hi))), V<R>(0, 0, 0));
      const Vec<R> lower_bound = scale3(delta, R(0.9998f));
      if (dot3(lower_bound, lower_bound) > maxr(best, R(0)) * maxr(best, R(0))) {
        cursor = node.info[3];
        continue;
      }
      if (node.info[2] == 1u) {
        if (node.info[0] != hint) {
          R d, vx;
          object_distance(p, node.info[0], need_spacing, &d, &vx);
          if (d < best) {
            best = d;
            voxel = vx;
            closest = node.info[0];
          }
        }
      }
      cursor++;
    }
    *best_out = best;
    *voxel_out = voxel;
    *closest_out = closest;
  }
  // visibility — sdf.wgsl:61-81.
  R visibility(Vec<R> p, Vec<R> n, Vec<R> target, Vec<R> end_normal) {
    cnt[6]++;
    const R eps = (R)F.sdf[1];
    R cd, cv;
    uint32_t closest;
    scene_distance(p, true, 0xffffffffu, &cd, &cv, &closest);
    const R epsilon = maxr(eps, cv * R(0.65f));
    const Vec<R> origin = add3(p, scale3(n, epsilon));
    Vec<R> endpoint = target;
    if (dot3(end_normal, end_normal) > R(0.5f)) {
      R ed, ev;
      uint32_t ec;
      scene_distance(target, true, 0xffffffffu, &ed, &ev, &ec);
      endpoint = add3(endpoint, scale3(end_normal, maxr(eps, ev * R(0.65f))));
    }
    const Vec<R> delta = sub3(endpoint, origin);
    const R limit = length3(delta) - eps;
    if (limit <= R(0)) return R(1);
    const Vec<R> direction = normalize3(delta);
    R t = eps, result = R(1), previous = R(1e20f);
    const uint32_t steps = to_u32((R)F.sdf[0]);
    for (uint32_t i = 0; i < steps; ++i) {
      if (t >= limit) return clampr(result, R(0), R(1));
      cnt[4]++;
      R h, hv;
      scene_distance(add3(origin, scale3(direction, t)), false, closest, &h, &hv, &closest);
      if (h < eps) return R(0);
      result = shadow_step(result, h, previous, t, (R)F.sdf[2]);
      if (result < R(0.001f)) return R(0);
      previous = h;
      t += h;
    }
    return R(0);
  }
  static R shadow_step(R result, R h, R previous, R t, R k) {  // sdf.wgsl:180-181
    const R y = h * h / (R(2) * previous);
    const R d = std::sqrt(maxr(h * h - y * y, R(0)));
    if (t > y) result = minr(result, k * d / (t - y));
    return result;
  }

  // ---------------------------------------------------------- lighting
  size_t idx(int64_t x, int64_t y) const { return (size_t)y * W() + (size_t)x; }
  bool valid(int64_t x, int64_t y) const { return (R)gb->depth[idx(x, y)] < R(1); }  // lighting.wgsl:35
  Surface<R> surface(int64_t x, int64_t y) const {                                    // lighting.wgsl:30-34
    const size_t i = idx(x, y);
    Surface<R> s;
    s.p = world(pixel_u(x), pixel_v(y), (R)gb->depth[i], F.inv_vp);
    s.n = normalize3(D3<R>(gb->normal + 4 * i));
    s.color = D3<R>(gb->color + 4 * i);
    s.rough = (R)gb->params[2 * i];
    s.metal = (R)gb->params[2 * i + 1];
    return s;
  }
  bool compatible(const Surface<R>& a, int64_t qx, int64_t qy) const {  // lighting.wgsl:36-40
    const size_t q = idx(qx, qy);
    const R d = (R)gb->depth[q];
    if (d >= R(1)) return false;
    const Vec<R> p = world(pixel_u(qx), pixel_v(qy), d, F.inv_vp);
    const Vec<R> n = normalize3(D3<R>(gb->normal + 4 * q));
    return dot3(a.n, n) > (R)F.temporal[2] &&
           length3(sub3(a.p, p)) < (R)F.temporal[1] * maxr(length3(sub3(a.p, camera())), R(1));
  }
  void reproject(int64_t x, int64_t y, int32_t* qx, int32_t* qy) const {  // lighting.wgsl:41-44
    const size_t i = idx(x, y);
    const R u = pixel_u(x) - (R)gb->motion[2 * i], v = pixel_v(y) - (R)gb->motion[2 * i + 1];
    *qx = to_i32(std::floor(u * (R)W()));
    *qy = to_i32(std::floor(v * (R)H()));
  }
  bool history_valid(int64_t x, int64_t y, int32_t qx, int32_t qy) const {  // lighting.wgsl:45-50
    if (F.flags[0] == 0u || F.flags[1] != F.size[3] || qx < 0 || qy < 0 || qx >= (int64_t)W() || qy >= (int64_t)H())
      return false;
    const size_t q = idx(qx, qy);
    const R d = (R)gb->previous_depth[q];
    if (d >= R(1)) return false;
    const Surface<R> a = surface(x, y);
    const Vec<R> oldp = world(pixel_u(qx), pixel_v(qy), d, F.previous_inv_vp);
    return dot3(a.n, D3<R>(gb->previous_normal + 4 * q)) > (R)F.temporal[2] &&
           length3(sub3(a.p, oldp)) < (R)F.temporal[1] * maxr(length3(sub3(a.p, camera())), R(1));
  }
  bool vpl_sample_valid(const Res<R>& r) const {  // lighting.wgsl:22-24
    return r.generation == F.size[3] && r.selected_target > R(0) && r.id < F.counts[1] &&
           F.size[2] - r.source <= to_u32((R)F.rr[2]);
  }
  const irt_vpl& selected_vpl(const Res<R>& r) const {  // lighting.wgsl:25-29
    if (r.source == F.size[2]) return io->current_vpls[r.id];
    if (r.source + 1u == F.size[2]) return io->previous_vpls[r.id];
    return io->archived_vpls[(size_t)(r.source % (to_u32((R)F.rr[2]) + 1u)) * F.counts[1] + r.id];
  }
  // lighting.wgsl:86-89: angle 2*PI*u, radius sqrt(u')*spatial_radius, round() half to even
  static void neighbor_offset(uint32_t* rng, uint32_t spatial_radius, int64_t* dx, int64_t* dy) {
    const R two_pi = (R)(2.0f * 3.14159265359f);  // 2.*PI folded in f32
    const R angle = two_pi * (R)random_f32(rng);
    const R radius = std::sqrt((R)random_f32(rng)) * (R)(float)spatial_radius;
    R cs, sn;
    if constexpr (sizeof(R) == sizeof(float)) {
      shared_sincos(angle, &sn, &cs);  // the device's float sin/cos
    } else {
      cs = std::cos(angle);
      sn = std::sin(angle);
    }
    *dx = to_i32(std::nearbyint(cs * radius));
    *dy = to_i32(std::nearbyint(sn * radius));
  }
  void neighbor(int64_t x, int64_t y, uint32_t* rng, int64_t* qx, int64_t* qy) {  // lighting.wgsl:86-89
    int64_t dx, dy;
    neighbor_offset(rng, F.restir[3], &dx, &dy);
    *qx = x + dx;
    *qy = y + dy;
  }

  void candidates_direct(int64_t x, int64_t y) {  // lighting.wgsl:52-58
    const uint32_t id = (uint32_t)idx(x, y);
    Res<R> r = empty_reservoir();
    uint32_t rng = seed_of(id, 11u, F.size[2]);
    if (valid(x, y) && F.counts[0] > 0u) {
      const Surface<R> s = surface(x, y);
      const R pdf = R(1) / (R)(float)F.counts[0];
      for (uint32_t i = 0; i < F.restir[0]; ++i) {
        const uint32_t light_id = minr(to_u32(random(&rng) * (R)(float)F.counts[0]), F.counts[0] - 1u);
        const R target = luminance(direct(s, light_id));
        add(&r, light_id, F.size[2], target, pdf, target / pdf, 1u, 0u, &rng);
      }
    }
    io->direct_out[id] = to_io(r);
  }
  void candidates_indirect(int64_t x, int64_t y) {  // lighting.wgsl:59-65
    const uint32_t id = (uint32_t)idx(x, y);
    Res<R> r = empty_reservoir();
    uint32_t rng = seed_of(id, 13u, F.size[2]);
    if (valid(x, y)) {
      const Surface<R> s = surface(x, y);
      const R pdf = R(1) / (R)(float)F.counts[1];
      for (uint32_t i = 0; i < F.restir[1]; ++i) {
        const uint32_t vpl_id = minr(to_u32(random(&rng) * (R)(float)F.counts[1]), F.counts[1] - 1u);
        const R target = luminance(indirect(s, io->current_vpls[vpl_id]));
        add(&r, vpl_id, F.size[2], target, pdf, target / pdf, 1u, 0u, &rng);
      }
    }
    io->indirect_out[id] = to_io(r);
  }
  void temporal_direct(int64_t x, int64_t y) {  // lighting.wgsl:66-74
    const uint32_t id = (uint32_t)idx(x, y);
    Res<R> r = from_io<R>(io->direct_in[id]);
    int32_t qx, qy;
    reproject(x, y, &qx, &qy);
    uint32_t rng = seed_of(id, 17u, F.size[2]);
    if (valid(x, y) && history_valid(x, y, qx, qy)) {
      const uint32_t hist = to_u32((R)F.rr[2]);
      const Res<R> previous = cap(from_io<R>(io->direct_history[idx(qx, qy)]), hist * F.restir[0]);
      if (previous.generation == F.size[3] && previous.selected_target > R(0) && previous.id < F.counts[0] &&
          previous.age < hist) {
        const R target = luminance(direct(surface(x, y), previous.id));
        add(&r, previous.id, previous.source, target, previous.pdf,
            target * normalization(previous) * (R)(float)previous.count, previous.count, previous.age + 1u, &rng);
      }
    }
    io->direct_out[id] = to_io(r);
  }
  void temporal_indirect(int64_t x, int64_t y) {  // lighting.wgsl:75-85
    const uint32_t id = (uint32_t)idx(x, y);
    Res<R> r = from_io<R>(io->indirect_in[id]);
    int32_t qx, qy;
    reproject(x, y, &qx, &qy);
    uint32_t rng = seed_of(id, 19u, F.size[2]);
    if (valid(x, y) && history_valid(x, y, qx, qy)) {
      const Res<R> previous = cap(from_io<R>(io->indirect_history[idx(qx, qy)]), to_u32((R)F.rr[2]) * F.restir[1]);
      if (vpl_sample_valid(previous)) {
        const irt_vpl& v = selected_vpl(previous);
        const R target = luminance(indirect(surface(x, y), v));
        add(&r, previous.id, previous.source, target, previous.pdf,
            target * normalization(previous) * (R)(float)previous.count, previous.count, previous.age + 1u, &rng);
      }
    }
    io->indirect_out[id] = to_io(r);
  }
  void spatial_direct(int64_t x, int64_t y) {  // lighting.wgsl:90-97
    const uint32_t id = (uint32_t)idx(x, y);
    Res<R> r = from_io<R>(io->direct_in[id]);
    uint32_t rng = seed_of(id, 23u, F.size[2]);
    if (valid(x, y)) {
      const Surface<R> s = surface(x, y);
      for (uint32_t i = 0; i < F.restir[2]; ++i) {
        int64_t qx, qy;
        neighbor(x, y, &rng, &qx, &qy);
        if (qx < 0 || qy < 0 || qx >= (int64_t)W() || qy >= (int64_t)H()) continue;
        if (!compatible(s, qx, qy)) continue;
        const Res<R> previous = from_io<R>(io->direct_in[idx(qx, qy)]);
        if (previous.selected_target > R(0) && previous.id < F.counts[0]) {
          const R target = luminance(direct(s, previous.id));
          add(&r, previous.id, previous.source, target, previous.pdf,
              target * normalization(previous) * (R)(float)previous.count, previous.count, previous.age, &rng);
        }
      }
    }
    io->direct_out[id] = to_io(cap(r, to_u32((R)F.rr[2]) * F.restir[0] * 8u));
  }
  void spatial_indirect(int64_t x, int64_t y) {  // lighting.wgsl:98-105
    const uint32_t id = (uint32_t)idx(x, y);
    Res<R> r = from_io<R>(io->indirect_in[id]);
    uint32_t rng = seed_of(id, 29u, F.size[2]);
    if (valid(x, y)) {
      const Surface<R> s = surface(x, y);
      for (uint32_t i = 0; i < F.restir[2]; ++i) {
        int64_t qx, qy;
        neighbor(x, y, &rng, &qx, &qy);
        if (qx < 0 || qy < 0 || qx >= (int64_t)W() || qy >= (int64_t)H()) continue;
        if (!compatible(s, qx, qy)) continue;
        const Res<R> previous = from_io<R>(io->indirect_in[idx(qx, qy)]);
        if (vpl_sample_valid(previous)) {
          const R target = luminance(indirect(s, selected_vpl(previous)));
          add(&r, previous.id, previous.source, target, previous.pdf,
              target * normalization(previous) * (R)(float)previous.count, previous.count, previous.age, &rng);
        }
      }
    }
    io->indirect_out[id] = to_io(cap(r, to_u32((R)F.rr[2]) * F.restir[1] * 8u));
  }
  void resolve_direct(int64_t x, int64_t y) {  // lighting.wgsl:106-122
    const size_t id = idx(x, y);
    Vec<R> result = V<R>(0, 0, 0);
    if (valid(x, y)) {
      const Surface<R> s = surface(x, y);
      const Res<R> d = from_io<R>(io->direct_in[id]);
      if (d.selected_target > R(0) && d.id < F.counts[0] && F.flags[2] != 2u && F.flags[2] < 3u) {
        const Vec<R> value = direct(s, d.id);
        if (luminance(value) > R(0)) {
          const R w = normalization(d);
          const R vis = visibility(s.p, s.n, F3<R>(sc->lights[d.id].position_strength), V<R>(0, 0, 0));
          result = add3(result, scale3(scale3(value, w), vis));
        }
      }
      if (F.flags[2] == 3u) result = s.color;
      if (F.flags[2] == 4u) result = add3(scale3(s.n, R(0.5f)), V<R>(R(0.5f), R(0.5f), R(0.5f)));
      if (F.flags[2] == 5u) {
        const R mx = (R)gb->motion[2 * id], my = (R)gb->motion[2 * id + 1];
        result = V<R>(clampr(R(0.5f) + mx * R(50), R(0), R(1)), clampr(R(0.5f) + my * R(50), R(0), R(1)), R(0.5f));
      }
    }
    double* o = io->direct_output + 4 * id;  // Rgba32F
    o[0] = (double)result.x;
    o[1] = (double)result.y;
    o[2] = (double)result.z;
    o[3] = 1.0;
  }
  void resolve_indirect(int64_t x, int64_t y) {  // lighting.wgsl:123-134
    const size_t id = idx(x, y);
    Vec<R> result = D3<R>(io->direct_lighting + 4 * id);
    if (valid(x, y) && F.flags[2] != 1u && F.flags[2] < 3u) {
      const Res<R> v = from_io<R>(io->indirect_in[id]);
      if (vpl_sample_valid(v)) {
        const Surface<R> s = surface(x, y);
        const irt_vpl& smp = selected_vpl(v);
        const Vec<R> value = indirect(s, smp);
        if (luminance(value) > R(0)) {
          const R w = normalization(v);
          const R vis = visibility(s.p, s.n, F3<R>(smp.position), F3<R>(smp.normal));
          result = add3(result, scale3(scale3(value, w), vis));
        }
      }
    }
    const R c = R(65000);
    double* o = io->output + 4 * id;  // Rgba16F
    o[0] = (double)store16(minr(result.x, c));
    o[1] = (double)store16(minr(result.y, c));
    o[2] = (double)store16(minr(result.z, c));
    o[3] = 1.0;
  }

  // ------------------------------------------------------------- SVGF
  bool inside(int64_t x, int64_t y) const { return x >= 0 && y >= 0 && x < (int64_t)W() && y < (int64_t)H(); }
  Vec<R> position(int64_t x, int64_t y) const {  // svgf.wgsl:19
    return world(pixel_u(x), pixel_v(y), (R)gb->depth[idx(x, y)], F.inv_vp);
  }
  bool history_ok(int64_t px, int64_t py, int64_t qx, int64_t qy) const {  // svgf.wgsl:20-26
    if (F.flags[0] == 0u || F.flags[1] != F.size[3] || !inside(qx, qy)) return false;
    const size_t q = idx(qx, qy);
    const R d = (R)gb->previous_depth[q];
    if (d >= R(1)) return false;
    const Vec<R> a = position(px, py);
    const Vec<R> b = world(pixel_u(qx), pixel_v(qy), d, F.previous_inv_vp);
    return length3(sub3(a, b)) < (R)F.temporal[1] * maxr(length3(sub3(a, camera())), R(1)) &&
           dot3(D3<R>(gb->normal + 4 * idx(px, py)), D3<R>(gb->previous_normal + 4 * q)) > (R)F.temporal[2];
  }
  struct Guide {
    Vec<R> position, normal, color;
    R m0, m1, pixel_scale;
  };
  Guide guide(int64_t x, int64_t y) const {  // svgf.wgsl:27-31
    const size_t i = idx(x, y);
    Guide g;
    g.position = position(x, y);
    g.normal = D3<R>(gb->normal + 4 * i);
    g.color = D3<R>(gb->color + 4 * i);
    g.m0 = (R)gb->params[2 * i];
    g.m1 = (R)gb->params[2 * i + 1];
    g.pixel_scale = length3(sub3(g.position, camera())) / (R)(float)F.size[1];
    return g;
  }
  R geometry_weight(const Guide& a, int64_t qx, int64_t qy, R stride) {  // svgf.wgsl:32-44
    cnt[5]++;
    if (!inside(qx, qy) || (R)gb->depth[idx(qx, qy)] >= R(1)) return R(0);
    const size_t q = idx(qx, qy);
    const Vec<R> nq = D3<R>(gb->normal + 4 * q);
    const R wn = pow64(maxr(dot3(a.normal, nq), R(0)));
    const Vec<R> wq = position(qx, qy);
    const R tolerance = maxr(R(0.0001f), a.pixel_scale * stride * R(0.5f));
    const R wz = std::exp(-std::fabs(dot3(sub3(wq, a.position), a.normal)) / tolerance);
    const R dm0 = a.m0 - (R)gb->params[2 * q], dm1 = a.m1 - (R)gb->params[2 * q + 1];
    const R wc = std::exp(-length3(sub3(a.color, D3<R>(gb->color + 4 * q))) * R(8) - std::sqrt(dm0 * dm0 + dm1 * dm1) * R(8));
    return wn * wz * wc;
  }
  void svgf_temporal(int64_t x, int64_t y) {  // svgf.wgsl:45-76
    const size_t i = idx(x, y);
    const Vec<R> c = D3<R>(io->noisy + 4 * i);
    const R l = luminance(c);
    Vec<R> history = V<R>(0, 0, 0);
    R hm0 = 0, hm1 = 0, length_history = 0, total = 0;
    if ((R)gb->depth[i] < R(1)) {
      const R px = (pixel_u(x) - (R)gb->motion[2 * i]) * (R)W() - R(0.5f);
      const R py = (pixel_v(y) - (R)gb->motion[2 * i + 1]) * (R)H() - R(0.5f);
      const int32_t bx = to_i32(std::floor(px)), by = to_i32(std::floor(py));
      const R fx = px - std::floor(px), fy = py - std::floor(py);  // fract
      for (int yy = 0; yy < 2; ++yy)
        for (int xx = 0; xx < 2; ++xx) {
          const int64_t qx = (int64_t)bx + xx, qy = (int64_t)by + yy;
          if (!history_ok(x, y, qx, qy)) continue;
          const double* m = io->old_moments + 4 * idx(qx, qy);
          if ((R)m[2] <= R(0)) continue;
          const R wx = xx == 1 ? fx : R(1) - fx;
          const R wy = yy == 1 ? fy : R(1) - fy;
          const R w = wx * wy;
          history = add3(history, scale3(D3<R>(io->old_radiance + 4 * idx(qx, qy)), w));
          hm0 += (R)m[0] * w;
          hm1 += (R)m[1] * w;
          length_history += (R)m[2] * w;
          total += w;
        }
    }
    R age = R(1);
    Vec<R> result = c;
    R m0 = l, m1 = l * l;
    if (total > R(0.001f)) {
      history = div3(history, total);
      hm0 /= total;
      hm1 /= total;
      age = minr(length_history / total + R(1), (R)F.temporal[3]);
      R lo = R(1e20f), hi = R(0);
      const Guide a = guide(x, y);
      for (int yy = -1; yy <= 1; ++yy)
        for (int xx = -1; xx <= 1; ++xx) {
          const int64_t qx = clampr<int64_t>(x + xx, 0, (int64_t)W() - 1);
          const int64_t qy = clampr<int64_t>(y + yy, 0, (int64_t)H() - 1);
          if (geometry_weight(a, qx, qy, R(1)) > R(0.01f)) {
            const R v = luminance(D3<R>(io->noisy + 4 * idx(qx, qy)));
            lo = minr(lo, v);
            hi = maxr(hi, v);
          }
        }
      const R lh = luminance(history);
      history = scale3(history, clampr(lh, lo * R(0.5f), hi * R(1.5f) + R(0.001f)) / maxr(lh, R(0.000001f)));
      const R alpha = maxr(R(1) - (R)F.temporal[0], R(1) / age);
      result = add3(scale3(history, R(1) - alpha), scale3(c, alpha));  // mix(history, c, alpha)
      const R beta = maxr(R(0.2f), R(1) / age);
      m0 = hm0 * (R(1) - beta) + m0 * beta;
      m1 = hm1 * (R(1) - beta) + m1 * beta;
    }
    const R variance = maxr(m1 - m0 * m0, R(0));
    double* o = io->radiance_out + 4 * i;  // Rgba16F
    o[0] = (double)store16(result.x);
    o[1] = (double)store16(result.y);
    o[2] = (double)store16(result.z);
    o[3] = (double)store16(minr(std::sqrt(variance), R(65000)));
    double* mo = io->moments_out + 4 * i;  // Rgba32F
    mo[0] = (double)m0;
    mo[1] = (double)m1;
    mo[2] = (double)age;
    mo[3] = 0.0;
  }
  void svgf_variance(int64_t x, int64_t y) {  // svgf.wgsl:77-89
    const size_t i = idx(x, y);
    const double* c = io->input_radiance + 4 * i;
    const R age = (R)io->moments[4 * i + 2];
    R sigma = (R)c[3];
    if (age < R(4) && (R)gb->depth[i] < R(1)) {
      R s0 = 0, s1 = 0, total = 0;
      const Guide a = guide(x, y);
      for (int yy = -3; yy <= 3; ++yy)
        for (int xx = -3; xx <= 3; ++xx) {
          const int64_t qx = x + xx, qy = y + yy;
          const R w = geometry_weight(a, qx, qy, R(1));
          if (w <= R(0)) continue;
          s0 += (R)io->moments[4 * idx(qx, qy)] * w;
          s1 += (R)io->moments[4 * idx(qx, qy) + 1] * w;
          total += w;
        }
      s0 /= maxr(total, R(0.00001f));
      s1 /= maxr(total, R(0.00001f));
      sigma = std::sqrt(maxr(s1 - s0 * s0, R(0)) * R(4) / maxr(age, R(1)));
    }
    double* o = io->radiance_out + 4 * i;
    o[0] = (double)store16((R)c[0]);
    o[1] = (double)store16((R)c[1]);
    o[2] = (double)store16((R)c[2]);
    o[3] = (double)store16(minr(sigma, R(65000)));
  }
  static R kernel(int x) {  // svgf.wgsl:90
    if (x == 0) return R(1);
    if (x == 1 || x == -1) return R(2.0f / 3.0f);
    return R(1.0f / 6.0f);
  }
  void svgf_atrous(int64_t x, int64_t y) {  // svgf.wgsl:91-110
    const size_t i = idx(x, y);
    const double* cin = io->input_radiance + 4 * i;
    double* o = io->radiance_out + 4 * i;
    if ((R)gb->depth[i] >= R(1)) {
      for (int k = 0; k < 4; ++k) o[k] = (double)store16((R)cin[k]);
      return;
    }
    const Vec<R> c = D3<R>(cin);
    R variance = 0, guide_weight = 0;
    const Guide a = guide(x, y);
    for (int yy = -1; yy <= 1; ++yy)
      for (int xx = -1; xx <= 1; ++xx) {
        const int64_t qx = clampr<int64_t>(x + xx, 0, (int64_t)W() - 1);
        const int64_t qy = clampr<int64_t>(y + yy, 0, (int64_t)H() - 1);
        const R w = (xx == 0 ? R(2) : R(1)) * (yy == 0 ? R(2) : R(1));
        const R s = (R)io->input_radiance[4 * idx(qx, qy) + 3];
        variance += s * s * w;
        guide_weight += w;
      }
    const R phi = R(4) * std::sqrt(variance / maxr(guide_weight, R(0.00001f))) + R(0.0001f);
    Vec<R> sum = V<R>(0, 0, 0);
    R total = 0, v = 0;
    const int64_t step = (int64_t)to_i32((R)(float)io->stride);
    for (int yy = -2; yy <= 2; ++yy)
      for (int xx = -2; xx <= 2; ++xx) {
        const int64_t qx = x + xx * step, qy = y + yy * step;
        const R geom = geometry_weight(a, qx, qy, (R)(float)step);
        if (geom <= R(0)) continue;
        const double* smp = io->input_radiance + 4 * idx(qx, qy);
        const Vec<R> srgb = D3<R>(smp);
        const R w = kernel(xx) * kernel(yy) * geom * std::exp(-std::fabs(luminance(srgb) - luminance(c)) / phi);
        sum = add3(sum, scale3(srgb, w));
        v += (R)smp[3] * (R)smp[3] * w * w;
        total += w;
      }
    const R t = maxr(total, R(0.00001f));
    o[0] = (double)store16(sum.x / t);
    o[1] = (double)store16(sum.y / t);
    o[2] = (double)store16(sum.z / t);
    o[3] = (double)store16(minr(std::sqrt(v) / t, R(65000)));
  }

  void run(int pass, int64_t x, int64_t y) {
    switch (pass) {
      case ORACLE_CANDIDATES_DIRECT: candidates_direct(x, y); break;
      case ORACLE_CANDIDATES_INDIRECT: candidates_indirect(x, y); break;
      case ORACLE_TEMPORAL_DIRECT: temporal_direct(x, y); break;
      case ORACLE_TEMPORAL_INDIRECT: temporal_indirect(x, y); break;
      case ORACLE_SPATIAL_DIRECT: spatial_direct(x, y); break;
      case ORACLE_SPATIAL_INDIRECT: spatial_indirect(x, y); break;
      case ORACLE_RESOLVE_DIRECT: resolve_direct(x, y); break;
      case ORACLE_RESOLVE_INDIRECT: resolve_indirect(x, y); break;
      case ORACLE_SVGF_TEMPORAL: svgf_temporal(x, y); break;
      case ORACLE_SVGF_VARIANCE: svgf_variance(x, y); break;
      case ORACLE_SVGF_ATROUS: svgf_atrous(x, y); break;
    }
  }
};

template <class R>
int run_pass(int pass, const irt_frame* frame, const oracle_scene* scene, const oracle_gbuffer* gb, oracle_io* io) {
  const uint32_t W = frame->size[0], H = frame->size[1];
  uint32_t y0 = io->row_begin, y1 = io->row_end;
  if (y0 == 0 && y1 == 0) y1 = H;
  if (y1 > H) y1 = H;
  uint64_t total[7] = {0, 0, 0, 0, 0, 0, 0};
  const int threads = io->threads > 0 ? io->threads : omp_get_max_threads();
#pragma omp parallel num_threads(threads)
  {
    Oracle<R> o(*frame, scene, gb, io);
#pragma omp for schedule(dynamic, 1)
    for (int64_t y = y0; y < (int64_t)y1; ++y)
      for (int64_t x = 0; x < (int64_t)W; ++x) o.run(pass, x, y);
#pragma omp critical
    for (int k = 0; k < 7; ++k) total[k] += o.cnt[k];
  }
  if (io->counters)
    for (int k = 0; k < 7; ++k) io->counters[k] += total[k];
  return 0;
}

template <class R>
Oracle<R> make(const irt_frame* f, const oracle_scene* s) {
  static oracle_gbuffer g{};
  static oracle_io i{};
  return Oracle<R>(*f, s, &g, &i);
}

const irt_frame& dummy_frame() {
  static irt_frame f{};
  return f;
}

template <class R>
int add_impl(oracle_reservoir* r, uint32_t id, uint32_t source, double target, double pdf, double weight,
                    uint32_t count, uint32_t age, double random, uint32_t* rng) {
  Res<R> a = from_io<R>(*r);
  a.count += count;
  int taken = 0;
  if (!((R)weight <= R(0) || (R)weight > R(1e30f))) {
    a.sum += (R)weight;
    const R u = random >= 0 ? (R)random : (R)random_f32(rng);
    if (u * a.sum < (R)weight) {
      a.id = id;
      a.source = source;
      a.selected_target = (R)target;
      a.pdf = (R)pdf;
      a.age = age;
      taken = 1;
    }
  }
  *r = to_io(a);
  return taken;
}

// ------------------------------------------------------------ VPL pass
// vpl.wgsl:2-47 in FP32.  The ray query's closest hit is found by testing
// every triangle of every object (no hierarchy): triangles both-sided, t in
// [0.0005, 10000] (vpl.wgsl:18), Moller-Trumbore intersection, equal
// distances to the lower (object, triangle) pair.
using Vf = Vec<float>;

Vf cross3(Vf a, Vf b) { return V<float>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }

// material_at — common.wgsl:50-66 — with textureSampleLevel on the
// repeat / linear sampler (renderer.rs:1024-1031): texel coordinates
// uv * size - 0.5, four texels with repeat wrapping, decoded (unorm c / 255,
// or the sRGB decode for the sRGB view), blended bilinearly.
struct Texel4 {
  float v[4];
};
float srgb_decode8(uint32_t k) {
  const double c = k / 255.0;
  return (float)(c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4));
}
Texel4 texel_at(const oracle_geometry* g, uint32_t layer, uint32_t x, uint32_t y, bool srgb) {
  const uint8_t* p = g->textures + 4 * (((size_t)layer * g->texture_size + y) * g->texture_size + x);
  Texel4 t;
  for (int k = 0; k < 3; ++k) t.v[k] = srgb ? srgb_decode8(p[k]) : (float)p[k] / 255.0f;
  t.v[3] = (float)p[3] / 255.0f;
  return t;
}
uint32_t wrap_i(int32_t i, uint32_t n) {
  const int32_t m = i % (int32_t)n;
  return (uint32_t)(m < 0 ? m + (int32_t)n : m);
}
Texel4 sample_layer(const oracle_geometry* g, uint32_t layer, float u, float v, bool srgb) {
  const uint32_t S = g->texture_size;
  const float x = u * (float)S - 0.5f, y = v * (float)S - 0.5f;
  const float fx0 = std::floor(x), fy0 = std::floor(y);
  const float fx = x - fx0, fy = y - fy0;
  const int32_t ix = (int32_t)std::fmin(std::fmax(fx0, -1073741824.0f), 1073741824.0f);
  const int32_t iy = (int32_t)std::fmin(std::fmax(fy0, -1073741824.0f), 1073741824.0f);
  const Texel4 a = texel_at(g, layer, wrap_i(ix, S), wrap_i(iy, S), srgb);
  const Texel4 b = texel_at(g, layer, wrap_i(ix + 1, S), wrap_i(iy, S), srgb);
  const Texel4 c = texel_at(g, layer, wrap_i(ix, S), wrap_i(iy + 1, S), srgb);
  const Texel4 d = texel_at(g, layer, wrap_i(ix + 1, S), wrap_i(iy + 1, S), srgb);
  const float gx = 1.0f - fx, gy = 1.0f - fy;
  Texel4 r;
  for (int k = 0; k < 4; ++k) r.v[k] = (a.v[k] * gx + b.v[k] * fx) * gy + (c.v[k] * gx + d.v[k] * fx) * fy;
  return r;
}
bool is_textured(const irt_material& m) {
  return m.textures[0] >= 0 || m.textures[1] >= 0 || m.textures[2] >= 0 || m.textures[3] >= 0;
}
// n: interpolated normal; tangent xyz + handedness w.
Surface<float> material_at(const oracle_geometry* g, const irt_material& m, float u, float v, Vf n, const float tangent[4]) {
  Surface<float> s;
  s.p = V<float>(0, 0, 0);
  const float tu = u, tv = 1.0f - v;
  s.color = F3<float>(m.color);
  s.rough = m.params[0];
  s.metal = m.params[1];
  s.n = normalize3(n);
  if (m.textures[0] >= 0) {
    const Texel4 c = sample_layer(g, (uint32_t)m.textures[0], tu, tv, m.params[3] > 0.5f);
    s.color = mul3(s.color, V<float>(c.v[0], c.v[1], c.v[2]));
  }
  if (m.textures[1] >= 0) s.rough = s.rough * sample_layer(g, (uint32_t)m.textures[1], tu, tv, false).v[m.channels[1]];
  if (m.textures[2] >= 0) s.metal = s.metal * sample_layer(g, (uint32_t)m.textures[2], tu, tv, false).v[m.channels[2]];
  if (m.textures[3] >= 0) {
    const Texel4 t4 = sample_layer(g, (uint32_t)m.textures[3], tu, tv, false);
    Vf nm = V<float>(t4.v[0] * 2.0f - 1.0f, t4.v[1] * 2.0f - 1.0f, t4.v[2] * 2.0f - 1.0f);
    nm = normalize3(V<float>(nm.x * m.params[2], nm.y * m.params[2], nm.z));
    const Vf tg = V<float>(tangent[0], tangent[1], tangent[2]);
    const Vf t = normalize3(sub3(tg, scale3(s.n, dot3(s.n, tg))));
    const Vf b = scale3(cross3(s.n, t), tangent[3]);
    s.n = normalize3(add3(add3(scale3(t, nm.x), scale3(b, nm.y)), scale3(s.n, nm.z)));
  }
  s.rough = std::fmin(std::fmax(s.rough, 0.045f), 1.0f);
  s.metal = std::fmin(std::fmax(s.metal, 0.0f), 1.0f);
  return s;
}

// column-major 4x4 times (x, y, z, w), row by row in column order
Vf xform(const float* m, float x, float y, float z, float w) {
  return V<float>(m[0] * x + m[4] * y + m[8] * z + m[12] * w, m[1] * x + m[5] * y + m[9] * z + m[13] * w,
                  m[2] * x + m[6] * y + m[10] * z + m[14] * w);
}

struct Hit {
  float t, u, v;
  uint32_t object, prim;
};

bool closer(float t, uint32_t object, uint32_t prim, const Hit& h) {
  return t < h.t || (t == h.t && (object < h.object || (object == h.object && prim < h.prim)));
}

void ray_triangle(Vf o, Vf d, Vf a, Vf b, Vf c, uint32_t object, uint32_t prim, Hit* h) {
  const Vf e1 = sub3(b, a), e2 = sub3(c, a);
  const Vf p = cross3(d, e2);
  const float det = dot3(e1, p);
  if (det == 0.0f) return;
  const float inv = 1.0f / det;
  const Vf s = sub3(o, a);
  const float u = dot3(s, p) * inv;
  if (!(u >= 0.0f && u <= 1.0f)) return;
  const Vf q = cross3(s, e1);
  const float v = dot3(d, q) * inv;
  if (!(v >= 0.0f && u + v <= 1.0f)) return;
  const float t = dot3(e2, q) * inv;
  if (!(t >= 0.0005f) || !closer(t, object, prim, *h)) return;
  *h = Hit{t, u, v, object, prim};
}

Hit trace_all(const oracle_scene* sc, const oracle_geometry* g, Vf o, Vf d) {
  Hit h{10000.0f, 0.0f, 0.0f, 0xffffffffu, 0xffffffffu};
  for (uint32_t i = 0; i < sc->n_objects; ++i) {
    const irt_object& ob = sc->objects[i];
    const irt_mesh& me = sc->meshes[ob.info[0]];
    const Vf lo = xform(ob.inverse, o.x, o.y, o.z, 1.0f);
    const Vf ld = xform(ob.inverse, d.x, d.y, d.z, 0.0f);
    const irt_vertex* vx = g->vertices + me.geometry[0];
    const uint32_t* ix = g->indices + me.geometry[2];
    for (uint32_t p = 0; p < me.geometry[3] / 3u; ++p)
      ray_triangle(lo, ld, F3<float>(vx[ix[3 * p]].position), F3<float>(vx[ix[3 * p + 1]].position),
                   F3<float>(vx[ix[3 * p + 2]].position), i, p, &h);
  }
  return h;
}

// sin / cos of the spatial neighbour's and the VPL pass's angles: quarter-turn reduction in two
// constants (after Cody & Waite 1980), Taylor polynomials to degree 11 / 12
// (truncation below 2e-9 on |r| <= pi/4), the device's float operations.
void shared_sincos(float x, float* sn, float* cs) {
  const float k = std::nearbyint(x * 0.63661977236758134f);
  const float r = (x - k * 1.57079637050628662f) - k * -4.37113900018624283e-8f;
  const float r2 = r * r;
  const float s = r + r * r2 * (-1.0f / 6.0f + r2 * (1.0f / 120.0f + r2 * (-1.0f / 5040.0f + r2 * (1.0f / 362880.0f + r2 * (-1.0f / 39916800.0f)))));
  const float c = 1.0f + r2 * (-0.5f + r2 * (1.0f / 24.0f + r2 * (-1.0f / 720.0f + r2 * (1.0f / 40320.0f + r2 * (-1.0f / 3628800.0f + r2 * (1.0f / 479001600.0f))))));
  const int q = ((int)k) & 3;
  *sn = q == 0 ? s : (q == 1 ? c : (q == 2 ? -s : -c));
  *cs = q == 0 ? c : (q == 1 ? -s : (q == 2 ? -c : s));
}
float shared_cos(float x) {
  float s, c;
  shared_sincos(x, &s, &c);
  return c;
}

// basis — common.wgsl:84, applied to v.
Vf basis_mul(Vf n, Vf v) {
  const Vf axis = std::fabs(n.z) > 0.9f ? V<float>(1, 0, 0) : V<float>(0, 0, 1);
  const Vf t = normalize3(cross3(axis, n));
  const Vf b = cross3(n, t);
  return V<float>(t.x * v.x + b.x * v.y + n.x * v.z, t.y * v.x + b.y * v.y + n.y * v.z,
                  t.z * v.x + b.z * v.y + n.z * v.z);
}

const float PI32 = 3.14159265359f;

// cosine_direction — common.wgsl:85.
Vf cosine_direction(Vf n, uint32_t* rng) {
  const float u = random_f32(rng);
  const float phi = 2.0f * PI32 * random_f32(rng);
  float sn, cs;
  shared_sincos(phi, &sn, &cs);
  return basis_mul(n, V<float>(std::sqrt(u) * cs, std::sqrt(u) * sn, std::sqrt(1.0f - u)));
}

// sample_bsdf — common.wgsl:86-97.  pow(rough, 4) as (rough^2)^2.
void sample_bsdf(Oracle<float>& o, const Surface<float>& s, Vf wo, uint32_t* rng, Vf* dir, Vf* weight) {
  const float spec_probability = 0.25f * (1.0f - s.metal) + 0.9f * s.metal;
  const float ra = s.rough * s.rough;
  const float a2 = ra * ra;
  Vf wi;
  if (random_f32(rng) < spec_probability) {
    const float u = random_f32(rng);
    const float cos_theta = std::sqrt((1.0f - u) / (1.0f + (a2 - 1.0f) * u));
    const float phi = 2.0f * PI32 * random_f32(rng);
    const float st = std::sqrt(1.0f - cos_theta * cos_theta);
    float sn, cs;
    shared_sincos(phi, &sn, &cs);
    const Vf h = basis_mul(s.n, V<float>(st * cs, st * sn, cos_theta));
    const Vf e = scale3(wo, -1.0f);
    wi = sub3(e, scale3(h, 2.0f * dot3(h, e)));  // reflect(-wo, h)
  } else {
    wi = cosine_direction(s.n, rng);
  }
  *dir = wi;
  const float nl = maxr(dot3(s.n, wi), 0.0f);
  if (nl <= 0.0f) {
    *weight = V<float>(0, 0, 0);
    return;
  }
  const Vf h = normalize3(add3(wo, wi));
  const float nh = maxr(dot3(s.n, h), 0.0f);
  const float q = nh * nh * (a2 - 1.0f) + 1.0f;
  const float pdf = (nl / PI32) * (1.0f - spec_probability) +
                    (a2 * nh / (PI32 * q * q * maxr(4.0f * std::fabs(dot3(wo, h)), 0.00001f))) * spec_probability;
  const Vf b = o.brdf(s, wo, wi);
  const float den = maxr(pdf, 0.00001f);
  *weight = V<float>(b.x * nl / den, b.y * nl / den, b.z * nl / den);
}

float clamp32(float x, float lo, float hi) { return std::fmin(std::fmax(x, lo), hi); }

// generate — vpl.wgsl:2-47, path `id`.
irt_vpl generate_one(Oracle<float>& o, const oracle_scene* sc, const oracle_geometry* g, uint32_t id,
                     uint64_t* rays) {
  const irt_frame& F = o.F;
  uint32_t rng = seed_of(id, 7u, F.size[2]);  // vpl.wgsl:4
  irt_vpl v;
  std::memset(&v, 0, sizeof(v));
  const uint32_t L = F.counts[0];
  if (L == 0u) return v;  // vpl.wgsl:5
  const uint32_t light_id = std::min(to_u32(random_f32(&rng) * (float)L), L - 1u);  // vpl.wgsl:6
  const irt_light& l = sc->lights[light_id];
  Vf origin = F3<float>(l.position_strength), direction;
  float omega = 4.0f * PI32;
  if (l.color_type[3] > 0.5f) {  // vpl.wgsl:8-10
    const float ch = shared_cos(l.direction_angle[3] * 0.5f);
    omega = 2.0f * PI32 * (1.0f - ch);
    const float r = random_f32(&rng);
    const float z = ch * (1.0f - r) + 1.0f * r;
    const float phi = 2.0f * PI32 * random_f32(&rng);
    float sn, cs;
    shared_sincos(phi, &sn, &cs);
    direction = basis_mul(normalize3(F3<float>(l.direction_angle)),
                          V<float>(std::sqrt(1.0f - z * z) * cs, std::sqrt(1.0f - z * z) * sn, z));
  } else {  // vpl.wgsl:11
    const float z = 2.0f * random_f32(&rng) - 1.0f;
    const float phi = 2.0f * PI32 * random_f32(&rng);
    float sn, cs;
    shared_sincos(phi, &sn, &cs);
    direction = V<float>(std::sqrt(1.0f - z * z) * cs, std::sqrt(1.0f - z * z) * sn, z);
  }
  const float Lf = (float)L;  // vpl.wgsl:14
  const Vf source_flux = V<float>(l.color_type[0] * l.position_strength[3] * omega * Lf,
                                  l.color_type[1] * l.position_strength[3] * omega * Lf,
                                  l.color_type[2] * l.position_strength[3] * omega * Lf);
  Vf throughput = V<float>(1, 1, 1);
  const float Vn = (float)F.counts[1];
  for (uint32_t bounce = 0; bounce < F.counts[2]; ++bounce) {  // vpl.wgsl:16-45
    const Hit hit = trace_all(sc, g, origin, direction);
    ++*rays;
    if (hit.object == 0xffffffffu) {
      v.flux[0] = v.flux[1] = v.flux[2] = v.flux[3] = 0.0f;
      break;
    }
    const irt_object& ob = sc->objects[hit.object];
    const irt_mesh& me = sc->meshes[ob.info[0]];
    const uint32_t base = me.geometry[2] + hit.prim * 3u;  // vpl.wgsl:23-28
    const irt_vertex& a = g->vertices[me.geometry[0] + g->indices[base]];
    const irt_vertex& b = g->vertices[me.geometry[0] + g->indices[base + 1u]];
    const irt_vertex& c = g->vertices[me.geometry[0] + g->indices[base + 2u]];
    const float bx = 1.0f - hit.u - hit.v, by = hit.u, bz = hit.v;
    const Vf ln = V<float>(a.normal[0] * bx + b.normal[0] * by + c.normal[0] * bz,
                           a.normal[1] * bx + b.normal[1] * by + c.normal[1] * bz,
                           a.normal[2] * bx + b.normal[2] * by + c.normal[2] * bz);
    const Vf n = normalize3(xform(ob.model, ln.x, ln.y, ln.z, 0.0f));
    const irt_material& m = g->materials[ob.info[1]];  // material_at, vpl.wgsl:27-29
    float tu = 0.0f, tv = 0.0f, tangent[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    if (is_textured(m)) {
      tu = a.uv[0] * bx + b.uv[0] * by + c.uv[0] * bz;
      tv = a.uv[1] * bx + b.uv[1] * by + c.uv[1] * bz;
      float t4[4];
      for (int k = 0; k < 4; ++k) t4[k] = a.tangent[k] * bx + b.tangent[k] * by + c.tangent[k] * bz;
      const Vf wt = xform(ob.model, t4[0], t4[1], t4[2], 0.0f);
      tangent[0] = wt.x;
      tangent[1] = wt.y;
      tangent[2] = wt.z;
      tangent[3] = t4[3] >= 0.0f ? 1.0f : -1.0f;
    }
    Surface<float> s = material_at(g, m, tu, tv, n, tangent);
    s.p = add3(origin, scale3(direction, hit.t));  // vpl.wgsl:30
    const Vf wo = scale3(direction, -1.0f);
    if (dot3(s.n, wo) <= 0.0f) {  // vpl.wgsl:32
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
    Vf sdir, sweight;
    sample_bsdf(o, s, wo, &rng, &sdir, &sweight);  // vpl.wgsl:36-38
    const Vf reflected = mul3(throughput, sweight);
    v.flux[0] = source_flux.x * reflected.x / Vn;
    v.flux[1] = source_flux.y * reflected.y / Vn;
    v.flux[2] = source_flux.z * reflected.z / Vn;
    v.flux[3] = 0.0f;
    if (bounce + 1u >= F.counts[2]) break;  // vpl.wgsl:39-44
    const float survive = clamp32(std::fmax(reflected.x, std::fmax(reflected.y, reflected.z)), F.rr[0], F.rr[1]);
    if (random_f32(&rng) > survive) break;
    if (o.luminance(sweight) <= 0.0f) break;
    throughput = div3(reflected, survive);
    origin = add3(s.p, scale3(s.n, 0.001f));
    direction = sdir;
  }
  return v;
}

// ------------------------------------------------------------ G-buffer pass
// gbuffer.wgsl:1-12 with the raster state of renderer.rs:1370-1415 (back
// faces culled, counter-clockwise front, depth Less then Equal): every
// instance triangle's covered pixels by 2-D homogeneous edge functions,
// the least (depth, ~draw index) per pixel, then the fragment stage.
void mat4_mul(const float* a, const float* b, float* out) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      out[c * 4 + r] = a[r] * b[c * 4] + a[4 + r] * b[c * 4 + 1] + a[8 + r] * b[c * 4 + 2] + a[12 + r] * b[c * 4 + 3];
}
struct Clip {
  float x, y, z, w;
};
Clip clip_of(const float* m, const float* p) {
  return Clip{m[0] * p[0] + m[4] * p[1] + m[8] * p[2] + m[12] * p[3], m[1] * p[0] + m[5] * p[1] + m[9] * p[2] + m[13] * p[3],
              m[2] * p[0] + m[6] * p[1] + m[10] * p[2] + m[14] * p[3],
              m[3] * p[0] + m[7] * p[1] + m[11] * p[2] + m[15] * p[3]};
}
float edge_planes(const Clip c[3], float e[9]) {
  const Vf V0 = V<float>(c[0].x, c[0].y, c[0].w), V1 = V<float>(c[1].x, c[1].y, c[1].w),
           V2 = V<float>(c[2].x, c[2].y, c[2].w);
  const Vf a = cross3(V1, V2), b = cross3(V2, V0), d = cross3(V0, V1);
  e[0] = a.x; e[1] = a.y; e[2] = a.z;
  e[3] = b.x; e[4] = b.y; e[5] = b.z;
  e[6] = d.x; e[7] = d.y; e[8] = d.z;
  return V0.x * a.x + V0.y * a.y + V0.z * a.z;
}
float plane_at(const float* e, float px, float py) { return e[0] * px + e[1] * py + e[2]; }
float ndc_x(uint32_t x, uint32_t W) { return ((float)x + 0.5f) / (float)W * 2.0f - 1.0f; }
float ndc_y(uint32_t y, uint32_t H) { return 1.0f - ((float)y + 0.5f) / (float)H * 2.0f; }

// Pixel bounds of the part of the triangle in front of the near plane
// (false: nothing to draw).
bool tri_bounds(const Clip c[3], uint32_t W, uint32_t H, uint32_t* x0, uint32_t* x1, uint32_t* y0, uint32_t* y1) {
  if (c[0].z >= c[0].w && c[1].z >= c[1].w && c[2].z >= c[2].w) return false;
  float lx = INFINITY, hx = -INFINITY, ly = INFINITY, hy = -INFINITY;
  bool any = false, whole = false;
  for (int k = 0; k < 3; ++k) {
    const Clip a = c[k], b = c[(k + 1) % 3];
    Clip q[2];
    int nq = 0;
    if (a.z >= 0.0f) q[nq++] = a;
    if ((a.z >= 0.0f) != (b.z >= 0.0f)) {
      const float f = a.z / (a.z - b.z);
      q[nq++] = Clip{a.x + (b.x - a.x) * f, a.y + (b.y - a.y) * f, 0.0f, a.w + (b.w - a.w) * f};
    }
    for (int j = 0; j < nq; ++j) {
      any = true;
      if (!(q[j].w > 0.0f)) {
        whole = true;
        continue;
      }
      const float sx = (q[j].x / q[j].w + 1.0f) * 0.5f * (float)W - 0.5f;
      const float sy = (1.0f - q[j].y / q[j].w) * 0.5f * (float)H - 0.5f;
      lx = std::fmin(lx, sx); hx = std::fmax(hx, sx); ly = std::fmin(ly, sy); hy = std::fmax(hy, sy);
    }
  }
  float fx0 = 0.0f, fx1 = (float)W - 1.0f, fy0 = 0.0f, fy1 = (float)H - 1.0f;
  if (!whole) {
    fx0 = std::fmax(fx0, std::floor(lx) - 1.0f); fx1 = std::fmin(fx1, std::ceil(hx) + 1.0f);
    fy0 = std::fmax(fy0, std::floor(ly) - 1.0f); fy1 = std::fmin(fy1, std::ceil(hy) + 1.0f);
  }
  if (!any || !(fx0 <= fx1 && fy0 <= fy1)) return false;
  *x0 = (uint32_t)fx0; *x1 = (uint32_t)fx1; *y0 = (uint32_t)fy0; *y1 = (uint32_t)fy1;
  return true;
}

// ------------------------------------------------------------ SDF bake
// tools/export_scene.py:121-158 with every triangle tested for the nearest
// point and for each parity cast (closest point: Ericson 2005, 5.1.5).
Vf closest_on_triangle(Vf p, Vf a, Vf b, Vf c) {
  const Vf ab = sub3(b, a), ac = sub3(c, a), ap = sub3(p, a);
  const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return a;
  const Vf bp = sub3(p, b);
  const float d3 = dot3(ab, bp), d4 = dot3(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return b;
  const float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return add3(a, scale3(ab, d1 / (d1 - d3)));
  const Vf cp = sub3(p, c);
  const float d5 = dot3(ab, cp), d6 = dot3(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return c;
  const float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return add3(a, scale3(ac, d2 / (d2 - d6)));
  const float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f)
    return add3(b, scale3(sub3(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
  const float denom = 1.0f / (va + vb + vc);
  const float v = vb * denom, w = vc * denom;
  return add3(add3(a, scale3(ab, v)), scale3(ac, w));
}

}  // namespace

extern "C" {

int oracle_run_pass(int bits, int pass, const irt_frame* frame, const oracle_scene* scene, const oracle_gbuffer* gb,
                    oracle_io* io) {
  if (!frame || !gb || !io || pass < 0 || pass > ORACLE_SVGF_ATROUS) return -1;
  if (pass <= ORACLE_RESOLVE_INDIRECT && !scene) return -1;
  return bits == 64 ? run_pass<double>(pass, frame, scene, gb, io) : run_pass<float>(pass, frame, scene, gb, io);
}

float oracle_random(uint32_t* state) { return random_f32(state); }
uint32_t oracle_seed(uint32_t i, uint32_t salt, uint32_t frame_index) { return seed_of(i, salt, frame_index); }

double oracle_signed_box_distance(int bits, const double p[3], const double lo[3], const double hi[3]) {
  if (bits == 64) return Oracle<double>::signed_box_distance(D3<double>(p), D3<double>(lo), D3<double>(hi));
  return (double)Oracle<float>::signed_box_distance(D3<float>(p), D3<float>(lo), D3<float>(hi));
}

double oracle_mesh_distance(int bits, const double p[3], const irt_mesh* mesh, const oracle_scene* scene) {
  if (bits == 64) {
    Oracle<double> o = make<double>(&dummy_frame(), scene);
    return o.mesh_distance(D3<double>(p), *mesh);
  }
  Oracle<float> o = make<float>(&dummy_frame(), scene);
  return (double)o.mesh_distance(D3<float>(p), *mesh);
}

double oracle_scene_distance(int bits, const irt_frame* frame, const oracle_scene* scene, const double p[3],
                             int need_spacing, uint32_t hint, double* voxel, uint32_t* closest) {
  uint32_t c = 0;
  if (bits == 64) {
    Oracle<double> o = make<double>(frame, scene);
    double d, v;
    o.scene_distance(D3<double>(p), need_spacing != 0, hint, &d, &v, &c);
    if (voxel) *voxel = v;
    if (closest) *closest = c;
    return d;
  }
  Oracle<float> o = make<float>(frame, scene);
  float d, v;
  o.scene_distance(D3<float>(p), need_spacing != 0, hint, &d, &v, &c);
  if (voxel) *voxel = v;
  if (closest) *closest = c;
  return d;
}

double oracle_visibility(int bits, const irt_frame* frame, const oracle_scene* scene, const double p[3],
                         const double n[3], const double target[3], const double end_normal[3], uint64_t* counters) {
  if (bits == 64) {
    Oracle<double> o = make<double>(frame, scene);
    double r = o.visibility(D3<double>(p), D3<double>(n), D3<double>(target), D3<double>(end_normal));
    if (counters)
      for (int k = 0; k < 7; ++k) counters[k] += o.cnt[k];
    return r;
  }
  Oracle<float> o = make<float>(frame, scene);
  float r = o.visibility(D3<float>(p), D3<float>(n), D3<float>(target), D3<float>(end_normal));
  if (counters)
    for (int k = 0; k < 7; ++k) counters[k] += o.cnt[k];
  return r;
}

void oracle_brdf(int bits, const double s[11], const double wo[3], const double wi[3], double out[3]) {
  if (bits == 64) {
    Oracle<double> o = make<double>(&dummy_frame(), nullptr);
    Surface<double> sf{D3<double>(s), D3<double>(s + 3), D3<double>(s + 6), s[9], s[10]};
    Vec<double> r = o.brdf(sf, D3<double>(wo), D3<double>(wi));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
    return;
  }
  Oracle<float> o = make<float>(&dummy_frame(), nullptr);
  Surface<float> sf{D3<float>(s), D3<float>(s + 3), D3<float>(s + 6), (float)s[9], (float)s[10]};
  Vec<float> r = o.brdf(sf, D3<float>(wo), D3<float>(wi));
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
}

int oracle_reservoir_add(int bits, oracle_reservoir* r, uint32_t id, uint32_t source, double target, double pdf,
                         double weight, uint32_t count, uint32_t age, double random, uint32_t* rng) {
  if (!r || (random < 0 && !rng)) return -1;
  return bits == 64 ? add_impl<double>(r, id, source, target, pdf, weight, count, age, random, rng)
                    : add_impl<float>(r, id, source, target, pdf, weight, count, age, random, rng);
}

double oracle_normalization(int bits, const oracle_reservoir* r) {
  if (bits == 64) return Oracle<double>::normalization(from_io<double>(*r));
  return (double)Oracle<float>::normalization(from_io<float>(*r));
}

double oracle_shadow_step(int bits, double result, double h, double previous, double t, double k) {
  if (bits == 64) return Oracle<double>::shadow_step(result, h, previous, t, k);
  return (double)Oracle<float>::shadow_step((float)result, (float)h, (float)previous, (float)t, (float)k);
}

void oracle_neighbor_offsets(int bits, uint32_t state, uint32_t spatial_radius, uint32_t count, int32_t* out) {
  for (uint32_t i = 0; i < count; ++i) {
    int64_t dx, dy;
    if (bits == 64)
      Oracle<double>::neighbor_offset(&state, spatial_radius, &dx, &dy);
    else
      Oracle<float>::neighbor_offset(&state, spatial_radius, &dx, &dy);
    out[2 * i] = (int32_t)dx;
    out[2 * i + 1] = (int32_t)dy;
  }
}

uint16_t oracle_f32_to_f16(float f) { return f32_to_f16(f); }
float oracle_f16_to_f32(uint16_t h) { return f16_to_f32(h); }

int oracle_generate_vpls(const irt_frame* frame, const oracle_scene* scene, const oracle_geometry* geometry,
                         irt_vpl* out, int threads, uint64_t* rays) {
  if (!frame || !scene || !geometry || !out) return -1;
  if (frame->counts[0] > scene->n_lights) return -1;
  Oracle<float> o = make<float>(frame, scene);
  const int n = (int)frame->counts[1];
  if (threads <= 0) threads = omp_get_max_threads();
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads) reduction(+ : total)
  for (int i = 0; i < n; ++i) {
    Oracle<float> local = o;
    uint64_t traced = 0;
    out[i] = generate_one(local, scene, geometry, (uint32_t)i, &traced);
    total += traced;
  }
  if (rays) *rays = total;
  return 0;
}

int oracle_trace(const oracle_scene* scene, const oracle_geometry* geometry, const float origin[3],
                 const float direction[3], float* t, float* u, float* v, uint32_t* object, uint32_t* prim) {
  const Hit h = trace_all(scene, geometry, F3<float>(origin), F3<float>(direction));
  *t = h.t;
  *u = h.u;
  *v = h.v;
  *object = h.object;
  *prim = h.prim;
  return h.object != 0xffffffffu;
}

void oracle_sample_bsdf(const float n[3], const float color[3], float rough, float metal, const float wo[3],
                        uint32_t* rng, float direction[3], float weight[3]) {
  Oracle<float> o = make<float>(&dummy_frame(), nullptr);
  Surface<float> s;
  s.p = V<float>(0, 0, 0);
  s.n = F3<float>(n);
  s.color = F3<float>(color);
  s.rough = rough;
  s.metal = metal;
  Vf d, w;
  sample_bsdf(o, s, F3<float>(wo), rng, &d, &w);
  direction[0] = d.x;
  direction[1] = d.y;
  direction[2] = d.z;
  weight[0] = w.x;
  weight[1] = w.y;
  weight[2] = w.z;
}

int oracle_render_gbuffer(const irt_frame* frame, const oracle_scene* scene, const oracle_geometry* g, float* depth,
                          uint16_t* normal, uint16_t* color, uint16_t* params, uint16_t* motion) {
  if (!frame || !scene || !g || !depth || !normal || !color || !params || !motion) return -1;
  const uint32_t W = frame->size[0], H = frame->size[1], n_obj = scene->n_objects;
  std::vector<float> mvp(16 * (size_t)n_obj), pmvp(16 * (size_t)n_obj);
  std::vector<uint32_t> base(n_obj + 1, 0);
  for (uint32_t o = 0; o < n_obj; ++o) {
    mat4_mul(frame->vp, scene->objects[o].model, &mvp[16 * o]);
    mat4_mul(frame->previous_vp, scene->objects[o].previous, &pmvp[16 * o]);
    base[o + 1] = base[o] + scene->meshes[scene->objects[o].info[0]].geometry[3] / 3u;
  }
  auto corners = [&](uint32_t o, uint32_t prim, uint32_t vi[3], Clip c[3]) {
    const irt_mesh& me = scene->meshes[scene->objects[o].info[0]];
    for (int k = 0; k < 3; ++k) {
      vi[k] = me.geometry[0] + g->indices[me.geometry[2] + 3u * prim + (uint32_t)k];
      c[k] = clip_of(&mvp[16 * o], g->vertices[vi[k]].position);
    }
  };
  const uint64_t clear = (uint64_t)0x3f800000u << 32 | 0xffffffffu;
  std::vector<uint64_t> vis((size_t)W * H, clear);
  for (uint32_t o = 0; o < n_obj; ++o)
    for (uint32_t prim = 0; prim + base[o] < base[o + 1]; ++prim) {
      const uint32_t t = base[o] + prim;
      uint32_t vi[3];
      Clip c[3];
      corners(o, prim, vi, c);
      float e[9], z[3];
      const float det = edge_planes(c, e);
      if (!(det > 0.0f)) continue;
      for (int k = 0; k < 3; ++k) z[k] = (c[0].z * e[k] + c[1].z * e[3 + k] + c[2].z * e[6 + k]) / det;
      uint32_t x0, x1, y0, y1;
      if (!tri_bounds(c, W, H, &x0, &x1, &y0, &y1)) continue;
      for (uint32_t y = y0; y <= y1; ++y)
        for (uint32_t x = x0; x <= x1; ++x) {
          const float px = ndc_x(x, W), py = ndc_y(y, H);
          const float e0 = plane_at(e, px, py), e1 = plane_at(e + 3, px, py), e2 = plane_at(e + 6, px, py);
          if (!(e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f && e0 + e1 + e2 > 0.0f)) continue;
          const float zz = plane_at(z, px, py) + 0.0f;
          if (!(zz >= 0.0f && zz < 1.0f)) continue;
          uint32_t zb;
          std::memcpy(&zb, &zz, 4);
          const uint64_t key = (uint64_t)zb << 32 | (uint64_t)(~t);
          uint64_t& slot = vis[(size_t)y * W + x];
          if (key < slot) slot = key;
        }
    }
#pragma omp parallel for schedule(static)
  for (int yy = 0; yy < (int)H; ++yy)
    for (uint32_t x = 0; x < W; ++x) {
      const uint32_t y = (uint32_t)yy;
      const size_t id = (size_t)y * W + x;
      const uint64_t key = vis[id];
      uint16_t* np = normal + 4 * id;
      uint16_t* cp = color + 4 * id;
      if (key == clear) {
        depth[id] = 1.0f;
        for (int k = 0; k < 4; ++k) np[k] = cp[k] = 0;
        params[2 * id] = params[2 * id + 1] = motion[2 * id] = motion[2 * id + 1] = 0;
        continue;
      }
      const uint32_t t = ~(uint32_t)(key & 0xffffffffu);
      uint32_t o = 0;
      while (base[o + 1] <= t) ++o;
      uint32_t vi[3];
      Clip c[3];
      corners(o, t - base[o], vi, c);
      float e[9];
      edge_planes(c, e);
      const float px = ndc_x(x, W), py = ndc_y(y, H);
      const float E0 = plane_at(e, px, py), E1 = plane_at(e + 3, px, py), E2 = plane_at(e + 6, px, py);
      const float sum = E0 + E1 + E2;
      const float l0 = E0 / sum, l1 = E1 / sum, l2 = E2 / sum;
      const irt_object& ob = scene->objects[o];
      Vf n[3];
      Clip pc[3];
      for (int k = 0; k < 3; ++k) {
        const irt_vertex& vx = g->vertices[vi[k]];
        n[k] = normalize3(xform(ob.model, vx.normal[0], vx.normal[1], vx.normal[2], 0.0f));
        pc[k] = clip_of(&pmvp[16 * o], vx.position);
      }
      const Vf ni = V<float>(n[0].x * l0 + n[1].x * l1 + n[2].x * l2, n[0].y * l0 + n[1].y * l1 + n[2].y * l2,
                             n[0].z * l0 + n[1].z * l1 + n[2].z * l2);
      const float cx = c[0].x * l0 + c[1].x * l1 + c[2].x * l2, cy = c[0].y * l0 + c[1].y * l1 + c[2].y * l2;
      const float cw = c[0].w * l0 + c[1].w * l1 + c[2].w * l2;
      const float qx = pc[0].x * l0 + pc[1].x * l1 + pc[2].x * l2, qy = pc[0].y * l0 + pc[1].y * l1 + pc[2].y * l2;
      const float qw = pc[0].w * l0 + pc[1].w * l1 + pc[2].w * l2;
      const irt_material& m = g->materials[ob.info[1]];
      float tu = 0.0f, tv = 0.0f, tangent[4] = {0.0f, 0.0f, 0.0f, 1.0f};
      if (is_textured(m)) {
        float ut[3], vt[3], tg[3][4];
        for (int k = 0; k < 3; ++k) {
          const irt_vertex& vx = g->vertices[vi[k]];
          ut[k] = vx.uv[0];
          vt[k] = vx.uv[1];
          const Vf wt = normalize3(xform(ob.model, vx.tangent[0], vx.tangent[1], vx.tangent[2], 0.0f));
          tg[k][0] = wt.x; tg[k][1] = wt.y; tg[k][2] = wt.z; tg[k][3] = vx.tangent[3];
        }
        tu = ut[0] * l0 + ut[1] * l1 + ut[2] * l2;
        tv = vt[0] * l0 + vt[1] * l1 + vt[2] * l2;
        for (int k = 0; k < 4; ++k) tangent[k] = tg[0][k] * l0 + tg[1][k] * l1 + tg[2][k] * l2;
      }
      const Surface<float> sf = material_at(g, m, tu, tv, ni, tangent);
      uint32_t zb = (uint32_t)(key >> 32);
      std::memcpy(&depth[id], &zb, 4);
      np[0] = f32_to_f16(sf.n.x); np[1] = f32_to_f16(sf.n.y); np[2] = f32_to_f16(sf.n.z); np[3] = f32_to_f16(1.0f);
      cp[0] = f32_to_f16(sf.color.x); cp[1] = f32_to_f16(sf.color.y); cp[2] = f32_to_f16(sf.color.z);
      cp[3] = f32_to_f16(1.0f);
      params[2 * id] = f32_to_f16(sf.rough);
      params[2 * id + 1] = f32_to_f16(sf.metal);
      motion[2 * id] = f32_to_f16((cx / cw - qx / qw) * 0.5f);
      motion[2 * id + 1] = f32_to_f16((cy / cw - qy / qw) * -0.5f);
    }
  return 0;
}

void oracle_present(const uint16_t* history, uint32_t n, uint8_t* out) {
  // decision thresholds of the 8-bit sRGB encode (IEC 61966-2-1): code k+0.5 to linear, rounded up to a float
  float t[255];
  for (int k = 0; k < 255; ++k) {
    const double c = (k + 0.5) / 255.0;
    const double lin = c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
    float f = (float)lin;
    if ((double)f < lin) f = std::nextafter(f, INFINITY);
    t[k] = f;
  }
  for (uint32_t i = 0; i < n; ++i) {
    for (int ch = 0; ch < 3; ++ch) {
      float v = f16_to_f32(history[4 * i + ch]);
      v = std::fmin(std::fmax(v, 0.0f), 1.0f);  // NaN -> 0
      int code = 0;
      while (code < 255 && t[code] <= v) ++code;
      out[4 * i + ch] = (uint8_t)code;
    }
    out[4 * i + 3] = 255;
  }
}

int oracle_bake_sdf(const float* positions, uint32_t stride, uint32_t nv, const uint32_t* indices, uint32_t ni,
                    uint32_t res, float* out, float lo_out[3], float hi_out[3], int threads) {
  if (!positions || !indices || !out || res < 4 || ni % 3u || stride < 3) return -1;
  const uint32_t nt = ni / 3u;
  std::vector<Vf> A(nt), B(nt), C(nt);
  for (uint32_t t = 0; t < nt; ++t) {
    for (int k = 0; k < 3; ++k)
      if (indices[3 * t + k] >= nv) return -1;
    A[t] = F3<float>(positions + (size_t)indices[3 * t] * stride);
    B[t] = F3<float>(positions + (size_t)indices[3 * t + 1] * stride);
    C[t] = F3<float>(positions + (size_t)indices[3 * t + 2] * stride);
  }
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < nv; ++v)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::fmin(lo[a], positions[(size_t)v * stride + a]);
      hi[a] = std::fmax(hi[a], positions[(size_t)v * stride + a]);
    }
  const float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const double padding = std::sqrt((double)e[0] * e[0] + (double)e[1] * e[1] + (double)e[2] * e[2]) / (double)(res - 1) * 2.0;
  const float pf = (float)padding;
  double glo[3], ghi[3];
  for (int a = 0; a < 3; ++a) {
    lo_out[a] = lo[a] - pf;
    hi_out[a] = hi[a] + pf;
    glo[a] = lo_out[a];
    ghi[a] = hi_out[a];
  }
  const float pe[3] = {hi_out[0] - lo_out[0], hi_out[1] - lo_out[1], hi_out[2] - lo_out[2]};
  const double extent = std::sqrt((double)pe[0] * pe[0] + (double)pe[1] * pe[1] + (double)pe[2] * pe[2]);
  const float tmax = (float)(extent * 4.0), step = std::fmax((float)(extent * 1e-6), 1e-7f);
  const Vf dirs[3] = {normalize3(V<float>(1.0f, 0.371f, 0.129f)), normalize3(V<float>(0.173f, 1.0f, 0.419f)),
                      normalize3(V<float>(0.319f, 0.217f, 1.0f))};
  int overflow = 0;
  if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads) reduction(| : overflow)
  for (int64_t i = 0; i < (int64_t)res * res * res; ++i) {
    const uint32_t c[3] = {(uint32_t)(i % res), (uint32_t)((i / res) % res), (uint32_t)(i / ((int64_t)res * res))};
    float pfv[3];
    for (int a = 0; a < 3; ++a) pfv[a] = (float)(glo[a] + (ghi[a] - glo[a]) * (double)c[a] / (double)(res - 1));
    const Vf p = V<float>(pfv[0], pfv[1], pfv[2]);
    float best = INFINITY;
    for (uint32_t t = 0; t < nt; ++t) {
      const Vf d = sub3(p, closest_on_triangle(p, A[t], B[t], C[t]));
      best = std::fmin(best, dot3(d, d));
    }
    const float dist = std::sqrt(best);
    int votes = 0;
    for (int k = 0; k < 3; ++k) {
      Vf start = p;
      uint32_t hits = 0;
      while (true) {
        Hit h{tmax, 0.0f, 0.0f, 0xffffffffu, 0xffffffffu};
        for (uint32_t t = 0; t < nt; ++t) {
          const Vf e1 = sub3(B[t], A[t]), e2 = sub3(C[t], A[t]);
          const Vf pv = cross3(dirs[k], e2);
          const float det = dot3(e1, pv);
          if (det == 0.0f) continue;
          const float inv = 1.0f / det;
          const Vf sv = sub3(start, A[t]);
          const float u = dot3(sv, pv) * inv;
          if (!(u >= 0.0f && u <= 1.0f)) continue;
          const Vf q = cross3(sv, e1);
          const float v = dot3(dirs[k], q) * inv;
          if (!(v >= 0.0f && u + v <= 1.0f)) continue;
          const float tt = dot3(e2, q) * inv;
          if (!(tt >= 0.0f) || !closer(tt, 0u, t, h)) continue;
          h = Hit{tt, u, v, 0u, t};
        }
        if (h.object == 0xffffffffu) break;
        if (++hits >= 256u) {
          overflow = 1;
          break;
        }
        start = add3(add3(start, scale3(dirs[k], h.t)), scale3(dirs[k], step));
      }
      votes += (int)(hits & 1u);
    }
    const double sign = votes >= 2 ? -1.0 : 1.0;
    out[i] = (float)(sign * (double)dist - extent * 2e-6);
  }
  return overflow ? -2 : 0;
}

}  // extern "C"
