// Synthetic scene producer for tests and benchmarks (input side of the path).
//
// The reference gets its inputs from a Blender export (meshes + SDF grids,
// tools/export_scene.py), a raster G-buffer (gbuffer.wgsl) and ray-query VPL
// paths (vpl.wgsl).  This file produces inputs of the same formats from
// analytic shapes so that no GPU ray tracing or Blender is needed:
//   - meshes are boxes, spheres and tori; each is baked to an SDF grid on the
//     exporter's convention (vertex AABB padded by 2*|diag|/(res-1), node-
//     centred samples, x fastest, minus extent*2e-6; export_scene.py:135-158)
//     and tessellated so the vertex-derived bounds are exact;
//   - the G-buffer is ray-cast analytically (depth = NDC z of the hit,
//     FP16-quantised normal/colour/params/motion as the Rgba16F/Rg16F
//     targets hold them, motion = (cur - prev NDC) * (0.5, -0.5), gbuffer.wgsl:11);
//   - VPLs are light paths traced against the same analytic shapes.
// All geometry arithmetic is double; results are rounded to the record types.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/irt.h"
#include "irt_host.h"

namespace {

struct d3 {
  double x, y, z;
};
inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline d3 operator*(d3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double len(d3 a) { return std::sqrt(dot(a, a)); }
inline d3 nrm(d3 a) { return a * (1.0 / len(a)); }
inline d3 cross(d3 a, d3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

struct M4 {
  double m[16];  // column-major
};
M4 mul(const M4& a, const M4& b) {
  M4 r;
  for (int c = 0; c < 4; ++c)
    for (int rr = 0; rr < 4; ++rr) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += a.m[k * 4 + rr] * b.m[c * 4 + k];
      r.m[c * 4 + rr] = s;
    }
  return r;
}
d3 xform_point(const M4& a, d3 p, double* w = nullptr) {
  double x = a.m[0] * p.x + a.m[4] * p.y + a.m[8] * p.z + a.m[12];
  double y = a.m[1] * p.x + a.m[5] * p.y + a.m[9] * p.z + a.m[13];
  double z = a.m[2] * p.x + a.m[6] * p.y + a.m[10] * p.z + a.m[14];
  double ww = a.m[3] * p.x + a.m[7] * p.y + a.m[11] * p.z + a.m[15];
  if (w) *w = ww;
  return {x, y, z};
}
// General 4x4 inverse (cofactor expansion), double precision.
bool inverse(const M4& a, M4* out) {
  const double* m = a.m;
  double inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  if (det == 0.0 || !std::isfinite(det)) return false;
  for (int i = 0; i < 16; ++i) out->m[i] = inv[i] / det;
  return true;
}
// glam from_scale_rotation_translation: T * R(q) * S, q = (x, y, z, w) normalised.
M4 srt(double s, const float q_in[4], const float t[3]) {
  double x = q_in[0], y = q_in[1], z = q_in[2], w = q_in[3];
  double n = std::sqrt(x * x + y * y + z * z + w * w);
  x /= n; y /= n; z /= n; w /= n;
  double xx = x * x, yy = y * y, zz = z * z, xy = x * y, xz = x * z, yz = y * z, wx = w * x, wy = w * y, wz = w * z;
  M4 r{};
  r.m[0] = (1 - 2 * (yy + zz)) * s; r.m[1] = 2 * (xy + wz) * s; r.m[2] = 2 * (xz - wy) * s; r.m[3] = 0;
  r.m[4] = 2 * (xy - wz) * s; r.m[5] = (1 - 2 * (xx + zz)) * s; r.m[6] = 2 * (yz + wx) * s; r.m[7] = 0;
  r.m[8] = 2 * (xz + wy) * s; r.m[9] = 2 * (yz - wx) * s; r.m[10] = (1 - 2 * (xx + yy)) * s; r.m[11] = 0;
  r.m[12] = t[0]; r.m[13] = t[1]; r.m[14] = t[2]; r.m[15] = 1;
  return r;
}
void to_f32(const M4& a, float* out) {
  for (int i = 0; i < 16; ++i) out[i] = (float)a.m[i];
}
M4 from_f32(const float* in) {
  M4 r;
  for (int i = 0; i < 16; ++i) r.m[i] = in[i];
  return r;
}

}  // namespace

extern "C" {

// Mesh kinds.
enum { SG_BOX = 0, SG_SPHERE = 1, SG_TORUS = 2 };

typedef struct sg_shape {
  uint32_t kind;        // SG_BOX: half extents p; SG_SPHERE: radius p[0]; SG_TORUS: major p[0], minor p[1] (axis y)
  uint32_t resolution;  // SDF grid resolution (>= 2)
  float p[3];
} sg_shape;

typedef struct sg_object {
  uint32_t mesh;
  float position[3];
  float rotation[4];  // quaternion x, y, z, w
  float scale;
  float albedo[3];
  float rough, metal;
} sg_object;

typedef struct sg_camera {
  float position[3];
  float rotation[4];  // quaternion x, y, z, w (camera looks down its local -Z, +Y up)
  float fov_y, near_plane, far_plane;
} sg_camera;

}  // extern "C"

namespace {

double sdf_local(const sg_shape& s, d3 p) {
  if (s.kind == SG_BOX) {
    d3 q = {std::fabs(p.x) - s.p[0], std::fabs(p.y) - s.p[1], std::fabs(p.z) - s.p[2]};
    d3 o = {std::max(q.x, 0.0), std::max(q.y, 0.0), std::max(q.z, 0.0)};
    return len(o) + std::min(std::max(q.x, std::max(q.y, q.z)), 0.0);
  }
  if (s.kind == SG_SPHERE) return len(p) - s.p[0];
  double a = std::sqrt(p.x * p.x + p.z * p.z) - s.p[0];
  return std::sqrt(a * a + p.y * p.y) - s.p[1];
}

d3 normal_local(const sg_shape& s, d3 p) {
  if (s.kind == SG_SPHERE) return nrm(p);
  if (s.kind == SG_TORUS) {
    double rxz = std::sqrt(p.x * p.x + p.z * p.z);
    d3 c = rxz > 0 ? d3{p.x / rxz * s.p[0], 0.0, p.z / rxz * s.p[0]} : d3{s.p[0], 0.0, 0.0};
    return nrm(p - c);
  }
  // box: axis of the largest normalised coordinate
  double ax = std::fabs(p.x) / s.p[0], ay = std::fabs(p.y) / s.p[1], az = std::fabs(p.z) / s.p[2];
  if (ax >= ay && ax >= az) return {p.x > 0 ? 1.0 : -1.0, 0, 0};
  if (ay >= az) return {0, p.y > 0 ? 1.0 : -1.0, 0};
  return {0, 0, p.z > 0 ? 1.0 : -1.0};
}

// Tessellation whose vertices include the shape's axis extremes (exact AABB).
void tessellate(const sg_shape& s, std::vector<float>& pos, std::vector<uint32_t>& idx) {
  pos.clear();
  idx.clear();
  if (s.kind == SG_BOX) {
    // 12 triangles, one vertex per triangle corner (36/36 like the exporter's cubes).
    static const int tri[36] = {0, 2, 6, 0, 6, 4, 1, 5, 7, 1, 7, 3, 0, 4, 5, 0, 5, 1,
                                2, 3, 7, 2, 7, 6, 0, 1, 3, 0, 3, 2, 4, 6, 7, 4, 7, 5};
    for (int k = 0; k < 36; ++k) {
      int c = tri[k];
      pos.push_back((c & 1) ? s.p[0] : -s.p[0]);
      pos.push_back((c & 2) ? s.p[1] : -s.p[1]);
      pos.push_back((c & 4) ? s.p[2] : -s.p[2]);
      idx.push_back((uint32_t)k);
    }
    return;
  }
  const int nu = 32, nv = 16;  // multiples of 4
  for (int j = 0; j <= nv; ++j)
    for (int i = 0; i < nu; ++i) {
      double u = 2.0 * M_PI * i / nu, v = (s.kind == SG_SPHERE ? M_PI * j / nv : 2.0 * M_PI * j / nv);
      d3 p;
      if (s.kind == SG_SPHERE) {
        double r = s.p[0];
        p = {r * std::sin(v) * std::cos(u), r * std::cos(v), r * std::sin(v) * std::sin(u)};
      } else {
        double R = s.p[0], r = s.p[1];
        double rr = R + r * std::cos(v);
        p = {rr * std::cos(u), r * std::sin(v), rr * std::sin(u)};
      }
      // Snap tiny values so the exact extremes are reproduced.
      pos.push_back((float)p.x);
      pos.push_back((float)p.y);
      pos.push_back((float)p.z);
    }
  for (int j = 0; j < nv; ++j)
    for (int i = 0; i < nu; ++i) {
      uint32_t a = j * nu + i, b = j * nu + (i + 1) % nu, c = (j + 1) * nu + i, d = (j + 1) * nu + (i + 1) % nu;
      idx.insert(idx.end(), {a, c, b, b, c, d});
    }
}

// Winds every triangle counter-clockwise seen from outside (the front face of
// the raster pass, renderer.rs:1384-1386): its geometric normal points along
// the shape's outward normal at the triangle's centroid.
void wind_outward(const sg_shape& s, const std::vector<float>& pos, std::vector<uint32_t>& idx) {
  for (size_t t = 0; t + 2 < idx.size(); t += 3) {
    d3 v[3];
    for (int k = 0; k < 3; ++k) v[k] = d3{pos[3 * idx[t + k]], pos[3 * idx[t + k] + 1], pos[3 * idx[t + k] + 2]};
    const d3 g = cross(v[1] - v[0], v[2] - v[0]);
    const d3 c = (v[0] + v[1] + v[2]) * (1.0 / 3.0);
    if (dot(g, normal_local(s, c)) < 0) std::swap(idx[t + 1], idx[t + 2]);
  }
}

// Ray / shape intersection in local space; returns t > tmin or -1.
double intersect_local(const sg_shape& s, d3 o, d3 d, double tmin) {
  // Slab test against the local AABB first (also the box answer).
  d3 h;
  if (s.kind == SG_BOX)
    h = {s.p[0], s.p[1], s.p[2]};
  else if (s.kind == SG_SPHERE)
    h = {s.p[0], s.p[0], s.p[0]};
  else
    h = {s.p[0] + s.p[1], s.p[1], s.p[0] + s.p[1]};
  double t0 = -1e300, t1 = 1e300;
  const double oo[3] = {o.x, o.y, o.z}, dd[3] = {d.x, d.y, d.z}, hh[3] = {h.x, h.y, h.z};
  for (int a = 0; a < 3; ++a) {
    if (std::fabs(dd[a]) < 1e-300) {
      if (oo[a] < -hh[a] || oo[a] > hh[a]) return -1;
      continue;
    }
    double ta = (-hh[a] - oo[a]) / dd[a], tb = (hh[a] - oo[a]) / dd[a];
    if (ta > tb) std::swap(ta, tb);
    t0 = std::max(t0, ta);
    t1 = std::min(t1, tb);
  }
  if (t0 > t1 || t1 < tmin) return -1;
  if (s.kind == SG_BOX) return t0 > tmin ? t0 : (t1 > tmin ? t1 : -1);
  if (s.kind == SG_SPHERE) {
    double b = dot(o, d), c = dot(o, o) - (double)s.p[0] * s.p[0];
    double disc = b * b - c;
    if (disc < 0) return -1;
    double sq = std::sqrt(disc);
    double ta = -b - sq, tb = -b + sq;
    if (ta > tmin) return ta;
    if (tb > tmin) return tb;
    return -1;
  }
  // Torus: sphere-trace the exact distance from the AABB entry.
  double t = std::max(t0, tmin);
  for (int i = 0; i < 512 && t <= t1 + 1e-9; ++i) {
    double dist = sdf_local(s, o + d * t);
    if (dist < 1e-9) return t;
    t += dist;
  }
  return -1;
}

struct Scene {
  std::vector<sg_shape> shapes;
  std::vector<sg_object> objects;
  std::vector<M4> model, model_inv, previous;
  std::vector<irt_mesh> meshes;
  std::vector<irt_light> lights;
  std::vector<irt_node> nodes;
  std::vector<float> atlas;
  uint32_t atlas_dims[3];
  uint32_t tile;
};

void place_objects(Scene& sc, const sg_object* objs, uint32_t n) {
  const bool first = sc.objects.empty();
  std::vector<M4> old = sc.model;
  sc.objects.assign(objs, objs + n);
  sc.model.resize(n);
  sc.model_inv.resize(n);
  sc.previous.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    sc.model[i] = srt(objs[i].scale, objs[i].rotation, objs[i].position);
    inverse(sc.model[i], &sc.model_inv[i]);
    sc.previous[i] = (first || i >= old.size()) ? sc.model[i] : old[i];
  }
  // hierarchy() over the world AABBs of the SDF boxes (renderer.rs:385-455).
  std::vector<float> lo(3 * n), hi(3 * n);
  for (uint32_t i = 0; i < n; ++i) {
    float m[16];
    to_f32(sc.model[i], m);
    const irt_mesh& mesh = sc.meshes[objs[i].mesh];
    irt_host::transform_box(m, mesh.lo, mesh.hi, &lo[3 * i], &hi[3 * i]);
  }
  sc.nodes.assign(2 * n - 1, irt_node{});
  irt_host::hierarchy(lo.data(), hi.data(), n, sc.nodes.data());
}

// Nearest analytic hit over all objects (world space).
bool trace(const Scene& sc, d3 o, d3 d, double tmin, double* t_hit, uint32_t* obj, d3* n_world) {
  double best = 1e300;
  uint32_t bo = 0;
  for (uint32_t i = 0; i < sc.objects.size(); ++i) {
    const M4& inv = sc.model_inv[i];
    const double s = sc.objects[i].scale;
    d3 ol = xform_point(inv, o);
    d3 dl = {inv.m[0] * d.x + inv.m[4] * d.y + inv.m[8] * d.z, inv.m[1] * d.x + inv.m[5] * d.y + inv.m[9] * d.z,
             inv.m[2] * d.x + inv.m[6] * d.y + inv.m[10] * d.z};
    // Uniform scale: |dl| = 1/s, local t = world t.
    d3 dn = dl * s;
    double tl = intersect_local(sc.shapes[sc.objects[i].mesh], ol, dn, tmin / s);
    if (tl > 0) {
      double tw = tl * s;
      if (tw < best && tw > tmin) {
        best = tw;
        bo = i;
      }
    }
  }
  if (best >= 1e300) return false;
  *t_hit = best;
  *obj = bo;
  d3 pw = o + d * best;
  const M4& inv = sc.model_inv[bo];
  d3 pl = xform_point(inv, pw);
  d3 nl = normal_local(sc.shapes[sc.objects[bo].mesh], pl);
  const M4& m = sc.model[bo];
  d3 nw = {m.m[0] * nl.x + m.m[4] * nl.y + m.m[8] * nl.z, m.m[1] * nl.x + m.m[5] * nl.y + m.m[9] * nl.z,
           m.m[2] * nl.x + m.m[6] * nl.y + m.m[10] * nl.z};
  *n_world = nrm(nw);
  return true;
}

// Camera::matrix — D3D-style RH perspective (depth 0..1) times inverse(T*R) (scene.rs:118-125).
M4 camera_matrix(const sg_camera& c, double aspect) {
  double sh = std::sin(0.5 * c.fov_y), ch = std::cos(0.5 * c.fov_y);
  double h = ch / sh, w = h / aspect, r = c.far_plane / (c.near_plane - c.far_plane);
  M4 p{};
  p.m[0] = w;
  p.m[5] = h;
  p.m[10] = r;
  p.m[11] = -1.0;
  p.m[14] = r * c.near_plane;
  float one_t[3] = {c.position[0], c.position[1], c.position[2]};
  M4 rt = srt(1.0, c.rotation, one_t);
  M4 view;
  inverse(rt, &view);
  return mul(p, view);
}

}  // namespace

extern "C" {

typedef struct sg_scene sg_scene;

// Builds meshes (SDF bake, tessellation, GpuMesh records), objects, BVH, atlas.
sg_scene* sg_create(const sg_shape* shapes, uint32_t n_shapes, const sg_object* objs, uint32_t n_objs,
                    const irt_light* lights, uint32_t n_lights) {
  if (!shapes || !objs || n_shapes == 0 || n_objs == 0) return nullptr;
  Scene* sc = new Scene();
  sc->shapes.assign(shapes, shapes + n_shapes);
  sc->lights.assign(lights, lights + n_lights);
  uint32_t tile = 2;
  for (auto& s : sc->shapes) tile = std::max(tile, s.resolution);
  sc->tile = tile;
  uint32_t grid[3];
  irt_host::atlas_grid(n_shapes, grid);
  sc->atlas_dims[0] = tile * grid[0];
  sc->atlas_dims[1] = tile * grid[1];
  sc->atlas_dims[2] = tile * grid[2];
  const size_t aw = sc->atlas_dims[0], ah = sc->atlas_dims[1];
  sc->atlas.assign(aw * ah * sc->atlas_dims[2], 0.0f);
  sc->meshes.resize(n_shapes);
  uint32_t vtx_off = 0, idx_off = 0;
  for (uint32_t m = 0; m < n_shapes; ++m) {
    const sg_shape& s = sc->shapes[m];
    const uint32_t res = s.resolution;
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    tessellate(s, pos, idx);
    wind_outward(s, pos, idx);
    const uint32_t nv = (uint32_t)(pos.size() / 3);
    float glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t v = 0; v < nv; ++v)
      for (int a = 0; a < 3; ++a) {
        glo[a] = std::min(glo[a], pos[v * 3 + a]);
        ghi[a] = std::max(ghi[a], pos[v * 3 + a]);
      }
    // Exporter grid convention (export_scene.py:135-158), computed in double then stored as float.
    double diag = std::sqrt((double)(ghi[0] - glo[0]) * (ghi[0] - glo[0]) + (double)(ghi[1] - glo[1]) * (ghi[1] - glo[1]) +
                            (double)(ghi[2] - glo[2]) * (ghi[2] - glo[2]));
    double pad = diag / (res - 1) * 2.0;
    double lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = glo[a] - pad;
      hi[a] = ghi[a] + pad;
    }
    double extent = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                              (hi[2] - lo[2]) * (hi[2] - lo[2]));
    float flo[3], fhi[3];
    for (int a = 0; a < 3; ++a) {
      flo[a] = (float)lo[a];
      fhi[a] = (float)hi[a];
    }
    std::vector<float> samples((size_t)res * res * res);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < (int)res; ++z)
      for (uint32_t y = 0; y < res; ++y)
        for (uint32_t x = 0; x < res; ++x) {
          d3 p = {lo[0] + (hi[0] - lo[0]) * x / (res - 1), lo[1] + (hi[1] - lo[1]) * y / (res - 1),
                  lo[2] + (hi[2] - lo[2]) * (double)z / (res - 1)};
          samples[x + (size_t)res * (y + (size_t)res * z)] = (float)(sdf_local(s, p) - extent * 2e-6);
        }
    const uint32_t origin[3] = {(m % grid[0]) * tile, ((m / grid[0]) % grid[1]) * tile, (m / (grid[0] * grid[1])) * tile};
    for (uint32_t z = 0; z < res; ++z)
      for (uint32_t y = 0; y < res; ++y)
        std::memcpy(&sc->atlas[(origin[0]) + aw * ((origin[1] + y) + ah * (size_t)(origin[2] + z))],
                    &samples[(size_t)res * (y + (size_t)res * z)], res * sizeof(float));
    // GpuMesh record (renderer.rs:818-885) in float arithmetic as glam would.
    float spacing[3];
    for (int a = 0; a < 3; ++a) spacing[a] = (fhi[a] - flo[a]) / (float)(res - 1);
    float sl = std::sqrt(spacing[0] * spacing[0] + spacing[1] * spacing[1] + spacing[2] * spacing[2]);
    const bool is_box = irt_host::is_box_geometry(pos.data(), nv, 3, idx.data(), (uint32_t)idx.size(), glo, ghi);
    const float lip = irt_host::field_lipschitz_bound(samples.data(), res, spacing);
    float center[3] = {(glo[0] + ghi[0]) * 0.5f, (glo[1] + ghi[1]) * 0.5f, (glo[2] + ghi[2]) * 0.5f};
    float radius = 0.f;
    for (uint32_t v = 0; v < nv; ++v) {
      float dx = pos[v * 3] - center[0], dy = pos[v * 3 + 1] - center[1], dz = pos[v * 3 + 2] - center[2];
      radius = std::max(radius, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
    radius = radius * 1.000001f + 0.000001f;
    irt_mesh& g = sc->meshes[m];
    std::memset(&g, 0, sizeof(g));
    for (int a = 0; a < 3; ++a) {
      g.lo[a] = flo[a];
      g.hi[a] = fhi[a];
      g.grid[a] = 1.0f / spacing[a];
      g.geometry_lo[a] = glo[a];
      g.geometry_hi[a] = ghi[a];
      g.enclosing_sphere[a] = center[a];
    }
    g.grid[3] = sl;
    g.geometry_lo[3] = is_box ? 1.0f : 0.0f;
    g.geometry_hi[3] = 1.0f / lip;
    g.enclosing_sphere[3] = radius;
    g.geometry[0] = vtx_off;
    g.geometry[1] = nv;
    g.geometry[2] = idx_off;
    g.geometry[3] = (uint32_t)idx.size();
    g.field[0] = origin[0];
    g.field[1] = origin[1];
    g.field[2] = origin[2];
    g.field[3] = res;
    vtx_off += nv;
    idx_off += (uint32_t)idx.size();
  }
  for (uint32_t i = 0; i < n_objs; ++i)
    if (objs[i].mesh >= n_shapes) {
      delete sc;
      return nullptr;
    }
  place_objects(*sc, objs, n_objs);
  return reinterpret_cast<sg_scene*>(sc);
}

void sg_destroy(sg_scene* s) { delete reinterpret_cast<Scene*>(s); }

// info: meshes, objects, nodes, lights, atlas w, h, d, tile
void sg_info(const sg_scene* s, uint32_t info[8]) {
  const Scene* sc = reinterpret_cast<const Scene*>(s);
  info[0] = (uint32_t)sc->meshes.size();
  info[1] = (uint32_t)sc->objects.size();
  info[2] = (uint32_t)sc->nodes.size();
  info[3] = (uint32_t)sc->lights.size();
  info[4] = sc->atlas_dims[0];
  info[5] = sc->atlas_dims[1];
  info[6] = sc->atlas_dims[2];
  info[7] = sc->tile;
}

// Scene records: objects() (renderer.rs:359-384), meshes, hierarchy, lights, atlas.
void sg_records(const sg_scene* s, irt_mesh* meshes, irt_object* objects, irt_node* nodes, irt_light* lights) {
  const Scene* sc = reinterpret_cast<const Scene*>(s);
  if (meshes) std::copy(sc->meshes.begin(), sc->meshes.end(), meshes);
  if (nodes) std::copy(sc->nodes.begin(), sc->nodes.end(), nodes);
  if (lights) std::copy(sc->lights.begin(), sc->lights.end(), lights);
  if (objects) {
    for (size_t i = 0; i < sc->objects.size(); ++i) {
      irt_object& o = objects[i];
      to_f32(sc->model[i], o.model);
      to_f32(sc->previous[i], o.previous);
      M4 inv;
      inverse(from_f32(o.model), &inv);
      to_f32(inv, o.inverse);
      const irt_mesh& mesh = sc->meshes[sc->objects[i].mesh];
      const float res1 = (float)(mesh.field[3] - 1);
      float sp[3];
      for (int a = 0; a < 3; ++a) sp[a] = (mesh.hi[a] - mesh.lo[a]) / res1;
      float scale = std::sqrt(o.model[0] * o.model[0] + o.model[1] * o.model[1] + o.model[2] * o.model[2]);
      float voxel = std::sqrt(sp[0] * sp[0] + sp[1] * sp[1] + sp[2] * sp[2]) * scale;
      o.info[0] = sc->objects[i].mesh;
      o.info[1] = (uint32_t)i;
      std::memcpy(&o.info[2], &scale, 4);
      std::memcpy(&o.info[3], &voxel, 4);
    }
  }
}

// Triangle geometry in the reference's vertex and material records (the
// vertex buffer and materials of renderer.rs:951-981, 1061-1106):
// info = vertex count, index count, material count.
void sg_geometry_info(const sg_scene* s, uint32_t info[3]) {
  const Scene* sc = reinterpret_cast<const Scene*>(s);
  uint32_t nv = 0, ni = 0;
  for (const irt_mesh& m : sc->meshes) {
    nv += m.geometry[1];
    ni += m.geometry[3];
  }
  info[0] = nv;
  info[1] = ni;
  info[2] = (uint32_t)sc->objects.size();
}

// Vertices carry the shape's analytic normal (boxes: the face normal of the
// one triangle each vertex belongs to), a unit tangent orthogonal to it and
// a (u, v) parameter; indices are mesh-local.  Material i is object i's
// untextured albedo / roughness / metalness (objects use material index i).
void sg_geometry(const sg_scene* s, irt_vertex* vertices, uint32_t* indices, irt_material* materials) {
  const Scene* sc = reinterpret_cast<const Scene*>(s);
  for (size_t m = 0; m < sc->shapes.size(); ++m) {
    const sg_shape& shape = sc->shapes[m];
    const irt_mesh& g = sc->meshes[m];
    std::vector<float> pos;
    std::vector<uint32_t> idx;
    tessellate(shape, pos, idx);
    wind_outward(shape, pos, idx);
    const uint32_t nv = (uint32_t)(pos.size() / 3);
    std::vector<d3> normal(nv);
    if (shape.kind == SG_BOX) {
      for (size_t t = 0; t + 2 < idx.size(); t += 3) {
        d3 c = {0, 0, 0};
        for (int k = 0; k < 3; ++k)
          c = c + d3{pos[3 * idx[t + k]], pos[3 * idx[t + k] + 1], pos[3 * idx[t + k] + 2]} * (1.0 / 3.0);
        const d3 n = normal_local(shape, c);
        for (int k = 0; k < 3; ++k) normal[idx[t + k]] = n;
      }
    } else {
      for (uint32_t v = 0; v < nv; ++v) normal[v] = normal_local(shape, d3{pos[3 * v], pos[3 * v + 1], pos[3 * v + 2]});
    }
    for (uint32_t v = 0; v < nv; ++v) {
      irt_vertex& out = vertices[g.geometry[0] + v];
      std::memset(&out, 0, sizeof(out));
      const d3 n = normal[v];
      const d3 axis = std::fabs(n.y) < 0.9 ? d3{0, 1, 0} : d3{1, 0, 0};
      const d3 t = nrm(cross(axis, n));
      for (int a = 0; a < 3; ++a) out.position[a] = pos[3 * v + a];
      out.position[3] = 1.0f;
      out.normal[0] = (float)n.x;
      out.normal[1] = (float)n.y;
      out.normal[2] = (float)n.z;
      out.tangent[0] = (float)t.x;
      out.tangent[1] = (float)t.y;
      out.tangent[2] = (float)t.z;
      out.tangent[3] = 1.0f;
      out.uv[0] = shape.kind == SG_BOX ? 0.0f : (float)((v % 32u) / 32.0);
      out.uv[1] = shape.kind == SG_BOX ? 0.0f : (float)((v / 32u) / 16.0);
    }
    std::copy(idx.begin(), idx.end(), indices + g.geometry[2]);
  }
  for (size_t i = 0; i < sc->objects.size(); ++i) {
    irt_material& mat = materials[i];
    std::memset(&mat, 0, sizeof(mat));
    const sg_object& o = sc->objects[i];
    for (int a = 0; a < 3; ++a) mat.color[a] = o.albedo[a];
    mat.color[3] = 1.0f;
    mat.params[0] = o.rough;
    mat.params[1] = o.metal;
    mat.params[2] = 1.0f;
    for (int k = 0; k < 4; ++k) mat.textures[k] = -1;
  }
}

void sg_atlas(const sg_scene* s, float* out) {
  const Scene* sc = reinterpret_cast<const Scene*>(s);
  std::copy(sc->atlas.begin(), sc->atlas.end(), out);
}

// Moves objects (keeps the previous transforms for motion) and rebuilds the BVH.
int sg_set_objects(sg_scene* s, const sg_object* objs, uint32_t n) {
  Scene* sc = reinterpret_cast<Scene*>(s);
  if (!objs || n == 0) return -1;
  for (uint32_t i = 0; i < n; ++i)
    if (objs[i].mesh >= sc->shapes.size()) return -1;
  place_objects(*sc, objs, n);
  return 0;
}
int sg_set_lights(sg_scene* s, const irt_light* lights, uint32_t n) {
  Scene* sc = reinterpret_cast<Scene*>(s);
  sc->lights.assign(lights, lights + n);
  return 0;
}

void sg_camera_matrix(const sg_camera* c, float aspect, float vp[16]) { to_f32(camera_matrix(*c, aspect), vp); }

// Analytic G-buffer of one frame (reference texture formats).
int sg_render_gbuffer(const sg_scene* s, const sg_camera* cam, const sg_camera* prev_cam, uint32_t W, uint32_t H,
                      float* depth, uint16_t* normal, uint16_t* color, uint16_t* params, uint16_t* motion) {
  const Scene* sc = reinterpret_cast<const Scene*>(s);
  const double aspect = (double)W / H;
  float vpf[16], pvpf[16];
  to_f32(camera_matrix(*cam, aspect), vpf);
  to_f32(camera_matrix(*prev_cam, aspect), pvpf);
  const M4 vp = from_f32(vpf), pvp = from_f32(pvpf);
  M4 inv_vp;
  inverse(vp, &inv_vp);
  const d3 eye = {cam->position[0], cam->position[1], cam->position[2]};
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < (int)H; ++y)
    for (uint32_t x = 0; x < W; ++x) {
      const size_t id = (size_t)y * W + x;
      double ndx = ((x + 0.5) / W) * 2.0 - 1.0, ndy = 1.0 - ((y + 0.5) / H) * 2.0;
      double w0, w1;
      d3 a = xform_point(inv_vp, {ndx, ndy, 0.0}, &w0);
      d3 b = xform_point(inv_vp, {ndx, ndy, 1.0}, &w1);
      a = a * (1.0 / w0);
      b = b * (1.0 / w1);
      d3 dir = nrm(b - a);
      double t;
      uint32_t obj;
      d3 n;
      float* dp = depth + id;
      uint16_t* np = normal + id * 4;
      uint16_t* cp = color + id * 4;
      uint16_t* pp = params + id * 2;
      uint16_t* mp = motion + id * 2;
      if (!trace(*sc, eye, dir, 1e-6, &t, &obj, &n)) {
        *dp = 1.0f;
        for (int k = 0; k < 4; ++k) np[k] = cp[k] = 0;
        pp[0] = pp[1] = mp[0] = mp[1] = 0;
        continue;
      }
      d3 p = eye + dir * t;
      double cw;
      d3 clip = xform_point(vp, p, &cw);
      double dz = clip.z / cw;
      if (!(dz < 1.0) || dz < 0.0) {  // beyond the far plane / before near: background
        *dp = 1.0f;
        for (int k = 0; k < 4; ++k) np[k] = cp[k] = 0;
        pp[0] = pp[1] = mp[0] = mp[1] = 0;
        continue;
      }
      *dp = (float)dz;
      if (!(*dp < 1.0f)) *dp = std::nextafter(1.0f, 0.0f);
      if (dot(n, dir) > 0) n = n * -1.0;  // facing the camera
      np[0] = irt_host::f32_to_f16((float)n.x);
      np[1] = irt_host::f32_to_f16((float)n.y);
      np[2] = irt_host::f32_to_f16((float)n.z);
      np[3] = irt_host::f32_to_f16(1.0f);
      const sg_object& o = sc->objects[obj];
      for (int k = 0; k < 3; ++k) cp[k] = irt_host::f32_to_f16(o.albedo[k]);
      cp[3] = irt_host::f32_to_f16(1.0f);
      pp[0] = irt_host::f32_to_f16(o.rough);
      pp[1] = irt_host::f32_to_f16(o.metal);
      // Motion: previous position of the surface point through the previous model.
      d3 pl = xform_point(sc->model_inv[obj], p);
      d3 pprev = xform_point(sc->previous[obj], pl);
      double pw;
      d3 pc = xform_point(pvp, pprev, &pw);
      double cx = clip.x / cw, cy = clip.y / cw, px = pc.x / pw, py = pc.y / pw;
      mp[0] = irt_host::f32_to_f16((float)((cx - px) * 0.5));
      mp[1] = irt_host::f32_to_f16((float)((cy - py) * -0.5));
    }
  return 0;
}

// VPL pool: `paths` light paths of `bounces` surface hits each; one VPL per
// hit (position, geometric normal, flux U[0,1]^3 * flux_scale).  Directions:
// uniform sphere from a uniformly chosen light, then cosine-weighted bounces.
int sg_vpls(const sg_scene* s, uint64_t seed, uint32_t paths, uint32_t bounces, float flux_scale, irt_vpl* out) {
  const Scene* sc = reinterpret_cast<const Scene*>(s);
  const uint32_t L = (uint32_t)sc->lights.size();
#pragma omp parallel for schedule(static)
  for (int pth = 0; pth < (int)paths; ++pth) {
    std::mt19937_64 rng(seed * 0x9E3779B97F4A7C15ull + (uint64_t)pth * 0xBF58476D1CE4E5B9ull + 1);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    irt_vpl* v = out + (size_t)pth * bounces;
    std::memset(v, 0, sizeof(irt_vpl) * bounces);
    if (L == 0) continue;
    const irt_light& l = sc->lights[std::min<uint32_t>((uint32_t)(U(rng) * L), L - 1)];
    d3 o = {l.position_strength[0], l.position_strength[1], l.position_strength[2]};
    double z = 2.0 * U(rng) - 1.0, phi = 2.0 * M_PI * U(rng), r = std::sqrt(std::max(0.0, 1.0 - z * z));
    d3 d = {r * std::cos(phi), r * std::sin(phi), z};
    for (uint32_t b = 0; b < bounces; ++b) {
      double t;
      uint32_t obj;
      d3 n;
      if (!trace(*sc, o, d, 1e-4, &t, &obj, &n)) break;
      if (dot(n, d) > 0) n = n * -1.0;
      d3 p = o + d * t;
      irt_vpl& out_v = v[b];
      out_v.position[0] = (float)p.x;
      out_v.position[1] = (float)p.y;
      out_v.position[2] = (float)p.z;
      out_v.position[3] = 1.0f;
      out_v.normal[0] = (float)n.x;
      out_v.normal[1] = (float)n.y;
      out_v.normal[2] = (float)n.z;
      for (int k = 0; k < 3; ++k) out_v.flux[k] = (float)(U(rng) * flux_scale);
      // cosine-weighted bounce about n
      d3 axis = std::fabs(n.z) > 0.9 ? d3{1, 0, 0} : d3{0, 0, 1};
      d3 tt = nrm(cross(axis, n)), bb = cross(n, tt);
      double u1 = U(rng), ph = 2.0 * M_PI * U(rng);
      double sr = std::sqrt(u1);
      d = nrm(tt * (sr * std::cos(ph)) + bb * (sr * std::sin(ph)) + n * std::sqrt(std::max(0.0, 1.0 - u1)));
      o = p + n * 1e-4;
    }
  }
  return 0;
}

}  // extern "C"
