// The G-buffer pass — gbuffer.wgsl:1-12 with the two raster pipelines of
// renderer.rs:1370-1415 (depth pre-pass: depth test Less with writes,
// cleared to 1; G-buffer pass: depth test Equal; back faces culled,
// counter-clockwise front faces; targets cleared to 0) — as compute kernels.
//
// Rasterisation without clipping: each instance triangle's clip-space
// corners V_k = (x_k, y_k, w_k) give three edge functions
// E_i(p) = (V_j x V_k) . (px, py, 1) over NDC pixel centres (2-D homogeneous
// rasterisation, Olano & Greer 1997, "Triangle scan conversion using 2D
// homogeneous coordinates").  A pixel is covered when all three are >= 0
// (the triangle faces the camera when det(V_0, V_1, V_2) > 0); then
// E_i / sum(E) are the perspective-correct barycentrics and the depth
// z_ndc = sum(E_i z_i) / det is affine in the pixel centre, as a depth
// buffer interpolates it.  Corners behind the eye need no clipping: such
// points get negative weights, and the near and far planes are the depth
// range test 0 <= z_ndc < 1.  An edge shared by two triangles has exactly
// negated edge functions in the two (same corners, swapped cross product),
// so no pixel centre falls between them.
//
// Depth test and draw order: a pixel's winner is the least 64-bit key
// (depth bits << 32 | ~draw index) over its covering fragments
// (atomicMin): the nearest depth, and among equal depths the fragment drawn
// last (objects in order, triangles in index order), which is the one the
// Equal-test G-buffer pass leaves in the targets.
//
// Work: k_gb_setup — one lane per instance triangle: corners, culling, the
// edge/depth planes, the 16x16 pixel tiles of its screen bounds and the
// vertex-stage outputs of the drawn ones;
// an inclusive scan of the tile counts; k_gb_raster — persistent
// workgroups, one 16x16 tile of one triangle per step, one pixel per lane;
// k_gb_resolve — one lane per pixel: the winner's attributes, interpolated
// and shaded as the fragment shader does, written in the reference formats;
// it leaves the visibility buffer cleared for the next frame.
#pragma once

#include "irt_material.h"

namespace irt {

constexpr uint32_t GB_TILE = 16;
constexpr uint64_t GB_CLEAR_KEY = (uint64_t)0x3f800000u << 32 | 0xffffffffu;  // depth 1.0, no triangle

// Per object: vp * model and previous_vp * previous (frame.vp*o.model,
// frame.previous_vp*o.previous of gbuffer.wgsl:3), column-major.
struct GbObject {
  float mvp[16];
  float pmvp[16];
};

// Per instance triangle, written by k_gb_setup.
struct __align__(16) GbTri {
  float e[9];      // edge functions: E_i(p) = e[3i] px + e[3i+1] py + e[3i+2]
  float z[3];      // z_ndc(p) = z[0] px + z[1] py + z[2]
  uint32_t bx0, by0, bw, pad;  // first tile and tiles per row of the screen bounds
};

// Per drawn triangle, written by k_gb_setup: the vertex-stage outputs the
// resolve interpolates (gbuffer.wgsl:2-6): world normals, current and
// previous clip x, y, w of the three corners, and the material.
struct __align__(16) GbAttr {
  float n[9];
  float cur[9];
  float prev[9];
  uint32_t material;
  float uv[6];       // corner uv
  float tangent[12]; // corner tangents: normalize(model * (t.xyz, 0)), t.w
};

struct GbScene {
  const irt_object* objects;
  const irt_mesh* meshes;
  const irt_vertex* vertices;
  const uint32_t* indices;
  const irt_material* materials;
  const uint32_t* obj_tri_base;  // n_objects + 1 prefix sums of instance triangles
  uint32_t n_objects;
  TextureView tex;
};

// (vp * m) with each entry a left-to-right sum over k of vp[k][r] * m[c][k].
__device__ __forceinline__ void mat4_mul(const float* a, const float* b, float* out) {
  for (int c = 0; c < 4; ++c)
    for (int r = 0; r < 4; ++r)
      out[c * 4 + r] = a[r] * b[c * 4] + a[4 + r] * b[c * 4 + 1] + a[8 + r] * b[c * 4 + 2] + a[12 + r] * b[c * 4 + 3];
}

__global__ void k_gb_objects(const irt_object* __restrict__ objects, uint32_t n, irt_frame F, GbObject* out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  mat4_mul(F.vp, objects[i].model, out[i].mvp);
  mat4_mul(F.previous_vp, objects[i].previous, out[i].pmvp);
}

// Object of an instance triangle: the last object whose base is <= t.
__device__ __forceinline__ uint32_t object_of(const uint32_t* base, uint32_t n_objects, uint32_t t) {
  uint32_t lo = 0, hi = n_objects;  // base[lo] <= t < base[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (base[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ float ndc_x(uint32_t x, uint32_t W) { return ((float)x + 0.5f) / (float)W * 2.0f - 1.0f; }
__device__ __forceinline__ float ndc_y(uint32_t y, uint32_t H) { return 1.0f - ((float)y + 0.5f) / (float)H * 2.0f; }

// Corners of instance triangle t in clip space (frame.vp*o.model*p) and the
// vertex indices it was built from.
__device__ __forceinline__ void tri_corners(const GbScene& sc, const GbObject* objs, uint32_t t, uint32_t& obj,
                                            uint32_t vi[3], float4 clip[3]) {
  obj = object_of(sc.obj_tri_base, sc.n_objects, t);
  const uint32_t prim = t - sc.obj_tri_base[obj];
  const irt_mesh& me = sc.meshes[sc.objects[obj].info[0]];
  const float* m = objs[obj].mvp;
  for (int k = 0; k < 3; ++k) {
    vi[k] = me.geometry[0] + sc.indices[me.geometry[2] + 3u * prim + (uint32_t)k];
    const float* p = sc.vertices[vi[k]].position;
    clip[k] = mat_mul(m, p[0], p[1], p[2], p[3]);
  }
}

// Edge functions (rows of adj(V)) and det(V) for V = [x; y; w] of the corners.
__device__ __forceinline__ float edge_planes(const float4 c[3], float e[9]) {
  const v3 V0 = v3{c[0].x, c[0].y, c[0].w}, V1 = v3{c[1].x, c[1].y, c[1].w}, V2 = v3{c[2].x, c[2].y, c[2].w};
  const v3 a = v3{V1.y * V2.z - V1.z * V2.y, V1.z * V2.x - V1.x * V2.z, V1.x * V2.y - V1.y * V2.x};
  const v3 b = v3{V2.y * V0.z - V2.z * V0.y, V2.z * V0.x - V2.x * V0.z, V2.x * V0.y - V2.y * V0.x};
  const v3 d = v3{V0.y * V1.z - V0.z * V1.y, V0.z * V1.x - V0.x * V1.z, V0.x * V1.y - V0.y * V1.x};
  e[0] = a.x; e[1] = a.y; e[2] = a.z;
  e[3] = b.x; e[4] = b.y; e[5] = b.z;
  e[6] = d.x; e[7] = d.y; e[8] = d.z;
  return V0.x * a.x + V0.y * a.y + V0.z * a.z;
}

__device__ __forceinline__ float plane_at(const float* e, float px, float py) { return e[0] * px + e[1] * py + e[2]; }

__global__ __launch_bounds__(256) void k_gb_setup(GbScene sc, const GbObject* __restrict__ objs, uint32_t n_tris,
                                                  uint32_t W, uint32_t H, GbTri* __restrict__ tris,
                                                  GbAttr* __restrict__ attrs, uint32_t* __restrict__ tiles) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_tris) return;
  uint32_t obj, vi[3];
  float4 c[3];
  tri_corners(sc, objs, t, obj, vi, c);
  GbTri g;
  const float det = edge_planes(c, g.e);
  uint32_t count = 0;
  if (det > 0.0f) {  // front-facing (counter-clockwise), not degenerate
    for (int k = 0; k < 3; ++k)
      g.z[k] = (c[0].z * g.e[k] + c[1].z * g.e[3 + k] + c[2].z * g.e[6 + k]) / det;
    // screen bounds of the part in front of the near plane (clip z >= 0; the
    // polygon clipped against it), one pixel of margin; the whole screen
    // when a clipped corner is not in front of the eye
    const bool beyond_far = c[0].z >= c[0].w && c[1].z >= c[1].w && c[2].z >= c[2].w;
    float lx = INFINITY, hx = -INFINITY, ly = INFINITY, hy = -INFINITY;
    bool any = false, whole = false;
    for (int k = 0; k < 3; ++k) {
      const float4 a = c[k], b = c[(k + 1) % 3];
      float4 q[2];
      int nq = 0;
      if (a.z >= 0.0f) q[nq++] = a;
      if ((a.z >= 0.0f) != (b.z >= 0.0f)) {
        const float f = a.z / (a.z - b.z);
        q[nq++] = make_float4(a.x + (b.x - a.x) * f, a.y + (b.y - a.y) * f, 0.0f, a.w + (b.w - a.w) * f);
      }
      for (int j = 0; j < nq; ++j) {
        any = true;
        if (!(q[j].w > 0.0f)) {
          whole = true;
          continue;
        }
        const float sx = (q[j].x / q[j].w + 1.0f) * 0.5f * (float)W - 0.5f;
        const float sy = (1.0f - q[j].y / q[j].w) * 0.5f * (float)H - 0.5f;
        lx = fminf(lx, sx); hx = fmaxf(hx, sx); ly = fminf(ly, sy); hy = fmaxf(hy, sy);
      }
    }
    float x0 = 0.0f, x1 = (float)W - 1.0f, y0 = 0.0f, y1 = (float)H - 1.0f;
    if (!whole) {
      x0 = fmaxf(x0, floorf(lx) - 1.0f); x1 = fminf(x1, ceilf(hx) + 1.0f);
      y0 = fmaxf(y0, floorf(ly) - 1.0f); y1 = fminf(y1, ceilf(hy) + 1.0f);
    }
    if (any && !beyond_far && x0 <= x1 && y0 <= y1) {
      g.bx0 = (uint32_t)x0 / GB_TILE;
      g.by0 = (uint32_t)y0 / GB_TILE;
      g.bw = (uint32_t)x1 / GB_TILE - g.bx0 + 1u;
      count = g.bw * ((uint32_t)y1 / GB_TILE - g.by0 + 1u);
    }
  }
  g.pad = 0;
  if (!count) g.bx0 = g.by0 = g.bw = 0;
  tris[t] = g;
  tiles[t] = count;
  if (!count) return;
  const irt_object& ob = sc.objects[obj];
  GbAttr a;
  for (int k = 0; k < 3; ++k) {
    const irt_vertex& vx = sc.vertices[vi[k]];
    const float4 wn = mat_mul(ob.model, vx.normal[0], vx.normal[1], vx.normal[2], 0.0f);
    const v3 nk = normalize(v3{wn.x, wn.y, wn.z});
    const float4 pk = mat_mul(objs[obj].pmvp, vx.position[0], vx.position[1], vx.position[2], vx.position[3]);
    a.n[3 * k] = nk.x; a.n[3 * k + 1] = nk.y; a.n[3 * k + 2] = nk.z;
    a.cur[3 * k] = c[k].x; a.cur[3 * k + 1] = c[k].y; a.cur[3 * k + 2] = c[k].w;
    a.prev[3 * k] = pk.x; a.prev[3 * k + 1] = pk.y; a.prev[3 * k + 2] = pk.w;
    a.uv[2 * k] = vx.uv[0]; a.uv[2 * k + 1] = vx.uv[1];
    const float4 wt = mat_mul(ob.model, vx.tangent[0], vx.tangent[1], vx.tangent[2], 0.0f);
    const v3 tk = normalize(v3{wt.x, wt.y, wt.z});
    a.tangent[4 * k] = tk.x; a.tangent[4 * k + 1] = tk.y; a.tangent[4 * k + 2] = tk.z;
    a.tangent[4 * k + 3] = vx.tangent[3];
  }
  a.material = ob.info[1];
  attrs[t] = a;
}

// Persistent: workgroup steps over (triangle, tile) items; `scan` is the
// inclusive prefix sum of the tile counts (its last entry is the total).
__global__ __launch_bounds__(256) void k_gb_raster(const GbTri* __restrict__ tris, const uint32_t* __restrict__ scan,
                                                   uint32_t n_tris, uint32_t W, uint32_t H,
                                                   unsigned long long* __restrict__ vis) {
  if (n_tris == 0) return;
  const uint32_t total = scan[n_tris - 1];
  const uint32_t lx = threadIdx.x % GB_TILE, ly = threadIdx.x / GB_TILE;
  for (uint32_t item = blockIdx.x; item < total; item += gridDim.x) {
    uint32_t lo = 0, hi = n_tris - 1;  // first t with scan[t] > item
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (scan[mid] > item) hi = mid;
      else lo = mid + 1;
    }
    const uint32_t t = lo;
    const uint32_t local = item - (t ? scan[t - 1] : 0u);
    const GbTri& g = tris[t];
    const uint32_t x = (g.bx0 + local % g.bw) * GB_TILE + lx;
    const uint32_t y = (g.by0 + local / g.bw) * GB_TILE + ly;
    if (x >= W || y >= H) continue;
    const float px = ndc_x(x, W), py = ndc_y(y, H);
    const float e0 = plane_at(g.e, px, py), e1 = plane_at(g.e + 3, px, py), e2 = plane_at(g.e + 6, px, py);
    if (!(e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f && e0 + e1 + e2 > 0.0f)) continue;
    const float z = plane_at(g.z, px, py) + 0.0f;  // + 0: no negative zero
    if (!(z >= 0.0f && z < 1.0f)) continue;
    const unsigned long long key = (unsigned long long)__float_as_uint(z) << 32 | (unsigned long long)(~t);
    atomicMin(vis + (size_t)y * W + x, key);
  }
}

// Inclusive scan of the per-triangle tile counts (the raster's work list):
// k_scan_sums writes the sum of each block of SCAN_BLOCK_ITEMS counts,
// k_scan_carries turns those sums into exclusive block offsets in place (one
// workgroup, 256 sums per step with a running carry), and k_scan_blocks scans
// each block and adds its offset.  Block scans are Hillis-Steele over 256
// per-thread totals in LDS; each thread owns SCAN_THREAD_ITEMS consecutive
// counts.
constexpr uint32_t SCAN_THREAD_ITEMS = 2, SCAN_BLOCK_ITEMS = 256 * SCAN_THREAD_ITEMS;

// Inclusive scan of v over the 256 threads of the block (s: 256 words of
// LDS); *total gets the block's sum.  Both are read before the last barrier,
// so a following call may overwrite s.
__device__ __forceinline__ uint32_t block_scan_256(uint32_t v, uint32_t* s, uint32_t* total) {
  const uint32_t t = threadIdx.x;
  s[t] = v;
  __syncthreads();
  for (uint32_t off = 1; off < 256; off <<= 1) {
    const uint32_t add = t >= off ? s[t - off] : 0u;
    __syncthreads();
    s[t] += add;
    __syncthreads();
  }
  const uint32_t r = s[t];
  *total = s[255];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void k_scan_sums(const uint32_t* __restrict__ in, uint32_t n,
                                                   uint32_t* __restrict__ sums) {
  __shared__ uint32_t s[256];
  const uint32_t base = blockIdx.x * SCAN_BLOCK_ITEMS + threadIdx.x * SCAN_THREAD_ITEMS;
  uint32_t v = 0;
  for (uint32_t k = 0; k < SCAN_THREAD_ITEMS; ++k)
    if (base + k < n) v += in[base + k];
  uint32_t total;
  block_scan_256(v, s, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(256) void k_scan_carries(uint32_t* sums, uint32_t m) {
  __shared__ uint32_t s[256];
  uint32_t carry = 0;
  for (uint32_t c0 = 0; c0 < m; c0 += 256) {
    const uint32_t i = c0 + threadIdx.x;
    const uint32_t v = i < m ? sums[i] : 0u;
    uint32_t total;
    const uint32_t incl = block_scan_256(v, s, &total);
    if (i < m) sums[i] = carry + incl - v;
    carry += total;
  }
}

__global__ __launch_bounds__(256) void k_scan_blocks(const uint32_t* __restrict__ in, uint32_t n,
                                                     const uint32_t* __restrict__ offsets,
                                                     uint32_t* __restrict__ out) {
  __shared__ uint32_t s[256];
  const uint32_t base = blockIdx.x * SCAN_BLOCK_ITEMS + threadIdx.x * SCAN_THREAD_ITEMS;
  uint32_t v[SCAN_THREAD_ITEMS];
  uint32_t total = 0;
  for (uint32_t k = 0; k < SCAN_THREAD_ITEMS; ++k) {
    v[k] = base + k < n ? in[base + k] : 0u;
    total += v[k];
  }
  uint32_t block_total;
  uint32_t run = offsets[blockIdx.x] + block_scan_256(total, s, &block_total) - total;
  for (uint32_t k = 0; k < SCAN_THREAD_ITEMS; ++k) {
    run += v[k];
    if (base + k < n) out[base + k] = run;
  }
}

__global__ void k_gb_clear(unsigned long long* vis, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) vis[i] = GB_CLEAR_KEY;
}

// Store of a value the reference writes to an Rg16F target.
__device__ __forceinline__ void st_rg16(float2* p, uint32_t i, float a, float b) {
  p[i] = make_float2(round_f16(a), round_f16(b));
}
__device__ __forceinline__ void st_rg16(h2* p, uint32_t i, float a, float b) { p[i] = h2{(_Float16)a, (_Float16)b}; }

// Fragment stage (gbuffer.wgsl:7-11) of each pixel's winning triangle: its
// vertex-stage outputs (k_gb_setup) interpolated with the perspective-correct
// barycentrics at the pixel centre.
// Values are rounded as the reference's targets hold them (Depth32Float,
// Rgba16F, Rg16F) and stored in the profile's G-buffer formats.
template <class P>
__global__ __launch_bounds__(256) void k_gb_resolve(const GbTri* __restrict__ tris, const GbAttr* __restrict__ attrs,
                                                    const irt_material* __restrict__ materials, TextureView tex,
                                                    unsigned long long* __restrict__ vis, uint32_t W,
                                                    uint32_t H, float* __restrict__ depth,
                                                    typename P::N4* __restrict__ normal,
                                                    typename P::N4* __restrict__ color,
                                                    typename P::P2* __restrict__ params,
                                                    typename P::P2* __restrict__ motion) {
  const uint32_t x = blockIdx.x * GB_TILE + threadIdx.x % GB_TILE;
  const uint32_t y = blockIdx.y * GB_TILE + threadIdx.x / GB_TILE;
  if (x >= W || y >= H) return;
  const uint32_t id = y * W + x;
  const unsigned long long key = vis[id];
  vis[id] = GB_CLEAR_KEY;  // cleared for the next frame's raster
  if (key == GB_CLEAR_KEY) {
    depth[id] = 1.0f;
    const float4 z4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    st_r16(normal, id, z4);
    st_r16(color, id, z4);
    st_rg16(params, id, 0.0f, 0.0f);
    st_rg16(motion, id, 0.0f, 0.0f);
    return;
  }
  const uint32_t t = ~(uint32_t)(key & 0xffffffffu);
  const GbTri& g = tris[t];
  const GbAttr& a = attrs[t];
  const float px = ndc_x(x, W), py = ndc_y(y, H);
  const float E0 = plane_at(g.e, px, py), E1 = plane_at(g.e + 3, px, py), E2 = plane_at(g.e + 6, px, py);
  const float sum = E0 + E1 + E2;
  const float l0 = E0 / sum, l1 = E1 / sum, l2 = E2 / sum;
  const v3 ni = v3{a.n[0] * l0 + a.n[3] * l1 + a.n[6] * l2, a.n[1] * l0 + a.n[4] * l1 + a.n[7] * l2,
                   a.n[2] * l0 + a.n[5] * l1 + a.n[8] * l2};
  const float cx = a.cur[0] * l0 + a.cur[3] * l1 + a.cur[6] * l2, cy = a.cur[1] * l0 + a.cur[4] * l1 + a.cur[7] * l2;
  const float cw = a.cur[2] * l0 + a.cur[5] * l1 + a.cur[8] * l2;
  const float qx = a.prev[0] * l0 + a.prev[3] * l1 + a.prev[6] * l2;
  const float qy = a.prev[1] * l0 + a.prev[4] * l1 + a.prev[7] * l2;
  const float qw = a.prev[2] * l0 + a.prev[5] * l1 + a.prev[8] * l2;
  const irt_material& m = materials[a.material];
  float tu = 0.0f, tv = 0.0f;
  float4 tangent = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
  if (textured(m)) {
    tu = a.uv[0] * l0 + a.uv[2] * l1 + a.uv[4] * l2;
    tv = a.uv[1] * l0 + a.uv[3] * l1 + a.uv[5] * l2;
    tangent = make_float4(a.tangent[0] * l0 + a.tangent[4] * l1 + a.tangent[8] * l2,
                          a.tangent[1] * l0 + a.tangent[5] * l1 + a.tangent[9] * l2,
                          a.tangent[2] * l0 + a.tangent[6] * l1 + a.tangent[10] * l2,
                          a.tangent[3] * l0 + a.tangent[7] * l1 + a.tangent[11] * l2);
  }
  Surface s;
  material_at(m, tex, tu, tv, ni, tangent, s);
  depth[id] = __uint_as_float((uint32_t)(key >> 32));
  st_r16(normal, id, make_float4(s.n.x, s.n.y, s.n.z, 1.0f));
  st_r16(color, id, make_float4(s.color.x, s.color.y, s.color.z, 1.0f));
  st_rg16(params, id, s.rough, s.metal);
  st_rg16(motion, id, (cx / cw - qx / qw) * 0.5f, (cy / cw - qy / qw) * -0.5f);
}

}  // namespace irt
