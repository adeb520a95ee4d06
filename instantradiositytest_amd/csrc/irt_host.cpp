// Host-side precompute of the path (the CPU work renderer.rs does before it
// uploads the scene): settings validation, box certification, Lipschitz
// bound of a sampled field, the object BVH with escape links, SDF atlas
// placement, binary16 conversion for the reference texture formats, and the
// per-mesh triangle hierarchies the VPL pass walks.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "irt_host.h"

namespace irt_host {

thread_local std::string g_error;

void set_global_error(const std::string& s) { g_error = s; }
const char* global_error() { return g_error.c_str(); }

// Settings::validate — src/settings.rs:42-106 (same messages).
int validate_settings(const irt_settings& s, std::string* msg) {
  auto fail = [&](const char* m) {
    *msg = m;
    return IRT_ERR_INVALID_SETTINGS;
  };
  if (!(s.render_width > 0 && s.render_height > 0 && s.render_width <= 8192 && s.render_height <= 8192))
    return fail("invalid render size");
  if (!(s.vpl_count > 0 && s.vpl_count <= 65536 && s.vpl_max_bounces >= 1 && s.vpl_max_bounces <= 32))
    return fail("invalid VPL budget");
  if (!(s.vpl_rr_min_survival > 0.f && s.vpl_rr_max_survival < 1.f &&
        s.vpl_rr_min_survival <= s.vpl_rr_max_survival))
    return fail("invalid roulette probabilities");
  if (!(s.restir_direct_candidates > 0 && s.restir_vpl_candidates > 0 && s.restir_direct_candidates <= 256 &&
        s.restir_vpl_candidates <= 256 && s.restir_spatial_neighbors <= 32 && s.restir_spatial_radius <= 256))
    return fail("invalid ReSTIR settings");
  if (!(s.restir_max_history > 0 && s.restir_max_history <= 32 && s.svgf_iterations > 0 && s.svgf_iterations <= 5))
    return fail("invalid ReSTIR history/SVGF iteration count");
  if (!(s.temporal_history_weight >= 0.f && s.temporal_history_weight < 1.f && s.temporal_max_history > 0 &&
        s.temporal_max_history <= 256))
    return fail("invalid temporal settings");
  if (!(s.sdf_max_steps > 0 && s.sdf_max_steps <= 1024 && s.sdf_hit_epsilon > 0.f && std::isfinite(s.sdf_hit_epsilon) &&
        s.sdf_soft_shadow_k > 0.f && std::isfinite(s.sdf_soft_shadow_k)))
    return fail("invalid SDF settings");
  if (!(s.depth_threshold > 0.f && std::isfinite(s.depth_threshold) && s.normal_threshold >= 0.f &&
        s.normal_threshold <= 1.f && s.target_frame_ms > 0.f && std::isfinite(s.target_frame_ms)))
    return fail("invalid rejection/frame budget");
  return IRT_OK;
}

// is_box_geometry — src/renderer.rs:287-329.  Every vertex sits on a corner
// of [lo, hi]; the 12 triangles pair up on the 6 faces, each pair sharing a
// face diagonal.
bool is_box_geometry(const float* pos, uint32_t nv, uint32_t stride, const uint32_t* idx, uint32_t ni,
                     const float lo[3], const float hi[3]) {
  if (ni != 36) return false;
  float ext = std::min(hi[0] - lo[0], std::min(hi[1] - lo[1], hi[2] - lo[2]));
  if (!(ext > 0.f)) return false;
  std::vector<uint8_t> corner(nv);
  for (uint32_t v = 0; v < nv; ++v) {
    uint8_t c = 0;
    for (int a = 0; a < 3; ++a) {
      float p = pos[(size_t)v * stride + a];
      if (p == hi[a])
        c |= (uint8_t)(1u << a);
      else if (p != lo[a])
        return false;
    }
    corner[v] = c;
  }
  std::vector<std::array<uint8_t, 3>> faces[6];
  for (uint32_t t = 0; t < 12; ++t) {
    std::array<uint8_t, 3> c;
    for (int k = 0; k < 3; ++k) {
      uint32_t i = idx[t * 3 + k];
      if (i >= nv) return false;
      c[k] = corner[i];
    }
    std::sort(c.begin(), c.end());
    if (c[0] == c[1] || c[1] == c[2]) return false;
    int axis = -1;
    for (int a = 0; a < 3 && axis < 0; ++a) {
      uint8_t m = (uint8_t)(1u << a);
      if ((c[0] & m) == (c[1] & m) && (c[1] & m) == (c[2] & m)) axis = a;
    }
    if (axis < 0) return false;
    int side = (c[0] >> axis) & 1;
    faces[axis * 2 + side].push_back(c);
  }
  for (auto& f : faces) {
    if (f.size() != 2) return false;
    std::vector<uint8_t> shared;
    for (uint8_t a : f[0])
      if (std::find(f[1].begin(), f[1].end(), a) != f[1].end()) shared.push_back(a);
    if (shared.size() != 2) return false;
    if (__builtin_popcount((unsigned)(shared[0] ^ shared[1])) != 2) return false;
  }
  return true;
}

// field_lipschitz_bound — src/renderer.rs:330-358.  Maximum corner-gradient
// norm over cells whose largest sample exceeds the interpolation allowance,
// at least 1, times 1.0001.
float field_lipschitz_bound(const float* samples, uint32_t res, const float spacing[3]) {
  const size_t n = res;
  const float sl = std::sqrt(spacing[0] * spacing[0] + spacing[1] * spacing[1] + spacing[2] * spacing[2]);
  const float threshold = sl * (0.5f + 1.0f / 128.0f);
  float maximum_squared = 1.0f;
  for (size_t z = 0; z + 1 < n; ++z)
    for (size_t y = 0; y + 1 < n; ++y)
      for (size_t x = 0; x + 1 < n; ++x) {
        float v[8];
        bool any = false;
        for (int i = 0; i < 8; ++i) {
          v[i] = samples[(x + (i & 1)) + n * ((y + ((i >> 1) & 1)) + n * (z + (i >> 2)))];
          any |= !(v[i] <= threshold);
        }
        if (!any) continue;
        for (int c = 0; c < 8; ++c) {
          float dx = (v[(c & 6) | 1] - v[c & 6]) / spacing[0];
          float dy = (v[(c & 5) | 2] - v[c & 5]) / spacing[1];
          float dz = (v[(c & 3) | 4] - v[c & 3]) / spacing[2];
          maximum_squared = std::max(maximum_squared, dx * dx + dy * dy + dz * dz);
        }
      }
  return std::sqrt(maximum_squared) * 1.0001f;
}

// hierarchy — src/renderer.rs:385-455.  Median split on the longest axis of
// the node box, sorted by box-centre sum; preorder layout, info.w = index of
// the node that follows the subtree (escape link).
namespace {
struct Item {
  uint32_t id;
  float lo[3], hi[3];
};
uint32_t build(Item* b, size_t n, std::vector<irt_node>& nodes) {
  uint32_t index = (uint32_t)nodes.size();
  nodes.push_back(irt_node{});
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], b[i].lo[a]);
      hi[a] = std::max(hi[a], b[i].hi[a]);
    }
  uint32_t info[4];
  if (n == 1) {
    info[0] = b[0].id;
    info[1] = 0;
    info[2] = 1;
  } else {
    float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    int axis = (e[0] > e[1] && e[0] > e[2]) ? 0 : (e[1] > e[2] ? 1 : 2);
    std::stable_sort(b, b + n, [axis](const Item& p, const Item& q) {
      return (p.lo[axis] + p.hi[axis]) < (q.lo[axis] + q.hi[axis]);
    });
    size_t half = n / 2;
    info[0] = build(b, half, nodes);
    info[1] = build(b + half, n - half, nodes);
    info[2] = 0;
  }
  info[3] = (uint32_t)nodes.size();
  irt_node& node = nodes[index];
  for (int a = 0; a < 3; ++a) {
    node.lo[a] = lo[a];
    node.hi[a] = hi[a];
  }
  node.lo[3] = 0.f;
  node.hi[3] = 0.f;
  std::memcpy(node.info, info, sizeof(info));
  return index;
}
}  // namespace

int hierarchy(const float* box_lo, const float* box_hi, uint32_t count, irt_node* out) {
  if (count == 0) return 0;
  std::vector<Item> items(count);
  for (uint32_t i = 0; i < count; ++i) {
    items[i].id = i;
    for (int a = 0; a < 3; ++a) {
      items[i].lo[a] = box_lo[i * 3 + a];
      items[i].hi[a] = box_hi[i * 3 + a];
    }
  }
  std::vector<irt_node> nodes;
  nodes.reserve(2 * count - 1);
  build(items.data(), items.size(), nodes);
  std::copy(nodes.begin(), nodes.end(), out);
  return (int)nodes.size();
}

void transform_box(const float m[16], const float lo[3], const float hi[3], float out_lo[3], float out_hi[3]) {
  for (int a = 0; a < 3; ++a) {
    out_lo[a] = INFINITY;
    out_hi[a] = -INFINITY;
  }
  for (int z = 0; z < 2; ++z)
    for (int y = 0; y < 2; ++y)
      for (int x = 0; x < 2; ++x) {
        float p[3] = {x ? hi[0] : lo[0], y ? hi[1] : lo[1], z ? hi[2] : lo[2]};
        float w[3];
        // glam transform_point3: (m * (p, 1)).xyz without division.
        for (int r = 0; r < 3; ++r) w[r] = m[r] * p[0] + m[4 + r] * p[1] + m[8 + r] * p[2] + m[12 + r];
        for (int a = 0; a < 3; ++a) {
          out_lo[a] = std::min(out_lo[a], w[a]);
          out_hi[a] = std::max(out_hi[a], w[a]);
        }
      }
}

// Atlas grid — src/renderer.rs:778-791.
void atlas_grid(uint32_t count, uint32_t grid[3]) {
  uint32_t gx = 1;
  while (gx * gx * gx < count) ++gx;
  uint32_t gy = 1;
  while (gx * gy * gy < count) ++gy;
  uint32_t gz = (count + gx * gy - 1) / (gx * gy);
  grid[0] = gx;
  grid[1] = gy;
  grid[2] = gz;
}

// 4x4 column-major inverse by cofactors in double.
bool inverse4(const float* a, float* out) {
  double m[16], inv[16];
  for (int i = 0; i < 16; ++i) m[i] = a[i];
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
  for (int i = 0; i < 16; ++i) out[i] = (float)(inv[i] / det);
  return true;
}

// Frame packing — src/renderer.rs:1687-1735.
int pack_frame(const irt_settings& s, const float* vp, const float* pvp, const float* cam, uint32_t frame_index,
               uint32_t generation, uint32_t previous_generation, uint32_t history_enabled, uint32_t debug_mode,
               uint32_t lights, uint32_t objects, irt_frame* f) {
  std::memset(f, 0, sizeof(*f));
  std::memcpy(f->vp, vp, 64);
  std::memcpy(f->previous_vp, pvp, 64);
  if (!inverse4(vp, f->inv_vp) || !inverse4(pvp, f->previous_inv_vp)) return IRT_ERR_INVALID_ARGUMENT;
  f->camera[0] = cam[0];
  f->camera[1] = cam[1];
  f->camera[2] = cam[2];
  f->camera[3] = 0.f;
  f->size[0] = s.render_width;
  f->size[1] = s.render_height;
  f->size[2] = frame_index;
  f->size[3] = generation;
  f->counts[0] = lights;
  f->counts[1] = s.vpl_count;
  f->counts[2] = s.vpl_max_bounces;
  f->counts[3] = objects;
  f->restir[0] = s.restir_direct_candidates;
  f->restir[1] = s.restir_vpl_candidates;
  f->restir[2] = s.restir_spatial_neighbors;
  f->restir[3] = s.restir_spatial_radius;
  f->temporal[0] = s.temporal_history_weight;
  f->temporal[1] = s.depth_threshold;
  f->temporal[2] = s.normal_threshold;
  f->temporal[3] = (float)s.temporal_max_history;
  f->sdf[0] = (float)s.sdf_max_steps;
  f->sdf[1] = s.sdf_hit_epsilon;
  f->sdf[2] = s.sdf_soft_shadow_k;
  f->rr[0] = s.vpl_rr_min_survival;
  f->rr[1] = s.vpl_rr_max_survival;
  f->rr[2] = (float)s.restir_max_history;
  f->flags[0] = history_enabled ? 1u : 0u;
  f->flags[1] = previous_generation;
  f->flags[2] = debug_mode;
  return IRT_OK;
}

// IEEE binary16 <-> binary32 (round to nearest even).
uint16_t f32_to_f16(float f) {
  uint32_t x;
  std::memcpy(&x, &f, 4);
  uint32_t sign = (x >> 16) & 0x8000u;
  uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) return (uint16_t)(sign | 0x7c00u | (ax > 0x7f800000u ? 0x200u : 0u));
  if (ax >= 0x477ff000u) return (uint16_t)(sign | 0x7c00u);  // rounds to infinity
  if (ax < 0x38800000u) {                                      // subnormal or zero in binary16
    if (ax < 0x33000000u) return (uint16_t)sign;               // < half of smallest subnormal
    uint32_t mant = (ax & 0x7fffffu) | 0x800000u;
    int shift = 126 - (int)(ax >> 23);  // 14..24
    uint32_t q = mant >> (shift);
    uint32_t rem = mant & ((1u << shift) - 1u);
    uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return (uint16_t)(sign | q);
  }
  uint32_t e = (ax >> 23) - 112u;  // rebias 127 -> 15
  uint32_t mant = ax & 0x7fffffu;
  uint32_t q = (e << 10) | (mant >> 13);
  uint32_t rem = mant & 0x1fffu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) ++q;
  return (uint16_t)(sign | q);
}
// binary16 rounded toward -infinity: the largest binary16 value <= f.
uint16_t f32_to_f16_down(float f) {
  if (!(f == f)) return 0x7e00u;
  uint16_t h = f32_to_f16(f);
  if (f16_to_f32(h) > f) {
    if (h == 0x0000u) return 0x8001u;       // +0 -> -min subnormal
    if (h & 0x8000u) return (uint16_t)(h + 1u);  // negative: larger magnitude
    return (uint16_t)(h - 1u);              // positive: smaller magnitude
  }
  return h;
}
float f16_to_f32(uint16_t h) {
  uint32_t sign = (uint32_t)(h & 0x8000u) << 16;
  uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {
      int k = 0;
      while (!(m & 0x400u)) {
        m <<= 1;
        ++k;
      }
      m &= 0x3ffu;
      x = sign | ((uint32_t)(113 - k) << 23) | (m << 13);
    }
  } else if (e == 31) {
    x = sign | 0x7f800000u | (m << 13);
  } else {
    x = sign | ((e + 112u) << 23) | (m << 13);
  }
  float f;
  std::memcpy(&f, &x, 4);
  return f;
}

// sRGB decision thresholds of the Rgba8UnormSrgb display (IEC 61966-2-1):
// code k + 0.5 decoded to linear in double, rounded up to the next float, so
// that "v >= threshold" holds exactly when the exact encode of v reaches
// k + 0.5.
void srgb_thresholds(float out[255]) {
  for (int k = 0; k < 255; ++k) {
    const double c = (k + 0.5) / 255.0;
    const double lin = c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
    float f = (float)lin;
    if ((double)f < lin) f = std::nextafter(f, INFINITY);
    out[k] = f;
  }
}

// ------------------------------------------------ mesh triangle hierarchy
// The software counterpart of a triangle BLAS (renderer.rs:1061-1106): a
// binary hierarchy over a mesh's triangles in preorder with escape links,
// walked without a stack by the VPL kernel.
namespace {
struct TriRef {
  float lo[3], hi[3], c[3];
  uint32_t prim;
};
constexpr uint32_t TRI_LEAF = 4;  // triangles per leaf at most
constexpr int TRI_BINS = 16;

float half_area(const float lo[3], const float hi[3]) {
  const float e0 = hi[0] - lo[0], e1 = hi[1] - lo[1], e2 = hi[2] - lo[2];
  return e0 * e1 + e1 * e2 + e2 * e0;
}

// Binned SAH split on the widest centroid axis (median split when the
// centroids coincide, the best bin leaves one side empty, or the tree is
// deep); leaves of at most TRI_LEAF triangles.  Returns the node's index.
uint32_t build_tris(TriRef* t, size_t n, int depth, std::vector<irt_tri_node>& nodes, std::vector<uint32_t>& order) {
  const uint32_t index = (uint32_t)nodes.size();
  nodes.push_back(irt_tri_node{});
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  float clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], t[i].lo[a]);
      hi[a] = std::max(hi[a], t[i].hi[a]);
      clo[a] = std::min(clo[a], t[i].c[a]);
      chi[a] = std::max(chi[a], t[i].c[a]);
    }
  uint32_t link = 0, count = 0;
  if (n <= TRI_LEAF) {
    link = (uint32_t)order.size();
    count = (uint32_t)n;
    for (size_t i = 0; i < n; ++i) order.push_back(t[i].prim);
  } else {
    const float ce[3] = {chi[0] - clo[0], chi[1] - clo[1], chi[2] - clo[2]};
    const int axis = (ce[0] >= ce[1] && ce[0] >= ce[2]) ? 0 : (ce[1] >= ce[2] ? 1 : 2);
    size_t mid = 0;
    if (ce[axis] > 0.f && depth < 48) {
      auto bin_of = [&](const TriRef& r) {
        int b = (int)((r.c[axis] - clo[axis]) / ce[axis] * (float)TRI_BINS);
        return std::min(std::max(b, 0), TRI_BINS - 1);
      };
      size_t cnt[TRI_BINS] = {};
      float blo[TRI_BINS][3], bhi[TRI_BINS][3];
      for (int b = 0; b < TRI_BINS; ++b)
        for (int a = 0; a < 3; ++a) {
          blo[b][a] = INFINITY;
          bhi[b][a] = -INFINITY;
        }
      for (size_t i = 0; i < n; ++i) {
        const int b = bin_of(t[i]);
        ++cnt[b];
        for (int a = 0; a < 3; ++a) {
          blo[b][a] = std::min(blo[b][a], t[i].lo[a]);
          bhi[b][a] = std::max(bhi[b][a], t[i].hi[a]);
        }
      }
      // cost of splitting after bin s: area(left) * n_left + area(right) * n_right
      float best = INFINITY;
      int best_s = -1;
      for (int s = 0; s < TRI_BINS - 1; ++s) {
        float llo[3] = {INFINITY, INFINITY, INFINITY}, lhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        float rlo[3] = {INFINITY, INFINITY, INFINITY}, rhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        size_t nl = 0, nr = 0;
        for (int b = 0; b < TRI_BINS; ++b) {
          if (!cnt[b]) continue;
          float* L = b <= s ? llo : rlo;
          float* Hh = b <= s ? lhi : rhi;
          for (int a = 0; a < 3; ++a) {
            L[a] = std::min(L[a], blo[b][a]);
            Hh[a] = std::max(Hh[a], bhi[b][a]);
          }
          (b <= s ? nl : nr) += cnt[b];
        }
        if (!nl || !nr) continue;
        const float cost = half_area(llo, lhi) * (float)nl + half_area(rlo, rhi) * (float)nr;
        if (cost < best) {
          best = cost;
          best_s = s;
        }
      }
      if (best_s >= 0)
        mid = (size_t)(std::partition(t, t + n, [&](const TriRef& r) { return bin_of(r) <= best_s; }) - t);
    }
    if (mid == 0 || mid == n) {
      mid = n / 2;
      std::nth_element(t, t + mid, t + n, [axis](const TriRef& p, const TriRef& q) {
        return p.c[axis] < q.c[axis] || (p.c[axis] == q.c[axis] && p.prim < q.prim);
      });
    }
    build_tris(t, mid, depth + 1, nodes, order);
    build_tris(t + mid, n - mid, depth + 1, nodes, order);
    link = (uint32_t)nodes.size();  // escape: the node after this subtree
  }
  irt_tri_node& nd = nodes[index];
  float mag = 0.f;
  for (int a = 0; a < 3; ++a) mag = std::max(mag, std::max(std::fabs(lo[a]), std::fabs(hi[a])));
  const float pad = mag * 4e-5f;
  for (int a = 0; a < 3; ++a) {
    nd.lo[a] = lo[a] - pad;
    nd.hi[a] = hi[a] + pad;
  }
  nd.link = link;
  nd.count = count;
  return index;
}
}  // namespace

int mesh_hierarchy(const irt_vertex* v, uint32_t nv, const uint32_t* idx, uint32_t ni,
                   std::vector<irt_tri_node>& nodes, std::vector<irt_tri>& tris, std::string* msg) {
  nodes.clear();
  tris.clear();
  if (ni % 3u) {
    *msg = "mesh index count is not a multiple of 3";
    return IRT_ERR_INVALID_ARGUMENT;
  }
  const uint32_t nt = ni / 3u;
  std::vector<TriRef> refs(nt);
  for (uint32_t p = 0; p < nt; ++p) {
    TriRef& r = refs[p];
    r.prim = p;
    for (int a = 0; a < 3; ++a) {
      r.lo[a] = INFINITY;
      r.hi[a] = -INFINITY;
    }
    for (int k = 0; k < 3; ++k) {
      const uint32_t i = idx[3 * p + k];
      if (i >= nv) {
        *msg = "mesh index refers past the mesh's vertices";
        return IRT_ERR_INVALID_ARGUMENT;
      }
      for (int a = 0; a < 3; ++a) {
        r.lo[a] = std::min(r.lo[a], v[i].position[a]);
        r.hi[a] = std::max(r.hi[a], v[i].position[a]);
      }
    }
    for (int a = 0; a < 3; ++a) {
      if (!std::isfinite(r.lo[a]) || !std::isfinite(r.hi[a])) {
        *msg = "vertex position is not finite";
        return IRT_ERR_INVALID_ARGUMENT;
      }
      r.c[a] = 0.5f * (r.lo[a] + r.hi[a]);
    }
  }
  if (nt == 0) return IRT_OK;
  std::vector<uint32_t> order;
  order.reserve(nt);
  nodes.reserve(2 * (size_t)nt);
  build_tris(refs.data(), refs.size(), 0, nodes, order);
  tris.resize(nt);
  for (uint32_t k = 0; k < nt; ++k) {
    const uint32_t p = order[k];
    irt_tri& t = tris[k];
    std::memset(&t, 0, sizeof(t));
    t.prim = p;
    for (int a = 0; a < 3; ++a) {
      t.a[a] = v[idx[3 * p]].position[a];
      t.b[a] = v[idx[3 * p + 1]].position[a];
      t.c[a] = v[idx[3 * p + 2]].position[a];
    }
  }
  return IRT_OK;
}

}  // namespace irt_host

// ------------------------------------------------------------- C ABI part
extern "C" {

int irt_settings_validate(const irt_settings* s) {
  if (!s) {
    irt_host::set_global_error("null settings");
    return IRT_ERR_INVALID_ARGUMENT;
  }
  std::string msg;
  int rc = irt_host::validate_settings(*s, &msg);
  if (rc) irt_host::set_global_error(msg);
  return rc;
}

int irt_host_is_box_geometry(const float* positions, uint32_t vertex_count, uint32_t stride_floats,
                             const uint32_t* indices, uint32_t index_count, const float lo[3], const float hi[3]) {
  if (!positions || !indices || !lo || !hi || stride_floats < 3) return 0;
  return irt_host::is_box_geometry(positions, vertex_count, stride_floats, indices, index_count, lo, hi) ? 1 : 0;
}

float irt_host_field_lipschitz_bound(const float* samples, uint32_t resolution, const float spacing[3]) {
  if (!samples || !spacing || resolution < 2) return 1.0001f;
  return irt_host::field_lipschitz_bound(samples, resolution, spacing);
}

int irt_host_hierarchy(const float* box_lo, const float* box_hi, uint32_t count, irt_node* nodes) {
  if (!box_lo || !box_hi || !nodes) return IRT_ERR_INVALID_ARGUMENT;
  return irt_host::hierarchy(box_lo, box_hi, count, nodes);
}

void irt_host_transform_box(const float model[16], const float lo[3], const float hi[3], float out_lo[3],
                            float out_hi[3]) {
  irt_host::transform_box(model, lo, hi, out_lo, out_hi);
}

void irt_host_atlas_grid(uint32_t mesh_count, uint32_t tile, uint32_t grid[3]) {
  (void)tile;
  irt_host::atlas_grid(mesh_count, grid);
}

int irt_host_pack_frame(const irt_settings* s, const float vp[16], const float previous_vp[16], const float camera[3],
                        uint32_t frame_index, uint32_t generation, uint32_t previous_generation,
                        uint32_t history_enabled, uint32_t debug_mode, uint32_t light_count, uint32_t object_count,
                        irt_frame* out) {
  if (!s || !vp || !previous_vp || !camera || !out) {
    irt_host::set_global_error("null argument");
    return IRT_ERR_INVALID_ARGUMENT;
  }
  int rc = irt_host::pack_frame(*s, vp, previous_vp, camera, frame_index, generation, previous_generation,
                                history_enabled, debug_mode, light_count, object_count, out);
  if (rc) irt_host::set_global_error("singular view-projection matrix");
  return rc;
}

int irt_host_inverse(const float m[16], float out[16]) {
  if (!m || !out) return 0;
  return irt_host::inverse4(m, out) ? 1 : 0;
}

void irt_host_srgb_thresholds(float out[255]) {
  if (out) irt_host::srgb_thresholds(out);
}

int irt_host_mesh_hierarchy(const irt_vertex* vertices, uint32_t vertex_count, const uint32_t* indices,
                            uint32_t index_count, irt_tri_node* nodes, irt_tri* tris) {
  if ((index_count && (!vertices || !indices || !nodes || !tris))) {
    irt_host::set_global_error("null argument");
    return IRT_ERR_INVALID_ARGUMENT;
  }
  std::vector<irt_tri_node> n;
  std::vector<irt_tri> t;
  std::string msg;
  int rc = irt_host::mesh_hierarchy(vertices, vertex_count, indices, index_count, n, t, &msg);
  if (rc) {
    irt_host::set_global_error(msg);
    return rc;
  }
  std::copy(n.begin(), n.end(), nodes);
  std::copy(t.begin(), t.end(), tris);
  return (int)n.size();
}

}  // extern "C"
