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

#include "irt_march.h"

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

// Tiles in column bands, one band per XCD.  The dispatcher deals the
// workgroups of a launch to the 8 XCDs in turn (workgroup L runs on XCD
// L % 8), so workgroup L takes a tile of band L % 8: XCD k walks the tile
// columns [k*bw, (k+1)*bw) row by row.  The neighbours a pixel gathers (up
// to 12 px away) then sit mostly in tiles of the same XCD, whose L2 already
// holds their lines.  The launch is one-dimensional: XCDS * bw * tiles_y
// workgroups of TILE x TILE threads.
constexpr uint32_t XCDS = 8;
__host__ __device__ __forceinline__ uint32_t band_width_tiles(uint32_t tiles_x) { return (tiles_x + XCDS - 1) / XCDS; }
__host__ __device__ __forceinline__ uint32_t band_width(uint32_t W) { return band_width_tiles((W + TILE - 1) / TILE); }
__device__ __forceinline__ Px pixel_banded(uint32_t W, uint32_t H) {
  const uint32_t bw = band_width(W);
  const uint32_t j = blockIdx.x / XCDS;
  const uint32_t tx = (blockIdx.x % XCDS) * bw + j % bw, ty = j / bw;
  Px p;
  p.x = (int32_t)(tx * TILE + threadIdx.x);
  p.y = (int32_t)(ty * TILE + threadIdx.y);
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
template <class P, class C>
__device__ __forceinline__ void body_reservoirs_a(const LightArgs<P>& a, const Px& p, irt_reservoir* direct0,
                                                  irt_reservoir* indirect0, C& cnt) {
  Surface s;
  bool valid = valid_surface(a, p, &s);
  if (!p.active) return;
  irt_reservoir rd = candidates_direct(a, p.id, valid, s, cnt);
  irt_reservoir ri = candidates_indirect(a, p.id, valid, s, cnt);
  if (direct0) direct0[p.id] = rd;
  if (indirect0) indirect0[p.id] = ri;
  uint32_t q;
  bool hv = temporal_gate(a, p, valid, s, &q);
  a.direct_out[p.id] = temporal_direct(a, rd, p.id, hv, q, s, cnt);
  a.indirect_out[p.id] = temporal_indirect(a, ri, p.id, hv, q, s, cnt);
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_reservoirs_a(LightArgs<P> a, irt_reservoir* direct0,
                                                     irt_reservoir* indirect0) {
  Px p = pixel(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  body_reservoirs_a(a, p, direct0, indirect0, cnt);
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

// ------------------------------------------------- decoupled resolve (fused)
// Reservoirs B split in three so that marches are scheduled independently of
// pixels: (1) spatial reuse + unshadowed contributions + ray set-up, rays
// appended to a queue; (2) a persistent march over the queue in which lanes
// whose ray finished take the next one; (3) the combine of contribution x
// visibility.  Arithmetic and its order are those of resolve_direct /
// resolve_indirect (value * normalization * visibility, then the sum).
struct ResolveRec {
  float cd[3];
  uint32_t flags;  // bit 0: direct contribution, bit 1: indirect contribution
  float ci[3];
  uint32_t pad;
};
static_assert(sizeof(ResolveRec) == 32, "resolve record is 32 bytes");

__device__ __forceinline__ RayReq make_req(v3 p, v3 n, v3 t, v3 e, uint32_t slot) {
  return RayReq{p.x, p.y, p.z, n.x, n.y, n.z, t.x, t.y, t.z, e.x, e.y, e.z, slot, 1u, 0u, 0u};
}

template <class P, class C>
__device__ __forceinline__ void body_spatial_setup(const LightArgs<P>& a, const Px& p, ResolveRec* rec, RayReq* reqs,
                                                   C& cnt) {
  Surface s;
  const irt_frame& F = a.F;
  bool valid = valid_surface(a, p, &s);
  bool want_d = false, want_i = false;
  RayReq ray_d{}, ray_i{};
  if (p.active) {
    irt_reservoir rd = a.direct_in[p.id], ri = a.indirect_in[p.id];
    spatial_both(a, rd, ri, p.x, p.y, valid, s, cnt);
    a.direct_out[p.id] = rd;
    a.indirect_out[p.id] = ri;
    ResolveRec r{};
    if (valid) {
      const uint32_t mode = F.flags[2];
      if (rd.selected_target > 0.0f && rd.id < F.counts[0] && mode != 2u && mode < 3u) {
        const irt_light& l = a.sc.lights[rd.id];
        v3 value = direct(s, l, cnt);
        if (luminance(value) > 0.0f) {
          v3 c = value * normalization(rd);
          r.cd[0] = c.x;
          r.cd[1] = c.y;
          r.cd[2] = c.z;
          r.flags |= 1u;
          want_d = true;
          ray_d = make_req(s.p, s.n, ld3(l.position_strength), v3{0.f, 0.f, 0.f}, 2 * p.id);
        }
      }
      if (mode != 1u && mode < 3u && vpl_sample_valid(F, ri)) {
        const irt_vpl& smp = selected_vpl(F, a.vpl_ring, ri);
        v3 value = indirect(s, smp, cnt);
        if (luminance(value) > 0.0f) {
          v3 c = value * normalization(ri);
          r.ci[0] = c.x;
          r.ci[1] = c.y;
          r.ci[2] = c.z;
          r.flags |= 2u;
          want_i = true;
          ray_i = make_req(s.p, s.n, ld3(smp.position), ld3(smp.normal), 2 * p.id + 1);
        }
      }
    }
    rec[p.id] = r;
    if (want_d) {
      reqs[2 * p.id] = ray_d;
    } else {
      reqs[2 * p.id].valid = 0u;
    }
    if (want_i) {
      reqs[2 * p.id + 1] = ray_i;
    } else {
      reqs[2 * p.id + 1].valid = 0u;
    }
  }
}
template <class P, bool CNT>
__global__ __launch_bounds__(64) void k_spatial_setup(LightArgs<P> a, ResolveRec* rec, RayReq* reqs) {
  Px p = pixel_banded(a.F.size[0], a.F.size[1]);
  Count<CNT> cnt;
  body_spatial_setup(a, p, rec, reqs, cnt);
  cnt.flush(a.counters);
}

template <class P>
__device__ __forceinline__ void body_resolve_combine(const LightArgs<P>& a, const Px& p, const ResolveRec* rec,
                                                     const float* vis) {
  if (!p.active) return;
  Surface s;
  bool valid = valid_surface(a, p, &s);
  const irt_frame& F = a.F;
  v3 dl = v3{0.f, 0.f, 0.f};
  ResolveRec r{};
  if (valid) {
    r = rec[p.id];
    if (r.flags & 1u) dl = dl + v3{r.cd[0], r.cd[1], r.cd[2]} * vis[2 * p.id];
    const uint32_t mode = F.flags[2];
    if (mode == 3u) dl = s.color;
    if (mode == 4u) dl = s.n * 0.5f + v3{0.5f, 0.5f, 0.5f};
    if (mode == 5u) {
      float2 m = ld2(a.motion, p.id);
      dl = v3{clampf(0.5f + m.x * 50.0f, 0.0f, 1.0f), clampf(0.5f + m.y * 50.0f, 0.0f, 1.0f), 0.5f};
    }
  }
  if (a.direct_lighting) st4(a.direct_lighting, p.id, make_float4(dl.x, dl.y, dl.z, 1.0f));
  v3 result = dl;
  if (valid && (r.flags & 2u)) result = result + v3{r.ci[0], r.ci[1], r.ci[2]} * vis[2 * p.id + 1];
  result = vmin(result, v3{65000.f, 65000.f, 65000.f});
  st_r16(a.lighting, p.id, make_float4(result.x, result.y, result.z, 1.0f));
}

template <class P>
__global__ __launch_bounds__(64) void k_resolve_combine(LightArgs<P> a, const ResolveRec* rec, const float* vis) {
  body_resolve_combine(a, pixel(a.F.size[0], a.F.size[1]), rec, vis);
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint32_t v, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Up to five device-to-device plane copies in one launch (the mixed
// profile's G-buffer ingest: the reference formats are its storage
// formats).  Each plane is copied in 16-byte units by a grid-stride loop;
// the host passes 16-byte aligned pointers and sizes in whole units, with
// any dword tail in `tail` (at most 3 dwords per plane, copied by block 0).
struct PlaneCopy {
  const uint4* src[5];
  uint4* dst[5];
  uint32_t units[5];
  uint32_t tail[5];
  uint32_t planes;
};
__global__ __launch_bounds__(256) void k_copy_planes(PlaneCopy pc) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = 0; k < pc.planes; ++k) {
    const uint4* __restrict__ s = pc.src[k];
    uint4* __restrict__ d = pc.dst[k];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < pc.units[k]; i += stride) d[i] = s[i];
    if (blockIdx.x == 0 && threadIdx.x < pc.tail[k])
      reinterpret_cast<uint32_t*>(d + pc.units[k])[threadIdx.x] =
          reinterpret_cast<const uint32_t*>(s + pc.units[k])[threadIdx.x];
  }
}

// ------------------------------------------------------------------- SVGF
template <class P, class C, bool FAST>
__device__ __forceinline__ void body_svgf_temporal(const SvgfArgs<P>& a, const Px& p, C& cnt) {
  if (!p.active) return;
  float4 rad, mom;
  svgf_temporal<P, C, FAST>(a, p.x, p.y, &rad, &mom, cnt);
  st_r16(a.out, p.id, rad);
  st_mom(a.moments_out, p.id, mom);
}
template <class P, class C, bool FAST>
__device__ __forceinline__ void body_svgf_variance(const SvgfArgs<P>& a, const Px& p, C& cnt) {
  if (p.active) st_r16(a.out, p.id, svgf_variance<P, C, FAST>(a, p.x, p.y, cnt));
}
template <class P, class C, bool FAST>
__device__ __forceinline__ void body_svgf_atrous(const SvgfArgs<P>& a, const Px& p, C& cnt) {
  if (!p.active) return;
  float4 r = svgf_atrous<P, C, FAST>(a, p.x, p.y, cnt);
  st_r16(a.out, p.id, r);
  if (a.out2) st_r16(a.out2, p.id, r);
}
template <class P, bool CNT, bool FAST = false>
__global__ __launch_bounds__(64) void k_svgf_temporal(SvgfArgs<P> a) {
  Count<CNT> cnt;
  body_svgf_temporal<P, Count<CNT>, FAST>(a, pixel(a.F.size[0], a.F.size[1]), cnt);
  cnt.flush(a.counters);
}
template <class P, bool CNT, bool FAST = false>
__global__ __launch_bounds__(64) void k_svgf_variance(SvgfArgs<P> a) {
  Count<CNT> cnt;
  body_svgf_variance<P, Count<CNT>, FAST>(a, pixel(a.F.size[0], a.F.size[1]), cnt);
  cnt.flush(a.counters);
}
template <class P, bool CNT, bool FAST = false>
__global__ __launch_bounds__(64) void k_svgf_atrous(SvgfArgs<P> a) {
  Count<CNT> cnt;
  body_svgf_atrous<P, Count<CNT>, FAST>(a, pixel(a.F.size[0], a.F.size[1]), cnt);
  cnt.flush(a.counters);
}

// Variance estimate with the 7 x 7 neighbourhood staged in LDS (guide
// records present).  A block renders a 16 x 16 tile; when any of its pixels
// is young enough for the 7 x 7 estimate (age < 4), it stages the guide
// records and first two moments of the 22 x 22 window around the tile once
// and every tap reads LDS; otherwise each pixel passes its variance through.
// Temporal accumulation with the 3 x 3 clamp window staged in LDS (guide
// records present): the 16 x 16 tile at (x0, y0), one pixel per thread of a
// 256-thread block (every thread of the block calls it); the guide records
// and noisy radiance of the 18 x 18 window around the tile are staged once.
constexpr int32_t TMP_TILE = 16, TMP_HALO = 1, TMP_SIDE = TMP_TILE + 2 * TMP_HALO;
template <class P, class C>
__device__ __forceinline__ void temporal_tile(const SvgfArgs<P>& a, int32_t x0, int32_t y0, uint4 (*s_g)[2],
                                              typename P::R4* s_r, C& cnt) {
  const int32_t W = (int32_t)a.F.size[0], H = (int32_t)a.F.size[1];
  for (int32_t e = (int32_t)threadIdx.x; e < TMP_SIDE * TMP_SIDE; e += 256) {
    const int32_t gx = x0 + e % TMP_SIDE - TMP_HALO, gy = y0 + e / TMP_SIDE - TMP_HALO;
    if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
      const uint32_t q = (uint32_t)gy * (uint32_t)W + (uint32_t)gx;
      const uint4* gp = reinterpret_cast<const uint4*>(a.guide + q);
      s_g[e][0] = gp[0];
      s_g[e][1] = gp[1];
      s_r[e] = a.noisy[q];
    } else {
      s_g[e][0] = make_uint4(0u, 0u, 0u, 0u);  // valid = 0: outside the image
      s_g[e][1] = make_uint4(0u, 0u, 0u, 0u);
      s_r[e] = typename P::R4{};
    }
  }
  __syncthreads();
  const int32_t i = (int32_t)threadIdx.x % TMP_TILE, j = (int32_t)threadIdx.x / TMP_TILE;
  const int32_t x = x0 + i, y = y0 + j;
  if (x < W && y < H) {
    const uint32_t id = (uint32_t)y * (uint32_t)W + (uint32_t)x;
    float4 rad, mom;
    svgf_temporal_staged<P>(a, x, y, s_g, s_r, i + TMP_HALO, j + TMP_HALO, TMP_SIDE, &rad, &mom, cnt);
    st_r16(a.out, id, rad);
    st_mom(a.moments_out, id, mom);
  }
}
template <class P, bool CNT>
__global__ __launch_bounds__(256) void k_svgf_temporal_lds(SvgfArgs<P> a) {
  __shared__ uint4 s_g[TMP_SIDE * TMP_SIDE][2];
  __shared__ typename P::R4 s_r[TMP_SIDE * TMP_SIDE];
  Count<CNT> cnt;
  temporal_tile(a, (int32_t)blockIdx.x * TMP_TILE, (int32_t)blockIdx.y * TMP_TILE, s_g, s_r, cnt);
  cnt.flush(a.counters);
}

constexpr int32_t VAR_TILE = 16, VAR_HALO = 3, VAR_SIDE = VAR_TILE + 2 * VAR_HALO;
// The 16 x 16 tile at (x0, y0) of the variance pass, one pixel per thread of
// a 256-thread block (every thread of the block calls it).
template <class P, class C>
__device__ __forceinline__ void variance_tile(const SvgfArgs<P>& a, int32_t x0, int32_t y0, uint4 (*s_g)[2],
                                              float2* s_m, C& cnt) {
  const int32_t W = (int32_t)a.F.size[0], H = (int32_t)a.F.size[1];
  const int32_t i = (int32_t)threadIdx.x % VAR_TILE, j = (int32_t)threadIdx.x / VAR_TILE;
  const int32_t x = x0 + i, y = y0 + j;
  const bool in = x < W && y < H;
  const uint32_t id = in ? (uint32_t)y * (uint32_t)W + (uint32_t)x : 0u;
  const bool young = in && ld_mom(a.moments, id).z < 4.0f;
  if (__syncthreads_or(young)) {
    for (int32_t e = (int32_t)threadIdx.x; e < VAR_SIDE * VAR_SIDE; e += 256) {
      const int32_t gx = x0 + e % VAR_SIDE - VAR_HALO, gy = y0 + e / VAR_SIDE - VAR_HALO;
      if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
        const uint32_t q = (uint32_t)gy * (uint32_t)W + (uint32_t)gx;
        const uint4* gp = reinterpret_cast<const uint4*>(a.guide + q);
        s_g[e][0] = gp[0];
        s_g[e][1] = gp[1];
        const float4 mo = ld_mom(a.moments, q);
        s_m[e] = make_float2(mo.x, mo.y);
      } else {
        s_g[e][0] = make_uint4(0u, 0u, 0u, 0u);  // valid = 0: outside the image
        s_g[e][1] = make_uint4(0u, 0u, 0u, 0u);
        s_m[e] = make_float2(0.f, 0.f);
      }
    }
    __syncthreads();
    if (in)
      st_r16(a.out, id, svgf_variance_staged<P, C, true>(a, x, y, s_g, s_m, i + VAR_HALO, j + VAR_HALO, VAR_SIDE, cnt));
  } else if (in) {
    const float4 c = ld4(a.input, id);
    st_r16(a.out, id, make_float4(c.x, c.y, c.z, fminf(c.w, 65000.0f)));
  }
}
template <class P, bool CNT>
__global__ __launch_bounds__(256) void k_svgf_variance_lds(SvgfArgs<P> a) {
  __shared__ uint4 s_g[VAR_SIDE * VAR_SIDE][2];
  __shared__ float2 s_m[VAR_SIDE * VAR_SIDE];
  Count<CNT> cnt;
  variance_tile(a, (int32_t)blockIdx.x * VAR_TILE, (int32_t)blockIdx.y * VAR_TILE, s_g, s_m, cnt);
  cnt.flush(a.counters);
}

// A-trous iteration with the tap neighbourhood staged in LDS (guide records
// present).  A block renders a 16 x 16 patch of one coset (cx, cy) of the
// stride-s lattice in patch region (bx, by), x = bx*16s + cx + i*s,
// y = by*16s + cy + j*s (atrous_slot places the block):
// all 5 x 5 taps of those pixels fall on the same lattice, so the block
// stages the 20 x 20 lattice points around the patch (guide record + input
// radiance) once and every tap reads LDS.  Arithmetic as svgf_atrous.
constexpr int32_t AT_TILE = 16, AT_HALO = 2, AT_SIDE = AT_TILE + 2 * AT_HALO;

// Launch slot L of an a-trous launch over (patch region, lattice coset)
// pairs.  The dispatcher deals slots to the eight XCDs round-robin, so slot
// L = 8 (r_hi C + c) + r_lo, with region r = 8 r_hi + r_lo and coset c < C,
// puts all cosets of a region on one XCD one after another: they read the
// same cache lines of the guide records and radiance.  Launches have
// 8 ceil(regions / 8) C slots; false for the padding slots.
__device__ __forceinline__ bool atrous_slot(uint32_t L, uint32_t regions, uint32_t cosets, uint32_t* region,
                                            uint32_t* coset) {
  const uint32_t r_lo = L % 8u, rest = L / 8u;
  *coset = rest % cosets;
  *region = (rest / cosets) * 8u + r_lo;
  return *region < regions;
}

template <class P, bool CNT>
__global__ __launch_bounds__(256) void k_svgf_atrous_lds(SvgfArgs<P> a) {
  __shared__ uint4 s_g[AT_SIDE * AT_SIDE][2];
  __shared__ typename P::R4 s_r[AT_SIDE * AT_SIDE];
  const int32_t W = (int32_t)a.F.size[0], H = (int32_t)a.F.size[1], st = (int32_t)a.stride;
  const uint32_t span = (uint32_t)(AT_TILE * st), nx = ((uint32_t)W + span - 1) / span, ny = ((uint32_t)H + span - 1) / span;
  uint32_t region, coset;
  if (!atrous_slot(blockIdx.x, nx * ny, (uint32_t)(st * st), &region, &coset)) return;
  const int32_t cx = (int32_t)coset % st, cy = (int32_t)coset / st;
  const int32_t x0 = (int32_t)(region % nx) * AT_TILE * st + cx, y0 = (int32_t)(region / nx) * AT_TILE * st + cy;
  for (int32_t e = (int32_t)threadIdx.x; e < AT_SIDE * AT_SIDE; e += 256) {
    const int32_t gx = x0 + (e % AT_SIDE - AT_HALO) * st, gy = y0 + (e / AT_SIDE - AT_HALO) * st;
    if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
      const uint32_t q = (uint32_t)gy * (uint32_t)W + (uint32_t)gx;
      const uint4* gp = reinterpret_cast<const uint4*>(a.guide + q);
      s_g[e][0] = gp[0];
      s_g[e][1] = gp[1];
      s_r[e] = a.input[q];
    } else {
      s_g[e][0] = make_uint4(0u, 0u, 0u, 0u);  // valid = 0: outside the image
      s_g[e][1] = make_uint4(0u, 0u, 0u, 0u);
      s_r[e] = typename P::R4{};
    }
  }
  __syncthreads();
  const int32_t i = (int32_t)threadIdx.x % AT_TILE, j = (int32_t)threadIdx.x / AT_TILE;
  const int32_t x = x0 + i * st, y = y0 + j * st;
  Count<CNT> cnt;
  if (x < W && y < H) {
    const uint32_t id = (uint32_t)y * (uint32_t)W + (uint32_t)x;
    const float4 r = svgf_atrous_staged<P>(a, x, y, s_g, s_r, i + AT_HALO, j + AT_HALO, AT_SIDE, st == 1, cnt);
    st_r16(a.out, id, r);
    if (a.out2) st_r16(a.out2, id, r);
  }
  cnt.flush(a.counters);
}

// Guide records of one frame: world position (world(), common.wgsl:46) and
// the binary16 channels of the G-buffer.
template <class P>
__device__ __forceinline__ void body_guide(const irt_frame& F, const float* depth, const typename P::N4* normal,
                                           const typename P::N4* color, const typename P::P2* params, GuidePx* out,
                                           int32_t x, int32_t y, uint32_t i) {
  const uint32_t W = F.size[0], H = F.size[1];
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
template <class P>
__global__ __launch_bounds__(256) void k_guide(irt_frame F, const float* depth, const typename P::N4* normal,
                                               const typename P::N4* color, const typename P::P2* params,
                                               GuidePx* out) {
  const uint32_t W = F.size[0], H = F.size[1];
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= W * H) return;
  body_guide<P>(F, depth, normal, color, params, out, (int32_t)(i % W), (int32_t)(i / W), i);
}

// --------------------------------------------------------- candidate grid
// Largest corner-gradient norm of a mesh's trilinear field over all of its
// cells (the trilinear gradient is a convex combination of the eight corner
// gradients).  One y-z slab of cells per thread; result in k2[m] as the bits
// of the squared norm (non-negative floats order like unsigned integers).
__global__ __launch_bounds__(256) void k_mesh_gradient(SceneView sc, uint32_t n_meshes, unsigned int* k2) {
  const uint32_t m = blockIdx.y;
  if (m >= n_meshes) return;
  const irt_mesh mesh = sc.meshes[m];
  const uint32_t res = mesh.field[3];
  const uint32_t ncell = res - 1u;
  float best = 1.0f;
  if (mesh.geometry_lo[3] <= 0.5f) {
    const size_t sy = sc.atlas_w, sz = (size_t)sc.atlas_w * sc.atlas_h;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell * ncell * ncell; c += gridDim.x * blockDim.x) {
      const uint32_t x = c % ncell, y = (c / ncell) % ncell, z = c / (ncell * ncell);
      const size_t off = ((size_t)(mesh.field[2] + z) * sc.atlas_h + (mesh.field[1] + y)) * sc.atlas_w +
                         (mesh.field[0] + x);
      float v[8];
      for (int i = 0; i < 8; ++i) {
        const size_t o = off + (i & 1) + ((i >> 1) & 1) * sy + (i >> 2) * sz;
        v[i] = sc.atlas16 ? (float)sc.atlas16[o] : sc.atlas[o];
      }
      for (int k = 0; k < 8; ++k) {
        float dx = (v[(k & 6) | 1] - v[k & 6]) * mesh.grid[0];
        float dy = (v[(k & 5) | 2] - v[k & 5]) * mesh.grid[1];
        float dz = (v[(k & 3) | 4] - v[k & 3]) * mesh.grid[2];
        best = fmaxf(best, dx * dx + dy * dy + dz * dz);
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) best = fmaxf(best, __shfl_xor(best, off, 64));
  if (threadIdx.x % 64 == 0) atomicMax(k2 + m, __float_as_uint(best));
}

// Candidate lists, one thread per cell.  With K_j a Lipschitz constant of
// object_distance (1 for certified boxes, sqrt(K_mesh^2 + 1) for grid
// meshes: the padded-box extension combines the field gradient with a unit
// outside term) and d_j the object distance at the cell centre:
//   lower bound of object j over the cell: the larger of 0.9998 x the gap
//     between the cell and j's world box (the hierarchy's own pruning bound;
//     -inf when they overlap) and d_j - K_j x the half diagonal;
//   upper bound: d_j + K_j x the half diagonal;
// both with a relative FP margin.  The second lower bound drops objects whose
// world box overlaps the cell but whose surface is far from it.
// K_mesh, the largest corner-gradient norm of the field, bounds the gradient
// everywhere in a cell: each gradient component is multilinear, so the
// squared norm is convex along each axis and peaks at a corner.
constexpr uint32_t GRID_MAX_OBJECTS = 64;
constexpr float GRID_CELLS_LONG_AXIS = 160.0f;  // cells on the root box's longest axis
__global__ __launch_bounds__(64) void k_build_cells(SceneView sc, uint32_t n_objects, float hit_eps,
                                                    const unsigned int* k2_mesh, CellList* cells) {
  const GridView& g = sc.grid;
  const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t total = g.nx * g.ny * g.nz;
  if (cell >= total) return;
  const uint32_t cx = cell % g.nx, cy = (cell / g.nx) % g.ny, cz = cell / (g.nx * g.ny);
  const float cs = 1.0f / g.inv_cell;
  const v3 lo = v3{g.x0 + cx * cs, g.y0 + cy * cs, g.z0 + cz * cs};
  const v3 hi = v3{lo.x + cs, lo.y + cs, lo.z + cs};
  const v3 center = v3{(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
  const float half_diag = 0.5f * cs * 1.7320508f * 1.0001f;
  // Box-gap lower bounds of every object, then the objects in increasing
  // gap order until the next gap exceeds the smallest upper bound so far.
  // An object's distance at the centre is at least its gap (the world box
  // holds the object; positive distances carry the 0.9999 factor, the gap
  // 0.9998), so an object left out has an upper bound above that minimum and
  // a lower bound above it too: it neither lowers the minimum nor becomes a
  // candidate, and the lists equal those of evaluating every object.
  float lower[GRID_MAX_OBJECTS];
  for (uint32_t j = 0; j < n_objects; ++j) {
    const float4 blo = g.boxes[2 * j], bhi = g.boxes[2 * j + 1];
    const float gx = fmaxf(fmaxf(blo.x - hi.x, lo.x - bhi.x), 0.0f);
    const float gy = fmaxf(fmaxf(blo.y - hi.y, lo.y - bhi.y), 0.0f);
    const float gz = fmaxf(fmaxf(blo.z - hi.z, lo.z - bhi.z), 0.0f);
    const float gap = sqrtf(gx * gx + gy * gy + gz * gz);
    lower[j] = gap > 0.0f ? 0.9998f * gap * (1.0f - 1e-6f) : -INFINITY;
  }
  float upper_min = INFINITY;
  uint64_t evaluated = 0;
  Count<false> cnt;
  while (true) {
    uint32_t next = n_objects;
    float next_gap = INFINITY;
    for (uint32_t j = 0; j < n_objects; ++j)
      if (!((evaluated >> j) & 1ull) && lower[j] < next_gap) {
        next_gap = lower[j];
        next = j;
      }
    if (next == n_objects || !(next_gap <= upper_min)) break;
    evaluated |= 1ull << next;
    const irt_object& obj = sc.objects[next];
    const irt_mesh& mesh = sc.meshes[obj.info[0]];
    float k = 1.0f;
    if (mesh.geometry_lo[3] <= 0.5f) {
      const float k2 = __uint_as_float(k2_mesh[obj.info[0]]);
      k = sqrtf(k2 + 1.0f);
    }
    float vx;
    const float dc = object_distance(sc, center, next, false, hit_eps, &vx, cnt);
    float lip_lower = dc - k * 1.001f * half_diag;
    lip_lower -= fabsf(lip_lower) * 1e-5f + 1e-6f;
    lower[next] = fmaxf(lower[next], lip_lower);
    float up = dc + k * 1.001f * half_diag;
    up += fabsf(up) * 1e-5f + 1e-6f;
    upper_min = fminf(upper_min, up);
  }
  // Candidates: lower bound <= smallest upper bound, sorted by lower bound.
  CellList out;
  out.count = 0;
  uint32_t n = 0;
  uint8_t ids[GRID_MAX_OBJECTS];
  for (uint32_t j = 0; j < n_objects; ++j) {
    if (!((evaluated >> j) & 1ull) || !(lower[j] <= upper_min)) continue;
    uint32_t k = n++;
    while (k > 0 && lower[ids[k - 1]] > lower[j]) {
      ids[k] = ids[k - 1];
      --k;
    }
    ids[k] = (uint8_t)j;
  }
  if (n > CELL_MAX) {
    out.count = CELL_OVERFLOW;
    for (uint32_t k = 0; k < CELL_MAX; ++k) out.ids[k] = 0;
  } else {
    out.count = (uint8_t)n;
    for (uint32_t k = 0; k < CELL_MAX; ++k) out.ids[k] = k < n ? ids[k] : 0;
  }
  cells[cell] = out;
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
