// The ray march of the fused and batched schedules: visibility requests
// drained by persistent waves (sdf.wgsl:61-86 per request).
//
// Two arithmetic contracts share this source.  The translation unit that
// holds the rest of the path (irt_api.hip) compiles it with one IEEE rounding
// per source operation and correctly rounded division and square root: the
// FP32 profile's march, bit for bit the FP32 restatement's.  The mixed
// profile's march is instantiated in irt_march_contracted.hip, compiled with
// multiply-add contraction inside each expression and the hardware's
// division and square root (1 ulp).  Its visibility differs from the exact
// march by a few ulps of the distances, far below what its 16-bit SDF
// samples already move them; visibility feeds only the final shading, never
// the reservoirs.  The template argument M names the contract and keeps the
// two kernels' symbols apart; a translation unit may only instantiate its own.
#pragma once

#include "irt_passes.h"

namespace irt {

__device__ __forceinline__ uint64_t lanes_below() {
  const uint32_t lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// A visibility query: surface point and normal, target (light or VPL
// position), the VPL normal (zero for lights) and the result slot.  Requests
// sit at slot 2*pixel + (0 direct, 1 indirect); `valid` marks the slots that
// hold one.  64 B.
struct RayReq {
  float px, py, pz, nx;
  float ny, nz, tx, ty;
  float tz, ex, ey, ez;
  uint32_t slot, valid, pad1, pad2;
};
static_assert(sizeof(RayReq) == 64, "ray request is 64 bytes");

constexpr uint32_t MARCH_CHUNK = 128;  // request slots claimed per atomic
constexpr uint32_t MARCH_REFILL = 12;   // idle lanes that trigger a refill

// The persistent march loop over one view's request slots (queue[1] = slots
// taken).  Every loop iteration runs one scene query per busy lane.
//
// A lane's ray lives in 15 registers: the request's twelve floats are reused
// as the ray proceeds.  Phase 0 (clearance at the surface point) reads the
// point (r[0..2]) and normal (r[3..5]) and leaves the ray origin in r[0..2];
// phase 1 (clearance at a VPL endpoint) reads the target (r[6..8]) and the
// VPL normal (r[9..11]); the set-up leaves the direction in r[3..5] and the
// march scalars in the target's and VPL normal's registers (r[6] limit, r[7]
// t, r[8] result, r[9] previous distance, r[10] step count), which the march
// steps no longer need.  The arithmetic is sdf.wgsl:61-81's, unchanged.
template <class SV, class C>
__device__ __forceinline__ void march_queue(const SV& lsc, const irt_frame& F, const RayReq* reqs,
                                            uint32_t n_rays, uint32_t* queue, float* vis, C& cnt) {
  const uint32_t lane = __lane_id();
  const float eps0 = F.sdf[1];
  const uint32_t max_steps = f2u_sat(F.sdf[0]);
  bool active = false, exhausted = false;
  uint32_t chunk_next = 0, chunk_end = 0;  // wave-uniform
  float r[12] = {};
  uint32_t slot = 0u, closest = 0u, phase = 0u;
  while (true) {
    const uint64_t idle = __ballot(!active);
    const uint32_t n_idle = (uint32_t)__popcll(idle);
    if (n_idle >= MARCH_REFILL || idle == ~0ull) {
      if (chunk_next == chunk_end && !exhausted) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(queue + 1, MARCH_CHUNK);
        base = __shfl(base, 0, 64);
        if (base >= n_rays) {
          exhausted = true;
        } else {
          chunk_next = base;
          chunk_end = min(base + MARCH_CHUNK, n_rays);
        }
      }
      const uint32_t take = min(n_idle, chunk_end - chunk_next);
      if (!active) {
        const uint32_t rank = (uint32_t)__popcll(idle & lanes_below());
        if (rank < take) {
          const RayReq rq = reqs[chunk_next + rank];
          r[0] = rq.px, r[1] = rq.py, r[2] = rq.pz, r[3] = rq.nx, r[4] = rq.ny, r[5] = rq.nz;
          r[6] = rq.tx, r[7] = rq.ty, r[8] = rq.tz, r[9] = rq.ex, r[10] = rq.ey, r[11] = rq.ez;
          slot = rq.slot;
          phase = 0;
          active = rq.valid != 0u;
          if (active) cnt.add(C_MARCH, 1);
        }
      }
      chunk_next += take;
    }
    if (__ballot(active) == 0ull) {
      if (exhausted && chunk_next == chunk_end) break;
      continue;
    }
    if (active) {
      v3 q;
      bool spacing = true;
      uint32_t hint = 0xffffffffu;
      if (phase == 0u) {
        q = v3{r[0], r[1], r[2]};
      } else if (phase == 1u) {
        q = v3{r[6], r[7], r[8]};
      } else {
        q = v3{r[0], r[1], r[2]} + v3{r[3], r[4], r[5]} * r[7];
        spacing = false;
        hint = closest;
        cnt.add(C_STEP, 1);
      }
      const Distance d = scene_distance(lsc, F.counts[3], eps0, q, spacing, hint, cnt);
      bool setup_done = false;
      v3 endpoint = v3{r[6], r[7], r[8]};
      float v = 0.0f;
      bool done = false;
      if (phase == 0u) {
        const float epsilon = fmaxf(eps0, d.voxel * 0.65f);
        const v3 origin = v3{r[0], r[1], r[2]} + v3{r[3], r[4], r[5]} * epsilon;
        r[0] = origin.x;
        r[1] = origin.y;
        r[2] = origin.z;
        closest = d.object_id;
        const v3 e = v3{r[9], r[10], r[11]};
        if (dot(e, e) > 0.5f) {
          phase = 1u;
        } else {
          setup_done = true;
        }
      } else if (phase == 1u) {
        endpoint = endpoint + v3{r[9], r[10], r[11]} * fmaxf(eps0, d.voxel * 0.65f);
        setup_done = true;
      } else {
        const float h = d.distance;
        closest = d.object_id;
        if (h < eps0) {
          v = 0.0f;
          done = true;
        } else {
          const float y = h * h / (2.0f * r[9]);
          const float dd = fsqrt(fmaxf(h * h - y * y, 0.0f));
          if (r[7] > y) r[8] = fminf(r[8], F.sdf[2] * dd / (r[7] - y));
          if (r[8] < 0.001f) {
            v = 0.0f;
            done = true;
          } else {
            r[9] = h;
            r[7] += h;
            r[10] = __uint_as_float(__float_as_uint(r[10]) + 1u);
          }
        }
      }
      if (setup_done) {
        const v3 delta = endpoint - v3{r[0], r[1], r[2]};
        const float limit = length(delta) - eps0;
        if (limit <= 0.0f) {
          v = 1.0f;
          done = true;
        } else {
          const v3 dir = normalize(delta);
          r[3] = dir.x;
          r[4] = dir.y;
          r[5] = dir.z;
          r[6] = limit;
          r[7] = eps0;                 // t
          r[8] = 1.0f;                 // result
          r[9] = 1e20f;                // previous distance
          r[10] = __uint_as_float(0u);  // steps taken
          phase = 2u;
        }
      }
      // loop-top checks of the march for the next iteration
      if (!done && phase == 2u) {
        if (__float_as_uint(r[10]) >= max_steps) {
          v = 0.0f;
          done = true;
        } else if (r[7] >= r[6]) {
          v = clampf(r[8], 0.0f, 1.0f);
          done = true;
        }
      }
      if (done) {
        vis[slot] = v;
        active = false;
      }
    }
  }
}

// Stages the query part of n object records (inverse and info: words 32-51
// of each 52-word record) into LDS, one word per thread and step.
__device__ __forceinline__ void stage_query_objects(QueryObject* dst, const irt_object* src, uint32_t n) {
  constexpr uint32_t QW = sizeof(QueryObject) / 4, OW = sizeof(irt_object) / 4;
  constexpr uint32_t OFF = offsetof(irt_object, inverse) / 4;
  static_assert(offsetof(irt_object, info) == offsetof(irt_object, inverse) + 64, "inverse, then info");
  for (uint32_t i = threadIdx.x; i < n * QW; i += blockDim.x)
    reinterpret_cast<uint32_t*>(dst)[i] = reinterpret_cast<const uint32_t*>(src)[(i / QW) * OW + OFF + i % QW];
}
// The scene view of the march: the global records except objects, meshes
// and object boxes, which it reads from LDS.
__device__ __forceinline__ QuerySceneView query_view(const SceneView& sc, const QueryObject* objects,
                                                     const irt_mesh* meshes, const float4* boxes) {
  QuerySceneView v;
  v.objects = objects;
  v.meshes = meshes;
  v.nodes = sc.nodes;
  v.lights = sc.lights;
  v.atlas = sc.atlas;
  v.atlas_w = sc.atlas_w;
  v.atlas_h = sc.atlas_h;
  v.grid = sc.grid;
  v.grid.boxes = boxes;
  v.atlas16 = sc.atlas16;
  return v;
}

// Persistent march over the request slots (queue[1] = slots taken).
// Scene records are staged in LDS once per workgroup.  Every wave leaves the
// loop once the queue is exhausted and its last march (<= sdf.x steps) ends.
constexpr uint32_t LDS_MAX_OBJECTS = 64;
template <bool CNT, int M>
__global__ __launch_bounds__(256) void k_march(SceneView sc, irt_frame F, uint32_t n_meshes, const RayReq* reqs,
                                               uint32_t n_rays, uint32_t* queue, float* vis,
                                               unsigned long long* counters) {
  static_assert(M == IRT_MARCH_MATH, "each translation unit instantiates the march of its own arithmetic");
  __shared__ QueryObject s_obj[LDS_MAX_OBJECTS];
  __shared__ irt_mesh s_mesh[LDS_MAX_OBJECTS];
  __shared__ float4 s_box[2 * LDS_MAX_OBJECTS];
  // The host launches this kernel only when the scene fits the LDS arrays
  // (objects and meshes <= LDS_MAX_OBJECTS) and the candidate grid exists.
  const uint32_t n_obj = F.counts[3];
  stage_query_objects(s_obj, sc.objects, n_obj);
  const uint32_t words_mesh = n_meshes * (uint32_t)(sizeof(irt_mesh) / 4);
  for (uint32_t i = threadIdx.x; i < words_mesh; i += blockDim.x)
    reinterpret_cast<uint32_t*>(s_mesh)[i] = reinterpret_cast<const uint32_t*>(sc.meshes)[i];
  for (uint32_t i = threadIdx.x; i < 2 * n_obj; i += blockDim.x) s_box[i] = sc.grid.boxes[i];
  __syncthreads();
  Count<CNT> cnt;
  const QuerySceneView lsc = query_view(sc, s_obj, s_mesh, s_box);
  // Each wave claims chunks of MARCH_CHUNK rays with one atomic and hands
  // them to its idle lanes; lanes refill as soon as MARCH_REFILL of them are
  // idle, so a wave is not held by its longest march.
  march_queue(lsc, F, reqs, n_rays, queue, vis, cnt);
  cnt.flush(counters);
}

// One view of a batched march: its scene, frame and request slots.
struct MarchView {
  SceneView sc;
  irt_frame F;
  const RayReq* reqs;
  uint32_t* queue;
  float* vis;
  uint32_t n_meshes, n_rays;
};

// The batched march: every workgroup walks the views in order, stages each
// view's scene in LDS and drains that view's request slots.
template <int M>
__global__ __launch_bounds__(256) void k_march_views(const MarchView* __restrict__ views, uint32_t n_views) {
  static_assert(M == IRT_MARCH_MATH, "each translation unit instantiates the march of its own arithmetic");
  __shared__ QueryObject s_obj[LDS_MAX_OBJECTS];
  __shared__ irt_mesh s_mesh[LDS_MAX_OBJECTS];
  __shared__ float4 s_box[2 * LDS_MAX_OBJECTS];
  Count<false> cnt;
  for (uint32_t v = 0; v < n_views; ++v) {
    const MarchView& mv = views[v];
    __syncthreads();  // the previous view's readers are done with the LDS records
    const uint32_t n_obj = mv.F.counts[3];
    stage_query_objects(s_obj, mv.sc.objects, n_obj);
    const uint32_t words_mesh = mv.n_meshes * (uint32_t)(sizeof(irt_mesh) / 4);
    for (uint32_t i = threadIdx.x; i < words_mesh; i += blockDim.x)
      reinterpret_cast<uint32_t*>(s_mesh)[i] = reinterpret_cast<const uint32_t*>(mv.sc.meshes)[i];
    for (uint32_t i = threadIdx.x; i < 2 * n_obj; i += blockDim.x) s_box[i] = mv.sc.grid.boxes[i];
    __syncthreads();
    const QuerySceneView lsc = query_view(mv.sc, s_obj, s_mesh, s_box);
    march_queue(lsc, mv.F, mv.reqs, mv.n_rays, mv.queue, mv.vis, cnt);
  }
}

// Launchers of the contracted instantiations (irt_march_contracted.hip).
hipError_t launch_march_contracted(bool counters, dim3 grid, hipStream_t stream, const SceneView& sc,
                                   const irt_frame& F, uint32_t n_meshes, const RayReq* reqs, uint32_t n_rays,
                                   uint32_t* queue, float* vis, unsigned long long* counter_out);
hipError_t launch_march_views_contracted(dim3 grid, hipStream_t stream, const MarchView* views, uint32_t n_views);

}  // namespace irt
