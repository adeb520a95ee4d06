// C ABI of the lighting + denoising path (include/irt.h): per-view context,
// device buffers, pass schedules, format conversion for read-back/injection.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <string>
#include <utility>
#include <vector>

#include "irt_host.h"
#include "irt_kernels.h"
#include "irt_vpl.h"
#include "irt_raster.h"
#include "irt_present.h"
#include "irt_bake.h"


using namespace irt;

namespace {

enum Fmt {
  FMT_RESERVOIR,      // irt_reservoir (same inside and outside)
  FMT_F32,            // float (Depth32Float)
  FMT_F32X4,          // float4 <-> Rgba32F
  FMT_F32X4_AS_F16X4, // internal float4 holding binary16 values <-> Rgba16F
  FMT_F16X4,          // internal binary16x4 <-> Rgba16F
  FMT_F32X2_AS_F16X2, // internal float2 <-> Rg16F
  FMT_F16X2,          // internal binary16x2 <-> Rg16F
  FMT_F16X4_AS_F32X4, // internal binary16x4 <-> Rgba32F
  FMT_MOM8,           // internal Mom8 (binary16 m1 and age, FP32 m2) <-> Rgba32F
  FMT_VPL             // irt_vpl
};

size_t internal_bytes(Fmt f) {
  switch (f) {
    case FMT_RESERVOIR: return 32;
    case FMT_F32: return 4;
    case FMT_F32X4: case FMT_F32X4_AS_F16X4: return 16;
    case FMT_F16X4: case FMT_F16X4_AS_F32X4: case FMT_MOM8: return 8;
    case FMT_F32X2_AS_F16X2: return 8;
    case FMT_F16X2: return 4;
    case FMT_VPL: return 48;
  }
  return 0;
}
size_t external_bytes(Fmt f) {
  switch (f) {
    case FMT_RESERVOIR: return 32;
    case FMT_F32: return 4;
    case FMT_F32X4: case FMT_F16X4_AS_F32X4: case FMT_MOM8: return 16;
    case FMT_F32X4_AS_F16X4: case FMT_F16X4: return 8;
    case FMT_F32X2_AS_F16X2: case FMT_F16X2: return 4;
    case FMT_VPL: return 48;
  }
  return 0;
}

struct Timer {
  std::string name;
  hipEvent_t a, b;
};

// ------------------------------------------------------ stream audit
// A host-side model of the ordering between a context's two streams (0: the
// lighting stream, 1: the SVGF stream), for checking the schedule in debug
// runs.  Every enqueued operation declares the device buffers it reads and
// writes; joins, event waits and host synchronisations record which of the
// other stream's operations a stream is ordered after.  A read after a write,
// a write after a read or a write after a write on different streams with no
// such ordering is a violation.  Operations on one stream are ordered by the
// stream itself.
class StreamAudit {
 public:
  struct Snapshot {  // a stream's knowledge when an event is recorded on it
    uint64_t epoch[2] = {0, 0};
    uint64_t seen[2] = {0, 0};
  };
  void op(int s, std::initializer_list<const void*> reads, std::initializer_list<const void*> writes,
          const char* what) {
    const uint64_t e = ++epoch_[s];
    for (const void* p : reads) {
      if (!p) continue;
      const Buf& b = bufs_[p];
      if (b.w >= 0 && !ordered(s, b.w, b.we)) report("read", what, s, p, b.wname, b.w);
    }
    for (const void* p : writes) {
      if (!p) continue;
      const Buf& b = bufs_[p];
      if (b.w >= 0 && !ordered(s, b.w, b.we)) report("write after write", what, s, p, b.wname, b.w);
      for (int t = 0; t < 2; ++t)
        if (t != s && b.r[t] && !ordered(s, t, b.r[t])) report("write after read", what, s, p, b.rname[t], t);
    }
    for (const void* p : reads)
      if (p) {
        bufs_[p].r[s] = e;
        bufs_[p].rname[s] = what;
      }
    for (const void* p : writes)
      if (p) {
        Buf& b = bufs_[p];
        b.w = s;
        b.we = e;
        b.wname = what;
        b.r[0] = b.r[1] = 0;
      }
  }
  // `to` waits for everything enqueued on `from` so far.
  void join(int from, int to) { wait(to, snapshot(from), from); }
  Snapshot snapshot(int s) const {
    Snapshot sn;
    sn.epoch[0] = epoch_[0];
    sn.epoch[1] = epoch_[1];
    sn.seen[0] = seen_[s][0];
    sn.seen[1] = seen_[s][1];
    return sn;
  }
  void wait(int to, const Snapshot& sn, int from) {
    seen_[to][from] = std::max(seen_[to][from], sn.epoch[from]);
    for (int x = 0; x < 2; ++x) seen_[to][x] = std::max(seen_[to][x], sn.seen[x]);
  }
  // The host waited for stream s: later operations on either stream follow
  // everything s had enqueued and everything s was ordered after.
  void host_sync(int s) {
    for (int x = 0; x < 2; ++x) {
      seen_[x][s] = std::max(seen_[x][s], epoch_[s]);
      for (int y = 0; y < 2; ++y) seen_[x][y] = std::max(seen_[x][y], seen_[s][y]);
    }
  }
  const std::vector<std::string>& violations() const { return log_; }
  uint64_t operations(int s) const { return epoch_[s]; }

 private:
  struct Buf {
    int w = -1;
    uint64_t we = 0;
    uint64_t r[2] = {0, 0};
    std::string wname, rname[2];
  };
  bool ordered(int s, int t, uint64_t e) const { return t == s || seen_[s][t] >= e; }
  void report(const char* kind, const char* what, int s, const void* p, const std::string& other, int t) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s: %s on stream %d touches buffer %p last used by %s on stream %d without "
                  "an ordering", kind, what, s, p, other.c_str(), t);
    log_.emplace_back(buf);
  }
  uint64_t epoch_[2] = {0, 0};
  uint64_t seen_[2][2] = {{0, 0}, {0, 0}};  // seen_[s][t]: later ops on s follow t's ops up to this epoch
  std::unordered_map<const void*, Buf> bufs_;
  std::vector<std::string> log_;
};

}  // namespace

struct irt_ctx {
  irt_settings s{};
  uint32_t precision = IRT_PRECISION_FP32;
  uint32_t schedule = IRT_SCHEDULE_PASSES;
  int device = 0;
  hipStream_t stream = nullptr;       // ingest, lighting, read-back
  hipStream_t svgf_stream = nullptr;  // SVGF of the fused schedule (overlaps the next frame's lighting)
  hipEvent_t ev_lit = nullptr;        // lighting-stream progress the SVGF stream waits for
  hipEvent_t ev_frame[2] = {nullptr, nullptr};  // SVGF work of frames, by frame parity
  uint64_t frames_ended = 0;
  uint32_t W = 0, H = 0, V = 0, slots = 0;
  size_t npx = 0;
  // scene
  irt_object* objects = nullptr;
  uint32_t n_objects = 0, cap_objects = 0;
  irt_node* nodes = nullptr;
  uint32_t n_nodes = 0, cap_nodes = 0;
  irt_mesh* meshes = nullptr;
  uint32_t n_meshes = 0, cap_meshes = 0;
  irt_light* lights = nullptr;
  uint32_t n_lights = 0, cap_lights = 0;
  float* atlas = nullptr;  // FP32 profile
  _Float16* atlas16 = nullptr;  // mixed profile
  uint32_t atlas_dims[3] = {0, 0, 0};
  irt_frame frame{};
  bool have_frame = false;
  std::vector<irt_mesh> host_meshes;     // host copies for launch-time validation
  uint32_t max_mesh_ref = 0;             // largest mesh index referenced by an object
  uint32_t max_material_ref = 0;         // largest material index referenced by an object
  // triangle geometry of the VPL pass (irt_set_geometry)
  irt_vertex* vertices = nullptr;
  uint32_t n_vertices = 0, cap_vertices = 0;
  uint32_t* indices = nullptr;
  uint32_t n_indices = 0, cap_indices = 0;
  irt_material* materials = nullptr;
  uint32_t n_materials = 0, cap_materials = 0;
  irt_tri_node* tri_nodes = nullptr;
  uint32_t n_tri_nodes = 0, cap_tri_nodes = 0;
  irt_tri* tris = nullptr;
  uint32_t n_tris = 0, cap_tris = 0;
  uint2* mesh_nodes = nullptr;  // per mesh: [first, end) of its triangle hierarchy
  uint32_t cap_mesh_nodes = 0;
  bool have_geometry = false;
  int32_t max_texture_layer = -1;  // largest texture layer a material names (-1: none)
  // material textures (irt_set_textures)
  uint32_t* tex = nullptr;
  size_t cap_tex = 0;
  uint32_t tex_size = 0, tex_layers = 0;
  float* srgb_lut = nullptr;
  uint32_t cap_srgb_lut = 0;
  // G-buffer pass (irt_render_gbuffer)
  std::vector<uint32_t> host_obj_mesh;   // mesh index of each object
  std::vector<irt_object> host_objects;  // the last uploaded object records and hierarchy
  std::vector<irt_node> host_nodes;
  std::vector<uint32_t> host_tri_base;   // instance-triangle prefix sums the device holds
  uint32_t* obj_tri_base = nullptr;
  uint32_t cap_obj_tri_base = 0;
  GbObject* gb_objs = nullptr;
  uint32_t cap_gb_objs = 0;
  GbTri* gb_tris = nullptr;
  uint32_t cap_gb_tris = 0;
  GbAttr* gb_attrs = nullptr;
  uint32_t cap_gb_attrs = 0;
  uint32_t* gb_tiles = nullptr;
  uint32_t cap_gb_tiles = 0;
  uint32_t* gb_scan = nullptr;
  uint32_t cap_gb_scan = 0;
  uint32_t* gb_scan_sums = nullptr;  // block sums, then block offsets, of the tile-count scan
  uint32_t cap_gb_scan_sums = 0;
  unsigned long long* gb_vis = nullptr;
  uint32_t cap_gb_vis = 0;
  bool gb_vis_clear = false;  // gb_vis holds the clear key (each resolve leaves it so)
  // presentation (irt_present)
  float* srgb_t = nullptr;  // sRGB decision thresholds
  uint32_t cap_srgb_t = 0;
  uchar4* present_buf = nullptr;
  uint32_t cap_present_buf = 0;
  // G-buffer slots (internal formats).  The current and previous frames'
  // G-buffers are slots gb_cur / gb_prev (the same slot right after a frame
  // ends, as when the reference copies the G-buffer); a write into the
  // current G-buffer first moves it to the slot that is neither gb_prev nor
  // gb_pp, the one no SVGF still in flight reads.  The pointers below are
  // the current bindings (bind_slots).
  float* gb_depth[3] = {nullptr, nullptr, nullptr};
  void* gb_normal[3] = {nullptr, nullptr, nullptr};
  void* gb_color[3] = {nullptr, nullptr, nullptr};
  void* gb_params[3] = {nullptr, nullptr, nullptr};
  void* gb_motion[3] = {nullptr, nullptr, nullptr};
  int gb_cur = 0, gb_prev = 0, gb_pp = 0;
  float* depth = nullptr;
  void* normal = nullptr;
  void* color = nullptr;
  void* params = nullptr;
  void* motion = nullptr;
  float* prev_depth = nullptr;
  void* prev_normal = nullptr;
  uint16_t* stage[4] = {nullptr, nullptr, nullptr, nullptr};  // binary16 staging (FP32 profile ingest)
  irt_reservoir* direct[3] = {nullptr, nullptr, nullptr};
  irt_reservoir* indirect[3] = {nullptr, nullptr, nullptr};
  irt_vpl* ring = nullptr;
  void* direct_lighting = nullptr;
  void* lit_slot[2] = {nullptr, nullptr};  // noisy radiance, alternating per frame
  int lit_cur = 0;
  void* lighting = nullptr;
  void* temporal = nullptr;
  void* scratch[2] = {nullptr, nullptr};
  void* feedback[2] = {nullptr, nullptr};
  void* history = nullptr;
  void* moments[2] = {nullptr, nullptr};  // SVGF moments: float4 (FP32 profile) or Mom8 (mixed)
  GuidePx* guide_slot[2] = {nullptr, nullptr};  // alternating per rebuild
  int guide_cur = 0;
  GuidePx* guide = nullptr;  // per-pixel guide records of the current G-buffer (fused schedule)
  bool guide_dirty = true;
  // candidate grid over the scene (fused schedule)
  CellList* cells = nullptr;
  size_t cells_cap = 0;
  float4* obj_boxes = nullptr;
  uint32_t obj_boxes_cap = 0;
  unsigned int* mesh_k2 = nullptr;
  uint32_t mesh_k2_cap = 0;
  GridView grid{};
  bool grid_dirty = true;
  bool mesh_k_dirty = true;
  // decoupled resolve (fused schedule)
  ResolveRec* resolve_rec = nullptr;
  RayReq* rays = nullptr;
  float* vis = nullptr;
  uint32_t* queue = nullptr;
  uint32_t march_blocks = 0;
  unsigned long long* counters = nullptr;
  bool counting = false;
  bool timing = false;
  bool timer_open = false;  // timer_begin recorded a start event for the pass in flight
  std::string timing_only;  // non-empty: time only the passes reported under this name
  std::vector<Timer> timers;
  size_t n_timers_used = 0;
  std::vector<std::pair<void*, size_t>> allocs;  // device allocations and their sizes
  size_t device_bytes = 0;
  std::string err;
  std::unique_ptr<StreamAudit> audit;  // IRT_STREAM_AUDIT=1 at irt_create
  StreamAudit::Snapshot ev_snap[2];    // the SVGF stream at the ev_frame records
  bool audit_dropped = false;          // the context joined a batch, whose streams the audit does not model

  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

#define HIP_CHECK(c, expr)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return (c)->fail(IRT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

namespace {

// Stream audit hooks (no-ops unless the audit is on).
int sid(const irt_ctx* c, hipStream_t h) { return h == c->svgf_stream ? 1 : 0; }
void aud(irt_ctx* c, hipStream_t h, std::initializer_list<const void*> reads,
         std::initializer_list<const void*> writes, const char* what) {
  if (c->audit) c->audit->op(sid(c, h), reads, writes, what);
}
hipError_t sync_stream(irt_ctx* c, hipStream_t h) {
  const hipError_t e = hipStreamSynchronize(h);
  if (e == hipSuccess && c->audit) c->audit->host_sync(sid(c, h));
  return e;
}

// Bytes dev_alloc allocates for a request (never zero).
size_t alloc_size(size_t bytes) { return bytes ? bytes : 256; }

// Waits for all work of both of the context's streams.  A batch step that
// used the context's buffers is included: after every step the context's
// streams wait for the step's events.
int sync_all(irt_ctx* c) {
  HIP_CHECK(c, sync_stream(c, c->svgf_stream));
  HIP_CHECK(c, sync_stream(c, c->stream));
  return IRT_OK;
}

// Allocates and zero-fills.  The fill has finished when this returns, so a
// first consumer on any stream (the SVGF stream, a batch stream) sees zeros:
// the fill is enqueued on the lighting stream, and nothing else would order
// it before work on the other streams.  Allocation happens at set-up and when
// a grow-only array grows, never per frame, so the wait costs nothing in
// steady state.
int dev_alloc(irt_ctx* c, void** p, size_t bytes) {
  bytes = alloc_size(bytes);
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return c->fail(IRT_ERR_OUT_OF_MEMORY, std::string("hipMalloc: ") + hipGetErrorString(e));
  c->allocs.emplace_back(*p, bytes);
  c->device_bytes += bytes;
  e = hipMemsetAsync(*p, 0, bytes, c->stream);
  if (e != hipSuccess) return c->fail(IRT_ERR_DEVICE, std::string("hipMemsetAsync: ") + hipGetErrorString(e));
  aud(c, c->stream, {}, {*p}, "allocation zero-fill");
  e = sync_stream(c, c->stream);
  if (e != hipSuccess) return c->fail(IRT_ERR_DEVICE, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
  return IRT_OK;
}

// Frees one allocation of the context (stream work using it must be done).
int dev_free(irt_ctx* c, void* p) {
  if (!p) return IRT_OK;
  for (auto& a : c->allocs)
    if (a.first == p) {
      c->device_bytes -= a.second;
      a = {nullptr, 0};
    }
  HIP_CHECK(c, hipFree(p));
  return IRT_OK;
}

// Grow-only device array for small scene records.
template <class T>
int ensure_array(irt_ctx* c, T** p, uint32_t* cap, uint32_t count) {
  uint32_t need = count > 0 ? count : 1;
  if (*p && *cap >= need) return IRT_OK;
  if (*p) {
    int rc = sync_all(c);  // readers may be on either stream (presentation runs on the SVGF stream)
    if (rc) return rc;
    rc = dev_free(c, *p);
    if (rc) return rc;
    *p = nullptr;
  }
  int rc = dev_alloc(c, (void**)p, sizeof(T) * need);
  if (rc) return rc;
  *cap = need;
  return IRT_OK;
}

bool mixed(const irt_ctx* c) { return c->precision == IRT_PRECISION_MIXED; }

// Current bindings of the slot buffers.
void bind_slots(irt_ctx* c) {
  c->depth = c->gb_depth[c->gb_cur];
  c->normal = c->gb_normal[c->gb_cur];
  c->color = c->gb_color[c->gb_cur];
  c->params = c->gb_params[c->gb_cur];
  c->motion = c->gb_motion[c->gb_cur];
  c->prev_depth = c->gb_depth[c->gb_prev];
  c->prev_normal = c->gb_normal[c->gb_prev];
  c->lighting = c->lit_slot[c->lit_cur];
  c->guide = c->guide_slot[c->guide_cur];
}

// Before the lighting stream writes a slot an SVGF still in flight may read
// (the G-buffer slot of two frames back, the other guide and noisy-radiance
// slots): wait for the SVGF work of the frame before the previous one.
int wait_for_slots(irt_ctx* c) {
  if (c->frames_ended >= 2) {
    HIP_CHECK(c, hipStreamWaitEvent(c->stream, c->ev_frame[c->frames_ended % 2], 0));
    if (c->audit) c->audit->wait(sid(c, c->stream), c->ev_snap[c->frames_ended % 2], 1);
  }
  return IRT_OK;
}

// Makes the current G-buffer a slot of its own (not the previous frame's
// and not the one before), copying the planes across when `keep` is set.
int make_current_writable(irt_ctx* c, bool keep) {
  if (c->gb_cur != c->gb_prev && c->gb_cur != c->gb_pp) return IRT_OK;
  int w = 0;
  while (w == c->gb_prev || w == c->gb_pp) ++w;
  int rc = wait_for_slots(c);
  if (rc) return rc;
  if (keep) {
    const size_t n = c->npx, n4 = c->precision == IRT_PRECISION_MIXED ? 8 : 16, p2 = n4 / 2;
    const int f = c->gb_cur;
    HIP_CHECK(c, hipMemcpyAsync(c->gb_depth[w], c->gb_depth[f], n * 4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->gb_normal[w], c->gb_normal[f], n * n4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->gb_color[w], c->gb_color[f], n * n4, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->gb_params[w], c->gb_params[f], n * p2, hipMemcpyDeviceToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->gb_motion[w], c->gb_motion[f], n * p2, hipMemcpyDeviceToDevice, c->stream));
    aud(c, c->stream, {c->gb_depth[f], c->gb_normal[f], c->gb_color[f], c->gb_params[f], c->gb_motion[f]},
        {c->gb_depth[w], c->gb_normal[w], c->gb_color[w], c->gb_params[w], c->gb_motion[w]}, "G-buffer slot copy");
  }
  c->gb_cur = w;
  bind_slots(c);
  return IRT_OK;
}

// The end of a frame (renderer.rs:1946-1960 copies the G-buffer into the
// previous-frame textures): the current slot becomes the previous one.
int end_frame_slots(irt_ctx* c) {
  HIP_CHECK(c, hipEventRecord(c->ev_frame[c->frames_ended % 2], c->svgf_stream));
  if (c->audit) c->ev_snap[c->frames_ended % 2] = c->audit->snapshot(1);
  c->frames_ended++;
  c->gb_pp = c->gb_prev;
  c->gb_prev = c->gb_cur;
  bind_slots(c);
  return IRT_OK;
}

// Orders `to` after everything enqueued so far on `from`.
int join(irt_ctx* c, hipStream_t from, hipStream_t to) {
  HIP_CHECK(c, hipEventRecord(c->ev_lit, from));
  HIP_CHECK(c, hipStreamWaitEvent(to, c->ev_lit, 0));
  if (c->audit) c->audit->join(sid(c, from), sid(c, to));
  return IRT_OK;
}

uint32_t compute_units(int device) {
  static int cached_device = -1;
  static uint32_t cached = 0;
  if (device != cached_device) {
    hipDeviceProp_t prop;
    cached = 256;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
      cached = (uint32_t)prop.multiProcessorCount;
    cached_device = device;
  }
  return cached;
}

// Every per-view device buffer (G-buffer, history, reservoirs, VPL ring,
// SVGF targets, guide records, ray requests) with its size; the scene
// records and the SDF atlas are allocated by their setters.
std::vector<std::pair<void**, size_t>> view_alloc_plan(irt_ctx* c, bool m) {
  const size_t n = c->npx;
  const size_t n4 = m ? 8 : 16, p2 = m ? 4 : 8, r4 = m ? 8 : 16, l4 = m ? 8 : 16;
  std::vector<std::pair<void**, size_t>> v;
  auto A = [&](void* slot, size_t bytes) { v.emplace_back((void**)slot, bytes); };
  for (int k = 0; k < 3; ++k) {
    A(&c->gb_depth[k], n * 4);
    A(&c->gb_normal[k], n * n4);
    A(&c->gb_color[k], n * n4);
    A(&c->gb_params[k], n * p2);
    A(&c->gb_motion[k], n * p2);
  }
  if (!m) {
    A(&c->stage[0], n * 8);
    A(&c->stage[1], n * 8);
    A(&c->stage[2], n * 4);
    A(&c->stage[3], n * 4);
  }
  for (int i = 0; i < 3; ++i) {
    A(&c->direct[i], n * 32);
    A(&c->indirect[i], n * 32);
  }
  A(&c->ring, (size_t)c->slots * c->V * sizeof(irt_vpl));
  A(&c->direct_lighting, n * l4);
  A(&c->lit_slot[0], n * r4);
  A(&c->lit_slot[1], n * r4);
  A(&c->temporal, n * r4);
  A(&c->scratch[0], n * r4);
  A(&c->scratch[1], n * r4);
  A(&c->feedback[0], n * r4);
  A(&c->feedback[1], n * r4);
  A(&c->history, n * r4);
  A(&c->moments[0], n * (m ? sizeof(Mom8) : 16));
  A(&c->moments[1], n * (m ? sizeof(Mom8) : 16));
  A(&c->counters, 8 * sizeof(unsigned long long));
  A(&c->guide_slot[0], n * sizeof(GuidePx));
  A(&c->guide_slot[1], n * sizeof(GuidePx));
  A(&c->resolve_rec, n * sizeof(ResolveRec));
  A(&c->rays, 2 * n * sizeof(RayReq));
  A(&c->vis, 2 * n * sizeof(float));
  A(&c->queue, 64);
  return v;
}

struct BufDesc {
  void* ptr;
  Fmt fmt;
  size_t count;
};

BufDesc describe(irt_ctx* c, uint32_t b) {
  const bool m = mixed(c);
  const uint32_t cur = c->frame.size[2] % 2u, old = 1u - cur;
  const Fmt n4 = m ? FMT_F16X4 : FMT_F32X4_AS_F16X4;
  const Fmt p2 = m ? FMT_F16X2 : FMT_F32X2_AS_F16X2;
  const Fmt r4 = m ? FMT_F16X4 : FMT_F32X4_AS_F16X4;
  const size_t n = c->npx;
  switch (b) {
    case IRT_BUF_DIRECT0: case IRT_BUF_DIRECT1: case IRT_BUF_DIRECT2:
      return {c->direct[b - IRT_BUF_DIRECT0], FMT_RESERVOIR, n};
    case IRT_BUF_INDIRECT0: case IRT_BUF_INDIRECT1: case IRT_BUF_INDIRECT2:
      return {c->indirect[b - IRT_BUF_INDIRECT0], FMT_RESERVOIR, n};
    case IRT_BUF_DIRECT_LIGHTING: return {c->direct_lighting, m ? FMT_F16X4_AS_F32X4 : FMT_F32X4, n};
    case IRT_BUF_LIGHTING: return {c->lighting, r4, n};
    case IRT_BUF_HISTORY: return {c->history, r4, n};
    case IRT_BUF_MOMENTS: return {c->moments[cur], m ? FMT_MOM8 : FMT_F32X4, n};
    case IRT_BUF_MOMENTS_OLD: return {c->moments[old], m ? FMT_MOM8 : FMT_F32X4, n};
    case IRT_BUF_FEEDBACK: return {c->feedback[cur], r4, n};
    case IRT_BUF_FEEDBACK_OLD: return {c->feedback[old], r4, n};
    case IRT_BUF_TEMPORAL: return {c->temporal, r4, n};
    case IRT_BUF_DEPTH: return {c->depth, FMT_F32, n};
    case IRT_BUF_NORMAL: return {c->normal, n4, n};
    case IRT_BUF_COLOR: return {c->color, n4, n};
    case IRT_BUF_PARAMS: return {c->params, p2, n};
    case IRT_BUF_MOTION: return {c->motion, p2, n};
    case IRT_BUF_PREVIOUS_DEPTH: return {c->prev_depth, FMT_F32, n};
    case IRT_BUF_PREVIOUS_NORMAL: return {c->prev_normal, n4, n};
    case IRT_BUF_VPL_ARCHIVE: return {c->ring, FMT_VPL, (size_t)c->slots * c->V};
    case IRT_BUF_SCRATCH0: return {c->scratch[0], r4, n};
    case IRT_BUF_SCRATCH1: return {c->scratch[1], r4, n};
  }
  return {nullptr, FMT_F32, 0};
}

// internal -> external (reference format), host side.
void to_external(Fmt f, const uint8_t* in, uint8_t* out, size_t count) {
  switch (f) {
    case FMT_F32X4_AS_F16X4: {
      const float* s = (const float*)in;
      uint16_t* d = (uint16_t*)out;
      for (size_t i = 0; i < count * 4; ++i) d[i] = irt_host::f32_to_f16(s[i]);
      break;
    }
    case FMT_F32X2_AS_F16X2: {
      const float* s = (const float*)in;
      uint16_t* d = (uint16_t*)out;
      for (size_t i = 0; i < count * 2; ++i) d[i] = irt_host::f32_to_f16(s[i]);
      break;
    }
    case FMT_F16X4_AS_F32X4: {
      const uint16_t* s = (const uint16_t*)in;
      float* d = (float*)out;
      for (size_t i = 0; i < count * 4; ++i) d[i] = irt_host::f16_to_f32(s[i]);
      break;
    }
    case FMT_MOM8: {  // (m1 f16, age f16, m2 f32) -> (m1, m2, age, 0)
      float* d = (float*)out;
      for (size_t i = 0; i < count; ++i) {
        uint16_t h[2];
        float m2;
        std::memcpy(h, in + 8 * i, 4);
        std::memcpy(&m2, in + 8 * i + 4, 4);
        d[4 * i] = irt_host::f16_to_f32(h[0]);
        d[4 * i + 1] = m2;
        d[4 * i + 2] = irt_host::f16_to_f32(h[1]);
        d[4 * i + 3] = 0.0f;
      }
      break;
    }
    default:
      std::memcpy(out, in, count * internal_bytes(f));
  }
}
void to_internal(Fmt f, const uint8_t* in, uint8_t* out, size_t count) {
  switch (f) {
    case FMT_F32X4_AS_F16X4: {
      const uint16_t* s = (const uint16_t*)in;
      float* d = (float*)out;
      for (size_t i = 0; i < count * 4; ++i) d[i] = irt_host::f16_to_f32(s[i]);
      break;
    }
    case FMT_F32X2_AS_F16X2: {
      const uint16_t* s = (const uint16_t*)in;
      float* d = (float*)out;
      for (size_t i = 0; i < count * 2; ++i) d[i] = irt_host::f16_to_f32(s[i]);
      break;
    }
    case FMT_F16X4_AS_F32X4: {
      const float* s = (const float*)in;
      uint16_t* d = (uint16_t*)out;
      for (size_t i = 0; i < count * 4; ++i) d[i] = irt_host::f32_to_f16(s[i]);
      break;
    }
    case FMT_MOM8: {
      const float* s = (const float*)in;
      for (size_t i = 0; i < count; ++i) {
        const uint16_t h[2] = {irt_host::f32_to_f16(s[4 * i]), irt_host::f32_to_f16(s[4 * i + 2])};
        std::memcpy(out + 8 * i, h, 4);
        std::memcpy(out + 8 * i + 4, &s[4 * i + 1], 4);
      }
      break;
    }
    default:
      std::memcpy(out, in, count * internal_bytes(f));
  }
}

// ------------------------------------------------------------- timing
void timer_begin(irt_ctx* c, const char* name) {
  c->timer_open = false;
  if (!c->timing || (!c->timing_only.empty() && c->timing_only != name)) return;
  c->timer_open = true;
  if (c->n_timers_used >= c->timers.size()) {
    Timer t;
    t.name = name;
    (void)hipEventCreate(&t.a);
    (void)hipEventCreate(&t.b);
    c->timers.push_back(t);
  }
  Timer& t = c->timers[c->n_timers_used];
  t.name = name;
  (void)hipEventRecord(t.a, c->stream);
}
void timer_end(irt_ctx* c) {
  if (!c->timer_open) return;
  c->timer_open = false;
  (void)hipEventRecord(c->timers[c->n_timers_used].b, c->stream);
  c->n_timers_used++;
}

// ------------------------------------------------------------- launches
dim3 grid_of(const irt_ctx* c) { return dim3((c->W + TILE - 1) / TILE, (c->H + TILE - 1) / TILE, 1); }
dim3 grid_banded(const irt_ctx* c) { return dim3(XCDS * band_width(c->W) * ((c->H + TILE - 1) / TILE), 1, 1); }

template <class P>
LightArgs<P> light_args(irt_ctx* c, int in, int out, int hist) {
  LightArgs<P> a;
  std::memset(&a, 0, sizeof(a));
  a.F = c->frame;
  a.sc = SceneView{c->objects, c->meshes, c->nodes, c->lights, c->atlas, c->atlas_dims[0], c->atlas_dims[1],
                   GridView{}, c->atlas16};
  a.depth = c->depth;
  a.normal = (const typename P::N4*)c->normal;
  a.color = (const typename P::N4*)c->color;
  a.params = (const typename P::P2*)c->params;
  a.motion = (const typename P::P2*)c->motion;
  a.prev_depth = c->prev_depth;
  a.prev_normal = (const typename P::N4*)c->prev_normal;
  a.direct_in = c->direct[in];
  a.indirect_in = c->indirect[in];
  a.direct_out = c->direct[out];
  a.indirect_out = c->indirect[out];
  a.direct_history = c->direct[hist];
  a.indirect_history = c->indirect[hist];
  a.vpl_ring = c->ring;
  a.direct_lighting = (typename P::L4*)c->direct_lighting;
  a.lighting = (typename P::R4*)c->lighting;
  a.counters = c->counters;
  return a;
}

template <class P>
SvgfArgs<P> svgf_args(irt_ctx* c) {
  SvgfArgs<P> a;
  std::memset(&a, 0, sizeof(a));
  a.F = c->frame;
  a.depth = c->depth;
  a.normal = (const typename P::N4*)c->normal;
  a.color = (const typename P::N4*)c->color;
  a.params = (const typename P::P2*)c->params;
  a.motion = (const typename P::P2*)c->motion;
  a.old_depth = c->prev_depth;
  a.old_normal = (const typename P::N4*)c->prev_normal;
  a.noisy = (const typename P::R4*)c->lighting;
  a.counters = c->counters;
  a.stride = 1;
  return a;
}

#define LAUNCH_SVGF(c, kern, P, fast, ...)                                                               \
  do {                                                                                                   \
    if ((c)->counting)                                                                                   \
      hipLaunchKernelGGL((kern<P, true, false>), grid_of(c), dim3(TILE, TILE, 1), 0, (c)->stream, __VA_ARGS__); \
    else if (fast)                                                                                       \
      hipLaunchKernelGGL((kern<P, false, true>), grid_of(c), dim3(TILE, TILE, 1), 0, (c)->stream, __VA_ARGS__); \
    else                                                                                                 \
      hipLaunchKernelGGL((kern<P, false, false>), grid_of(c), dim3(TILE, TILE, 1), 0, (c)->stream, __VA_ARGS__); \
    HIP_CHECK(c, hipGetLastError());                                                                     \
  } while (0)

#define LAUNCH_GRID(c, grid, kern, P, ...)                                                               \
  do {                                                                                                   \
    if ((c)->counting)                                                                                   \
      hipLaunchKernelGGL((kern<P, true>), grid, dim3(TILE, TILE, 1), 0, (c)->stream, __VA_ARGS__);       \
    else                                                                                                 \
      hipLaunchKernelGGL((kern<P, false>), grid, dim3(TILE, TILE, 1), 0, (c)->stream, __VA_ARGS__);      \
    HIP_CHECK(c, hipGetLastError());                                                                     \
  } while (0)
#define LAUNCH(c, kern, P, ...) LAUNCH_GRID(c, grid_of(c), kern, P, __VA_ARGS__)

int check_ready(irt_ctx* c) {
  if (!c->have_frame) return c->fail(IRT_ERR_STATE, "no frame set (irt_set_frame)");
  if (!c->objects || c->n_objects == 0) return c->fail(IRT_ERR_STATE, "no objects set (irt_set_objects)");
  if (!c->meshes || c->n_meshes == 0) return c->fail(IRT_ERR_STATE, "no meshes set (irt_set_meshes)");
  if (!c->atlas && !c->atlas16) return c->fail(IRT_ERR_STATE, "no SDF atlas set (irt_set_sdf_atlas)");
  if (c->max_mesh_ref >= c->n_meshes) return c->fail(IRT_ERR_STATE, "an object references a missing mesh");
  for (const irt_mesh& m : c->host_meshes) {
    const uint32_t res = m.field[3];
    if (m.field[0] + res > c->atlas_dims[0] || m.field[1] + res > c->atlas_dims[1] ||
        m.field[2] + res > c->atlas_dims[2])
      return c->fail(IRT_ERR_STATE, "mesh SDF tile lies outside the atlas");
  }
  if (c->frame.counts[0] > c->n_lights) return c->fail(IRT_ERR_STATE, "frame light count exceeds uploaded lights");
  if (c->frame.counts[3] != c->n_objects) return c->fail(IRT_ERR_STATE, "frame object count differs from uploaded objects");
  return IRT_OK;
}

// Rebuilds the guide records when the G-buffer or the frame changed.
template <class P>
int ensure_guide(irt_ctx* c) {
  if (!c->guide_dirty) return IRT_OK;
  // the other slot: an SVGF of the previous frame may still read this one
  int rc = wait_for_slots(c);
  if (rc) return rc;
  c->guide_cur ^= 1;
  bind_slots(c);
  const uint32_t n = (uint32_t)c->npx;
  hipLaunchKernelGGL((k_guide<P>), dim3((n + 255) / 256), dim3(256), 0, c->stream, c->frame, c->depth,
                     (const typename P::N4*)c->normal, (const typename P::N4*)c->color,
                     (const typename P::P2*)c->params, c->guide);
  HIP_CHECK(c, hipGetLastError());
  aud(c, c->stream, {c->depth, c->normal, c->color, c->params}, {c->guide}, "guide records");
  c->guide_dirty = false;
  return IRT_OK;
}

// Rebuilds the candidate grid when objects, meshes or the atlas changed.
// Leaves c->grid.cells null (hierarchy walk) for scenes above the grid's
// object limit.
int ensure_grid(irt_ctx* c) {
  if (!c->grid_dirty) return IRT_OK;
  c->grid.cells = nullptr;
  if (c->n_objects == 0 || c->n_objects > GRID_MAX_OBJECTS || (!c->atlas && !c->atlas16) || !c->meshes) {
    c->grid_dirty = false;
    return IRT_OK;
  }
  SceneView sc{c->objects, c->meshes, c->nodes, c->lights, c->atlas, c->atlas_dims[0], c->atlas_dims[1], GridView{},
               c->atlas16};
  if (c->mesh_k_dirty) {
    int rc = ensure_array(c, &c->mesh_k2, &c->mesh_k2_cap, c->n_meshes);
    if (rc) return rc;
    HIP_CHECK(c, hipMemsetAsync(c->mesh_k2, 0, sizeof(unsigned int) * c->n_meshes, c->stream));
    hipLaunchKernelGGL(k_mesh_gradient, dim3(64, c->n_meshes), dim3(256), 0, c->stream, sc, c->n_meshes, c->mesh_k2);
    HIP_CHECK(c, hipGetLastError());
    c->mesh_k_dirty = false;
  }
  const size_t total = (size_t)c->grid.nx * c->grid.ny * c->grid.nz;
  if (c->cells_cap < total) {
    HIP_CHECK(c, sync_stream(c, c->stream));
    int rc = dev_free(c, c->cells);
    if (rc) return rc;
    c->cells = nullptr;
    rc = dev_alloc(c, (void**)&c->cells, total * sizeof(CellList));
    if (rc) return rc;
    c->cells_cap = total;
  }
  GridView g = c->grid;
  g.cells = c->cells;
  g.boxes = c->obj_boxes;
  sc.grid = g;
  hipLaunchKernelGGL(k_build_cells, dim3((uint32_t)((total + 63) / 64)), dim3(64), 0, c->stream, sc, c->n_objects,
                     c->s.sdf_hit_epsilon, (const unsigned int*)c->mesh_k2, c->cells);
  HIP_CHECK(c, hipGetLastError());
  c->grid = g;
  c->grid_dirty = false;
  return IRT_OK;
}

// One reservoir/resolve pass with the reference's bindings (pass_group, renderer.rs:1617-1674).
template <class P>
int lighting_pass(irt_ctx* c, uint32_t pass) {
  switch (pass) {
    case IRT_PASS_CANDIDATES_DIRECT: LAUNCH(c, k_candidates_direct, P, light_args<P>(c, 2, 0, 2)); break;
    case IRT_PASS_CANDIDATES_INDIRECT: LAUNCH(c, k_candidates_indirect, P, light_args<P>(c, 2, 0, 2)); break;
    case IRT_PASS_TEMPORAL_DIRECT: LAUNCH(c, k_temporal_direct, P, light_args<P>(c, 0, 1, 2)); break;
    case IRT_PASS_TEMPORAL_INDIRECT: LAUNCH(c, k_temporal_indirect, P, light_args<P>(c, 0, 1, 2)); break;
    case IRT_PASS_SPATIAL_DIRECT: LAUNCH(c, k_spatial_direct, P, light_args<P>(c, 1, 2, 0)); break;
    case IRT_PASS_SPATIAL_INDIRECT: LAUNCH(c, k_spatial_indirect, P, light_args<P>(c, 1, 2, 0)); break;
    case IRT_PASS_RESOLVE_DIRECT: LAUNCH(c, k_resolve_direct, P, light_args<P>(c, 2, 0, 2)); break;
    case IRT_PASS_RESOLVE_INDIRECT: LAUNCH(c, k_resolve_indirect, P, light_args<P>(c, 2, 0, 2)); break;
    default: return c->fail(IRT_ERR_INVALID_ARGUMENT, "not a lighting pass");
  }
  return IRT_OK;
}

template <class P>
int run_lighting(irt_ctx* c) {
  int rc = check_ready(c);
  if (rc) return rc;
  // this frame's noisy radiance goes to the other slot (the previous frame's
  // SVGF may still read the current one)
  rc = wait_for_slots(c);
  if (rc) return rc;
  c->lit_cur ^= 1;
  bind_slots(c);
  if (c->schedule == IRT_SCHEDULE_FUSED) {
    timer_begin(c, "reservoirs_a");
    rc = ensure_guide<P>(c);
    if (!rc) rc = ensure_grid(c);
    if (rc) return rc;
    LightArgs<P> a = light_args<P>(c, 2, 1, 2);
    a.guide = c->guide;
    a.sc.grid = c->grid;
    LAUNCH(c, k_reservoirs_a, P, a, c->direct[0], c->indirect[0]);
    aud(c, c->stream, {c->guide, c->motion, c->prev_depth, c->prev_normal, c->direct[2], c->indirect[2], c->ring},
        {c->direct[0], c->direct[1], c->indirect[0], c->indirect[1]}, "candidates + temporal reuse");
    timer_end(c);
    LightArgs<P> b = light_args<P>(c, 1, 2, 0);
    b.guide = c->guide;
    b.sc.grid = c->grid;
    if (!c->grid.cells || c->n_objects > LDS_MAX_OBJECTS || c->n_meshes > LDS_MAX_OBJECTS) {
      // scene beyond the LDS staging limits: one kernel per pixel
      timer_begin(c, "reservoirs_b_resolve");
      LAUNCH(c, k_reservoirs_b, P, b);
      aud(c, c->stream, {c->direct[1], c->indirect[1], c->guide, c->motion, c->ring},
          {c->direct[2], c->indirect[2], c->direct_lighting, c->lighting}, "spatial reuse + resolves");
      timer_end(c);
      return IRT_OK;
    }
    timer_begin(c, "spatial_setup");
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->stream, c->queue, 16u);
    LAUNCH_GRID(c, grid_banded(c), k_spatial_setup, P, b, c->resolve_rec, c->rays);
    aud(c, c->stream, {c->direct[1], c->indirect[1], c->guide, c->ring},
        {c->direct[2], c->indirect[2], c->resolve_rec, c->rays, c->queue}, "spatial reuse + ray set-up");
    timer_end(c);
    timer_begin(c, "march");
    if (mixed(c))
      HIP_CHECK(c, launch_march_contracted(c->counting, dim3(c->march_blocks), c->stream, b.sc, c->frame, c->n_meshes,
                                           c->rays, (uint32_t)(2 * c->npx), c->queue, c->vis, c->counters));
    else if (c->counting)
      hipLaunchKernelGGL((k_march<true, IRT_MARCH_EXACT>), dim3(c->march_blocks), dim3(256), 0, c->stream, b.sc,
                         c->frame, c->n_meshes, c->rays, (uint32_t)(2 * c->npx), c->queue, c->vis, c->counters);
    else
      hipLaunchKernelGGL((k_march<false, IRT_MARCH_EXACT>), dim3(c->march_blocks), dim3(256), 0, c->stream, b.sc,
                         c->frame, c->n_meshes, c->rays, (uint32_t)(2 * c->npx), c->queue, c->vis, c->counters);
    HIP_CHECK(c, hipGetLastError());
    aud(c, c->stream, {c->rays, c->queue}, {c->vis, c->queue}, "march");
    timer_end(c);
    timer_begin(c, "resolve_combine");
    hipLaunchKernelGGL((k_resolve_combine<P>), grid_of(c), dim3(TILE, TILE, 1), 0, c->stream, b, c->resolve_rec,
                       c->vis);
    HIP_CHECK(c, hipGetLastError());
    aud(c, c->stream, {c->resolve_rec, c->vis, c->guide, c->motion}, {c->direct_lighting, c->lighting},
        "resolve combine");
    timer_end(c);
    return IRT_OK;
  }
  static const char* names[8] = {"candidates_direct", "candidates_indirect", "temporal_direct",
                                 "temporal_indirect", "spatial_direct", "spatial_indirect",
                                 "visibility_direct", "visibility_indirect"};
  for (uint32_t p = 0; p < 8; ++p) {
    timer_begin(c, names[p]);
    rc = lighting_pass<P>(c, p);
    if (rc) return rc;
    aud(c, c->stream, {c->depth, c->normal, c->color, c->params, c->motion, c->prev_depth, c->prev_normal,
                       c->direct[0], c->direct[1], c->direct[2], c->indirect[0], c->indirect[1], c->indirect[2], c->ring},
        {c->direct[0], c->direct[1], c->direct[2], c->indirect[0], c->indirect[1], c->indirect[2], c->direct_lighting,
         c->lighting}, names[p]);
    timer_end(c);
  }
  return IRT_OK;
}

template <class P>
int svgf_temporal_pass(irt_ctx* c, const GuidePx* guide = nullptr) {
  const uint32_t cur = c->frame.size[2] % 2u, old = 1u - cur;
  aud(c, c->stream, {c->lighting, c->feedback[old], c->moments[old], guide, c->depth, c->normal, c->color, c->params,
                     c->motion, c->prev_depth, c->prev_normal}, {c->temporal, c->moments[cur]}, "SVGF temporal");
  SvgfArgs<P> a = svgf_args<P>(c);
  a.guide = guide;
  a.old_radiance = (const typename P::R4*)c->feedback[old];
  a.old_moments = (const typename P::M4*)c->moments[old];
  a.input = (const typename P::R4*)c->feedback[old];
  a.moments = (const typename P::M4*)c->moments[old];
  a.out = (typename P::R4*)c->temporal;
  a.moments_out = (typename P::M4*)c->moments[cur];
  if (guide) {
    const dim3 grid((c->W + TMP_TILE - 1) / TMP_TILE, (c->H + TMP_TILE - 1) / TMP_TILE);
    if (c->counting)
      hipLaunchKernelGGL((k_svgf_temporal_lds<P, true>), grid, dim3(256), 0, c->stream, a);
    else
      hipLaunchKernelGGL((k_svgf_temporal_lds<P, false>), grid, dim3(256), 0, c->stream, a);
    HIP_CHECK(c, hipGetLastError());
    return IRT_OK;
  }
  LAUNCH_SVGF(c, k_svgf_temporal, P, false, a);
  return IRT_OK;
}
template <class P>
int svgf_variance_pass(irt_ctx* c, const GuidePx* guide = nullptr) {
  const uint32_t cur = c->frame.size[2] % 2u;
  aud(c, c->stream, {c->temporal, c->moments[cur], guide, c->depth, c->normal, c->color, c->params},
      {c->scratch[0]}, "SVGF variance");
  SvgfArgs<P> a = svgf_args<P>(c);
  a.guide = guide;
  a.input = (const typename P::R4*)c->temporal;
  a.moments = (const typename P::M4*)c->moments[cur];
  a.out = (typename P::R4*)c->scratch[0];
  if (guide) {
    const dim3 grid((c->W + VAR_TILE - 1) / VAR_TILE, (c->H + VAR_TILE - 1) / VAR_TILE);
    if (c->counting)
      hipLaunchKernelGGL((k_svgf_variance_lds<P, true>), grid, dim3(256), 0, c->stream, a);
    else
      hipLaunchKernelGGL((k_svgf_variance_lds<P, false>), grid, dim3(256), 0, c->stream, a);
    HIP_CHECK(c, hipGetLastError());
    return IRT_OK;
  }
  LAUNCH_SVGF(c, k_svgf_variance, P, false, a);
  return IRT_OK;
}
template <class P>
int svgf_atrous_pass(irt_ctx* c, uint32_t i, void* in, void* out, void* out2, const GuidePx* guide = nullptr) {
  const uint32_t cur = c->frame.size[2] % 2u;
  aud(c, c->stream, {in, guide, c->depth, c->normal, c->color, c->params}, {out, out2}, "SVGF a-trous");
  SvgfArgs<P> a = svgf_args<P>(c);
  a.guide = guide;
  a.input = (const typename P::R4*)in;
  a.moments = (const typename P::M4*)c->moments[cur];
  a.out = (typename P::R4*)out;
  a.out2 = (typename P::R4*)out2;
  a.stride = 1u << i;
  if (guide) {
    // fused schedule: taps staged in LDS per 16 x 16 patch of a stride-lattice coset
    const uint32_t span = (uint32_t)AT_TILE << i;
    const uint32_t regions = ((c->W + span - 1) / span) * ((c->H + span - 1) / span);
    const dim3 grid(((regions + 7) / 8) * 8 * (1u << (2 * i)));
    if (c->counting)
      hipLaunchKernelGGL((k_svgf_atrous_lds<P, true>), grid, dim3(256), 0, c->stream, a);
    else
      hipLaunchKernelGGL((k_svgf_atrous_lds<P, false>), grid, dim3(256), 0, c->stream, a);
    HIP_CHECK(c, hipGetLastError());
    return IRT_OK;
  }
  LAUNCH_SVGF(c, k_svgf_atrous, P, false, a);
  return IRT_OK;
}

size_t r4_bytes(const irt_ctx* c) { return mixed(c) ? 8 : 16; }

// Points c->stream at the SVGF stream for its lifetime (launch helpers and
// timers use c->stream).
struct StreamSwap {
  irt_ctx* c;
  hipStream_t saved;
  explicit StreamSwap(irt_ctx* ctx) : c(ctx), saved(ctx->stream) { c->stream = c->svgf_stream; }
  ~StreamSwap() { c->stream = saved; }
};

template <class P>
int run_svgf_on(irt_ctx* c, uint32_t iterations, uint32_t debug);

// Svgf::encode — svgf.rs:140-261.
template <class P>
int run_svgf(irt_ctx* c, uint32_t iterations, uint32_t debug) {
  if (!c->have_frame) return c->fail(IRT_ERR_STATE, "no frame set (irt_set_frame)");
  if (iterations == 0 || iterations > 5) return c->fail(IRT_ERR_INVALID_ARGUMENT, "SVGF iterations must be 1..5");
  if (c->schedule == IRT_SCHEDULE_FUSED) {
    // The fused schedule runs SVGF on its own stream, after everything the
    // lighting stream has queued, so it overlaps the next frame's lighting.
    int rc = ensure_guide<P>(c);
    if (!rc) rc = join(c, c->stream, c->svgf_stream);
    if (rc) return rc;
    StreamSwap swap(c);
    return run_svgf_on<P>(c, iterations, debug);
  }
  return run_svgf_on<P>(c, iterations, debug);
}

// The SVGF passes on c->stream (the SVGF stream while a StreamSwap holds).
template <class P>
int run_svgf_on(irt_ctx* c, uint32_t iterations, uint32_t debug) {
  const uint32_t cur = c->frame.size[2] % 2u;
  const size_t rb = r4_bytes(c) * c->npx;
  const bool fused = c->schedule == IRT_SCHEDULE_FUSED;
  if (debug >= 3) {
    timer_begin(c, "svgf");
    HIP_CHECK(c, hipMemcpyAsync(c->history, c->lighting, rb, hipMemcpyDeviceToDevice, c->stream));
    aud(c, c->stream, {c->lighting}, {c->history}, "SVGF bypass copy");
    timer_end(c);
    return IRT_OK;
  }
  int rc = IRT_OK;
  if (fused) {
    // One timer per kernel; feedback and final copies folded into the
    // a-trous stores.
    static const char* atrous_names[5] = {"svgf_atrous0", "svgf_atrous1", "svgf_atrous2", "svgf_atrous3",
                                          "svgf_atrous4"};
    const GuidePx* guide = c->guide;
    timer_begin(c, "svgf_temporal");
    rc = svgf_temporal_pass<P>(c, guide);
    timer_end(c);
    if (rc) return rc;
    timer_begin(c, "svgf_variance");
    rc = svgf_variance_pass<P>(c, guide);
    timer_end(c);
    if (rc) return rc;
    void* in = c->scratch[0];
    for (uint32_t i = 0; i < iterations; ++i) {
      const bool last = i + 1 == iterations;
      void* out;
      void* out2 = nullptr;
      if (i == 0) {
        out = c->feedback[cur];
        if (last) out2 = c->history;
      } else if (last) {
        out = c->history;
      } else {
        out = in == c->scratch[1] ? c->scratch[0] : c->scratch[1];
      }
      timer_begin(c, atrous_names[i]);
      rc = svgf_atrous_pass<P>(c, i, in, out, out2, guide);
      timer_end(c);
      if (rc) return rc;
      in = out;
    }
    return IRT_OK;
  }
  timer_begin(c, "svgf");
  rc = svgf_temporal_pass<P>(c);
  if (!rc) rc = svgf_variance_pass<P>(c);
  if (rc) return rc;
  for (uint32_t i = 0; i < iterations; ++i) {
    rc = svgf_atrous_pass<P>(c, i, c->scratch[i % 2], c->scratch[1 - i % 2], nullptr);
    if (rc) return rc;
    if (i == 0) {
      HIP_CHECK(c, hipMemcpyAsync(c->feedback[cur], c->scratch[1], rb, hipMemcpyDeviceToDevice, c->stream));
      aud(c, c->stream, {c->scratch[1]}, {c->feedback[cur]}, "SVGF feedback copy");
    }
  }
  HIP_CHECK(c, hipMemcpyAsync(c->history, c->scratch[iterations % 2], rb, hipMemcpyDeviceToDevice, c->stream));
  aud(c, c->stream, {c->scratch[iterations % 2]}, {c->history}, "SVGF history copy");
  timer_end(c);
  return IRT_OK;
}

template <class P>
int run_pass(irt_ctx* c, uint32_t pass, uint32_t arg) {
  int jrc = join(c, c->svgf_stream, c->stream);  // after any SVGF still in flight
  if (jrc) return jrc;
  if (pass < 8) {
    int rc = check_ready(c);
    if (rc) return rc;
    return lighting_pass<P>(c, pass);
  }
  if (!c->have_frame) return c->fail(IRT_ERR_STATE, "no frame set (irt_set_frame)");
  if (pass == IRT_PASS_SVGF_TEMPORAL) return svgf_temporal_pass<P>(c);
  if (pass == IRT_PASS_SVGF_VARIANCE) return svgf_variance_pass<P>(c);
  if (pass == IRT_PASS_SVGF_ATROUS) {
    if (arg > 4) return c->fail(IRT_ERR_INVALID_ARGUMENT, "a-trous iteration must be 0..4");
    int rc = svgf_atrous_pass<P>(c, arg, c->scratch[arg % 2], c->scratch[1 - arg % 2], nullptr);
    if (rc) return rc;
    if (arg == 0) {
      const uint32_t cur = c->frame.size[2] % 2u;
      HIP_CHECK(c, hipMemcpyAsync(c->feedback[cur], c->scratch[1], r4_bytes(c) * c->npx, hipMemcpyDeviceToDevice,
                                  c->stream));
    }
    return IRT_OK;
  }
  return c->fail(IRT_ERR_INVALID_ARGUMENT, "unknown pass id");
}

template <class P>
int ingest(irt_ctx* c, const float* depth, const uint16_t* n, const uint16_t* col, const uint16_t* prm,
           const uint16_t* mot) {
  const uint32_t count = (uint32_t)c->npx;
  hipLaunchKernelGGL((k_ingest_gbuffer<P>), dim3((count + 255) / 256), dim3(256), 0, c->stream, count, depth, n, col,
                     prm, mot, c->depth, (typename P::N4*)c->normal, (typename P::N4*)c->color,
                     (typename P::P2*)c->params, (typename P::P2*)c->motion);
  HIP_CHECK(c, hipGetLastError());
  aud(c, c->stream, {n, col, prm, mot}, {c->depth, c->normal, c->color, c->params, c->motion}, "G-buffer ingest");
  return IRT_OK;
}

}  // namespace

// ====================================================================== ABI
namespace {

// ---------------------------------------------------------------- SDF bake
// irt_bake_sdf and irt_bake_sdf_noalloc share everything but where the
// device scratch comes from: the hierarchy nodes and triangles, the sample
// grid the kernels write before it is copied out, and two words (the
// parity-crossing overflow flag and the bits of the Lipschitz reduction's
// squared maximum).

struct BakePlan {
  std::vector<irt_tri_node> nodes;
  std::vector<irt_tri> tris;
  BakeGrid g{};
  uint64_t samples = 0;
  float sp[3] = {0, 0, 0};
  float threshold = 0;
};

int bake_fail(int code, const std::string& m) {
  irt_host::set_global_error(m);
  return code;
}

// Validation, the host hierarchy and the grid of export_scene.py:135-158.
int bake_prepare(const float* positions, uint32_t stride_floats, uint32_t vertex_count, const uint32_t* indices,
                 uint32_t index_count, uint32_t resolution, const float* samples, float bounds_lo[3],
                 float bounds_hi[3], const float* lipschitz, BakePlan& plan) {
  if (!positions || !indices || !samples || !bounds_lo || !bounds_hi || !lipschitz || stride_floats < 3)
    return bake_fail(IRT_ERR_INVALID_ARGUMENT, "null argument or position stride below 3");
  if (resolution < 4 || resolution > 256) return bake_fail(IRT_ERR_INVALID_ARGUMENT, "SDF resolution must be in [4,256]");
  if (index_count == 0 || vertex_count == 0) return bake_fail(IRT_ERR_INVALID_ARGUMENT, "mesh has no triangles");
  std::vector<irt_vertex> verts(vertex_count);
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (uint32_t v = 0; v < vertex_count; ++v) {
    std::memset(&verts[v], 0, sizeof(irt_vertex));
    for (int a = 0; a < 3; ++a) {
      const float x = positions[(size_t)v * stride_floats + a];
      verts[v].position[a] = x;
      lo[a] = std::min(lo[a], x);
      hi[a] = std::max(hi[a], x);
    }
    verts[v].position[3] = 1.0f;
  }
  std::string msg;
  int rc = irt_host::mesh_hierarchy(verts.data(), vertex_count, indices, index_count, plan.nodes, plan.tris, &msg);
  if (rc) return bake_fail(rc, msg);
  // export_scene.py:136-139: padding = |hi - lo| / (res - 1) * 2 on every axis
  const float e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const double padding = std::sqrt((double)e[0] * e[0] + (double)e[1] * e[1] + (double)e[2] * e[2]) /
                         (double)(resolution - 1) * 2.0;
  const float pf = (float)padding;
  BakeGrid& g = plan.g;
  for (int a = 0; a < 3; ++a) {
    bounds_lo[a] = lo[a] - pf;
    bounds_hi[a] = hi[a] + pf;
    g.lo[a] = bounds_lo[a];
    g.hi[a] = bounds_hi[a];
  }
  const float pe[3] = {bounds_hi[0] - bounds_lo[0], bounds_hi[1] - bounds_lo[1], bounds_hi[2] - bounds_lo[2]};
  g.extent = std::sqrt((double)pe[0] * pe[0] + (double)pe[1] * pe[1] + (double)pe[2] * pe[2]);
  g.res = resolution;
  g.n_nodes = (uint32_t)plan.nodes.size();
  plan.samples = (uint64_t)resolution * resolution * resolution;
  // spacing and threshold of field_lipschitz_bound, in float as the host computes them
  for (int a = 0; a < 3; ++a) plan.sp[a] = (bounds_hi[a] - bounds_lo[a]) / (float)(resolution - 1);
  const float sl = std::sqrt(plan.sp[0] * plan.sp[0] + plan.sp[1] * plan.sp[1] + plan.sp[2] * plan.sp[2]);
  plan.threshold = sl * (0.5f + 1.0f / 128.0f);
  return IRT_OK;
}

// Uploads the hierarchy, clears the two words, runs the bake and the
// Lipschitz reduction on the null stream and copies the samples and words
// back.  Every sample of the grid is written by k_bake_sdf, so the sample
// buffer needs no clearing.
hipError_t bake_run(const BakePlan& plan, irt_tri_node* d_nodes, irt_tri* d_tris, float* d_out, uint32_t* d_flags,
                    float* samples, uint32_t flags[2]) {
  const uint32_t init[2] = {0u, 0x3f800000u};  // overflow flag, bits of the squared maximum (starts at 1)
  hipError_t e1 = hipMemcpy(d_nodes, plan.nodes.data(), sizeof(irt_tri_node) * plan.nodes.size(), hipMemcpyHostToDevice);
  if (e1 == hipSuccess) e1 = hipMemcpy(d_tris, plan.tris.data(), sizeof(irt_tri) * plan.tris.size(), hipMemcpyHostToDevice);
  if (e1 == hipSuccess) e1 = hipMemcpy(d_flags, init, sizeof(init), hipMemcpyHostToDevice);
  const uint64_t n = plan.samples;
  if (e1 == hipSuccess) {
    hipLaunchKernelGGL(k_bake_sdf, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, nullptr, d_nodes, d_tris, plan.g,
                       d_out, d_flags);
    e1 = hipGetLastError();
  }
  const uint64_t r1 = plan.g.res - 1, cells = r1 * r1 * r1;
  if (e1 == hipSuccess) {
    const uint32_t lip_blocks = (uint32_t)std::min<uint64_t>((cells + 255) / 256, 1024);
    hipLaunchKernelGGL(k_lipschitz, dim3(lip_blocks), dim3(256), 0, nullptr, (const float*)d_out, plan.g.res,
                       plan.sp[0], plan.sp[1], plan.sp[2], plan.threshold, d_flags + 1);
    e1 = hipGetLastError();
  }
  if (e1 == hipSuccess) e1 = hipMemcpy(samples, d_out, sizeof(float) * n, hipMemcpyDeviceToHost);
  if (e1 == hipSuccess) e1 = hipMemcpy(flags, d_flags, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost);
  return e1;
}

int bake_finish(hipError_t e1, const uint32_t flags[2], float* lipschitz) {
  if (e1 != hipSuccess) return bake_fail(IRT_ERR_DEVICE, std::string("SDF bake: ") + hipGetErrorString(e1));
  if (flags[0]) return bake_fail(IRT_ERR_INVALID_ARGUMENT, "SDF classification exceeded 256 intersections");
  float max_sq;
  std::memcpy(&max_sq, &flags[1], 4);
  *lipschitz = std::sqrt(max_sq) * 1.0001f;
  return IRT_OK;
}

// Fixed scratch of irt_bake_sdf_noalloc, laid out as
// [flags 256 B | nodes | triangles | samples], each part 256-byte aligned.
// A mesh of T triangles has at most 2T - 1 hierarchy nodes.
constexpr size_t bake_align(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t BAKE_MAX_NODES = 2 * (size_t)IRT_BAKE_NOALLOC_MAX_TRIANGLES;
constexpr size_t BAKE_OFF_FLAGS = 0;
constexpr size_t BAKE_OFF_NODES = 256;
constexpr size_t BAKE_OFF_TRIS = BAKE_OFF_NODES + bake_align(sizeof(irt_tri_node) * BAKE_MAX_NODES);
constexpr size_t BAKE_OFF_SAMPLES = BAKE_OFF_TRIS + bake_align(sizeof(irt_tri) * IRT_BAKE_NOALLOC_MAX_TRIANGLES);
constexpr size_t BAKE_SCRATCH_BYTES =
    BAKE_OFF_SAMPLES + bake_align(sizeof(float) * (size_t)IRT_BAKE_NOALLOC_MAX_RESOLUTION *
                                  IRT_BAKE_NOALLOC_MAX_RESOLUTION * IRT_BAKE_NOALLOC_MAX_RESOLUTION);
static_assert(BAKE_SCRATCH_BYTES == IRT_BAKE_NOALLOC_SCRATCH_BYTES, "irt.h states the scratch size");

constexpr int BAKE_MAX_DEVICES = 64;
std::mutex bake_scratch_lock[BAKE_MAX_DEVICES];
std::mutex bake_symbol_lock;
unsigned char* bake_scratch_ptr[BAKE_MAX_DEVICES] = {};

}  // namespace

// The array lives in the library's code object; every device the library
// runs on holds its own copy until the process ends.  Host code never uses
// this name as an address: the device address comes from
// hipGetSymbolAddress, once per device.
static __device__ __align__(256) unsigned char scratch[BAKE_SCRATCH_BYTES];

namespace {

// The current device's address of `scratch`, looked up on the first call on
// that device with the device current.
int bake_scratch(unsigned char** base, int* device) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return bake_fail(IRT_ERR_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= BAKE_MAX_DEVICES) return bake_fail(IRT_ERR_DEVICE, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lock(bake_symbol_lock);
  if (!bake_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return bake_fail(IRT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    bake_scratch_ptr[dev] = (unsigned char*)ptr;
  }
  *base = bake_scratch_ptr[dev];
  *device = dev;
  return IRT_OK;
}

}  // namespace

extern "C" {

uint32_t irt_abi_version(void) { return IRT_ABI_VERSION; }
const char* irt_global_error(void) { return irt_host::global_error(); }

// Every binary32 input whose top byte is `hi` (one launch per top byte):
// rcp_rn, div_pi_rn and sqrt_rn against the correctly rounded sequences on
// their domains, and div_ranged for that numerator over one divisor of its
// domain.
__global__ __launch_bounds__(256) void k_check_fast_division(uint32_t hi, unsigned long long* mismatches) {
  const uint32_t bits = (hi << 24) | (blockIdx.x * 256u + threadIdx.x);
  const float x = __uint_as_float(bits);
  const float ax = fabsf(x);
  const bool special = !(ax > 0.0f) || ax == __builtin_inff();  // zero, NaN or infinite
  bool bad_rcp = false, bad_pi = false, bad_sqrt = false;
  if (special || (ax >= 0x1p-100f && ax <= 0x1p100f))
    bad_rcp = __float_as_uint(rcp_rn(x)) != __float_as_uint(1.0f / x);
  if (special || (ax >= 0x1p-79f && ax <= 0x1p126f))
    bad_pi = __float_as_uint(div_pi_rn(x)) != __float_as_uint(x / IRT_PI);
  if (x != x || x >= 0x1p-96f)  // NaN, [2^-96, +inf]
    bad_sqrt = __float_as_uint(sqrt_rn(x)) != __float_as_uint(sqrtf(x));
  // div_ranged: numerator x (every binary32 value) over a divisor drawn from
  // the input's hash, exponent in [-100, 99], any significand, either sign;
  // compared where the pair is inside div_ranged's contract
  uint32_t hsh = bits * 0x9e3779b1u;
  hsh ^= hsh >> 15;
  hsh *= 0x85ebca77u;
  hsh ^= hsh >> 13;
  const uint32_t eb = 27u + (hsh >> 24) % 200u;  // biased exponent 27 .. 226
  const float b = __uint_as_float((hsh & 0x80000000u) | (eb << 23) | (hsh & 0x007fffffu));
  const int ea = __builtin_amdgcn_frexp_expf(x), ed = ea - __builtin_amdgcn_frexp_expf(b);
  const bool in_contract = special || (ax >= 0x1p-79f && ed >= -125 && ed <= 126);
  bool bad_div = false;
  if (in_contract)
    bad_div = __float_as_uint(div_ranged(x, b, rcp_refined(b))) != __float_as_uint(x / b);
  if (bad_rcp) atomicAdd(&mismatches[0], 1ull);
  if (bad_pi) atomicAdd(&mismatches[1], 1ull);
  if (bad_sqrt) atomicAdd(&mismatches[2], 1ull);
  if (bad_div) atomicAdd(&mismatches[3], 1ull);
}

int irt_check_fast_division(uint64_t mismatches[4]) {
  if (!mismatches) {
    irt_host::set_global_error("null argument");
    return IRT_ERR_INVALID_ARGUMENT;
  }
  unsigned long long* d = nullptr;
  unsigned long long h[4] = {0ull, 0ull, 0ull, 0ull};
  hipError_t e = hipMalloc(&d, sizeof(h));
  if (e == hipSuccess) e = hipMemcpy(d, h, sizeof(h), hipMemcpyHostToDevice);
  for (uint32_t hi = 0; e == hipSuccess && hi < 256u; ++hi) {
    hipLaunchKernelGGL(k_check_fast_division, dim3(1u << 16), dim3(256), 0, nullptr, hi, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(h, d, sizeof(h), hipMemcpyDeviceToHost);
  if (d) (void)hipFree(d);
  if (e != hipSuccess) {
    irt_host::set_global_error(std::string("irt_check_fast_division: ") + hipGetErrorString(e));
    return IRT_ERR_DEVICE;
  }
  for (int k = 0; k < 4; ++k) mismatches[k] = h[k];
  return IRT_OK;
}

int irt_create(const irt_settings* s, uint32_t precision, irt_ctx** out) {
  if (!s || !out) {
    irt_host::set_global_error("null argument");
    return IRT_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  std::string msg;
  int rc = irt_host::validate_settings(*s, &msg);
  if (rc) {
    irt_host::set_global_error(msg);
    return rc;
  }
  if (precision > IRT_PRECISION_MIXED) {
    irt_host::set_global_error("unknown precision profile");
    return IRT_ERR_INVALID_ARGUMENT;
  }
  const uint64_t pixels = (uint64_t)s->render_width * s->render_height;
  // Adapter-limit checks of renderer.rs:750-766 against a 4 GiB binding.
  const uint64_t max_binding = 1ull << 32;
  if ((uint64_t)(s->restir_max_history + 1) * s->vpl_count * 48 > max_binding) {
    irt_host::set_global_error("VPL history budget exceeds adapter storage binding limit");
    return IRT_ERR_LIMIT;
  }
  if (pixels * 32 > max_binding) {
    irt_host::set_global_error("render size exceeds adapter reservoir binding limit");
    return IRT_ERR_LIMIT;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    irt_host::set_global_error("no HIP device available");
    return IRT_ERR_DEVICE;
  }
  irt_ctx* c = new irt_ctx();
  c->s = *s;
  c->precision = precision;
  const char* audit_env = std::getenv("IRT_STREAM_AUDIT");
  if (audit_env && audit_env[0] == '1') c->audit.reset(new StreamAudit());
  (void)hipGetDevice(&c->device);
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&c->svgf_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_lit, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_frame[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_frame[1], hipEventDisableTiming) != hipSuccess) {
    irt_host::set_global_error("hipStreamCreate / hipEventCreate failed");
    irt_destroy(c);
    return IRT_ERR_DEVICE;
  }
  c->W = s->render_width;
  c->H = s->render_height;
  c->npx = pixels;
  c->V = s->vpl_count;
  c->slots = s->restir_max_history + 1;
  rc = IRT_OK;
  for (const auto& a : view_alloc_plan(c, precision == IRT_PRECISION_MIXED))
    if (!rc) rc = dev_alloc(c, a.first, a.second);
  if (!rc) c->march_blocks = compute_units(c->device) * 8u;
  // Depth targets start cleared to 1 (far plane) as the reference's depth pass clears them.
  if (!rc) {
    bind_slots(c);
    const uint32_t n = (uint32_t)c->npx, one = 0x3f800000u;
    for (int k = 0; k < 3; ++k)
      hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, c->stream, (uint32_t*)c->gb_depth[k], one, n);
    if (hipGetLastError() != hipSuccess) rc = c->fail(IRT_ERR_DEVICE, "initial depth clear failed");
  }
  if (rc) {
    irt_host::set_global_error(c->err);
    irt_destroy(c);
    return rc;
  }
  *out = c;
  return IRT_OK;
}

void irt_destroy(irt_ctx* c) {
  if (!c) return;
  if (c->svgf_stream) (void)hipStreamSynchronize(c->svgf_stream);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& a : c->allocs)
    if (a.first) (void)hipFree(a.first);
  for (auto& t : c->timers) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  if (c->ev_lit) (void)hipEventDestroy(c->ev_lit);
  for (hipEvent_t e : c->ev_frame)
    if (e) (void)hipEventDestroy(e);
  if (c->svgf_stream) (void)hipStreamDestroy(c->svgf_stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* irt_last_error(const irt_ctx* c) { return c ? c->err.c_str() : irt_host::global_error(); }

int irt_set_schedule(irt_ctx* c, uint32_t schedule) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (schedule > IRT_SCHEDULE_FUSED) return c->fail(IRT_ERR_INVALID_ARGUMENT, "unknown schedule");
  c->schedule = schedule;
  return IRT_OK;
}

int irt_set_sdf_atlas(irt_ctx* c, const float* texels, uint32_t w, uint32_t h, uint32_t d) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!texels || w < 2 || h < 2 || d < 2) return c->fail(IRT_ERR_INVALID_ARGUMENT, "invalid SDF atlas");
  const size_t count = (size_t)w * h * d;
  HIP_CHECK(c, sync_stream(c, c->stream));
  for (void** slot : {(void**)&c->atlas, (void**)&c->atlas16}) {
    if (*slot) {
      int rc = dev_free(c, *slot);
      if (rc) return rc;
      *slot = nullptr;
    }
  }
  if (mixed(c)) {
    // 16-bit samples rounded toward -inf: every stored sample is <= the FP32
    // sample, so the conservative bounds built on it stay lower bounds.
    std::vector<uint16_t> h16(count);
    for (size_t i = 0; i < count; ++i) h16[i] = irt_host::f32_to_f16_down(texels[i]);
    int rc = dev_alloc(c, (void**)&c->atlas16, count * 2);
    if (rc) return rc;
    HIP_CHECK(c, hipMemcpyAsync(c->atlas16, h16.data(), count * 2, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, sync_stream(c, c->stream));
  } else {
    int rc = dev_alloc(c, (void**)&c->atlas, count * 4);
    if (rc) return rc;
    HIP_CHECK(c, hipMemcpyAsync(c->atlas, texels, count * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, sync_stream(c, c->stream));
  }
  c->atlas_dims[0] = w;
  c->atlas_dims[1] = h;
  c->atlas_dims[2] = d;
  c->mesh_k_dirty = true;
  c->grid_dirty = true;
  return IRT_OK;
}

int irt_set_meshes(irt_ctx* c, const irt_mesh* meshes, uint32_t count) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!meshes || count == 0) return c->fail(IRT_ERR_INVALID_ARGUMENT, "no meshes");
  for (uint32_t i = 0; i < count; ++i) {
    const irt_mesh& m = meshes[i];
    const uint32_t res = m.field[3];
    if (res < 2) return c->fail(IRT_ERR_INVALID_ARGUMENT, "mesh SDF resolution must be >= 2");
    if (c->atlas && (m.field[0] + res > c->atlas_dims[0] || m.field[1] + res > c->atlas_dims[1] ||
                     m.field[2] + res > c->atlas_dims[2]))
      return c->fail(IRT_ERR_INVALID_ARGUMENT, "mesh SDF tile lies outside the atlas");
  }
  int rc = ensure_array(c, &c->meshes, &c->cap_meshes, count);
  if (rc) return rc;
  HIP_CHECK(c, hipMemcpyAsync(c->meshes, meshes, sizeof(irt_mesh) * count, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(c, sync_stream(c, c->stream));
  c->n_meshes = count;
  c->host_meshes.assign(meshes, meshes + count);
  c->have_geometry = false;  // the triangle hierarchies follow the mesh records
  c->mesh_k_dirty = true;
  c->grid_dirty = true;
  return IRT_OK;
}

int irt_set_objects(irt_ctx* c, const irt_object* objects, uint32_t count, const irt_node* nodes,
                    uint32_t node_count) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!objects || !nodes || count == 0) return c->fail(IRT_ERR_INVALID_ARGUMENT, "no objects");
  if (node_count != 2 * count - 1) return c->fail(IRT_ERR_INVALID_ARGUMENT, "node count must be 2*objects-1");
  uint32_t max_ref = 0, max_material = 0;
  for (uint32_t i = 0; i < count; ++i) {
    if (c->n_meshes && objects[i].info[0] >= c->n_meshes)
      return c->fail(IRT_ERR_INVALID_ARGUMENT, "object references a missing mesh");
    max_ref = std::max(max_ref, objects[i].info[0]);
    max_material = std::max(max_material, objects[i].info[1]);
  }
  for (uint32_t i = 0; i < node_count; ++i) {
    const irt_node& nd = nodes[i];
    if (nd.info[3] > node_count || (nd.info[2] == 1u && nd.info[0] >= count) ||
        (nd.info[2] != 1u && (nd.info[0] >= node_count || nd.info[1] >= node_count)))
      return c->fail(IRT_ERR_INVALID_ARGUMENT, "malformed hierarchy node");
    if (nd.info[3] <= i) return c->fail(IRT_ERR_INVALID_ARGUMENT, "hierarchy escape link must point forward");
  }
  // A host that uploads its objects every frame (renderer.rs:1736-1782) mostly
  // sends the records it sent last time: keep the device copies and the
  // candidate grid built over them.
  if (c->objects && c->host_objects.size() == count && c->host_nodes.size() == node_count &&
      std::memcmp(c->host_objects.data(), objects, sizeof(irt_object) * count) == 0 &&
      std::memcmp(c->host_nodes.data(), nodes, sizeof(irt_node) * node_count) == 0)
    return IRT_OK;
  int rc = ensure_array(c, &c->objects, &c->cap_objects, count);
  if (!rc) rc = ensure_array(c, &c->nodes, &c->cap_nodes, node_count);
  if (rc) return rc;
  c->host_objects.clear();  // stays empty unless the whole upload succeeds
  c->host_nodes.clear();
  HIP_CHECK(c, hipMemcpyAsync(c->objects, objects, sizeof(irt_object) * count, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(c, hipMemcpyAsync(c->nodes, nodes, sizeof(irt_node) * node_count, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(c, sync_stream(c, c->stream));
  c->n_objects = count;
  c->n_nodes = node_count;
  c->max_mesh_ref = max_ref;
  c->max_material_ref = max_material;
  c->host_obj_mesh.resize(count);
  for (uint32_t i = 0; i < count; ++i) c->host_obj_mesh[i] = objects[i].info[0];
  // Object world boxes (the hierarchy's leaves) and the grid over the root box.
  std::vector<float4> boxes(2 * (size_t)count, make_float4(0.f, 0.f, 0.f, 0.f));
  for (uint32_t i = 0; i < node_count; ++i)
    if (nodes[i].info[2] == 1u) {
      const uint32_t j = nodes[i].info[0];
      boxes[2 * j] = make_float4(nodes[i].lo[0], nodes[i].lo[1], nodes[i].lo[2], 0.f);
      boxes[2 * j + 1] = make_float4(nodes[i].hi[0], nodes[i].hi[1], nodes[i].hi[2], 0.f);
    }
  rc = ensure_array(c, &c->obj_boxes, &c->obj_boxes_cap, 2 * count);
  if (rc) return rc;
  HIP_CHECK(c, hipMemcpyAsync(c->obj_boxes, boxes.data(), sizeof(float4) * boxes.size(), hipMemcpyHostToDevice,
                              c->stream));
  HIP_CHECK(c, sync_stream(c, c->stream));
  {
    const float margin = 1e-3f;
    float lo[3], ext = 0.f, e[3];
    for (int a = 0; a < 3; ++a) {
      lo[a] = nodes[0].lo[a] - margin;
      e[a] = nodes[0].hi[a] + margin - lo[a];
      ext = std::max(ext, e[a]);
    }
    const float cell = std::max(ext / GRID_CELLS_LONG_AXIS, 1e-4f);
    c->grid.x0 = lo[0];
    c->grid.y0 = lo[1];
    c->grid.z0 = lo[2];
    c->grid.inv_cell = 1.0f / cell;
    c->grid.nx = std::max(1u, (uint32_t)std::ceil(e[0] / cell));
    c->grid.ny = std::max(1u, (uint32_t)std::ceil(e[1] / cell));
    c->grid.nz = std::max(1u, (uint32_t)std::ceil(e[2] / cell));
  }
  c->grid_dirty = true;
  c->host_objects.assign(objects, objects + count);
  c->host_nodes.assign(nodes, nodes + node_count);
  return IRT_OK;
}

int irt_set_lights(irt_ctx* c, const irt_light* lights, uint32_t count) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (count && !lights) return c->fail(IRT_ERR_INVALID_ARGUMENT, "null lights");
  int rc = ensure_array(c, &c->lights, &c->cap_lights, count);
  if (rc) return rc;
  if (count) {
    HIP_CHECK(c, hipMemcpyAsync(c->lights, lights, sizeof(irt_light) * count, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, sync_stream(c, c->stream));
  }
  c->n_lights = count;
  return IRT_OK;
}

int irt_copy_scene(irt_ctx* dst, irt_ctx* src) {
  if (!dst || !src) return IRT_ERR_INVALID_ARGUMENT;
  if (dst == src) return IRT_OK;
  if (dst->precision != src->precision) return dst->fail(IRT_ERR_INVALID_ARGUMENT, "scene copy across precision profiles");
  if ((!src->atlas && !src->atlas16) || !src->meshes || !src->objects)
    return dst->fail(IRT_ERR_STATE, "source context has no complete scene");
  int rc = ensure_grid(src);
  if (rc) return dst->fail(rc, std::string("source scene: ") + src->err);
  HIP_CHECK(dst, sync_stream(src, src->stream));
  HIP_CHECK(dst, sync_stream(dst, dst->stream));
  const size_t texels = (size_t)src->atlas_dims[0] * src->atlas_dims[1] * src->atlas_dims[2];
  const size_t tb = src->atlas16 ? 2 : 4;
  void** slot = src->atlas16 ? (void**)&dst->atlas16 : (void**)&dst->atlas;
  const void* from = src->atlas16 ? (const void*)src->atlas16 : (const void*)src->atlas;
  const size_t have = (dst->atlas16 || dst->atlas) ? (size_t)dst->atlas_dims[0] * dst->atlas_dims[1] * dst->atlas_dims[2] : 0;
  if (!*slot || have != texels) {
    rc = dev_free(dst, *slot);
    if (!rc) *slot = nullptr;
    if (!rc) rc = dev_alloc(dst, slot, texels * tb);
    if (rc) return rc;
  }
  HIP_CHECK(dst, hipMemcpyAsync(*slot, from, texels * tb, hipMemcpyDeviceToDevice, dst->stream));
  for (int a = 0; a < 3; ++a) dst->atlas_dims[a] = src->atlas_dims[a];
  rc = ensure_array(dst, &dst->meshes, &dst->cap_meshes, src->n_meshes);
  if (!rc) rc = ensure_array(dst, &dst->objects, &dst->cap_objects, src->n_objects);
  if (!rc) rc = ensure_array(dst, &dst->nodes, &dst->cap_nodes, src->n_nodes);
  if (!rc) rc = ensure_array(dst, &dst->lights, &dst->cap_lights, src->n_lights);
  if (!rc) rc = ensure_array(dst, &dst->obj_boxes, &dst->obj_boxes_cap, 2 * src->n_objects);
  if (!rc) rc = ensure_array(dst, &dst->mesh_k2, &dst->mesh_k2_cap, src->n_meshes);
  if (rc) return rc;
  HIP_CHECK(dst, hipMemcpyAsync(dst->meshes, src->meshes, sizeof(irt_mesh) * src->n_meshes, hipMemcpyDeviceToDevice,
                                dst->stream));
  HIP_CHECK(dst, hipMemcpyAsync(dst->objects, src->objects, sizeof(irt_object) * src->n_objects,
                                hipMemcpyDeviceToDevice, dst->stream));
  HIP_CHECK(dst, hipMemcpyAsync(dst->nodes, src->nodes, sizeof(irt_node) * src->n_nodes, hipMemcpyDeviceToDevice,
                                dst->stream));
  if (src->n_lights)
    HIP_CHECK(dst, hipMemcpyAsync(dst->lights, src->lights, sizeof(irt_light) * src->n_lights,
                                  hipMemcpyDeviceToDevice, dst->stream));
  HIP_CHECK(dst, hipMemcpyAsync(dst->obj_boxes, src->obj_boxes, sizeof(float4) * 2 * src->n_objects,
                                hipMemcpyDeviceToDevice, dst->stream));
  HIP_CHECK(dst, hipMemcpyAsync(dst->mesh_k2, src->mesh_k2, sizeof(unsigned int) * src->n_meshes,
                                hipMemcpyDeviceToDevice, dst->stream));
  dst->n_meshes = src->n_meshes;
  dst->n_objects = src->n_objects;
  dst->n_nodes = src->n_nodes;
  dst->n_lights = src->n_lights;
  dst->host_meshes = src->host_meshes;
  dst->max_mesh_ref = src->max_mesh_ref;
  dst->max_material_ref = src->max_material_ref;
  dst->host_obj_mesh = src->host_obj_mesh;
  dst->host_objects = src->host_objects;
  dst->host_nodes = src->host_nodes;
  dst->have_geometry = false;
  if (src->have_geometry) {
    rc = ensure_array(dst, &dst->vertices, &dst->cap_vertices, src->n_vertices);
    if (!rc) rc = ensure_array(dst, &dst->indices, &dst->cap_indices, src->n_indices);
    if (!rc) rc = ensure_array(dst, &dst->materials, &dst->cap_materials, src->n_materials);
    if (!rc) rc = ensure_array(dst, &dst->tri_nodes, &dst->cap_tri_nodes, src->n_tri_nodes);
    if (!rc) rc = ensure_array(dst, &dst->tris, &dst->cap_tris, src->n_tris);
    if (!rc) rc = ensure_array(dst, &dst->mesh_nodes, &dst->cap_mesh_nodes, src->n_meshes);
    if (rc) return rc;
    const std::pair<void*, std::pair<const void*, size_t>> parts[] = {
        {dst->vertices, {src->vertices, sizeof(irt_vertex) * src->n_vertices}},
        {dst->indices, {src->indices, sizeof(uint32_t) * src->n_indices}},
        {dst->materials, {src->materials, sizeof(irt_material) * src->n_materials}},
        {dst->tri_nodes, {src->tri_nodes, sizeof(irt_tri_node) * src->n_tri_nodes}},
        {dst->tris, {src->tris, sizeof(irt_tri) * src->n_tris}},
        {dst->mesh_nodes, {src->mesh_nodes, sizeof(uint2) * src->n_meshes}}};
    for (const auto& p : parts)
      if (p.second.second)
        HIP_CHECK(dst, hipMemcpyAsync(p.first, p.second.first, p.second.second, hipMemcpyDeviceToDevice, dst->stream));
    dst->n_vertices = src->n_vertices;
    dst->n_indices = src->n_indices;
    dst->n_materials = src->n_materials;
    dst->n_tri_nodes = src->n_tri_nodes;
    dst->n_tris = src->n_tris;
    dst->max_texture_layer = src->max_texture_layer;
    dst->have_geometry = true;
  }
  if (src->tex_layers) {
    const size_t texels = (size_t)src->tex_size * src->tex_size * src->tex_layers;
    if (dst->cap_tex < texels) {
      rc = dev_free(dst, dst->tex);
      dst->tex = nullptr;
      dst->cap_tex = 0;
      if (!rc) rc = dev_alloc(dst, (void**)&dst->tex, texels * 4);
      if (rc) return rc;
      dst->cap_tex = texels;
    }
    rc = ensure_array(dst, &dst->srgb_lut, &dst->cap_srgb_lut, 256);
    if (rc) return rc;
    HIP_CHECK(dst, hipMemcpyAsync(dst->tex, src->tex, texels * 4, hipMemcpyDeviceToDevice, dst->stream));
    HIP_CHECK(dst, hipMemcpyAsync(dst->srgb_lut, src->srgb_lut, 256 * sizeof(float), hipMemcpyDeviceToDevice,
                                  dst->stream));
    dst->tex_size = src->tex_size;
    dst->tex_layers = src->tex_layers;
  }
  dst->grid = src->grid;
  dst->grid.cells = nullptr;
  dst->grid.boxes = dst->obj_boxes;
  if (src->grid.cells) {
    const size_t total = (size_t)src->grid.nx * src->grid.ny * src->grid.nz;
    if (dst->cells_cap < total) {
      rc = dev_free(dst, dst->cells);
      dst->cells = nullptr;
      if (!rc) rc = dev_alloc(dst, (void**)&dst->cells, total * sizeof(CellList));
      if (rc) return rc;
      dst->cells_cap = total;
    }
    HIP_CHECK(dst, hipMemcpyAsync(dst->cells, src->cells, total * sizeof(CellList), hipMemcpyDeviceToDevice,
                                  dst->stream));
    dst->grid.cells = dst->cells;
  }
  dst->mesh_k_dirty = src->mesh_k_dirty;
  dst->grid_dirty = false;
  dst->guide_dirty = true;
  return IRT_OK;
}

int irt_reset(irt_ctx* c) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  int rc = sync_all(c);
  if (rc) return rc;
  for (const auto& a : view_alloc_plan(c, mixed(c))) {
    HIP_CHECK(c, hipMemsetAsync(*a.first, 0, alloc_size(a.second), c->stream));
    aud(c, c->stream, {}, {*a.first}, "irt_reset clear");
  }
  const uint32_t n = (uint32_t)c->npx, one = 0x3f800000u;
  for (int k = 0; k < 3; ++k)
    hipLaunchKernelGGL(k_fill_u32, dim3((n + 255) / 256), dim3(256), 0, c->stream, (uint32_t*)c->gb_depth[k], one, n);
  HIP_CHECK(c, hipGetLastError());
  c->gb_cur = c->gb_prev = c->gb_pp = 0;
  c->lit_cur = c->guide_cur = 0;
  c->frames_ended = 0;
  bind_slots(c);
  std::memset(&c->frame, 0, sizeof(c->frame));
  c->have_frame = false;
  c->guide_dirty = true;
  c->counting = false;
  c->n_timers_used = 0;
  return IRT_OK;
}

size_t irt_device_bytes(const irt_ctx* c) { return c ? c->device_bytes : 0; }

size_t irt_scene_bytes(const irt_ctx* c) {
  if (!c) return 0;
  size_t view = 0;
  for (const auto& a : view_alloc_plan(const_cast<irt_ctx*>(c), mixed(c))) view += alloc_size(a.second);
  return c->device_bytes > view ? c->device_bytes - view : 0;
}

size_t irt_estimate_bytes(const irt_settings* s, uint32_t precision) {
  if (!s) return 0;
  irt_ctx tmp;
  tmp.npx = (size_t)s->render_width * s->render_height;
  tmp.V = s->vpl_count;
  tmp.slots = s->restir_max_history + 1;
  size_t total = 0;
  for (const auto& a : view_alloc_plan(&tmp, precision == IRT_PRECISION_MIXED)) total += alloc_size(a.second);
  return total;
}

int irt_set_frame(irt_ctx* c, const irt_frame* f) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!f) return c->fail(IRT_ERR_INVALID_ARGUMENT, "null frame");
  if (f->size[0] != c->W || f->size[1] != c->H) return c->fail(IRT_ERR_INVALID_ARGUMENT, "frame size differs from context size");
  if (f->counts[1] != c->V) return c->fail(IRT_ERR_INVALID_ARGUMENT, "frame VPL count differs from settings");
  if (f->rr[2] != (float)c->s.restir_max_history)
    return c->fail(IRT_ERR_INVALID_ARGUMENT, "frame ReSTIR history differs from settings");
  if (f->counts[3] == 0) return c->fail(IRT_ERR_INVALID_ARGUMENT, "frame must have at least one object");
  c->frame = *f;
  c->have_frame = true;
  c->guide_dirty = true;
  return IRT_OK;
}

int irt_set_gbuffer(irt_ctx* c, const float* depth, const uint16_t* normal, const uint16_t* color,
                    const uint16_t* params, const uint16_t* motion) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!depth || !normal || !color || !params || !motion) return c->fail(IRT_ERR_INVALID_ARGUMENT, "null G-buffer plane");
  int wrc = make_current_writable(c, false);
  if (wrc) return wrc;
  const size_t n = c->npx;
  c->guide_dirty = true;
  HIP_CHECK(c, hipMemcpyAsync(c->depth, depth, n * 4, hipMemcpyHostToDevice, c->stream));
  aud(c, c->stream, {}, {c->depth, c->normal, c->color, c->params, c->motion, c->stage[0], c->stage[1], c->stage[2],
                         c->stage[3]}, "G-buffer upload");
  if (mixed(c)) {
    HIP_CHECK(c, hipMemcpyAsync(c->normal, normal, n * 8, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->color, color, n * 8, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->params, params, n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, hipMemcpyAsync(c->motion, motion, n * 4, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(c, sync_stream(c, c->stream));
    return IRT_OK;
  }
  HIP_CHECK(c, hipMemcpyAsync(c->stage[0], normal, n * 8, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(c, hipMemcpyAsync(c->stage[1], color, n * 8, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(c, hipMemcpyAsync(c->stage[2], params, n * 4, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(c, hipMemcpyAsync(c->stage[3], motion, n * 4, hipMemcpyHostToDevice, c->stream));
  int rc = ingest<ProfileF32>(c, c->depth, c->stage[0], c->stage[1], c->stage[2], c->stage[3]);
  if (rc) return rc;
  HIP_CHECK(c, sync_stream(c, c->stream));
  return IRT_OK;
}

int irt_set_gbuffer_device(irt_ctx* c, const float* depth, const uint16_t* normal, const uint16_t* color,
                           const uint16_t* params, const uint16_t* motion) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!depth || !normal || !color || !params || !motion) return c->fail(IRT_ERR_INVALID_ARGUMENT, "null G-buffer plane");
  int wrc = make_current_writable(c, false);
  if (wrc) return wrc;
  c->guide_dirty = true;
  if (mixed(c)) {
    // the reference formats are this profile's storage: plain copies, one
    // launch for the five planes when every pointer is 16-byte aligned
    const size_t n = c->npx;
    const void* src[5] = {depth, normal, color, params, motion};
    void* dst[5] = {c->depth, c->normal, c->color, c->params, c->motion};
    const size_t bytes[5] = {n * 4, n * 8, n * 8, n * 4, n * 4};
    bool aligned = bytes[1] / 16 <= 0xffffffffull;
    for (int k = 0; k < 5; ++k) aligned = aligned && ((uintptr_t)src[k] % 16 == 0) && ((uintptr_t)dst[k] % 16 == 0);
    if (aligned) {
      PlaneCopy pc;
      pc.planes = 5;
      size_t most = 0;
      for (int k = 0; k < 5; ++k) {
        pc.src[k] = (const uint4*)src[k];
        pc.dst[k] = (uint4*)dst[k];
        pc.units[k] = (uint32_t)(bytes[k] / 16);
        pc.tail[k] = (uint32_t)(bytes[k] % 16 / 4);
        most = std::max(most, bytes[k] / 16);
      }
      const uint32_t blocks = (uint32_t)std::min<size_t>((most + 255) / 256, 8u * compute_units(c->device));
      hipLaunchKernelGGL(k_copy_planes, dim3(std::max(blocks, 1u)), dim3(256), 0, c->stream, pc);
      HIP_CHECK(c, hipGetLastError());
    } else {
      for (int k = 0; k < 5; ++k)
        HIP_CHECK(c, hipMemcpyAsync(dst[k], src[k], bytes[k], hipMemcpyDeviceToDevice, c->stream));
    }
    aud(c, c->stream, {depth, normal, color, params, motion}, {c->depth, c->normal, c->color, c->params, c->motion},
        "G-buffer ingest");
    return IRT_OK;
  }
  return ingest<ProfileF32>(c, depth, normal, color, params, motion);
}

static int set_vpls_impl(irt_ctx* c, const irt_vpl* v, uint32_t count, hipMemcpyKind kind) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!v) return c->fail(IRT_ERR_INVALID_ARGUMENT, "null VPL pool");
  if (count != c->V) return c->fail(IRT_ERR_INVALID_ARGUMENT, "VPL pool size differs from settings.vpl_count");
  if (!c->have_frame) return c->fail(IRT_ERR_STATE, "set the frame before its VPL pool");
  const uint32_t slot = c->frame.size[2] % c->slots;
  HIP_CHECK(c, hipMemcpyAsync(c->ring + (size_t)slot * c->V, v, sizeof(irt_vpl) * c->V, kind, c->stream));
  aud(c, c->stream, {kind == hipMemcpyDeviceToDevice ? (const void*)v : nullptr}, {c->ring}, "VPL pool upload");
  if (kind == hipMemcpyHostToDevice) HIP_CHECK(c, sync_stream(c, c->stream));
  return IRT_OK;
}
int irt_set_vpls(irt_ctx* c, const irt_vpl* v, uint32_t count) { return set_vpls_impl(c, v, count, hipMemcpyHostToDevice); }
int irt_set_vpls_device(irt_ctx* c, const irt_vpl* v, uint32_t count) {
  return set_vpls_impl(c, v, count, hipMemcpyDeviceToDevice);
}

int irt_set_geometry(irt_ctx* c, const irt_vertex* vertices, uint32_t vertex_count, const uint32_t* indices,
                     uint32_t index_count, const irt_material* materials, uint32_t material_count) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if ((vertex_count && !vertices) || (index_count && !indices) || !materials || material_count == 0)
    return c->fail(IRT_ERR_INVALID_ARGUMENT, "null geometry array or no materials");
  if (c->host_meshes.empty()) return c->fail(IRT_ERR_STATE, "set the meshes before their geometry");
  int32_t max_layer = -1;
  for (uint32_t i = 0; i < material_count; ++i) {
    for (int k = 0; k < 4; ++k) max_layer = std::max(max_layer, materials[i].textures[k]);
    if ((materials[i].textures[1] >= 0 && materials[i].channels[1] > 3) ||
        (materials[i].textures[2] >= 0 && materials[i].channels[2] > 3))
      return c->fail(IRT_ERR_INVALID_ARGUMENT, "material texture channel must be 0..3");
  }
  std::vector<irt_tri_node> all_nodes, nodes;
  std::vector<irt_tri> all_tris, tris;
  std::vector<uint2> ranges(c->n_meshes);
  for (uint32_t m = 0; m < c->n_meshes; ++m) {
    const uint32_t* g = c->host_meshes[m].geometry;
    if ((uint64_t)g[0] + g[1] > vertex_count || (uint64_t)g[2] + g[3] > index_count)
      return c->fail(IRT_ERR_INVALID_ARGUMENT, "mesh geometry lies outside the vertex or index arrays");
    std::string msg;
    const int rc = irt_host::mesh_hierarchy(vertices + g[0], g[1], indices + g[2], g[3], nodes, tris, &msg);
    if (rc) return c->fail(rc, msg);
    const uint32_t node_base = (uint32_t)all_nodes.size(), tri_base = (uint32_t)all_tris.size();
    for (irt_tri_node nd : nodes) {
      nd.link += nd.count ? tri_base : node_base;
      all_nodes.push_back(nd);
    }
    all_tris.insert(all_tris.end(), tris.begin(), tris.end());
    ranges[m] = make_uint2(node_base, (uint32_t)all_nodes.size());
  }
  int rc = ensure_array(c, &c->vertices, &c->cap_vertices, vertex_count);
  if (!rc) rc = ensure_array(c, &c->indices, &c->cap_indices, index_count);
  if (!rc) rc = ensure_array(c, &c->materials, &c->cap_materials, material_count);
  if (!rc) rc = ensure_array(c, &c->tri_nodes, &c->cap_tri_nodes, (uint32_t)all_nodes.size());
  if (!rc) rc = ensure_array(c, &c->tris, &c->cap_tris, (uint32_t)all_tris.size());
  if (!rc) rc = ensure_array(c, &c->mesh_nodes, &c->cap_mesh_nodes, c->n_meshes);
  if (rc) return rc;
  if (vertex_count)
    HIP_CHECK(c, hipMemcpyAsync(c->vertices, vertices, sizeof(irt_vertex) * vertex_count, hipMemcpyHostToDevice,
                                c->stream));
  if (index_count)
    HIP_CHECK(c, hipMemcpyAsync(c->indices, indices, sizeof(uint32_t) * index_count, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(c, hipMemcpyAsync(c->materials, materials, sizeof(irt_material) * material_count, hipMemcpyHostToDevice,
                              c->stream));
  if (!all_nodes.empty())
    HIP_CHECK(c, hipMemcpyAsync(c->tri_nodes, all_nodes.data(), sizeof(irt_tri_node) * all_nodes.size(),
                                hipMemcpyHostToDevice, c->stream));
  if (!all_tris.empty())
    HIP_CHECK(c, hipMemcpyAsync(c->tris, all_tris.data(), sizeof(irt_tri) * all_tris.size(), hipMemcpyHostToDevice,
                                c->stream));
  HIP_CHECK(c, hipMemcpyAsync(c->mesh_nodes, ranges.data(), sizeof(uint2) * ranges.size(), hipMemcpyHostToDevice,
                              c->stream));
  HIP_CHECK(c, sync_stream(c, c->stream));
  c->n_vertices = vertex_count;
  c->n_indices = index_count;
  c->n_materials = material_count;
  c->n_tri_nodes = (uint32_t)all_nodes.size();
  c->n_tris = (uint32_t)all_tris.size();
  c->max_texture_layer = max_layer;
  c->have_geometry = true;
  return IRT_OK;
}

static TextureView texture_view(const irt_ctx* c) {
  return TextureView{c->tex, c->srgb_lut, c->tex_size, c->tex_layers};
}

static int check_textures(irt_ctx* c) {
  if (c->max_texture_layer >= 0 && (uint32_t)c->max_texture_layer >= c->tex_layers)
    return c->fail(IRT_ERR_STATE, "a material names texture layer " + std::to_string(c->max_texture_layer) +
                                      " but " + std::to_string(c->tex_layers) + " layers are set (irt_set_textures)");
  return IRT_OK;
}

int irt_set_textures(irt_ctx* c, const uint8_t* rgba8, uint32_t size, uint32_t layers) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!rgba8 || size == 0 || layers == 0) return c->fail(IRT_ERR_INVALID_ARGUMENT, "no texture data");
  if (size > 8192 || (uint64_t)size * size * layers > (1ull << 31))
    return c->fail(IRT_ERR_LIMIT, "texture array too large");
  const size_t texels = (size_t)size * size * layers;
  if (c->cap_tex < texels) {
    HIP_CHECK(c, sync_stream(c, c->stream));
    int rc = dev_free(c, c->tex);
    c->tex = nullptr;
    c->cap_tex = 0;
    if (!rc) rc = dev_alloc(c, (void**)&c->tex, texels * 4);
    if (rc) return rc;
    c->cap_tex = texels;
  }
  int rc = ensure_array(c, &c->srgb_lut, &c->cap_srgb_lut, 256);
  if (rc) return rc;
  float lut[256];
  for (int k = 0; k < 256; ++k) {
    const double v = k / 255.0;
    lut[k] = (float)(v <= 0.04045 ? v / 12.92 : std::pow((v + 0.055) / 1.055, 2.4));
  }
  HIP_CHECK(c, hipMemcpyAsync(c->tex, rgba8, texels * 4, hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(c, hipMemcpyAsync(c->srgb_lut, lut, sizeof(lut), hipMemcpyHostToDevice, c->stream));
  HIP_CHECK(c, sync_stream(c, c->stream));
  c->tex_size = size;
  c->tex_layers = layers;
  return IRT_OK;
}

int irt_generate_vpls(irt_ctx* c) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!c->have_frame) return c->fail(IRT_ERR_STATE, "set the frame before generating its VPL pool");
  if (!c->have_geometry) return c->fail(IRT_ERR_STATE, "no triangle geometry (irt_set_geometry)");
  if (!c->n_objects || !c->n_nodes) return c->fail(IRT_ERR_STATE, "no objects");
  const irt_frame& F = c->frame;
  if (F.counts[1] != c->V) return c->fail(IRT_ERR_INVALID_ARGUMENT, "frame VPL count differs from settings.vpl_count");
  if (F.counts[0] > c->n_lights) return c->fail(IRT_ERR_INVALID_ARGUMENT, "frame light count exceeds the lights set");
  if (F.counts[2] < 1 || F.counts[2] > 32) return c->fail(IRT_ERR_INVALID_ARGUMENT, "invalid VPL bounce count");
  if (c->max_mesh_ref >= c->n_meshes) return c->fail(IRT_ERR_STATE, "object references a missing mesh");
  if (c->max_material_ref >= c->n_materials) return c->fail(IRT_ERR_STATE, "object references a missing material");
  int trc = check_textures(c);
  if (trc) return trc;
  const uint32_t slot = F.size[2] % c->slots;
  TriScene sc{c->objects, c->meshes,  c->mesh_nodes, c->tri_nodes, c->tris,
              c->vertices, c->indices, c->materials,  c->lights,   texture_view(c)};
  // lanes per path: enough for one object per lane up to a wave
  uint32_t group = 8;
  while (group < 64 && group < c->n_objects) group <<= 1;
  const uint32_t blocks = (uint32_t)(((uint64_t)c->V * group + 255) / 256);
  irt_vpl* pool = c->ring + (size_t)slot * c->V;
  timer_begin(c, "vpl_generate");
  switch (group) {
    case 8:
      hipLaunchKernelGGL(k_generate_vpls<8>, dim3(blocks), dim3(256), 0, c->stream, sc, F, c->obj_boxes, c->n_objects,
                         pool);
      break;
    case 16:
      hipLaunchKernelGGL(k_generate_vpls<16>, dim3(blocks), dim3(256), 0, c->stream, sc, F, c->obj_boxes,
                         c->n_objects, pool);
      break;
    case 32:
      hipLaunchKernelGGL(k_generate_vpls<32>, dim3(blocks), dim3(256), 0, c->stream, sc, F, c->obj_boxes,
                         c->n_objects, pool);
      break;
    default:
      hipLaunchKernelGGL(k_generate_vpls<64>, dim3(blocks), dim3(256), 0, c->stream, sc, F, c->obj_boxes,
                         c->n_objects, pool);
  }
  HIP_CHECK(c, hipGetLastError());
  timer_end(c);
  aud(c, c->stream, {}, {c->ring}, "VPL pass");
  return IRT_OK;
}

int irt_render_gbuffer(irt_ctx* c) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!c->have_frame) return c->fail(IRT_ERR_STATE, "set the frame before rendering its G-buffer");
  if (!c->have_geometry) return c->fail(IRT_ERR_STATE, "no triangle geometry (irt_set_geometry)");
  if (!c->n_objects) return c->fail(IRT_ERR_STATE, "no objects");
  if (c->max_mesh_ref >= c->n_meshes) return c->fail(IRT_ERR_STATE, "object references a missing mesh");
  if (c->max_material_ref >= c->n_materials) return c->fail(IRT_ERR_STATE, "object references a missing material");
  int trc = check_textures(c);
  if (trc) return trc;
  // instance triangles: objects in order, each its mesh's triangles in index order (the draw order)
  std::vector<uint32_t> base(c->n_objects + 1, 0);
  for (uint32_t o = 0; o < c->n_objects; ++o) {
    const uint64_t next = (uint64_t)base[o] + c->host_meshes[c->host_obj_mesh[o]].geometry[3] / 3u;
    if (next >= 0xffffffffull) return c->fail(IRT_ERR_LIMIT, "too many instance triangles");
    base[o + 1] = (uint32_t)next;
  }
  const uint32_t n = base[c->n_objects];
  int rc = ensure_array(c, &c->obj_tri_base, &c->cap_obj_tri_base, c->n_objects + 1);
  if (!rc) rc = ensure_array(c, &c->gb_objs, &c->cap_gb_objs, c->n_objects);
  if (!rc) rc = ensure_array(c, &c->gb_tris, &c->cap_gb_tris, n);
  if (!rc) rc = ensure_array(c, &c->gb_attrs, &c->cap_gb_attrs, n);
  if (!rc) rc = ensure_array(c, &c->gb_tiles, &c->cap_gb_tiles, n);
  if (!rc) rc = ensure_array(c, &c->gb_scan, &c->cap_gb_scan, n);
  unsigned long long* vis_before = c->gb_vis;
  if (!rc) rc = ensure_array(c, &c->gb_vis, &c->cap_gb_vis, (uint32_t)c->npx);
  if (rc) return rc;
  if (c->gb_vis != vis_before) c->gb_vis_clear = false;
  const uint32_t scan_blocks = (n + SCAN_BLOCK_ITEMS - 1) / SCAN_BLOCK_ITEMS;
  rc = ensure_array(c, &c->gb_scan_sums, &c->cap_gb_scan_sums, scan_blocks);
  if (rc) return rc;
  if (base != c->host_tri_base || c->cap_obj_tri_base == 0) {
    HIP_CHECK(c, hipMemcpyAsync(c->obj_tri_base, base.data(), sizeof(uint32_t) * base.size(), hipMemcpyHostToDevice,
                                c->stream));
    HIP_CHECK(c, sync_stream(c, c->stream));
    c->host_tri_base = base;
  }
  rc = make_current_writable(c, false);
  if (rc) return rc;
  c->guide_dirty = true;
  const uint32_t W = c->W, H = c->H, npx = (uint32_t)c->npx;
  GbScene sc{c->objects,      c->meshes,    c->vertices,       c->indices,
             c->materials,    c->obj_tri_base, c->n_objects,   texture_view(c)};
  timer_begin(c, "gbuffer");
  hipLaunchKernelGGL(k_gb_objects, dim3((c->n_objects + 63) / 64), dim3(64), 0, c->stream, c->objects, c->n_objects,
                     c->frame, c->gb_objs);
  if (!c->gb_vis_clear) {
    hipLaunchKernelGGL(k_gb_clear, dim3((npx + 255) / 256), dim3(256), 0, c->stream, c->gb_vis, npx);
    c->gb_vis_clear = true;
  }
  if (n) {
    hipLaunchKernelGGL(k_gb_setup, dim3((n + 255) / 256), dim3(256), 0, c->stream, sc, c->gb_objs, n, W, H,
                       c->gb_tris, c->gb_attrs, c->gb_tiles);
    hipLaunchKernelGGL(k_scan_sums, dim3(scan_blocks), dim3(256), 0, c->stream, c->gb_tiles, n, c->gb_scan_sums);
    hipLaunchKernelGGL(k_scan_carries, dim3(1), dim3(256), 0, c->stream, c->gb_scan_sums, scan_blocks);
    hipLaunchKernelGGL(k_scan_blocks, dim3(scan_blocks), dim3(256), 0, c->stream, c->gb_tiles, n, c->gb_scan_sums,
                       c->gb_scan);
    hipLaunchKernelGGL(k_gb_raster, dim3(c->march_blocks), dim3(256), 0, c->stream, c->gb_tris, c->gb_scan, n, W, H,
                       c->gb_vis);
  }
  const dim3 tiles((W + GB_TILE - 1) / GB_TILE, (H + GB_TILE - 1) / GB_TILE);
  if (mixed(c))
    hipLaunchKernelGGL(k_gb_resolve<ProfileMixed>, tiles, dim3(256), 0, c->stream, c->gb_tris, c->gb_attrs,
                       c->materials, texture_view(c), c->gb_vis, W, H, c->depth, (h4*)c->normal, (h4*)c->color, (h2*)c->params,
                       (h2*)c->motion);
  else
    hipLaunchKernelGGL(k_gb_resolve<ProfileF32>, tiles, dim3(256), 0, c->stream, c->gb_tris, c->gb_attrs,
                       c->materials, texture_view(c), c->gb_vis, W, H, c->depth, (float4*)c->normal, (float4*)c->color,
                       (float2*)c->params, (float2*)c->motion);
  HIP_CHECK(c, hipGetLastError());
  aud(c, c->stream, {}, {c->depth, c->normal, c->color, c->params, c->motion}, "G-buffer pass");
  timer_end(c);
  return IRT_OK;
}

int irt_present_device(irt_ctx* c, void* rgba8) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!rgba8) return c->fail(IRT_ERR_INVALID_ARGUMENT, "null display buffer");
  const bool fresh = !c->srgb_t;
  int rc = ensure_array(c, &c->srgb_t, &c->cap_srgb_t, SRGB_THRESHOLDS);
  if (rc) return rc;
  if (fresh) {
    float t[SRGB_THRESHOLDS];
    irt_host::srgb_thresholds(t);
    HIP_CHECK(c, hipMemcpyAsync(c->srgb_t, t, sizeof(t), hipMemcpyHostToDevice, c->stream));
    aud(c, c->stream, {}, {c->srgb_t}, "sRGB threshold upload");
    HIP_CHECK(c, sync_stream(c, c->stream));
  }
  // after the frame's SVGF, on the stream that runs it
  hipStream_t s = c->schedule == IRT_SCHEDULE_FUSED ? c->svgf_stream : c->stream;
  aud(c, s, {c->history, c->srgb_t}, {rgba8}, "presentation");
  const uint32_t n = (uint32_t)c->npx;
  if (mixed(c))
    hipLaunchKernelGGL(k_present<ProfileMixed>, dim3((n + 255) / 256), dim3(256), 0, s, (const h4*)c->history, n,
                       c->srgb_t, (uchar4*)rgba8);
  else
    hipLaunchKernelGGL(k_present<ProfileF32>, dim3((n + 255) / 256), dim3(256), 0, s, (const float4*)c->history, n,
                       c->srgb_t, (uchar4*)rgba8);
  HIP_CHECK(c, hipGetLastError());
  return IRT_OK;
}

int irt_present(irt_ctx* c, uint8_t* rgba8) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!rgba8) return c->fail(IRT_ERR_INVALID_ARGUMENT, "null display buffer");
  int rc = ensure_array(c, &c->present_buf, &c->cap_present_buf, (uint32_t)c->npx);
  if (!rc) rc = irt_present_device(c, c->present_buf);
  if (!rc) rc = sync_all(c);
  if (rc) return rc;
  HIP_CHECK(c, hipMemcpy(rgba8, c->present_buf, c->npx * 4, hipMemcpyDeviceToHost));
  return IRT_OK;
}

int irt_bake_sdf(const float* positions, uint32_t stride_floats, uint32_t vertex_count, const uint32_t* indices,
                 uint32_t index_count, uint32_t resolution, float* samples, float bounds_lo[3], float bounds_hi[3],
                 float* lipschitz) {
  BakePlan plan;
  int rc = bake_prepare(positions, stride_floats, vertex_count, indices, index_count, resolution, samples, bounds_lo,
                        bounds_hi, lipschitz, plan);
  if (rc) return rc;
  void *d_nodes = nullptr, *d_tris = nullptr, *d_out = nullptr, *d_flags = nullptr;
  hipError_t e1 = hipMalloc(&d_nodes, sizeof(irt_tri_node) * plan.nodes.size());
  if (e1 == hipSuccess) e1 = hipMalloc(&d_tris, sizeof(irt_tri) * plan.tris.size());
  if (e1 == hipSuccess) e1 = hipMalloc(&d_out, sizeof(float) * plan.samples);
  if (e1 == hipSuccess) e1 = hipMalloc(&d_flags, 2 * sizeof(uint32_t));
  uint32_t flags[2] = {0, 0};
  if (e1 == hipSuccess)
    e1 = bake_run(plan, (irt_tri_node*)d_nodes, (irt_tri*)d_tris, (float*)d_out, (uint32_t*)d_flags, samples, flags);
  for (void* p : {d_nodes, d_tris, d_out, d_flags})
    if (p) (void)hipFree(p);
  return bake_finish(e1, flags, lipschitz);
}

int irt_bake_sdf_noalloc(const float* positions, uint32_t stride_floats, uint32_t vertex_count,
                         const uint32_t* indices, uint32_t index_count, uint32_t resolution, float* samples,
                         float bounds_lo[3], float bounds_hi[3], float* lipschitz) {
  if (resolution > IRT_BAKE_NOALLOC_MAX_RESOLUTION || index_count / 3 > IRT_BAKE_NOALLOC_MAX_TRIANGLES)
    return bake_fail(IRT_ERR_INVALID_ARGUMENT,
                     "irt_bake_sdf_noalloc: resolution above " + std::to_string(IRT_BAKE_NOALLOC_MAX_RESOLUTION) +
                         " or more than " + std::to_string(IRT_BAKE_NOALLOC_MAX_TRIANGLES) +
                         " triangles exceeds its fixed scratch; use irt_bake_sdf");
  BakePlan plan;
  int rc = bake_prepare(positions, stride_floats, vertex_count, indices, index_count, resolution, samples, bounds_lo,
                        bounds_hi, lipschitz, plan);
  if (rc) return rc;
  if (plan.nodes.size() > BAKE_MAX_NODES || plan.tris.size() > IRT_BAKE_NOALLOC_MAX_TRIANGLES)
    return bake_fail(IRT_ERR_INVALID_ARGUMENT, "irt_bake_sdf_noalloc: triangle hierarchy exceeds its fixed scratch");
  unsigned char* base = nullptr;
  int dev = 0;
  rc = bake_scratch(&base, &dev);
  if (rc) return rc;
  // One scratch array per device: calls on the same device take turns.
  std::lock_guard<std::mutex> lock(bake_scratch_lock[dev]);
  uint32_t flags[2] = {0, 0};
  const hipError_t e1 = bake_run(plan, (irt_tri_node*)(base + BAKE_OFF_NODES), (irt_tri*)(base + BAKE_OFF_TRIS),
                                 (float*)(base + BAKE_OFF_SAMPLES), (uint32_t*)(base + BAKE_OFF_FLAGS), samples, flags);
  return bake_finish(e1, flags, lipschitz);
}

int irt_audit_report(irt_ctx* c, char* out, size_t capacity) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (!c->audit)
    return c->fail(IRT_ERR_STATE, c->audit_dropped ? "stream audit stopped: the context joined a batch"
                                                   : "stream audit is off (set IRT_STREAM_AUDIT=1 before irt_create)");
  const auto& v = c->audit->violations();
  if (out && capacity) {
    std::string text;
    for (const auto& line : v) text += line + "\n";
    text += "audited " + std::to_string(c->audit->operations(0)) + " operations on the lighting stream and " +
            std::to_string(c->audit->operations(1)) + " on the SVGF stream\n";
    const size_t n = std::min(text.size(), capacity - 1);
    std::memcpy(out, text.data(), n);
    out[n] = '\0';
  }
  return (int)std::min<size_t>(v.size(), 0x7fffffff);
}

int irt_host_audit_selftest(void) {
  // Scripted schedules with known answers; returns 0, or the number of the
  // first case the audit judges wrongly.
  int x[8];
  const void* A = &x[0];
  const void* B = &x[1];
  const void* C = &x[2];
  const void* D = &x[3];
  const void* E = &x[4];
  const void* F = &x[5];
  const void* G = &x[6];
  struct Case {
    int expect;
    void (*run)(StreamAudit&, const void* const*);
  };
  const void* bufs[7] = {A, B, C, D, E, F, G};
  const Case cases[] = {
      // 1: a write on stream 0 read on stream 1 without a join
      {1, [](StreamAudit& a, const void* const* b) { a.op(0, {}, {b[0]}, "w"); a.op(1, {b[0]}, {}, "r"); }},
      // 2: the same with a join
      {0, [](StreamAudit& a, const void* const* b) {
         a.op(0, {}, {b[0]}, "w");
         a.join(0, 1);
         a.op(1, {b[0]}, {}, "r");
       }},
      // 3: a read on stream 1 overwritten on stream 0 without waiting for it
      {1, [](StreamAudit& a, const void* const* b) { a.op(1, {b[2]}, {}, "r"); a.op(0, {}, {b[2]}, "w"); }},
      // 4: the same after an event recorded behind the read and waited for
      {0, [](StreamAudit& a, const void* const* b) {
         a.op(1, {b[3]}, {}, "r");
         const StreamAudit::Snapshot s = a.snapshot(1);
         a.wait(0, s, 1);
         a.op(0, {}, {b[1]}, "w other");
         a.op(0, {}, {b[3]}, "w");
       }},
      // 4b: a second read enqueued after the record is not covered by the wait
      {1, [](StreamAudit& a, const void* const* b) {
         a.op(1, {b[3]}, {}, "r");
         const StreamAudit::Snapshot s = a.snapshot(1);
         a.op(1, {b[3]}, {}, "r2");
         a.wait(0, s, 1);
         a.op(0, {}, {b[3]}, "w");
       }},
      // 5: an event recorded before the read does not order the overwrite
      {1, [](StreamAudit& a, const void* const* b) {
         const StreamAudit::Snapshot s = a.snapshot(1);
         a.op(1, {b[3]}, {}, "r");
         a.wait(0, s, 1);
         a.op(0, {}, {b[3]}, "w");
       }},
      // 6: a host wait for stream 0 orders its write before stream 1
      {0, [](StreamAudit& a, const void* const* b) {
         a.op(0, {}, {b[4]}, "w");
         a.host_sync(0);
         a.op(1, {b[4]}, {}, "r");
       }},
      // 7: a zero-fill on stream 0 and a store on stream 1 into the same buffer
      {1, [](StreamAudit& a, const void* const* b) {
         a.op(0, {}, {b[5]}, "fill");
         a.op(1, {b[6]}, {b[5]}, "store");
       }},
      // 8: ordering carried through a third party: 0 -> 1 joined, 1 -> 0 joined
      {0, [](StreamAudit& a, const void* const* b) {
         a.op(0, {}, {b[0]}, "w0");
         a.join(0, 1);
         a.op(1, {b[0]}, {b[1]}, "rw1");
         a.join(1, 0);
         a.op(0, {b[1]}, {b[0]}, "rw0");
       }},
  };
  int k = 0;
  for (const Case& cs : cases) {
    ++k;
    StreamAudit a;
    cs.run(a, bufs);
    if ((int)a.violations().size() != cs.expect) return k;
  }
  return 0;
}

int irt_lighting(irt_ctx* c) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  return mixed(c) ? run_lighting<ProfileMixed>(c) : run_lighting<ProfileF32>(c);
}

int irt_svgf(irt_ctx* c, uint32_t iterations, uint32_t debug) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  return mixed(c) ? run_svgf<ProfileMixed>(c, iterations, debug) : run_svgf<ProfileF32>(c, iterations, debug);
}

int irt_end_frame(irt_ctx* c) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  return end_frame_slots(c);
}

int irt_encode(irt_ctx* c) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  int rc = irt_lighting(c);
  if (!rc) rc = irt_svgf(c, c->s.svgf_iterations, c->frame.flags[2]);
  if (!rc) rc = irt_end_frame(c);
  return rc;
}

int irt_pass(irt_ctx* c, uint32_t pass, uint32_t arg) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (pass >= IRT_PASS_COUNT) return c->fail(IRT_ERR_INVALID_ARGUMENT, "unknown pass id");
  return mixed(c) ? run_pass<ProfileMixed>(c, pass, arg) : run_pass<ProfileF32>(c, pass, arg);
}

size_t irt_buffer_bytes(const irt_ctx* c, uint32_t b) {
  if (!c || b >= IRT_BUF_COUNT) return 0;
  BufDesc d = describe(const_cast<irt_ctx*>(c), b);
  return d.count * external_bytes(d.fmt);
}

int irt_read(irt_ctx* c, uint32_t b, void* dst, size_t bytes) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (b >= IRT_BUF_COUNT || !dst) return c->fail(IRT_ERR_INVALID_ARGUMENT, "bad buffer id or null destination");
  int rrc = sync_all(c);
  if (rrc) return rrc;
  BufDesc d = describe(c, b);
  if (bytes != d.count * external_bytes(d.fmt)) return c->fail(IRT_ERR_INVALID_ARGUMENT, "read size mismatch");
  std::vector<uint8_t> tmp(d.count * internal_bytes(d.fmt));
  HIP_CHECK(c, hipMemcpyAsync(tmp.data(), d.ptr, tmp.size(), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(c, sync_stream(c, c->stream));
  to_external(d.fmt, tmp.data(), (uint8_t*)dst, d.count);
  return IRT_OK;
}

int irt_write(irt_ctx* c, uint32_t b, const void* src, size_t bytes) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  if (b >= IRT_BUF_COUNT || !src) return c->fail(IRT_ERR_INVALID_ARGUMENT, "bad buffer id or null source");
  int wrc = sync_all(c);
  if (!wrc && (b == IRT_BUF_DEPTH || b == IRT_BUF_NORMAL || b == IRT_BUF_COLOR || b == IRT_BUF_PARAMS ||
               b == IRT_BUF_MOTION || b == IRT_BUF_PREVIOUS_DEPTH || b == IRT_BUF_PREVIOUS_NORMAL))
    wrc = make_current_writable(c, true);  // the previous frame's planes stay as they were
  if (wrc) return wrc;
  BufDesc d = describe(c, b);
  if (bytes != d.count * external_bytes(d.fmt)) return c->fail(IRT_ERR_INVALID_ARGUMENT, "write size mismatch");
  std::vector<uint8_t> tmp(d.count * internal_bytes(d.fmt));
  to_internal(d.fmt, (const uint8_t*)src, tmp.data(), d.count);
  if (b == IRT_BUF_DEPTH || b == IRT_BUF_NORMAL || b == IRT_BUF_COLOR || b == IRT_BUF_PARAMS) c->guide_dirty = true;
  HIP_CHECK(c, hipMemcpyAsync(d.ptr, tmp.data(), tmp.size(), hipMemcpyHostToDevice, c->stream));
  aud(c, c->stream, {}, {d.ptr}, "irt_write");
  HIP_CHECK(c, sync_stream(c, c->stream));
  return IRT_OK;
}

int irt_sync(irt_ctx* c) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  int rc = sync_all(c);
  if (rc) return rc;
  return IRT_OK;
}

void* irt_stream(irt_ctx* c) { return c ? (void*)c->stream : nullptr; }
void* irt_stream_svgf(irt_ctx* c) { return c ? (void*)c->svgf_stream : nullptr; }

int irt_enable_timing(irt_ctx* c, int enable) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  c->timing = enable != 0;
  c->timing_only.clear();
  c->n_timers_used = 0;
  return IRT_OK;
}

int irt_enable_timing_for(irt_ctx* c, const char* pass) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  c->timing = true;
  c->timing_only = pass ? pass : "";
  c->n_timers_used = 0;
  return IRT_OK;
}

int irt_pass_times(irt_ctx* c, char names[][32], float* ms, int capacity) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  int rc = sync_all(c);
  if (rc) return rc;
  int n = 0;
  for (size_t i = 0; i < c->n_timers_used && n < capacity; ++i, ++n) {
    float t = 0.f;
    HIP_CHECK(c, hipEventElapsedTime(&t, c->timers[i].a, c->timers[i].b));
    if (names) {
      std::snprintf(names[n], 32, "%s", c->timers[i].name.c_str());
    }
    if (ms) ms[n] = t;
  }
  c->n_timers_used = 0;
  return n;
}

int irt_enable_counters(irt_ctx* c, int enable) {
  if (!c) return IRT_ERR_INVALID_ARGUMENT;
  int rc = join(c, c->svgf_stream, c->stream);
  if (rc) return rc;
  c->counting = enable != 0;
  HIP_CHECK(c, hipMemsetAsync(c->counters, 0, 8 * sizeof(unsigned long long), c->stream));
  return IRT_OK;
}

int irt_read_counters(irt_ctx* c, uint64_t counts[IRT_COUNTER_COUNT]) {
  if (!c || !counts) return IRT_ERR_INVALID_ARGUMENT;
  unsigned long long tmp[8];
  int rc = join(c, c->svgf_stream, c->stream);
  if (rc) return rc;
  HIP_CHECK(c, hipMemcpyAsync(tmp, c->counters, sizeof(tmp), hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(c, sync_stream(c, c->stream));
  for (int i = 0; i < IRT_COUNTER_COUNT; ++i) counts[i] = tmp[i];
  return IRT_OK;
}

}  // extern "C"

#include "irt_batch_impl.h"
#include "irt_server_impl.h"
