// Batched frames: one launch per pass covers the tiles of many views.
//
// Included at the end of irt_api.hip (it uses that translation unit's
// context helpers).  Views keep their own device state; a batch holds, per
// view, the pass arguments the fused schedule would use (uploaded each frame
// from double-buffered pinned memory) and tile tables (view, tile x, tile y)
// for three tile shapes.  Pixel passes are flat launches over the tile table
// (the dispatcher balances tiles across CUs); the march is one persistent
// kernel that walks the views in order, stages each view's scene in LDS and
// drains that view's request slots with the per-wave refill of k_march.
// The launch sequence can be captured once in a HIP graph and replayed while
// the membership stays the same: the kernels read everything that changes
// per frame from device memory.

namespace {

enum BatchKind {
  BK_GUIDE = 0,
  BK_RES_A = 1,
  BK_SETUP = 2,
  BK_COMBINE = 3,
  BK_SVGF_T = 4,
  BK_SVGF_V = 5,
  BK_SVGF_A = 6,
  BK_END = 7,
  BK_COUNT = 8
};
constexpr int kShapes[3][2] = {{8, 8}, {16, 8}, {16, 16}};
constexpr uint32_t kMaxBatchViews = 255;  // 8-bit view index in a tile entry
constexpr uint32_t kMaxTileCoord = 4096;  // 12-bit tile coordinates

template <class P>
struct BatchView {
  LightArgs<P> la;  // reservoirs A: in 2, out 1, history 2
  LightArgs<P> lb;  // spatial set-up / combine: in 1, out 2
  SvgfArgs<P> st, sv, sa[5];
  irt_reservoir* direct0;
  irt_reservoir* indirect0;
  ResolveRec* rec;
  RayReq* reqs;
  float* vis;
  GuidePx* guide;
  uint32_t W, H, iterations, pad;
};

template <int K, class P, int TW, int TH>
__global__ __launch_bounds__(TW* TH) void k_batch(const BatchView<P>* __restrict__ views,
                                                  const uint32_t* __restrict__ tiles, uint32_t iter) {
  const uint32_t e = __builtin_amdgcn_readfirstlane(tiles[blockIdx.x]);
  const uint32_t v = e >> 24, tx = (e >> 12) & 0xfffu, ty = e & 0xfffu;
  const BatchView<P>& bv = views[v];
  if constexpr (K == BK_SVGF_A) {
    if (iter >= bv.iterations) return;  // this view has finished its a-trous iterations
  }
  Px p;
  p.x = (int32_t)(tx * TW + threadIdx.x % TW);
  p.y = (int32_t)(ty * TH + threadIdx.x / TW);
  p.active = (uint32_t)p.x < bv.W && (uint32_t)p.y < bv.H;
  p.id = p.active ? (uint32_t)p.y * bv.W + (uint32_t)p.x : 0u;
  Count<false> cnt;
  if constexpr (K == BK_GUIDE) {
    if (p.active)
      body_guide<P>(bv.la.F, bv.la.depth, bv.la.normal, bv.la.color, bv.la.params, bv.guide, p.x, p.y, p.id);
  } else if constexpr (K == BK_RES_A) {
    body_reservoirs_a(bv.la, p, bv.direct0, bv.indirect0, cnt);
  } else if constexpr (K == BK_SETUP) {
    body_spatial_setup(bv.lb, p, bv.rec, bv.reqs, cnt);
  } else if constexpr (K == BK_COMBINE) {
    body_resolve_combine(bv.lb, p, bv.rec, bv.vis);
  } else if constexpr (K == BK_SVGF_T) {
    body_svgf_temporal<P, Count<false>, true>(bv.st, p, cnt);
  } else if constexpr (K == BK_SVGF_V) {
    body_svgf_variance<P, Count<false>, true>(bv.sv, p, cnt);
  } else if constexpr (K == BK_SVGF_A) {
    body_svgf_atrous<P, Count<false>, true>(bv.sa[iter], p, cnt);
  }
}

// Batched a-trous iteration `iter` with LDS-staged taps: tile entries are
// 16s x 16s pixel patches (s = 2^iter) of a view, placed with their lattice
// cosets by atrous_slot; the body is k_svgf_atrous_lds's.
template <class P>
__global__ __launch_bounds__(256) void k_batch_atrous(const BatchView<P>* __restrict__ views,
                                                     const uint32_t* __restrict__ tiles, uint32_t n_regions,
                                                     uint32_t iter) {
  __shared__ uint4 s_g[AT_SIDE * AT_SIDE][2];
  __shared__ typename P::R4 s_r[AT_SIDE * AT_SIDE];
  uint32_t region, coset;
  if (!atrous_slot(blockIdx.x, n_regions, 1u << (2 * iter), &region, &coset)) return;
  const uint32_t e = __builtin_amdgcn_readfirstlane(tiles[region]);
  const BatchView<P>& bv = views[e >> 24];
  if (iter >= bv.iterations) return;  // this view has finished its a-trous iterations
  const SvgfArgs<P>& a = bv.sa[iter];
  const int32_t W = (int32_t)bv.W, H = (int32_t)bv.H, st = (int32_t)a.stride;
  const int32_t cx = (int32_t)coset % st, cy = (int32_t)coset / st;
  const int32_t x0 = (int32_t)((e >> 12) & 0xfffu) * AT_TILE * st + cx;
  const int32_t y0 = (int32_t)(e & 0xfffu) * AT_TILE * st + cy;
  for (int32_t k = (int32_t)threadIdx.x; k < AT_SIDE * AT_SIDE; k += 256) {
    const int32_t gx = x0 + (k % AT_SIDE - AT_HALO) * st, gy = y0 + (k / AT_SIDE - AT_HALO) * st;
    if (gx >= 0 && gy >= 0 && gx < W && gy < H) {
      const uint32_t q = (uint32_t)gy * (uint32_t)W + (uint32_t)gx;
      const uint4* gp = reinterpret_cast<const uint4*>(a.guide + q);
      s_g[k][0] = gp[0];
      s_g[k][1] = gp[1];
      s_r[k] = a.input[q];
    } else {
      s_g[k][0] = make_uint4(0u, 0u, 0u, 0u);
      s_g[k][1] = make_uint4(0u, 0u, 0u, 0u);
      s_r[k] = typename P::R4{};
    }
  }
  __syncthreads();
  const int32_t i = (int32_t)threadIdx.x % AT_TILE, j = (int32_t)threadIdx.x / AT_TILE;
  const int32_t x = x0 + i * st, y = y0 + j * st;
  Count<false> cnt;
  if (x < W && y < H) {
    const uint32_t id = (uint32_t)y * (uint32_t)W + (uint32_t)x;
    const float4 r = svgf_atrous_staged<P>(a, x, y, s_g, s_r, i + AT_HALO, j + AT_HALO, AT_SIDE, st == 1, cnt);
    st_r16(a.out, id, r);
    if (a.out2) st_r16(a.out2, id, r);
  }
}

// Batched temporal accumulation: the 16 x 16 tile table, LDS-staged clamp
// windows; the body is k_svgf_temporal_lds's.
template <class P>
__global__ __launch_bounds__(256) void k_batch_temporal(const BatchView<P>* __restrict__ views,
                                                       const uint32_t* __restrict__ tiles) {
  __shared__ uint4 s_g[TMP_SIDE * TMP_SIDE][2];
  __shared__ typename P::R4 s_r[TMP_SIDE * TMP_SIDE];
  const uint32_t e = __builtin_amdgcn_readfirstlane(tiles[blockIdx.x]);
  const BatchView<P>& bv = views[e >> 24];
  Count<false> cnt;
  temporal_tile(bv.st, (int32_t)((e >> 12) & 0xfffu) * TMP_TILE, (int32_t)(e & 0xfffu) * TMP_TILE, s_g, s_r, cnt);
}

// Batched variance estimate: the 16 x 16 tile table, LDS-staged 7 x 7
// neighbourhoods; the body is k_svgf_variance_lds's.
template <class P>
__global__ __launch_bounds__(256) void k_batch_variance(const BatchView<P>* __restrict__ views,
                                                       const uint32_t* __restrict__ tiles) {
  __shared__ uint4 s_g[VAR_SIDE * VAR_SIDE][2];
  __shared__ float2 s_m[VAR_SIDE * VAR_SIDE];
  const uint32_t e = __builtin_amdgcn_readfirstlane(tiles[blockIdx.x]);
  const BatchView<P>& bv = views[e >> 24];
  Count<false> cnt;
  variance_tile(bv.sv, (int32_t)((e >> 12) & 0xfffu) * VAR_TILE, (int32_t)(e & 0xfffu) * VAR_TILE, s_g, s_m, cnt);
}

__global__ __launch_bounds__(64) void k_zero_queues(const MarchView* __restrict__ views, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n * 16u) views[i / 16u].queue[i % 16u] = 0u;
}

}  // namespace

struct irt_batch {
  std::vector<irt_ctx*> views;
  std::vector<std::array<uint32_t, 3>> view_shape;  // W, H, SVGF iterations at set_views
  uint32_t precision = 0;
  int device = 0;
  hipStream_t stream = nullptr;       // lighting passes
  hipStream_t svgf_stream = nullptr;  // SVGF passes (overlap the next step's lighting)
  size_t view_size = 0;       // sizeof(BatchView<P>) of the precision profile
  uint32_t cap_views = 0;
  // device arguments by step parity: a step's SVGF still reads its own while
  // the next step uploads
  void* d_views[2] = {nullptr, nullptr};
  void* h_views[2] = {nullptr, nullptr};
  MarchView* d_march[2] = {nullptr, nullptr};
  MarchView* h_march[2] = {nullptr, nullptr};
  uint64_t step = 0;
  hipEvent_t ev_lit = nullptr;                   // a step's lighting done
  hipEvent_t ev_svgf[2] = {nullptr, nullptr};    // a step's SVGF done, by step parity
  hipEvent_t staged[2] = {nullptr, nullptr};
  bool staged_used[2] = {false, false};
  int flip = 0;
  // tile tables of the three shapes, concatenated
  uint32_t* d_tiles = nullptr;
  size_t cap_tiles = 0;
  uint32_t* h_tiles[2] = {nullptr, nullptr};
  size_t cap_h_tiles[2] = {0, 0};
  hipEvent_t tiles_staged[2] = {nullptr, nullptr};
  bool tiles_used[2] = {false, false};
  int tflip = 0;
  // tables 0-2: row-major tiles of shape 0-2; tables 3-5: the same tiles in
  // column bands per XCD (the spatial set-up's order, tile_bands)
  uint32_t tile_off[6] = {0, 0, 0, 0, 0, 0};
  uint32_t n_tiles[6] = {0, 0, 0, 0, 0, 0};
  uint32_t at_off[5] = {0, 0, 0, 0, 0};  // a-trous patch tables, iteration i: 16*2^i pixel patches
  uint32_t at_n[5] = {0, 0, 0, 0, 0};
  uint32_t shape[BK_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint32_t max_iterations = 0;
  uint32_t march_blocks = 0;
  bool use_graph = false;
  hipGraphExec_t exec_l[2] = {nullptr, nullptr};  // lighting launches, by step parity
  hipGraphExec_t exec_s[2] = {nullptr, nullptr};  // SVGF launches, by step parity
  hipEvent_t done = nullptr;
  hipEvent_t view_ready = nullptr;
  bool timing = false;
  hipEvent_t tev[BK_COUNT + 1][2];
  float last_ms[BK_COUNT + 1] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  std::string err;
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

namespace {

#define BHIP(b, expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return (b)->fail(IRT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class P>
void fill_view(irt_ctx* c, BatchView<P>& bv, MarchView& mv) {
  std::memset(&bv, 0, sizeof(bv));
  const uint32_t cur = c->frame.size[2] % 2u, old = 1u - cur;
  bv.la = light_args<P>(c, 2, 1, 2);
  bv.la.guide = c->guide;
  bv.la.sc.grid = c->grid;
  bv.lb = light_args<P>(c, 1, 2, 0);
  bv.lb.guide = c->guide;
  bv.lb.sc.grid = c->grid;
  bv.direct0 = c->direct[0];
  bv.indirect0 = c->indirect[0];
  bv.rec = c->resolve_rec;
  bv.reqs = c->rays;
  bv.vis = c->vis;
  bv.guide = c->guide;
  bv.W = c->W;
  bv.H = c->H;
  bv.iterations = c->s.svgf_iterations;
  SvgfArgs<P> base = svgf_args<P>(c);
  base.guide = c->guide;
  bv.st = base;
  bv.st.old_radiance = (const typename P::R4*)c->feedback[old];
  bv.st.old_moments = (const typename P::M4*)c->moments[old];
  bv.st.input = (const typename P::R4*)c->feedback[old];
  bv.st.moments = (const typename P::M4*)c->moments[old];
  bv.st.out = (typename P::R4*)c->temporal;
  bv.st.moments_out = (typename P::M4*)c->moments[cur];
  bv.sv = base;
  bv.sv.input = (const typename P::R4*)c->temporal;
  bv.sv.moments = (const typename P::M4*)c->moments[cur];
  bv.sv.out = (typename P::R4*)c->scratch[0];
  // a-trous with the feedback / final copies folded into the stores (run_svgf, fused)
  void* in = c->scratch[0];
  const uint32_t it = c->s.svgf_iterations;
  for (uint32_t i = 0; i < it; ++i) {
    const bool last = i + 1 == it;
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
    bv.sa[i] = base;
    bv.sa[i].input = (const typename P::R4*)in;
    bv.sa[i].moments = (const typename P::M4*)c->moments[cur];
    bv.sa[i].out = (typename P::R4*)out;
    bv.sa[i].out2 = (typename P::R4*)out2;
    bv.sa[i].stride = 1u << i;
    in = out;
  }
  mv.sc = bv.lb.sc;
  mv.F = c->frame;
  mv.reqs = c->rays;
  mv.queue = c->queue;
  mv.vis = c->vis;
  mv.n_meshes = c->n_meshes;
  mv.n_rays = (uint32_t)(2 * c->npx);
}

template <int K, class P>
int launch_kind(irt_batch* b, int q, hipStream_t st, uint32_t iter) {
  const uint32_t sh = b->shape[K];
  const uint32_t table = K == BK_SETUP ? 3 + sh : sh;  // the set-up gathers neighbours: banded tiles
  const BatchView<P>* views = (const BatchView<P>*)b->d_views[q];
  const uint32_t n = b->n_tiles[table];
  const uint32_t* tiles = b->d_tiles + b->tile_off[table];
  if (n == 0) return IRT_OK;
  if (sh == 0)
    hipLaunchKernelGGL((k_batch<K, P, 8, 8>), dim3(n), dim3(64), 0, st, views, tiles, iter);
  else if (sh == 1)
    hipLaunchKernelGGL((k_batch<K, P, 16, 8>), dim3(n), dim3(128), 0, st, views, tiles, iter);
  else
    hipLaunchKernelGGL((k_batch<K, P, 16, 16>), dim3(n), dim3(256), 0, st, views, tiles, iter);
  BHIP(b, hipGetLastError());
  return IRT_OK;
}

void bmark(irt_batch* b, hipStream_t st, int slot, int end) {
  if (b->timing) (void)hipEventRecord(b->tev[slot][end], st);
}

// The lighting launches of one batched frame (b->stream; what graph
// exec_l[q] holds).
template <class P>
int batch_lighting(irt_batch* b, int q) {
  const uint32_t nv = (uint32_t)b->views.size();
  hipStream_t st = b->stream;
  int rc = IRT_OK;
  bmark(b, st, BK_GUIDE, 0);
  rc = launch_kind<BK_GUIDE, P>(b, q, st, 0);
  bmark(b, st, BK_GUIDE, 1);
  if (!rc) {
    bmark(b, st, BK_RES_A, 0);
    rc = launch_kind<BK_RES_A, P>(b, q, st, 0);
    bmark(b, st, BK_RES_A, 1);
  }
  if (!rc) {
    bmark(b, st, BK_SETUP, 0);
    hipLaunchKernelGGL(k_zero_queues, dim3((nv * 16 + 63) / 64), dim3(64), 0, st, b->d_march[q], nv);
    rc = launch_kind<BK_SETUP, P>(b, q, st, 0);
    bmark(b, st, BK_SETUP, 1);
  }
  if (!rc) {
    bmark(b, st, BK_COUNT, 0);  // march
    if (b->precision == IRT_PRECISION_MIXED) {
      BHIP(b, launch_march_views_contracted(dim3(b->march_blocks), st, b->d_march[q], nv));
    } else {
      hipLaunchKernelGGL((k_march_views<IRT_MARCH_EXACT>), dim3(b->march_blocks), dim3(256), 0, st, b->d_march[q], nv);
      BHIP(b, hipGetLastError());
    }
    bmark(b, st, BK_COUNT, 1);
    bmark(b, st, BK_COMBINE, 0);
    rc = launch_kind<BK_COMBINE, P>(b, q, st, 0);
    bmark(b, st, BK_COMBINE, 1);
  }
  return rc;
}

// The SVGF launches of one batched frame (b->svgf_stream; graph exec_s[q]).
template <class P>
int batch_svgf(irt_batch* b, int q) {
  hipStream_t st = b->svgf_stream;
  int rc = IRT_OK;
  bmark(b, st, BK_SVGF_T, 0);
  // LDS-staged clamp windows need the 16 x 16 tiles (table of shape 2)
  if (b->n_tiles[2]) {
    hipLaunchKernelGGL((k_batch_temporal<P>), dim3(b->n_tiles[2]), dim3(256), 0, st,
                       (const BatchView<P>*)b->d_views[q], b->d_tiles + b->tile_off[2]);
    BHIP(b, hipGetLastError());
  }
  bmark(b, st, BK_SVGF_T, 1);
  if (!rc) {
    bmark(b, st, BK_SVGF_V, 0);
    // LDS-staged neighbourhoods need the 16 x 16 tiles (table of shape 2)
    if (b->n_tiles[2]) {
      hipLaunchKernelGGL((k_batch_variance<P>), dim3(b->n_tiles[2]), dim3(256), 0, st,
                         (const BatchView<P>*)b->d_views[q], b->d_tiles + b->tile_off[2]);
      BHIP(b, hipGetLastError());
    }
    bmark(b, st, BK_SVGF_V, 1);
  }
  bmark(b, st, BK_SVGF_A, 0);
  for (uint32_t i = 0; !rc && i < b->max_iterations; ++i) {
    if (b->at_n[i] == 0) continue;
    hipLaunchKernelGGL((k_batch_atrous<P>), dim3(((b->at_n[i] + 7) / 8) * 8 * (1u << (2 * i))), dim3(256), 0, st,
                       (const BatchView<P>*)b->d_views[q], b->d_tiles + b->at_off[i], b->at_n[i], i);
    BHIP(b, hipGetLastError());
  }
  bmark(b, st, BK_SVGF_A, 1);
  // BK_END: the previous-frame G-buffer is a slot rotation on the host
  bmark(b, st, BK_END, 0);
  bmark(b, st, BK_END, 1);
  return rc;
}

// Captures `launch` on `st` into *exec (first use) and replays it.
template <class F>
int replay(irt_batch* b, hipStream_t st, hipGraphExec_t* exec, F launch) {
  if (!*exec) {
    hipGraph_t graph = nullptr;
    BHIP(b, hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = launch();
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return b->fail(IRT_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return b->fail(IRT_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  BHIP(b, hipGraphLaunch(*exec, st));
  return IRT_OK;
}

void drop_graph(irt_batch* b) {
  for (int q = 0; q < 2; ++q) {
    if (b->exec_l[q]) (void)hipGraphExecDestroy(b->exec_l[q]);
    if (b->exec_s[q]) (void)hipGraphExecDestroy(b->exec_s[q]);
    b->exec_l[q] = b->exec_s[q] = nullptr;
  }
}

template <class P>
int batch_encode(irt_batch* b) {
  const uint32_t nv = (uint32_t)b->views.size();
  if (nv == 0) return b->fail(IRT_ERR_STATE, "batch has no views");
  // per-view preconditions and lazy structures on the views' own streams
  for (irt_ctx* c : b->views) {
    int rc = check_ready(c);
    if (!rc) rc = ensure_grid(c);
    if (rc) return b->fail(rc, std::string("view: ") + c->err);
    if (!c->grid.cells || c->n_objects > LDS_MAX_OBJECTS || c->n_meshes > LDS_MAX_OBJECTS)
      return b->fail(IRT_ERR_LIMIT, "batched views need scenes of at most 64 objects and meshes");
    if (c->frame.flags[2] >= 3u) return b->fail(IRT_ERR_STATE, "batched views render debug modes < 3");
    // this step writes the view's other guide and noisy-radiance slots and a
    // G-buffer slot: wait for the SVGF (batched or not) of the frame before
    // the previous one, their last reader
    rc = wait_for_slots(c);
    if (rc) return b->fail(rc, std::string("view: ") + c->err);
    // the step's SVGF follows the view's earlier SVGF work (moments and
    // feedback ping-pong); its lighting follows the view's ingest.  The
    // previous step's SVGF is not waited for: it overlaps this step's lighting.
    BHIP(b, hipEventRecord(b->view_ready, c->svgf_stream));
    BHIP(b, hipStreamWaitEvent(b->svgf_stream, b->view_ready, 0));
    BHIP(b, hipEventRecord(b->view_ready, c->stream));
    BHIP(b, hipStreamWaitEvent(b->stream, b->view_ready, 0));
  }
  // The argument buffers of this step's parity: the SVGF of two steps back
  // read them.
  const int q = (int)(b->step % 2);
  BHIP(b, hipStreamWaitEvent(b->stream, b->ev_svgf[q], 0));
  // per-frame arguments into the free pinned buffer, then to the device
  const int f = b->flip;
  if (b->staged_used[f]) BHIP(b, hipEventSynchronize(b->staged[f]));
  BatchView<P>* hv = (BatchView<P>*)b->h_views[f];
  for (uint32_t i = 0; i < nv; ++i) {
    irt_ctx* c = b->views[i];
    c->lit_cur ^= 1;
    c->guide_cur ^= 1;
    bind_slots(c);
    fill_view<P>(c, hv[i], b->h_march[f][i]);
  }
  BHIP(b, hipMemcpyAsync(b->d_views[q], hv, sizeof(BatchView<P>) * nv, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipMemcpyAsync(b->d_march[q], b->h_march[f], sizeof(MarchView) * nv, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipEventRecord(b->staged[f], b->stream));
  b->staged_used[f] = true;
  b->flip = 1 - f;
  int rc = IRT_OK;
  const bool graph = b->use_graph && !b->timing;
  if (graph)
    rc = replay(b, b->stream, &b->exec_l[q], [&] { return batch_lighting<P>(b, q); });
  else
    rc = batch_lighting<P>(b, q);
  if (rc) return rc;
  BHIP(b, hipEventRecord(b->ev_lit, b->stream));
  BHIP(b, hipStreamWaitEvent(b->svgf_stream, b->ev_lit, 0));
  if (graph)
    rc = replay(b, b->svgf_stream, &b->exec_s[q], [&] { return batch_svgf<P>(b, q); });
  else
    rc = batch_svgf<P>(b, q);
  if (rc) return rc;
  BHIP(b, hipEventRecord(b->ev_svgf[q], b->svgf_stream));
  BHIP(b, hipEventRecord(b->done, b->svgf_stream));
  for (irt_ctx* c : b->views) {
    // the view's next ingest and VPL upload wait for this step's lighting,
    // its frame end for this step's SVGF
    BHIP(b, hipStreamWaitEvent(c->stream, b->ev_lit, 0));
    BHIP(b, hipStreamWaitEvent(c->svgf_stream, b->done, 0));
    c->guide_dirty = false;  // the batch built this frame's guide records
    int rc2 = end_frame_slots(c);  // the frame's G-buffer becomes the previous one
    if (rc2) return b->fail(rc2, std::string("view: ") + c->err);
  }
  b->step++;
  return IRT_OK;
}

// Replaces the batch's views: argument buffers grow as needed and the tile
// tables are rebuilt and uploaded in stream order (no host wait unless a
// device buffer has to grow).
int set_views(irt_batch* b, irt_ctx* const* views, uint32_t count) {
  if (count == 0 || count > kMaxBatchViews) return b->fail(IRT_ERR_INVALID_ARGUMENT, "batch needs 1..255 views");
  uint32_t max_it = 0;
  for (uint32_t i = 0; i < count; ++i) {
    const irt_ctx* c = views[i];
    if (!c) return b->fail(IRT_ERR_INVALID_ARGUMENT, "null view");
    if (c->precision != b->precision)
      return b->fail(IRT_ERR_INVALID_ARGUMENT, "batched views must share a precision profile");
    if (c->device != b->device) return b->fail(IRT_ERR_INVALID_ARGUMENT, "batched views must share a device");
    if ((c->W + 7) / 8 > kMaxTileCoord || (c->H + 7) / 8 > kMaxTileCoord)
      return b->fail(IRT_ERR_LIMIT, "view too large for the tile table");
    for (uint32_t j = 0; j < i; ++j)
      if (views[j] == c) return b->fail(IRT_ERR_INVALID_ARGUMENT, "a view appears twice in the batch");
    max_it = std::max(max_it, c->s.svgf_iterations);
  }
  if (count > b->cap_views) {
    BHIP(b, hipStreamSynchronize(b->svgf_stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    for (int q = 0; q < 2; ++q) {
      if (b->d_views[q]) BHIP(b, hipFree(b->d_views[q]));
      if (b->d_march[q]) BHIP(b, hipFree(b->d_march[q]));
      b->d_views[q] = nullptr;
      b->d_march[q] = nullptr;
    }
    for (int f = 0; f < 2; ++f) {
      if (b->h_views[f]) BHIP(b, hipHostFree(b->h_views[f]));
      if (b->h_march[f]) BHIP(b, hipHostFree(b->h_march[f]));
      b->h_views[f] = nullptr;
      b->h_march[f] = nullptr;
      b->staged_used[f] = false;
    }
    const uint32_t cap = std::max(count, std::min<uint32_t>(kMaxBatchViews, 2 * b->cap_views));
    for (int q = 0; q < 2; ++q) {
      BHIP(b, hipMalloc(&b->d_views[q], b->view_size * cap));
      BHIP(b, hipMalloc((void**)&b->d_march[q], sizeof(MarchView) * cap));
    }
    for (int f = 0; f < 2; ++f) {
      BHIP(b, hipHostMalloc(&b->h_views[f], b->view_size * cap, hipHostMallocDefault));
      BHIP(b, hipHostMalloc((void**)&b->h_march[f], sizeof(MarchView) * cap, hipHostMallocDefault));
    }
    b->cap_views = cap;
  }
  // tile tables: table t occupies [tile_off[t], tile_off[t] + n_tiles[t]),
  // then the a-trous patch tables of iterations 0..4
  size_t total = 0;
  uint32_t n[6], na[5];
  for (int t = 0; t < 6; ++t) {
    const uint32_t tw = kShapes[t % 3][0], th = kShapes[t % 3][1];
    size_t k = 0;
    for (uint32_t v = 0; v < count; ++v) {
      const uint32_t nx = (views[v]->W + tw - 1) / tw, ny = (views[v]->H + th - 1) / th;
      if (t >= 3 && XCDS * band_width_tiles(nx) > kMaxTileCoord)
        return b->fail(IRT_ERR_LIMIT, "view too wide for the banded tile table");
      k += (size_t)(t < 3 ? nx : XCDS * band_width_tiles(nx)) * ny;
    }
    if (k > 0x7fffffffull) return b->fail(IRT_ERR_LIMIT, "too many tiles in one batch");
    n[t] = (uint32_t)k;
    total += k;
  }
  for (uint32_t i = 0; i < 5; ++i) {
    const uint32_t span = (uint32_t)AT_TILE << i;
    size_t k = 0;
    for (uint32_t v = 0; v < count; ++v)
      if (i < views[v]->s.svgf_iterations) k += (size_t)((views[v]->W + span - 1) / span) * ((views[v]->H + span - 1) / span);
    na[i] = (uint32_t)k;
    total += k;
  }
  const int f = b->tflip;
  if (b->tiles_used[f]) BHIP(b, hipEventSynchronize(b->tiles_staged[f]));
  if (b->cap_h_tiles[f] < total) {
    if (b->h_tiles[f]) BHIP(b, hipHostFree(b->h_tiles[f]));
    b->h_tiles[f] = nullptr;
    BHIP(b, hipHostMalloc((void**)&b->h_tiles[f], total * 4, hipHostMallocDefault));
    b->cap_h_tiles[f] = total;
  }
  if (b->cap_tiles < total) {
    BHIP(b, hipStreamSynchronize(b->svgf_stream));
    BHIP(b, hipStreamSynchronize(b->stream));
    if (b->d_tiles) BHIP(b, hipFree(b->d_tiles));
    b->d_tiles = nullptr;
    BHIP(b, hipMalloc((void**)&b->d_tiles, total * 4));
    b->cap_tiles = total;
  }
  uint32_t* h = b->h_tiles[f];
  // the a-trous launches of the last step (SVGF stream) still read the tables
  BHIP(b, hipStreamWaitEvent(b->stream, b->done, 0));
  size_t k = 0;
  for (int t = 0; t < 6; ++t) {
    b->tile_off[t] = (uint32_t)k;
    const uint32_t tw = kShapes[t % 3][0], th = kShapes[t % 3][1];
    for (uint32_t v = 0; v < count; ++v) {
      const uint32_t nx = (views[v]->W + tw - 1) / tw, ny = (views[v]->H + th - 1) / th;
      if (t < 3) {
        for (uint32_t ty = 0; ty < ny; ++ty)
          for (uint32_t tx = 0; tx < nx; ++tx) h[k++] = (v << 24) | (tx << 12) | ty;
      } else {
        // entry 8j + x of the view's block (the launch deals it to XCD x, as
        // every view's block starts at a multiple of 8): tile j of band x, the
        // band's columns [x*bw, (x+1)*bw) walked row by row; columns past the
        // view's last are empty tiles whose pixels all fall outside it
        const uint32_t bw = band_width_tiles(nx);
        for (uint32_t j = 0; j < bw * ny; ++j)
          for (uint32_t x = 0; x < XCDS; ++x) h[k++] = (v << 24) | ((x * bw + j % bw) << 12) | (j / bw);
      }
    }
    b->n_tiles[t] = n[t];
  }
  for (uint32_t i = 0; i < 5; ++i) {
    b->at_off[i] = (uint32_t)k;
    const uint32_t span = (uint32_t)AT_TILE << i;
    for (uint32_t v = 0; v < count; ++v) {
      if (i >= views[v]->s.svgf_iterations) continue;
      const uint32_t nx = (views[v]->W + span - 1) / span, ny = (views[v]->H + span - 1) / span;
      for (uint32_t ty = 0; ty < ny; ++ty)
        for (uint32_t tx = 0; tx < nx; ++tx) h[k++] = (v << 24) | (tx << 12) | ty;
    }
    b->at_n[i] = na[i];
  }
  BHIP(b, hipMemcpyAsync(b->d_tiles, h, total * 4, hipMemcpyHostToDevice, b->stream));
  BHIP(b, hipEventRecord(b->tiles_staged[f], b->stream));
  b->tiles_used[f] = true;
  b->tflip = 1 - f;
  b->views.assign(views, views + count);
  // The stream audit models a context's own two streams; batched steps run
  // on the batch's streams, so a context in a batch leaves the audit.
  for (irt_ctx* v : b->views)
    if (v->audit) {
      v->audit.reset();
      v->audit_dropped = true;
    }
  b->view_shape.resize(count);
  for (uint32_t v = 0; v < count; ++v)
    b->view_shape[v] = {views[v]->W, views[v]->H, views[v]->s.svgf_iterations};
  b->max_iterations = max_it;
  drop_graph(b);
  return IRT_OK;
}

// True when `views` are the batch's members with the sizes and SVGF
// iteration counts its tile tables were built for (a destroyed context's
// address can come back as a new context of another size).
bool same_views(const irt_batch* b, irt_ctx* const* views, uint32_t count) {
  if (b->views.size() != count) return false;
  for (uint32_t i = 0; i < count; ++i) {
    const irt_ctx* c = views[i];
    if (b->views[i] != c || b->view_shape[i][0] != c->W || b->view_shape[i][1] != c->H ||
        b->view_shape[i][2] != c->s.svgf_iterations)
      return false;
  }
  return true;
}

}  // namespace

extern "C" {

int irt_batch_create(irt_ctx* const* views, uint32_t count, irt_batch** out) {
  if (!views || !out || count == 0 || count > kMaxBatchViews || !views[0]) {
    irt_host::set_global_error("batch needs 1..255 views");
    return IRT_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  irt_batch* b = new irt_batch();
  for (auto& pair : b->tev)
    for (auto& e : pair) e = nullptr;
  b->precision = views[0]->precision;
  b->device = views[0]->device;
  b->view_size = b->precision == IRT_PRECISION_MIXED ? sizeof(BatchView<ProfileMixed>) : sizeof(BatchView<ProfileF32>);
  b->march_blocks = views[0]->march_blocks;
  auto bail = [&](int code, const std::string& m) {
    irt_host::set_global_error(m);
    irt_batch_destroy(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->svgf_stream, hipStreamNonBlocking) != hipSuccess)
    return bail(IRT_ERR_DEVICE, "hipStreamCreate failed");
  for (int f = 0; f < 2; ++f)
    if (hipEventCreateWithFlags(&b->staged[f], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->tiles_staged[f], hipEventDisableTiming) != hipSuccess)
      return bail(IRT_ERR_DEVICE, "hipEventCreate failed");
  if (hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->view_ready, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_lit, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_svgf[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_svgf[1], hipEventDisableTiming) != hipSuccess)
    return bail(IRT_ERR_DEVICE, "hipEventCreate failed");
  for (auto& pair : b->tev)
    for (auto& e : pair)
      if (hipEventCreate(&e) != hipSuccess) return bail(IRT_ERR_DEVICE, "hipEventCreate failed");
  int rc = set_views(b, views, count);
  if (rc) return bail(rc, b->err);
  *out = b;
  return IRT_OK;
}

void irt_batch_destroy(irt_batch* b) {
  if (!b) return;
  if (b->svgf_stream) (void)hipStreamSynchronize(b->svgf_stream);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_graph(b);
  if (b->d_tiles) (void)hipFree(b->d_tiles);
  for (int f = 0; f < 2; ++f) {
    if (b->d_views[f]) (void)hipFree(b->d_views[f]);
    if (b->d_march[f]) (void)hipFree(b->d_march[f]);
    if (b->ev_svgf[f]) (void)hipEventDestroy(b->ev_svgf[f]);
    if (b->h_views[f]) (void)hipHostFree(b->h_views[f]);
    if (b->h_march[f]) (void)hipHostFree(b->h_march[f]);
    if (b->h_tiles[f]) (void)hipHostFree(b->h_tiles[f]);
    if (b->staged[f]) (void)hipEventDestroy(b->staged[f]);
    if (b->tiles_staged[f]) (void)hipEventDestroy(b->tiles_staged[f]);
  }
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->view_ready) (void)hipEventDestroy(b->view_ready);
  if (b->ev_lit) (void)hipEventDestroy(b->ev_lit);
  for (auto& pair : b->tev)
    for (auto& e : pair)
      if (e) (void)hipEventDestroy(e);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  if (b->svgf_stream) (void)hipStreamDestroy(b->svgf_stream);
  delete b;
}

const char* irt_batch_last_error(const irt_batch* b) { return b ? b->err.c_str() : irt_host::global_error(); }

int irt_batch_set_views(irt_batch* b, irt_ctx* const* views, uint32_t count) {
  if (!b) return IRT_ERR_INVALID_ARGUMENT;
  if (!views) return b->fail(IRT_ERR_INVALID_ARGUMENT, "null view list");
  if (same_views(b, views, count)) return IRT_OK;
  return set_views(b, views, count);
}

int irt_batch_encode(irt_batch* b) {
  if (!b) return IRT_ERR_INVALID_ARGUMENT;
  return b->precision == IRT_PRECISION_MIXED ? batch_encode<ProfileMixed>(b) : batch_encode<ProfileF32>(b);
}

int irt_batch_set_tile(irt_batch* b, uint32_t kind, uint32_t shape) {
  if (!b) return IRT_ERR_INVALID_ARGUMENT;
  if (kind >= BK_COUNT || shape > 2) return b->fail(IRT_ERR_INVALID_ARGUMENT, "bad pass kind or tile shape");
  if (b->shape[kind] != shape) drop_graph(b);
  b->shape[kind] = shape;
  return IRT_OK;
}

int irt_batch_use_graph(irt_batch* b, int enable) {
  if (!b) return IRT_ERR_INVALID_ARGUMENT;
  b->use_graph = enable != 0;
  if (!b->use_graph) drop_graph(b);
  return IRT_OK;
}

int irt_batch_sync(irt_batch* b) {
  if (!b) return IRT_ERR_INVALID_ARGUMENT;
  BHIP(b, hipStreamSynchronize(b->svgf_stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  return IRT_OK;
}

void* irt_batch_stream(irt_batch* b) { return b ? (void*)b->stream : nullptr; }

void* irt_batch_stream_svgf(irt_batch* b) { return b ? (void*)b->svgf_stream : nullptr; }

int irt_batch_pass_times(irt_batch* b, float ms[9]) {
  if (!b || !ms) return IRT_ERR_INVALID_ARGUMENT;
  for (int k = 0; k <= BK_COUNT; ++k) ms[k] = b->last_ms[k];
  return BK_COUNT + 1;
}

// Per-pass GPU times of one batched frame (no graph), for profiling.
int irt_batch_encode_timed(irt_batch* b, float ms[9]) {
  if (!b || !ms) return IRT_ERR_INVALID_ARGUMENT;
  b->timing = true;
  int rc = irt_batch_encode(b);
  b->timing = false;
  if (rc) return rc;
  BHIP(b, hipStreamSynchronize(b->svgf_stream));
  BHIP(b, hipStreamSynchronize(b->stream));
  for (int k = 0; k <= BK_COUNT; ++k) {
    float t = 0.f;
    BHIP(b, hipEventElapsedTime(&t, b->tev[k][0], b->tev[k][1]));
    ms[k] = t;
    b->last_ms[k] = t;
  }
  return IRT_OK;
}

// Times one batched frame per tile shape (all pixel passes at that shape,
// `repeats` frames each) and keeps, per pass, the fastest shape.  Each trial
// renders a frame of every view from its current inputs: run it during
// warm-up.
int irt_batch_autotune(irt_batch* b, uint32_t repeats, uint32_t shapes_out[8]) {
  if (!b) return IRT_ERR_INVALID_ARGUMENT;
  if (repeats == 0) repeats = 1;
  float best[BK_COUNT];
  uint32_t pick[BK_COUNT];
  for (int k = 0; k < BK_COUNT; ++k) {
    best[k] = INFINITY;
    pick[k] = 0;
  }
  for (uint32_t s = 0; s < 3; ++s) {
    for (int k = 0; k < BK_COUNT; ++k) b->shape[k] = s;
    float acc[BK_COUNT] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t r = 0; r < repeats; ++r) {
      float ms[BK_COUNT + 1];
      int rc = irt_batch_encode_timed(b, ms);
      if (rc) return rc;
      for (int k = 0; k < BK_COUNT; ++k) acc[k] += ms[k];
    }
    for (int k = 0; k < BK_COUNT; ++k)
      if (acc[k] < best[k]) {
        best[k] = acc[k];
        pick[k] = s;
      }
  }
  for (int k = 0; k < BK_COUNT; ++k) b->shape[k] = pick[k];
  drop_graph(b);
  if (shapes_out)
    for (int k = 0; k < BK_COUNT; ++k) shapes_out[k] = pick[k];
  return IRT_OK;
}

}  // extern "C"
