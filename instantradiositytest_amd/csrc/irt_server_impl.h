// Multi-session frame server: device-resident per-session state, admission
// by slot count and HBM budget, continuous batching of frame requests.
//
// Included at the end of irt_api.hip after the batch implementation.  Each
// session is one context (reservoirs, SVGF history, previous G-buffer, VPL
// ring and its own copy of a scene).  Clients submit frame requests whose
// G-buffer and VPL pool already sit in device memory.  Sessions are dealt
// to `max_in_flight` step groups (the group with the fewest pixels takes a
// new session), each with its own batch; a step takes the oldest pending
// request of every session of one group, the groups taking turns, and
// renders them with one batched frame.  Requests that arrive while steps
// are on the GPU join their group's next step.  Up to `max_in_flight` steps
// are queued on the device at once, and because consecutive steps belong to
// different groups, one group's SVGF overlaps the next group's lighting
// even when every client waits for its previous frame.  A session's frames
// are rendered in submission order and its results do not depend on which
// other sessions share its steps.
//
// The frame bookkeeping of Renderer::encode (renderer.rs:1687-1735 packing,
// 1973-1982 previous_vp / previous_generation / previous_debug /
// frame_index) is kept per session.

#include <algorithm>
#include <chrono>
#include <deque>

namespace {

double server_clock() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct ServerRequest {
  irt_frame frame;
  irt_frame_request in;
  double submit_s;
};

struct ServerSession {
  irt_ctx* ctx = nullptr;
  bool open = false;
  uint32_t frame_index = 0;
  float previous_vp[16];
  bool have_previous = false;
  uint32_t previous_generation = 0;
  uint32_t previous_debug = 0;
  std::deque<ServerRequest> pending;
  uint32_t in_flight = 0;
  uint32_t group = 0;
  uint64_t pixels = 0;
};

struct ServerStep {
  uint64_t seq = 0;
  hipEvent_t start = nullptr, done = nullptr;
  double launch_s = 0;
  std::vector<irt_completion> members;
};

}  // namespace

struct irt_server {
  irt_server_config cfg{};
  std::vector<ServerSession> sessions;
  std::vector<irt_ctx*> idle;  // closed sessions' contexts kept for reuse
  std::vector<irt_batch*> batches;      // one per step group (max_in_flight)
  std::vector<uint64_t> group_pixels;   // pixels of the open sessions of each group
  uint32_t next_group = 0;              // the group whose turn is next
  std::deque<ServerStep> in_flight;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> spare_events;
  std::deque<irt_completion> completed;
  uint64_t next_seq = 1;
  std::string err;
  int fail(int code, const std::string& m) {
    err = m;
    return code;
  }
};

namespace {

size_t server_bytes(const irt_server* s) {
  size_t b = 0;
  for (const auto& x : s->sessions)
    if (x.ctx) b += irt_device_bytes(x.ctx);
  for (irt_ctx* c : s->idle) b += irt_device_bytes(c);
  return b;
}

bool same_settings(const irt_settings& a, const irt_settings& b) { return std::memcmp(&a, &b, sizeof(a)) == 0; }

int retire_front(irt_server* s) {
  ServerStep& st = s->in_flight.front();
  const double now = server_clock();
  float gpu_ms = 0.f;
  (void)hipEventElapsedTime(&gpu_ms, st.start, st.done);
  for (irt_completion c : st.members) {
    c.done_s = now;
    c.gpu_ms = gpu_ms;
    ServerSession& ss = s->sessions[c.session];
    if (ss.in_flight) --ss.in_flight;
    s->completed.push_back(c);
  }
  s->spare_events.emplace_back(st.start, st.done);
  s->in_flight.pop_front();
  return IRT_OK;
}

// Retires finished steps (in order); with `wait`, blocks on the oldest one
// first.
int drain(irt_server* s, bool wait) {
  if (wait && !s->in_flight.empty()) {
    hipError_t e = hipEventSynchronize(s->in_flight.front().done);
    if (e != hipSuccess) return s->fail(IRT_ERR_DEVICE, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
    retire_front(s);
  }
  while (!s->in_flight.empty()) {
    hipError_t e = hipEventQuery(s->in_flight.front().done);
    if (e == hipErrorNotReady) break;
    if (e != hipSuccess) return s->fail(IRT_ERR_DEVICE, std::string("hipEventQuery: ") + hipGetErrorString(e));
    retire_front(s);
  }
  return IRT_OK;
}

// Waits until no step holds a frame of session `id`.
int drain_session(irt_server* s, uint32_t id) {
  while (s->sessions[id].in_flight) {
    int rc = drain(s, true);
    if (rc) return rc;
  }
  return IRT_OK;
}

}  // namespace

extern "C" {

double irt_server_now(void) { return server_clock(); }

int irt_server_create(const irt_server_config* cfg, irt_server** out) {
  if (!cfg || !out) {
    irt_host::set_global_error("null argument");
    return IRT_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  if (cfg->max_sessions == 0 || cfg->max_sessions > 255 || cfg->precision > IRT_PRECISION_MIXED ||
      cfg->max_in_flight == 0 || cfg->max_in_flight > 8) {
    irt_host::set_global_error("server needs 1..255 sessions, a known precision and 1..8 steps in flight");
    return IRT_ERR_INVALID_ARGUMENT;
  }
  irt_server* s = new irt_server();
  s->cfg = *cfg;
  s->sessions.resize(cfg->max_sessions);
  s->batches.assign(cfg->max_in_flight, nullptr);
  s->group_pixels.assign(cfg->max_in_flight, 0);
  *out = s;
  return IRT_OK;
}

void irt_server_destroy(irt_server* s) {
  if (!s) return;
  while (!s->in_flight.empty()) {
    (void)hipEventSynchronize(s->in_flight.front().done);
    retire_front(s);
  }
  for (irt_batch* b : s->batches)
    if (b) irt_batch_destroy(b);
  for (auto& x : s->sessions)
    if (x.ctx) irt_destroy(x.ctx);
  for (irt_ctx* c : s->idle) irt_destroy(c);
  for (auto& e : s->spare_events) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
  delete s;
}

const char* irt_server_last_error(const irt_server* s) { return s ? s->err.c_str() : irt_host::global_error(); }

int irt_server_open(irt_server* s, const irt_settings* settings, irt_ctx* scene, uint32_t* session) {
  if (!s) return IRT_ERR_INVALID_ARGUMENT;
  if (!settings || !scene || !session) return s->fail(IRT_ERR_INVALID_ARGUMENT, "null argument");
  if (scene->precision != s->cfg.precision) return s->fail(IRT_ERR_INVALID_ARGUMENT, "scene precision differs from the server's");
  std::string msg;
  int rc = irt_host::validate_settings(*settings, &msg);
  if (rc) return s->fail(rc, msg);
  uint32_t id = 0;
  while (id < s->sessions.size() && s->sessions[id].open) ++id;
  if (id == s->sessions.size()) return s->fail(IRT_ERR_LIMIT, "no free session slot");
  // A slot freed by a close may still own a context with frames in flight.
  rc = drain_session(s, id);
  if (rc) return rc;
  if (s->sessions[id].ctx) {
    s->idle.push_back(s->sessions[id].ctx);
    s->sessions[id].ctx = nullptr;
  }
  // Reuse an idle context of the same settings, else create one within the budget.
  irt_ctx* c = nullptr;
  for (size_t i = 0; i < s->idle.size(); ++i)
    if (same_settings(s->idle[i]->s, *settings)) {
      c = s->idle[i];
      s->idle.erase(s->idle.begin() + (ptrdiff_t)i);
      break;
    }
  // Bytes the session will hold: its own state plus its scene copy (a
  // reused context keeps its allocations, so count the larger figure).
  const size_t need = irt_estimate_bytes(settings, s->cfg.precision) + irt_scene_bytes(scene);
  const size_t hold = c ? std::max(need, irt_device_bytes(c)) : need;
  size_t used = server_bytes(s);  // excludes c, taken off the idle list
  while (used + hold > s->cfg.hbm_budget_bytes && !s->idle.empty()) {
    irt_destroy(s->idle.back());
    s->idle.pop_back();
    used = server_bytes(s);
  }
  if (used + hold > s->cfg.hbm_budget_bytes) {
    if (c) s->idle.push_back(c);
    return s->fail(IRT_ERR_LIMIT, "HBM budget exhausted");
  }
  if (c) {
    rc = irt_reset(c);
    if (rc) {
      std::string m = c->err;
      irt_destroy(c);
      return s->fail(rc, "reset: " + m);
    }
  } else {
    rc = irt_create(settings, s->cfg.precision, &c);
    if (rc) return s->fail(rc, std::string("create: ") + irt_host::global_error());
  }
  c->schedule = IRT_SCHEDULE_FUSED;
  rc = irt_copy_scene(c, scene);
  if (rc) {
    std::string m = c->err;
    s->idle.push_back(c);
    return s->fail(rc, "scene copy: " + m);
  }
  ServerSession& ss = s->sessions[id];
  ss = ServerSession{};
  ss.ctx = c;
  ss.open = true;
  ss.pixels = (uint64_t)c->W * c->H;
  ss.group = (uint32_t)(std::min_element(s->group_pixels.begin(), s->group_pixels.end()) - s->group_pixels.begin());
  s->group_pixels[ss.group] += ss.pixels;
  *session = id;
  return IRT_OK;
}

irt_ctx* irt_server_context(irt_server* s, uint32_t session) {
  if (!s || session >= s->sessions.size() || !s->sessions[session].open) return nullptr;
  return s->sessions[session].ctx;
}

int irt_server_close(irt_server* s, uint32_t session) {
  if (!s) return IRT_ERR_INVALID_ARGUMENT;
  if (session >= s->sessions.size() || !s->sessions[session].open)
    return s->fail(IRT_ERR_INVALID_ARGUMENT, "no such session");
  ServerSession& ss = s->sessions[session];
  ss.open = false;
  ss.pending.clear();
  s->group_pixels[ss.group] -= ss.pixels;
  // the context stays with the slot until its frames in flight are done
  if (!ss.in_flight) {
    s->idle.push_back(ss.ctx);
    ss.ctx = nullptr;
  }
  return IRT_OK;
}

int irt_server_submit(irt_server* s, uint32_t session, const irt_frame_request* req) {
  if (!s) return IRT_ERR_INVALID_ARGUMENT;
  if (!req) return s->fail(IRT_ERR_INVALID_ARGUMENT, "null request");
  if (session >= s->sessions.size() || !s->sessions[session].open)
    return s->fail(IRT_ERR_INVALID_ARGUMENT, "no such session");
  ServerSession& ss = s->sessions[session];
  irt_ctx* c = ss.ctx;
  // null planes: the G-buffer pass renders them; a null pool: the VPL pass generates it
  const int planes = !!req->depth + !!req->normal + !!req->color + !!req->params + !!req->motion;
  if (planes != 0 && planes != 5)
    return s->fail(IRT_ERR_INVALID_ARGUMENT, "G-buffer planes must be all given or all null");
  if ((planes == 0 || !req->vpls) && !c->have_geometry)
    return s->fail(IRT_ERR_STATE, "the session's scene has no triangle geometry to render its inputs from");
  if (req->vpl_count != c->V) return s->fail(IRT_ERR_INVALID_ARGUMENT, "VPL pool size differs from the session's settings");
  if (req->debug_mode >= 3u) return s->fail(IRT_ERR_INVALID_ARGUMENT, "served sessions render debug modes 0..2");
  if (req->light_count > c->n_lights) return s->fail(IRT_ERR_INVALID_ARGUMENT, "light count exceeds the scene's lights");
  ServerRequest r;
  r.in = *req;
  r.submit_s = server_clock();
  const float* pvp = ss.have_previous ? ss.previous_vp : req->vp;
  const uint32_t history = ss.frame_index > 0 && ss.previous_debug == req->debug_mode ? 1u : 0u;
  int rc = irt_host::pack_frame(c->s, req->vp, pvp, req->camera, ss.frame_index, req->generation,
                                ss.previous_generation, history, req->debug_mode, req->light_count, c->n_objects,
                                &r.frame);
  if (rc) return s->fail(rc, "singular view-projection matrix");
  ss.pending.push_back(r);
  std::memcpy(ss.previous_vp, req->vp, sizeof(ss.previous_vp));
  ss.have_previous = true;
  ss.previous_generation = req->generation;
  ss.previous_debug = req->debug_mode;
  ++ss.frame_index;
  return IRT_OK;
}

int irt_server_step(irt_server* s) {
  if (!s) return IRT_ERR_INVALID_ARGUMENT;
  int rc = drain(s, false);
  if (rc) return rc;
  // contexts of closed sessions whose last frames have finished go idle
  for (auto& ss : s->sessions)
    if (!ss.open && ss.ctx && !ss.in_flight) {
      s->idle.push_back(ss.ctx);
      ss.ctx = nullptr;
    }
  if (s->in_flight.size() >= s->cfg.max_in_flight) return 0;
  // the first group in turn with pending requests
  const uint32_t groups = (uint32_t)s->batches.size();
  std::vector<irt_ctx*> views;
  std::vector<uint32_t> ids;
  uint32_t g = 0;
  for (uint32_t t = 0; t < groups && views.empty(); ++t) {
    g = (s->next_group + t) % groups;
    for (uint32_t i = 0; i < s->sessions.size(); ++i)
      if (s->sessions[i].open && s->sessions[i].group == g && !s->sessions[i].pending.empty()) {
        views.push_back(s->sessions[i].ctx);
        ids.push_back(i);
      }
  }
  if (views.empty()) return 0;
  s->next_group = (g + 1) % groups;
  ServerStep st;
  st.seq = s->next_seq++;
  for (size_t k = 0; k < ids.size(); ++k) {
    ServerSession& ss = s->sessions[ids[k]];
    const ServerRequest& r = ss.pending.front();
    irt_ctx* c = ss.ctx;
    rc = irt_set_frame(c, &r.frame);
    if (!rc)
      rc = r.in.depth ? irt_set_gbuffer_device(c, r.in.depth, r.in.normal, r.in.color, r.in.params, r.in.motion)
                      : irt_render_gbuffer(c);
    if (!rc) rc = r.in.vpls ? irt_set_vpls_device(c, r.in.vpls, r.in.vpl_count) : irt_generate_vpls(c);
    if (rc) return s->fail(rc, "session " + std::to_string(ids[k]) + ": " + c->err);
    irt_completion cm{};
    cm.tag = r.in.tag;
    cm.session = ids[k];
    cm.frame_index = r.frame.size[2];
    cm.submit_s = r.submit_s;
    cm.pixels = (uint64_t)c->W * c->H;
    st.members.push_back(cm);
  }
  irt_batch*& batch = s->batches[g];
  if (!batch) {
    rc = irt_batch_create(views.data(), (uint32_t)views.size(), &batch);
    if (rc) return s->fail(rc, std::string("batch: ") + irt_host::global_error());
    irt_batch_use_graph(batch, s->cfg.use_graph ? 1 : 0);
    for (uint32_t k = 0; k < BK_COUNT; ++k) batch->shape[k] = s->cfg.tile_shapes[k] <= 2 ? s->cfg.tile_shapes[k] : 0;
  } else {
    rc = irt_batch_set_views(batch, views.data(), (uint32_t)views.size());
    if (rc) return s->fail(rc, std::string("batch: ") + batch->err);
  }
  if (s->spare_events.empty()) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
      return s->fail(IRT_ERR_DEVICE, "hipEventCreate failed");
    s->spare_events.emplace_back(a, b);
  }
  st.start = s->spare_events.back().first;
  st.done = s->spare_events.back().second;
  s->spare_events.pop_back();
  hipError_t e = hipEventRecord(st.start, batch->stream);
  if (e != hipSuccess) return s->fail(IRT_ERR_DEVICE, "hipEventRecord failed");
  rc = irt_batch_encode(batch);
  if (rc) return s->fail(rc, std::string("batch: ") + batch->err);
  e = hipEventRecord(st.done, batch->svgf_stream);  // a step's last work
  if (e != hipSuccess) return s->fail(IRT_ERR_DEVICE, "hipEventRecord failed");
  st.launch_s = server_clock();
  for (auto& m : st.members) {
    m.launch_s = st.launch_s;
    m.batch_views = (uint32_t)ids.size();
  }
  for (uint32_t id : ids) {
    ServerSession& ss = s->sessions[id];
    ss.pending.pop_front();
    ++ss.in_flight;
  }
  const int launched = (int)ids.size();
  s->in_flight.push_back(std::move(st));
  return launched;
}

int irt_server_poll(irt_server* s, irt_completion* out, uint32_t capacity, int wait) {
  if (!s) return IRT_ERR_INVALID_ARGUMENT;
  int rc = drain(s, wait != 0 && s->completed.empty());
  if (rc) return rc;
  uint32_t n = 0;
  while (n < capacity && !s->completed.empty()) {
    if (out) out[n] = s->completed.front();
    s->completed.pop_front();
    ++n;
  }
  return (int)n;
}

int irt_server_sync(irt_server* s) {
  if (!s) return IRT_ERR_INVALID_ARGUMENT;
  while (!s->in_flight.empty()) {
    int rc = drain(s, true);
    if (rc) return rc;
  }
  return IRT_OK;
}

int irt_server_stats(const irt_server* s, irt_server_status* out) {
  if (!s || !out) return IRT_ERR_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  for (const auto& ss : s->sessions) {
    if (ss.open) ++out->open_sessions;
    if (ss.open) out->pending_requests += (uint32_t)ss.pending.size();
  }
  out->idle_contexts = (uint32_t)s->idle.size();
  out->steps_in_flight = (uint32_t)s->in_flight.size();
  out->device_bytes = server_bytes(s);
  out->budget_bytes = s->cfg.hbm_budget_bytes;
  return IRT_OK;
}

}  // extern "C"

static_assert(sizeof(irt_server_config) == 56, "irt_server_config layout");
static_assert(sizeof(irt_frame_request) == 152, "irt_frame_request layout");
static_assert(sizeof(irt_completion) == 56, "irt_completion layout");
static_assert(sizeof(irt_server_status) == 32, "irt_server_status layout");
