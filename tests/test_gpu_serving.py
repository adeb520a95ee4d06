"""The frame server on the GPU.

* Alone vs in the mix: a session's buffers after N frames are bit-identical
  whether it was the only session or shared steps with sessions of other
  sizes that arrived, left and were replaced (a reused, reset context)
  while it ran -- and identical to the per-view fused encode of the same
  frames.
* Admission: the slot count and the HBM budget are enforced; a closed
  session's context is reused.
* A short Poisson-arrival run completes frames with finite latencies.
* Sessions whose requests carry no inputs get them from the G-buffer and
  VPL passes: bit-identical to per-view frames with the same passes.
"""
from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

from instantradiositytest_amd.renderer import Renderer
from instantradiositytest_amd.scenes import CONFIGS
from instantradiositytest_amd.serving import Arrival, SceneStream, Server, simulate

pytestmark = pytest.mark.gpu

BUFS = ("direct2", "indirect2", "lighting", "history", "feedback", "moments", "previous_depth", "previous_normal",
        "vpl_archive")
FRAMES = 6


def streams(precision: str):
    cfgs = [CONFIGS["plain-640x400"].scaled(96, 64, FRAMES),
            replace(CONFIGS["plain-1920x1200"], seed=0x51).scaled(128, 80, FRAMES),
            replace(CONFIGS["plain-640x400"], seed=0x52).scaled(80, 48, FRAMES)]
    return [SceneStream(c, precision, FRAMES) for c in cfgs]


def submit(srv, sid, st, k, tag):
    srv.submit(sid, st.frames[k], st.cs.scene.generation, len(st.cs.scene.lights), st.cfg.vpl_count, tag)


def flush(srv):
    """Steps until no request is pending, then waits for the device."""
    while True:
        n = srv.step()
        if n == 0 and srv.stats()["pending_requests"] == 0:
            break
        if n == 0:
            srv.sync()
    srv.sync()


def run_alone(precision, st, frames):
    srv = Server(max_sessions=1, precision=precision, budget_bytes=1e12)
    sid = srv.open(st.settings, st.scene_ctx)
    for k in range(frames):
        submit(srv, sid, st, k, k)
        flush(srv)
    out = {b: srv.read(sid, b) for b in BUFS}
    srv.close_server()
    return out


def same(a: dict, b: dict, what: str):
    for k in a:
        if a[k].tobytes() != b[k].tobytes():
            raise AssertionError(f"{what}: buffer {k} differs")


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_session_alone_equals_session_in_the_mix(require_gpu, precision):
    print(f"[serving] alone vs mix precision={precision}", flush=True)
    A, B, C = streams(precision)
    alone_a = run_alone(precision, A, FRAMES)
    alone_c = run_alone(precision, C, 4)
    alone_b = run_alone(precision, B, 3)

    srv = Server(max_sessions=3, precision=precision, budget_bytes=1e12, max_in_flight=2)
    a = srv.open(A.settings, A.scene_ctx)
    b = srv.open(B.settings, B.scene_ctx)
    submit(srv, a, A, 0, 1)
    submit(srv, b, B, 0, 2)
    # two step groups (one per step in flight): A and B land in different
    # groups, so the first two steps take one session each and overlap
    assert srv.step() == 1
    assert srv.step() == 1
    c = srv.open(C.settings, C.scene_ctx)
    submit(srv, a, A, 1, 3)
    submit(srv, c, C, 0, 4)
    flush(srv)
    submit(srv, b, B, 1, 5)
    srv.step()
    submit(srv, a, A, 2, 6)
    submit(srv, b, B, 2, 7)
    submit(srv, c, C, 1, 8)
    srv.step()  # at most two steps in flight: later requests wait for the next calls
    flush(srv)
    done_b = {k: srv.read(b, k) for k in BUFS}
    srv.close(b)
    # a new session with B's settings reuses B's context after a reset
    d = srv.open(B.settings, B.scene_ctx)
    assert d is not None
    st = srv.stats()
    print("stats after reuse", st, flush=True)
    for k in range(3, FRAMES):
        submit(srv, a, A, k, 100 + k)
        if k - 3 < 3:
            submit(srv, d, B, k - 3, 200 + k)
        if k - 1 < 4:
            submit(srv, c, C, k - 1, 300 + k)
        srv.step()
        srv.poll(wait=False)
    flush(srv)
    same(alone_a, {k: srv.read(a, k) for k in BUFS}, "session A in the mix")
    same(alone_c, {k: srv.read(c, k) for k in BUFS}, "session C in the mix")
    same(alone_b, done_b, "session B in the mix")
    same(alone_b, {k: srv.read(d, k) for k in BUFS}, "session D on a reused context")

    # the served session equals the per-view fused encode of the same frames
    r = Renderer(A.cs.scene, A.settings, precision=precision, schedule="fused")
    for k in range(FRAMES):
        r.encode(A.cs.scene, A.cs.frame_inputs(k))
    r.sync()
    same(alone_a, {k: r.read(k) for k in BUFS}, "server vs per-view encode")


def test_admission_by_slots_and_budget(require_gpu):
    print("[serving] admission", flush=True)
    A, B, _ = streams("mixed")
    srv = Server(max_sessions=8, precision="mixed", budget_bytes=1e12)
    s0 = srv.open(A.settings, A.scene_ctx)
    per = srv.stats()["device_bytes"]
    srv.close_server()
    # room for two sessions of A, not three
    srv = Server(max_sessions=8, precision="mixed", budget_bytes=int(2.5 * per))
    s0 = srv.open(A.settings, A.scene_ctx)
    s1 = srv.open(A.settings, A.scene_ctx)
    assert s0 is not None and s1 is not None
    assert srv.open(A.settings, A.scene_ctx) is None
    srv.close(s0)
    s2 = srv.open(A.settings, A.scene_ctx)  # reuses s0's context
    assert s2 is not None
    st = srv.stats()
    assert st["open_sessions"] == 2 and st["device_bytes"] <= st["budget_bytes"]
    # a larger session does not fit until an idle context is evicted
    srv.close(s1)
    assert srv.stats()["idle_contexts"] == 1
    s3 = srv.open(B.settings, B.scene_ctx)
    st = srv.stats()
    print("after evict", st, flush=True)
    if s3 is not None:
        assert st["device_bytes"] <= st["budget_bytes"]
    srv.close_server()
    one = Server(max_sessions=1, precision="mixed", budget_bytes=1e12)
    assert one.open(A.settings, A.scene_ctx) is not None
    assert one.open(A.settings, A.scene_ctx) is None


def test_short_poisson_run(require_gpu):
    print("[serving] short poisson run", flush=True)
    sts = streams("mixed")
    srv = Server(max_sessions=4, precision="mixed", budget_bytes=1e12)
    arrivals = [Arrival(0.01 * i, i % 3, 5 + i, i % FRAMES) for i in range(6)]
    rep = simulate(srv, sts, arrivals, duration_s=3.0, warmup_s=0.0, cadence_hz=60.0)
    print(rep.as_dict(), flush=True)
    assert rep.sessions_admitted == 6 and rep.sessions_finished == 6
    assert rep.frames == sum(a.lifetime for a in arrivals)
    assert np.isfinite(rep.p99_ms) and rep.p99_ms > 0


def test_sessions_render_their_own_inputs(require_gpu):
    print("[serving] device-produced inputs", flush=True)
    A = SceneStream(CONFIGS["plain-640x400"].scaled(96, 64, FRAMES), "mixed", FRAMES, device_inputs=True)
    B = SceneStream(replace(CONFIGS["plain-640x400"], seed=0x53).scaled(80, 48, FRAMES), "mixed", FRAMES,
                    device_inputs=True)
    srv = Server(max_sessions=2, precision="mixed", budget_bytes=1e12, max_in_flight=2)
    a = srv.open(A.settings, A.scene_ctx)
    b = srv.open(B.settings, B.scene_ctx)
    for k in range(FRAMES):
        submit(srv, a, A, k, k)
        submit(srv, b, B, k, 100 + k)
        srv.step()
    flush(srv)
    for st, sid in ((A, a), (B, b)):
        r = Renderer(st.cs.scene, st.settings, precision="mixed", schedule="fused")
        r.set_geometry(*st.cs.scene.geometry())
        for k in range(FRAMES):
            inp = st.cs.frame_inputs(k)
            inp.depth = inp.vpls = None
            r.encode(st.cs.scene, inp)
        r.sync()
        same({k: r.read(k) for k in BUFS}, {k: srv.read(sid, k) for k in BUFS}, "served with device inputs")
        r.close()
    srv.close_server()
