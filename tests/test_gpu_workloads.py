"""The batched and serving workloads at their full sizes on the GPU.

* 16 views (4 x 1920x1200, 4 x 1280x720, 8 x 640x360, each with its own
  seeded scene of 8-48 meshes at 64^3-128^3 and 300-2000 VPLs) rendered by
  one launch per pass with HIP-graph replay and mixed tile shapes, mixed
  storage, three frames.  Every view: finite radiance and history, reservoir
  M <= 256 (the spatial cap, lighting.wgsl:97/105), reservoir age <= 4,
  SVGF age in [1, 32] (svgf.wgsl:61).  One view of each resolution class
  against the oracle's closed loop of the same frames, run on the same
  16-bit SDF samples, to the mixed profile's bar (relative L2 <= 1e-3, >= 99 %
  of pixels within 1e-3 relative).
* The serving mix: 32 sessions drawn 640x360 / 1280x720 / 1920x1200 at
  0.4 / 0.4 / 0.2, each with its own seeded scene (8-48 meshes, 500-2000
  VPLs, 3-5 bounces) and its G-buffer and VPL pool produced on the device,
  admitted together and served in two step groups; one session of each
  class run again alone must give bit-identical buffers.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from instantradiositytest_amd.renderer import Batch, Renderer, half_to_float
from instantradiositytest_amd.scenes import ConfigScene
from instantradiositytest_amd.serving import SERVING_MIX, SceneStream, Server, batched_view_configs, session_config
from parity import compare_images, rgba16_to_f64
from test_gpu_parity import atlas_as_stored

pytestmark = pytest.mark.gpu

SHAPES = ["16x8", "16x16", "8x8", "16x16", "16x8", "8x8", "16x16", "8x8"]


def log(*a):
    print(*a, flush=True)


def test_batched_sixteen_views_full_size(require_gpu):
    frames = 3
    cfgs = batched_view_configs(frames)
    log(f"[c4] building {len(cfgs)} scenes")
    css = [ConfigScene(c) for c in cfgs]
    views = [Renderer(cs.scene, c.settings(), precision="mixed", schedule="fused") for cs, c in zip(css, cfgs)]
    batch = Batch(views, graph=True)
    for i, s in enumerate(SHAPES):
        batch.set_tile(i, s)
    checked = {0: None, 4: None, 8: None}  # one view per resolution class
    orc = {}
    for vi in checked:
        c, cs = cfgs[vi], css[vi]
        st = c.settings()
        sc = cs.scene
        orc[vi] = (oracle.OracleRenderer(32, c.width, c.height, c.vpl_count, st.restir_max_history),
                   oracle.SceneRecords(sc.objects, sc.meshes, sc.nodes, sc.lights, atlas_as_stored(sc, "mixed")), st)
    for fi in range(frames):
        inputs = [cs.frame_inputs(fi) for cs in css]
        recs = [v.begin_frame(cs.scene, inp) for v, cs, inp in zip(views, css, inputs)]
        batch.encode_current()
        batch.finish()
        for vi, (o, rec, st) in orc.items():
            o.encode(recs[vi], rec, inputs[vi], st.svgf_iterations, 0)
        log(f"[c4] frame {fi} done")
    batch.sync()
    for vi, (v, c) in enumerate(zip(views, cfgs)):
        lt = half_to_float(v.read("lighting")).astype(np.float64)
        hist = half_to_float(v.read("history")).astype(np.float64)
        mom = v.read("moments")
        depth = v.read("depth")
        assert np.isfinite(lt).all() and np.isfinite(hist).all(), c.name
        for dom in ("direct2", "indirect2"):
            res = v.read(dom)
            assert int(res["count"].max()) <= 256, (c.name, dom)
            assert int(res["age"].max()) <= 4, (c.name, dom)
            assert int(res["age"].max()) >= 1, (c.name, dom)
        ages = mom[..., 2][depth < 1]
        assert ages.min() >= 1 and ages.max() <= 32 and ages.max() == frames, c.name
        if vi in orc:
            o = orc[vi][0]
            s_lt = compare_images(lt, o.lighting, 1e-3, 1e-4)
            s_h = compare_images(hist, o.histories[(frames - 1) % 2], 1e-3, 1e-4)
            log(f"[c4] {c.name}: lighting {s_lt}, history {s_h}")
            for s in (s_lt, s_h):
                assert s["rel_l2"] <= 1e-3 and s["within"] >= 0.99, (c.name, s)
    batch.close()
    for v in views:
        v.close()


BUFS = ("direct2", "indirect2", "lighting", "history", "feedback", "moments", "previous_depth", "previous_normal",
        "vpl_archive")


def test_serving_mix_thirty_two_sessions_alone_equals_mix(require_gpu):
    frames = 3
    rng = np.random.default_rng(0xC5)
    classes = [0, 1, 2] + list(rng.choice(3, size=29, p=[p for _, _, p in SERVING_MIX]))
    log(f"[c5] building 32 session scenes, classes {np.bincount(classes).tolist()}")
    streams = []
    for i, k in enumerate(classes):
        w, h, _ = SERVING_MIX[k]
        cfg = session_config(0xC500 + i, w, h, frames, rng)
        streams.append(SceneStream(cfg, "mixed", frames, device_inputs=True))
    srv = Server(max_sessions=32, precision="mixed", budget_bytes=200e9, max_in_flight=2)
    sids = [srv.open(st.settings, st.scene_ctx) for st in streams]
    assert all(s is not None for s in sids)
    for k in range(frames):
        for i, (sid, st) in enumerate(zip(sids, streams)):
            srv.submit(sid, st.frames[k], st.cs.scene.generation, len(st.cs.scene.lights), st.cfg.vpl_count,
                       k * 100 + i)
            if i % 5 == 4:
                srv.step()  # requests keep arriving while steps run
        srv.step()
        log(f"[c5] frame {k} submitted for every session")
    while True:
        n = srv.step()
        if n == 0 and srv.stats()["pending_requests"] == 0:
            break
        if n == 0:
            srv.sync()
    srv.sync()
    st_all = srv.stats()
    log(f"[c5] {st_all}")
    mix = {i: {b: srv.read(sids[i], b) for b in BUFS} for i in (0, 1, 2)}
    srv.close_server()
    for i in (0, 1, 2):
        st = streams[i]
        one = Server(max_sessions=1, precision="mixed", budget_bytes=200e9)
        sid = one.open(st.settings, st.scene_ctx)
        for k in range(frames):
            one.submit(sid, st.frames[k], st.cs.scene.generation, len(st.cs.scene.lights), st.cfg.vpl_count, k)
            while one.step() > 0 or one.stats()["pending_requests"] > 0:
                one.sync()
            one.sync()
        for b in BUFS:
            assert one.read(sid, b).tobytes() == mix[i][b].tobytes(), (st.cfg.name, b)
        log(f"[c5] session {st.cfg.name} ({st.cfg.width}x{st.cfg.height}, {len(st.cs.scene.meshes)} meshes, "
            f"{st.cfg.vpl_count} VPLs) alone == in the mix")
        one.close_server()
