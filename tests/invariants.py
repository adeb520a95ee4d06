"""Behavioural invariants of the reference's GPU integration tests
(src/gpu_tests.rs), written once and run against two backends:
the CPU restatement (``OracleBackend``) and the HIP path (``GpuBackend``).

Scenes: the reference uses its exported test scene and an analytic cube
fixture (tools/test_export.py:22-32: 2 m cube at the origin, camera at
(0,0,5), point light at (0,0,3) with 10 W, SDF resolution 16).  The cube is
restated here exactly; the test scene is replaced by the synthetic room with
a static camera.
"""
from __future__ import annotations

import math

import numpy as np

import oracle
from instantradiositytest_amd.renderer import Renderer, float_to_half, half_to_float, pack_frame
from instantradiositytest_amd.scenes import CONFIGS, ConfigScene, FrameInputs, Scene, look_at_camera
from instantradiositytest_amd.types import LIGHT_DT, SG_BOX, SGOBJECT_DT, SHAPE_DT


# ------------------------------------------------------------------ backends
class GpuBackend:
    def __init__(self, scene: Scene, settings, schedule="fused", precision="fp32"):
        self.r = Renderer(scene, settings, precision=precision, schedule=schedule)
        self.settings = settings

    @property
    def debug_mode(self):
        return self.r.debug_mode

    @debug_mode.setter
    def debug_mode(self, v):
        self.r.debug_mode = v

    def encode(self, scene, inputs):
        self.r.encode(scene, inputs)

    def lighting(self):
        return half_to_float(self.r.read("lighting"))

    def moments(self):
        return self.r.read("moments")

    def history(self):
        return half_to_float(self.r.read("history"))

    def reservoirs(self, indirect):
        return self.r.read("indirect2" if indirect else "direct2")

    def vpl_archive(self):
        return self.r.read("vpl_archive")

    def svgf_signal(self, signal, material_edge=False, disoccluded=False):
        """Renderer::test_svgf_signal (renderer.rs:557-630)."""
        H, W = signal.shape[:2]
        self.r.write("lighting", float_to_half(signal))
        if material_edge:
            cols = np.zeros((H, W, 4), np.float32)
            cols[:, : W // 2] = (1, 0, 0, 1)
            cols[:, W // 2:] = (0, 1, 0, 1)
            self.r.write("color", float_to_half(cols))
        if disoccluded:
            self.r.write("motion", float_to_half(np.full((H, W, 2), 2.0, np.float32)))
        self.r.svgf(self.settings.svgf_iterations, 0)
        self.r.sync()
        return self.history(), self.moments(), self.r.read("depth")


class OracleBackend:
    def __init__(self, scene: Scene, settings, bits=32):
        self.settings = settings
        self.W, self.H = settings.render_width, settings.render_height
        self.o = oracle.OracleRenderer(bits, self.W, self.H, settings.vpl_count, settings.restir_max_history)
        self.frame_index = 0
        self.previous_vp = None
        self.previous_generation = 0
        self.previous_debug = 0
        self.debug_mode = 0
        self.last = None

    def encode(self, scene, inputs):
        vp = inputs.vp
        pvp = vp if self.previous_vp is None else self.previous_vp
        history = self.frame_index > 0 and self.previous_debug == self.debug_mode
        f = pack_frame(self.settings, vp, pvp, inputs.camera, self.frame_index, scene.generation,
                       self.previous_generation, history, self.debug_mode, len(scene.lights), len(scene.objects))
        rec = oracle.SceneRecords.from_scene(scene)
        gb = self.o.encode(f, rec, inputs, self.settings.svgf_iterations, self.debug_mode)
        self.last = (f, inputs, gb)
        self.previous_vp = vp
        self.previous_generation = scene.generation
        self.previous_debug = self.debug_mode
        self.frame_index += 1

    def cur(self):
        return (self.frame_index - 1) % 2

    def lighting(self):
        return self.o.lighting.astype(np.float32)

    def moments(self):
        return self.o.moments[self.cur()].astype(np.float32)

    def history(self):
        return self.o.histories[self.cur()].astype(np.float32)

    def reservoirs(self, indirect):
        return (self.o.indirect if indirect else self.o.direct)[2]

    def vpl_archive(self):
        return self.o.archive

    def svgf_signal(self, signal, material_edge=False, disoccluded=False):
        f, inputs, gb = self.last
        H, W = signal.shape[:2]
        self.o.lighting[:] = half_to_float(float_to_half(signal))
        color = gb.color.copy()
        motion = gb.motion.copy()
        if material_edge:
            color[:, : W // 2] = (1, 0, 0, 1)
            color[:, W // 2:] = (0, 1, 0, 1)
        if disoccluded:
            motion[:] = 2.0
        g2 = oracle.GBufferPlanes(gb.depth, gb.normal, color, gb.params, motion, self.o.previous_depth,
                                  self.o.previous_normal)
        self.o.svgf(f, g2, self.settings.svgf_iterations, 0)
        return self.history(), self.moments(), gb.depth.astype(np.float32)


# -------------------------------------------------------------------- scenes
def static_room(width=160, height=100):
    cfg = CONFIGS["plain-640x400"].scaled(width, height, 64)
    cfg.orbit_deg = 0.0
    return cfg, ConfigScene(cfg)


def cube_scene(occluder_at=None):
    """The exporter test's cube fixture: 2 m cube, SDF resolution 16, 10 W point light at (0,0,3)."""
    S = np.zeros(1, SHAPE_DT)
    S[0]["kind"], S[0]["resolution"], S[0]["p"] = SG_BOX, 16, (1, 1, 1)
    objs = [dict(mesh=0, position=(0, 0, 0), scale=1.0)]
    if occluder_at is not None:
        objs.append(dict(mesh=0, position=occluder_at, scale=0.25))
    O = np.zeros(len(objs), SGOBJECT_DT)
    for i, o in enumerate(objs):
        O[i]["mesh"] = o["mesh"]
        O[i]["position"] = o["position"]
        O[i]["rotation"] = (0, 0, 0, 1)
        O[i]["scale"] = o["scale"]
        O[i]["albedo"] = (0.8, 0.8, 0.8)
        O[i]["rough"] = 0.5
        O[i]["metal"] = 0.0
    Lt = np.zeros(1, LIGHT_DT)
    Lt[0]["position_strength"] = (0, 0, 3, 10.0 / (4 * math.pi))
    Lt[0]["direction_angle"] = (0, 0, -1, 0)
    Lt[0]["color_type"] = (1, 1, 1, 0)
    return Scene(S, O, Lt)


def cube_inputs(scene: Scene, settings, vpl_bounces=1, seed=1) -> FrameInputs:
    W, H = settings.render_width, settings.render_height
    cam = look_at_camera((0, 0, 5), (0, 0, 0), fov_y=0.69, near=0.1, far=100.0)
    depth, normal, color, params, motion = scene.gbuffer(cam, cam, W, H)
    paths = max(settings.vpl_count // vpl_bounces, 1)
    vpls = scene.vpl_pool(seed, paths, vpl_bounces)
    if len(vpls) < settings.vpl_count:
        vpls = np.concatenate([vpls, np.zeros(settings.vpl_count - len(vpls), vpls.dtype)])
    return FrameInputs(depth, normal, color, params, motion, vpls[: settings.vpl_count],
                       scene.camera_matrix(cam, W / H), np.asarray(cam["position"], np.float32))


# ---------------------------------------------------------------- invariants
def svgf_noise_edges_disocclusion(make_backend):
    """gpu_tests.rs:81-181 on the static room at 160x100."""
    cfg, cs = static_room()
    settings = cfg.settings()
    b = make_backend(cs.scene, settings)
    rng = 123456789
    last = None
    W, H = cfg.width, cfg.height
    for fi in range(12):
        b.encode(cs.scene, cs.frame_inputs(fi))
        vals = np.empty(W * H, np.float32)
        for i in range(W * H):
            rng ^= (rng << 13) & 0xFFFFFFFF
            rng ^= rng >> 17
            rng ^= (rng << 5) & 0xFFFFFFFF
            vals[i] = 0.0 if rng & 1 == 0 else 2.0
        sig = np.repeat(vals.reshape(H, W, 1), 4, axis=2)
        sig[..., 3] = 1.0
        last = b.svgf_signal(sig)
    filtered, moments, depth = last
    m = depth < 1.0
    v = filtered[..., 0][m]
    assert np.isfinite(v).all() and np.isfinite(filtered[..., 3][m]).all()
    assert (moments[..., 1][m] >= 0).all() and np.isfinite(moments[..., 1][m]).all()
    mse = float(np.mean((v - 1.0) ** 2))
    mean = float(np.mean(v))
    assert mse < 0.1, f"SVGF must remove at least 90% of the unit input MSE: {mse}"
    assert abs(mean - 1.0) < 0.04, f"SVGF must preserve mean radiance: {mean}"
    # material edge
    cs.scene.changed()
    b.encode(cs.scene, cs.frame_inputs(12))
    sig = np.where((np.arange(W) < W // 2)[None, :, None], 0.5, 2.0) * np.ones((H, W, 4), np.float32)
    sig[..., 3] = 1.0
    filtered, _, depth = b.svgf_signal(sig.astype(np.float32), material_edge=True)
    cols = np.arange(W)
    edge = (depth < 1.0) & ((cols >= W // 2 - 2) & (cols <= W // 2 + 1))[None, :]
    assert edge.sum() > 20
    assert np.all(np.abs(filtered[..., 0][edge] - sig[..., 0][edge]) < 0.02), "material edge bled"
    # disocclusion
    b.encode(cs.scene, cs.frame_inputs(13))
    _, moments, depth = b.svgf_signal(sig.astype(np.float32), disoccluded=True)
    assert np.all(moments[..., 2][depth < 1.0] == 1.0)
    # HDR
    cs.scene.changed()
    b.encode(cs.scene, cs.frame_inputs(14))
    hdr = np.where((np.arange(W * H) % 2 == 0).reshape(H, W), 0.0, 60000.0)[..., None] * np.ones(4)
    hdr[..., 3] = 1.0
    filtered, moments, _ = b.svgf_signal(hdr.astype(np.float32))
    assert np.isfinite(filtered).all(), "HDR wavelet radiance/variance must stay finite"
    assert np.isfinite(moments).all(), "HDR squared luminance must not overflow"
    assert (moments[..., 1] > 65504.0).any(), "moments need FP32 range"


def lighting_history_and_motion(make_backend, frames=40):
    """gpu_tests.rs:185-252 on the static room at 160x100 (debug mode 1)."""
    cfg, cs = static_room()
    settings = cfg.settings()
    b = make_backend(cs.scene, settings)
    b.debug_mode = 1
    for fi in range(frames):
        b.encode(cs.scene, cs.frame_inputs(fi))
    ages = b.moments()[..., 2]
    assert ages.max() >= 16.0, "static surfaces must accumulate SVGF history"
    assert (ages >= 16).mean() > 0.25
    for indirect in (False, True):
        r = b.reservoirs(indirect)
        assert int(r["count"].max()) >= settings.restir_max_history * settings.restir_direct_candidates
        age = int(r["age"].max())
        assert 1 <= age <= settings.restir_max_history
        assert (r["age"] >= 1).mean() > 0.1
    # transport resources (renderer.rs:457-500)
    lt = b.lighting()
    assert np.isfinite(lt).all() and (lt >= 0).all()
    fi = frames - 1
    slots = settings.restir_max_history + 1
    arch = b.vpl_archive()
    pool = cs.frame_inputs(fi).vpls
    s = fi % slots
    assert arch[s * settings.vpl_count:(s + 1) * settings.vpl_count].tobytes() == pool.tobytes()
    assert (b.reservoirs(False)["count"] > settings.restir_direct_candidates).any()
    assert (b.reservoirs(True)["count"] > settings.restir_vpl_candidates).any()
    assert (lt[..., :3] > 0.05).any()
    # lights off -> black immediately
    lights = cs.scene.lights.copy()
    lights["position_strength"][:, 3] = 0.0
    cs.scene.set_lights(lights)
    cs.scene.changed()
    b.encode(cs.scene, cs.frame_inputs(frames))
    assert np.all(b.lighting()[..., :3] == 0), "changed lighting must invalidate history immediately"
    assert np.all(b.history()[..., :3] == 0)
    # debug change -> SVGF history reset
    b.debug_mode = 2
    b.encode(cs.scene, cs.frame_inputs(frames + 1))
    assert b.moments()[..., 2].max() == 1.0


def inverse_square_and_spot(make_backend):
    """gpu_tests.rs:283-317 with the analytic cube fixture, 64x64, debug 1."""
    from instantradiositytest_amd.types import make_settings
    sc = cube_scene()
    settings = make_settings(render_width=64, render_height=64)
    b = make_backend(sc, settings)
    b.debug_mode = 1

    def center():
        b.encode(sc, cube_inputs(sc, settings))
        return float(b.lighting()[32, 32, 0])

    near = center()
    assert near > 0
    lights = sc.lights.copy()
    lights["position_strength"][0, :3] = (0, 0, 6)
    sc.set_lights(lights)
    sc.changed()
    far = center()
    assert abs(near / far - 6.25) < 0.05, near / far
    lights["color_type"][0, 3] = 1.0
    lights["direction_angle"][0] = (0, 0, -1, 0.4)
    sc.set_lights(lights)
    sc.changed()
    assert abs(center() - far) < 0.002
    lights["direction_angle"][0] = (0, 0, 1, 0.4)
    sc.set_lights(lights)
    sc.changed()
    assert center() == 0.0


def moving_occluder(make_backend):
    """gpu_tests.rs:321-352: a uniformly scaled occluder moved into the light ray."""
    from instantradiositytest_amd.types import make_settings
    sc = cube_scene(occluder_at=(20, 0, 2))
    lights = sc.lights.copy()
    lights["position_strength"][0, :3] = (3, 0, 3)
    sc.set_lights(lights)
    settings = make_settings(render_width=64, render_height=64, vpl_max_bounces=1)
    b = make_backend(sc, settings)
    b.debug_mode = 1
    b.encode(sc, cube_inputs(sc, settings))
    assert float(b.lighting()[32, 32, 0]) > 0.001
    objs = sc.sg_objects.copy()
    objs[1]["position"] = (1.5, 0, 2)
    sc.set_objects(objs)
    sc.changed()
    b.encode(sc, cube_inputs(sc, settings))
    assert float(b.lighting()[32, 32, 0]) == 0.0, "occluder must update the SDF hierarchy"
