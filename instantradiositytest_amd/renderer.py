"""Host side of the lighting + denoising path over the C ABI.

``Renderer`` plays the part of the reference's ``Renderer`` for this path
(src/renderer.rs:713-766 construction, 1675-1984 ``encode``): it owns one
library context (device state of one view), uploads the scene records,
packs the Frame record each frame with the same bookkeeping
(frame_index, previous_vp, previous_generation, previous_debug,
flags.x = frame_index > 0 && previous_debug == debug_mode) and exposes the
reference's test hooks (renderer.rs:457-712) as read-backs.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from ._lib import check
from .scenes import FrameInputs, Scene
from .types import FRAME_DT, RESERVOIR_DT, VPL_DT, Settings, ptr

PRECISIONS = {"fp32": 0, "mixed": 1}
SCHEDULES = {"passes": 0, "fused": 1}

BUFFERS = {
    "direct0": 0, "direct1": 1, "direct2": 2, "indirect0": 3, "indirect1": 4, "indirect2": 5,
    "direct_lighting": 6, "lighting": 7, "history": 8, "moments": 9, "moments_old": 10,
    "feedback": 11, "feedback_old": 12, "temporal": 13, "depth": 14, "normal": 15, "color": 16,
    "params": 17, "motion": 18, "previous_depth": 19, "previous_normal": 20, "vpl_archive": 21,
    "scratch0": 22, "scratch1": 23,
}
PASSES = {
    "candidates_direct": 0, "candidates_indirect": 1, "temporal_direct": 2, "temporal_indirect": 3,
    "spatial_direct": 4, "spatial_indirect": 5, "resolve_direct": 6, "resolve_indirect": 7,
    "svgf_temporal": 8, "svgf_variance": 9, "svgf_atrous": 10,
}
# (numpy dtype, channels) of each buffer in the reference format.
_FORMAT = {
    "reservoir": (RESERVOIR_DT, None),
    "rgba32f": (np.float32, 4),
    "rgba16f": (np.uint16, 4),
    "rg16f": (np.uint16, 2),
    "depth": (np.float32, None),
    "vpl": (VPL_DT, None),
}
_BUFFER_FORMAT = {
    **{k: "reservoir" for k in ("direct0", "direct1", "direct2", "indirect0", "indirect1", "indirect2")},
    "direct_lighting": "rgba32f", "lighting": "rgba16f", "history": "rgba16f", "moments": "rgba32f",
    "moments_old": "rgba32f", "feedback": "rgba16f", "feedback_old": "rgba16f", "temporal": "rgba16f",
    "depth": "depth", "normal": "rgba16f", "color": "rgba16f", "params": "rg16f", "motion": "rg16f",
    "previous_depth": "depth", "previous_normal": "rgba16f", "vpl_archive": "vpl", "scratch0": "rgba16f",
    "scratch1": "rgba16f",
}


def read_buffer(ctx, name: str, width: int, height: int) -> np.ndarray:
    """Read-back of one buffer of a context in the reference format
    (irt_read), shaped [H, W(, C)] for images."""
    L = _lib.irt()
    fmt = _BUFFER_FORMAT[name]
    dt, ch = _FORMAT[fmt]
    nbytes = L.irt_buffer_bytes(ctx, BUFFERS[name])
    itemsize = np.dtype(dt).itemsize * (ch or 1)
    out = np.empty(nbytes // itemsize * (ch or 1), dtype=dt)
    check(L.irt_read(ctx, BUFFERS[name], ptr(out), nbytes), ctx)
    if fmt == "vpl":
        return out
    return out.reshape((height, width) + ((ch,) if ch else ()))


def pack_frame(settings: Settings, vp, previous_vp, camera, frame_index: int, generation: int,
               previous_generation: int, history_enabled: bool, debug_mode: int, light_count: int,
               object_count: int) -> np.ndarray:
    """Frame record of renderer.rs:1687-1735 (via the library's host helper)."""
    f = np.zeros(1, dtype=FRAME_DT)
    vp = np.ascontiguousarray(vp, dtype=np.float32)
    pvp = np.ascontiguousarray(previous_vp, dtype=np.float32)
    cam = np.ascontiguousarray(camera, dtype=np.float32)
    check(_lib.irt().irt_host_pack_frame(ctypes.byref(settings), ptr(vp), ptr(pvp), ptr(cam), frame_index,
                                         generation, previous_generation, int(history_enabled), debug_mode,
                                         light_count, object_count, ptr(f)))
    return f


class Renderer:
    def __init__(self, scene: Scene, settings: Settings, precision: str = "fp32", schedule: str = "passes",
                 initial_vp=None):
        L = _lib.irt()
        self.L = L
        self.settings = settings
        self.precision = precision
        self.schedule_name = schedule
        self.timing = False
        ctx = ctypes.c_void_p()
        check(L.irt_create(ctypes.byref(settings), PRECISIONS[precision], ctypes.byref(ctx)))
        self.ctx = ctx
        check(L.irt_set_schedule(ctx, SCHEDULES[schedule]), ctx)
        self.W, self.H = settings.render_width, settings.render_height
        atlas = scene.atlas()
        d, h, w = atlas.shape
        check(L.irt_set_sdf_atlas(ctx, ptr(atlas), w, h, d), ctx)
        check(L.irt_set_meshes(ctx, ptr(scene.meshes), len(scene.meshes)), ctx)
        self.upload_scene(scene)
        self.frame_index = 0
        self.previous_vp = None if initial_vp is None else np.asarray(initial_vp, np.float32)
        self.previous_generation = 0
        self.previous_debug = 0
        self.debug_mode = 0
        self.last_frame = None

    # -- scene ------------------------------------------------------------
    def upload_scene(self, scene: Scene):
        """Per-frame object / hierarchy / light uploads (renderer.rs:1736-1782)."""
        L, ctx = self.L, self.ctx
        check(L.irt_set_objects(ctx, ptr(scene.objects), len(scene.objects), ptr(scene.nodes), len(scene.nodes)), ctx)
        lights = scene.lights if len(scene.lights) else np.zeros(0, dtype=scene.lights.dtype)
        check(L.irt_set_lights(ctx, ptr(lights) if len(lights) else None, len(lights)), ctx)

    def set_geometry(self, vertices: np.ndarray, indices: np.ndarray, materials: np.ndarray):
        """Triangle geometry and materials of the VPL pass (irt_set_geometry)."""
        v = np.ascontiguousarray(vertices)
        i = np.ascontiguousarray(indices, dtype=np.uint32)
        m = np.ascontiguousarray(materials)
        check(self.L.irt_set_geometry(self.ctx, ptr(v), len(v), ptr(i), len(i), ptr(m), len(m)), self.ctx)

    def present(self) -> np.ndarray:
        """The display image of the last frame (irt_present): uint8 [H, W, 4] RGBA, sRGB."""
        out = np.empty((self.H, self.W, 4), np.uint8)
        check(self.L.irt_present(self.ctx, ptr(out)), self.ctx)
        return out

    def capture(self, path) -> np.ndarray:
        """Renderer::capture (renderer.rs:2109-2163): the display image as an RGBA PNG."""
        from .images import write_png
        img = self.present()
        write_png(path, img)
        return img

    def render_gbuffer(self):
        """The G-buffer pass of the frame set last (irt_render_gbuffer, gbuffer.wgsl)."""
        check(self.L.irt_render_gbuffer(self.ctx), self.ctx)

    def set_textures(self, textures: np.ndarray):
        """Material texture array (irt_set_textures): uint8 [layers, size, size, 4] RGBA8."""
        t = np.ascontiguousarray(textures, np.uint8)
        if t.ndim != 4 or t.shape[1] != t.shape[2] or t.shape[3] != 4:
            raise ValueError("textures must be [layers, size, size, 4]")
        check(self.L.irt_set_textures(self.ctx, ptr(t), t.shape[1], t.shape[0]), self.ctx)

    def generate_vpls(self):
        """The VPL pass of the frame set last (irt_generate_vpls, vpl.wgsl:2-47)."""
        check(self.L.irt_generate_vpls(self.ctx), self.ctx)

    def schedule(self, name: str):
        check(self.L.irt_set_schedule(self.ctx, SCHEDULES[name]), self.ctx)
        self.schedule_name = name

    # -- frame ------------------------------------------------------------
    def begin_frame(self, scene: Scene, inputs: FrameInputs) -> np.ndarray:
        """Upload the frame's records and inputs; returns the packed Frame.
        inputs.depth None: the G-buffer pass renders the G-buffer (set_geometry
        first); inputs.vpls None: the VPL pass generates the pool."""
        L, ctx = self.L, self.ctx
        self.upload_scene(scene)
        vp = np.asarray(inputs.vp, np.float32)
        pvp = vp if self.previous_vp is None else self.previous_vp
        history = self.frame_index > 0 and self.previous_debug == self.debug_mode
        f = pack_frame(self.settings, vp, pvp, inputs.camera, self.frame_index, scene.generation,
                       self.previous_generation, history, self.debug_mode, len(scene.lights), len(scene.objects))
        check(L.irt_set_frame(ctx, ptr(f)), ctx)
        if inputs.depth is None:
            self.render_gbuffer()
        else:
            check(L.irt_set_gbuffer(ctx, ptr(inputs.depth), ptr(inputs.normal), ptr(inputs.color),
                                    ptr(inputs.params), ptr(inputs.motion)), ctx)
        if inputs.vpls is None:
            self.generate_vpls()
        else:
            check(L.irt_set_vpls(ctx, ptr(inputs.vpls), len(inputs.vpls)), ctx)
        self.last_frame = f
        self._pending = (vp, scene.generation)
        return f

    def begin_frame_device(self, bufs, vp, camera, generation: int, light_count: int, object_count: int):
        """begin_frame with the G-buffer planes and VPL pool already in device
        memory (`bufs`: objects with a `.ptr` c_void_p, in the order depth,
        normal, color, params, motion, vpls); the scene records stay as
        uploaded."""
        L, ctx = self.L, self.ctx
        vp = np.asarray(vp, np.float32)
        pvp = vp if self.previous_vp is None else self.previous_vp
        history = self.frame_index > 0 and self.previous_debug == self.debug_mode
        f = pack_frame(self.settings, vp, pvp, camera, self.frame_index, generation, self.previous_generation,
                       history, self.debug_mode, light_count, object_count)
        check(L.irt_set_frame(ctx, ptr(f)), ctx)
        d = bufs
        check(L.irt_set_gbuffer_device(ctx, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, d[4].ptr), ctx)
        check(L.irt_set_vpls_device(ctx, d[5].ptr, self.settings.vpl_count), ctx)
        self.last_frame = f
        self._pending = (vp, generation)
        return f

    def finish_frame(self):
        """Bookkeeping at the end of Renderer::encode (renderer.rs:1973-1982)."""
        vp, generation = self._pending
        self.previous_vp = vp
        self.previous_generation = generation
        self.previous_debug = self.debug_mode
        self.frame_index += 1

    def encode(self, scene: Scene, inputs: FrameInputs) -> np.ndarray:
        """One frame of the path: lighting, SVGF, history (Renderer::encode's hot-path part)."""
        f = self.begin_frame(scene, inputs)
        check(self.L.irt_encode(self.ctx), self.ctx)
        self.finish_frame()
        return f

    def run_pass(self, name: str, arg: int = 0):
        check(self.L.irt_pass(self.ctx, PASSES[name], arg), self.ctx)

    def lighting(self):
        check(self.L.irt_lighting(self.ctx), self.ctx)

    def svgf(self, iterations: int | None = None, debug: int | None = None):
        it = self.settings.svgf_iterations if iterations is None else iterations
        dbg = self.debug_mode if debug is None else debug
        check(self.L.irt_svgf(self.ctx, it, dbg), self.ctx)

    def end_frame(self):
        check(self.L.irt_end_frame(self.ctx), self.ctx)

    def sync(self):
        check(self.L.irt_sync(self.ctx), self.ctx)

    # -- buffers ------------------------------------------------------------
    def read(self, name: str) -> np.ndarray:
        return read_buffer(self.ctx, name, self.W, self.H)

    def write(self, name: str, data: np.ndarray):
        data = np.ascontiguousarray(data)
        nbytes = self.L.irt_buffer_bytes(self.ctx, BUFFERS[name])
        if data.nbytes != nbytes:
            raise ValueError(f"{name}: expected {nbytes} bytes, got {data.nbytes}")
        check(self.L.irt_write(self.ctx, BUFFERS[name], ptr(data), nbytes), self.ctx)

    # -- timing / counters ---------------------------------------------------
    def enable_timing(self, on: bool = True, only: str | None = None):
        """Record pass GPU times (irt_pass_times); `only` names the single pass
        to time (irt_enable_timing_for)."""
        if on and only:
            check(self.L.irt_enable_timing_for(self.ctx, only.encode()), self.ctx)
        else:
            check(self.L.irt_enable_timing(self.ctx, int(on)), self.ctx)
        self.timing = on

    def pass_times(self) -> dict:
        names = (ctypes.c_char * 32 * 32)()
        ms = (ctypes.c_float * 32)()
        n = check(self.L.irt_pass_times(self.ctx, ctypes.cast(names, ctypes.c_void_p), ms, 32), self.ctx)
        out = {}
        for i in range(n):
            key = bytes(names[i]).split(b"\0", 1)[0].decode()
            out[key] = out.get(key, 0.0) + float(ms[i])
        return out

    def enable_counters(self, on: bool = True):
        check(self.L.irt_enable_counters(self.ctx, int(on)), self.ctx)

    def counters(self) -> dict:
        c = (ctypes.c_uint64 * 7)()
        check(self.L.irt_read_counters(self.ctx, c), self.ctx)
        keys = ("brdf", "node", "obj_box", "obj_grid", "step", "tap", "march")
        return {k: int(c[i]) for i, k in enumerate(keys)}

    # -- reference test hooks (renderer.rs:457-712) ----------------------------
    def history_max_age(self) -> float:
        return float(self.read("moments")[..., 2].max())

    def history_mature_fraction(self, minimum_age: float) -> float:
        ages = self.read("moments")[..., 2]
        return float((ages >= minimum_age).sum()) / ages.size

    def reservoir_maturity(self, indirect: bool) -> tuple:
        r = self.read("indirect2" if indirect else "direct2")
        return int(r["count"].max()), int(r["age"].max())

    def reservoir_mature_fraction(self, indirect: bool, minimum_age: int) -> float:
        r = self.read("indirect2" if indirect else "direct2")
        return float((r["age"] >= minimum_age).sum()) / r.size

    def center_radiance(self) -> float:
        v = half_to_float(self.read("lighting"))
        return float(v[self.H // 2, self.W // 2, 0])

    def close(self):
        if getattr(self, "ctx", None):
            self.L.irt_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_KINDS = ("guide", "reservoirs_a", "spatial_setup", "resolve_combine", "svgf_temporal", "svgf_variance",
               "svgf_atrous", "history")
TILE_SHAPES = ("8x8", "16x8", "16x16")


class Batch:
    """Many views rendered by one launch per pass (irt_batch_*).

    Each view is a ``Renderer`` (its own device state, fused arithmetic);
    ``encode`` performs ``begin_frame`` on every view, one batched frame and
    the per-view bookkeeping, so each view's buffers afterwards hold what its
    own ``encode`` would have produced.
    """

    def __init__(self, renderers: list[Renderer], graph: bool = False):
        L = _lib.irt()
        self.L = L
        self.views = list(renderers)
        arr = (ctypes.c_void_p * len(self.views))(*[r.ctx.value for r in self.views])
        b = ctypes.c_void_p()
        check(L.irt_batch_create(arr, len(self.views), ctypes.byref(b)))
        self.b = b
        self.use_graph(graph)

    def _check(self, rc: int) -> int:
        if rc < 0:
            msg = self.L.irt_batch_last_error(self.b) or b""
            raise _lib.IrtError(rc, msg.decode())
        return rc

    def use_graph(self, on: bool = True):
        self._check(self.L.irt_batch_use_graph(self.b, int(on)))

    def set_tile(self, kind: str | int, shape: str | int):
        k = BATCH_KINDS.index(kind) if isinstance(kind, str) else kind
        s = TILE_SHAPES.index(shape) if isinstance(shape, str) else shape
        self._check(self.L.irt_batch_set_tile(self.b, k, s))

    def begin(self, scenes, inputs):
        for r, sc, inp in zip(self.views, scenes, inputs):
            r.begin_frame(sc, inp)

    def finish(self):
        for r in self.views:
            r.finish_frame()

    def encode(self, scenes, inputs):
        """One frame of every view; ``scenes``/``inputs`` are per view."""
        self.begin(scenes, inputs)
        self.encode_current()
        self.finish()

    def encode_current(self):
        """One batched frame from the inputs already set on every view."""
        self._check(self.L.irt_batch_encode(self.b))

    def encode_timed(self, scenes, inputs) -> dict:
        self.begin(scenes, inputs)
        t = self.encode_timed_current()
        self.finish()
        return t

    def encode_timed_current(self) -> dict:
        ms = (ctypes.c_float * 9)()
        self._check(self.L.irt_batch_encode_timed(self.b, ms))
        return {**{k: float(ms[i]) for i, k in enumerate(BATCH_KINDS)}, "march": float(ms[8])}

    def autotune(self, scenes, inputs, repeats: int = 2) -> dict:
        """Picks a tile shape per pass kind; renders 3 * repeats frames of
        every view from the given inputs (frame bookkeeping advances once)."""
        self.begin(scenes, inputs)
        picks = self.autotune_current(repeats)
        self.finish()
        return picks

    def autotune_current(self, repeats: int = 2) -> dict:
        out = (ctypes.c_uint32 * 8)()
        self._check(self.L.irt_batch_autotune(self.b, repeats, out))
        return {k: TILE_SHAPES[out[i]] for i, k in enumerate(BATCH_KINDS)}

    def stream(self):
        """The lighting stream (a step's first work)."""
        return self.L.irt_batch_stream(self.b)

    def stream_svgf(self):
        """The SVGF stream (a step's last work)."""
        return self.L.irt_batch_stream_svgf(self.b)

    def sync(self):
        self._check(self.L.irt_batch_sync(self.b))

    def close(self):
        if getattr(self, "b", None):
            self.L.irt_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def half_to_float(a: np.ndarray) -> np.ndarray:
    """binary16 bit patterns -> float32."""
    return np.ascontiguousarray(a, dtype=np.uint16).view(np.float16).astype(np.float32)


def float_to_half(a: np.ndarray) -> np.ndarray:
    """float -> binary16 bit patterns (round to nearest even)."""
    return np.asarray(a, dtype=np.float32).astype(np.float16).view(np.uint16)
