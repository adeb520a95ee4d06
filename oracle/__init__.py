"""Python driver of the CPU restatement (test infrastructure only).

Only tests/, bench.py's cpu_baseline leg and __graft_entry__.smoke() use
this package, and only as the checker.  ``OracleRenderer`` runs the passes
of one frame with the reference's buffer schedule (renderer.rs:1873-1960,
svgf.rs:140-261): triple-buffered reservoirs, current/previous VPL pools plus
the archive ring, SVGF moments/feedback ping-pong, previous depth/normal.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

ORES_DT = np.dtype([
    ("id", "<u4"), ("source", "<u4"), ("count", "<u4"), ("age", "<u4"),
    ("sum", "<f8"), ("selected_target", "<f8"), ("pdf", "<f8"), ("generation", "<u4"), ("pad", "<u4"),
])
assert ORES_DT.itemsize == 48

PASS = {
    "candidates_direct": 0, "candidates_indirect": 1, "temporal_direct": 2, "temporal_indirect": 3,
    "spatial_direct": 4, "spatial_indirect": 5, "resolve_direct": 6, "resolve_indirect": 7,
    "svgf_temporal": 8, "svgf_variance": 9, "svgf_atrous": 10,
}
COUNTER_KEYS = ("brdf", "node", "obj_box", "obj_grid", "step", "tap", "march")


class Scene(ctypes.Structure):
    _fields_ = [("objects", c_void_p), ("meshes", c_void_p), ("nodes", c_void_p), ("lights", c_void_p),
                ("atlas", c_void_p), ("atlas_w", c_uint32), ("atlas_h", c_uint32), ("atlas_d", c_uint32),
                ("n_objects", c_uint32), ("n_meshes", c_uint32), ("n_nodes", c_uint32), ("n_lights", c_uint32)]


class GBuffer(ctypes.Structure):
    _fields_ = [(n, c_void_p) for n in ("depth", "normal", "color", "params", "motion", "previous_depth",
                                        "previous_normal")]


class IO(ctypes.Structure):
    _fields_ = [(n, c_void_p) for n in (
        "direct_in", "indirect_in", "direct_out", "indirect_out", "direct_history", "indirect_history",
        "current_vpls", "previous_vpls", "archived_vpls", "direct_output", "direct_lighting", "output",
        "noisy", "old_radiance", "old_moments", "input_radiance", "moments", "radiance_out", "moments_out")] + [
        ("stride", c_uint32), ("counters", c_void_p), ("row_begin", c_uint32), ("row_end", c_uint32),
        ("threads", c_int)]


_L = None


def build() -> Path:
    """Compile the restatement (make in oracle/)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _L
    if _L is not None:
        return _L
    if not LIB.exists():
        build()
    L = ctypes.CDLL(str(LIB))
    P = c_void_p
    d = c_double
    L.oracle_run_pass.restype = c_int
    L.oracle_run_pass.argtypes = [c_int, c_int, P, P, P, P]
    L.oracle_random.restype = c_float
    L.oracle_random.argtypes = [POINTER(c_uint32)]
    L.oracle_seed.restype = c_uint32
    L.oracle_seed.argtypes = [c_uint32, c_uint32, c_uint32]
    L.oracle_signed_box_distance.restype = d
    L.oracle_signed_box_distance.argtypes = [c_int, P, P, P]
    L.oracle_mesh_distance.restype = d
    L.oracle_mesh_distance.argtypes = [c_int, P, P, P]
    L.oracle_scene_distance.restype = d
    L.oracle_scene_distance.argtypes = [c_int, P, P, P, c_int, c_uint32, POINTER(d), POINTER(c_uint32)]
    L.oracle_visibility.restype = d
    L.oracle_visibility.argtypes = [c_int, P, P, P, P, P, P, P]
    L.oracle_brdf.restype = None
    L.oracle_brdf.argtypes = [c_int, P, P, P, P]
    L.oracle_reservoir_add.restype = c_int
    L.oracle_reservoir_add.argtypes = [c_int, P, c_uint32, c_uint32, d, d, d, c_uint32, c_uint32, d, POINTER(c_uint32)]
    L.oracle_normalization.restype = d
    L.oracle_normalization.argtypes = [c_int, P]
    L.oracle_shadow_step.restype = d
    L.oracle_shadow_step.argtypes = [c_int, d, d, d, d, d]
    L.oracle_neighbor_offsets.restype = None
    L.oracle_neighbor_offsets.argtypes = [c_int, c_uint32, c_uint32, c_uint32, P]
    L.oracle_f32_to_f16.restype = ctypes.c_uint16
    L.oracle_f32_to_f16.argtypes = [c_float]
    L.oracle_f16_to_f32.restype = c_float
    L.oracle_f16_to_f32.argtypes = [ctypes.c_uint16]
    L.oracle_generate_vpls.restype = c_int
    L.oracle_generate_vpls.argtypes = [P, P, P, P, c_int, POINTER(c_uint64)]
    L.oracle_trace.restype = c_int
    L.oracle_trace.argtypes = [P, P, P, P, POINTER(c_float), POINTER(c_float), POINTER(c_float),
                               POINTER(c_uint32), POINTER(c_uint32)]
    L.oracle_render_gbuffer.restype = c_int
    L.oracle_render_gbuffer.argtypes = [P, P, P, P, P, P, P, P]
    L.oracle_bake_sdf.restype = c_int
    L.oracle_bake_sdf.argtypes = [P, c_uint32, c_uint32, P, c_uint32, c_uint32, P, P, P, c_int]
    L.oracle_present.restype = None
    L.oracle_present.argtypes = [P, c_uint32, P]
    L.oracle_sample_bsdf.restype = None
    L.oracle_sample_bsdf.argtypes = [P, P, c_float, c_float, P, POINTER(c_uint32), P, P]
    _L = L
    return L


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return c_void_p(a.ctypes.data)


def half_bits_to_f64(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, np.uint16).view(np.float16).astype(np.float64)


class SceneRecords:
    """Scene records held for the oracle (keeps the arrays alive)."""

    def __init__(self, objects, meshes, nodes, lights, atlas: np.ndarray):
        self.objects = np.ascontiguousarray(objects)
        self.meshes = np.ascontiguousarray(meshes)
        self.nodes = np.ascontiguousarray(nodes)
        self.lights = np.ascontiguousarray(lights) if len(lights) else np.zeros(1, dtype=lights.dtype)
        self.n_lights = len(lights)
        self.atlas = np.ascontiguousarray(atlas, dtype=np.float32)
        d, h, w = self.atlas.shape
        self.c = Scene(_p(self.objects), _p(self.meshes), _p(self.nodes), _p(self.lights), _p(self.atlas), w, h, d,
                       len(self.objects), len(self.meshes), len(self.nodes), self.n_lights)

    @staticmethod
    def from_scene(scene) -> "SceneRecords":
        return SceneRecords(scene.objects, scene.meshes, scene.nodes, scene.lights, scene.atlas())


class Geometry(ctypes.Structure):
    _fields_ = [("vertices", c_void_p), ("indices", c_void_p), ("materials", c_void_p), ("n_vertices", c_uint32),
                ("n_indices", c_uint32), ("n_materials", c_uint32), ("textures", c_void_p),
                ("texture_size", c_uint32), ("texture_layers", c_uint32)]


class GeometryRecords:
    """Triangle geometry of the VPL pass held for the oracle: vertex records
    (64 B: position, normal, tangent, uv), mesh-local uint32 indices and
    material records (64 B)."""

    def __init__(self, vertices, indices, materials, textures: np.ndarray | None = None):
        self.vertices = np.ascontiguousarray(vertices)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.materials = np.ascontiguousarray(materials)
        assert self.vertices.itemsize == 64 and self.materials.itemsize == 64
        # textures: uint8 [layers, size, size, 4]
        self.textures = None if textures is None else np.ascontiguousarray(textures, np.uint8)
        layers, size = (0, 0) if self.textures is None else self.textures.shape[:2]
        self.c = Geometry(_p(self.vertices), _p(self.indices), _p(self.materials), len(self.vertices),
                          len(self.indices), len(self.materials), _p(self.textures), size, layers)


def generate_vpls(frame: np.ndarray, scene: SceneRecords, geometry: GeometryRecords, threads: int = 0,
                  with_rays: bool = False):
    """vpl.wgsl:2-47 for the frame: float32 [counts.y, 12] records (position,
    normal, flux), and the number of rays traced when `with_rays` is set."""
    f = np.ascontiguousarray(frame)
    n = int(np.frombuffer(f.tobytes(), np.uint32)[73])  # counts.y (counts at byte 288)
    out = np.zeros((n, 12), dtype=np.float32)
    rays = c_uint64()
    rc = lib().oracle_generate_vpls(_p(f), ctypes.byref(scene.c), ctypes.byref(geometry.c), _p(out), threads,
                                    ctypes.byref(rays))
    if rc != 0:
        raise ValueError("oracle_generate_vpls: bad arguments")
    return (out, rays.value) if with_rays else out


def render_gbuffer(frame: np.ndarray, scene: SceneRecords, geometry: GeometryRecords):
    """gbuffer.wgsl for the frame: (depth float32 [H, W], normal/color uint16
    [H, W, 4], params/motion uint16 [H, W, 2]), binary16 bit patterns."""
    f = np.ascontiguousarray(frame)
    words = np.frombuffer(f.tobytes(), np.uint32)
    W, H = int(words[68]), int(words[69])  # size.xy (size at byte 272)
    depth = np.zeros((H, W), np.float32)
    normal = np.zeros((H, W, 4), np.uint16)
    color = np.zeros((H, W, 4), np.uint16)
    params = np.zeros((H, W, 2), np.uint16)
    motion = np.zeros((H, W, 2), np.uint16)
    rc = lib().oracle_render_gbuffer(_p(f), ctypes.byref(scene.c), ctypes.byref(geometry.c), _p(depth), _p(normal),
                                     _p(color), _p(params), _p(motion))
    if rc != 0:
        raise ValueError("oracle_render_gbuffer: bad arguments")
    return depth, normal, color, params, motion


def bake_sdf(positions: np.ndarray, indices: np.ndarray, res: int, threads: int = 0):
    """export_scene.py:121-158 for a mesh: (samples float32 [res, res, res] as
    [z, y, x], bounds lo, hi)."""
    pos = np.ascontiguousarray(positions, np.float32)
    idx = np.ascontiguousarray(indices, np.uint32)
    out = np.zeros((res, res, res), np.float32)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    rc = lib().oracle_bake_sdf(_p(pos), pos.shape[1], len(pos), _p(idx), len(idx), res, _p(out), _p(lo), _p(hi),
                               threads)
    if rc != 0:
        raise ValueError("oracle_bake_sdf: " + ("crossing limit" if rc == -2 else "bad arguments"))
    return out, lo, hi


def present(history_rgba16f: np.ndarray) -> np.ndarray:
    """present.wgsl into the sRGB display: uint16 [H, W, 4] binary16 history
    -> uint8 [H, W, 4]."""
    h = np.ascontiguousarray(history_rgba16f, np.uint16)
    out = np.zeros(h.shape[:-1] + (4,), np.uint8)
    lib().oracle_present(_p(h), h.size // 4, _p(out))
    return out


def trace(scene: SceneRecords, geometry: GeometryRecords, origin, direction):
    """Closest hit of one world-space ray: (hit, t, u, v, object, prim)."""
    o = np.ascontiguousarray(origin, np.float32)
    d = np.ascontiguousarray(direction, np.float32)
    t, u, v = c_float(), c_float(), c_float()
    obj, prim = c_uint32(), c_uint32()
    hit = lib().oracle_trace(ctypes.byref(scene.c), ctypes.byref(geometry.c), _p(o), _p(d), ctypes.byref(t),
                             ctypes.byref(u), ctypes.byref(v), ctypes.byref(obj), ctypes.byref(prim))
    return bool(hit), t.value, u.value, v.value, obj.value, prim.value


def sample_bsdf(n, color, rough: float, metal: float, wo, state: int):
    """sample_bsdf (common.wgsl:86-97): (direction, weight, next PCG state)."""
    nn = np.ascontiguousarray(n, np.float32)
    cc = np.ascontiguousarray(color, np.float32)
    w = np.ascontiguousarray(wo, np.float32)
    st = c_uint32(state & 0xFFFFFFFF)
    direction = np.zeros(3, np.float32)
    weight = np.zeros(3, np.float32)
    lib().oracle_sample_bsdf(_p(nn), _p(cc), rough, metal, _p(w), ctypes.byref(st), _p(direction), _p(weight))
    return direction, weight, st.value


def neighbor_offsets(bits: int, state: int, spatial_radius: int, count: int) -> np.ndarray:
    """[count, 2] spatial-reuse neighbour offsets drawn in sequence from one PCG state."""
    out = np.zeros((count, 2), np.int32)
    lib().oracle_neighbor_offsets(bits, state, spatial_radius, count, _p(out))
    return out


class GBufferPlanes:
    """Current and previous G-buffer planes as float64."""

    def __init__(self, depth, normal, color, params, motion, previous_depth, previous_normal):
        H, W = depth.shape
        self.depth = np.ascontiguousarray(depth, np.float64)
        self.normal = np.ascontiguousarray(normal, np.float64).reshape(H, W, 4)
        self.color = np.ascontiguousarray(color, np.float64).reshape(H, W, 4)
        self.params = np.ascontiguousarray(params, np.float64).reshape(H, W, 2)
        self.motion = np.ascontiguousarray(motion, np.float64).reshape(H, W, 2)
        self.previous_depth = np.ascontiguousarray(previous_depth, np.float64)
        self.previous_normal = np.ascontiguousarray(previous_normal, np.float64).reshape(H, W, 4)
        self.c = GBuffer(*(_p(a) for a in (self.depth, self.normal, self.color, self.params, self.motion,
                                           self.previous_depth, self.previous_normal)))


def run_pass(bits: int, name: str, frame: np.ndarray, scene: SceneRecords | None, gb: GBufferPlanes,
             threads: int = 0, rows=None, counters: np.ndarray | None = None, **bufs) -> None:
    """Run one pass over the frame.  ``bufs`` are the oracle_io fields by name."""
    io = IO()
    keep = []
    for k, v in bufs.items():
        if k == "stride":
            io.stride = int(v)
        else:
            arr = np.ascontiguousarray(v)
            if arr is not v:
                raise ValueError(f"{k} must be C-contiguous")
            keep.append(arr)
            setattr(io, k, _p(arr))
    if counters is not None:
        io.counters = _p(counters)
    if rows is not None:
        io.row_begin, io.row_end = rows
    io.threads = threads
    f = np.ascontiguousarray(frame)
    rc = lib().oracle_run_pass(bits, PASS[name], _p(f), ctypes.byref(scene.c) if scene else None,
                               ctypes.byref(gb.c), ctypes.byref(io))
    if rc != 0:
        raise RuntimeError(f"oracle pass {name} failed")


class OracleRenderer:
    """The reference's frame schedule on the CPU restatement.

    State mirrors renderer.rs / svgf.rs resources: direct[3], indirect[3],
    vpls[2] + archive ring, direct_lighting, lighting, moments[2],
    feedback[2], temporal, scratch[2], histories[2], previous depth/normal."""

    def __init__(self, bits: int, W: int, H: int, vpl_count: int, restir_max_history: int, threads: int = 0):
        self.bits, self.W, self.H, self.V = bits, W, H, vpl_count
        self.slots = restir_max_history + 1
        self.threads = threads
        n = W * H
        self.direct = [np.zeros(n, ORES_DT) for _ in range(3)]
        self.indirect = [np.zeros(n, ORES_DT) for _ in range(3)]
        vdt = np.dtype([("position", "<f4", (4,)), ("normal", "<f4", (4,)), ("flux", "<f4", (4,))])
        self.vpls = [np.zeros(vpl_count, vdt), np.zeros(vpl_count, vdt)]
        self.archive = np.zeros(self.slots * vpl_count, vdt)
        z4 = lambda: np.zeros((H, W, 4), np.float64)
        self.direct_lighting = z4()
        self.lighting = z4()
        self.moments = [z4(), z4()]
        self.feedback = [z4(), z4()]
        self.temporal = z4()
        self.scratch = [z4(), z4()]
        self.histories = [z4(), z4()]
        self.previous_depth = np.ones((H, W), np.float64)
        self.previous_normal = np.zeros((H, W, 4), np.float64)
        self.counters = np.zeros(7, np.uint64)

    def gbuffer(self, depth, normal_bits, color_bits, params_bits, motion_bits) -> GBufferPlanes:
        return GBufferPlanes(depth, half_bits_to_f64(normal_bits), half_bits_to_f64(color_bits),
                             half_bits_to_f64(params_bits), half_bits_to_f64(motion_bits), self.previous_depth,
                             self.previous_normal)

    def _p(self, name, frame, scene, gb, **bufs):
        run_pass(self.bits, name, frame, scene, gb, threads=self.threads, counters=self.counters, **bufs)

    def set_vpls(self, frame_index: int, vpls: np.ndarray):
        """VPL pool of this frame into vpls[current] and the archive slot (renderer.rs:1873-1884)."""
        cur = frame_index % 2
        self.vpls[cur][:] = vpls
        slot = frame_index % self.slots
        self.archive[slot * self.V:(slot + 1) * self.V] = vpls

    def lighting_passes(self, frame, scene, gb, passes=None):
        f = frame
        fi = int(f["size"][0][2])
        cur, old = fi % 2, 1 - fi % 2
        V = dict(current_vpls=self.vpls[cur], previous_vpls=self.vpls[old], archived_vpls=self.archive)
        sched = [
            ("candidates_direct", 2, 0, 2), ("candidates_indirect", 2, 0, 2),
            ("temporal_direct", 0, 1, 2), ("temporal_indirect", 0, 1, 2),
            ("spatial_direct", 1, 2, 0), ("spatial_indirect", 1, 2, 0),
        ]
        for name, i, o, h in sched:
            if passes is not None and name not in passes:
                continue
            self._p(name, f, scene, gb, direct_in=self.direct[i], indirect_in=self.indirect[i],
                    direct_out=self.direct[o], indirect_out=self.indirect[o], direct_history=self.direct[h],
                    indirect_history=self.indirect[h], **V)
        if passes is None or "resolve_direct" in passes:
            self._p("resolve_direct", f, scene, gb, direct_in=self.direct[2], indirect_in=self.indirect[2],
                    direct_output=self.direct_lighting, **V)
        if passes is None or "resolve_indirect" in passes:
            self._p("resolve_indirect", f, scene, gb, direct_in=self.direct[2], indirect_in=self.indirect[2],
                    direct_lighting=self.direct_lighting, output=self.lighting, **V)

    def svgf(self, frame, gb, iterations: int, debug: int):
        fi = int(frame["size"][0][2])
        cur, old = fi % 2, 1 - fi % 2
        if debug >= 3:
            self.histories[cur][:] = self.lighting
            return
        self._p("svgf_temporal", frame, None, gb, noisy=self.lighting, old_radiance=self.feedback[old],
                old_moments=self.moments[old], input_radiance=self.feedback[old], moments=self.moments[old],
                radiance_out=self.temporal, moments_out=self.moments[cur])
        self._p("svgf_variance", frame, None, gb, noisy=self.lighting, input_radiance=self.temporal,
                moments=self.moments[cur], radiance_out=self.scratch[0])
        for i in range(iterations):
            self._p("svgf_atrous", frame, None, gb, noisy=self.lighting, input_radiance=self.scratch[i % 2],
                    moments=self.moments[cur], radiance_out=self.scratch[1 - i % 2], stride=1 << i)
            if i == 0:
                self.feedback[cur][:] = self.scratch[1]
        self.histories[cur][:] = self.scratch[iterations % 2]

    def end_frame(self, gb: GBufferPlanes):
        self.previous_depth = gb.depth.copy()
        self.previous_normal = gb.normal.copy()

    def encode(self, frame, scene, inputs, iterations: int, debug: int) -> GBufferPlanes:
        fi = int(frame["size"][0][2])
        self.set_vpls(fi, inputs.vpls)
        gb = self.gbuffer(inputs.depth, inputs.normal, inputs.color, inputs.params, inputs.motion)
        self.lighting_passes(frame, scene, gb)
        self.svgf(frame, gb, iterations, debug)
        self.end_frame(gb)
        return gb
