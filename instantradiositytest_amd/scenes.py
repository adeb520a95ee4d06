"""Synthetic scenes and per-frame inputs of the path.

The reference renders Blender exports; here a seeded procedural generator
builds scenes of the same record formats (meshes with SDF grids on the
exporter's convention, objects, point lights), and every frame's inputs (an
analytically ray-cast G-buffer and a VPL pool) come from libirt_scenegen.so.

Configurations (sizes of the plain, precision, batched and serving runs):
  plain-640x400   : 4x3x5 m room, 6 meshes (3 wall slabs, box x2, sphere,
                    torus) at 64^3, 4 point lights, 300 VPLs (100 paths x 3
                    bounces), orbit 2 deg/frame, 64 frames
  plain-1920x1200 : same generator, 24 meshes at 128^3, 16 lights, 1000 VPLs
                    (250 x 4), orbit + 0.5 m dolly, 128 frames
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import _lib
from .types import (CAMERA_DT, LIGHT_DT, MATERIAL_DT, MESH_DT, NODE_DT, OBJECT_DT, SG_BOX, SG_SPHERE, SG_TORUS,
                    SGOBJECT_DT, SHAPE_DT, VERTEX_DT, VPL_DT, make_settings, ptr)

ROOM_HALF = (2.0, 1.5, 2.5)  # x half-width, half-height, z half-depth of the 4x3x5 m room
ROOM_CENTER_Y = 1.5


def quat_y(angle: float) -> tuple:
    return (0.0, math.sin(angle / 2), 0.0, math.cos(angle / 2))


def quat_from_matrix(m: np.ndarray) -> tuple:
    """Quaternion (x, y, z, w) of a 3x3 rotation matrix (columns = basis)."""
    t = m[0, 0] + m[1, 1] + m[2, 2]
    if t > 0:
        s = math.sqrt(t + 1.0) * 2
        w, x, y, z = 0.25 * s, (m[2, 1] - m[1, 2]) / s, (m[0, 2] - m[2, 0]) / s, (m[1, 0] - m[0, 1]) / s
    elif m[0, 0] > m[1, 1] and m[0, 0] > m[2, 2]:
        s = math.sqrt(1.0 + m[0, 0] - m[1, 1] - m[2, 2]) * 2
        w, x, y, z = (m[2, 1] - m[1, 2]) / s, 0.25 * s, (m[0, 1] + m[1, 0]) / s, (m[0, 2] + m[2, 0]) / s
    elif m[1, 1] > m[2, 2]:
        s = math.sqrt(1.0 + m[1, 1] - m[0, 0] - m[2, 2]) * 2
        w, x, y, z = (m[0, 2] - m[2, 0]) / s, (m[0, 1] + m[1, 0]) / s, 0.25 * s, (m[1, 2] + m[2, 1]) / s
    else:
        s = math.sqrt(1.0 + m[2, 2] - m[0, 0] - m[1, 1]) * 2
        w, x, y, z = (m[1, 0] - m[0, 1]) / s, (m[0, 2] + m[2, 0]) / s, (m[1, 2] + m[2, 1]) / s, 0.25 * s
    return (x, y, z, w)


def look_at_camera(eye, target, fov_y=1.0471976, near=0.05, far=30.0) -> np.ndarray:
    """Camera record looking from eye to target (local -Z forward, +Y up)."""
    eye = np.asarray(eye, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - eye
    f /= np.linalg.norm(f)
    right = np.cross(f, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, f)
    rot = np.stack([right, up, -f], axis=1)
    cam = np.zeros((), dtype=CAMERA_DT)
    cam["position"] = eye
    cam["rotation"] = quat_from_matrix(rot)
    cam["fov_y"] = fov_y
    cam["near_plane"] = near
    cam["far_plane"] = far
    return cam


@dataclass
class SceneConfig:
    name: str
    seed: int
    width: int
    height: int
    n_meshes: int
    sdf_res: tuple  # (min, max) grid resolution
    n_lights: int
    vpl_paths: int
    vpl_bounces: int
    frames: int
    orbit_deg: float = 2.0
    dolly: float = 0.0
    flux_scale: float = 1.0

    @property
    def vpl_count(self) -> int:
        return self.vpl_paths * self.vpl_bounces

    def settings(self, **overrides):
        values = dict(render_width=self.width, render_height=self.height, vpl_count=self.vpl_count,
                      vpl_max_bounces=self.vpl_bounces)
        values.update(overrides)
        return make_settings(**values)

    def scaled(self, width: int, height: int, frames: int | None = None) -> "SceneConfig":
        return replace(self, width=width, height=height, frames=frames if frames is not None else self.frames)


CONFIGS = {
    "plain-640x400": SceneConfig("plain-640x400", 0xC1, 640, 400, 6, (64, 64), 4, 100, 3, 64),
    "plain-1920x1200": SceneConfig("plain-1920x1200", 0xC2, 1920, 1200, 24, (128, 128), 16, 250, 4, 128, dolly=0.5),
}


def place_prop(rng, rad: float, placed: list, hx: float, hy: float, hz: float):
    """A free spot for a prop of bounding radius `rad` (4000 tries): the
    centre zone, the ring near the walls or the upper band; None if none."""
    for _ in range(4000):
        zone = rng.integers(0, 3)
        x = rng.uniform(-hx + rad, hx - rad)
        z = rng.uniform(-hz + rad, hz - rad)
        r = math.hypot(x, z)
        if zone == 2:
            y = rng.uniform(1.95 + rad, 2 * hy - rad) if 1.95 + rad < 2 * hy - rad else None
        else:
            if zone == 0 and r > 0.95 - rad:
                continue
            if zone == 1 and r < 1.95 + rad:
                continue
            y = rad if rng.uniform() < 0.6 else rng.uniform(rad, 2 * hy - rad)
        if y is None:
            continue
        if any(math.dist((x, y, z), q[:3]) < rad + q[3] + 0.02 for q in placed):
            continue
        return (x, y, z)
    return None


def build_room(cfg: SceneConfig):
    """Seeded room: 3 wall-slab meshes instanced twice each, then props.

    Returns (shapes, objects, lights) record arrays."""
    rng = np.random.default_rng(cfg.seed)
    hx, hy, hz = ROOM_HALF
    t = 0.05
    shapes = [(SG_BOX, (hx + 2 * t, t, hz + 2 * t)),  # floor / ceiling
              (SG_BOX, (t, hy + 2 * t, hz + 2 * t)),  # side walls
              (SG_BOX, (hx + 2 * t, hy + 2 * t, t))]  # end walls
    kinds = [SG_BOX, SG_SPHERE, SG_TORUS]
    for m in range(3, cfg.n_meshes):
        k = kinds[(m - 3) % 3]
        if k == SG_BOX:
            shapes.append((k, tuple(rng.uniform(0.12, 0.3, 3))))
        elif k == SG_SPHERE:
            shapes.append((k, (rng.uniform(0.12, 0.3), 0.0, 0.0)))
        else:
            shapes.append((k, (rng.uniform(0.15, 0.25), rng.uniform(0.04, 0.08), 0.0)))
    lo_res, hi_res = cfg.sdf_res
    S = np.zeros(len(shapes), dtype=SHAPE_DT)
    for i, (k, p) in enumerate(shapes):
        S[i]["kind"] = k
        S[i]["p"] = p
        S[i]["resolution"] = lo_res if lo_res == hi_res else int(rng.integers(lo_res, hi_res + 1))
    # objects
    objs = []

    def wall(mesh, pos):
        objs.append(dict(mesh=mesh, position=pos, rotation=(0, 0, 0, 1), scale=1.0,
                         albedo=tuple(rng.uniform(0.1, 0.9, 3)), rough=0.6, metal=0.0))

    wall(0, (0.0, -t, 0.0))
    wall(0, (0.0, 2 * hy + t, 0.0))
    wall(1, (-hx - t, hy, 0.0))
    wall(1, (hx + t, hy, 0.0))
    wall(2, (0.0, hy, -hz - t))
    wall(2, (0.0, hy, hz + t))
    placed = []  # (x, y, z, radius)
    prop_meshes = list(range(3, cfg.n_meshes))
    if prop_meshes:
        prop_meshes.insert(1, 3)  # the first box is instanced twice
    for mesh in prop_meshes:
        k, p = shapes[mesh]
        scale = float(rng.uniform(0.8, 1.2))
        if k == SG_BOX:
            extent = math.sqrt(sum(v * v for v in p))
        elif k == SG_SPHERE:
            extent = p[0]
        else:
            extent = p[0] + p[1]
        # crowded rooms (many props): shrink a prop that finds no free spot
        for _shrink in range(12):
            rad = extent * scale
            spot = place_prop(rng, rad, placed, hx, hy, hz)
            if spot is not None:
                break
            scale *= 0.8
        else:
            raise RuntimeError(f"could not place prop mesh {mesh}")
        x, y, z = spot
        placed.append((x, y, z, rad))
        metal = 1.0 if rng.uniform() < 0.2 else 0.0
        objs.append(dict(mesh=mesh, position=(x, y, z), rotation=quat_y(rng.uniform(0, 2 * math.pi)), scale=scale,
                         albedo=tuple(rng.uniform(0.1, 0.9, 3)), rough=float(rng.uniform(0.3, 0.8)), metal=metal))
    O = np.zeros(len(objs), dtype=SGOBJECT_DT)
    for i, o in enumerate(objs):
        for key, v in o.items():
            O[i][key] = v
    # point lights, inside the room and away from props
    L = np.zeros(cfg.n_lights, dtype=LIGHT_DT)
    for i in range(cfg.n_lights):
        for _ in range(4000):
            pos = (rng.uniform(-hx + 0.3, hx - 0.3), rng.uniform(0.5, 2 * hy - 0.3), rng.uniform(-hz + 0.3, hz - 0.3))
            if all(math.dist(pos, q[:3]) > q[3] + 0.1 for q in placed):
                break
        flux = rng.uniform(5.0, 50.0, 3)
        L[i]["position_strength"] = (*pos, 1.0 / (4.0 * math.pi))
        L[i]["direction_angle"] = (0.0, -1.0, 0.0, 0.0)
        L[i]["color_type"] = (*flux, 0.0)
    return S, O, L


def orbit_camera(cfg: SceneConfig, frame: int) -> np.ndarray:
    """Camera orbiting the room centre at cfg.orbit_deg per frame; the dolly
    moves it inward by up to cfg.dolly metres over the run."""
    angle = math.radians(cfg.orbit_deg * frame) + 0.3
    radius = 1.5 - cfg.dolly * (frame / max(cfg.frames - 1, 1))
    eye = (radius * math.cos(angle), 1.6, radius * math.sin(angle))
    target = (-0.6 * math.cos(angle), 1.0, -0.6 * math.sin(angle))
    return look_at_camera(eye, target)


@dataclass
class FrameInputs:
    """What the reference's raster + VPL passes hand to the lighting path."""
    depth: np.ndarray   # float32 [H, W]
    normal: np.ndarray  # uint16 [H, W, 4] binary16
    color: np.ndarray   # uint16 [H, W, 4]
    params: np.ndarray  # uint16 [H, W, 2]
    motion: np.ndarray  # uint16 [H, W, 2]
    vpls: np.ndarray    # VPL_DT [V]
    vp: np.ndarray      # float32 [16]
    camera: np.ndarray  # float32 [3]


class Scene:
    """A built synthetic scene (records + SDF atlas) with a generation counter."""

    def __init__(self, shapes: np.ndarray, objects: np.ndarray, lights: np.ndarray):
        self._sg = _lib.scenegen()
        self.shapes = np.ascontiguousarray(shapes, dtype=SHAPE_DT)
        self.sg_objects = np.ascontiguousarray(objects, dtype=SGOBJECT_DT)
        self.lights = np.ascontiguousarray(lights, dtype=LIGHT_DT)
        self.handle = self._sg.sg_create(ptr(self.shapes), len(self.shapes), ptr(self.sg_objects),
                                         len(self.sg_objects), ptr(self.lights), len(self.lights))
        if not self.handle:
            raise ValueError("invalid scene description")
        self.generation = 0
        self._refresh()

    def _refresh(self):
        info = np.zeros(8, dtype=np.uint32)
        self._sg.sg_info(self.handle, ptr(info))
        n_mesh, n_obj, n_node, n_light = (int(v) for v in info[:4])
        self.atlas_dims = tuple(int(v) for v in info[4:7])
        self.meshes = np.zeros(n_mesh, dtype=MESH_DT)
        self.objects = np.zeros(n_obj, dtype=OBJECT_DT)
        self.nodes = np.zeros(n_node, dtype=NODE_DT)
        lights = np.zeros(max(n_light, 1), dtype=LIGHT_DT)
        self._sg.sg_records(self.handle, ptr(self.meshes), ptr(self.objects), ptr(self.nodes), ptr(lights))
        self.lights = lights[:n_light].copy()

    def atlas(self) -> np.ndarray:
        w, h, d = self.atlas_dims
        a = np.empty((d, h, w), dtype=np.float32)
        self._sg.sg_atlas(self.handle, ptr(a))
        return a

    def changed(self):
        """Scene::changed (src/scene.rs:383-385): bump the generation, never 0."""
        self.generation = max((self.generation + 1) & 0xFFFFFFFF, 1)

    def set_objects(self, objects: np.ndarray):
        self.sg_objects = np.ascontiguousarray(objects, dtype=SGOBJECT_DT)
        if self._sg.sg_set_objects(self.handle, ptr(self.sg_objects), len(self.sg_objects)) != 0:
            raise ValueError("invalid objects")
        self._refresh()

    def set_lights(self, lights: np.ndarray):
        lights = np.ascontiguousarray(lights, dtype=LIGHT_DT)
        self._sg.sg_set_lights(self.handle, ptr(lights), len(lights))
        self._refresh()

    def camera_matrix(self, camera: np.ndarray, aspect: float) -> np.ndarray:
        vp = np.zeros(16, dtype=np.float32)
        cam = np.ascontiguousarray(camera, dtype=CAMERA_DT)
        self._sg.sg_camera_matrix(ptr(cam), ctypes.c_float(aspect), ptr(vp))
        return vp

    def gbuffer(self, camera, prev_camera, width: int, height: int):
        cam = np.ascontiguousarray(camera, dtype=CAMERA_DT)
        pcam = np.ascontiguousarray(prev_camera, dtype=CAMERA_DT)
        depth = np.empty((height, width), dtype=np.float32)
        normal = np.empty((height, width, 4), dtype=np.uint16)
        color = np.empty((height, width, 4), dtype=np.uint16)
        params = np.empty((height, width, 2), dtype=np.uint16)
        motion = np.empty((height, width, 2), dtype=np.uint16)
        self._sg.sg_render_gbuffer(self.handle, ptr(cam), ptr(pcam), width, height, ptr(depth), ptr(normal),
                                   ptr(color), ptr(params), ptr(motion))
        return depth, normal, color, params, motion

    def geometry(self):
        """Triangle geometry in the reference's records: (vertices VERTEX_DT, mesh-local
        indices uint32, materials MATERIAL_DT); object i uses material i."""
        info = np.zeros(3, dtype=np.uint32)
        self._sg.sg_geometry_info(self.handle, ptr(info))
        vertices = np.zeros(int(info[0]), dtype=VERTEX_DT)
        indices = np.zeros(int(info[1]), dtype=np.uint32)
        materials = np.zeros(int(info[2]), dtype=MATERIAL_DT)
        self._sg.sg_geometry(self.handle, ptr(vertices), ptr(indices), ptr(materials))
        return vertices, indices, materials

    def vpl_pool(self, seed: int, paths: int, bounces: int, flux_scale: float = 1.0) -> np.ndarray:
        out = np.zeros(paths * bounces, dtype=VPL_DT)
        self._sg.sg_vpls(self.handle, ctypes.c_uint64(seed & 0xFFFFFFFFFFFFFFFF), paths, bounces,
                         ctypes.c_float(flux_scale), ptr(out))
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self._sg.sg_destroy(h)
            self.handle = None


class ConfigScene:
    """A SceneConfig's scene plus its camera path and per-frame inputs."""

    def __init__(self, cfg: SceneConfig):
        self.cfg = cfg
        shapes, objects, lights = build_room(cfg)
        self.scene = Scene(shapes, objects, lights)

    def camera(self, frame: int) -> np.ndarray:
        return orbit_camera(self.cfg, frame)

    def frame_inputs(self, frame: int) -> FrameInputs:
        cfg = self.cfg
        cam = self.camera(frame)
        prev = self.camera(frame - 1) if frame > 0 else cam
        depth, normal, color, params, motion = self.scene.gbuffer(cam, prev, cfg.width, cfg.height)
        vpls = self.scene.vpl_pool((cfg.seed << 20) ^ (frame * 0x9E3779B1 + 7), cfg.vpl_paths, cfg.vpl_bounces,
                                   cfg.flux_scale)
        vp = self.scene.camera_matrix(cam, cfg.width / cfg.height)
        return FrameInputs(depth, normal, color, params, motion, vpls, vp, np.asarray(cam["position"], np.float32))
