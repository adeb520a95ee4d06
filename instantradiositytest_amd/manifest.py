"""Scenes in the reference's exported format: a JSON manifest plus binary mesh
and SDF payloads (src/scene.rs:128-260, written by tools/export_scene.py).

``ManifestScene.load(path)`` reads ``scene.json`` and builds the records the
reference uploads (src/renderer.rs): the GpuMesh table with its SDF atlas
placement, box certification, Lipschitz bound and enclosing sphere
(renderer.rs:776-888), the GpuObject records (renderer.rs:359-384), the
object hierarchy (renderer.rs:385-455), GpuLight (renderer.rs:1743-1755) and
GpuMaterial (renderer.rs:950-982) records, and the vertex and index buffers
the G-buffer and VPL passes read.  The host precompute runs in the library
(irt_host_*), the camera matrix in the scene generator's implementation of
Camera::matrix (scene.rs:118-125).

Payloads are read from the files the manifest names, as Scene::load does, or
from a ``payload.npz`` beside the manifest that holds the same arrays
(mesh<i>_vertices / mesh<i>_indices, or, for a flat-shaded mesh listed
triangle by triangle, mesh<i>_positions / _corners / _triangle_normals /
_tangent; mesh<i>_sdf).  The object works as a scene for ``Renderer``:
meshes, objects, nodes, lights, atlas(), geometry(), generation.
"""
from __future__ import annotations

import ctypes
import json
import math
from pathlib import Path

import numpy as np

from . import _lib
from .types import CAMERA_DT, LIGHT_DT, MATERIAL_DT, MESH_DT, NODE_DT, OBJECT_DT, VERTEX_DT, ptr

F32 = np.float32


def _expand_flat(positions, corners, normals, tangent) -> np.ndarray:
    v = np.zeros((len(corners), 16), F32)
    v[:, :3] = positions[corners]
    v[:, 3] = 1.0
    v[:, 4:7] = np.repeat(normals, 3, axis=0)
    v[:, 8:12] = tangent
    return v


def _quat_matrix(q) -> np.ndarray:
    """3x3 rotation of a unit quaternion (x, y, z, w) in float32, columns = axes."""
    x, y, z, w = (F32(c) for c in q)
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz = x * x2, x * y2, x * z2
    yy, yz, zz = y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    one = F32(1.0)
    return np.array([[one - (yy + zz), xy - wz, xz + wy],
                     [xy + wz, one - (xx + zz), yz - wx],
                     [xz - wy, yz + wx, one - (xx + yy)]], F32)


def transform_matrix(t: dict) -> np.ndarray:
    """Transform::matrix (scene.rs:19-25): scale, normalised rotation, translation; column-major 16 floats."""
    q = np.asarray(t["rotation"], F32)
    q = (q / F32(math.sqrt(float(np.dot(q.astype(np.float64), q.astype(np.float64)))))).astype(F32)
    m = np.eye(4, dtype=F32)
    m[:3, :3] = _quat_matrix(q) * F32(t["scale"])
    m[:3, 3] = np.asarray(t["position"], F32)
    return np.ascontiguousarray(m.T).reshape(16)


class ManifestScene:
    """A manifest scene and the reference's GPU records of it."""

    def __init__(self, manifest: dict, vertices: list, indices: list, fields: list, root: Path | None = None):
        self.manifest = manifest
        self.root = root
        self.mesh_vertices = vertices   # per mesh: float32 [V, 16] vertex records
        self.mesh_indices = indices     # per mesh: uint32 mesh-local indices
        self.fields = fields            # per mesh: float32 res^3 samples, x fastest
        self.generation = 1             # Scene::load starts at 1 (scene.rs:194)
        self._check()
        self.mesh_ids = {m["id"]: i for i, m in enumerate(manifest["meshes"])}
        self.material_ids = {m["id"]: i for i, m in enumerate(manifest["materials"])}
        self._build_meshes()
        self._models = [transform_matrix(o["transform"]) for o in manifest["objects"]]
        self._previous = [m.copy() for m in self._models]
        self._build_objects()
        self._build_lights()

    # ------------------------------------------------------------------ load
    @staticmethod
    def load(path) -> "ManifestScene":
        path = Path(path)
        manifest = json.loads(path.read_text())
        npz = path.parent / "payload.npz"
        vertices, indices, fields = [], [], []
        if npz.exists():
            with np.load(npz, allow_pickle=False) as z:
                for i, _ in enumerate(manifest["meshes"]):
                    if f"mesh{i}_vertices" in z:
                        vertices.append(z[f"mesh{i}_vertices"].astype(F32))
                        indices.append(z[f"mesh{i}_indices"].astype(np.uint32))
                    else:
                        v = _expand_flat(z[f"mesh{i}_positions"], z[f"mesh{i}_corners"],
                                         z[f"mesh{i}_triangle_normals"], z[f"mesh{i}_tangent"])
                        vertices.append(v)
                        indices.append(np.arange(len(v), dtype=np.uint32))
                    fields.append(z[f"mesh{i}_sdf"].astype(F32))
        else:
            blobs = {}
            for m in manifest["meshes"]:
                g, s = m["geometry"], m["sdf"]
                geo = blobs.setdefault(g["path"], np.fromfile(path.parent / g["path"], np.uint8))
                vb = geo[g["vertex_offset"]:g["vertex_offset"] + 64 * g["vertex_count"]]
                ib = geo[g["index_offset"]:g["index_offset"] + 4 * g["index_count"]]
                if len(vb) != 64 * g["vertex_count"] or len(ib) != 4 * g["index_count"]:
                    raise ValueError(f"mesh {m['id']}: binary range exceeds {g['path']}")
                vertices.append(vb.view(F32).reshape(-1, 16).copy())
                indices.append(ib.view(np.uint32).copy())
                fld = blobs.setdefault(s["path"], np.fromfile(path.parent / s["path"], np.uint8))
                n = 4 * s["resolution"] ** 3
                fb = fld[s["offset"]:s["offset"] + n]
                if len(fb) != n:
                    raise ValueError(f"mesh {m['id']}: SDF range exceeds {s['path']}")
                fields.append(fb.view(F32).copy())
        return ManifestScene(manifest, vertices, indices, fields, path.parent)

    def _check(self):
        """The payload checks of Scene::load (scene.rs:205-240)."""
        for m, v, i, f in zip(self.manifest["meshes"], self.mesh_vertices, self.mesh_indices, self.fields):
            if len(v) == 0 or len(i) == 0 or len(i) % 3:
                raise ValueError(f"mesh {m['id']}: geometry must have triangles")
            if int(i.max()) >= len(v):
                raise ValueError(f"mesh {m['id']}: index outside vertex range")
            if not np.isfinite(v).all():
                raise ValueError(f"mesh {m['id']}: nonfinite vertex")
            if len(f) != m["sdf"]["resolution"] ** 3 or not np.isfinite(f).all():
                raise ValueError(f"mesh {m['id']}: bad SDF")

    # --------------------------------------------------------------- records
    def _build_meshes(self):
        L = _lib.irt()
        meshes = self.manifest["meshes"]
        n = len(meshes)
        self.tile = max(m["sdf"]["resolution"] for m in meshes)
        grid = np.zeros(3, np.uint32)
        L.irt_host_atlas_grid(n, self.tile, ptr(grid))
        self.atlas_grid = tuple(int(g) for g in grid)
        self.atlas_dims = (self.tile * self.atlas_grid[0], self.tile * self.atlas_grid[1], self.tile * self.atlas_grid[2])
        self.meshes = np.zeros(n, MESH_DT)
        self.is_box = []
        self.lipschitz = []
        vtx = idx = 0
        gx, gy, _ = self.atlas_grid
        for k, (m, v, ind, f) in enumerate(zip(meshes, self.mesh_vertices, self.mesh_indices, self.fields)):
            s = m["sdf"]
            res = int(s["resolution"])
            bmin, bmax = np.asarray(s["bounds_min"], F32), np.asarray(s["bounds_max"], F32)
            spacing = ((bmax - bmin) / F32(res - 1)).astype(F32)
            pos = np.ascontiguousarray(v[:, :3])
            glo, ghi = pos.min(axis=0).astype(F32), pos.max(axis=0).astype(F32)
            center = ((glo + ghi) * F32(0.5)).astype(F32)
            d = pos - center
            radius = F32(np.sqrt((d * d).sum(axis=1, dtype=F32)).max())
            radius = F32(radius * F32(1.000001) + F32(0.000001))
            vv = np.ascontiguousarray(v)
            box = bool(L.irt_host_is_box_geometry(ptr(vv), len(vv), 16, ptr(ind), len(ind), ptr(glo), ptr(ghi)))
            lip = float(L.irt_host_field_lipschitz_bound(ptr(f), res, ptr(spacing)))
            self.is_box.append(box)
            self.lipschitz.append(lip)
            origin = ((k % gx) * self.tile, ((k // gx) % gy) * self.tile, (k // (gx * gy)) * self.tile)
            r = self.meshes[k]
            r["lo"][:3], r["hi"][:3] = bmin, bmax
            r["geometry"] = (vtx, len(v), idx, len(ind))
            r["field"] = (*origin, res)
            r["grid"][:3] = (F32(1.0) / spacing).astype(F32)
            r["grid"][3] = F32(math.sqrt(float(np.dot(spacing, spacing))))
            r["geometry_lo"] = (*glo, 1.0 if box else 0.0)
            r["geometry_hi"] = (*ghi, F32(1.0) / F32(lip))
            r["enclosing_sphere"] = (*center, radius)
            vtx += len(v)
            idx += len(ind)

    def atlas(self) -> np.ndarray:
        """The R32Float SDF atlas (renderer.rs:778-845) as [depth, height, width]."""
        w, h, d = self.atlas_dims
        a = np.zeros((d, h, w), F32)
        for r, f in zip(self.meshes, self.fields):
            ox, oy, oz, res = (int(c) for c in r["field"])
            a[oz:oz + res, oy:oy + res, ox:ox + res] = f.reshape(res, res, res)
        return a

    def _build_objects(self):
        """objects() (renderer.rs:359-384) and hierarchy() (renderer.rs:385-455)."""
        L = _lib.irt()
        objs = self.manifest["objects"]
        self.objects = np.zeros(len(objs), OBJECT_DT)
        lo = np.zeros((len(objs), 3), F32)
        hi = np.zeros((len(objs), 3), F32)
        for i, o in enumerate(objs):
            mi = self.mesh_ids[o["mesh"]]
            mesh = self.meshes[mi]
            model = self._models[i]
            inv = np.zeros(16, F32)
            if not L.irt_host_inverse(ptr(model), ptr(inv)):
                raise ValueError(f"object {o['id']}: singular transform")
            scale = F32(math.sqrt(float(np.dot(model[:3].astype(np.float64), model[:3].astype(np.float64)))))
            spacing = ((mesh["hi"][:3] - mesh["lo"][:3]) / F32(mesh["field"][3] - 1)).astype(F32)
            voxel = F32(F32(math.sqrt(float(np.dot(spacing, spacing)))) * scale)
            rec = self.objects[i]
            rec["model"], rec["previous"], rec["inverse"] = model, self._previous[i], inv
            rec["info"] = (mi, self.material_ids[o["material"]], np.asarray(scale, F32).view(np.uint32),
                           np.asarray(voxel, F32).view(np.uint32))
            mlo = np.ascontiguousarray(mesh["lo"][:3])
            mhi = np.ascontiguousarray(mesh["hi"][:3])
            blo, bhi = np.zeros(3, F32), np.zeros(3, F32)
            L.irt_host_transform_box(ptr(np.ascontiguousarray(model)), ptr(mlo), ptr(mhi), ptr(blo), ptr(bhi))
            lo[i], hi[i] = blo, bhi
        self.nodes = np.zeros(max(2 * len(objs) - 1, 1), NODE_DT)
        L.irt_host_hierarchy(ptr(lo), ptr(hi), len(objs), ptr(self.nodes))

    def _build_lights(self):
        lights = self.manifest["lights"]
        self.lights = np.zeros(len(lights), LIGHT_DT)
        for rec, l in zip(self.lights, lights):
            rec["position_strength"] = (*l["position"], l["strength"])
            rec["direction_angle"] = (*l.get("direction", (0.0, 0.0, 0.0)), l.get("angle_radians", 0.0))
            rec["color_type"] = (*l["color"], 1.0 if l["type"] == "spot" else 0.0)

    def geometry(self):
        """(vertices VERTEX_DT, mesh-local indices, materials MATERIAL_DT) for irt_set_geometry."""
        verts = np.concatenate(self.mesh_vertices).astype(F32).view(VERTEX_DT).reshape(-1)
        inds = np.concatenate(self.mesh_indices).astype(np.uint32)
        mats = np.zeros(len(self.manifest["materials"]), MATERIAL_DT)
        for rec, m in zip(mats, self.manifest["materials"]):
            rec["color"] = m["base_color"]
            rec["params"] = (m["roughness"], m["metallic"], m.get("normal_strength", 1.0),
                             1.0 if m.get("base_color_srgb", True) else 0.0)
            if any(m.get(k) for k in ("base_color_texture", "roughness_texture", "metallic_texture",
                                      "normal_texture")):
                raise NotImplementedError("material textures: load them with Renderer.set_textures")
            rec["textures"] = (-1, -1, -1, -1)
            rec["channels"] = m.get("texture_channels", (0, 0, 0, 0))
        return verts, inds, mats

    # --------------------------------------------------------------- editing
    def changed(self):
        """Scene::changed (scene.rs:383-385): bump the generation, never 0."""
        self.generation = max((self.generation + 1) & 0xFFFFFFFF, 1)

    def set_light_strengths(self, strength: float):
        for l in self.manifest["lights"]:
            l["strength"] = strength
        self._build_lights()

    def move_object(self, index: int, offset):
        """Translates one object; its previous model becomes the model before the move."""
        o = self.manifest["objects"][index]
        o["transform"]["position"] = [float(a) + float(b) for a, b in zip(o["transform"]["position"], offset)]
        self._previous = [m.copy() for m in self._models]
        self._models[index] = transform_matrix(o["transform"])
        self._build_objects()

    def settle(self):
        """After a frame: every object's previous model is its current one."""
        self._previous = [m.copy() for m in self._models]
        self._build_objects()

    # ---------------------------------------------------------------- camera
    def camera(self) -> np.ndarray:
        c = self.manifest["camera"]
        cam = np.zeros((), CAMERA_DT)
        cam["position"] = c["position"]
        cam["rotation"] = c["rotation"]
        cam["fov_y"] = c["fov_y_radians"]
        cam["near_plane"] = c["near"]
        cam["far_plane"] = c["far"]
        return cam

    def camera_matrix(self, camera: np.ndarray, aspect: float) -> np.ndarray:
        """Camera::matrix (scene.rs:118-125): perspective_rh * inverse(rotation, translation)."""
        vp = np.zeros(16, F32)
        cam = np.ascontiguousarray(camera, dtype=CAMERA_DT)
        _lib.scenegen().sg_camera_matrix(ptr(cam), ctypes.c_float(aspect), ptr(vp))
        return vp
