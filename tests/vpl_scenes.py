"""Small hand-built triangle scenes for the VPL pass (oracle and GPU tests).

Each scene carries the records the C ABI takes (objects, meshes with their
geometry ranges, the instance hierarchy, lights, vertices, mesh-local
indices, materials) and a 2x2x2 placeholder SDF atlas: the VPL pass reads
no SDF, but the mesh records must name a valid atlas tile.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

import oracle
from instantradiositytest_amd import _lib
from instantradiositytest_amd.renderer import pack_frame
from instantradiositytest_amd.types import (LIGHT_DT, MATERIAL_DT, MESH_DT, NODE_DT, OBJECT_DT, VERTEX_DT,
                                            make_settings, ptr)


def model_matrix(scale=1.0, axis=(0.0, 1.0, 0.0), angle=0.0, translate=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Column-major T * R * S as float32[16]."""
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    x, y, z = a
    c, s = math.cos(angle), math.sin(angle)
    R = np.array([[c + x * x * (1 - c), x * y * (1 - c) - z * s, x * z * (1 - c) + y * s],
                  [y * x * (1 - c) + z * s, c + y * y * (1 - c), y * z * (1 - c) - x * s],
                  [z * x * (1 - c) - y * s, z * y * (1 - c) + x * s, c + z * z * (1 - c)]])
    M = np.eye(4)
    M[:3, :3] = R * scale
    M[:3, 3] = translate
    return M


def uv_sphere(radius: float, nu: int = 48, nv: int = 24, inward: bool = True):
    """Positions, normals (inward or outward) and indices of a UV sphere."""
    pos, nrm = [], []
    for j in range(nv + 1):
        th = math.pi * j / nv
        for i in range(nu):
            ph = 2 * math.pi * i / nu
            p = np.array([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)])
            pos.append(p * radius)
            nrm.append(-p if inward else p)
    idx = []
    for j in range(nv):
        for i in range(nu):
            a, b = j * nu + i, j * nu + (i + 1) % nu
            c, d = (j + 1) * nu + i, (j + 1) * nu + (i + 1) % nu
            idx += [a, c, b, b, c, d]
    return np.array(pos), np.array(nrm), np.array(idx, np.uint32)


def light(position, color=(1.0, 1.0, 1.0), strength=1.0, axis=None, angle=0.0):
    L = np.zeros(1, LIGHT_DT)
    L[0]["position_strength"] = (*position, strength)
    if axis is not None:
        L[0]["direction_angle"] = (*axis, angle)
        L[0]["color_type"] = (*color, 1.0)
    else:
        L[0]["color_type"] = (*color, 0.0)
    return L


@dataclass
class TriScene:
    objects: np.ndarray
    meshes: np.ndarray
    nodes: np.ndarray
    lights: np.ndarray
    vertices: np.ndarray
    indices: np.ndarray
    materials: np.ndarray
    atlas: np.ndarray

    def oracle_scene(self) -> oracle.SceneRecords:
        return oracle.SceneRecords(self.objects, self.meshes, self.nodes, self.lights, self.atlas)

    def oracle_geometry(self) -> oracle.GeometryRecords:
        return oracle.GeometryRecords(self.vertices, self.indices, self.materials)


def build(meshes, objects, materials, lights) -> TriScene:
    """meshes: [(positions Nx3, normals Nx3, indices)]; objects: [(mesh, 4x4 model, material)];
    materials: [(rgb, roughness, metallic)]; lights: LIGHT_DT array."""
    M = np.zeros(len(meshes), MESH_DT)
    verts, idx = [], []
    v_off = i_off = 0
    local_lo, local_hi = [], []
    for k, (p, n, ix) in enumerate(meshes):
        p = np.asarray(p, np.float32)
        V = np.zeros(len(p), VERTEX_DT)
        V["position"][:, :3] = p
        V["position"][:, 3] = 1
        V["normal"][:, :3] = n
        V["tangent"][:, 3] = 1
        verts.append(V)
        idx.append(np.asarray(ix, np.uint32))
        M[k]["geometry"] = (v_off, len(p), i_off, len(ix))
        M[k]["field"] = (0, 0, 0, 2)
        lo, hi = p.min(0), p.max(0)
        M[k]["lo"][:3], M[k]["hi"][:3] = lo, hi
        M[k]["geometry_lo"][:3], M[k]["geometry_hi"][:3] = lo, hi
        M[k]["grid"] = (1, 1, 1, 1)
        local_lo.append(lo)
        local_hi.append(hi)
        v_off += len(p)
        i_off += len(ix)
    O = np.zeros(len(objects), OBJECT_DT)
    box_lo = np.zeros((len(objects), 3), np.float32)
    box_hi = np.zeros((len(objects), 3), np.float32)
    for k, (mesh, model, mat) in enumerate(objects):
        O[k]["model"] = np.asarray(model, np.float64).T.reshape(-1)
        O[k]["previous"] = O[k]["model"]
        O[k]["inverse"] = np.linalg.inv(np.asarray(model, np.float64)).T.reshape(-1)
        O[k]["info"] = (mesh, mat, 0, 0)
        lo, hi = local_lo[mesh], local_hi[mesh]
        corners = np.array([[(hi if (c >> a) & 1 else lo)[a] for a in range(3)] + [1.0] for c in range(8)])
        w = (np.asarray(model, np.float64) @ corners.T).T[:, :3]
        pad = 1e-3 + 1e-3 * np.abs(w).max()
        box_lo[k], box_hi[k] = w.min(0) - pad, w.max(0) + pad
    N = np.zeros(2 * len(objects) - 1, NODE_DT)
    _lib.irt().irt_host_hierarchy(ptr(box_lo), ptr(box_hi), len(objects), ptr(N))
    MAT = np.zeros(len(materials), MATERIAL_DT)
    for k, (rgb, rough, metal) in enumerate(materials):
        MAT[k]["color"] = (*rgb, 1)
        MAT[k]["params"] = (rough, metal, 1, 0)
        MAT[k]["textures"] = -1
    return TriScene(O, M, N, np.ascontiguousarray(lights, LIGHT_DT), np.concatenate(verts), np.concatenate(idx), MAT,
                    np.zeros((2, 2, 2), np.float32))


def vpl_frame(V: int, bounces: int, n_lights: int, n_objects: int, frame_index: int = 0, rr=(0.1, 0.95),
              width: int = 16, height: int = 16) -> np.ndarray:
    s = make_settings(render_width=width, render_height=height, vpl_count=V, vpl_max_bounces=bounces,
                      vpl_rr_min_survival=rr[0], vpl_rr_max_survival=rr[1])
    vp = np.eye(4, dtype=np.float32).reshape(-1)
    return pack_frame(s, vp, vp, np.zeros(3, np.float32), frame_index, 1, 0, False, 0, n_lights, n_objects)


def closed_sphere(lights, radius=2.0, rgb=(0.5, 0.5, 0.5), rough=1.0, metal=0.0) -> TriScene:
    """One inward-facing UV sphere around the origin (every emitted ray hits it)."""
    p, n, ix = uv_sphere(radius)
    return build([(p, n, ix)], [(0, np.eye(4), 0)], [(rgb, rough, metal)], lights)


def device_pool(sc: TriScene, V: int, bounces: int, frame_index: int = 0, precision: int = 0,
                rr=(0.1, 0.95)) -> np.ndarray:
    """irt_generate_vpls for a hand-built scene on a fresh 16x16 context; the
    generated pool as float32 [V, 12]."""
    import ctypes

    from instantradiositytest_amd._lib import check
    from instantradiositytest_amd.renderer import read_buffer

    L = _lib.irt()
    s = make_settings(render_width=16, render_height=16, vpl_count=V, vpl_max_bounces=bounces,
                      vpl_rr_min_survival=rr[0], vpl_rr_max_survival=rr[1])
    ctx = ctypes.c_void_p()
    check(L.irt_create(ctypes.byref(s), precision, ctypes.byref(ctx)))
    try:
        check(L.irt_set_sdf_atlas(ctx, ptr(sc.atlas), 2, 2, 2), ctx)
        check(L.irt_set_meshes(ctx, ptr(sc.meshes), len(sc.meshes)), ctx)
        check(L.irt_set_objects(ctx, ptr(sc.objects), len(sc.objects), ptr(sc.nodes), len(sc.nodes)), ctx)
        check(L.irt_set_lights(ctx, ptr(sc.lights), len(sc.lights)), ctx)
        check(L.irt_set_geometry(ctx, ptr(sc.vertices), len(sc.vertices), ptr(sc.indices), len(sc.indices),
                                 ptr(sc.materials), len(sc.materials)), ctx)
        f = vpl_frame(V, bounces, len(sc.lights), len(sc.objects), frame_index, rr)
        check(L.irt_set_frame(ctx, ptr(f)), ctx)
        check(L.irt_generate_vpls(ctx), ctx)
        ring = read_buffer(ctx, "vpl_archive", 16, 16)
        slot = frame_index % (s.restir_max_history + 1)
        return ring[slot * V:(slot + 1) * V].view(np.float32).reshape(V, 12).copy()
    finally:
        L.irt_destroy(ctx)
