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
    textures: np.ndarray | None = None  # uint8 [layers, size, size, 4]

    def oracle_scene(self) -> oracle.SceneRecords:
        return oracle.SceneRecords(self.objects, self.meshes, self.nodes, self.lights, self.atlas)

    def oracle_geometry(self) -> oracle.GeometryRecords:
        return oracle.GeometryRecords(self.vertices, self.indices, self.materials, self.textures)


def build(meshes, objects, materials, lights, textures=None) -> TriScene:
    """meshes: [(positions Nx3, normals Nx3, indices[, uv Nx2, tangents Nx4])]; objects: [(mesh, 4x4 model,
    material)]; materials: [(rgb, roughness, metallic[, options])] where options may give "textures"
    (4 layer indices, -1 none), "channels", "srgb" and "normal_strength"; lights: LIGHT_DT array;
    textures: uint8 [layers, size, size, 4] or None."""
    M = np.zeros(len(meshes), MESH_DT)
    verts, idx = [], []
    v_off = i_off = 0
    local_lo, local_hi = [], []
    for k, mesh in enumerate(meshes):
        p, n, ix = mesh[:3]
        p = np.asarray(p, np.float32)
        V = np.zeros(len(p), VERTEX_DT)
        V["position"][:, :3] = p
        V["position"][:, 3] = 1
        V["normal"][:, :3] = n
        V["tangent"][:, 3] = 1
        if len(mesh) > 3:
            V["uv"][:, :2] = mesh[3]
        if len(mesh) > 4:
            V["tangent"][:] = mesh[4]
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
    for k, mat in enumerate(materials):
        rgb, rough, metal = mat[:3]
        opt = mat[3] if len(mat) > 3 else {}
        MAT[k]["color"] = (*rgb, 1)
        MAT[k]["params"] = (rough, metal, opt.get("normal_strength", 1.0), 1.0 if opt.get("srgb") else 0.0)
        MAT[k]["textures"] = opt.get("textures", (-1, -1, -1, -1))
        MAT[k]["channels"] = opt.get("channels", (0, 0, 0, 0))
    tex = None if textures is None else np.ascontiguousarray(textures, np.uint8)
    return TriScene(O, M, N, np.ascontiguousarray(lights, LIGHT_DT), np.concatenate(verts), np.concatenate(idx), MAT,
                    np.zeros((2, 2, 2), np.float32), tex)


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
        if sc.textures is not None:
            check(L.irt_set_textures(ctx, ptr(sc.textures), sc.textures.shape[1], sc.textures.shape[0]), ctx)
        f = vpl_frame(V, bounces, len(sc.lights), len(sc.objects), frame_index, rr)
        check(L.irt_set_frame(ctx, ptr(f)), ctx)
        check(L.irt_generate_vpls(ctx), ctx)
        ring = read_buffer(ctx, "vpl_archive", 16, 16)
        slot = frame_index % (s.restir_max_history + 1)
        return ring[slot * V:(slot + 1) * V].view(np.float32).reshape(V, 12).copy()
    finally:
        L.irt_destroy(ctx)


def textured_sphere(lights, size: int = 16, seed: int = 7) -> TriScene:
    """An inward UV sphere with spherical uv and tangents, an sRGB base-colour
    texture, a roughness texture (green channel) and a normal map."""
    p, n, ix = uv_sphere(2.0)
    d = p / np.linalg.norm(p, axis=1, keepdims=True)
    uv = np.stack([np.arctan2(d[:, 2], d[:, 0]) / (2 * np.pi) + 0.5, np.arccos(np.clip(d[:, 1], -1, 1)) / np.pi], 1)
    tang = np.stack([-d[:, 2], np.zeros(len(d)), d[:, 0], np.ones(len(d))], 1)
    tang[np.linalg.norm(tang[:, :3], axis=1) < 1e-6, :3] = (1, 0, 0)
    rng = np.random.default_rng(seed)
    tex = rng.integers(0, 256, (3, size, size, 4), dtype=np.uint8)
    tex[2, ..., 2] = np.maximum(tex[2, ..., 2], 160)  # normal map leaning along +z
    mat = ((0.9, 0.8, 0.7), 0.9, 0.3, {"textures": (0, 1, -1, 2), "channels": (0, 1, 0, 0), "srgb": True,
                                       "normal_strength": 0.7})
    return build([(p, n, ix, uv, tang)], [(0, np.eye(4), 0)], [mat], lights, tex)


def textured_quads(size: int = 8, seed: int = 9) -> TriScene:
    """Two overlapping quads in NDC space (identity view-projection), each with
    its own textures, uv spanning two texture repeats, and a tangent."""
    rng = np.random.default_rng(seed)
    tex = rng.integers(0, 256, (4, size, size, 4), dtype=np.uint8)
    quad = [0, 1, 2, 0, 2, 3]
    n = np.tile([0, 0, 1.0], (4, 1))
    a = np.array([[-1, -1, 0.6], [0.6, -1, 0.6], [0.6, 0.7, 0.6], [-1, 0.7, 0.6]])
    b = np.array([[-0.3, -0.5, 0.4], [1, -0.6, 0.45], [1, 1, 0.5], [-0.3, 1, 0.42]])
    uv = np.array([[0, 0], [2, 0], [2, 2], [0, 2]], np.float64)
    t1 = np.tile([1.0, 0, 0, 1.0], (4, 1))
    t2 = np.tile([0.0, 1.0, 0, -1.0], (4, 1))
    mats = [((0.8, 0.7, 0.6), 0.7, 0.2, {"textures": (0, 1, 2, 3), "channels": (0, 2, 1, 0), "srgb": True}),
            ((0.5, 0.9, 0.4), 0.4, 0.6, {"textures": (1, -1, 0, 3), "channels": (0, 0, 3, 0)})]
    return build([(a, n, quad, uv, t1), (b, n, quad, uv * 1.5, t2)],
                 [(0, np.eye(4), 0), (1, np.eye(4), 1)], mats, light((0, 0, 0)), tex)
