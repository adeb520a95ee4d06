"""The candidate-list rule behind the fused path's scene queries, checked on
the CPU restatement: for every point, the minimum object distance over its
cell's candidates equals the hierarchy walk's result (sdf.wgsl:42-60).

Rule (as built on the GPU): lower bound of object j over a cell = 0.9998 x
the cell-to-world-box gap (-inf when they overlap); upper bound = distance at
the cell centre + K_j x half diagonal, K_j = 1 for certified boxes and
sqrt(K_mesh^2 + 1) for grid meshes, K_mesh the largest corner gradient of
the field; candidates = lower <= min upper."""
from __future__ import annotations

import ctypes

import numpy as np

import oracle
from instantradiositytest_amd.scenes import CONFIGS, ConfigScene
from instantradiositytest_amd.types import FRAME_DT, NODE_DT

L = oracle.lib()


def mesh_gradient_bound(atlas, mesh):
    ox, oy, oz, res = (int(v) for v in mesh["field"])
    f = atlas[oz:oz + res, oy:oy + res, ox:ox + res].astype(np.float64)
    g = mesh["grid"][:3].astype(np.float64)
    best = 1.0
    for cz in (0, 1):
        for cy in (0, 1):
            for cx in (0, 1):
                # corner (cx, cy, cz) gradient of every cell
                dx = (f[cz:res - 1 + cz, cy:res - 1 + cy, 1:] - f[cz:res - 1 + cz, cy:res - 1 + cy, :-1]) * g[0]
                dy = (f[cz:res - 1 + cz, 1:, cx:res - 1 + cx] - f[cz:res - 1 + cz, :-1, cx:res - 1 + cx]) * g[1]
                dz = (f[1:, cy:res - 1 + cy, cx:res - 1 + cx] - f[:-1, cy:res - 1 + cy, cx:res - 1 + cx]) * g[2]
                n2 = dx[:, :, :] ** 2 + dy ** 2 + dz ** 2
                best = max(best, float(n2.max()))
    return best


def test_candidate_lists_preserve_the_minimum():
    cs = ConfigScene(CONFIGS["plain-1920x1200"].scaled(64, 40, 1))
    rec = oracle.SceneRecords.from_scene(cs.scene)
    n = len(rec.objects)
    f = np.zeros(1, FRAME_DT)
    f["counts"][0] = (rec.n_lights, 1, 1, n)
    f["sdf"][0] = (96, 0.001, 16, 0)
    one = np.zeros(1, FRAME_DT)
    one["counts"][0] = (rec.n_lights, 1, 1, 1)
    one["sdf"][0] = (96, 0.001, 16, 0)
    leaf = np.zeros(1, NODE_DT)
    leaf[0]["info"] = (0, 0, 1, 1)
    singles = [oracle.SceneRecords(rec.objects[[j]].copy(), rec.meshes, leaf, rec.lights[:rec.n_lights], rec.atlas)
               for j in range(n)]
    boxes = {}
    for nd in rec.nodes:
        if nd["info"][2] == 1:
            boxes[int(nd["info"][0])] = (nd["lo"][:3].astype(np.float64), nd["hi"][:3].astype(np.float64))
    K = []
    for j in range(n):
        mesh = rec.meshes[int(rec.objects[j]["info"][0])]
        K.append(1.0 if mesh["geometry_lo"][3] > 0.5 else float(np.sqrt(mesh_gradient_bound(rec.atlas, mesh) + 1.0)))

    def dist(s, frame, p):
        return L.oracle_scene_distance(32, frame.ctypes.data, ctypes.byref(s.c),
                                       np.ascontiguousarray(p, np.float64).ctypes.data, 0, 0xFFFFFFFF, None, None)

    rng = np.random.default_rng(9)
    root_lo = rec.nodes[0]["lo"][:3].astype(np.float64)
    root_hi = rec.nodes[0]["hi"][:3].astype(np.float64)
    cell = (root_hi - root_lo).max() / 24
    checked = 0
    total_cands = 0
    for _ in range(60):
        c0 = root_lo + np.floor(rng.uniform(0, 1, 3) * (root_hi - root_lo) / cell) * cell
        c1 = c0 + cell
        center = (c0 + c1) / 2
        half = 0.5 * cell * np.sqrt(3) * 1.0001
        lower, upper = [], []
        for j in range(n):
            blo, bhi = boxes[j]
            gap = np.linalg.norm(np.maximum(np.maximum(blo - c1, c0 - bhi), 0))
            lower.append(0.9998 * gap if gap > 0 else -np.inf)
            up = dist(singles[j], one, center) + K[j] * 1.001 * half
            upper.append(up + abs(up) * 1e-5 + 1e-6)
        cand = [j for j in range(n) if lower[j] <= min(upper)]
        total_cands += len(cand)
        for p in rng.uniform(c0, c1, (25, 3)):
            full = dist(rec, f, p)
            best = min(dist(singles[j], one, p) for j in cand)
            assert best == full
            checked += 1
    assert checked == 1500
    assert total_cands / 60 < n / 2  # the lists actually cull
