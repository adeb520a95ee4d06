"""Writes the reference's exported test scene (assets/test_scene) as a
compact, exact fixture: tests/golden/test_scene/scene.json (the manifest,
unchanged) and tests/golden/test_scene/payload.npz.

The manifest's meshes point into meshes_9e1a11caa7e0d0ee.bin (64-byte vertex
records and u32 indices, src/scene.rs:128-260) and fields_ab55580bb9dab0a9.bin
(48^3 float32 SDF samples).  Only the referenced ranges are kept.  A mesh whose
triangles are listed in order (index i = i), flat shaded (one normal, w = 0, and
one tangent per triangle) and without texture coordinates -- the exported
Suzanne, 778 566 vertex records of 130 251 distinct positions -- is stored as
its distinct positions, the position index of every corner and the normal of
every triangle; every other mesh keeps its records and indices as they are.
The script rebuilds every record from the stored arrays and checks it against
the original bytes before writing.

Run in the container that holds the reference (it is not needed on the GPU
box): python tests/golden/make_test_scene_fixture.py [reference root]
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "test_scene"


def flat_encoding(v: np.ndarray, idx: np.ndarray):
    """(positions, corners, triangle normals, tangent) when the mesh allows the compact form, else None."""
    if not np.array_equal(idx, np.arange(len(idx), dtype=np.uint32)) or len(v) % 3:
        return None
    tri = v.reshape(-1, 3, 16)
    if np.any(tri[:, :, 4:8] != tri[:, :1, 4:8]) or np.any(v[:, 12:16] != 0) or np.any(v[:, 3] != 1):
        return None
    tangent = v[0, 8:12]
    if np.any(v[:, 8:12] != tangent) or np.any(v[:, 7] != 0):
        return None
    pos = np.ascontiguousarray(v[:, :3])
    uniq, corner = np.unique(pos.view(np.dtype((np.void, 12))).reshape(-1), return_inverse=True)
    return uniq.view(np.float32).reshape(-1, 3), corner.reshape(-1).astype(np.uint32), tri[:, 0, 4:7].copy(), tangent


def expand_flat(pos, corner, normal, tangent) -> np.ndarray:
    v = np.zeros((len(corner), 16), np.float32)
    v[:, :3] = pos[corner]
    v[:, 3] = 1.0
    v[:, 4:7] = np.repeat(normal, 3, axis=0)
    v[:, 8:12] = tangent
    return v


def main(root: Path) -> None:
    scene = root / "assets" / "test_scene"
    manifest = json.loads((scene / "scene.json").read_text())
    blobs = {}
    arrays = {}
    for i, m in enumerate(manifest["meshes"]):
        g, s = m["geometry"], m["sdf"]
        geo = blobs.setdefault(g["path"], np.fromfile(scene / g["path"], np.uint8))
        v = geo[g["vertex_offset"]:g["vertex_offset"] + 64 * g["vertex_count"]].view(np.float32).reshape(-1, 16)
        idx = geo[g["index_offset"]:g["index_offset"] + 4 * g["index_count"]].view(np.uint32)
        flat = flat_encoding(v, idx)
        if flat is not None:
            pos, corner, normal, tangent = flat
            assert expand_flat(pos, corner, normal, tangent).tobytes() == v.tobytes()
            arrays[f"mesh{i}_positions"], arrays[f"mesh{i}_corners"] = pos, corner
            arrays[f"mesh{i}_triangle_normals"], arrays[f"mesh{i}_tangent"] = normal, tangent
        else:
            arrays[f"mesh{i}_vertices"], arrays[f"mesh{i}_indices"] = v.copy(), idx.copy()
        fld = blobs.setdefault(s["path"], np.fromfile(scene / s["path"], np.uint8))
        r = s["resolution"]
        arrays[f"mesh{i}_sdf"] = fld[s["offset"]:s["offset"] + 4 * r ** 3].view(np.float32).copy()
    OUT.mkdir(parents=True, exist_ok=True)
    (OUT / "scene.json").write_bytes((scene / "scene.json").read_bytes())
    np.savez_compressed(OUT / "payload.npz", **arrays)
    print(f"wrote {OUT}: {len(manifest['meshes'])} meshes, {(OUT / 'payload.npz').stat().st_size / 1e6:.2f} MB")


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else Path("/root/reference"))
