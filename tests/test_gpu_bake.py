"""SDF baking on the GPU (irt_bake_sdf) against the oracle.

Nearest distances and parity casts use the same arithmetic and tie rule on
both sides and the minimum over triangles does not depend on the walk, so
the baked fields must be identical bit for bit; the GPU's Lipschitz bound
must equal the host's field_lipschitz_bound of the same field.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from instantradiositytest_amd import _lib
from instantradiositytest_amd.types import ptr
from test_oracle_bake import box_mesh, torus_mesh
from vpl_scenes import uv_sphere

pytestmark = pytest.mark.gpu


def gpu_bake(pos, idx, res):
    pos = np.ascontiguousarray(pos, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    out = np.zeros((res, res, res), np.float32)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lip = ctypes.c_float()
    rc = _lib.irt().irt_bake_sdf(ptr(pos), pos.shape[1], len(pos), ptr(idx), len(idx), res, ptr(out), ptr(lo), ptr(hi),
                                 ctypes.byref(lip))
    return rc, out, lo, hi, lip.value


@pytest.mark.parametrize("shape,res", [("box", 16), ("sphere", 24), ("torus", 32), ("sphere", 48)])
def test_gpu_bake_matches_oracle(require_gpu, shape, res):
    print(f"[bake] {shape} {res}^3", flush=True)
    if shape == "box":
        pos, idx = box_mesh(np.array([0.3, 0.2, 0.45]))
    elif shape == "sphere":
        pos, _, idx = uv_sphere(0.4, 48, 24, inward=False)
    else:
        pos, idx = torus_mesh(0.3, 0.1)
    rc, field, lo, hi, lip = gpu_bake(pos, idx, res)
    assert rc == 0, _lib.irt().irt_global_error()
    ref, rlo, rhi = oracle.bake_sdf(pos, idx, res)
    assert np.array_equal(lo, rlo) and np.array_equal(hi, rhi)
    diff = field != ref
    print(f"[bake] {shape} {res}: {diff.mean():.6f} of samples differ", flush=True)
    assert not diff.any()
    sp = ((rhi - rlo) / np.float32(res - 1)).astype(np.float32)
    host_lip = _lib.irt().irt_host_field_lipschitz_bound(ptr(ref), res, ptr(sp))
    assert lip == host_lip


def test_gpu_bake_rejects_bad_input(require_gpu):
    pos, idx = box_mesh(np.array([0.3, 0.2, 0.45]))
    assert gpu_bake(pos, idx, 3)[0] == -1
    quads, tri = [], []
    for k in range(257):
        z = -0.5 + k / 256.0
        base = len(quads)
        quads += [[-2, -2, z], [2, -2, z], [2, 2, z], [-2, 2, z]]
        tri += [base, base + 1, base + 2, base, base + 2, base + 3]
    rc = gpu_bake(np.array(quads), np.array(tri, np.uint32), 4)[0]
    assert rc == -1 and b"256 intersections" in _lib.irt().irt_global_error()
