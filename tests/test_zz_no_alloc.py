"""irt_bake_sdf_noalloc (fixed device scratch, no allocation) against
irt_bake_sdf (which allocates its scratch per call), through the C ABI.

Both run the same kernels on the same plan, so samples, padded bounds and
the Lipschitz bound must be identical bit for bit.  Calls follow each other
with a smaller mesh and grid second, so stale contents of the shared scratch
from the larger call must not leak into the smaller one; an overflowing mesh
between good calls checks that the flag word is set afresh by every call.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from instantradiositytest_amd import _lib
from instantradiositytest_amd.types import ptr
from test_oracle_bake import box_mesh, torus_mesh

pytestmark = pytest.mark.gpu


def bake(entry, pos, idx, res):
    pos = np.ascontiguousarray(pos, np.float32)
    idx = np.ascontiguousarray(idx, np.uint32)
    out = np.full((res, res, res), np.nan, np.float32)
    lo = np.zeros(3, np.float32)
    hi = np.zeros(3, np.float32)
    lip = ctypes.c_float(-1.0)
    rc = getattr(_lib.irt(), entry)(ptr(pos), pos.shape[1], len(pos), ptr(idx), len(idx), res, ptr(out), ptr(lo),
                                    ptr(hi), ctypes.byref(lip))
    return rc, out, lo, hi, lip.value


def stacked_quads(n):
    quads, tri = [], []
    for k in range(n):
        z = -0.5 + k / 256.0
        base = len(quads)
        quads += [[-2, -2, z], [2, -2, z], [2, 2, z], [-2, 2, z]]
        tri += [base, base + 1, base + 2, base, base + 2, base + 3]
    return np.array(quads, np.float32), np.array(tri, np.uint32)


def assert_same(a, b):
    assert a[0] == 0 and b[0] == 0, _lib.irt().irt_global_error()
    assert np.array_equal(a[1], b[1])
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert a[4] == b[4] and a[4] >= 1.0


def test_noalloc_bake_equals_allocating_bake(require_gpu):
    torus = torus_mesh(0.3, 0.1)
    box = box_mesh(np.array([0.3, 0.2, 0.45]))
    print("[noalloc] torus 48^3", flush=True)
    assert_same(bake("irt_bake_sdf_noalloc", *torus, 48), bake("irt_bake_sdf", *torus, 48))
    print("[noalloc] box 16^3 after the torus", flush=True)
    assert_same(bake("irt_bake_sdf_noalloc", *box, 16), bake("irt_bake_sdf", *box, 16))
    print("[noalloc] overflowing quads, then the box again", flush=True)
    rc = bake("irt_bake_sdf_noalloc", *stacked_quads(257), 4)[0]
    assert rc == -1 and b"256 intersections" in _lib.irt().irt_global_error()
    assert_same(bake("irt_bake_sdf_noalloc", *box, 16), bake("irt_bake_sdf", *box, 16))


def test_noalloc_bake_at_its_limits(require_gpu):
    torus = torus_mesh(0.3, 0.1)
    print("[noalloc] torus 128^3 (largest grid)", flush=True)
    assert_same(bake("irt_bake_sdf_noalloc", *torus, 128), bake("irt_bake_sdf", *torus, 128))
    rc = bake("irt_bake_sdf_noalloc", *torus, 129)[0]
    assert rc == -1 and b"fixed scratch" in _lib.irt().irt_global_error()
    pos, idx = stacked_quads(8193)  # 16386 triangles
    rc = bake("irt_bake_sdf_noalloc", pos, idx, 8)[0]
    assert rc == -1 and b"fixed scratch" in _lib.irt().irt_global_error()
