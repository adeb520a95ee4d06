"""Pins the CPU restatement (oracle) before it is trusted as the checker.

Sources of truth, none of which re-types the oracle's own formula:
  - golden PCG vectors from an exact big-integer evaluation (tests/golden);
  - the reference's known-answer tests in src/math.rs:78-199;
  - closed forms (BRDF at normal incidence, inverse-square ratio);
  - invariants: conservative SDF bounds never overstep the analytic
    distance; the BVH walk equals a brute-force minimum over objects; the
    march agrees with a dense fixed-step scan on binary visibility.
"""
from __future__ import annotations

import ctypes
import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from instantradiositytest_amd.scenes import CONFIGS, ConfigScene, Scene
from instantradiositytest_amd.types import LIGHT_DT, NODE_DT, SG_SPHERE, SG_TORUS, SG_BOX, SGOBJECT_DT, SHAPE_DT

L = oracle.lib()
GOLDEN = Path(__file__).resolve().parent / "golden"


def d3(v):
    return np.ascontiguousarray(v, np.float64)


# ----------------------------------------------------------------- PCG stream
def test_pcg_stream_matches_golden_vectors():
    cases = json.loads((GOLDEN / "pcg_vectors.json").read_text())
    for c in cases:
        assert L.oracle_seed(c["i"] & 0xFFFFFFFF, c["salt"], c["frame"] & 0xFFFFFFFF) == c["seed"]
        st = ctypes.c_uint32(c["seed"])
        got = [L.oracle_random(ctypes.byref(st)) for _ in c["values"]]
        assert got == c["values"]
        assert st.value == c["final_state"]


# ------------------------------------------------------- reservoirs (math.rs)
def reservoir():
    return np.zeros(1, oracle.ORES_DT)


def add(bits, r, id_, target, pdf, weight, count, random):
    return L.oracle_reservoir_add(bits, r.ctypes.data, id_, 0, target, pdf, weight, count, 0, random, None)


@pytest.mark.parametrize("bits", [32, 64])
def test_constant_target_reservoir_recovers_light_sum(bits):
    # math.rs:109-116: 8 inserts of target 2 with pdf 1/4 -> M = 8, 2 * W = 8
    r = reservoir()
    for i in range(8):
        add(bits, r, i, 2.0, 0.25, 2.0 / 0.25, 1, 0.1)
    assert r["count"][0] == 8
    assert abs(2 * L.oracle_normalization(bits, r.ctypes.data) - 8) < 1e-6


@pytest.mark.parametrize("bits", [32, 64])
def test_reuse_reweights_target(bits):
    # math.rs:118-130
    src = reservoir()
    add(bits, src, 0, 2.0, 0.25, 8.0, 1, 0.0)
    w = L.oracle_normalization(bits, src.ctypes.data)
    r = reservoir()
    add(bits, r, int(src["id"][0]), 4.0, 0.25, 4.0 * w * float(src["count"][0]), int(src["count"][0]), 0.0)
    assert abs(L.oracle_normalization(bits, r.ctypes.data) - 4.0) < 1e-6


def test_add_range_rule_follows_wgsl():
    # common.wgsl:100-106: weight <= 0 or > 1e30 is rejected (count still grows); NaN passes the test
    r = reservoir()
    assert add(32, r, 1, 1.0, 1.0, 0.0, 1, 0.5) == 0
    assert add(32, r, 2, 1.0, 1.0, 1e31, 1, 0.5) == 0
    assert add(32, r, 3, 1.0, 1.0, float("inf"), 1, 0.5) == 0
    assert r["count"][0] == 3 and r["sum"][0] == 0
    assert add(32, r, 4, 1.0, 1.0, float("nan"), 1, 0.5) == 0
    assert r["count"][0] == 4 and math.isnan(r["sum"][0])


# ------------------------------------------------------- soft shadow (math.rs)
@pytest.mark.parametrize("bits", [32, 64])
def test_shadow_step_known_answer_and_monotone(bits):
    # math.rs:158-163 (first step: previous = 1e20 in the WGSL, sdf.wgsl:175)
    a = L.oracle_shadow_step(bits, 1.0, 0.1, 1e20, 2.0, 16.0)
    assert abs(a - 0.8) < 1e-5
    b = L.oracle_shadow_step(bits, a, 0.01, 0.1, 2.1, 16.0)
    assert 0 <= b <= a


# -------------------------------------------------- signed box distance (math.rs)
@pytest.mark.parametrize("bits", [32, 64])
def test_signed_box_distance_known_answers(bits):
    lo, hi = d3([-1, -1, -1]), d3([1, 1, 1])
    f = lambda p: L.oracle_signed_box_distance(bits, d3(p).ctypes.data, lo.ctypes.data, hi.ctypes.data)
    assert f([0, 0, 0]) == -1
    assert f([2, 0, 0]) == 1
    assert abs(f([2, 2, 2]) - math.sqrt(3)) < 1e-6
    assert abs(f([1, 0, 0])) < 1e-6


def test_box_and_sphere_bounds_never_overstep():
    # math.rs:81-102: max(box, sphere) is a lower bound of a unit sphere and a torus (0.7, 0.3) and 1-Lipschitz
    lo, hi = d3([-1, -1, -1]), d3([1, 1, 1])
    g = np.arange(-15, 16) / 10.0
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X, Y, Z], -1).reshape(-1, 3)
    box = np.array([L.oracle_signed_box_distance(64, d3(p).ctypes.data, lo.ctypes.data, hi.ctypes.data) for p in P])
    bound = np.maximum(box, np.linalg.norm(P, axis=1) - 1)
    sphere = np.linalg.norm(P, axis=1) - 1
    torus = np.hypot(np.hypot(P[:, 0], P[:, 1]) - 0.7, P[:, 2]) - 0.3
    assert np.all(bound <= sphere + 1e-6)
    assert np.all(bound <= torus + 1e-6)
    for off in (np.array([1, 0, 0]), np.array([0, 1, 0]), np.array([0, 0, 1]), np.ones(3) / math.sqrt(3)):
        Q = P + off * 0.031
        qb = np.array([L.oracle_signed_box_distance(64, d3(q).ctypes.data, lo.ctypes.data, hi.ctypes.data) for q in Q])
        other = np.maximum(qb, np.linalg.norm(Q, axis=1) - 1)
        assert np.all(np.abs(bound - other) <= 0.031 + 1e-6)


# -------------------------------------------------------- BRDF closed forms
@pytest.mark.parametrize("bits", [32, 64])
def test_brdf_normal_incidence_closed_forms(bits):
    n = [0, 0, 1]
    color = [0.5, 0.25, 0.75]
    out = np.zeros(3)
    # dielectric, rough = 1: D = 1/pi, G = 1, F = 0.04 -> (0.96 * c + 0.01) / pi
    s = d3([0, 0, 0, *n, *color, 1.0, 0.0])
    L.oracle_brdf(bits, s.ctypes.data, d3(n).ctypes.data, d3(n).ctypes.data, out.ctypes.data)
    expect = (0.96 * np.array(color) + 0.01) / math.pi
    assert np.allclose(out, expect, rtol=2e-6)
    # metal, rough = 0.5: diffuse 0, F = c, D = 1/(pi a2) with a2 = 0.0625 -> c / (4 pi a2)
    s = d3([0, 0, 0, *n, *color, 0.5, 1.0])
    L.oracle_brdf(bits, s.ctypes.data, d3(n).ctypes.data, d3(n).ctypes.data, out.ctypes.data)
    assert np.allclose(out, np.array(color) / (4 * math.pi * 0.0625), rtol=2e-6)
    # light below the horizon -> 0
    L.oracle_brdf(bits, s.ctypes.data, d3(n).ctypes.data, d3([0, 0, -1]).ctypes.data, out.ctypes.data)
    assert np.all(out == 0)


# ------------------------------------------------------ SDF bounds vs analytic
def single_object_scene(kind, params, res, position=(0, 0, 0), scale=1.0, angle=0.3):
    S = np.zeros(1, SHAPE_DT)
    S[0]["kind"], S[0]["resolution"], S[0]["p"] = kind, res, params
    O = np.zeros(1, SGOBJECT_DT)
    O[0]["mesh"] = 0
    O[0]["position"] = position
    O[0]["rotation"] = (0, math.sin(angle / 2), 0, math.cos(angle / 2))
    O[0]["scale"] = scale
    return Scene(S, O, np.zeros(0, LIGHT_DT))


def analytic(kind, params, p):
    if kind == SG_SPHERE:
        return np.linalg.norm(p, axis=-1) - params[0]
    if kind == SG_TORUS:
        return np.hypot(np.hypot(p[..., 0], p[..., 2]) - params[0], p[..., 1]) - params[1]
    q = np.abs(p) - np.array(params)
    return np.linalg.norm(np.maximum(q, 0), axis=-1) + np.minimum(q.max(axis=-1), 0)


@pytest.mark.parametrize("kind,params,res", [(SG_SPHERE, (0.4, 0, 0), 16), (SG_SPHERE, (0.3, 0, 0), 48),
                                              (SG_TORUS, (0.3, 0.08, 0), 32), (SG_BOX, (0.3, 0.1, 0.2), 24)])
@pytest.mark.parametrize("bits", [32, 64])
def test_mesh_distance_is_conservative(kind, params, res, bits):
    """Every mesh_distance value (grid, Jensen allowance, Lipschitz scale,
    padded-box extension, box and sphere bounds; sdf.wgsl:5-33) is a lower
    bound of the analytic distance, inside and outside the grid."""
    sc = single_object_scene(kind, params, res)
    rec = oracle.SceneRecords.from_scene(sc)
    rng = np.random.default_rng(res)
    P = rng.uniform(-1.2, 1.2, (3000, 3))
    got = np.array([L.oracle_mesh_distance(bits, d3(p).ctypes.data, rec.meshes[0:1].ctypes.data, ctypes.byref(rec.c))
                    for p in P])
    exact = analytic(kind, params, P)
    assert np.all(got <= exact + 1e-5), float((got - exact).max())
    # and it is useful: outside the object the bound is positive away from the surface
    far = exact > 0.2
    assert np.all(got[far] > 0)


def scene_records_subset(rec: oracle.SceneRecords, ids):
    objs = rec.objects[list(ids)].copy()
    nodes = np.zeros(1, NODE_DT)
    nodes[0]["info"] = (0, 0, 1, 1)
    return oracle.SceneRecords(objs, rec.meshes, nodes, rec.lights[:rec.n_lights], rec.atlas)


def frame_for(rec: oracle.SceneRecords, n_objects):
    from instantradiositytest_amd.types import FRAME_DT
    f = np.zeros(1, FRAME_DT)
    f["counts"][0] = (rec.n_lights, 1, 1, n_objects)
    f["sdf"][0] = (96, 0.001, 16, 0)
    return f


def test_bvh_walk_equals_brute_force_minimum():
    cs = ConfigScene(CONFIGS["plain-640x400"].scaled(64, 40, 1))
    rec = oracle.SceneRecords.from_scene(cs.scene)
    n = len(rec.objects)
    full = frame_for(rec, n)
    singles = [scene_records_subset(rec, [i]) for i in range(n)]
    one = frame_for(singles[0], 1)
    rng = np.random.default_rng(3)
    P = np.column_stack([rng.uniform(-2.2, 2.2, 400), rng.uniform(-0.2, 3.2, 400), rng.uniform(-2.7, 2.7, 400)])
    for bits in (32, 64):
        for p in P:
            v = ctypes.c_double()
            c = ctypes.c_uint32()
            d = L.oracle_scene_distance(bits, full.ctypes.data, ctypes.byref(rec.c), d3(p).ctypes.data, 1, 0xFFFFFFFF,
                                        ctypes.byref(v), ctypes.byref(c))
            each = [L.oracle_scene_distance(bits, one.ctypes.data, ctypes.byref(s.c), d3(p).ctypes.data, 1,
                                            0xFFFFFFFF, None, None) for s in singles]
            assert d == min(each)
            assert each[c.value] == d


def test_march_agrees_with_dense_scan():
    """Binary visibility of the sphere-traced march (sdf.wgsl:61-81) against a
    1 mm fixed-step scan of the same conservative distance field."""
    cs = ConfigScene(CONFIGS["plain-640x400"].scaled(64, 40, 1))
    rec = oracle.SceneRecords.from_scene(cs.scene)
    f = frame_for(rec, len(rec.objects))
    rng = np.random.default_rng(11)
    agree = checked = 0
    for _ in range(60):
        a = np.array([rng.uniform(-1.8, 1.8), rng.uniform(0.2, 2.8), rng.uniform(-2.3, 2.3)])
        b = np.array([rng.uniform(-1.8, 1.8), rng.uniform(0.2, 2.8), rng.uniform(-2.3, 2.3)])
        da = L.oracle_scene_distance(64, f.ctypes.data, ctypes.byref(rec.c), d3(a).ctypes.data, 0, 0xFFFFFFFF, None, None)
        db = L.oracle_scene_distance(64, f.ctypes.data, ctypes.byref(rec.c), d3(b).ctypes.data, 0, 0xFFFFFFFF, None, None)
        if da < 0.05 or db < 0.05:
            continue
        n = np.zeros(3)
        vis = L.oracle_visibility(64, f.ctypes.data, ctypes.byref(rec.c), d3(a).ctypes.data, d3(n).ctypes.data,
                                  d3(b).ctypes.data, d3(n).ctypes.data, None)
        ts = np.linspace(0, 1, int(np.linalg.norm(b - a) / 0.001) + 2)
        blocked = False
        for t in ts[::1]:
            p = a + (b - a) * t
            if L.oracle_scene_distance(64, f.ctypes.data, ctypes.byref(rec.c), d3(p).ctypes.data, 0, 0xFFFFFFFF,
                                       None, None) < 0:
                blocked = True
                break
        checked += 1
        if blocked:
            assert vis == 0.0, "march passed through geometry"
        agree += 1
    assert checked >= 20
