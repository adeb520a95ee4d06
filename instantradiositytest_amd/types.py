"""Record layouts of the C ABI (include/irt.h) as numpy dtypes and ctypes structs.

The layouts are those of the reference's GPU records (src/renderer.rs:8-66,
src/shaders/common.wgsl:3-16), so arrays built here are byte-identical to
what renderer.rs uploads.
"""
from __future__ import annotations

import ctypes

import numpy as np

F4x16 = ("<f4", (16,))
F4x4 = ("<f4", (4,))
U4x4 = ("<u4", (4,))

FRAME_DT = np.dtype([
    ("vp", F4x16), ("inv_vp", F4x16), ("previous_vp", F4x16), ("previous_inv_vp", F4x16),
    ("camera", F4x4), ("size", U4x4), ("counts", U4x4), ("restir", U4x4),
    ("temporal", F4x4), ("sdf", F4x4), ("rr", F4x4), ("flags", U4x4),
])
OBJECT_DT = np.dtype([("model", F4x16), ("previous", F4x16), ("inverse", F4x16), ("info", U4x4)])
MESH_DT = np.dtype([
    ("lo", F4x4), ("hi", F4x4), ("geometry", U4x4), ("field", U4x4), ("grid", F4x4),
    ("geometry_lo", F4x4), ("geometry_hi", F4x4), ("enclosing_sphere", F4x4),
])
LIGHT_DT = np.dtype([("position_strength", F4x4), ("direction_angle", F4x4), ("color_type", F4x4)])
NODE_DT = np.dtype([("lo", F4x4), ("hi", F4x4), ("info", U4x4)])
VPL_DT = np.dtype([("position", F4x4), ("normal", F4x4), ("flux", F4x4)])
RESERVOIR_DT = np.dtype([
    ("id", "<u4"), ("source", "<u4"), ("count", "<u4"), ("age", "<u4"),
    ("sum", "<f4"), ("selected_target", "<f4"), ("pdf", "<f4"), ("generation", "<u4"),
])

VERTEX_DT = np.dtype([("position", F4x4), ("normal", F4x4), ("tangent", F4x4), ("uv", F4x4)])
MATERIAL_DT = np.dtype([("color", F4x4), ("params", F4x4), ("textures", ("<i4", (4,))), ("channels", U4x4)])
TRI_NODE_DT = np.dtype([("lo", "<f4", (3,)), ("link", "<u4"), ("hi", "<f4", (3,)), ("count", "<u4")])
TRI_DT = np.dtype([("a", "<f4", (3,)), ("prim", "<u4"), ("b", "<f4", (3,)), ("pad0", "<u4"),
                   ("c", "<f4", (3,)), ("pad1", "<u4")])

assert FRAME_DT.itemsize == 384
assert VERTEX_DT.itemsize == 64
assert MATERIAL_DT.itemsize == 64
assert TRI_NODE_DT.itemsize == 32
assert TRI_DT.itemsize == 48
assert OBJECT_DT.itemsize == 208
assert MESH_DT.itemsize == 128
assert LIGHT_DT.itemsize == 48
assert NODE_DT.itemsize == 48
assert VPL_DT.itemsize == 48
assert RESERVOIR_DT.itemsize == 32

# Synthetic-scene records (scenegen.cpp).
SHAPE_DT = np.dtype([("kind", "<u4"), ("resolution", "<u4"), ("p", "<f4", (3,))])
SGOBJECT_DT = np.dtype([
    ("mesh", "<u4"), ("position", "<f4", (3,)), ("rotation", "<f4", (4,)), ("scale", "<f4"),
    ("albedo", "<f4", (3,)), ("rough", "<f4"), ("metal", "<f4"),
])
CAMERA_DT = np.dtype([
    ("position", "<f4", (3,)), ("rotation", "<f4", (4,)), ("fov_y", "<f4"), ("near_plane", "<f4"),
    ("far_plane", "<f4"),
])
SG_BOX, SG_SPHERE, SG_TORUS = 0, 1, 2


class Settings(ctypes.Structure):
    """Mirror of the reference Settings record (src/settings.rs:3-28)."""

    _fields_ = [
        ("render_width", ctypes.c_uint32),
        ("render_height", ctypes.c_uint32),
        ("vpl_count", ctypes.c_uint32),
        ("vpl_max_bounces", ctypes.c_uint32),
        ("vpl_rr_min_survival", ctypes.c_float),
        ("vpl_rr_max_survival", ctypes.c_float),
        ("restir_direct_candidates", ctypes.c_uint32),
        ("restir_vpl_candidates", ctypes.c_uint32),
        ("restir_spatial_neighbors", ctypes.c_uint32),
        ("restir_spatial_radius", ctypes.c_uint32),
        ("depth_threshold", ctypes.c_float),
        ("normal_threshold", ctypes.c_float),
        ("temporal_history_weight", ctypes.c_float),
        ("temporal_max_history", ctypes.c_uint32),
        ("restir_max_history", ctypes.c_uint32),
        ("svgf_iterations", ctypes.c_uint32),
        ("sdf_max_steps", ctypes.c_uint32),
        ("sdf_hit_epsilon", ctypes.c_float),
        ("sdf_soft_shadow_k", ctypes.c_float),
        ("target_frame_ms", ctypes.c_float),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


# The reference's settings.json (defaults of the renderer).
DEFAULT_SETTINGS = {
    "render_width": 1920,
    "render_height": 1200,
    "vpl_count": 100,
    "vpl_max_bounces": 3,
    "vpl_rr_min_survival": 0.1,
    "vpl_rr_max_survival": 0.95,
    "restir_direct_candidates": 8,
    "restir_vpl_candidates": 8,
    "restir_spatial_neighbors": 4,
    "restir_spatial_radius": 12,
    "restir_max_history": 4,
    "svgf_iterations": 4,
    "depth_threshold": 0.03,
    "normal_threshold": 0.85,
    "temporal_history_weight": 0.9,
    "temporal_max_history": 32,
    "sdf_max_steps": 96,
    "sdf_hit_epsilon": 0.001,
    "sdf_soft_shadow_k": 16.0,
    "target_frame_ms": 8.0,
}


def make_settings(**overrides) -> Settings:
    """Settings from the reference defaults with keyword overrides.

    Unknown keys are rejected, as serde's deny_unknown_fields does
    (src/settings.rs:4)."""
    values = dict(DEFAULT_SETTINGS)
    for k, v in overrides.items():
        if k not in values:
            raise KeyError(f"unknown settings field {k!r}")
        values[k] = v
    s = Settings()
    for k, v in values.items():
        setattr(s, k, v)
    return s


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    """Data pointer of a C-contiguous numpy array."""
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return ctypes.c_void_p(a.ctypes.data)


class ServerConfig(ctypes.Structure):
    """irt_server_config (include/irt.h)."""
    _fields_ = [("max_sessions", ctypes.c_uint32), ("precision", ctypes.c_uint32),
                ("hbm_budget_bytes", ctypes.c_uint64), ("max_in_flight", ctypes.c_uint32),
                ("use_graph", ctypes.c_uint32), ("tile_shapes", ctypes.c_uint32 * 8)]


class FrameRequest(ctypes.Structure):
    """irt_frame_request: one frame of one session; G-buffer planes and VPLs are device pointers."""
    _fields_ = [("vp", ctypes.c_float * 16), ("camera", ctypes.c_float * 3), ("generation", ctypes.c_uint32),
                ("light_count", ctypes.c_uint32), ("debug_mode", ctypes.c_uint32), ("vpl_count", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("depth", ctypes.c_void_p), ("normal", ctypes.c_void_p),
                ("color", ctypes.c_void_p), ("params", ctypes.c_void_p), ("motion", ctypes.c_void_p),
                ("vpls", ctypes.c_void_p), ("tag", ctypes.c_uint64)]


class Completion(ctypes.Structure):
    """irt_completion: a finished frame."""
    _fields_ = [("tag", ctypes.c_uint64), ("session", ctypes.c_uint32), ("frame_index", ctypes.c_uint32),
                ("pixels", ctypes.c_uint64), ("submit_s", ctypes.c_double), ("launch_s", ctypes.c_double),
                ("done_s", ctypes.c_double), ("gpu_ms", ctypes.c_float), ("batch_views", ctypes.c_uint32)]


class ServerStatus(ctypes.Structure):
    """irt_server_status."""
    _fields_ = [("open_sessions", ctypes.c_uint32), ("idle_contexts", ctypes.c_uint32),
                ("steps_in_flight", ctypes.c_uint32), ("pending_requests", ctypes.c_uint32),
                ("device_bytes", ctypes.c_uint64), ("budget_bytes", ctypes.c_uint64)]
