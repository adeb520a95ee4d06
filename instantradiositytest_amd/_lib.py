"""Loading of the in-tree native libraries and their C prototypes.

libirt.so is the product (gfx950 kernels + C ABI); libirt_scenegen.so is the
CPU synthetic-scene producer.  Both are built by ``build.py`` into
``instantradiositytest_amd/_lib``.  A missing library is an error: there is
no fallback path.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_size_t, c_uint32, c_uint64, c_void_p
from pathlib import Path

from .types import Completion, FrameRequest, ServerConfig, ServerStatus, Settings

LIB_DIR = Path(__file__).resolve().parent / "_lib"

_irt = None
_sg = None


class NativeLibraryMissing(RuntimeError):
    pass


def _load(name: str) -> ctypes.CDLL:
    path = LIB_DIR / name
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python instantradiositytest_amd/build.py`")
    return ctypes.CDLL(str(path))


def irt() -> ctypes.CDLL:
    """libirt.so with prototypes of include/irt.h."""
    global _irt
    if _irt is not None:
        return _irt
    L = _load("libirt.so")
    P = c_void_p
    u32 = c_uint32
    sigs = {
        "irt_abi_version": (u32, []),
        "irt_global_error": (c_char_p, []),
        "irt_settings_validate": (c_int, [POINTER(Settings)]),
        "irt_create": (c_int, [POINTER(Settings), u32, POINTER(c_void_p)]),
        "irt_destroy": (None, [P]),
        "irt_last_error": (c_char_p, [P]),
        "irt_set_schedule": (c_int, [P, u32]),
        "irt_set_sdf_atlas": (c_int, [P, P, u32, u32, u32]),
        "irt_set_meshes": (c_int, [P, P, u32]),
        "irt_set_objects": (c_int, [P, P, u32, P, u32]),
        "irt_set_lights": (c_int, [P, P, u32]),
        "irt_set_frame": (c_int, [P, P]),
        "irt_set_gbuffer": (c_int, [P, P, P, P, P, P]),
        "irt_set_gbuffer_device": (c_int, [P, P, P, P, P, P]),
        "irt_set_vpls": (c_int, [P, P, u32]),
        "irt_set_vpls_device": (c_int, [P, P, u32]),
        "irt_set_geometry": (c_int, [P, P, u32, P, u32, P, u32]),
        "irt_generate_vpls": (c_int, [P]),
        "irt_set_textures": (c_int, [P, P, u32, u32]),
        "irt_render_gbuffer": (c_int, [P]),
        "irt_present": (c_int, [P, P]),
        "irt_present_device": (c_int, [P, P]),
        "irt_host_srgb_thresholds": (None, [P]),
        "irt_bake_sdf": (c_int, [P, u32, u32, P, u32, u32, P, P, P, POINTER(c_float)]),
        "irt_bake_sdf_noalloc": (c_int, [P, u32, u32, P, u32, u32, P, P, P, POINTER(c_float)]),
        "irt_audit_report": (c_int, [P, c_char_p, c_size_t]),
        "irt_host_audit_selftest": (c_int, []),
        "irt_check_fast_division": (c_int, [P]),
        "irt_lighting": (c_int, [P]),
        "irt_svgf": (c_int, [P, u32, u32]),
        "irt_end_frame": (c_int, [P]),
        "irt_encode": (c_int, [P]),
        "irt_pass": (c_int, [P, u32, u32]),
        "irt_buffer_bytes": (c_size_t, [P, u32]),
        "irt_read": (c_int, [P, u32, P, c_size_t]),
        "irt_write": (c_int, [P, u32, P, c_size_t]),
        "irt_sync": (c_int, [P]),
        "irt_stream": (c_void_p, [P]),
        "irt_stream_svgf": (c_void_p, [P]),
        "irt_enable_timing": (c_int, [P, c_int]),
        "irt_enable_timing_for": (c_int, [P, c_char_p]),
        "irt_pass_times": (c_int, [P, P, POINTER(c_float), c_int]),
        "irt_enable_counters": (c_int, [P, c_int]),
        "irt_read_counters": (c_int, [P, POINTER(c_uint64)]),
        "irt_copy_scene": (c_int, [P, P]),
        "irt_device_bytes": (c_size_t, [P]),
        "irt_scene_bytes": (c_size_t, [P]),
        "irt_estimate_bytes": (c_size_t, [POINTER(Settings), u32]),
        "irt_reset": (c_int, [P]),
        "irt_server_now": (c_double, []),
        "irt_server_create": (c_int, [POINTER(ServerConfig), POINTER(c_void_p)]),
        "irt_server_destroy": (None, [P]),
        "irt_server_last_error": (c_char_p, [P]),
        "irt_server_open": (c_int, [P, POINTER(Settings), P, POINTER(c_uint32)]),
        "irt_server_context": (c_void_p, [P, u32]),
        "irt_server_close": (c_int, [P, u32]),
        "irt_server_submit": (c_int, [P, u32, POINTER(FrameRequest)]),
        "irt_server_step": (c_int, [P]),
        "irt_server_poll": (c_int, [P, POINTER(Completion), u32, c_int]),
        "irt_server_sync": (c_int, [P]),
        "irt_server_stats": (c_int, [P, POINTER(ServerStatus)]),
        "irt_batch_create": (c_int, [P, u32, POINTER(c_void_p)]),
        "irt_batch_destroy": (None, [P]),
        "irt_batch_last_error": (c_char_p, [P]),
        "irt_batch_set_views": (c_int, [P, P, u32]),
        "irt_batch_encode": (c_int, [P]),
        "irt_batch_set_tile": (c_int, [P, u32, u32]),
        "irt_batch_autotune": (c_int, [P, u32, POINTER(c_uint32)]),
        "irt_batch_use_graph": (c_int, [P, c_int]),
        "irt_batch_sync": (c_int, [P]),
        "irt_batch_stream": (c_void_p, [P]),
        "irt_batch_stream_svgf": (c_void_p, [P]),
        "irt_batch_encode_timed": (c_int, [P, POINTER(c_float)]),
        "irt_batch_pass_times": (c_int, [P, POINTER(c_float)]),
        "irt_host_is_box_geometry": (c_int, [P, u32, u32, P, u32, P, P]),
        "irt_host_field_lipschitz_bound": (c_float, [P, u32, P]),
        "irt_host_hierarchy": (c_int, [P, P, u32, P]),
        "irt_host_transform_box": (None, [P, P, P, P, P]),
        "irt_host_atlas_grid": (None, [u32, u32, P]),
        "irt_host_pack_frame": (c_int, [POINTER(Settings), P, P, P, u32, u32, u32, u32, u32, u32, u32, P]),
        "irt_host_inverse": (c_int, [P, P]),
        "irt_host_mesh_hierarchy": (c_int, [P, u32, P, u32, P, P]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _irt = L
    return L


def scenegen() -> ctypes.CDLL:
    """libirt_scenegen.so (synthetic scenes, G-buffers, VPL pools)."""
    global _sg
    if _sg is not None:
        return _sg
    L = _load("libirt_scenegen.so")
    P = c_void_p
    u32 = c_uint32
    sigs = {
        "sg_create": (c_void_p, [P, u32, P, u32, P, u32]),
        "sg_destroy": (None, [P]),
        "sg_info": (None, [P, P]),
        "sg_records": (None, [P, P, P, P, P]),
        "sg_atlas": (None, [P, P]),
        "sg_set_objects": (c_int, [P, P, u32]),
        "sg_set_lights": (c_int, [P, P, u32]),
        "sg_camera_matrix": (None, [P, c_float, P]),
        "sg_render_gbuffer": (c_int, [P, P, P, u32, u32, P, P, P, P, P]),
        "sg_vpls": (c_int, [P, c_uint64, u32, u32, c_float, P]),
        "sg_geometry_info": (None, [P, P]),
        "sg_geometry": (None, [P, P, P, P]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _sg = L
    return L


class IrtError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"irt error {code}: {message}")
        self.code = code
        self.message = message


def check(rc: int, ctx=None) -> int:
    if rc < 0:
        L = irt()
        msg = (L.irt_last_error(ctx) if ctx else L.irt_global_error()) or b""
        raise IrtError(rc, msg.decode())
    return rc
