"""MI355X (gfx950) per-pixel lighting + denoising path of the instant-radiosity renderer.

ReSTIR direct/VPL reservoirs, the conservative SDF soft-shadow resolve and
SVGF, as hand-written HIP kernels behind a C ABI (include/irt.h).  This
package is the host side over that ABI: record layouts, the Renderer
bookkeeping of the reference, synthetic scenes for tests and benchmarks.
"""
from .types import (DEFAULT_SETTINGS, FRAME_DT, LIGHT_DT, MESH_DT, NODE_DT, OBJECT_DT, RESERVOIR_DT, VPL_DT,
                    Settings, make_settings)

__all__ = ["DEFAULT_SETTINGS", "FRAME_DT", "LIGHT_DT", "MESH_DT", "NODE_DT", "OBJECT_DT", "RESERVOIR_DT",
           "VPL_DT", "Settings", "make_settings"]
