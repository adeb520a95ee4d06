"""Device-side G-buffer ingest (irt_set_gbuffer_device) against its source.

The planes arrive in the reference formats (Depth32Float, Rgba16F, Rg16F).
Read back in those formats they must equal the input bit for bit, in both
profiles: the mixed profile stores the reference formats as they are (one
launch copies the five planes, or plain copies for unaligned pointers), the
FP32 profile widens binary16 to FP32 and narrows it again on read-back.
Odd pixel counts exercise the copy's dword tails.
"""
from __future__ import annotations

import numpy as np
import pytest

from instantradiositytest_amd import hip
from instantradiositytest_amd.renderer import Renderer
from instantradiositytest_amd.scenes import CONFIGS, ConfigScene

pytestmark = pytest.mark.gpu

PLANES = ("depth", "normal", "color", "params", "motion")


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
@pytest.mark.parametrize("w,h", [(37, 11), (64, 32)])
def test_device_gbuffer_ingest_is_exact(require_gpu, precision, w, h):
    print(f"[ingest] {precision} {w}x{h}", flush=True)
    cfg = CONFIGS["plain-640x400"].scaled(w, h, 2)
    cs = ConfigScene(cfg)
    r = Renderer(cs.scene, cfg.settings(), precision=precision, schedule="fused")
    try:
        for fi in range(2):
            inp = cs.frame_inputs(fi)
            arrays = (inp.depth, inp.normal, inp.color, inp.params, inp.motion, inp.vpls)
            bufs = tuple(hip.DeviceBuffer.from_array(a) for a in arrays)
            r.begin_frame_device(bufs, inp.vp, inp.camera, cs.scene.generation, len(cs.scene.lights),
                                 len(cs.scene.objects))
            r.sync()
            for name, src in zip(PLANES, arrays):
                got = r.read(name)
                assert got.tobytes() == np.ascontiguousarray(src).tobytes(), f"{name} differs after ingest"
            r.lighting()
            r.svgf()
            r.end_frame()
            r.finish_frame()
            r.sync()
    finally:
        r.close()
