"""Pass timers on the GPU: every pass of the fused schedule reports its GPU
time (read_gpu_time, renderer.rs:2066-2108), and irt_enable_timing_for
restricts the record to one pass."""
from __future__ import annotations

import pytest

from instantradiositytest_amd.renderer import Renderer
from instantradiositytest_amd.scenes import CONFIGS, ConfigScene

pytestmark = pytest.mark.gpu


def test_single_pass_timing_records_only_that_pass(require_gpu):
    print("[timing] fused 160x100, all passes then the march only", flush=True)
    cfg = CONFIGS["plain-640x400"].scaled(160, 100, 6)
    cs = ConfigScene(cfg)
    r = Renderer(cs.scene, cfg.settings(), precision="mixed", schedule="fused")
    try:
        r.encode(cs.scene, cs.frame_inputs(0))
        r.enable_timing(True)
        for fi in (1, 2):
            r.encode(cs.scene, cs.frame_inputs(fi))
        every = r.pass_times()
        assert {"reservoirs_a", "spatial_setup", "march", "resolve_combine", "svgf_temporal", "svgf_variance",
                "svgf_atrous0"} <= set(every), every
        r.enable_timing(True, only="march")
        for fi in (3, 4, 5):
            r.encode(cs.scene, cs.frame_inputs(fi))
        only = r.pass_times()
        print(every, only, flush=True)
        assert set(only) == {"march"}
        r.enable_timing(False)
        r.encode(cs.scene, cs.frame_inputs(5))
        assert r.pass_times() == {}
    finally:
        r.close()
