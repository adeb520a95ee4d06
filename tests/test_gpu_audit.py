"""The per-view schedules under the stream-ordering audit (IRT_STREAM_AUDIT=1
at context creation): every operation a context enqueues declares what it
reads and writes, and no read or write may touch a buffer the other stream
used without a join, event wait or host synchronisation in between.  Covers
the fused schedule's lighting and SVGF streams, slot rotation across frames,
the G-buffer and VPL passes, presentation into a new buffer (the case of a
zero-fill on one stream and a store on the other), the SVGF bypass and the
per-pass schedule."""
from __future__ import annotations

import ctypes
import os

import numpy as np
import pytest

from instantradiositytest_amd import hip
from instantradiositytest_amd.renderer import Batch, Renderer
from instantradiositytest_amd.scenes import CONFIGS, ConfigScene

pytestmark = pytest.mark.gpu


@pytest.fixture
def audited():
    os.environ["IRT_STREAM_AUDIT"] = "1"
    try:
        yield
    finally:
        os.environ.pop("IRT_STREAM_AUDIT", None)


def report(r: Renderer):
    buf = ctypes.create_string_buffer(1 << 16)
    n = r.L.irt_audit_report(r.ctx, buf, len(buf))
    return n, buf.value.decode()


@pytest.mark.parametrize("precision,schedule,device_inputs", [("mixed", "fused", False), ("fp32", "fused", True),
                                                              ("fp32", "passes", False)])
def test_schedules_have_no_unordered_access(require_gpu, audited, precision, schedule, device_inputs):
    print(f"[audit] {precision} {schedule} device_inputs={device_inputs}", flush=True)
    cfg = CONFIGS["plain-640x400"].scaled(96, 64, 8)
    cs = ConfigScene(cfg)
    r = Renderer(cs.scene, cfg.settings(), precision=precision, schedule=schedule)
    try:
        if device_inputs:
            r.set_geometry(*cs.scene.geometry())
        for fi in range(6):
            inp = cs.frame_inputs(fi)
            if device_inputs:
                inp.depth = inp.vpls = None
            r.encode(cs.scene, inp)
            if fi == 2:
                buf = hip.DeviceBuffer(96 * 64 * 4)
                assert r.L.irt_present_device(r.ctx, buf.ptr) == 0
        r.present()
        r.debug_mode = 3  # the SVGF bypass
        r.encode(cs.scene, cs.frame_inputs(6))
        r.present()
        r.read("history")
        r.sync()
        n, text = report(r)
        print(text, flush=True)
        assert n == 0, text
        lit, svgf = (int(w) for w in text.split("audited ")[1].split() if w.isdigit())
        assert lit > 40  # every frame's ingest, guide, reservoirs, ... are seen
        assert svgf > 20 if schedule == "fused" else svgf == 0
    finally:
        r.close()


def test_audit_stops_when_a_context_joins_a_batch(require_gpu, audited):
    cfg = CONFIGS["plain-640x400"].scaled(64, 48, 2)
    cs = ConfigScene(cfg)
    r = Renderer(cs.scene, cfg.settings(), precision="mixed", schedule="fused")
    try:
        assert report(r)[0] == 0
        b = Batch([r])
        assert report(r)[0] == -4  # IRT_ERR_STATE: the batch's streams are not modelled
        b.close()
    finally:
        r.close()
