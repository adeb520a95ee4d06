#!/usr/bin/env python3
"""Benchmark of the per-pixel lighting + denoising path on one MI355X.

A step is one frame of the path for one view: the ReSTIR candidate /
temporal / spatial passes, both resolves with the SDF soft-shadow march,
and SVGF (temporal, variance, four a-trous) at 1920x1200 with the
reference's default settings, on the synthetic 24-mesh / 128^3-SDF /
16-light / 1000-VPL room (orbit + 0.5 m dolly).  Every frame's G-buffer and
VPL pool are generated up front and staged in HBM; the timed region starts
with them resident and covers G-buffer ingest + lighting + SVGF + history
per frame.

Legs (one JSON line carries all of them; the G-buffer and VPL passes are
reported beside the headline and together with it in full_frame, because the
headline frame takes its G-buffer and VPL pool as staged inputs):
  headline   fused kernels, mixed storage (16-bit G-buffer / radiance / SDF
             samples, FP32 reservoirs, moments and march), single view
  plain      one kernel per reference pass, FP32 storage: the plain FP32
             path the headline is compared with
  fused_fp32 fused kernels, FP32 storage
  vpls       the VPL pass (vpl.wgsl) on the same scene's triangles: launch
             time, rays/s, and the oracle's rate for the same pool
  gbuffer    the G-buffer pass (gbuffer.wgsl raster) of the same scene:
             launch time and the oracle's time for the same frame
  full       the whole reference frame per view on the GPU: G-buffer pass,
             VPL pass, lighting, SVGF, history, presentation (nothing staged)
  bake       SDF bakes of all the scene's meshes at their resolutions
             (the exporter's algorithm on the GPU, for scene admission)
  batched    16 views (4 x 1920x1200, 4 x 1280x720, 8 x 640x360), one
             launch per pass, autotuned tiles, HIP graph replay
  serving    Poisson arrivals (40/s) of sessions (640x360 / 1280x720 /
             1920x1200 at 0.4 / 0.4 / 0.2, lifetimes U[120,600] frames), each
             with its own seeded scene whose G-buffer and VPL pool are rendered
             on the device every frame, up to 32 resident sessions in a
             200 GB budget, 60 Hz clients, continuous batching: p50 / p99
             frame latency (the line's p50_ms / p99_ms) and Mpix/s
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "shaded Mpix/s at 1920x1200 + p50/p99 frame ms @32 sessions vs plain FP32"

# Algorithmic FLOPs per event of the reference WGSL and the peaks; the
# op-by-op count behind each weight is bench/roofline.md.
_ROOFLINE = json.loads((ROOT / "bench" / "roofline.json").read_text())
FLOP_WEIGHTS = _ROOFLINE["weights"]
MARCH_EVENTS = ("node", "obj_box", "obj_grid", "step", "march")
FP32_VALU_PEAK_TFLOPS = _ROOFLINE["fp32_vector_peak_tflops"]
HBM_PEAK_GBS = _ROOFLINE["hbm_peak_gbs"]


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def flops_of(counts: dict, keys=None) -> float:
    keys = FLOP_WEIGHTS if keys is None else keys
    return float(sum(FLOP_WEIGHTS[k] * counts.get(k, 0) for k in keys))


def cpu_baseline(cs, settings, frames: int, threads: int) -> dict:
    """The FP32 CPU restatement over the first `frames` frames of the workload."""
    import oracle
    from instantradiositytest_amd.renderer import pack_frame
    cfg = cs.cfg
    rec = oracle.SceneRecords.from_scene(cs.scene)
    o = oracle.OracleRenderer(32, cfg.width, cfg.height, cfg.vpl_count, settings.restir_max_history, threads=threads)
    prev_vp = None
    elapsed = 0.0
    for fi in range(frames):
        inp = cs.frame_inputs(fi)
        vp = inp.vp
        f = pack_frame(settings, vp, vp if prev_vp is None else prev_vp, inp.camera, fi, 0, 0, fi > 0, 0,
                       len(cs.scene.lights), len(cs.scene.objects))
        t0 = time.perf_counter()
        o.encode(f, rec, inp, settings.svgf_iterations, 0)
        elapsed += time.perf_counter() - t0
        prev_vp = vp
        log(f"[cpu] frame {fi} cumulative {elapsed:.1f}s")
    px = cfg.width * cfg.height * frames
    return {"value": round(px / elapsed / 1e6, 4), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": f"FP32 CPU restatement (oracle, OpenMP) of frames 0-{frames - 1} of the same workload, "
                      f"{cfg.width}x{cfg.height}, all lighting passes + SVGF, {elapsed:.1f} s"}


def pass_time_entries(r) -> list:
    cap = 8192
    names = (ctypes.c_char * 32 * cap)()
    ms = (ctypes.c_float * cap)()
    n = r.L.irt_pass_times(r.ctx, ctypes.cast(names, ctypes.c_void_p), ms, cap)
    return [(bytes(names[i]).split(b"\0", 1)[0].decode(), float(ms[i])) for i in range(max(n, 0))]


def run_single(cs, settings, staged, host_frames, precision: str, schedule: str, warmup: int, steps: int):
    """Frames of one view from staged inputs; per-frame GPU time from events
    on the view's stream, per-kernel times from the library's timers."""
    from instantradiositytest_amd import hip
    from instantradiositytest_amd._lib import check
    from instantradiositytest_amd.renderer import Renderer
    r = Renderer(cs.scene, settings, precision=precision, schedule=schedule)
    L, ctx = r.L, r.ctx
    stream = L.irt_stream(ctx)
    # a frame's last work: SVGF's stream under the fused schedule
    done_stream = L.irt_stream_svgf(ctx) if schedule == "fused" else stream
    gen, nl, no = cs.scene.generation, len(cs.scene.lights), len(cs.scene.objects)

    def submit(fi: int):
        vp, cam = host_frames[fi]
        r.begin_frame_device(staged[fi], vp, cam, gen, nl, no)
        check(L.irt_encode(ctx), ctx)
        r.finish_frame()

    for fi in range(warmup):
        submit(fi)
    r.sync()
    # The timed frames time the march only (the fused schedule's dominant
    # kernel, two events per frame); the other passes' times come from a few
    # frames after the timed region with every pass timed.
    timed_only = "march" if schedule == "fused" else None
    r.enable_timing(True, only=timed_only)
    starts = [hip.Event() for _ in range(steps)]
    ends = [hip.Event() for _ in range(steps)]
    t0 = time.perf_counter()
    for k in range(steps):
        starts[k].record(stream)
        submit(warmup + k)
        ends[k].record(done_stream)
    r.sync()
    wall = time.perf_counter() - t0
    frame_ms = np.array([starts[k].elapsed_ms(ends[k]) for k in range(steps)])
    per = {}
    for name, ms in pass_time_entries(r):
        per.setdefault(name, []).append(ms)
    if timed_only:
        extra = min(steps, 10)
        r.enable_timing(True)
        for k in range(extra):
            submit(warmup + k)
        r.sync()
        for name, ms in pass_time_entries(r):
            if name != timed_only:
                per.setdefault(name, []).append(ms)
        launches = {k: len(v) // (steps if k == timed_only else extra) for k, v in per.items()}
    else:
        launches = {k: len(v) // steps for k, v in per.items()}
    r.enable_timing(False)
    W, H = settings.render_width, settings.render_height
    out = {"value": round(W * H * steps / wall / 1e6, 3), "ms_per_step": round(wall * 1e3 / steps, 4),
           "p50_ms": round(float(np.percentile(frame_ms, 50)), 4),
           "p99_ms": round(float(np.percentile(frame_ms, 99)), 4),
           "kernel_ms": {k: round(float(np.mean(v)), 4) for k, v in per.items()},
           "kernel_launches_per_frame": launches}
    if timed_only:
        out["kernel_ms_note"] = (f"{timed_only}: HIP events in the timed frames; other passes: {extra} frames after "
                                 f"the timed region with every pass timed")
    log(f"[bench] {schedule}/{precision}: {out['value']} Mpix/s, {out['ms_per_step']} ms/frame, kernels {out['kernel_ms']}")
    return out, r


def run_host_inputs(cs, settings, host_inputs, warmup: int, steps: int) -> dict:
    """The headline workload with each frame's G-buffer and VPL pool handed over in (pageable) host
    memory through irt_set_gbuffer / irt_set_vpls: the PCIe-inclusive rate.  The uploads are
    synchronous (the caller may reuse its arrays once the call returns), so a frame's upload does not
    overlap the previous frame's work on the lighting stream."""
    from instantradiositytest_amd.renderer import Renderer
    r = Renderer(cs.scene, settings, precision="mixed", schedule="fused")
    n = len(host_inputs)
    for k in range(warmup):
        r.encode(cs.scene, host_inputs[k % n])
    r.sync()
    t0 = time.perf_counter()
    for k in range(steps):
        r.encode(cs.scene, host_inputs[(warmup + k) % n])
    r.sync()
    wall = time.perf_counter() - t0
    r.close()
    W, H = settings.render_width, settings.render_height
    per_frame = sum(a.nbytes for a in (host_inputs[0].depth, host_inputs[0].normal, host_inputs[0].color,
                                       host_inputs[0].params, host_inputs[0].motion, host_inputs[0].vpls))
    out = {"mpix_s": round(W * H * steps / wall / 1e6, 3), "ms_per_step": round(wall * 1e3 / steps, 4),
           "upload_bytes_per_frame": per_frame, "frames": steps,
           "note": f"headline workload with inputs uploaded from pageable host memory each frame "
                   f"({per_frame / 1e6:.1f} MB: G-buffer planes + VPL pool; {n} distinct frames cycled); "
                   f"never the headline value"}
    log(f"[bench] host inputs: {out}")
    return out


def reference_counts(r, cs, settings, staged, host_frames, fi: int) -> dict:
    """Reference-algorithm event counts of one frame, pass by pass (the
    per-pass kernels walk the hierarchy as sdf.wgsl does), on the frame
    after the timed ones."""
    from instantradiositytest_amd._lib import check
    L, ctx = r.L, r.ctx
    r.schedule("passes")
    vp, cam = host_frames[fi]
    r.begin_frame_device(staged[fi], vp, cam, cs.scene.generation, len(cs.scene.lights), len(cs.scene.objects))
    out = {}
    names = ["candidates_direct", "candidates_indirect", "temporal_direct", "temporal_indirect", "spatial_direct",
             "spatial_indirect", "resolve_direct", "resolve_indirect"]
    for i, name in enumerate(names):
        r.enable_counters(True)
        check(L.irt_pass(ctx, i, 0), ctx)
        out[name] = r.counters()
    r.enable_counters(True)
    check(L.irt_pass(ctx, 8, 0), ctx)
    out["svgf_temporal"] = r.counters()
    r.enable_counters(True)
    check(L.irt_pass(ctx, 9, 0), ctx)
    out["svgf_variance"] = r.counters()
    for it in range(settings.svgf_iterations):
        r.enable_counters(True)
        check(L.irt_pass(ctx, 10, it), ctx)
        out[f"svgf_atrous{it}"] = r.counters()
    r.enable_counters(False)
    r.sync()
    return out


def kernel_flops(ref: dict) -> dict:
    """Algorithmic FLOPs of each kernel of the fused schedule (and of the
    per-pass kernels) from the reference-algorithm counts."""
    add = lambda *ds: {k: sum(d.get(k, 0) for d in ds) for k in FLOP_WEIGHTS}
    res_d, res_i = ref["resolve_direct"], ref["resolve_indirect"]
    out = {
        "reservoirs_a": flops_of(add(ref["candidates_direct"], ref["candidates_indirect"], ref["temporal_direct"],
                                     ref["temporal_indirect"])),
        "spatial_setup": flops_of(add(ref["spatial_direct"], ref["spatial_indirect"])) +
        FLOP_WEIGHTS["brdf"] * (res_d.get("brdf", 0) + res_i.get("brdf", 0)),
        "march": flops_of(add(res_d, res_i), MARCH_EVENTS),
        "visibility_direct": flops_of(res_d),
        "visibility_indirect": flops_of(res_i),
    }
    for k, v in ref.items():
        if k.startswith("svgf") or k.startswith("candidates") or k.startswith("temporal") or k.startswith("spatial"):
            out[k] = flops_of(v)
    return out


def run_batched(frames: int, warmup: int, steps: int, graph: bool = True) -> dict:
    """16 views of the batched workload in one launch per pass."""
    from instantradiositytest_amd.renderer import Batch, Renderer
    from instantradiositytest_amd.serving import SceneStream, batched_view_configs
    t0 = time.perf_counter()
    cfgs = batched_view_configs(frames)
    streams = [SceneStream(c, "mixed", frames) for c in cfgs]
    log(f"[batch] 16 streams staged in {time.perf_counter() - t0:.1f}s")
    views = []
    for st in streams:
        r = Renderer(st.cs.scene, st.settings, precision="mixed", schedule="fused")
        views.append(r)
    batch = Batch(views, graph=graph)
    px = sum(st.pixels for st in streams)

    def begin(fi):
        for r, st in zip(views, streams):
            f = st.frames[fi % frames]
            r.begin_frame_device(f.bufs, f.vp, f.camera, st.cs.scene.generation, len(st.cs.scene.lights),
                                 len(st.cs.scene.objects))

    begin(0)
    picks = batch.autotune_current(repeats=2)
    batch.finish()
    for fi in range(1, warmup):
        begin(fi)
        batch.encode_current()
        batch.finish()
    batch.sync()
    timed = None
    begin(warmup)
    timed = batch.encode_timed_current()
    batch.finish()
    t0 = time.perf_counter()
    for k in range(steps):
        begin(warmup + 1 + k)
        batch.encode_current()
        batch.finish()
    batch.sync()
    wall = time.perf_counter() - t0
    out = {"workload": "16 views: 4x1920x1200 + 4x1280x720 + 8x640x360, own seeds, 8-48 meshes at 64^3-128^3, "
                       "300-2000 VPLs; mixed storage; one launch per pass; HIP graph replay",
           "pixels_per_frame": px, "frames": steps, "mpix_s": round(px * steps / wall / 1e6, 3),
           "ms_per_batched_frame": round(wall * 1e3 / steps, 4), "tile_shapes": picks,
           "kernel_ms": {k: round(v, 4) for k, v in timed.items()}, "graph": graph}
    log(f"[batch] {out['mpix_s']} Mpix/s, {out['ms_per_batched_frame']} ms per batched frame, tiles {picks}")
    batch.close()
    for r in views:
        r.close()
    return out


def run_vpls(cs, settings, host_frames, steps: int, threads: int, cpu: bool) -> dict:
    """The VPL pass (irt_generate_vpls) on the scene's triangles: mean launch
    time over `steps` launches of frame 0's pool, rays traced (from the
    oracle's run of the same pool), and the oracle's time for that pool."""
    import oracle
    from instantradiositytest_amd._lib import check
    from instantradiositytest_amd.renderer import Renderer, pack_frame
    from instantradiositytest_amd.types import ptr
    r = Renderer(cs.scene, settings, precision="mixed", schedule="fused")
    try:
        vertices, indices, materials = cs.scene.geometry()
        r.set_geometry(vertices, indices, materials)
        vp, cam = host_frames[0]
        f = pack_frame(settings, vp, vp, cam, 0, cs.scene.generation, 0, False, 0, len(cs.scene.lights),
                       len(cs.scene.objects))
        check(r.L.irt_set_frame(r.ctx, ptr(f)), r.ctx)
        for _ in range(3):
            r.generate_vpls()
        r.sync()
        r.enable_timing(True)
        for _ in range(steps):
            r.generate_vpls()
        r.sync()
        ms = [t for name, t in pass_time_entries(r) if name == "vpl_generate"]
        r.enable_timing(False)
        t0 = time.perf_counter()
        _, rays = oracle.generate_vpls(f, oracle.SceneRecords.from_scene(cs.scene),
                                       oracle.GeometryRecords(vertices, indices, materials), threads=threads,
                                       with_rays=True)
        cpu_s = time.perf_counter() - t0
    finally:
        r.close()
    launch_ms = float(np.mean(ms))
    out = {"launch_ms": round(launch_ms, 4), "launches": len(ms), "paths": settings.vpl_count,
           "max_bounces": settings.vpl_max_bounces, "rays": int(rays), "triangles": int(len(indices) // 3),
           "objects": len(cs.scene.objects), "mrays_s": round(rays / launch_ms / 1e3, 2)}
    if cpu:
        out["cpu_baseline"] = {"value": round(rays / cpu_s / 1e6, 4), "unit": "Mrays/s", "cores": threads,
                               "kind": "port", "sample": f"the oracle's VPL pass (every ray tested against every "
                                                         f"triangle) for the same pool, {cpu_s:.2f} s"}
    log(f"[bench] vpls: {out}")
    return out


def run_gbuffer(cs, settings, host_frames, steps: int, threads: int, cpu: bool) -> dict:
    """The G-buffer pass (irt_render_gbuffer) of frame 0 at the workload's size:
    mean time over `steps` launches, and the oracle's time for the same frame."""
    import oracle
    from instantradiositytest_amd._lib import check
    from instantradiositytest_amd.renderer import Renderer, pack_frame
    from instantradiositytest_amd.types import ptr
    r = Renderer(cs.scene, settings, precision="mixed", schedule="fused")
    try:
        geo = cs.scene.geometry()
        r.set_geometry(*geo)
        vp, cam = host_frames[1]
        pvp, _ = host_frames[0]
        f = pack_frame(settings, vp, pvp, cam, 1, cs.scene.generation, cs.scene.generation, True, 0,
                       len(cs.scene.lights), len(cs.scene.objects))
        check(r.L.irt_set_frame(r.ctx, ptr(f)), r.ctx)
        for _ in range(3):
            r.render_gbuffer()
        r.sync()
        r.enable_timing(True)
        for _ in range(steps):
            r.render_gbuffer()
        r.sync()
        ms = [t for name, t in pass_time_entries(r) if name == "gbuffer"]
        r.enable_timing(False)
        cpu_s = None
        if cpu:
            t0 = time.perf_counter()
            oracle.render_gbuffer(f, oracle.SceneRecords.from_scene(cs.scene), oracle.GeometryRecords(*geo))
            cpu_s = time.perf_counter() - t0
    finally:
        r.close()
    W, H = settings.render_width, settings.render_height
    objs = cs.scene.objects
    inst = int(sum(cs.scene.meshes[o["info"][0]]["geometry"][3] // 3 for o in objs))
    launch_ms = float(np.mean(ms))
    out = {"launch_ms": round(launch_ms, 4), "launches": len(ms), "width": W, "height": H,
           "instance_triangles": inst, "mpix_s": round(W * H / launch_ms / 1e3, 1)}
    if cpu_s is not None:
        out["cpu_baseline"] = {"value": round(W * H / cpu_s / 1e6, 3), "unit": "Mpix/s", "cores": 1, "kind": "port",
                               "sample": f"the oracle's G-buffer pass (raster loop on one core, resolve on "
                                         f"{threads} threads) for the same frame, {cpu_s:.2f} s"}
    log(f"[bench] gbuffer: {out}")
    return out


def run_bake(cs, threads: int, cpu: bool) -> dict:
    """SDF bakes of every mesh of the scene at its own resolution (irt_bake_sdf:
    tools/export_scene.py:121-158 plus the Lipschitz bound), wall time of the
    blocking calls; the oracle's rate on the largest mesh at 24^3."""
    import oracle
    from instantradiositytest_amd import _lib
    from instantradiositytest_amd.types import ptr
    L = _lib.irt()
    vertices, indices, _ = cs.scene.geometry()
    meshes = []
    for m in cs.scene.meshes:
        v0, nv, i0, ni = (int(x) for x in m["geometry"])
        pos = np.ascontiguousarray(vertices["position"][v0:v0 + nv, :3], np.float32)
        meshes.append((pos, np.ascontiguousarray(indices[i0:i0 + ni], np.uint32), int(m["field"][3])))
    lip = ctypes.c_float()

    def bake(pos, idx, res):
        out = np.empty(res ** 3, np.float32)
        lo, hi = np.zeros(3, np.float32), np.zeros(3, np.float32)
        rc = L.irt_bake_sdf(ptr(pos), 3, len(pos), ptr(idx), len(idx), res, ptr(out), ptr(lo), ptr(hi),
                            ctypes.byref(lip))
        if rc:
            raise RuntimeError(L.irt_global_error().decode())
        return out

    bake(*meshes[0])  # first use: code objects
    t0 = time.perf_counter()
    samples = 0
    for pos, idx, res in meshes:
        bake(pos, idx, res)
        samples += res ** 3
    wall = time.perf_counter() - t0
    out = {"meshes": len(meshes), "samples": samples, "triangles": int(sum(len(i) // 3 for _, i, _ in meshes)),
           "wall_ms": round(wall * 1e3, 2), "msamples_s": round(samples / wall / 1e6, 1)}
    if cpu:
        pos, idx, _ = max(meshes, key=lambda m: len(m[1]))
        t0 = time.perf_counter()
        oracle.bake_sdf(pos, idx, 24, threads=threads)
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(24 ** 3 / cpu_s / 1e6, 4), "unit": "Msamples/s", "cores": threads,
                               "kind": "port", "sample": f"the oracle's bake (every triangle tested) of the largest "
                                                         f"mesh ({len(idx) // 3} triangles) at 24^3, {cpu_s:.2f} s"}
    log(f"[bench] bake: {out}")
    return out


def run_full_frame(cs, settings, host_frames, warmup: int, steps: int) -> dict:
    """The whole reference frame on the GPU for one view: G-buffer pass, VPL
    pass, lighting, SVGF, history and presentation into the sRGB display
    (Renderer::encode's passes), fused kernels, mixed storage; the host
    uploads only the objects, lights and the Frame record each frame."""
    from instantradiositytest_amd._lib import check
    from instantradiositytest_amd.renderer import Renderer
    from instantradiositytest_amd.scenes import FrameInputs
    from instantradiositytest_amd import hip
    r = Renderer(cs.scene, settings, precision="mixed", schedule="fused")
    display = hip.DeviceBuffer(settings.render_width * settings.render_height * 4)
    try:
        r.set_geometry(*cs.scene.geometry())

        def submit(fi: int):
            vp, cam = host_frames[fi]
            r.begin_frame(cs.scene, FrameInputs(None, None, None, None, None, None, vp, cam))
            check(r.L.irt_encode(r.ctx), r.ctx)
            check(r.L.irt_present_device(r.ctx, display.ptr), r.ctx)  # renderer.rs:1961
            r.finish_frame()

        for fi in range(warmup):
            submit(fi)
        r.sync()
        r.enable_timing(True)
        t0 = time.perf_counter()
        for k in range(steps):
            submit(warmup + k)
        r.sync()
        wall = time.perf_counter() - t0
        per = {}
        for name, ms in pass_time_entries(r):
            per.setdefault(name, []).append(ms)
        r.enable_timing(False)
    finally:
        r.close()
    W, H = settings.render_width, settings.render_height
    out = {"mpix_s": round(W * H * steps / wall / 1e6, 3), "ms_per_step": round(wall * 1e3 / steps, 4),
           "kernel_ms": {k: round(float(np.mean(v)), 4) for k, v in per.items()}}
    log(f"[bench] full frame: {out}")
    return out


def build_session_scenes(classes: list, frames: int, seed: int) -> list:
    """One seeded scene per arriving session (arrival order), each with its G-buffer and VPL pool produced on
    the device; the host copy of each scene is dropped once its template context holds it."""
    from concurrent.futures import ThreadPoolExecutor

    from instantradiositytest_amd.scenes import ConfigScene
    from instantradiositytest_amd.serving import SERVING_MIX, SceneStream, session_config
    rng = np.random.default_rng(seed)
    cfgs = [session_config(seed * 1000 + i, SERVING_MIX[k][0], SERVING_MIX[k][1], frames, rng)
            for i, k in enumerate(classes)]
    t0 = time.perf_counter()
    streams = []
    with ThreadPoolExecutor(4) as ex:  # the SDF bakes inside run OpenMP loops; a few in parallel fill the cores
        for cs in ex.map(ConfigScene, cfgs):
            st = SceneStream(cs.cfg, "mixed", frames, device_inputs=True, cs=cs)
            st.release_host()
            streams.append(st)
            if len(streams) % 16 == 0:
                log(f"[serve] {len(streams)}/{len(cfgs)} session scenes built ({time.perf_counter() - t0:.1f}s)")
    log(f"[serve] {len(streams)} session scenes built in {time.perf_counter() - t0:.1f}s")
    return streams


def run_serving(streams, arrivals, seconds: float, warmup_s: float, sessions: int, tile_shapes=None,
                max_in_flight: int = 2, cadence_hz: float = 60.0) -> dict:
    """The serving simulation; arrival i uses scene i, and arrivals beyond the built scenes reuse them
    cyclically (counted in scenes_reused)."""
    from instantradiositytest_amd.serving import Arrival, Server, simulate
    n = len(streams)
    arrivals = [Arrival(a.t, i % n, a.lifetime, a.offset) for i, a in enumerate(arrivals)]
    srv = Server(max_sessions=sessions, precision="mixed", budget_bytes=200e9, max_in_flight=max_in_flight,
                 tile_shapes=tile_shapes)
    rep = simulate(srv, streams, arrivals, seconds, warmup_s, cadence_hz, log=log)
    d = rep.as_dict()
    d["workload"] = (f"Poisson arrivals 40/s over {seconds:.0f} s, sessions 640x360/1280x720/1920x1200 at "
                     f"0.4/0.4/0.2, lifetimes U[120,600] frames, <= {sessions} resident sessions, 200 GB budget, "
                     f"{cadence_hz:g} Hz clients with one request outstanding, continuous batching ({max_in_flight} steps in "
                     f"flight); every session its own seeded scene (8-48 meshes at 64^3-128^3, 500-2000 VPLs, "
                     f"3-5 bounces) with its G-buffer and VPL pool rendered on the device each frame; latency = "
                     f"submit to completion; window after {warmup_s:.0f} s warm-up")
    d["max_sessions"] = sessions
    d["distinct_scenes"] = n
    d["scenes_reused"] = max(0, rep.sessions_admitted - n)
    srv.close_server()
    log(f"[serve] {d}")
    return d


def measured_traffic(config: str, kernel: str):
    """Bytes per launch of the kernel between L2 and memory, from the two
    rocprofv3 counter passes summarised in profiles/ (FETCH_SIZE and
    WRITE_SIZE, calibrated on a copy kernel of known size); None when no
    summary covers this workload."""
    path = Path(__file__).resolve().parent / "profiles" / "hbm_traffic.json"
    if config != "plain-1920x1200" or not path.exists():
        return None
    per = json.loads(path.read_text())["per_launch"]
    rec = per.get("k_" + kernel)
    return rec["traffic_bytes"] if rec else None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--config", default="plain-1920x1200")
    ap.add_argument("--cpu-frames", type=int, default=3)
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--serve-seconds", type=float, default=25.0)
    ap.add_argument("--serve-warmup", type=float, default=5.0)
    ap.add_argument("--sessions", type=int, default=32)
    ap.add_argument("--session-scenes", type=int, default=128,
                    help="distinct seeded session scenes built for the serving run (arrival order)")
    ap.add_argument("--serve-sweep", type=lambda v: [int(x) for x in v.split(",") if x], default=[4, 8, 16],
                    help="resident-session caps of the short serving runs after the main one")
    ap.add_argument("--sweep-seconds", type=float, default=10.0)
    ap.add_argument("--serve-cadences", type=lambda v: [float(x) for x in v.split(",") if x], default=[30.0, 24.0, 20.0],
                    help="client frame rates (Hz) of short serving runs at the full session count")
    ap.add_argument("--only", default="",
                    help="comma list of legs: headline,plain,fused_fp32,host_inputs,vpls,gbuffer,full,bake,batched,serving")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-counters", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the path renders each view and each session on one GPU with no "
              f"cross-GPU work; --gpus {args.gpus} is not supported", file=sys.stderr)
        return 2
    legs = set(args.only.split(",")) if args.only else {"headline", "plain", "fused_fp32", "host_inputs", "vpls",
                                                        "gbuffer", "full", "bake", "batched", "serving"}

    from instantradiositytest_amd import hip
    from instantradiositytest_amd.scenes import CONFIGS, ConfigScene

    if hip.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1
    total = args.warmup + args.steps
    cfg = CONFIGS[args.config]
    cfg = cfg.scaled(cfg.width, cfg.height, total + 1)
    settings = cfg.settings()
    W, H = cfg.width, cfg.height
    t0 = time.perf_counter()
    cs = ConfigScene(cfg)
    log(f"[bench] scene {args.config}: {len(cs.scene.meshes)} meshes, {len(cs.scene.objects)} objects, "
        f"atlas {cs.scene.atlas_dims}, built in {time.perf_counter() - t0:.1f}s")
    staged, host_frames, host_inputs = [], [], []
    t0 = time.perf_counter()
    for fi in range(total + 1):
        inp = cs.frame_inputs(fi)
        staged.append(tuple(hip.DeviceBuffer.from_array(a) for a in
                            (inp.depth, inp.normal, inp.color, inp.params, inp.motion, inp.vpls)))
        host_frames.append((inp.vp, inp.camera))
        if "host_inputs" in legs and fi < 16:
            host_inputs.append(inp)
        if fi % 16 == 0:
            log(f"[bench] staged frame {fi}/{total + 1} ({time.perf_counter() - t0:.1f}s)")

    results = {}
    head_r = None
    for leg, prec, sched in (("plain", "fp32", "passes"), ("fused_fp32", "fp32", "fused"),
                             ("headline", "mixed", "fused")):
        if leg not in legs:
            continue
        res, r = run_single(cs, settings, staged, host_frames, prec, sched, args.warmup, args.steps)
        results[leg] = res
        if leg == "headline":
            head_r = r
        else:
            r.close()

    line = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic"}
    head = results.get("headline")
    if head:
        line.update({"value": head["value"], "unit": "Mpix/s", "ms_per_step": head["ms_per_step"]})
        line["single_view_latency_ms"] = {"p50": head["p50_ms"], "p99": head["p99_ms"],
                                          "note": "one 1920x1200 view, start to finish of each timed frame"}
        line["config"] = {
            "workload": f"precision variant of the {W}x{H} workload ({args.config} scene in 16-bit storage): "
                        f"single view, fused kernels, 16-bit G-buffer, radiance and SDF samples with FP32 "
                        f"arithmetic, reservoirs, moments and march, reference default settings; inputs staged "
                        f"in HBM.  The same workload in FP32 storage is fused_fp32; the plain FP32 path is "
                        f"plain_fp32",
            "width": W, "height": H, "meshes": len(cs.scene.meshes), "objects": len(cs.scene.objects),
            "sdf_res": int(cfg.sdf_res[1]), "lights": len(cs.scene.lights), "vpls": cfg.vpl_count,
            "storage": "mixed", "schedule": "fused"}
        line["kernel_ms"] = head["kernel_ms"]
    if "plain" in results:
        line["plain_fp32"] = results["plain"]
        if head:
            line["speedup_vs_plain_fp32"] = round(head["value"] / results["plain"]["value"], 3)
    if "fused_fp32" in results:
        line["fused_fp32"] = results["fused_fp32"]
        if "plain" in results:
            line["fused_fp32"]["speedup_vs_plain_fp32"] = round(
                results["fused_fp32"]["value"] / results["plain"]["value"], 3)

    # roofline of the headline path's dominant kernel
    if head and head_r is not None and not args.no_counters:
        ref = reference_counts(head_r, cs, settings, staged, host_frames, total)
        kf = kernel_flops(ref)
        launches = head["kernel_launches_per_frame"]
        per_launch = {k: v for k, v in head["kernel_ms"].items() if launches.get(k, 1) == 1}
        dominant = max(per_launch, key=per_launch.get)
        flop = kf.get(dominant)
        if flop:
            achieved = flop / (head["kernel_ms"][dominant] * 1e-3) / 1e12
            line["roofline"] = {
                "bound": "fp32-valu", "compute": "FP32 vector ALU: no matrix work on this path; peak is the FP32 "
                                                 "vector figure; achieved counts the reference algorithm's "
                                                 "operations (bench/roofline.md)",
                "kernel": dominant, "achieved": round(achieved, 3), "peak": FP32_VALU_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 5),
                "traffic": measured_traffic(args.config, dominant),
                "flops_per_launch": flop, "launch_ms": head["kernel_ms"][dominant]}
        line["kernel_tflops"] = {k: round(kf[k] / (ms * 1e-3) / 1e12, 3) for k, ms in per_launch.items() if kf.get(k)}
        line["reference_events_per_frame"] = ref
    if head_r is not None:
        head_r.close()
    del staged
    if "host_inputs" in legs and host_inputs:
        line["pcie_inclusive"] = run_host_inputs(cs, settings, host_inputs, min(args.warmup, 4), min(args.steps, 20))
        del host_inputs
    threads = min(16, len(os.sched_getaffinity(0)))
    if "vpls" in legs:
        line["vpl_pass"] = run_vpls(cs, settings, host_frames, 50, threads, not args.no_cpu_baseline)
        if head:
            line["vpl_pass"]["share_of_headline_frame"] = round(line["vpl_pass"]["launch_ms"] / head["ms_per_step"], 4)
    if "gbuffer" in legs:
        line["gbuffer_pass"] = run_gbuffer(cs, settings, host_frames, 50, threads, not args.no_cpu_baseline)
        if head:
            line["gbuffer_pass"]["share_of_headline_frame"] = round(
                line["gbuffer_pass"]["launch_ms"] / head["ms_per_step"], 4)
    if "full" in legs:
        line["full_frame"] = run_full_frame(cs, settings, host_frames, args.warmup, args.steps)
    if "bake" in legs:
        line["sdf_bake"] = run_bake(cs, threads, not args.no_cpu_baseline)

    if "batched" in legs:
        line["batched"] = run_batched(12, 4, args.batch_steps)
    if "serving" in legs:
        shapes = None
        if "batched" in line:
            order = ("guide", "reservoirs_a", "spatial_setup", "resolve_combine", "svgf_temporal", "svgf_variance",
                     "svgf_atrous", "history")
            shapes = [("8x8", "16x8", "16x16").index(line["batched"]["tile_shapes"][k]) for k in order]
        from instantradiositytest_amd.serving import SERVING_MIX, session_arrivals
        rng = np.random.default_rng(0xC5)
        probs = [p for _, _, p in SERVING_MIX]
        arrivals, classes = session_arrivals(40.0, args.serve_seconds, probs, (120, 600), rng)
        streams = build_session_scenes(classes[:args.session_scenes], 600, 0xC5)
        line["serving"] = run_serving(streams, arrivals, args.serve_seconds, args.serve_warmup, args.sessions, shapes)
        if "plain" in results:
            line["serving"]["speedup_vs_plain_fp32"] = round(line["serving"]["mpix_s"] / results["plain"]["value"], 3)
        # the metric's latency half: p50 / p99 frame latency at 32 resident sessions
        line["p50_ms"] = line["serving"]["p50_ms"]
        line["p99_ms"] = line["serving"]["p99_ms"]
        line["latency_note"] = (f"p50_ms / p99_ms: request-to-completion frame latency of the serving workload at "
                                f"<= {args.sessions} resident sessions (serving); target p99 <= 16.7 ms")
        # the same workload with fewer resident sessions: where 60 Hz demand fits the GPU
        sweep = {}
        for n in args.serve_sweep:
            arr, _ = session_arrivals(40.0, args.sweep_seconds, probs, (120, 600), rng)
            d = run_serving(streams, arr, args.sweep_seconds, 3.0, n, shapes)
            sweep[str(n)] = {k: d[k] for k in ("mpix_s", "p50_ms", "p99_ms", "mean_batch", "frames")}
        if sweep:
            line["serving"]["by_sessions"] = sweep
        # the full session count with slower clients: the rate at which the latency target holds
        by_rate = {}
        for hz in args.serve_cadences:
            arr, _ = session_arrivals(40.0, args.sweep_seconds, probs, (120, 600), rng)
            d = run_serving(streams, arr, args.sweep_seconds, 3.0, args.sessions, shapes, cadence_hz=hz)
            by_rate[f"{hz:g}"] = {k: d[k] for k in ("mpix_s", "p50_ms", "p99_ms", "mean_batch", "frames")}
        if by_rate:
            line["serving"][f"by_client_hz_at_{args.sessions}_sessions"] = by_rate
        del streams
    if not args.no_cpu_baseline:
        threads = min(16, len(os.sched_getaffinity(0)))
        line["cpu_baseline"] = cpu_baseline(cs, settings, args.cpu_frames, threads)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
