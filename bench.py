#!/usr/bin/env python3
"""Benchmark of the per-pixel lighting + denoising path on one MI355X.

A step is one frame of the path for one view: the eight ReSTIR/resolve
passes and SVGF (temporal, variance, four a-trous) at 1920x1200 with the
reference's default settings, on the synthetic 24-mesh / 128^3-SDF /
16-light / 1000-VPL room (orbit + 0.5 m dolly).  Every frame's G-buffer and
VPL pool are generated up front and staged in HBM; the timed region starts
with them resident and covers ingest + lighting + SVGF + history per frame.

Prints one JSON line (see the repository's measurement notes for fields).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "shaded Mpix/s at 1920x1200 + p50/p99 frame ms @32 sessions vs plain FP32"

# Algorithmic FLOPs per event, from an operation count of the reference WGSL
# (one FLOP per add/mul/compare/min/max/div/sqrt/exp, two per FMA-able
# multiply-add).  Events are counted by the instrumented kernels.
FLOP_WEIGHTS = {
    "brdf": 160,     # direct()/indirect() + brdf() + luminance (common.wgsl:67-83, 45)
    "node": 23,      # one BVH AABB test (sdf.wgsl:53-59)
    "obj_box": 47,   # object_distance with the certified-box shortcut (sdf.wgsl:1-15, 35-41)
    "obj_grid": 133, # object_distance with the trilinear grid bound (sdf.wgsl:5-41, common.wgsl:31-38)
    "step": 23,      # march iteration bookkeeping (sdf.wgsl:70-78)
    "tap": 85,       # SVGF geometry_weight incl. position reconstruction (svgf.wgsl:32-44)
    "march": 35,     # visibility() set-up (sdf.wgsl:61-69)
}
FP32_VALU_PEAK_TFLOPS = 157.3  # MI355X published FP32 vector peak
HBM_PEAK_GBS = 8000.0


def log(msg: str):
    print(msg, file=sys.stderr, flush=True)


def flops_of(counts: dict) -> float:
    return float(sum(FLOP_WEIGHTS[k] * counts.get(k, 0) for k in FLOP_WEIGHTS))


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
            "sample": f"FP32 CPU restatement (oracle) of frames 0-{frames - 1} of the same workload, "
                      f"{cfg.width}x{cfg.height}, all 8 lighting passes + SVGF, {elapsed:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=8)
    ap.add_argument("--config", default="plain-1920x1200")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "mixed"])
    ap.add_argument("--schedule", default="passes", choices=["passes", "fused"])
    ap.add_argument("--cpu-frames", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-counters", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this path runs one view per GPU with no cross-GPU work; "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        return 2

    from instantradiositytest_amd import hip
    from instantradiositytest_amd.renderer import Renderer, pack_frame
    from instantradiositytest_amd.scenes import CONFIGS, ConfigScene
    from instantradiositytest_amd._lib import check

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
    # Generate and stage every frame's inputs in HBM.
    staged = []
    host_frames = []
    t0 = time.perf_counter()
    for fi in range(total + 1):
        inp = cs.frame_inputs(fi)
        staged.append(tuple(hip.DeviceBuffer.from_array(a) for a in
                            (inp.depth, inp.normal, inp.color, inp.params, inp.motion, inp.vpls)))
        host_frames.append((inp.vp, inp.camera))
        if fi % 16 == 0:
            log(f"[bench] staged frame {fi}/{total + 1} ({time.perf_counter() - t0:.1f}s)")
    r = Renderer(cs.scene, settings, precision=args.precision, schedule=args.schedule)
    L, ctx = r.L, r.ctx
    stream = L.irt_stream(ctx)

    def submit(fi: int):
        vp, cam = host_frames[fi]
        pvp = vp if r.previous_vp is None else r.previous_vp
        f = pack_frame(settings, vp, pvp, cam, r.frame_index, cs.scene.generation, r.previous_generation,
                       r.frame_index > 0 and r.previous_debug == r.debug_mode, r.debug_mode,
                       len(cs.scene.lights), len(cs.scene.objects))
        check(L.irt_set_frame(ctx, f.ctypes.data), ctx)
        d = staged[fi]
        check(L.irt_set_gbuffer_device(ctx, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, d[4].ptr), ctx)
        check(L.irt_set_vpls_device(ctx, d[5].ptr, cfg.vpl_count), ctx)
        check(L.irt_encode(ctx), ctx)
        r.previous_vp = vp
        r.previous_generation = cs.scene.generation
        r.previous_debug = r.debug_mode
        r.frame_index += 1

    for fi in range(args.warmup):
        submit(fi)
    r.sync()
    log(f"[bench] warmup {args.warmup} frames done")
    r.enable_timing(True)
    starts = [hip.Event() for _ in range(args.steps)]
    ends = [hip.Event() for _ in range(args.steps)]
    r.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record(stream)
        submit(args.warmup + k)
        ends[k].record(stream)
    r.sync()
    wall = time.perf_counter() - t0
    frame_ms = np.array([starts[k].elapsed_ms(ends[k]) for k in range(args.steps)])
    pass_entries = pass_time_entries(r)
    r.enable_timing(False)
    log(f"[bench] timed {args.steps} frames in {wall:.3f}s")
    value = W * H * args.steps / wall / 1e6
    # per-kernel mean durations over the timed region
    per_pass = {}
    for name, ms in pass_entries:
        per_pass.setdefault(name, []).append(ms)
    pass_ms = {k: float(np.mean(v)) for k, v in per_pass.items()}
    dominant = max(pass_ms, key=pass_ms.get) if pass_ms else None
    roofline = None
    counts_by_pass = {}
    if not args.no_counters and dominant is not None:
        counts_by_pass = instrumented_counts(r, cs, settings, staged, host_frames, total)
        flop = flops_of(counts_by_pass.get(dominant, {}))
        achieved = flop / (pass_ms[dominant] * 1e-3) / 1e12
        roofline = {"bound": "mfma", "compute": "fp32 vector ALU (no matrix work in this path)",
                    "kernel": dominant, "achieved": round(achieved, 3), "peak": FP32_VALU_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / FP32_VALU_PEAK_TFLOPS, 5), "traffic": None,
                    "flops_per_launch": flop, "launch_ms": round(pass_ms[dominant], 4)}
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.config}: single view {W}x{H}, {args.precision} storage, "
                               f"{args.schedule} schedule, reference default settings",
                   "width": W, "height": H, "meshes": len(cs.scene.meshes), "objects": len(cs.scene.objects),
                   "sdf_res": int(cfg.sdf_res[1]), "lights": len(cs.scene.lights), "vpls": cfg.vpl_count,
                   "precision": args.precision, "schedule": args.schedule},
        "p50_ms": round(float(np.percentile(frame_ms, 50)), 4),
        "p99_ms": round(float(np.percentile(frame_ms, 99)), 4),
        "pass_ms": {k: round(v, 4) for k, v in pass_ms.items()},
        "roofline": roofline,
    }
    if counts_by_pass:
        executed = counts_by_pass.pop("_executed", None)
        line["events_per_frame"] = {k: v for k, v in counts_by_pass.items()}
        if executed:
            line["executed_events_per_frame"] = executed
            if roofline is not None and dominant in executed:
                ex = flops_of(executed[dominant]) / (pass_ms[dominant] * 1e-3) / 1e12
                roofline["executed_tflops"] = round(ex, 3)
                roofline["executed_frac"] = round(ex / FP32_VALU_PEAK_TFLOPS, 5)
    if not args.no_cpu_baseline:
        threads = min(16, len(os.sched_getaffinity(0)))
        line["cpu_baseline"] = cpu_baseline(cs, settings, args.cpu_frames, threads)
    print(json.dumps(line), flush=True)
    return 0


def pass_time_entries(r) -> list:
    import ctypes
    cap = 8192
    names = (ctypes.c_char * 32 * cap)()
    ms = (ctypes.c_float * cap)()
    n = r.L.irt_pass_times(r.ctx, ctypes.cast(names, ctypes.c_void_p), ms, cap)
    out = []
    for i in range(max(n, 0)):
        out.append((bytes(names[i]).split(b"\0", 1)[0].decode(), float(ms[i])))
    return out


FUSED_MEMBERS = {
    "reservoirs_a": ["candidates_direct", "candidates_indirect", "temporal_direct", "temporal_indirect"],
    "reservoirs_b_resolve": ["spatial_direct", "spatial_indirect", "resolve_direct", "resolve_indirect"],
}
PASS_TIMER_NAME = {"resolve_direct": "visibility_direct", "resolve_indirect": "visibility_indirect"}


def instrumented_counts(r, cs, settings, staged, host_frames, last: int) -> dict:
    """Event counts per kernel for one more frame, run pass by pass with the
    instrumented kernels (outside the timed region)."""
    from instantradiositytest_amd.renderer import pack_frame
    from instantradiositytest_amd._lib import check
    L, ctx = r.L, r.ctx
    vp, cam = host_frames[last]
    f = pack_frame(settings, vp, r.previous_vp, cam, r.frame_index, cs.scene.generation, r.previous_generation,
                   True, r.debug_mode, len(cs.scene.lights), len(cs.scene.objects))
    check(L.irt_set_frame(ctx, f.ctypes.data), ctx)
    d = staged[last]
    check(L.irt_set_gbuffer_device(ctx, d[0].ptr, d[1].ptr, d[2].ptr, d[3].ptr, d[4].ptr), ctx)
    check(L.irt_set_vpls_device(ctx, d[5].ptr, cs.cfg.vpl_count), ctx)
    out = {}
    names = ["candidates_direct", "candidates_indirect", "temporal_direct", "temporal_indirect", "spatial_direct",
             "spatial_indirect", "resolve_direct", "resolve_indirect"]
    for i, name in enumerate(names):
        r.enable_counters(True)
        check(L.irt_pass(ctx, i, 0), ctx)
        out[PASS_TIMER_NAME.get(name, name)] = r.counters()
    # leave the passes' results in place for the executed-work count below
    r.enable_counters(True)
    check(L.irt_svgf(ctx, settings.svgf_iterations, 0), ctx)
    out["svgf"] = r.counters()
    r.enable_counters(False)
    for fused, members in FUSED_MEMBERS.items():
        agg = {}
        for m in members:
            for k, v in out[PASS_TIMER_NAME.get(m, m)].items():
                agg[k] = agg.get(k, 0) + v
        out[fused] = agg
    if r.schedule_name == "fused":
        # Work the fused kernels execute (candidate-grid queries count their
        # candidate tests as node tests), on the same frame's inputs.
        executed = {}
        r.enable_counters(True)
        r.schedule("fused")
        check(L.irt_lighting(ctx), ctx)
        executed["reservoirs"] = r.counters()
        r.enable_counters(True)
        check(L.irt_svgf(ctx, settings.svgf_iterations, 0), ctx)
        executed["svgf"] = r.counters()
        r.enable_counters(False)
        # split the lighting counts between the two fused kernels: march events belong to reservoirs_b
        lit = executed.pop("reservoirs")
        a_brdf = out["reservoirs_a"]["brdf"]
        executed["reservoirs_a"] = {k: (a_brdf if k == "brdf" else 0) for k in lit}
        executed["reservoirs_b_resolve"] = {k: (v - a_brdf if k == "brdf" else v) for k, v in lit.items()}
        out["_executed"] = executed
    return out


if __name__ == "__main__":
    sys.exit(main())
