#!/usr/bin/env python3
"""Headline benchmark: RK4 geodesic steps/s of the batched persistent Kerr
march at 1920x1080, spin 0.9, 4096-step budget, observer at r = 50,
inclination 60 degrees, 60 degree field of view, through the 1M-particle /
128^3 disk volume (settled 600 substeps), against the plain
one-launch-per-frame march on the same GPU in the same run.

One bench step is one frame: every ray of the 1080p image marched once.
Inputs (particles, volume, tables) are resident in HBM before timing; the
timed region is bracketed by device synchronization and HIP events on the
stream the kernels run on. Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 geodesic steps/s at 1080p a=0.9 vs plain march; p50/p99 render latency"
# Reference-form FP32 flops of one Kerr RK4 step of a live ray (4 Hamiltonian
# evaluations, the RK4 update, step length, position, path length and the
# Carter / null-residual drift gauges); the count is derived in DESIGN.md.
FLOPS_PER_KERR_STEP = 509
FP32_PEAK_TFLOPS = 157.3          # MI355X published FP32 vector peak (packed)
FP32_PEAK_NONPACKED_TFLOPS = 78.6


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed frames")
    ap.add_argument("--warmup", type=int, default=3, help="untimed frames")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=0, help="march every k-th row on the CPU (0 = auto)")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--config", default="", help="batch config as block,waves,period,threshold")
    return ap.parse_args()


def build(ctx, api):
    """1M particles from the annulus generator (sigma_z 0.3), settled 600
    substeps, exact CIC into 128^3 spanning |x|,|y| <= 30, |z| <= 3; the
    camera and 1080p targets of the headline frame."""
    p = api.Particles(ctx, 1_000_000, seed=1)
    p.seed(api.GENERATOR_ANNULUS, sigma_z=0.3)
    p.step(300)
    vol = api.Volume(ctx, 128, 30.0, 3.0)
    vol.clear()
    vol.splat(p, exact=True)
    vol.resolve(clear=True)
    ctx.sync()
    inclination, distance, fov = 60.0, 50.0, 60.0
    cam = api.orbit_camera(azimuth=0.0, elevation=math.radians(90.0 - inclination), distance=distance,
                           fov_y_deg=fov)
    u = api.uniforms(cam, 1920, 1080, api.settings("still", step_cap=4096, escape_radius=max(60.0, distance + 10)),
                     spin=0.9, compensated=True, star_seed=1, volume_half_extent=30.0)
    targets = api.Targets(ctx, 1920, 1080)
    return p, vol, u, targets


def time_frames(ctx, api, run, warmup, steps):
    for _ in range(warmup):
        run()
    ctx.sync()
    start, end = ctx.event(), ctx.event()
    marks = [ctx.event() for _ in range(steps + 1)]
    start.record()
    marks[0].record()
    for i in range(steps):
        run()
        marks[i + 1].record()
    end.record()
    end.sync()
    ctx.sync()
    total_ms = start.elapsed_ms(end)
    frames = [marks[i].elapsed_ms(marks[i + 1]) for i in range(steps)]
    return total_ms, frames


def cpu_baseline(u, vol, every: int):
    """The FP64 oracle's Kerr march (march.metal:123-184 restated) on every
    `every`-th row of the same rays and volume, on the host's cores."""
    import oracle as O
    from tests import scenes
    threads = max(1, min(16, os.cpu_count() or 1))
    field = scenes.field_from_volume(vol)
    p = scenes.oracle_params(u)
    w, h = int(u.resolution[0]), int(u.resolution[1])
    ys = np.arange(0, h, every)
    xs = np.arange(w)
    xy = np.stack(np.meshgrid(xs, ys), -1).reshape(-1, 2)
    t0 = time.perf_counter()
    _, total = scenes.oracle_pixels(p, xy, field, threads)
    dt = time.perf_counter() - t0
    return {"value": total / dt, "unit": "geodesic steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle Kerr march with gas on every {every}th row of the 1920x1080 frame "
                      f"({len(xy)} rays, {total} steps, {dt:.1f} s)",
            "cpu": _cpu_model()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for {args.gpus}).", file=sys.stderr)
        sys.exit(2)
    from gargantua_amd import api
    ctx = api.Context()
    info = ctx.info()
    p, vol, u, targets = build(ctx, api)

    cfg = ctx.batch_config()
    if args.config:
        b, w, per, thr = (int(x) for x in args.config.split(","))
        cfg = ctx.batch_config(block=b, waves_per_simd=w, compaction_period=per, refill_threshold=thr)
    request = [{"uniforms": u, "targets": targets, "volume": vol}]

    # Steps per frame are a property of the scene: count them once.
    targets.reset_counters()
    totals = ctx.march_batch(request, cfg)
    steps_per_frame = int(totals.steps)
    rays = int(totals.rays)

    batched_ms, batched_frames = time_frames(ctx, api, lambda: _batch_nosync(ctx, api, request, cfg), args.warmup,
                                             args.steps)
    ms_per_step = batched_ms / args.steps
    value = steps_per_frame / (ms_per_step * 1e-3)

    plain = None
    if not args.skip_plain:
        plain_targets = api.Targets(ctx, 1920, 1080)
        plain_targets.reset_counters()
        ctx.march_image(u, plain_targets, vol, None, api.INTEGRATOR_REFERENCE)
        plain_steps = int(plain_targets.counters().steps)
        plain_ms, plain_frames = time_frames(
            ctx, api, lambda: ctx.march_image(u, plain_targets, vol, None, api.INTEGRATOR_REFERENCE),
            max(1, args.warmup // 2), max(3, args.steps // 2))
        plain_ms_per = plain_ms / max(3, args.steps // 2)
        plain = {"ms_per_step": plain_ms_per, "value": plain_steps / (plain_ms_per * 1e-3),
                 "steps_per_frame": plain_steps, "integrator": "reference", "launch": "one per frame, 8x8 waves",
                 "p50_ms": float(np.percentile(plain_frames, 50)), "p99_ms": float(np.percentile(plain_frames, 99))}
        # Decomposition: this library's integrator with the plain launch, so
        # the gain splits into arithmetic (plain_fast / plain) and scheduling
        # (batched / plain_fast).
        plain_targets.reset_counters()
        ctx.march_image(u, plain_targets, vol, None, api.INTEGRATOR_FAST)
        fast_steps = int(plain_targets.counters().steps)
        fast_ms, _ = time_frames(ctx, api, lambda: ctx.march_image(u, plain_targets, vol, None, api.INTEGRATOR_FAST),
                                 max(1, args.warmup // 2), max(3, args.steps // 2))
        fast_ms_per = fast_ms / max(3, args.steps // 2)
        plain["fast_integrator"] = {"ms_per_step": fast_ms_per, "value": fast_steps / (fast_ms_per * 1e-3),
                                    "steps_per_frame": fast_steps}
        plain_targets.close()

    achieved = steps_per_frame * FLOPS_PER_KERR_STEP / (ms_per_step * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "geodesic steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "kerr_march_1920x1080_a0.9_4096", "resolution": [1920, 1080], "spin": 0.9,
                   "step_cap": 4096, "observer_radius": 50.0, "inclination_deg": 60.0, "fov_deg": 60.0,
                   "particles": 1_000_000, "volume": 128, "volume_half_extent_z": 3.0, "settle_substeps": 600,
                   "rays": rays, "steps_per_frame": steps_per_frame, "compensated": True,
                   "batch": {"block": cfg.block, "waves_per_simd": cfg.waves_per_simd,
                             "compaction_period": cfg.compaction_period, "refill_threshold": cfg.refill_threshold}},
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        "latency_ms": {"p50": float(np.percentile(batched_frames, 50)),
                       "p99": float(np.percentile(batched_frames, 99))},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None,
                     "peak_nonpacked": FP32_PEAK_NONPACKED_TFLOPS,
                     "flops_per_step": FLOPS_PER_KERR_STEP, "kernel": "march_batch_kernel<true,256>"},
        "device": {"name": info.name.decode(), "arch": info.arch.decode(), "compute_units": info.compute_units},
    }
    if not args.no_cpu_baseline:
        every = args.cpu_rows or 1
        line["cpu_baseline"] = cpu_baseline(u, vol, every)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    for o in (targets, vol, p):
        o.close()
    ctx.close()


def _batch_nosync(ctx, api, request, cfg):
    """One frame of the batched march, enqueued without reading counters back."""
    import ctypes as C
    arr = (api.MarchRequest * len(request))()
    for i, r in enumerate(request):
        arr[i].uniforms = r["uniforms"]
        arr[i].targets = r["targets"].h
        arr[i].volume = r["volume"].h if r.get("volume") else None
    st = api.load().garg_march_batch(ctx.h, None, arr, len(request), C.byref(cfg), None)
    if st != 0:
        raise api.GargError(st, api.load().garg_last_error().decode())


if __name__ == "__main__":
    main()
