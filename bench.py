#!/usr/bin/env python3
"""Headline benchmark: RK4 geodesic steps/s of the batched persistent Kerr
march at 1920x1080, spin 0.9, 4096-step budget, observer at r = 50,
inclination 60 degrees, 60 degree field of view, through the 1M-particle /
128^3 disk volume (settled 600 substeps), against the plain
one-launch-per-frame march on the same GPU in the same run. The headline
form is the one that keeps every escaping ray's drift under the 1e-5 still
budget: the polar form (Boyer-Lindquist fast steps away from the spin axis,
Kerr-Schild steps near it); the Boyer-Lindquist fast form and the all
Kerr-Schild form are side legs.

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
# Reference-form FP32 flops of one Kerr RK4 step of a live ray, counted from
# the formulas of Kerr.h:11-186 with +, -, *, /, sqrt, sin, cos = 1 and
# FMA = 2: four Hamiltonian evaluations of 92 (368), the RK4 combination of
# five components (68), the step length with its pole factor (10), the
# Boyer-Lindquist to Cartesian position (13), the path length (10) and the
# Carter and null-residual drift gauges (40).
FLOPS_PER_KERR_STEP = 509
FP32_PEAK_TFLOPS = 157.3          # MI355X published FP32 vector peak
# HBM bytes per launch of the batched polar-form Kerr march at this frame,
# from two rocprofv3 counter runs (FETCH_SIZE x 2.0, WRITE_SIZE x 1.0: the
# factors a 256 MiB device copy reads); profiles/march_kerr_1080p_a0.9_counters.md.
TRAFFIC_BYTES_PER_LAUNCH = 415.5e6


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed frames")
    ap.add_argument("--warmup", type=int, default=3, help="untimed frames")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=0, help="march every k-th row on the CPU (0 = auto)")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--config", default="", help="batch config as block,waves,period,threshold")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 64-request mixed batch")
    ap.add_argument("--no-passes", action="store_true", help="skip the sim/splat/resolve and FP64 legs")
    ap.add_argument("--no-schwarzschild", action="store_true", help="skip the spin-0 legs (small and default presets)")
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


def progress(text):
    """One line per finished leg on stderr (stdout carries only the JSON line)."""
    print(f"bench: {text}", file=sys.stderr, flush=True)


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


def settled_ms(ctx, api, run, reps, warmup=1):
    """Time per call of a side leg, timed twice with the second pass kept:
    a leg of frames of a millisecond or two would otherwise be charged for
    the clocks settling after the leg before it, whichever leg runs first."""
    time_frames(ctx, api, run, warmup, reps)
    ms, _ = time_frames(ctx, api, run, 0, reps)
    return ms / reps


def host_threads():
    """Cores this process may use: those it is allowed to run on, within the
    thread count the host grants it (OMP_NUM_THREADS: a GPU box gives one
    GPU's process a share of 16 cores)."""
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return max(1, min(allowed, share) if share > 0 else allowed), allowed


def cpu_baseline(u, vol, every: int):
    """The FP64 oracle's Kerr march (march.metal:123-184 restated) on every
    `every`-th row of the same rays and volume, on the host's cores."""
    import oracle as O
    from tests import scenes
    threads, allowed = host_threads()
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
            "sample": f"FP64 oracle Kerr march with gas on {'every row' if every == 1 else f'every {every}th row'} of the 1920x1080 frame "
                      f"({len(xy)} rays, {total} steps, {dt:.1f} s)",
            "cpu": _cpu_model(), "nproc": os.cpu_count(), "cores_allowed": allowed,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


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

    from gargantua_amd import autotune
    base = autotune.best(ctx, "headline")
    if args.config:
        b, w, per, thr = (int(x) for x in args.config.split(","))
        base = ctx.batch_config(block=b, waves_per_simd=w, compaction_period=per, refill_threshold=thr)

    def form_config(kerr_schild):
        return ctx.batch_config(block=base.block, waves_per_simd=base.waves_per_simd,
                                compaction_period=base.compaction_period, refill_threshold=base.refill_threshold,
                                kerr_schild=kerr_schild)

    cfg = form_config(2)  # the polar form
    request = [{"uniforms": u, "targets": targets, "volume": vol}]

    def frame_drift(t, counters):
        """The frame's drift (maximum over each ray's steps of the Carter,
        null-residual and axial-momentum drifts) against the still budget:
        over the rays that escape or run out of budget, and the counter's
        figure that includes captured rays (whose gauges diverge at the
        horizon)."""
        debug = t.read()[1]
        kept = debug[..., 2] != api.RAY_CAPTURED
        return {"budget": float(u.drift_budget), "max_escaped_or_exhausted": float(debug[..., 0][kept].max()),
                "rays_over_budget_escaped_or_exhausted": int((debug[..., 0][kept] > u.drift_budget).sum()),
                "rays_over_budget_all": int(counters.over_drift_budget), "captured": int(counters.captured),
                "exhausted": int(counters.exhausted)}

    # Steps per frame are a property of the scene and form: count them once.
    targets.reset_counters()
    totals = ctx.march_batch(request, cfg)
    steps_per_frame = int(totals.steps)
    rays = int(totals.rays)
    drift = frame_drift(targets, totals)

    batched_ms, batched_frames = time_frames(ctx, api, lambda: _batch_nosync(ctx, api, request, cfg), args.warmup,
                                             args.steps)
    ms_per_step = batched_ms / args.steps
    value = steps_per_frame / (ms_per_step * 1e-3)
    progress(f"headline polar batched {ms_per_step:.3f} ms")

    # The other batched forms on the same frame: the Boyer-Lindquist fast
    # form (the reference's coordinates, over the drift budget on near-polar
    # rays as the FP64 oracle is) and the all Kerr-Schild form.
    side = {}
    side_steps = max(3, args.steps // 2)
    for name, ks in (("boyer_lindquist_fast", 0), ("kerr_schild", 1)):
        c = form_config(ks)
        targets.reset_counters()
        t = ctx.march_batch(request, c)
        d = frame_drift(targets, t)
        ms, frames = time_frames(ctx, api, lambda: _batch_nosync(ctx, api, request, c), max(1, args.warmup // 2),
                                 side_steps)
        side[name] = {"ms_per_step": ms / side_steps, "value": int(t.steps) / (ms / side_steps * 1e-3),
                      "steps_per_frame": int(t.steps), "drift": d,
                      "p50_ms": float(np.percentile(frames, 50)), "p99_ms": float(np.percentile(frames, 99))}

    progress("other batched forms")
    plain = None
    if not args.skip_plain:
        plain_targets = api.Targets(ctx, 1920, 1080)
        n_plain = max(3, args.steps // 2)

        def plain_leg(integrator):
            plain_targets.reset_counters()
            ctx.march_image(u, plain_targets, vol, None, integrator)
            c = plain_targets.counters()
            d = frame_drift(plain_targets, c)
            ms, frames = time_frames(ctx, api, lambda: ctx.march_image(u, plain_targets, vol, None, integrator),
                                     max(1, args.warmup // 2), n_plain)
            per = ms / n_plain
            return {"ms_per_step": per, "value": int(c.steps) / (per * 1e-3), "steps_per_frame": int(c.steps),
                    "drift": d, "p50_ms": float(np.percentile(frames, 50)), "p99_ms": float(np.percentile(frames, 99))}

        # The baseline the metric names: the reference's marchImage formulas,
        # one thread per pixel, 8x8 workgroups, one launch per frame.
        plain = plain_leg(api.INTEGRATOR_REFERENCE)
        plain.update({"integrator": "reference", "launch": "one per frame, 8x8 waves"})
        # This library's forms with the same plain launch, so the gain splits
        # into arithmetic (plain form / plain reference) and scheduling
        # (batched / plain of the same form).
        plain["polar_integrator"] = plain_leg(api.INTEGRATOR_POLAR)
        plain["fast_integrator"] = plain_leg(api.INTEGRATOR_FAST)
        plain["kerr_schild_integrator"] = plain_leg(api.INTEGRATOR_KERR_SCHILD)
        plain_targets.close()
        progress("plain launches")

    achieved = steps_per_frame * FLOPS_PER_KERR_STEP / (ms_per_step * 1e-3) / 1e12
    fma_forms = {name: max(ctx.valu_rate(form) for _ in range(3)) / 1e12
                 for name, form in (("literal_operand", api.VALU_FMA_LITERAL), ("three_vgprs", api.VALU_FMA_VGPR),
                                    ("sgpr_operands", api.VALU_FMA_SGPR))}
    fma_ceiling = fma_forms["literal_operand"]
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
                   "form": "polar (Boyer-Lindquist fast steps until sin(theta) < 0.25 or a plunge inside r+ + 0.5, Kerr-Schild after)",
                   "batch": {"block": cfg.block, "waves_per_simd": cfg.waves_per_simd,
                             "compaction_period": cfg.compaction_period, "refill_threshold": cfg.refill_threshold,
                             "kerr_schild": cfg.kerr_schild}},
        "drift": drift,
        "drift_clause_met": drift["rays_over_budget_escaped_or_exhausted"] == 0,
        "batched_other_forms": side,
        "plain": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain else None,
        # Equal integrator: the batched persistent launch against the plain
        # launch of the same (polar) form; the rest of speedup_vs_plain is
        # the integrator's arithmetic.
        "speedup_scheduling": (plain["polar_integrator"]["ms_per_step"] / ms_per_step) if plain else None,
        "latency_ms": {"p50": float(np.percentile(batched_frames, 50)),
                       "p99": float(np.percentile(batched_frames, 99))},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": TRAFFIC_BYTES_PER_LAUNCH,
                     "traffic_unit": "bytes per launch",
                     # FP32 ceilings measured on this GPU in the same run:
                     # independent non-packed fused multiply-adds in three
                     # operand forms (garg_valu_rate), TFLOP/s; the fastest
                     # form is the achievable FP32 ceiling.
                     "fma_measured_tflops": fma_forms,
                     "frac_of_measured": achieved / fma_ceiling if fma_ceiling else None,
                     "flops_per_step": FLOPS_PER_KERR_STEP,
                     "kernel": ("march_tiles_kernel" if cfg.refill_threshold == 0 else "march_batch_kernel")
                     + f"<3, {cfg.block}, true>"},
        "device": {"name": info.name.decode(), "arch": info.arch.decode(), "compute_units": info.compute_units},
    }
    if not args.no_mixed:
        line["mixed"] = mixed_leg(ctx, api, vol, args)
        progress("mixed batch")
    if not args.no_passes:
        line["passes"] = pass_leg(ctx, api, p, vol)
        line["fp64"] = fp64_leg(ctx)
        progress("disk passes and FP64 kernels")
    if not args.no_mixed:
        line["kerr_640x360_x8"] = kerr_views_leg(ctx, api, vol, cfg)
        line["kerr_presets"] = kerr_presets_leg(ctx, api, vol, cfg)
        progress("Kerr views and presets")
    if not args.no_schwarzschild:
        line["schwarzschild"] = schwarzschild_legs(ctx, api, vol, cfg, not args.no_cpu_baseline)
        progress("Schwarzschild legs")
    if not args.no_cpu_baseline:
        every = args.cpu_rows or 1
        line["cpu_baseline"] = cpu_baseline(u, vol, every)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    for o in (targets, vol, p):
        o.close()
    ctx.close()


def schwarzschild_legs(ctx, api, default_volume, cfg, cpu: bool):
    """The spin-0 presets with the far-field jump: small (250k particles
    from the annulus generator, 1000 substeps, exact CIC into 96^3 with
    |z| <= 3, 640x360, 256-step interactive budget, observer r = 50,
    inclination 80 deg, 60 deg field of view) and default (the 1M / 128^3
    volume above, 960x540 at 256 steps, then a 3840x2160 still at 4096 steps
    with compensated accumulation). Each: the batched launch against the
    plain one in both integrator forms; the small frame also on the FP64
    oracle over every row."""
    small_p = api.Particles(ctx, 250_000, seed=1)
    small_p.seed(api.GENERATOR_ANNULUS, sigma_z=0.3)
    small_p.step(500)
    small_v = api.Volume(ctx, 96, 30.0, 3.0)
    small_v.clear()
    small_v.splat(small_p, exact=True)
    small_v.resolve(clear=True)
    ctx.sync()
    tables = api.Tables(ctx)
    tables.update(50.0)
    cam = api.orbit_camera(azimuth=0.0, elevation=math.radians(10.0), distance=50.0, fov_y_deg=60.0)
    out = {}
    for name, w, h, kind, cap, vol in (("small_640x360_256", 640, 360, "interactive", 256, small_v),
                                       ("default_960x540_256", 960, 540, "interactive", 256, default_volume),
                                       ("default_3840x2160_4096", 3840, 2160, "still", 4096, default_volume)):
        u = api.uniforms(cam, w, h, api.settings(kind, step_cap=cap), compensated=kind == "still", star_seed=1,
                         drift_budget=1e-5 if kind == "still" else 1e-4, volume_half_extent=30.0)
        tables.apply(u)
        t = api.Targets(ctx, w, h)
        req = [{"uniforms": u, "targets": t, "volume": vol, "tables": tables}]
        t.reset_counters()
        totals = ctx.march_batch(req, cfg)
        steps = int(totals.steps)
        reps = 3 if cap > 256 else 10
        batched_ms = settled_ms(ctx, api, lambda: _batch_nosync(ctx, api, req, cfg), reps)
        leg = {"rays": int(totals.rays), "steps": steps, "over_drift_budget": int(totals.over_drift_budget),
               "exhausted": int(totals.exhausted), "batched_ms": batched_ms,
               "batched_steps_per_s": steps / (batched_ms * 1e-3)}
        for form, integrator in (("plain_reference", api.INTEGRATOR_REFERENCE), ("plain_fast", api.INTEGRATOR_FAST)):
            leg[form + "_ms"] = settled_ms(ctx, api, lambda: ctx.march_image(u, t, vol, tables, integrator), reps)
        leg["speedup_vs_plain_reference"] = leg["plain_reference_ms"] / leg["batched_ms"]
        if cpu and name.startswith("small"):
            import oracle as O
            from tests import scenes
            field = scenes.field_from_volume(vol)
            params = scenes.oracle_params(u, scenes.oracle_tables(50.0))
            xs, ys = np.meshgrid(np.arange(w), np.arange(h))
            xy = np.stack([xs.ravel(), ys.ravel()], 1)
            threads = max(1, min(16, os.cpu_count() or 1))
            t0 = time.perf_counter()
            _, total = scenes.oracle_pixels(params, xy, field, threads)
            dt = time.perf_counter() - t0
            leg["cpu_oracle"] = {"steps_per_s": total / dt, "seconds": dt, "cores": threads, "steps": total}
        out[name] = leg
        t.close()
    out["small_640x360_256_x8"] = small_batch_leg(ctx, api, small_v, tables, cfg)
    out["bake_walk_960x540"] = bake_leg(ctx, api, default_volume, tables)
    tables.close()
    small_v.close()
    small_p.close()
    return out


def kerr_views_leg(ctx, api, vol, cfg, views=8):
    """Eight 640x360 views of the metric's scene (a = 0.9, 4096-step still
    budget, r_obs = 50, inclination 60 deg, the 1M / 128^3 volume) from
    eight azimuths: one batched launch against eight plain launches in each
    integrator form. A 640x360 frame (3,600 tiles) holds less than the GPU's
    5,120 wave slots, so one frame alone leaves most of the GPU idle while
    its slowest tiles finish."""
    reqs = []
    for k in range(views):
        cam = api.orbit_camera(azimuth=2.0 * math.pi * k / views, elevation=math.radians(30.0), distance=50.0,
                               fov_y_deg=60.0)
        u = api.uniforms(cam, 640, 360, api.settings("still", step_cap=4096, escape_radius=60.0), spin=0.9,
                         compensated=True, star_seed=1, volume_half_extent=30.0)
        reqs.append({"uniforms": u, "targets": api.Targets(ctx, 640, 360), "volume": vol})
    for r in reqs:
        r["targets"].reset_counters()
    totals = ctx.march_batch(reqs, cfg)
    steps = int(totals.steps)
    batched_ms = settled_ms(ctx, api, lambda: _batch_nosync(ctx, api, reqs, cfg), 10)
    out = {"requests": views, "rays": int(totals.rays), "steps": steps, "batched_ms": batched_ms,
           "batched_steps_per_s": steps / (batched_ms * 1e-3), "batched_form": _form_name(cfg)}
    for form, integrator in (("plain_reference", api.INTEGRATOR_REFERENCE), ("plain_fast", api.INTEGRATOR_FAST),
                             ("plain_polar", api.INTEGRATOR_POLAR)):
        def run_all():
            for r in reqs:
                ctx.march_image(r["uniforms"], r["targets"], vol, None, integrator)
        out[form + "_ms"] = settled_ms(ctx, api, run_all, 5)
    out["speedup_vs_plain_reference"] = out["plain_reference_ms"] / out["batched_ms"]
    out["speedup_vs_plain_fast"] = out["plain_fast_ms"] / out["batched_ms"]
    out["speedup_vs_plain_polar"] = out["plain_polar_ms"] / out["batched_ms"]
    for r in reqs:
        r["targets"].close()
    return out


def kerr_presets_leg(ctx, api, vol, cfg):
    """The reference's own Kerr operating points on the metric's scene
    (a = 0.9, r_obs = 50, inclination 60 deg, 60 deg field of view, the
    1M / 128^3 volume): the 960x540 interactive frame at the 256-step budget
    (the reference reports "near 8 fps" for the whole interactive frame on
    an M1 Pro, in its README.md:136-137) and the 1280x720 still at 4096 steps with
    compensated accumulation ("under a second", docs/lab/07-kerr.md:49-50):
    the batched march of one frame against the plain launch in each form."""
    cam = api.orbit_camera(azimuth=0.0, elevation=math.radians(30.0), distance=50.0, fov_y_deg=60.0)
    out = {}
    for name, w, h, kind, cap in (("interactive_960x540_256", 960, 540, "interactive", 256),
                                  ("still_1280x720_4096", 1280, 720, "still", 4096)):
        u = api.uniforms(cam, w, h, api.settings(kind, step_cap=cap, escape_radius=60.0), spin=0.9,
                         compensated=kind == "still", star_seed=1, volume_half_extent=30.0,
                         drift_budget=1e-5 if kind == "still" else 1e-4)
        t = api.Targets(ctx, w, h)
        req = [{"uniforms": u, "targets": t, "volume": vol}]
        t.reset_counters()
        totals = ctx.march_batch(req, cfg)
        leg = {"rays": int(totals.rays), "steps": int(totals.steps),
               "batched_ms": settled_ms(ctx, api, lambda: _batch_nosync(ctx, api, req, cfg), 10),
               "batched_form": _form_name(cfg)}
        for form, integrator in (("plain_reference", api.INTEGRATOR_REFERENCE), ("plain_fast", api.INTEGRATOR_FAST),
                                 ("plain_polar", api.INTEGRATOR_POLAR)):
            leg[form + "_ms"] = settled_ms(ctx, api, lambda: ctx.march_image(u, t, vol, None, integrator), 5)
        leg["speedup_vs_plain_reference"] = leg["plain_reference_ms"] / leg["batched_ms"]
        leg["speedup_vs_plain_polar"] = leg["plain_polar_ms"] / leg["batched_ms"]
        t.close()
        out[name] = leg
    return out


def bake_leg(ctx, api, vol, tables):
    """The bake strategy at 960x540 on the default volume: one still-policy
    Schwarzschild march storing up to 96 samples per ray at half-voxel
    spacing (bakeRays), then the walk that composites the stored samples
    against the volume (walkRays), each timed over five calls."""
    w, h = 960, 540
    cam = api.orbit_camera(azimuth=0.0, elevation=math.radians(10.0), distance=50.0, fov_y_deg=60.0)
    u = api.uniforms(cam, w, h, api.settings("still"), star_seed=1, volume_half_extent=30.0,
                     bake_spacing=60.0 / 128 * 0.5)
    tables.apply(u)
    bake = api.Bake(ctx, w, h, 96)
    t = api.Targets(ctx, w, h)
    bake_ms = settled_ms(ctx, api, lambda: bake.rows(u, tables), 5)
    walk_ms = settled_ms(ctx, api, lambda: bake.walk(u, vol, t), 5)
    headers, _ = bake.read()
    # The walk's algorithmic HBM bytes per ray: the 16-byte header, 8 bytes
    # per stored sample it reads (at most the ray's count; it stops at the
    # transmittance floor) and 12 bytes written (colour and depth). The
    # volume's texels are L2-resident reads, not counted.
    walk_bytes = int(w * h * (16 + 12) + 8 * int(headers[:, 0].sum()))
    out = {"bake_ms": bake_ms, "walk_ms": walk_ms, "rays": w * h, "overflow_rays": bake.overflow(),
           "mean_samples_per_ray": float(headers[:, 0].mean()),
           "bake_bytes": w * h * (8 * 96 + 16),
           "walk_bytes_upper": walk_bytes, "walk_GB_per_s": walk_bytes / (walk_ms * 1e-3) / 1e9,
           "walk_hbm_frac": walk_bytes / (walk_ms * 1e-3) / 8e12,
           # The bake writes its samples and headers: 8 bytes per stored sample, 16 per ray.
           "bake_written_bytes": int(16 * w * h + 8 * int(headers[:, 0].sum()))}
    bake.close()
    t.close()
    return out


def small_batch_leg(ctx, api, vol, tables, cfg, views=8):
    """Eight small frames (640x360, 256 steps, the small preset's volume)
    from eight azimuths around the hole: one batched launch against eight
    plain launches. A single small frame is latency-bound (its slowest tiles
    set the time); marching several together fills the machine."""
    reqs = []
    for k in range(views):
        cam = api.orbit_camera(azimuth=2.0 * math.pi * k / views, elevation=math.radians(10.0), distance=50.0,
                               fov_y_deg=60.0)
        u = api.uniforms(cam, 640, 360, api.settings("interactive", step_cap=256), star_seed=1, drift_budget=1e-4,
                         volume_half_extent=30.0)
        tables.apply(u)
        reqs.append({"uniforms": u, "targets": api.Targets(ctx, 640, 360), "volume": vol, "tables": tables})
    for r in reqs:
        r["targets"].reset_counters()
    totals = ctx.march_batch(reqs, cfg)
    steps = int(totals.steps)
    batched_ms = settled_ms(ctx, api, lambda: _batch_nosync(ctx, api, reqs, cfg), 10)
    out = {"requests": views, "rays": int(totals.rays), "steps": steps, "batched_ms": batched_ms,
           "batched_steps_per_s": steps / (batched_ms * 1e-3)}
    for form, integrator in (("plain_reference", api.INTEGRATOR_REFERENCE), ("plain_fast", api.INTEGRATOR_FAST)):
        def run_all():
            for r in reqs:
                ctx.march_image(r["uniforms"], r["targets"], vol, tables, integrator)
        out[form + "_ms"] = settled_ms(ctx, api, run_all, 10)
    out["speedup_vs_plain_reference"] = out["plain_reference_ms"] / out["batched_ms"]
    out["speedup_vs_plain_fast"] = out["plain_fast_ms"] / out["batched_ms"]
    for r in reqs:
        r["targets"].close()
    return out


def mixed_leg(ctx, api, vol, args):
    """64 requests (resolution, spin, inclination, observer radius and
    budget drawn from seed 4) in one batched launch against the same
    requests marched one plain launch each, both integrator forms."""
    from gargantua_amd import autotune, workloads
    reqs = workloads.mixed(ctx, vol)
    tuned = autotune.best(ctx, "mixed")
    cfg = ctx.batch_config(block=tuned.block, waves_per_simd=tuned.waves_per_simd,
                           compaction_period=tuned.compaction_period, refill_threshold=tuned.refill_threshold,
                           kerr_schild=0)
    for r in reqs:
        r["targets"].reset_counters()
    totals = ctx.march_batch(reqs, cfg)
    steps = int(totals.steps)
    batched_ms, _ = time_frames(ctx, api, lambda: _batch_nosync(ctx, api, reqs, cfg), 1, 3)
    out = {"requests": len(reqs), "rays": int(totals.rays), "steps": steps, "captured": int(totals.captured),
           "exhausted": int(totals.exhausted),
           "batch": {"block": cfg.block, "waves_per_simd": cfg.waves_per_simd,
                     "compaction_period": cfg.compaction_period, "refill_threshold": cfg.refill_threshold},
           "batched_ms": batched_ms / 3, "batched_steps_per_s": steps / (batched_ms / 3 * 1e-3)}
    # The same batch in the polar form, with each request's drift budget.
    polar = ctx.batch_config(block=cfg.block, waves_per_simd=cfg.waves_per_simd,
                             compaction_period=cfg.compaction_period, refill_threshold=cfg.refill_threshold,
                             kerr_schild=2)
    for r in reqs:
        r["targets"].reset_counters()
    ptotals = ctx.march_batch(reqs, polar)
    # Escaping or exhausted rays over their request's drift budget, split by
    # policy: stills (4096 steps, compensated, 1e-5) and interactive requests
    # (256 steps, uncompensated, 1e-4, where most rays end exhausted).
    over = {"still": 0, "interactive": 0}
    non_finite = 0
    for r in reqs:
        u = r["uniforms"]
        debug = r["targets"].read()[1].astype(np.float64)
        kept = debug[..., 2] != api.RAY_CAPTURED
        over["still" if u.compensated else "interactive"] += int((debug[..., 0][kept] > u.drift_budget).sum())
        non_finite += int((~np.isfinite(debug[..., 0][kept])).sum())
    polar_ms, _ = time_frames(ctx, api, lambda: _batch_nosync(ctx, api, reqs, polar), 1, 3)
    out["polar"] = {"batched_ms": polar_ms / 3, "steps": int(ptotals.steps),
                    "batched_steps_per_s": int(ptotals.steps) / (polar_ms / 3 * 1e-3),
                    "rays_over_budget_escaped_or_exhausted": over["still"] + over["interactive"],
                    "rays_over_budget_by_policy": over, "rays_non_finite_drift": non_finite}
    for name, integrator in (("plain_reference", api.INTEGRATOR_REFERENCE), ("plain_fast", api.INTEGRATOR_FAST),
                             ("plain_polar", api.INTEGRATOR_POLAR)):
        def run_all():
            for r in reqs:
                ctx.march_image(r["uniforms"], r["targets"], vol, None, integrator)
        ms, _ = time_frames(ctx, api, run_all, 1, 2)
        out[name + "_ms"] = ms / 2
        out[name + "_steps_per_s"] = steps / (ms / 2 * 1e-3)
    out["speedup_vs_plain_reference"] = out["plain_reference_ms"] / out["batched_ms"]
    out["speedup_vs_plain_fast"] = out["plain_fast_ms"] / out["batched_ms"]
    out["polar"]["speedup_vs_plain_reference"] = out["plain_reference_ms"] / out["polar"]["batched_ms"]
    out["polar"]["speedup_vs_plain_polar"] = out["plain_polar_ms"] / out["polar"]["batched_ms"]
    workloads.close(reqs)
    return out


def pass_leg(ctx, api, particles, vol):
    """Disk passes at 1M particles into 128^3: one sim dispatch (two
    substeps; also fifty in one call), the exact splat bin (sorted, and the direct path it replaces)
    and the resolve with its fused clear, against their algorithmic bytes."""
    n = 1_000_000
    voxels = 128 ** 3

    def timed(fn, reps):
        fn()
        ctx.sync()
        a, b = ctx.event(), ctx.event()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        b.sync()
        return a.elapsed_ms(b) / reps

    sim_ms = timed(lambda: particles.step(1), 50)
    # 50 consecutive dispatches in one call: one launch with the particles
    # in registers (the settle and any multi-dispatch step take this path).
    fused_ms = timed(lambda: particles.step(50), 4) / 50
    from gargantua_amd import autotune
    brick = autotune.best_brick(ctx, "splat_1M_128")
    vol.brick = brick
    bin_ms = timed(lambda: vol.splat(particles, exact=True), 10)
    # The direct exact deposit (one global atomic per corner sum): inside a
    # graph capture the exact bin takes that path, so time its replays.
    ctx.graph_begin()
    try:
        vol.splat(particles, exact=True)
    finally:
        direct = ctx.graph_end()
    direct_ms = timed(lambda: direct.launch(), 10)
    direct.close()
    # The resolve after each of its own deposits (it clears only the sums a
    # deposit reached), timed alone; it leaves the volume as the other legs
    # expect it.
    vol.resolve(clear=True)
    resolve_ms, touched = resolve_after(ctx, vol, lambda: vol.splat(particles, exact=True), 20)
    sim_bytes = 64 * n
    resolve_bytes = (24 + 16) * voxels + 24 * touched
    # The ceiling for the bin: the same count of no-return atomics (8 corners
    # x 6 sums per particle) to random words of the sums' footprint.
    ceiling = ctx.atomic_rate(6 * voxels, 48, n)
    max_preset = max_preset_passes(ctx, api, timed)
    return {"max_preset_4M_256": max_preset,
            "sim": {"ms": sim_ms, "particles_per_s": n / (sim_ms * 1e-3), "bytes": sim_bytes,
                    "GB_per_s": sim_bytes / (sim_ms * 1e-3) / 1e9, "hbm_frac": sim_bytes / (sim_ms * 1e-3) / 8e12,
                    "fused_50": {"ms_per_dispatch": fused_ms, "particle_dispatches_per_s": n / (fused_ms * 1e-3)}},
            # Exact bin: 8 corners x 6 integer sums per particle. Sorted and
            # deposited per voxel brick, it issues no global atomics, so its
            # corner-sum rate is compared with what direct atomics reach.
            "splat_bin_exact": {"ms": bin_ms, "brick": list(brick), "deposits_per_s": n / (bin_ms * 1e-3),
                                "corner_sums_per_s": 8 * 6 * n / (bin_ms * 1e-3),
                                "direct_atomic_ceiling_per_s": ceiling,
                                "vs_direct_atomic_ceiling": 8 * 6 * n / (bin_ms * 1e-3) / ceiling,
                                "direct_path_ms": direct_ms,
                                "direct_path_corner_sums_per_s": 8 * 6 * n / (direct_ms * 1e-3)},
            "resolve": {"ms": resolve_ms, "voxels_per_s": voxels / (resolve_ms * 1e-3), "voxels_reached": touched,
                        "bytes": resolve_bytes,
                        "GB_per_s": resolve_bytes / (resolve_ms * 1e-3) / 1e9,
                        "hbm_frac": resolve_bytes / (resolve_ms * 1e-3) / 8e12}}


def resolve_after(ctx, vol, deposit, reps):
    """The resolve with its fused clear, each time after a deposit of its
    own (the clear writes only the voxels a deposit reached), timed alone
    with events around it; also the number of voxels one deposit reaches."""
    deposit()
    sums = vol.read_sums()
    n = len(sums[2])
    reached = int(np.count_nonzero((sums[0].reshape(n, 2) != 0).any(1) | (sums[1].reshape(n, 3) != 0).any(1)
                                   | (sums[2] != 0)))
    vol.resolve(clear=True)
    total = 0.0
    for _ in range(reps):
        deposit()
        a, b = ctx.event(), ctx.event()
        a.record()
        vol.resolve(clear=True)
        b.record()
        b.sync()
        total += a.elapsed_ms(b)
    return total / reps, reached


def max_preset_passes(ctx, api, timed):
    """The serving disk state (4M particles, 256^3): one sim dispatch (64
    bytes per particle), the stochastic splat that serving runs every frame
    (one corner per particle: sorted by voxel brick, the default, and the
    reference's pass of six no-return atomics per particle; 28 bytes of
    particle read either way) and the resolve with its fused clear (64
    bytes per voxel)."""
    n, size = 4_000_000, 256
    voxels = size ** 3
    p = api.Particles(ctx, n, seed=1)
    p.seed(api.GENERATOR_DISK)
    p.step(20)
    v = api.Volume(ctx, size)
    v.clear()
    sim_ms = timed(lambda: p.step(1), 20)
    frame = [0]

    def stochastic():
        frame[0] += 1
        v.splat(p, exact=False, frame=frame[0])

    splat_ms = timed(stochastic, 20)
    v.method = api.SPLAT_ATOMIC
    atomic_ms = timed(stochastic, 20)
    v.method = api.SPLAT_SORTED
    v.resolve(clear=True)
    resolve_ms, touched = resolve_after(ctx, v, stochastic, 20)
    resolve_bytes = 40 * voxels + 24 * touched
    ceiling = ctx.atomic_rate(6 * voxels, 6, n)
    out = {"sim": {"ms": sim_ms, "bytes": 64 * n, "GB_per_s": 64 * n / (sim_ms * 1e-3) / 1e9,
                   "hbm_frac": 64 * n / (sim_ms * 1e-3) / 8e12},
           # Sorted by brick: count, column scan, scatter of 32-byte records
           # and one LDS gather per brick, no global atomics; against it the
           # atomic pass and the rate of the same count of atomics.
           "splat_stochastic": {"ms": splat_ms, "method": "counting sort by 16x16x8 brick",
                                "deposits_per_s": n / (splat_ms * 1e-3),
                                "bytes_read": 28 * n, "GB_per_s_read": 28 * n / (splat_ms * 1e-3) / 1e9,
                                "speedup_vs_atomic_pass": atomic_ms / splat_ms,
                                "atomic_pass": {"ms": atomic_ms, "atomics_per_s": 6 * n / (atomic_ms * 1e-3),
                                                "atomic_ceiling_per_s": ceiling,
                                                "vs_atomic_ceiling": 6 * n / (atomic_ms * 1e-3) / ceiling}},
           "resolve": {"ms": resolve_ms, "voxels_reached": touched, "bytes": resolve_bytes,
                       "GB_per_s": resolve_bytes / (resolve_ms * 1e-3) / 1e9,
                       "hbm_frac": resolve_bytes / (resolve_ms * 1e-3) / 8e12}}
    v.close()
    p.close()
    return out


def fp64_leg(ctx):
    """FP64 GPU kernels of the spin 0.9 view: the capture boundary along 360
    screen azimuths (multi-section) and the 4096-entry far-field table on
    log-spaced b in [5.2, 1000]."""
    t0 = time.perf_counter()
    edge, _ = ctx.shadow_boundary(0.9, 50.0, 60.0, 360)
    shadow_ms = (time.perf_counter() - t0) * 1e3
    phi = np.arange(360) * (2 * math.pi / 360)
    x = edge * np.cos(phi)
    t0 = time.perf_counter()
    ctx.deflection_table(50.0, 4096, 5.2, 1000.0)
    table_ms = (time.perf_counter() - t0) * 1e3
    return {"shadow_boundary": {"azimuths": 360, "ms": shadow_ms, "width_rad": float(x.max() - x.min()),
                                "centre_offset_rad": float(0.5 * (x.max() + x.min()))},
            "far_field_table": {"entries": 4096, "ms": table_ms}}


def _form_name(cfg):
    return {0: "boyer_lindquist_fast", 1: "kerr_schild", 2: "polar"}.get(int(cfg.kerr_schild), "?")


def _batch_nosync(ctx, api, request, cfg):
    """One frame of the batched march, enqueued without reading counters back."""
    import ctypes as C
    arr = (api.MarchRequest * len(request))()
    for i, r in enumerate(request):
        arr[i].uniforms = r["uniforms"]
        arr[i].targets = r["targets"].h
        arr[i].volume = r["volume"].h if r.get("volume") else None
        arr[i].tables = r["tables"].h if r.get("tables") else None
        arr[i].ray_begin = r.get("ray_begin", 0)
        arr[i].ray_count = r.get("ray_count", 0)
    st = api.load().garg_march_batch(ctx.h, None, arr, len(request), C.byref(cfg), None)
    if st != 0:
        raise api.GargError(st, api.load().garg_last_error().decode())


if __name__ == "__main__":
    main()
