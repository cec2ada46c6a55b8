"""The autotuner. For the persistent march it sweeps block size, resident
waves per SIMD, compaction period and refill threshold on a workload; for
the exact (sorted, atomic-free) splat it sweeps the LDS brick shape of the
gather (garg_volume_set_splat_brick). The fastest configuration per (arch,
workload) persists as <table>/<arch>/<workload>.json = {"config": {...},
"ms": ..., "date": ...}, which callers load with `best()` and
`best_brick()`. The table is GARG_AUTOTUNE_DIR, else gargantua_amd/autotune;
`gargantua bench --sweep-batch` writes the march's files in the same form.

python -m gargantua_amd.autotune --workload headline|mixed|splat_1M_128|splat_4M_256
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import time

from . import api, workloads

def table_dir() -> str:
    return os.environ.get("GARG_AUTOTUNE_DIR") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune")


def entry_path(arch: str, workload: str) -> str:
    return os.path.join(table_dir(), arch, f"{workload}.json")

# refill_threshold 0 selects whole-tile claims, where the compaction period
# plays no part; the grid visits it once per block size and wave count.
GRID = {
    "block": [64, 128, 256],
    "waves_per_simd": [4, 5, 8],
    "compaction_period": [16, 32, 64, 128],
    "refill_threshold": [0, 16, 32, 48, 63],
}


def time_config(ctx: api.Context, requests: list, cfg: api.BatchConfig, repeats: int) -> float:
    ctx.march_batch(requests, cfg)  # warm
    a, b = ctx.event(), ctx.event()
    a.record()
    for _ in range(repeats):
        _enqueue(ctx, requests, cfg)
    b.record()
    b.sync()
    return a.elapsed_ms(b) / repeats


def _enqueue(ctx, requests, cfg):
    import ctypes as C
    arr = (api.MarchRequest * len(requests))()
    for i, r in enumerate(requests):
        arr[i].uniforms = r["uniforms"]
        arr[i].targets = r["targets"].h
        arr[i].tables = r["tables"].h if r.get("tables") else None
        arr[i].volume = r["volume"].h if r.get("volume") else None
    st = api.load().garg_march_batch(ctx.h, None, arr, len(requests), C.byref(cfg), None)
    if st != 0:
        raise api.GargError(st, api.load().garg_last_error().decode())


FORMS = {"fast": 0, "kerr-schild": 1, "polar": 2}  # garg_batch_config.kerr_schild


def sweep(ctx: api.Context, requests: list, grid: dict | None = None, repeats: int = 2, log=print,
          kerr_schild: int = 2) -> list:
    grid = grid or GRID
    results = []
    seen = set()
    for block, waves, period, thr in itertools.product(*(grid[k] for k in GRID)):
        if thr == 0:
            if (block, waves) in seen:
                continue
            seen.add((block, waves))
        cfg = ctx.batch_config(block=block, waves_per_simd=waves, compaction_period=period, refill_threshold=thr,
                               kerr_schild=kerr_schild)
        ms = time_config(ctx, requests, cfg, repeats)
        results.append(({"block": block, "waves_per_simd": waves, "compaction_period": period,
                          "refill_threshold": thr, "kerr_schild": kerr_schild}, ms))
        log(f"block {block:3d} waves {waves} period {period:3d} threshold {thr:2d}: {ms:9.3f} ms")
    results.sort(key=lambda r: r[1])
    return results


def load_entry(arch: str, workload: str) -> dict | None:
    try:
        with open(entry_path(arch, workload)) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


def best(ctx: api.Context, workload: str) -> api.BatchConfig:
    """The persisted winner for this device and workload, else the default."""
    entry = load_entry(ctx.info().arch.decode(), workload)
    if entry and isinstance(entry.get("config"), dict):
        return ctx.batch_config(**entry["config"])
    return ctx.batch_config()


def save(arch: str, workload: str, config: dict, ms: float):
    path = entry_path(arch, workload)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"config": config, "ms": ms, "date": time.strftime("%Y-%m-%d")}, f, indent=1, sort_keys=True)


# Exact-splat workloads: particles settled on the annulus, volume size.
SPLAT_WORKLOADS = {"splat_1M_128": (1_000_000, 128), "splat_4M_256": (4_000_000, 256)}


def sweep_bricks(ctx: api.Context, particles: api.Particles, volume: api.Volume, repeats: int = 5,
                 log=print) -> list:
    """Every brick shape of the exact deposit, timed over `repeats` deposits
    (sort, stage and gather) with device events; fastest first."""
    results = []
    for shape in api.Volume.BRICKS:
        volume.brick = shape
        volume.splat(particles, exact=True)  # warm
        a, b = ctx.event(), ctx.event()
        a.record()
        for _ in range(repeats):
            volume.splat(particles, exact=True)
        b.record()
        b.sync()
        ms = a.elapsed_ms(b) / repeats
        results.append((list(shape), ms))
        log(f"brick {shape[0]:2d}x{shape[1]:2d}x{shape[2]}: {ms:8.4f} ms")
    volume.brick = api.Volume.BRICKS[0]
    results.sort(key=lambda r: r[1])
    return results


def brick_for(arch: str, workload: str) -> tuple:
    """The persisted brick shape for this arch and splat workload, else the
    default (also for a shape the library does not build)."""
    entry = load_entry(arch, workload)
    shape = entry.get("config", {}).get("brick") if entry else None
    if isinstance(shape, list) and tuple(shape) in api.Volume.BRICKS:
        return tuple(shape)
    return api.Volume.BRICKS[0]


def best_brick(ctx: api.Context, workload: str) -> tuple:
    return brick_for(ctx.info().arch.decode(), workload)


def main():
    ap = argparse.ArgumentParser(description="persistent march and exact splat autotuner")
    ap.add_argument("--workload", choices=["headline", "mixed", *SPLAT_WORKLOADS], default="headline")
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--quick", action="store_true", help="a reduced grid")
    ap.add_argument("--form", choices=sorted(FORMS), default="polar", help="integrator form of spinning requests")
    args = ap.parse_args()
    ctx = api.Context()
    if args.workload in SPLAT_WORKLOADS:
        count, size = SPLAT_WORKLOADS[args.workload]
        p, v = workloads.disk(ctx, particles=count, volume=size)
        results = sweep_bricks(ctx, p, v, max(args.repeats, 5))
        shape, ms = results[0]
        save(ctx.info().arch.decode(), args.workload, {"brick": shape}, ms)
        print(json.dumps({"workload": args.workload, "best": {"brick": shape}, "ms": ms, "worst_ms": results[-1][1]}))
        v.close()
        p.close()
        ctx.close()
        return
    p, v = workloads.disk(ctx)
    reqs = [workloads.headline(ctx, v)] if args.workload == "headline" else workloads.mixed(ctx, v)
    grid = GRID
    if args.quick:
        grid = {"block": [64, 256], "waves_per_simd": [4, 5], "compaction_period": [16, 64],
                "refill_threshold": [0, 32, 63]}
    results = sweep(ctx, reqs, grid, args.repeats, kerr_schild=FORMS[args.form])
    cfg, ms = results[0]
    save(ctx.info().arch.decode(), args.workload, cfg, ms)
    print(json.dumps({"workload": args.workload, "best": cfg, "ms": ms, "worst_ms": results[-1][1]}))
    workloads.close(reqs)
    v.close()
    p.close()
    ctx.close()


if __name__ == "__main__":
    main()
