"""Synthetic scenes of the benchmark configurations, built on the device
through the C ABI: the disk (particles settled and splatted into a volume)
and the render requests marched against it.

* headline: 1920x1080, spin 0.9, 4096-step budget, observer r = 50,
  inclination 60 degrees, 60 degree vertical field of view;
* mixed(n, seed): n requests with resolution from {256x144, 640x360,
  960x540, 1920x1080, 3840x2160}, spin ~ U[0, 0.99], inclination ~
  U[10, 89] degrees, observer r ~ U[30, 100], budget 256 or 4096.
"""
from __future__ import annotations

import math

import numpy as np

from . import api

RESOLUTIONS = [(256, 144), (640, 360), (960, 540), (1920, 1080), (3840, 2160)]


def disk(ctx: api.Context, particles: int = 1_000_000, volume: int = 128, seed: int = 1, settle_substeps: int = 600,
         half_extent_z: float = 3.0):
    """Annulus generator (sigma_z 0.3), settled, exact CIC into `volume`^3
    spanning |x|,|y| <= 30 and |z| <= half_extent_z, resolved."""
    p = api.Particles(ctx, particles, seed=seed)
    p.seed(api.GENERATOR_ANNULUS, sigma_z=0.3)
    p.step(settle_substeps // 2)
    v = api.Volume(ctx, volume, 30.0, half_extent_z)
    v.clear()
    v.splat(p, exact=True)
    v.resolve(clear=True)
    ctx.sync()
    return p, v


def view(width: int, height: int, spin: float, inclination_deg: float, distance: float, cap: int,
         fov_deg: float = 60.0, azimuth: float = 0.0) -> api.MarchUniforms:
    cam = api.orbit_camera(azimuth=azimuth, elevation=math.radians(90.0 - inclination_deg), distance=distance,
                           fov_y_deg=fov_deg)
    policy = "still" if cap > 256 else "interactive"
    geo = api.settings(policy, step_cap=cap, escape_radius=max(60.0, distance + 10.0))
    return api.uniforms(cam, width, height, geo, spin=spin, compensated=policy == "still", star_seed=1,
                        drift_budget=1e-5 if policy == "still" else 1e-4)


def headline(ctx: api.Context, volume: api.Volume) -> dict:
    u = view(1920, 1080, 0.9, 60.0, 50.0, 4096)
    return {"uniforms": u, "targets": api.Targets(ctx, 1920, 1080), "volume": volume}


def mixed(ctx: api.Context, volume: api.Volume, n: int = 64, seed: int = 4) -> list[dict]:
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        w, h = RESOLUTIONS[int(rng.integers(0, len(RESOLUTIONS)))]
        spin = float(rng.uniform(0.0, 0.99))
        incl = float(rng.uniform(10.0, 89.0))
        dist = float(rng.uniform(30.0, 100.0))
        cap = 256 if rng.random() < 0.5 else 4096
        u = view(w, h, spin, incl, dist, cap)
        out.append({"uniforms": u, "targets": api.Targets(ctx, w, h), "volume": volume})
    return out


def close(requests):
    for r in requests:
        r["targets"].close()
