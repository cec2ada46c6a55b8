"""Scene builders shared by the GPU tests: the reference's probe launch sets
(Validation/Probes.swift:45-66), cameras, and oracle parameter blocks
built from the exact float32 uniforms the GPU receives."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import oracle as O
from gargantua_amd import api


def aimed(radius: float, b: float):
    """Probes.aimed: launch toward the hole from the x axis with impact parameter b."""
    ray = O.schw_launch_b(radius, b)
    return O.schw_position(ray), O.schw_direction(ray)


def camera_grid(camera: dict, columns: int, rows: int, aspect: float, as_float32: bool = True):
    """Probes.cameraGrid, optionally rounded to the float32 launch values both sides use."""
    tan_half = float(np.tan(np.float32(camera["fov_y"]) * np.float32(0.5), dtype=np.float32))
    origin = camera["position"].astype(np.float64)
    r, u, f = (camera[k].astype(np.float64) for k in ("right", "up", "forward"))
    out = []
    for row in range(rows):
        for col in range(columns):
            x = (2.0 * (col + 0.5) / columns - 1.0) * tan_half * aspect
            y = (1.0 - 2.0 * (row + 0.5) / rows) * tan_half
            d = f + x * r + y * u
            d = d / np.linalg.norm(d)
            if as_float32:
                out.append((origin.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)))
            else:
                out.append((origin, d))
    return out


def look_at(distance: float, inclination_deg: float, fov_deg: float, azimuth: float = 0.0) -> dict:
    """A camera at `distance` whose line of sight makes `inclination_deg`
    with the spin axis, looking at the hole, vertical field of view fov_deg."""
    el = math.radians(90.0 - inclination_deg)
    return api.orbit_camera(azimuth=azimuth, elevation=el, distance=distance, fov_y_deg=fov_deg)


def oracle_params(u: api.MarchUniforms, tables: tuple | None = None) -> O.MarchParams:
    """MarchParams for the oracle from the float32 uniforms the GPU got.
    `tables` are FP64 (sphere, camera, sky) tables computed by the oracle."""
    p = O.MarchParams()
    for name in ("camera_position", "camera_right", "camera_up", "camera_forward"):
        getattr(p, name)[:] = [float(x) for x in getattr(u, name)]
    p.tan_half_fov[:] = [float(u.tan_half_fov[0]), float(u.tan_half_fov[1])]
    p.width, p.height = int(u.resolution[0]), int(u.resolution[1])
    p.volume_half_extent = float(u.volume_half_extent)
    p.volume_half_extent_z = float(u.volume_half_extent_z or u.volume_half_extent)
    p.opacity_scale = float(u.opacity_scale)
    p.star_brightness = float(u.star_brightness)
    p.drift_budget = float(u.drift_budget)
    p.star_seed = int(u.star_seed)
    p.redshift = int(u.redshift)
    p.sphere_radius = float(u.sphere_radius)
    p.camera_radius = float(u.camera_radius)
    p.sphere_max_b = float(u.sphere_max_b)
    p.sphere_table_max_b = float(u.sphere_table_max_b)
    p.camera_max_b = float(u.camera_max_b)
    p.sky_min_b = float(u.sky_min_b)
    p.sky_max_b = float(u.sky_max_b)
    p.spin = float(u.spin)
    g = O.Settings()
    for f in ("step_factor", "step_min", "step_max", "capture_radius", "escape_radius", "empty_step_factor",
              "empty_step_max"):
        setattr(g, f, float(getattr(u.geodesic, f)))
    g.step_cap = int(u.geodesic.step_cap)
    p.geodesic = g
    bb = np.ascontiguousarray(O.blackbody_table())
    p.blackbody = O.ptr(bb)
    keep = [bb]
    if tables is not None:
        arrs = [np.ascontiguousarray(t, dtype=np.float64) for t in tables]
        p.sphere_table, p.camera_table, p.sky_table = (O.ptr(a) for a in arrs)
        keep += arrs
    p._keep = keep
    return p


def oracle_tables(camera_radius: float, sphere_radius: float = 30.0):
    """The three sweep tables in FP64 (DeflectionTables.swift:29-64)."""
    sphere = O.sweep_table(sphere_radius, 1024, 2048)
    camera = O.sweep_table(camera_radius, 1024, 256)
    sky = O.sky_table(camera_radius, sphere_radius - 1.0, 1024, 256)
    return sphere, camera, sky


def oracle_pixels(p: O.MarchParams, xy: np.ndarray, field=None, threads: int = 8):
    xy = np.ascontiguousarray(xy, dtype=np.int32)
    n = len(xy)
    out = (O.Pixel * n)()
    total = O.lib.orc_march_pixels(C.byref(p), C.byref(field) if field is not None else None,
                                   xy.ctypes.data_as(C.POINTER(C.c_int32)), n, threads, out)
    return out, total


def field_from_volume(volume: api.Volume):
    """The GPU's resolved volume as an oracle field (Validate.swift:279-281)."""
    e, v = volume.read_fields()
    return O.make_field(e.astype(np.float64), v[..., :3].astype(np.float64), volume.half_extent,
                        volume.half_extent_z)
