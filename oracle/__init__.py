"""ctypes bindings for the FP64 CPU oracle (oracle/liboracle.so).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package, as the checker or the timed CPU
baseline. The product (gargantua_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build(force: bool = False) -> str:
    """Compile liboracle.so with its Makefile (g++)."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _load() -> C.CDLL:
    if not os.path.exists(LIB_PATH):
        build()
    return C.CDLL(LIB_PATH)


lib = _load()

d = C.c_double
i64 = C.c_int64
u32 = C.c_uint32
u64 = C.c_uint64
dp = C.POINTER(C.c_double)


class RayState(C.Structure):
    _fields_ = [("v", d), ("r", d), ("phi", d), ("vdot", d), ("rdot", d), ("phidot", d)]


class Ray(C.Structure):
    _fields_ = [("s", RayState), ("e1", d * 3), ("e2", d * 3), ("energy", d), ("angular_momentum", d)]


class Settings(C.Structure):
    _fields_ = [("step_factor", d), ("step_min", d), ("step_max", d), ("capture_radius", d),
                ("escape_radius", d), ("step_cap", i64), ("empty_step_factor", d), ("empty_step_max", d)]


class Result(C.Structure):
    _fields_ = [("ray", Ray), ("outcome", C.c_int), ("steps", i64), ("drift", d)]


class KerrState(C.Structure):
    _fields_ = [("r", d), ("theta", d), ("phi", d), ("pr", d), ("ptheta", d), ("energy", d),
                ("angular_momentum", d)]


class KerrSettings(C.Structure):
    _fields_ = [("step_factor", d), ("step_min", d), ("step_max", d), ("capture_margin", d),
                ("escape_radius", d), ("step_cap", i64)]


class KerrResult(C.Structure):
    _fields_ = [("state", KerrState), ("outcome", C.c_int), ("steps", i64), ("drift", d)]


class KsState(C.Structure):
    _fields_ = [("x", d), ("y", d), ("z", d), ("px", d), ("py", d), ("pz", d), ("energy", d), ("phase", d)]


class Field(C.Structure):
    _fields_ = [("size", C.c_int), ("half_extent", d), ("half_extent_z", d), ("emission", dp),
                ("velocity", dp)]


class Camera(C.Structure):
    _fields_ = [("position", d * 3), ("right", d * 3), ("up", d * 3), ("forward", d * 3),
                ("tan_half_fov", d * 2)]


class MarchParams(C.Structure):
    _fields_ = [("camera_position", d * 3), ("camera_right", d * 3), ("camera_up", d * 3),
                ("camera_forward", d * 3), ("tan_half_fov", d * 2), ("width", C.c_int), ("height", C.c_int),
                ("volume_half_extent", d), ("volume_half_extent_z", d), ("opacity_scale", d),
                ("star_brightness", d), ("drift_budget", d), ("star_seed", u32), ("redshift", C.c_int),
                ("sphere_radius", d), ("camera_radius", d), ("sphere_max_b", d), ("sphere_table_max_b", d),
                ("camera_max_b", d), ("sky_min_b", d), ("sky_max_b", d), ("spin", d), ("geodesic", Settings),
                ("sphere_table", dp), ("camera_table", dp), ("sky_table", dp), ("blackbody", dp)]


class Pixel(C.Structure):
    _fields_ = [("color", d * 3), ("transmittance", d), ("drift", d), ("depth", d), ("steps", i64),
                ("outcome", C.c_int), ("falls_in", C.c_int), ("sky", d * 3)]


def _sig(name, res, *args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = list(args)
    return f


P = C.POINTER
vp = C.c_void_p
_sig("orc_pcg4d", None, P(u32), P(u32))
_sig("orc_hash_draw", None, u64, u32, u32, u32, P(u32))
_sig("orc_unit", d, u32)
_sig("orc_gaussian_pair", None, d, d, dp)
_sig("orc_fnv1a", u64, vp, u64)
for n in ("orc_potential_energy", "orc_potential_gradient", "orc_circular_speed", "orc_angular_frequency",
          "orc_circular_energy", "orc_circular_angular_momentum"):
    _sig(n, d, d, C.c_int)
_sig("orc_particle_step", None, dp, C.c_int, d, d)
_sig("orc_particle_removed", C.c_int, dp, dp)
_sig("orc_specific_energy", d, dp, C.c_int)
_sig("orc_disk_shape", d, d, d)
_sig("orc_disk_temperature", d, d, d)
_sig("orc_steady_surface_density", d, d)
_sig("orc_radial_number_density", d, d)
_sig("orc_radial_inverse_cdf", None, C.c_int, C.c_int, dp)
_sig("orc_spawn", None, u32, u32, u64, dp, C.c_int, C.c_int, C.c_int, d, dp)
_sig("orc_inner_edge", d, dp, i64, d, d, d)
_sig("orc_disk_run", None, u32, u64, C.c_int, C.c_int, dp, P(u64))
_sig("orc_blackbody_xyz", None, d, dp)
_sig("orc_blackbody_chromaticity", None, d, dp)
_sig("orc_blackbody_color", None, d, dp)
_sig("orc_blackbody_table", None, C.c_int, d, d, dp)
_sig("orc_blackbody_coordinate", d, d)
_sig("orc_settings_interactive", None, P(Settings))
_sig("orc_settings_still", None, P(Settings))
_sig("orc_schw_f", d, d)
_sig("orc_schw_launch", None, dp, dp, P(Ray))
_sig("orc_schw_launch_b", None, d, d, P(Ray))
_sig("orc_schw_step", None, P(RayState), d)
_sig("orc_schw_null_residual", d, P(RayState))
_sig("orc_schw_energy", d, P(RayState))
_sig("orc_schw_angular_momentum", d, P(RayState))
_sig("orc_schw_position", None, P(Ray), dp)
_sig("orc_schw_direction", None, P(Ray), dp)
_sig("orc_schw_integrate", None, P(Ray), P(Settings), P(Result))
_sig("orc_schw_deflection", d, P(Ray), P(Result))
_sig("orc_schw_critical_b", d, d, P(Settings), C.c_int)
_sig("orc_redshift_factor", d, d, dp, dp)
_sig("orc_max_impact_parameter", d, d)
_sig("orc_sweep", d, d, d, C.c_int)
_sig("orc_table_b", d, d, d, d)
_sig("orc_table_coordinate", d, d, d, d)
_sig("orc_periapsis", d, d)
_sig("orc_turning_sweep", d, d, d, C.c_int)
_sig("orc_escape_sweep", d, d, d, C.c_int)
_sig("orc_sweep_table", None, d, C.c_int, C.c_int, dp)
_sig("orc_sky_table", None, d, d, C.c_int, C.c_int, dp)
_sig("orc_far_field_table", None, d, C.c_int, d, d, C.c_int, dp, dp, dp)
_sig("orc_strong_deflection", d, d)
_sig("orc_kerr_settings_default", None, P(KerrSettings), i64)
_sig("orc_kerr_outer_horizon", d, d)
_sig("orc_kerr_launch", None, d, dp, dp, P(KerrState))
_sig("orc_kerr_step", None, d, P(KerrState), d)
_sig("orc_kerr_step_length", d, d, P(KerrState), P(KerrSettings))
_sig("orc_kerr_carter", d, d, P(KerrState))
_sig("orc_kerr_null_residual", d, d, P(KerrState))
_sig("orc_kerr_position", None, d, P(KerrState), dp)
_sig("orc_kerr_direction", None, d, P(KerrState), dp)
_sig("orc_kerr_integrate", None, d, P(KerrState), P(KerrSettings), P(KerrResult))
_sig("orc_kerr_radius_of", d, d, dp)
_sig("orc_kerr_bisect_critical_xi", d, d, d, d, P(KerrSettings), C.c_int)
_sig("orc_kerr_critical_xi", d, d, d)
_sig("orc_kerr_critical_eta", d, d, d)
_sig("orc_kerr_photon_orbits", None, d, dp)
_sig("orc_ks_metric", None, d, d, d, d, dp)
_sig("orc_ks_radius", d, d, d, d, d)
_sig("orc_ks_from_bl", None, d, P(KerrState), P(KsState))
_sig("orc_ks_to_bl", None, d, P(KsState), P(KerrState))
_sig("orc_ks_null_residual", d, d, P(KsState))
_sig("orc_ks_step", None, d, P(KsState), d)
_sig("orc_ks_step_length", d, d, P(KsState), P(KerrSettings))
_sig("orc_ks_position", None, d, P(KsState), dp)
_sig("orc_ks_drift", d, d, P(KsState), d, d, d)
_sig("orc_ks_integrate", None, d, P(KsState), P(KerrSettings), P(KerrResult))
_sig("orc_field_sample", None, P(Field), dp, dp, dp)
_sig("orc_luminance", d, dp)
_sig("orc_side_ratio", d, dp, C.c_int, C.c_int)
_sig("orc_camera_direction", None, P(Camera), C.c_int, C.c_int, C.c_int, C.c_int, dp)
_sig("orc_trace", None, dp, dp, P(Field), P(Settings), d, d, C.c_int, dp)
_sig("orc_render", None, C.c_int, C.c_int, P(Camera), P(Field), P(Settings), d, C.c_int, C.c_int, dp)
_sig("orc_march_pixel", None, P(MarchParams), P(Field), C.c_int, C.c_int, P(Pixel))
_sig("orc_march_pixels", i64, P(MarchParams), P(Field), P(C.c_int32), i64, C.c_int, P(Pixel))
_sig("orc_probe", None, P(MarchParams), dp, dp, dp, dp, dp, P(C.c_int), P(i64), dp, dp)


class BakedRay(C.Structure):
    _fields_ = [("count", i64), ("falls_in", C.c_int), ("overflowed", C.c_int), ("direction", d * 3),
                ("depth", d)]


_sig("orc_bake_pixel", None, P(MarchParams), C.c_int, C.c_int, i64, d, dp, P(BakedRay))
_sig("orc_walk_pixel", None, P(MarchParams), P(Field), dp, i64, C.c_int, dp, dp)
_sig("orc_starfield", None, dp, u32, dp, dp)
_sig("orc_pbr_neutral", None, dp, dp)
_sig("orc_splat_replay", None, vp, u32, u32, C.c_float, C.c_float, C.c_float, C.c_float, u32, u32, u64,
     P(u32), P(C.c_int32), P(u32))
_sig("orc_sim_replay_step", None, vp, u32, C.c_int, C.c_float, C.c_float, C.c_float, P(C.c_uint8))
_sig("orc_srgb_encode", d, d)


def vec(values) -> C.Array:
    a = (d * len(values))(*[float(v) for v in values])
    return a


def ptr(a: np.ndarray, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def settings(kind: str = "still") -> Settings:
    s = Settings()
    (lib.orc_settings_still if kind == "still" else lib.orc_settings_interactive)(C.byref(s))
    return s


def kerr_settings(step_cap: int = 4096) -> KerrSettings:
    s = KerrSettings()
    lib.orc_kerr_settings_default(C.byref(s), step_cap)
    return s


def schw_launch(origin, direction) -> Ray:
    r = Ray()
    lib.orc_schw_launch(vec(origin), vec(direction), C.byref(r))
    return r


def schw_launch_b(radius: float, b: float) -> Ray:
    r = Ray()
    lib.orc_schw_launch_b(radius, b, C.byref(r))
    return r


def schw_integrate(ray: Ray, s: Settings) -> Result:
    out = Result()
    lib.orc_schw_integrate(C.byref(ray), C.byref(s), C.byref(out))
    return out


def schw_position(ray: Ray) -> np.ndarray:
    out = vec([0, 0, 0])
    lib.orc_schw_position(C.byref(ray), out)
    return np.array(out[:])


def schw_direction(ray: Ray) -> np.ndarray:
    out = vec([0, 0, 0])
    lib.orc_schw_direction(C.byref(ray), out)
    return np.array(out[:])


def kerr_launch(spin: float, origin, direction) -> KerrState:
    s = KerrState()
    lib.orc_kerr_launch(spin, vec(origin), vec(direction), C.byref(s))
    return s


def kerr_integrate(spin: float, launch: KerrState, s: KerrSettings) -> KerrResult:
    out = KerrResult()
    lib.orc_kerr_integrate(spin, C.byref(launch), C.byref(s), C.byref(out))
    return out


def kerr_position(spin: float, s: KerrState) -> np.ndarray:
    out = vec([0, 0, 0])
    lib.orc_kerr_position(spin, C.byref(s), out)
    return np.array(out[:])


def kerr_direction(spin: float, s: KerrState) -> np.ndarray:
    out = vec([0, 0, 0])
    lib.orc_kerr_direction(spin, C.byref(s), out)
    return np.array(out[:])


def hash_draw(seed: int, index: int, event: int, lane: int = 0) -> tuple:
    out = (u32 * 4)()
    lib.orc_hash_draw(seed, index, event, lane, out)
    return tuple(out)


def pcg4d(v) -> tuple:
    a = (u32 * 4)(*v)
    out = (u32 * 4)()
    lib.orc_pcg4d(a, out)
    return tuple(out)


def inverse_cdf(count: int = 1024, resolution: int = 8192) -> np.ndarray:
    out = np.zeros(count)
    lib.orc_radial_inverse_cdf(count, resolution, ptr(out))
    return out


def blackbody_table(count: int = 256, tmin: float = 1000.0, tmax: float = 40000.0) -> np.ndarray:
    out = np.zeros((count, 3))
    lib.orc_blackbody_table(count, tmin, tmax, ptr(out))
    return out


def sweep_table(r: float, count: int, intervals: int = 2048) -> np.ndarray:
    out = np.zeros(count)
    lib.orc_sweep_table(r, count, intervals, ptr(out))
    return out


def sky_table(camera_r: float, sphere_r: float, count: int, intervals: int = 2048) -> np.ndarray:
    out = np.zeros(count)
    lib.orc_sky_table(camera_r, sphere_r, count, intervals, ptr(out))
    return out


def far_field_table(r_obs: float, count: int, b_lo: float, b_hi: float, intervals: int = 2048):
    b = np.zeros(count)
    alpha = np.zeros(count)
    sweep = np.zeros(count)
    lib.orc_far_field_table(r_obs, count, b_lo, b_hi, intervals, ptr(b), ptr(alpha), ptr(sweep))
    return b, alpha, sweep


def make_field(emission: np.ndarray, velocity: np.ndarray, half_extent: float = 30.0,
               half_extent_z: float | None = None) -> Field:
    """emission: (n, n, n, 4) z-major float64, velocity: (n, n, n, 3)."""
    n = emission.shape[0]
    e = np.ascontiguousarray(emission, dtype=np.float64)
    v = np.ascontiguousarray(velocity, dtype=np.float64)
    f = Field(n, half_extent, half_extent if half_extent_z is None else half_extent_z, ptr(e), ptr(v))
    f._keep = (e, v)
    return f


def field_sample(field: Field, position):
    e = vec([0, 0, 0, 0])
    v = vec([0, 0, 0])
    lib.orc_field_sample(C.byref(field), vec(position), e, v)
    return np.array(e[:]), np.array(v[:])
