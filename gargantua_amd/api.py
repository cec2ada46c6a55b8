"""ctypes bindings of libgargantua_hip.so (include/gargantua.h).

The host-side mirror of the reference's pass objects for Python callers
(tests, bench.py): a context owns the device, objects own their device
buffers, every call raises GargError with the library's sentence on
failure. Loading fails loudly when the HIP library has not been built: there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgargantua_hip.so")

u32 = C.c_uint32
u64 = C.c_uint64
f32 = C.c_float
f64 = C.c_double
vp = C.c_void_p

GARG_OK = 0
RAY_CAPTURED, RAY_ESCAPED, RAY_EXHAUSTED = 0, 1, 2
POTENTIAL_NEWTONIAN, POTENTIAL_PACZYNSKI_WIITA = 0, 1
GENERATOR_DISK, GENERATOR_ANNULUS = 0, 1
INTEGRATOR_REFERENCE, INTEGRATOR_FAST, INTEGRATOR_KERR_SCHILD, INTEGRATOR_POLAR = 0, 1, 2, 3
GARG_E_ARGUMENT, GARG_E_DEVICE, GARG_E_BUDGET, GARG_E_STATE = 1, 2, 3, 4
FRAME_AUTO = 0xFFFFFFFF  # splat frame from the particle store's dispatch count
PROBE_NOALLOC_MAX = 16384  # GARG_PROBE_NOALLOC_MAX: probes per garg_probe_noalloc call
VALU_FMA_LITERAL, VALU_FMA_VGPR, VALU_FMA_SGPR = 0, 1, 2  # garg_valu_rate operand forms
SPLAT_SORTED, SPLAT_ATOMIC = 0, 1  # garg_volume_set_splat_method


class GargError(RuntimeError):
    def __init__(self, status: int, sentence: str):
        super().__init__(f"gargantua status {status}: {sentence}")
        self.status = status


class Particle(C.Structure):
    _fields_ = [("position", f32 * 3), ("velocity", f32 * 3), ("temperature", f32), ("respawns", u32)]


class GeodesicSettings(C.Structure):
    _fields_ = [("step_factor", f32), ("step_min", f32), ("step_max", f32), ("capture_radius", f32),
                ("escape_radius", f32), ("step_cap", u32), ("empty_step_factor", f32), ("empty_step_max", f32)]


class MarchUniforms(C.Structure):
    _fields_ = [("camera_position", f32 * 3), ("camera_right", f32 * 3), ("camera_up", f32 * 3),
                ("camera_forward", f32 * 3), ("tan_half_fov", f32 * 2), ("resolution", u32 * 2),
                ("tile_origin", u32 * 2), ("volume_half_extent", f32), ("opacity_scale", f32),
                ("star_brightness", f32), ("drift_budget", f32), ("star_seed", u32), ("redshift", u32),
                ("bake_spacing", f32), ("bake_capacity", u32), ("sphere_radius", f32), ("camera_radius", f32),
                ("sphere_max_b", f32), ("sphere_table_max_b", f32), ("camera_max_b", f32), ("sky_min_b", f32),
                ("sky_max_b", f32), ("spin", f32), ("jitter", f32 * 2), ("volume_half_extent_z", f32),
                ("compensated", u32), ("geodesic", GeodesicSettings)]


class RayProbe(C.Structure):
    _fields_ = [("origin", f32 * 3), ("direction", f32 * 3)]


class RayProbeResult(C.Structure):
    _fields_ = [("position", f32 * 3), ("direction", f32 * 3), ("drift", f32), ("outcome", u32), ("steps", u32),
                ("energy", f32), ("angular_momentum", f32)]


class RayProbeFp64Result(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("direction", C.c_double * 3), ("drift", C.c_double),
                ("steps", u64), ("outcome", u32), ("reserved", u32)]


class RenderCameraFp64(C.Structure):
    _fields_ = [("position", C.c_double * 3), ("right", C.c_double * 3), ("up", C.c_double * 3),
                ("forward", C.c_double * 3), ("tan_half_fov", C.c_double * 2)]


class MarchCounters(C.Structure):
    _fields_ = [("rays", u32), ("over_drift_budget", u32), ("exhausted", u32), ("captured", u32), ("steps", u64)]


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", u32), ("clock_khz", u32),
                ("total_bytes", u64), ("free_bytes", u64), ("budget_bytes", u64)]


class MarchRequest(C.Structure):
    _fields_ = [("uniforms", MarchUniforms), ("targets", vp), ("tables", vp), ("volume", vp), ("ray_begin", u32),
                ("ray_count", u32)]


class BatchConfig(C.Structure):
    _fields_ = [("block", u32), ("waves_per_simd", u32), ("compaction_period", u32), ("refill_threshold", u32),
                ("kerr_schild", u32)]


# Every entry point of include/gargantua.h with its argument types.
P = C.POINTER
SIGNATURES = {
    "garg_last_error": (C.c_char_p, []),
    "garg_version": (C.c_char_p, []),
    "garg_ctx_create": (C.c_int, [C.c_int, u64, P(vp)]),
    "garg_ctx_destroy": (C.c_int, [vp]),
    "garg_device_info_get": (C.c_int, [vp, P(DeviceInfo)]),
    "garg_ctx_allocated": (C.c_int, [vp, P(u64)]),
    "garg_stream_create": (C.c_int, [vp, P(vp)]),
    "garg_stream_destroy": (C.c_int, [vp]),
    "garg_stream_sync": (C.c_int, [vp, vp]),
    "garg_event_create": (C.c_int, [vp, P(vp)]),
    "garg_event_destroy": (C.c_int, [vp]),
    "garg_event_record": (C.c_int, [vp, vp]),
    "garg_event_sync": (C.c_int, [vp]),
    "garg_event_elapsed_ms": (C.c_int, [vp, vp, P(f32)]),
    "garg_particles_create": (C.c_int, [vp, u32, u64, u32, P(vp)]),
    "garg_particles_destroy": (C.c_int, [vp]),
    "garg_particles_seed": (C.c_int, [vp, vp, u32, f32]),
    "garg_particles_step": (C.c_int, [vp, vp, u32]),
    "garg_particles_reseed": (C.c_int, [vp, u64]),
    "garg_particles_read": (C.c_int, [vp, vp, u32]),
    "garg_particles_write": (C.c_int, [vp, vp, u32]),
    "garg_particles_checksum": (C.c_int, [vp, P(u64)]),
    "garg_volume_create": (C.c_int, [vp, u32, f32, f32, P(vp)]),
    "garg_volume_destroy": (C.c_int, [vp]),
    "garg_volume_clear": (C.c_int, [vp, vp]),
    "garg_volume_set_splat_brick": (C.c_int, [vp, u32, u32, u32]),
    "garg_volume_splat_brick": (C.c_int, [vp, P(u32)]),
    "garg_volume_set_splat_method": (C.c_int, [vp, u32]),
    "garg_volume_splat_method": (C.c_int, [vp, P(u32)]),
    "garg_splat_bin": (C.c_int, [vp, vp, vp, u32, u32]),
    "garg_splat_resolve": (C.c_int, [vp, vp, u32]),
    "garg_volume_read_sums": (C.c_int, [vp, vp, vp, vp]),
    "garg_volume_read_fields": (C.c_int, [vp, vp, vp]),
    "garg_volume_write_fields": (C.c_int, [vp, vp, vp]),
    "garg_tables_create": (C.c_int, [vp, vp, f32, P(vp)]),
    "garg_tables_destroy": (C.c_int, [vp]),
    "garg_tables_update": (C.c_int, [vp, vp, f32]),
    "garg_tables_apply": (C.c_int, [vp, P(MarchUniforms)]),
    "garg_tables_read": (C.c_int, [vp, vp, vp, vp]),
    "garg_targets_create": (C.c_int, [vp, u32, u32, P(vp)]),
    "garg_targets_destroy": (C.c_int, [vp]),
    "garg_targets_read": (C.c_int, [vp, vp, vp, vp]),
    "garg_targets_read_counters": (C.c_int, [vp, P(MarchCounters)]),
    "garg_targets_reset_counters": (C.c_int, [vp, vp]),
    "garg_march_image": (C.c_int, [vp, vp, P(MarchUniforms), vp, vp, vp, u32, u32, u32]),
    "garg_probe": (C.c_int, [vp, vp, P(RayProbe), u32, P(MarchUniforms), vp, u32, P(RayProbeResult)]),
    "garg_probe_noalloc": (C.c_int, [vp, vp, P(RayProbe), u32, P(MarchUniforms), vp, u32, P(RayProbeResult)]),
    "garg_probe_fp64": (C.c_int, [vp, vp, P(RayProbe), u32, P(GeodesicSettings), P(RayProbeFp64Result)]),
    "garg_render_fp64": (C.c_int, [vp, vp, vp, P(RenderCameraFp64), u32, u32, P(GeodesicSettings), C.c_double, u32,
                                   P(C.c_double)]),
    "garg_request_rays": (u32, [u32, u32]),
    "garg_batch_config_default": (C.c_int, [vp, P(BatchConfig)]),
    "garg_march_batch": (C.c_int, [vp, vp, P(MarchRequest), u32, P(BatchConfig), P(MarchCounters)]),
    "garg_bake_create": (C.c_int, [vp, u32, u32, u32, P(vp)]),
    "garg_bake_destroy": (C.c_int, [vp]),
    "garg_bake_bytes": (u64, [u32, u32, u32]),
    "garg_bake_rows": (C.c_int, [vp, vp, P(MarchUniforms), vp, u32, u32]),
    "garg_bake_walk": (C.c_int, [vp, vp, P(MarchUniforms), vp, vp]),
    "garg_bake_overflow": (C.c_int, [vp, P(u32)]),
    "garg_bake_read": (C.c_int, [vp, vp, vp]),
    "garg_tonemap": (C.c_int, [vp, vp, f32, u32, f32, vp]),
    "garg_deflection_table": (C.c_int, [vp, vp, f64, u32, f64, f64, u32, vp, vp, vp]),
    "garg_shadow_boundary": (C.c_int, [vp, vp, f64, f64, f64, u32, u32, u32, u32, vp, vp]),
    "garg_particles_dispatches": (C.c_int, [vp, P(u32)]),
    "garg_atomic_rate": (C.c_int, [vp, vp, u32, u32, u32, P(f64)]),
    "garg_valu_rate": (C.c_int, [vp, vp, u32, u32, u32, P(f64)]),
    "garg_graph_begin": (C.c_int, [vp, vp]),
    "garg_graph_end": (C.c_int, [vp, vp, P(vp)]),
    "garg_graph_launch": (C.c_int, [vp, vp]),
    "garg_graph_destroy": (C.c_int, [vp]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()'). "
                           "There is no CPU fallback.")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(status: int):
    if status != GARG_OK:
        raise GargError(status, load().garg_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(vp)


# --- uniforms, cameras, policies (MarchPass.swift:61-98, Camera.swift) -----

def settings(kind: str = "still", **overrides) -> GeodesicSettings:
    """Schwarzschild.Settings.interactive / .still (Schwarzschild.swift:167-172)."""
    s = GeodesicSettings(0.02, 0.02, 0.5, 2.02, 60.0, 256, 0.1, 3.0)
    if kind == "still":
        s.step_cap = 4096
        s.empty_step_factor = 0.06
        s.empty_step_max = 1.5
    for k, v in overrides.items():
        setattr(s, k, v)
    return s


def orbit_camera(azimuth: float = 0.6, elevation: float = 0.32, distance: float = 52.0,
                 fov_y_deg: float = 50.0):
    """OrbitCamera basis (Camera.swift:26-33) in float32 like the reference."""
    az, el, dist = np.float32(azimuth), np.float32(elevation), np.float32(distance)
    c = np.cos(el, dtype=np.float32)
    pos = (dist * np.array([c * np.cos(az), c * np.sin(az), np.sin(el)], dtype=np.float32)).astype(np.float32)
    fwd = (-pos / np.float32(np.linalg.norm(pos))).astype(np.float32)
    right = np.cross(fwd, np.array([0, 0, 1], dtype=np.float32)).astype(np.float32)
    right = (right / np.float32(np.linalg.norm(right))).astype(np.float32)
    up = np.cross(right, fwd).astype(np.float32)
    return {"position": pos, "right": right, "up": up, "forward": fwd,
            "fov_y": np.float32(fov_y_deg * math.pi / 180.0), "distance": float(dist)}


def uniforms(camera: dict, width: int, height: int, geodesic: GeodesicSettings | None = None, *,
             drift_budget: float = 1e-5, redshift: bool = True, star_seed: int = 0, stars: bool = True,
             spin: float = 0.0, compensated: bool = False, tile_origin=(0, 0), bake_spacing: float = 0.0,
             volume_half_extent: float = 30.0) -> MarchUniforms:
    """MarchPass.uniforms (MarchPass.swift:67-98) without the table fields."""
    u = MarchUniforms()
    for name in ("position", "right", "up", "forward"):
        getattr(u, "camera_" + name)[:] = [float(x) for x in camera[name]]
    tan_half = np.tan(np.float32(camera["fov_y"]) * np.float32(0.5), dtype=np.float32)
    u.tan_half_fov[0] = float(tan_half * np.float32(width) / np.float32(height))
    u.tan_half_fov[1] = float(tan_half)
    u.resolution[:] = [width, height]
    u.tile_origin[:] = list(tile_origin)
    u.volume_half_extent = volume_half_extent
    u.opacity_scale = 0.0005
    u.star_brightness = 1.5 if stars else 0.0
    u.drift_budget = drift_budget
    u.star_seed = star_seed & 0xFFFFFFFF
    u.redshift = 1 if redshift else 0
    u.bake_spacing = bake_spacing
    u.bake_capacity = 96
    u.spin = spin
    u.compensated = 1 if compensated else 0
    u.geodesic = geodesic if geodesic is not None else settings("still")
    return u


def half_to_float(a: np.ndarray) -> np.ndarray:
    return a.view(np.float16).astype(np.float32)


# --- objects ------------------------------------------------------------------

class Context:
    def __init__(self, device: int = 0, budget_bytes: int = 0):
        self.lib = load()
        h = vp()
        _check(self.lib.garg_ctx_create(device, budget_bytes, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.lib.garg_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def info(self) -> DeviceInfo:
        i = DeviceInfo()
        _check(self.lib.garg_device_info_get(self.h, C.byref(i)))
        return i

    def allocated(self) -> int:
        b = u64()
        _check(self.lib.garg_ctx_allocated(self.h, C.byref(b)))
        return b.value

    def sync(self, stream=None):
        _check(self.lib.garg_stream_sync(self.h, stream))

    def event(self) -> "Event":
        return Event(self)

    def stream(self) -> "Stream":
        return Stream(self)

    def atomic_rate(self, words: int, adds_per_thread: int, threads: int) -> float:
        """No-return 32-bit atomic adds per second to random words (garg_atomic_rate)."""
        r = f64()
        _check(self.lib.garg_atomic_rate(self.h, None, words, adds_per_thread, threads, C.byref(r)))
        return r.value

    def valu_rate(self, form: int = 0, iterations: int = 2048, threads: int = 1 << 21) -> float:
        """Non-packed FP32 fused multiply-add flops per second in one operand
        form (VALU_FMA_LITERAL, VALU_FMA_VGPR, VALU_FMA_SGPR; garg_valu_rate)."""
        r = f64()
        _check(self.lib.garg_valu_rate(self.h, None, form, iterations, threads, C.byref(r)))
        return r.value

    def graph_begin(self, stream=None):
        """Record every launch enqueued on `stream` until graph_end."""
        _check(self.lib.garg_graph_begin(self.h, stream))

    def graph_end(self, stream=None) -> "Graph":
        h = vp()
        _check(self.lib.garg_graph_end(self.h, stream, C.byref(h)))
        return Graph(self, h)

    def probe_fp64(self, probes, settings: GeodesicSettings):
        """The reference's FP64 Schwarzschild integrator on the GPU
        (garg_probe_fp64) for the float launches of `probes`."""
        n = len(probes)
        arr = (RayProbe * n)()
        for i, (o, d) in enumerate(probes):
            arr[i].origin[:] = [float(x) for x in o]
            arr[i].direction[:] = [float(x) for x in d]
        out = (RayProbeFp64Result * n)()
        _check(self.lib.garg_probe_fp64(self.h, None, arr, n, C.byref(settings), out))
        return out

    def render_fp64(self, volume, cam, width: int, height: int, settings: GeodesicSettings,
                    opacity: float = 0.0005, redshift: bool = True) -> np.ndarray:
        """The reference Oracle's lensed render of `volume` on the GPU in
        double precision (garg_render_fp64): linear rgb, (height, width, 3).
        `cam` is an orbit camera as orbit_camera() returns it."""
        c = RenderCameraFp64()
        for k in ("position", "right", "up", "forward"):
            getattr(c, k)[:] = [float(x) for x in cam[k]]
        th = math.tan(float(cam["fov_y"]) * 0.5)
        c.tan_half_fov[:] = [th * width / height, th]
        out = np.zeros((height, width, 3), np.float64)
        _check(self.lib.garg_render_fp64(self.h, None, volume.h, C.byref(c), width, height, C.byref(settings),
                                         opacity, 1 if redshift else 0, out.ctypes.data_as(P(C.c_double))))
        return out

    def probe(self, probes, u: MarchUniforms, tables=None, integrator: int = INTEGRATOR_FAST, noalloc: bool = False):
        """probeRays; noalloc=True goes through garg_probe_noalloc (at most
        PROBE_NOALLOC_MAX probes, staging in the library's fixed array)."""
        n = len(probes)
        arr = (RayProbe * n)()
        for i, (o, d) in enumerate(probes):
            arr[i].origin[:] = [float(x) for x in o]
            arr[i].direction[:] = [float(x) for x in d]
        out = (RayProbeResult * n)()
        fn = self.lib.garg_probe_noalloc if noalloc else self.lib.garg_probe
        _check(fn(self.h, None, arr, n, C.byref(u), tables.h if tables else None, integrator, out))
        return out

    def march_image(self, u: MarchUniforms, targets, volume=None, tables=None, integrator: int = INTEGRATOR_FAST,
                    extent=(0, 0), stream=None):
        _check(self.lib.garg_march_image(self.h, stream, C.byref(u), volume.h if volume else None,
                                         tables.h if tables else None, targets.h, extent[0], extent[1], integrator))

    def batch_config(self, **overrides) -> BatchConfig:
        c = BatchConfig()
        _check(self.lib.garg_batch_config_default(self.h, C.byref(c)))
        for k, v in overrides.items():
            setattr(c, k, v)
        return c

    def march_batch(self, requests, config: BatchConfig | None = None, stream=None) -> MarchCounters:
        n = len(requests)
        arr = (MarchRequest * n)()
        for i, r in enumerate(requests):
            arr[i].uniforms = r["uniforms"]
            arr[i].targets = r["targets"].h
            arr[i].tables = r["tables"].h if r.get("tables") else None
            arr[i].volume = r["volume"].h if r.get("volume") else None
            arr[i].ray_begin = r.get("ray_begin", 0)
            arr[i].ray_count = r.get("ray_count", 0)
        total = MarchCounters()
        cfg = config or self.batch_config()
        _check(self.lib.garg_march_batch(self.h, stream, arr, n, C.byref(cfg), C.byref(total)))
        return total

    def deflection_table(self, r_obs: float, count: int, b_lo: float, b_hi: float, intervals: int = 2048):
        b = np.zeros(count)
        alpha = np.zeros(count)
        sweep = np.zeros(count)
        _check(self.lib.garg_deflection_table(self.h, None, r_obs, count, b_lo, b_hi, intervals, _ptr(b), _ptr(alpha),
                                              _ptr(sweep)))
        return b, alpha, sweep

    def shadow_boundary(self, spin: float, r_obs: float, inclination_deg: float, azimuths: int = 360,
                        rounds: int = 6, step_cap: int = 4096):
        edge = np.zeros(azimuths)
        xi = np.zeros(azimuths)
        _check(self.lib.garg_shadow_boundary(self.h, None, spin, r_obs, inclination_deg, azimuths, rounds, 64,
                                             step_cap, _ptr(edge), _ptr(xi)))
        return edge, xi


class Event:
    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = vp()
        _check(ctx.lib.garg_event_create(ctx.h, C.byref(h)))
        self.h = h

    def record(self, stream=None):
        _check(self.ctx.lib.garg_event_record(self.h, stream))

    def sync(self):
        _check(self.ctx.lib.garg_event_sync(self.h))

    def elapsed_ms(self, end: "Event") -> float:
        ms = f32()
        _check(self.ctx.lib.garg_event_elapsed_ms(self.h, end.h, C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            self.ctx.lib.garg_event_destroy(self.h)
            self.h = None


class Stream(C.c_void_p):
    """A HIP stream of the context. The object is the garg_stream handle
    itself, so it passes wherever an entry point takes `stream`."""

    def __init__(self, ctx: Context):
        super().__init__()
        self.ctx = ctx
        _check(ctx.lib.garg_stream_create(ctx.h, C.byref(self)))

    def sync(self):
        _check(self.ctx.lib.garg_stream_sync(self.ctx.h, self))

    def close(self):
        if self.value:
            self.ctx.lib.garg_stream_destroy(self)
            self.value = None


class Graph:
    """An instantiated capture (garg_graph_end); launch replays it."""

    def __init__(self, ctx: Context, h):
        self.ctx, self.h = ctx, h

    def launch(self, stream=None):
        _check(self.ctx.lib.garg_graph_launch(self.h, stream))

    def close(self):
        if self.h:
            self.ctx.lib.garg_graph_destroy(self.h)
            self.h = None


class Particles:
    def __init__(self, ctx: Context, count: int, seed: int = 1, potential: int = POTENTIAL_PACZYNSKI_WIITA):
        self.ctx, self.count = ctx, count
        h = vp()
        _check(ctx.lib.garg_particles_create(ctx.h, count, seed, potential, C.byref(h)))
        self.h = h

    def seed(self, generator: int = GENERATOR_DISK, sigma_z: float = 0.3, stream=None):
        _check(self.ctx.lib.garg_particles_seed(self.h, stream, generator, sigma_z))

    def step(self, dispatches: int = 1, stream=None):
        _check(self.ctx.lib.garg_particles_step(self.h, stream, dispatches))

    def read(self) -> np.ndarray:
        out = np.zeros(self.count, dtype=np.dtype([("position", "<f4", 3), ("velocity", "<f4", 3),
                                                  ("temperature", "<f4"), ("respawns", "<u4")]))
        _check(self.ctx.lib.garg_particles_read(self.h, _ptr(out), self.count))
        return out

    def write(self, records: np.ndarray):
        assert records.dtype.itemsize == 32 and len(records) == self.count
        _check(self.ctx.lib.garg_particles_write(self.h, _ptr(np.ascontiguousarray(records)), self.count))

    def checksum(self) -> int:
        v = u64()
        _check(self.ctx.lib.garg_particles_checksum(self.h, C.byref(v)))
        return v.value

    def dispatches(self) -> int:
        v = u32()
        _check(self.ctx.lib.garg_particles_dispatches(self.h, C.byref(v)))
        return v.value

    def close(self):
        if self.h:
            self.ctx.lib.garg_particles_destroy(self.h)
            self.h = None


class Volume:
    def __init__(self, ctx: Context, size: int, half_extent: float = 30.0, half_extent_z: float = 0.0):
        self.ctx, self.size = ctx, size
        h = vp()
        _check(ctx.lib.garg_volume_create(ctx.h, size, half_extent, half_extent_z, C.byref(h)))
        self.h = h
        self.half_extent = half_extent
        self.half_extent_z = half_extent_z or half_extent

    def clear(self, stream=None):
        _check(self.ctx.lib.garg_volume_clear(self.h, stream))

    # The exact deposit's LDS brick shapes (garg_volume_set_splat_brick); the first is the default.
    BRICKS = ((16, 8, 4), (8, 8, 4), (8, 8, 8), (16, 16, 4), (32, 8, 4), (16, 8, 8), (32, 16, 4), (16, 16, 8))

    @property
    def brick(self) -> tuple:
        shape = (u32 * 3)()
        _check(self.ctx.lib.garg_volume_splat_brick(self.h, shape))
        return tuple(shape[:])

    @brick.setter
    def brick(self, shape):
        _check(self.ctx.lib.garg_volume_set_splat_brick(self.h, *[int(v) for v in shape]))

    @property
    def method(self) -> int:
        """SPLAT_SORTED (the default) or SPLAT_ATOMIC: how splat() deposits;
        both give the same sums."""
        m = u32()
        _check(self.ctx.lib.garg_volume_splat_method(self.h, C.byref(m)))
        return m.value

    @method.setter
    def method(self, method: int):
        _check(self.ctx.lib.garg_volume_set_splat_method(self.h, int(method)))

    def splat(self, particles: Particles, exact: bool = True, frame: int = 0, stream=None):
        _check(self.ctx.lib.garg_splat_bin(self.h, particles.h, stream, 1 if exact else 0, frame))

    def resolve(self, clear: bool = True, stream=None):
        _check(self.ctx.lib.garg_splat_resolve(self.h, stream, 1 if clear else 0))

    def read_sums(self):
        v = self.size ** 3
        e = np.zeros(2 * v, np.uint32)
        m = np.zeros(3 * v, np.int32)
        a = np.zeros(v, np.uint32)
        _check(self.ctx.lib.garg_volume_read_sums(self.h, _ptr(e), _ptr(m), _ptr(a)))
        return e, m, a

    def read_fields(self):
        v = self.size ** 3
        e = np.zeros(4 * v, np.uint16)
        w = np.zeros(4 * v, np.uint16)
        _check(self.ctx.lib.garg_volume_read_fields(self.h, _ptr(e), _ptr(w)))
        return half_to_float(e).reshape(self.size, self.size, self.size, 4), \
            half_to_float(w).reshape(self.size, self.size, self.size, 4)

    def write_fields(self, emission: np.ndarray, velocity: np.ndarray):
        e = np.ascontiguousarray(emission, np.float32).astype(np.float16).view(np.uint16)
        v = np.ascontiguousarray(velocity, np.float32).astype(np.float16).view(np.uint16)
        _check(self.ctx.lib.garg_volume_write_fields(self.h, _ptr(e), _ptr(v)))

    def close(self):
        if self.h:
            self.ctx.lib.garg_volume_destroy(self.h)
            self.h = None


class Tables:
    def __init__(self, ctx: Context, sphere_radius: float = 30.0):
        self.ctx = ctx
        h = vp()
        _check(ctx.lib.garg_tables_create(ctx.h, None, sphere_radius, C.byref(h)))
        self.h = h

    def update(self, camera_radius: float, stream=None):
        _check(self.ctx.lib.garg_tables_update(self.h, stream, camera_radius))

    def apply(self, u: MarchUniforms):
        _check(self.ctx.lib.garg_tables_apply(self.h, C.byref(u)))

    def read(self):
        s, c, k = (np.zeros(1024, np.float32) for _ in range(3))
        _check(self.ctx.lib.garg_tables_read(self.h, _ptr(s), _ptr(c), _ptr(k)))
        return s, c, k

    def close(self):
        if self.h:
            self.ctx.lib.garg_tables_destroy(self.h)
            self.h = None


class Targets:
    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx, self.width, self.height = ctx, width, height
        h = vp()
        _check(ctx.lib.garg_targets_create(ctx.h, width, height, C.byref(h)))
        self.h = h

    def read(self):
        n = self.width * self.height
        color = np.zeros(4 * n, np.uint16)
        debug = np.zeros(4 * n, np.uint16)
        depth = np.zeros(n, np.float32)
        _check(self.ctx.lib.garg_targets_read(self.h, _ptr(color), _ptr(debug), _ptr(depth)))
        shape = (self.height, self.width, 4)
        return half_to_float(color).reshape(shape), half_to_float(debug).reshape(shape), \
            depth.reshape(self.height, self.width)

    def counters(self) -> MarchCounters:
        c = MarchCounters()
        _check(self.ctx.lib.garg_targets_read_counters(self.h, C.byref(c)))
        return c

    def reset_counters(self, stream=None):
        _check(self.ctx.lib.garg_targets_reset_counters(self.h, stream))

    def tonemap(self, exposure: float = 1.0, debug_view: bool = False, drift_budget: float = 1e-4) -> np.ndarray:
        out = np.zeros(self.width * self.height * 4, np.uint8)
        _check(self.ctx.lib.garg_tonemap(self.h, None, exposure, 1 if debug_view else 0, drift_budget, _ptr(out)))
        return out.reshape(self.height, self.width, 4)

    def close(self):
        if self.h:
            self.ctx.lib.garg_targets_destroy(self.h)
            self.h = None


class Bake:
    def __init__(self, ctx: Context, width: int, height: int, capacity: int = 96):
        self.ctx, self.width, self.height, self.capacity = ctx, width, height, capacity
        h = vp()
        _check(ctx.lib.garg_bake_create(ctx.h, width, height, capacity, C.byref(h)))
        self.h = h

    def rows(self, u: MarchUniforms, tables=None, row0: int = 0, rows: int | None = None, stream=None):
        _check(self.ctx.lib.garg_bake_rows(self.h, stream, C.byref(u), tables.h if tables else None, row0,
                                           self.height if rows is None else rows))

    def walk(self, u: MarchUniforms, volume, targets: Targets, stream=None):
        _check(self.ctx.lib.garg_bake_walk(self.h, stream, C.byref(u), volume.h if volume else None, targets.h))

    def overflow(self) -> int:
        v = u32()
        _check(self.ctx.lib.garg_bake_overflow(self.h, C.byref(v)))
        return v.value

    def read(self):
        n = self.width * self.height
        headers = np.zeros(4 * n, np.uint32)
        samples = np.zeros(4 * n * self.capacity, np.uint16)
        _check(self.ctx.lib.garg_bake_read(self.h, _ptr(headers), _ptr(samples)))
        return headers.reshape(n, 4), half_to_float(samples).reshape(self.capacity, n, 4)

    def close(self):
        if self.h:
            self.ctx.lib.garg_bake_destroy(self.h)
            self.h = None


def request_rays(width: int, height: int) -> int:
    return load().garg_request_rays(width, height)
