"""Checks that pin the oracle functions the closed-form and golden tests
leave open: the fp32 replays of the sim dispatch and of the integer deposit,
the starfield, the sphere jump of the march restatement, the camera rays
and the sRGB curve. Each is checked against something independent of its
own formula: the FP64 dynamics, conservation laws of the deposit, the
hash's threshold probability, the fully integrated march, geometry, or the
published sRGB values."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import oracle as O

PARTICLE = np.dtype([("position", "<f4", 3), ("velocity", "<f4", 3), ("temperature", "<f4"), ("respawns", "<u4")])


def _ring(n, r_lo, r_hi, seed, z_sigma=0.2, perturb=0.01):
    rng = np.random.default_rng(seed)
    r = rng.uniform(r_lo, r_hi, n)
    phi = rng.uniform(0, 2 * math.pi, n)
    rec = np.zeros(n, PARTICLE)
    rec["position"][:, 0] = r * np.cos(phi)
    rec["position"][:, 1] = r * np.sin(phi)
    rec["position"][:, 2] = rng.normal(0, z_sigma, n)
    v = np.array([O.lib.orc_circular_speed(float(x), 1) for x in r])
    rec["velocity"][:, 0] = -v * np.sin(phi) + perturb * rng.normal(size=n)
    rec["velocity"][:, 1] = v * np.cos(phi) + perturb * rng.normal(size=n)
    rec["velocity"][:, 2] = perturb * rng.normal(size=n)
    return rec


def test_sim_replay_agrees_with_fp64_dynamics():
    """The fp32 dispatch replay (sim.metal:44-79) against two FP64 substeps
    of ParticleDynamics.step and its removal rule: a dropped term, a wrong
    sign or a swapped axis in either moves the orbit by far more than fp32
    rounding."""
    n, dt, alpha, scale = 2000, 0.1, 0.003, 8000.0
    rec = _ring(n, 7.0, 24.0, 5)
    before = rec.copy()
    flags = np.zeros(n, np.uint8)
    O.lib.orc_sim_replay_step(rec.ctypes.data_as(C.c_void_p), n, 1, C.c_float(dt), C.c_float(alpha),
                              C.c_float(scale), flags.ctypes.data_as(C.POINTER(C.c_uint8)))
    worst = 0.0
    for i in range(n):
        s = (C.c_double * 6)(*[float(x) for x in before["position"][i]], *[float(x) for x in before["velocity"][i]])
        removed = False
        for _ in range(2):
            b = (C.c_double * 3)(s[0], s[1], s[2])
            O.lib.orc_particle_step(s, 1, dt, alpha)
            if O.lib.orc_particle_removed(b, s):
                removed = True
                break
        assert bool(flags[i]) == removed
        if removed:
            continue
        want = np.array(s[:])
        got = np.concatenate([rec["position"][i], rec["velocity"][i]]).astype(np.float64)
        worst = max(worst, float(np.max(np.abs(got - want) / np.maximum(np.abs(want), 1.0))))
        r = float(np.linalg.norm(want[:3]))
        t = scale * O.lib.orc_disk_shape(r, 0.01)
        assert abs(rec["temperature"][i] - t) <= 1e-4 * t + 1e-3
    assert worst < 2e-5


def test_sim_replay_removes_what_the_rule_removes():
    """Swept capture through the hole, escape past r = 60 and the speed limit."""
    rec = np.zeros(3, PARTICLE)
    rec["position"][0] = (2.5, 0, 0)
    rec["velocity"][0] = (-5.0, 0.0, 0.0)    # sweeps through r < 2.02 within the step
    rec["position"][1] = (59.95, 0, 0)
    rec["velocity"][1] = (3.0, 0.0, 0.0)     # leaves r = 60
    rec["position"][2] = (20.0, 0, 0)
    rec["velocity"][2] = (0.0, 0.25, 0.0)    # slow, bound: stays
    flags = np.zeros(3, np.uint8)
    O.lib.orc_sim_replay_step(rec.ctypes.data_as(C.c_void_p), 3, 1, C.c_float(0.1), C.c_float(0.003),
                              C.c_float(8000.0), flags.ctypes.data_as(C.POINTER(C.c_uint8)))
    assert list(flags) == [1, 1, 0]


def _splat(rec, size, exact, half=30.0, half_z=30.0, frame=0, seed=3):
    e = np.zeros(2 * size ** 3, np.uint32)
    m = np.zeros(3 * size ** 3, np.int32)
    a = np.zeros(size ** 3, np.uint32)
    O.lib.orc_splat_replay(rec.ctypes.data_as(C.c_void_p), len(rec), size, C.c_float(half), C.c_float(half_z),
                           C.c_float(8000.0), C.c_float(0.0005), 1 if exact else 0, frame, seed,
                           e.ctypes.data_as(C.POINTER(C.c_uint32)), m.ctypes.data_as(C.POINTER(C.c_int32)),
                           a.ctypes.data_as(C.POINTER(C.c_uint32)))
    return e, m, a


def test_splat_replay_voxel_centre_lands_in_one_voxel():
    """A particle on a texel centre deposits everything into that voxel,
    with the fixed-point scales of splat.metal (2^24 energy, 2^20 energy x
    temperature, 4096 momentum, 256 mass): pins the index order (x fastest,
    then y, then z) and the voxel-centre grid."""
    size, half = 16, 30.0
    voxel = 2 * half / size
    i, j, k = 3, 11, 6
    rec = np.zeros(1, PARTICLE)
    rec["position"][0] = ((i + 0.5) * voxel - half, (j + 0.5) * voxel - half, (k + 0.5) * voxel - half)
    rec["velocity"][0] = (0.25, -0.5, 0.125)
    rec["temperature"][0] = 8000.0
    e, m, a = _splat(rec, size, True)
    index = (k * size + j) * size + i
    assert a[index] == 256 and a.sum() == 256
    assert e[2 * index] == round(0.0005 * 2 ** 24) and e[2 * index + 1] == round(0.0005 * 2 ** 20)
    assert list(m[3 * index:3 * index + 3]) == [1024, -2048, 512]


def test_splat_replay_conserves_mass_momentum_and_centroid():
    """Cloud-in-cell conserves mass and momentum and keeps the first moment
    of every particle (up to the round-to-nearest of each corner), and the
    stochastic one-corner deposit conserves mass exactly."""
    size, half = 32, 30.0
    voxel = 2 * half / size
    rec = _ring(4000, 6.0, 24.0, 9, z_sigma=1.0, perturb=0.02)
    rec["temperature"] = 6000.0
    e, m, a = _splat(rec, size, True)
    n = len(rec)
    assert abs(int(a.sum()) - 256 * n) <= 4 * n
    mom = m.reshape(-1, 3).sum(0).astype(np.float64) / 4096.0
    want = rec["velocity"].astype(np.float64).sum(0)
    assert np.all(np.abs(mom - want) <= 0.002 * n / 4096.0 * 8 + 1e-9 + 1e-3 * np.abs(want).max())
    centres = (np.arange(size) + 0.5) * voxel - half
    grid = a.reshape(size, size, size).astype(np.float64)
    total = grid.sum()
    centroid = np.array([(grid.sum((0, 1)) * centres).sum(), (grid.sum((0, 2)) * centres).sum(),
                         (grid.sum((1, 2)) * centres).sum()]) / total
    assert np.allclose(centroid, rec["position"].astype(np.float64).mean(0), atol=0.01)
    _, _, s = _splat(rec, size, False, frame=5)
    assert int(s.sum()) == 256 * n


def test_starfield_threshold_fraction_and_faces():
    """pcg4d(cell, face, seed).x <= 0x30000000 lights a cell, so 3/16 of
    the directions fall on a star cell; the six faces of the cube map are
    separate streams, and the seed changes the sky."""
    rng = np.random.default_rng(2)
    dirs = rng.normal(size=(20000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    bb = np.ascontiguousarray(O.blackbody_table())
    out = (C.c_double * 3)()
    lit = []
    for d in dirs:
        O.lib.orc_starfield(O.vec(d), 1, O.ptr(bb), out)
        lit.append(out[0] + out[1] + out[2] > 0.0)
    frac = float(np.mean(lit))
    assert abs(frac - (0x30000000 + 1) / 2 ** 32) < 0.01
    O.lib.orc_starfield(O.vec(dirs[0]), 1, O.ptr(bb), out)
    a = tuple(out)
    O.lib.orc_starfield(O.vec(dirs[0]), 1, O.ptr(bb), out)
    assert tuple(out) == a
    differs = 0
    for d in dirs[:2000]:
        O.lib.orc_starfield(O.vec(d), 1, O.ptr(bb), out)
        x = out[0]
        O.lib.orc_starfield(O.vec(d), 2, O.ptr(bb), out)
        differs += x != out[0]
    assert differs > 200


def test_camera_rays_geometry():
    """Pixel centres at (p + 0.5): the rays of mirrored pixels are mirror
    images about the forward axis and the corner ray makes the half field
    of view with it."""
    cam = O.Camera()
    cam.position[:] = [0.0, 0.0, 50.0]
    cam.right[:] = [1.0, 0.0, 0.0]
    cam.up[:] = [0.0, 1.0, 0.0]
    cam.forward[:] = [0.0, 0.0, -1.0]
    th = math.tan(math.radians(30.0))
    cam.tan_half_fov[:] = [th * 16 / 9, th]
    w, h = 160, 90
    a, b = (C.c_double * 3)(), (C.c_double * 3)()
    O.lib.orc_camera_direction(C.byref(cam), 10, 20, w, h, a)
    O.lib.orc_camera_direction(C.byref(cam), w - 11, h - 21, w, h, b)
    assert abs(a[0] + b[0]) < 1e-15 and abs(a[1] + b[1]) < 1e-15 and abs(a[2] - b[2]) < 1e-15
    O.lib.orc_camera_direction(C.byref(cam), 0, 0, w, h, a)
    x = (2 * 0.5 / w - 1) * th * 16 / 9
    y = (1 - 2 * 0.5 / h) * th
    assert abs(math.acos(-a[2]) - math.atan(math.hypot(x, y))) < 1e-12
    assert a[0] < 0 and a[1] > 0  # top-left pixel looks left and up


def test_srgb_encode_published_values():
    """IEC 61966-2-1: linear segment 12.92 x below 0.0031308, then
    1.055 x^(1/2.4) - 0.055."""
    assert O.lib.orc_srgb_encode(0.0) == 0.0
    assert abs(O.lib.orc_srgb_encode(1.0) - 1.0) < 1e-12
    assert abs(O.lib.orc_srgb_encode(0.0031308) - 0.040449936) < 1e-8
    assert abs(O.lib.orc_srgb_encode(0.5) - 0.7353569830524495) < 1e-12
    assert abs(O.lib.orc_srgb_encode(0.18) - 0.4613561295004) < 1e-9


def _disk_field():
    from tests.test_oracle_geodesic import synthetic_disk
    e, v = synthetic_disk(48, True)
    return O.make_field(e, v)


def test_march_restatement_jump_matches_full_integration():
    """The sphere jump of the march restatement (Geodesic.h:102-172) against
    the same march integrating every ray: the jump golden's bound (1% of
    luminance) on a synthetic moving disk, with sweep tables from the oracle."""
    from tests import scenes
    w, h = 96, 54
    dist, el = 52.0, 0.32
    pos = np.array([dist * math.cos(el), 0.0, dist * math.sin(el)])
    fwd = -pos / dist
    right = np.array([fwd[1], -fwd[0], 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = math.tan(math.radians(25.0))
    field = _disk_field()
    xs, ys = np.meshgrid(np.arange(w), np.arange(h))
    xy = np.stack([xs.ravel(), ys.ravel()], 1)
    lums = []
    for jump in (False, True):
        p = O.MarchParams()
        p.camera_position[:] = pos
        p.camera_right[:] = right
        p.camera_up[:] = up
        p.camera_forward[:] = fwd
        p.tan_half_fov[:] = [th * w / h, th]
        p.width, p.height = w, h
        p.volume_half_extent = p.volume_half_extent_z = 30.0
        p.opacity_scale, p.drift_budget, p.redshift = 5e-4, 1e-5, 1
        p.geodesic = O.settings("still")
        bb = np.ascontiguousarray(O.blackbody_table())
        p.blackbody = O.ptr(bb)
        keep = [bb]
        if jump:
            tables = [np.ascontiguousarray(t) for t in scenes.oracle_tables(dist)]
            keep += tables
            p.sphere_table, p.camera_table, p.sky_table = (O.ptr(t) for t in tables)
            p.sphere_radius, p.camera_radius = 30.0, dist
            p.sphere_max_b = p.sky_min_b = O.lib.orc_max_impact_parameter(29.0)
            p.sphere_table_max_b = O.lib.orc_max_impact_parameter(30.0)
            p.camera_max_b = p.sky_max_b = O.lib.orc_max_impact_parameter(dist)
        pix, _ = scenes.oracle_pixels(p, xy, field)
        c = np.array([px.color[:] for px in pix])
        lums.append(0.2126 * c[:, 0] + 0.7152 * c[:, 1] + 0.0722 * c[:, 2])
        del keep
    assert lums[0].sum() > 0
    measured = np.abs(lums[0] - lums[1]).sum() / lums[0].sum()
    print(f"jump vs full integration: {measured:.2e}")
    assert measured < 0.01
