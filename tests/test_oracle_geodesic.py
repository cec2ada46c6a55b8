"""Pins of the FP64 oracle's geodesics and tables: the reference's
GeodesicTests, KerrTests, DeflectionTests and LensedRenderTests restated,
plus closed forms (b_c = 3 sqrt 3, 4/b, Bardeen's xi and eta, the Bozza
strong-deflection limit) and 4th-order convergence. CPU only."""
import ctypes as C
import math

import numpy as np

import oracle as O
from tests.conftest import golden_passes, load_golden

B_C = 3.0 * math.sqrt(3.0)


# ---- Schwarzschild (GeodesicTests.swift) -------------------------------------

def test_launch_reproduces_origin_direction_null_condition():
    origin, direction = np.array([30.0, -10.0, 12.0]), np.array([-0.8, 0.3, -0.1])
    ray = O.schw_launch(origin, direction)
    unit = direction / np.linalg.norm(direction)
    assert np.max(np.abs(O.schw_position(ray) - origin)) < 1e-12
    assert abs(O.schw_direction(ray) @ unit - 1.0) < 1e-12
    assert abs(O.lib.orc_schw_null_residual(C.byref(ray.s))) < 1e-12
    r0 = np.linalg.norm(origin)
    assert abs(ray.energy - math.sqrt(1 - 2 / r0)) < 1e-12


def test_impact_parameter_launch():
    ray = O.schw_launch_b(50.0, 5.0)
    assert abs(ray.angular_momentum / ray.energy - 5.0) < 1e-12 and ray.s.rdot < 0


def test_conserved_quantities_and_fourth_order():
    ray = O.schw_launch([40.0, 5.0, 15.0], [-0.95, -0.2, -0.3])
    s = O.settings("still")
    s.step_factor, s.empty_step_factor = 0.02, 0.02
    s.empty_step_max = 0.5
    res = O.schw_integrate(ray, s)
    assert res.outcome != 2 and res.drift < 1e-6
    assert abs(O.lib.orc_schw_null_residual(C.byref(res.ray.s))) < 1e-6
    finer = O.settings("still")
    for f in ("step_factor", "step_min", "step_max", "empty_step_factor", "empty_step_max"):
        setattr(finer, f, getattr(s, f) * 0.5)
    finer.step_cap = 100_000
    ref = O.schw_integrate(ray, finer)
    assert ref.drift < res.drift / 8.0


def test_critical_impact_parameter_is_three_root_three():
    s = O.settings("still")
    b = O.lib.orc_schw_critical_b(50.0, C.byref(s), 60)
    assert abs(b - B_C) / B_C < 1e-4


def test_capture_and_escape():
    s = O.settings("still")
    assert O.schw_integrate(O.schw_launch_b(50.0, 4.0), s).outcome == 0
    assert O.schw_integrate(O.schw_launch_b(50.0, 7.0), s).outcome == 1


def test_weak_field_deflection_is_four_over_b():
    s = O.settings("still")
    s.step_max = s.empty_step_max = 1e300
    s.escape_radius = 50_000.0
    s.step_factor = s.empty_step_factor = 0.02
    launch = O.schw_launch_b(50_000.0, 1000.0)
    res = O.schw_integrate(launch, s)
    assert res.outcome == 1
    d = O.lib.orc_schw_deflection(C.byref(launch), C.byref(res))
    assert abs(d - 0.004) / 0.004 < 0.005


def test_radial_ray_falls_straight_in():
    ray = O.schw_launch([20.0, 0, 0], [-1.0, 0, 0])
    res = O.schw_integrate(ray, O.settings("interactive"))
    assert res.outcome == 0 and abs(res.ray.s.phi) < 1e-12 and abs(ray.angular_momentum) < 1e-12


def test_redshift_limits():
    g = lambda f, d, v: O.lib.orc_redshift_factor(f, O.vec(d), O.vec(v))
    assert abs(g(1 - 2 / 1e6, [1, 0, 0], [0, 0, 0]) - 1.0) < 1e-5
    assert abs(g(1 - 2 / 2.5, [1, 0, 0], [0, 0, 0]) - math.sqrt(1 - 2 / 2.5)) < 1e-12
    assert g(1 - 2e-6, [1, 0, 0], [-0.3, 0, 0]) > 1.0 > g(1 - 2e-6, [1, 0, 0], [0.3, 0, 0])
    fast = g(1 - 2e-6, [1, 0, 0], [-5.0, 0, 0])
    assert fast < (1 + 0.85) / math.sqrt(1 - 0.85 ** 2) * 1.0001 and math.isfinite(fast)
    plunge = g(1 - 2 / 2.05, [0, 1, 0], [0, -89.0, 0])
    assert math.isfinite(plunge) and plunge > 0


# ---- deflection tables (DeflectionTests.swift) --------------------------------

def test_nearly_radial_sweep_is_b_over_r():
    assert abs(O.lib.orc_sweep(100.0, 0.01, 2048) - 0.01 / 100.0) < 1e-9


def test_periapsis_root():
    for b in (6.0, 10.0, 30.0, 1000.0):
        p = O.lib.orc_periapsis(b)
        assert abs(1.0 - b * b / (p * p) * (1.0 - 2.0 / p)) < 1e-12 and 3.0 < p < b


def test_turning_sweep_weak_field():
    b = 1000.0
    p = O.lib.orc_periapsis(b)
    d = 2.0 * O.lib.orc_sweep(p, b, 2048) - math.pi
    assert abs(d - 4.0 / b) / (4.0 / b) < 0.005
    total = O.lib.orc_escape_sweep(1e7, b, 2048)
    assert abs(total - (d + math.pi - b / 1e7)) < 1e-7


def test_table_coordinate_round_trip():
    for t in (0.0, 0.1, 0.5, 0.9, 1.0):
        b = O.lib.orc_table_b(2.0, 30.0, t)
        assert abs(O.lib.orc_table_coordinate(2.0, 30.0, b) - t) < 1e-12


def test_sweep_agrees_with_integrator():
    b = 10.0
    launch = O.schw_launch_b(50_000.0, b)
    s = O.settings("still")
    s.step_factor = s.empty_step_factor = 0.01
    s.step_max = s.empty_step_max = 2.0
    s.escape_radius = 50_000.0
    s.step_cap = 200_000
    res = O.schw_integrate(launch, s)
    assert res.outcome == 1
    assert abs(res.ray.s.phi - O.lib.orc_turning_sweep(50_000.0, b, 2048)) < 1e-5


def test_tables_monotone_bounded():
    sphere = O.sweep_table(30.0, 256)
    assert sphere[0] == 0.0 and np.all(np.diff(sphere) >= 0) and sphere[-1] < math.pi
    sky = O.sky_table(52.0, 30.0, 256)
    assert np.all((sky > 0) & (sky < 2 * math.pi))


def test_far_field_table_matches_strong_and_weak_limits():
    """alpha(b) -> 4/b far out and -ln(b/b_c - 1) + ln[216(7 - 4 sqrt 3)] - pi
    near b_c (Bozza 2002); S(r_obs, b) -> b / r_obs for b << r_obs."""
    b, alpha, sweep = O.far_field_table(50.0, 64, 5.2, 1000.0, 2048)
    assert np.all(np.diff(alpha) < 0) and np.all(np.diff(b) > 0)
    # Weak field to second order, 4/b + 15 pi / (4 b^2) (Epstein and Shapiro 1980).
    weak = 4.0 / 1000.0 + 15.0 * math.pi / 4.0 / 1000.0 ** 2
    assert abs(alpha[-1] - weak) / weak < 1e-4
    # The strong-deflection limit is exact as b -> b_c; its error is linear in b - b_c.
    for bb in (5.197, 5.2):
        a = O.far_field_table(50.0, 2, bb, bb + 1.0, 2048)[1][0]
        assert abs(a - O.lib.orc_strong_deflection(bb)) < 0.7 * (bb - B_C)
    # Flat-space sweep from r_obs to infinity is asin(b / r_obs) far out.
    small = O.far_field_table(5000.0, 2, 5.2, 6.0, 2048)[2]
    assert abs(small[0] - math.asin(5.2 / 5000.0)) < 1e-6


# ---- Kerr (KerrTests.swift) ------------------------------------------------------

def test_photon_orbits_and_bardeen_eta():
    o = (C.c_double * 2)()
    O.lib.orc_kerr_photon_orbits(0.9, o)
    assert abs(o[0] - 2 * (1 + math.cos(2 / 3 * math.acos(-0.9)))) < 1e-12
    assert o[0] < 3.0 < o[1]
    for rc in (o[0], o[1]):
        assert abs(O.lib.orc_kerr_critical_eta(0.9, rc)) < 1e-9
    O.lib.orc_kerr_photon_orbits(1e-9, o)
    assert abs(o[0] - 3) < 1e-3 and abs(o[1] - 3) < 1e-3


def test_zero_spin_reduces_to_schwarzschild():
    for aim in (0.15, 0.3):
        k = O.kerr_launch(0.0, [50.0, 0, 0], [-1.0, aim, 0])
        kr = O.kerr_integrate(0.0, k, O.kerr_settings(100_000))
        s = O.settings("still")
        s.step_cap = 100_000
        s.empty_step_factor, s.empty_step_max = s.step_factor, s.step_max
        sr = O.schw_integrate(O.schw_launch([50.0, 0, 0], [-1.0, aim, 0]), s)
        assert kr.outcome == 1 and sr.outcome == 1
        assert abs(kr.state.phi - sr.ray.s.phi) < 2e-4


def test_carter_and_null_residual_hold():
    for d in ([-1.0, 0.12, 0.0], [-1.0, -0.09, 0.05], [-0.9, 0.05, -0.2]):
        k = O.kerr_launch(0.9, [45.0, 12.0, 9.0], d)
        assert abs(O.lib.orc_kerr_null_residual(0.9, C.byref(k))) < 1e-10
        r = O.kerr_integrate(0.9, k, O.kerr_settings(100_000))
        assert r.outcome != 2 and r.drift < 1e-7


def test_shadow_edges_match_bardeen():
    s = O.KerrSettings()
    O.lib.orc_kerr_settings_default(C.byref(s), 40_000)
    s.step_max = 1e9
    s.escape_radius = 5000.0
    o = (C.c_double * 2)()
    O.lib.orc_kerr_photon_orbits(0.9, o)
    pro, retro = O.lib.orc_kerr_critical_xi(0.9, o[0]), O.lib.orc_kerr_critical_xi(0.9, o[1])
    mp = O.lib.orc_kerr_bisect_critical_xi(0.9, 5000.0, 1.0, C.byref(s), 48)
    mr = O.lib.orc_kerr_bisect_critical_xi(0.9, 5000.0, -1.0, C.byref(s), 48)
    assert abs(mp - pro) / abs(pro) < 1e-3 and abs(mr - retro) / abs(retro) < 1e-3
    assert abs(mr / mp) > 2.0


def test_zero_angular_momentum_rays_are_dragged():
    k = O.kerr_launch(0.9, [20.0, 0, 0], [-1.0, 0, 0])
    k.angular_momentum = 0.0
    r = O.kerr_integrate(0.9, k, O.kerr_settings(100_000))
    assert r.outcome == 0 and r.state.phi > 0.05


def test_shadow_kerr_golden_oracle_side():
    """The oracle side of goldens/shadow_kerr.json: the asymmetry the GPU is held to."""
    g = load_golden("shadow_kerr")
    p = g["parameters"]
    s = O.kerr_settings(4096)
    pro = O.lib.orc_kerr_bisect_critical_xi(p["spin"], p["launchRadius"], 1.0, C.byref(s), int(p["iterations"]))
    retro = O.lib.orc_kerr_bisect_critical_xi(p["spin"], p["launchRadius"], -1.0, C.byref(s), int(p["iterations"]))
    assert abs(abs(retro / pro) - 2.4020) < 5e-4


# ---- volume field and lensed render (LensedRenderTests.swift) -------------------

def test_volume_field_texel_centres_and_outside():
    e = np.zeros((4, 4, 4, 4))
    v = np.tile(np.array([1.0, 2.0, 3.0]), (4, 4, 4, 1))
    e[1, 2, 3] = [5.0, 6.0, 7.0, 8.0]
    f = O.make_field(e, v, 2.0)
    centre = np.array([3.5, 2.5, 1.5]) - 2.0
    se, sv = O.field_sample(f, centre)
    assert np.all(se == [5, 6, 7, 8]) and np.all(sv == [1, 2, 3])
    se, _ = O.field_sample(f, centre - [0.5, 0, 0])
    assert abs(se[3] - 4.0) < 1e-12
    se, sv = O.field_sample(f, [3.0, 0, 0])
    assert se[3] == 0.0 and np.all(sv == 0)


def synthetic_disk(size: int, moving: bool):
    """SyntheticDisk.field (LensedRenderTests.swift:7-53)."""
    half = 30.0
    voxel = 2 * half / size
    c = (np.arange(size) + 0.5) * voxel - half
    z, y, x = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(x * x + y * y)
    inside = (r > 6.0) & (r < 24.0) & (np.abs(z) < 0.75)
    e = np.zeros((size, size, size, 4))
    v = np.zeros((size, size, size, 3))
    idx = np.argwhere(inside)
    for (k, j, i) in idx:
        rr = r[k, j, i]
        t = O.lib.orc_disk_temperature(rr, 8000.0)
        col = (C.c_double * 3)()
        O.lib.orc_blackbody_color(t, col)
        w = (t / 8000.0) ** 4
        e[k, j, i] = [col[0] * w, col[1] * w, col[2] * w, 100.0]
        if moving:
            s = O.lib.orc_circular_speed(rr, 1)
            v[k, j, i] = [-s * y[k, j, i] / rr, s * x[k, j, i] / rr, 0.0]
    return e, v


def synthetic_camera():
    el, az, dist = 0.32, 0.6, 52.0
    pos = np.array([dist * math.cos(el) * math.cos(az), dist * math.cos(el) * math.sin(az), dist * math.sin(el)])
    fwd = -pos / dist
    right = np.array([fwd[1], -fwd[0], 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    th = math.tan(50.0 * math.pi / 360.0)
    cam = O.Camera()
    cam.position[:], cam.right[:], cam.up[:], cam.forward[:] = list(pos), list(right), list(up), list(fwd)
    cam.tan_half_fov[:] = [th * 16 / 9, th]
    return cam


def test_static_disk_symmetric_moving_disk_beamed():
    w, h = 64, 36
    ratios = []
    for moving in (False, True):
        e, v = synthetic_disk(60, moving)
        f = O.make_field(e, v)
        img = np.zeros((h, w, 3))
        O.lib.orc_render(w, h, C.byref(synthetic_camera()), C.byref(f), C.byref(O.settings("still")), 0.0005, 1, 8,
                         O.ptr(img))
        lum = np.ascontiguousarray(0.2126 * img[..., 0] + 0.7152 * img[..., 1] + 0.0722 * img[..., 2])
        assert lum.sum() > 0
        ratios.append(O.lib.orc_side_ratio(O.ptr(lum), w, h))
    assert abs(ratios[0] - 1.0) < 0.05
    assert ratios[1] > 1.3


def test_pbr_neutral_fixed_points():
    out = (C.c_double * 3)()
    O.lib.orc_pbr_neutral(O.vec([0.5, 0.5, 0.5]), out)  # below compression: offset only
    assert abs(out[0] - 0.46) < 1e-12
    O.lib.orc_pbr_neutral(O.vec([0.0, 0.0, 0.0]), out)
    assert out[0] == 0.0
    O.lib.orc_pbr_neutral(O.vec([100.0, 100.0, 100.0]), out)
    assert 0.99 < out[0] <= 1.0


def test_goldens_parse_and_tolerance_semantics():
    names = ["beaming", "conservation_interactive", "conservation_still", "deflection", "determinism", "isco", "jump",
             "parity", "shadow", "shadow_kerr", "sky"]
    for n in names:
        g = load_golden(n)
        assert g["name"] == n and g["claim"] and g["method"]
    g = load_golden("isco")
    assert golden_passes(g, 6.29) and not golden_passes(g, 6.31) and golden_passes(g, 5.71)
