"""Pins of the FP64 oracle: the reference's own known-answer tests
(Tests/OracleTests/*.swift) restated against the C++ oracle, plus closed
forms chosen so that a dropped term, a wrong sign or a swapped operand fails
one of them. CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from tests.conftest import golden_passes, load_golden

PW, NEWTON = 1, 0


# ---- potentials (OracleTests.swift:5-35) -------------------------------------

def test_pw_angular_momentum_minimal_at_isco():
    r = np.arange(3.0, 20.0, 1e-4)
    L = np.array([O.lib.orc_circular_angular_momentum(x, PW) for x in r[::10]])
    best = r[::10][np.argmin(L)]
    assert abs(best - 6.0) < 1e-3


def test_pw_marginally_bound_has_zero_energy():
    assert abs(O.lib.orc_circular_energy(4.0, PW)) < 1e-12


def test_circular_speeds_closed_forms():
    assert abs(O.lib.orc_circular_speed(6.0, PW) - math.sqrt(6.0) / 4.0) < 1e-12
    assert abs(O.lib.orc_circular_speed(4.0, NEWTON) - 0.5) < 1e-12
    assert abs(O.lib.orc_angular_frequency(6.0, PW) - 1.0 / math.sqrt(96.0)) < 1e-12


def test_acceleration_points_inward_with_pw_magnitude():
    s = (C.c_double * 6)(3.0, 4.0, 0.0, 0.0, 0.0, 0.0)
    O.lib.orc_particle_step(s, PW, 1.0, 0.0)  # after one unit step, v = g
    # |g| = 1 / (r - 2)^2 = 1 / 9 at r = 5, pointing to the origin.
    assert s[3] < 0 and s[4] < 0 and s[5] == 0.0
    assert abs(math.hypot(s[3], s[4]) - 1.0 / 9.0) < 1e-12


# ---- dynamics (OracleTests.swift:37-104) ---------------------------------------

def _circular(r, phi, potential=PW):
    v = O.lib.orc_circular_speed(r, potential)
    return (C.c_double * 6)(r * math.cos(phi), r * math.sin(phi), 0.0, -v * math.sin(phi), v * math.cos(phi), 0.0)


def test_circular_orbit_stays_circular_without_drag():
    s = _circular(8.0, 0.3)
    period = 2 * math.pi / O.lib.orc_angular_frequency(8.0, PW)
    e0 = O.lib.orc_specific_energy(s, PW)
    lz0 = s[0] * s[4] - s[1] * s[3]
    lo = hi = 8.0
    for _ in range(int(100 * period / 0.1)):
        O.lib.orc_particle_step(s, PW, 0.1, 0.0)
        r = math.sqrt(s[0] ** 2 + s[1] ** 2 + s[2] ** 2)
        lo, hi = min(lo, r), max(hi, r)
    assert hi - lo < 0.1
    assert abs(O.lib.orc_specific_energy(s, PW) - e0) / abs(e0) < 1e-4
    assert abs((s[0] * s[4] - s[1] * s[3]) - lz0) / abs(lz0) < 1e-9


def test_drag_spirals_inward():
    s = _circular(12.0, 0.0)
    period = 2 * math.pi / O.lib.orc_angular_frequency(12.0, PW)
    for _ in range(int(2 * period / 0.1)):
        O.lib.orc_particle_step(s, PW, 0.1, 0.003)
    r = math.sqrt(s[0] ** 2 + s[1] ** 2)
    assert 9.5 < r < 11.5


def test_orbit_inside_isco_plunges_outside_survives():
    inside = _circular(5.0, 0.0)
    outside = _circular(8.0, 0.0)
    for s in (inside, outside):
        for i in (3, 4):
            s[i] *= 0.99
    captured = False
    for _ in range(20000):
        O.lib.orc_particle_step(inside, PW, 0.1, 0.0)
        if math.sqrt(inside[0] ** 2 + inside[1] ** 2) < 2.02:
            captured = True
            break
    assert captured
    low = 8.0
    for _ in range(20000):
        O.lib.orc_particle_step(outside, PW, 0.1, 0.0)
        low = min(low, math.sqrt(outside[0] ** 2 + outside[1] ** 2))
    assert low > 5.0


def test_tunnelling_grazing_and_runaway():
    d3 = C.c_double * 3
    d6 = C.c_double * 6
    assert O.lib.orc_particle_removed(d3(3.0, 0.1, 0.0), d6(-3.0, 0.1, 0.0, -60.0, 0.0, 0.0))
    assert not O.lib.orc_particle_removed(d3(2.6, 2.4, 0.0), d6(2.5, 2.5, 0.0, 0.5, 0.0, 0.0))
    assert O.lib.orc_particle_removed(d3(9.0, 0.0, 0.0), d6(10.0, 0.0, 0.0, 40.0, 0.0, 0.0))


# ---- disk model and edge (OracleTests.swift:106-157, GoldenTests.swift:66-77) --

def test_temperature_peaks_at_49_over_36_isco():
    peak = O.lib.orc_disk_temperature(6.0 * 49.0 / 36.0, 8000.0)
    assert abs(peak - 8000.0) < 1e-9
    for r in (7.5, 9.0):
        assert O.lib.orc_disk_temperature(r, 8000.0) < peak
    assert O.lib.orc_disk_temperature(24.0, 8000.0) < 0.7 * peak
    edge = O.lib.orc_disk_temperature(6.0, 8000.0)
    assert 0 < edge < 0.8 * 8000.0 and O.lib.orc_disk_temperature(3.0, 8000.0) > edge


def test_steady_profile_vanishes_and_scales():
    assert O.lib.orc_steady_surface_density(6.0) == 0.0
    assert O.lib.orc_steady_surface_density(24.0) == 0.0
    ratio = O.lib.orc_steady_surface_density(6.2) / O.lib.orc_steady_surface_density(6.1)
    assert abs(ratio - 2.0 * (6.1 / 6.2) ** 1.5) < 1e-9


def test_inverse_cdf_monotone_and_spans():
    t = O.inverse_cdf(1024)
    assert t[0] == 6.0 and t[-1] == 24.0
    assert np.all(np.diff(t) >= 0)
    assert 10.0 < t[512] < 20.0


def _draw_radii(table, seed, n):
    radii = np.empty(n)
    for i in range(n):
        u = O.lib.orc_unit(O.hash_draw(seed, i, 0)[0])
        pos = u * (len(table) - 1)
        k = min(int(pos), len(table) - 2)
        radii[i] = table[k] + (pos - k) * (table[k + 1] - table[k])
    return radii


def test_inverse_cdf_reproduces_profile():
    table = O.inverse_cdf(4096)
    radii = _draw_radii(table, 7, 200_000)
    counts, _ = np.histogram(np.minimum(((radii - 6.0) / 0.5).astype(int), 35), bins=np.arange(37) - 0.5)
    expected = np.array([O.lib.orc_radial_number_density(6.0 + (b + 0.5) * 0.5) for b in range(36)])
    expected *= 200_000 / expected.sum()
    for b in range(1, 35):
        if expected[b] > 2000:
            assert abs(counts[b] - expected[b]) / expected[b] < 0.05


def test_edge_recovers_analytic_profile():
    table = O.inverse_cdf(4096)
    radii = np.ascontiguousarray(_draw_radii(table, 5, 400_000))
    edge = O.lib.orc_inner_edge(O.ptr(radii), len(radii), 0.25, 6.5, 9.0)
    assert abs(edge - 6.0) < 0.1


def test_isco_golden_in_the_oracle():
    """GoldenTests.swift:39-64: 24k particles, seed 11, 1500 substeps."""
    g = load_golden("isco")
    n = 24_000
    radii = np.zeros(n)
    respawns = C.c_uint64()
    O.lib.orc_disk_run(n, 11, int(g["parameters"]["settlingSteps"]), 8, O.ptr(radii), C.byref(respawns))
    p = g["parameters"]
    edge = O.lib.orc_inner_edge(O.ptr(radii), n, p["binWidth"], p["fitInner"], p["fitOuter"])
    assert golden_passes(g, edge), edge
    assert respawns.value > 0


# ---- blackbody (OracleTests.swift:159-181) ---------------------------------------

def test_6500k_on_planckian_locus():
    c = (C.c_double * 2)()
    O.lib.orc_blackbody_chromaticity(6500.0, c)
    assert abs(c[0] - 0.3135) < 0.01 and abs(c[1] - 0.3237) < 0.01


def test_cool_red_hot_blue_and_table():
    cool, hot = (C.c_double * 3)(), (C.c_double * 3)()
    O.lib.orc_blackbody_color(2000.0, cool)
    O.lib.orc_blackbody_color(20000.0, hot)
    assert cool[0] == 1.0 and cool[1] < cool[0] and cool[2] < cool[1]
    assert hot[2] == 1.0 and hot[0] < hot[2]
    t = O.blackbody_table()
    assert t.shape == (256, 3) and np.all(t.max(axis=1) == 1.0) and t.min() >= 0.0
    assert O.lib.orc_blackbody_coordinate(1000.0) == 0.0 and O.lib.orc_blackbody_coordinate(40000.0) == 1.0


# ---- hash (OracleTests.swift:183-222) ----------------------------------------------

def test_pcg4d_known_vectors():
    # Independent restatement of Jarzynski and Olano's pcg4d in numpy, so an
    # error in either implementation's mixing order fails here.
    def ref(v):
        v = np.array(v, dtype=np.uint64)
        m = np.uint64(0xFFFFFFFF)
        v = (v * np.uint64(1664525) + np.uint64(1013904223)) & m
        x, y, z, w = (int(t) for t in v)
        M = 0xFFFFFFFF
        x = (x + y * w) & M
        y = (y + z * x) & M
        z = (z + x * y) & M
        w = (w + y * z) & M
        x ^= x >> 16
        y ^= y >> 16
        z ^= z >> 16
        w ^= w >> 16
        x = (x + y * w) & M
        y = (y + z * x) & M
        z = (z + x * y) & M
        w = (w + y * z) & M
        return (x, y, z, w)

    for v in [(0, 0, 0, 0), (1, 2, 3, 4), (0xFFFFFFFF, 7, 0x9E3779B9, 123456789)]:
        assert O.pcg4d(v) == ref(v)


def test_hash_determinism_and_streams():
    a = O.hash_draw(42, 7, 3)
    assert a == O.hash_draw(42, 7, 3)
    assert O.hash_draw(43, 7, 3) != a and O.hash_draw(42, 8, 3) != a
    assert O.hash_draw(42, 7, 4) != a and O.hash_draw(42, 7, 3, 1) != a


def test_uniform_and_gaussian_moments():
    u = np.array([O.lib.orc_unit(O.hash_draw(1, i, 0)[1]) for i in range(100_000)])
    assert abs(u.mean() - 0.5) < 0.005 and u.min() >= 0.0 and u.max() < 1.0
    g = (C.c_double * 2)()
    vals = []
    for i in range(100_000):
        h = O.hash_draw(3, i, 1)
        O.lib.orc_gaussian_pair(O.lib.orc_unit(h[0]), O.lib.orc_unit(h[1]), g)
        vals += [g[0], g[1]]
    vals = np.array(vals)
    assert abs(vals.mean()) < 0.01 and abs(vals.var() - 1.0) < 0.02


def test_fnv1a_known_answers():
    # FNV-1a 64-bit published test vectors (Fowler, Noll, Vo).
    assert O.lib.orc_fnv1a(b"", 0) == 0xcbf29ce484222325
    assert O.lib.orc_fnv1a(b"a", 1) == 0xaf63dc4c8601ec8c
    assert O.lib.orc_fnv1a(b"foobar", 6) == 0x85944171f73967e8
