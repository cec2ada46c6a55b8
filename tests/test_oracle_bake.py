"""Pins of the oracle's bake and walk restatements (orc_bake_pixel,
orc_walk_pixel in oracle/render.cpp; bake.metal:17-92 and 94-144), which
judge the GPU bake and walk kernels.

The bake stores a ray's in-slab points at least `spacing` of path apart, up
to the capacity; the walk composites the stored points against a field with
the full Doppler factor cubed at f(|position|) along the chord from the
previous point. On a straight ray far from the hole (the uniform-slab scene
of test_oracle_gather.py) both have closed forms: with constant path h per
step and spacing between 5h and 6h a point is stored every 6 steps, so the
count is the number of whole 6h intervals in the path, each point sits 6h
further along the ray with 6h stored as its path, the depth is the path to
the first point, and the walk is Beer-Lambert with a 6h step. The walk over
the bake also reproduces the march restatement's image of the SyntheticDisk
fixture within the bound of the reference's walk-vs-march check.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from tests.test_oracle_gather import (EMISSION, ESCAPE, FLOOR, HALF, HALF_Z, L, LUMA, OPACITY, R_ESCAPE, X0, Y0,
                                      _params, _settings, _slab, _speed_for_beta)
from tests.test_oracle_geodesic import synthetic_camera, synthetic_disk

H, SPACING, EVERY = 0.1, 0.55, 6


def _bake(p, px, py, capacity, spacing):
    samples = np.zeros(4 * max(capacity, 1))
    header = O.BakedRay()
    O.lib.orc_bake_pixel(C.byref(p), px, py, capacity, spacing, O.ptr(samples), C.byref(header))
    return samples.reshape(-1, 4)[:header.count], header


def _walk(p, field, samples, header):
    s = np.ascontiguousarray(samples, dtype=np.float64)
    out = (C.c_double * 3)()
    direction = (C.c_double * 3)(*header.direction)
    O.lib.orc_walk_pixel(C.byref(p), C.byref(field) if field is not None else None, O.ptr(s) if len(s) else None,
                         len(s), header.falls_in, direction, out)
    return np.array(out[:])


def _slab_params(redshift=False, z=0.0):
    return _params((X0, Y0, z), (1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), (0.1, 0.1), 1, 1,
                   _settings(H, ESCAPE), 0.0, HALF, HALF_Z, redshift)


def test_bake_stores_a_point_every_spacing_along_the_ray():
    p = _slab_params()
    samples, header = _bake(p, 0, 0, 4000, SPACING)
    steps = L / H
    assert abs(header.count - steps // EVERY) <= 1
    assert not header.overflowed and not header.falls_in
    k = np.arange(header.count)
    np.testing.assert_allclose(samples[:, 0], X0 + EVERY * H * (k + 1), atol=0.05)
    np.testing.assert_allclose(samples[:, 1], Y0, atol=2.0)  # bending toward the hole, 4/b over the path
    np.testing.assert_allclose(samples[:, 2], 0.0, atol=1e-9)
    np.testing.assert_allclose(samples[:, 3], EVERY * H, rtol=1e-3)
    assert abs(header.depth - EVERY * H / (2.0 * R_ESCAPE)) < 1e-6
    np.testing.assert_allclose(header.direction, [1.0, 0.0, 0.0], atol=1e-3)


def test_bake_caps_at_the_capacity_and_flags_the_overflow():
    p = _slab_params()
    capped, header = _bake(p, 0, 0, 96, SPACING)
    full, _ = _bake(p, 0, 0, 4000, SPACING)
    assert header.count == 96 and header.overflowed
    np.testing.assert_array_equal(capped, full[:96])


def test_bake_outside_the_slab_stores_nothing_and_reports_the_whole_path():
    p = _slab_params(z=10.0)  # above |z| < 6
    samples, header = _bake(p, 0, 0, 96, SPACING)
    assert header.count == 0 and not header.overflowed
    assert header.depth == 1.0  # min(path / (2 R_ESCAPE), 1) with a path of 600


def test_walk_is_beer_lambert_with_the_spacing_as_its_step():
    k = 1.0 / L
    p = _slab_params()
    samples, header = _bake(p, 0, 0, 4000, SPACING)
    color = _walk(p, _slab(k / OPACITY), samples, header)
    delta = EVERY * H
    path = delta * header.count
    want = EMISSION * (1.0 - math.exp(-k * path)) / k * (1.0 + 0.5 * k * delta)
    np.testing.assert_allclose(color, want, rtol=1e-3)


def test_walk_stops_at_the_transmittance_floor():
    k = 20.0 / L
    p = _slab_params()
    samples, header = _bake(p, 0, 0, 4000, SPACING)
    color = _walk(p, _slab(k / OPACITY), samples, header)
    delta = EVERY * H
    q = math.exp(-k * delta)
    # The walk stops after the first point that takes T to the floor or below.
    n = math.ceil(math.log(FLOOR) / math.log(q))
    want = EMISSION * delta * (1.0 - q ** n) / (1.0 - q)
    np.testing.assert_allclose(color, want, rtol=1e-5)


@pytest.mark.parametrize("sign", [-1.0, 1.0], ids=["approaching", "receding"])
def test_walk_takes_the_full_doppler_factor_at_any_density(sign):
    """Unlike the march, the walk has no beaming floor: at density 10 the
    factor is f^(3/2) D^3, f at the point's radius."""
    beta = 0.5
    k = 10.0 * OPACITY
    p = _slab_params(redshift=True)
    samples, header = _bake(p, 0, 0, 4000, SPACING)
    color = _walk(p, _slab(10.0, (sign * _speed_for_beta(beta), 0.0, 0.0)), samples, header)
    doppler = math.sqrt((1 + beta) / (1 - beta)) if sign < 0 else math.sqrt((1 - beta) / (1 + beta))
    delta = EVERY * H
    q = math.exp(-k * delta)
    n = min(header.count, math.ceil(math.log(FLOOR) / math.log(q)))
    f = 1.0 - 2.0 / math.hypot(X0, Y0)
    want = EMISSION * f ** 1.5 * doppler ** 3 * delta * (1.0 - q ** n) / (1.0 - q)
    np.testing.assert_allclose(color, want, rtol=2e-4)


def test_walk_over_the_bake_reproduces_the_march_image():
    """The walked image of the SyntheticDisk against the march restatement's
    image of the same view, no stars: the bake's spacing and the walk's
    chord Doppler direction keep them within 5% of the total luminance."""
    w, h = 48, 27
    e, v = synthetic_disk(48, True)
    field = O.make_field(e, v)
    cam = synthetic_camera()
    s = O.settings("still")
    p = _params(cam.position, cam.forward, cam.right, cam.up, cam.tan_half_fov, w, h, s, 0.0, 30.0, 30.0, True)
    xs, ys = np.meshgrid(np.arange(w), np.arange(h))
    xy = np.ascontiguousarray(np.stack([xs.ravel(), ys.ravel()], 1), dtype=np.int32)
    pix = (O.Pixel * len(xy))()
    O.lib.orc_march_pixels(C.byref(p), C.byref(field), xy.ctypes.data_as(C.POINTER(C.c_int32)), len(xy), 8, pix)
    march = np.array([px.color[:] for px in pix]) @ LUMA
    walked = np.zeros(w * h)
    for i, (x, y) in enumerate(xy):
        samples, header = _bake(p, int(x), int(y), 96, 0.5 * 60.0 / 128)
        walked[i] = _walk(p, field, samples, header) @ LUMA
        assert bool(header.falls_in) == bool(pix[i].falls_in)
    diff = np.abs(walked - march).sum() / march.sum()
    print(f"walk vs march: {diff:.3%}")
    assert march.sum() > 0 and diff < 0.05
