"""Pins of the oracle's march restatement gather (orc_march_pixel's Gather::add
and its Schwarzschild and Kerr loops, oracle/render.cpp), which judges every
per-pixel colour test of the GPU march. The gather restates
march.metal:81-102 (Schwarzschild) and :151-170 (Kerr). Three independent
checks:

* at spin 0 with no sphere jump and no stars the restatement is the
  LensedRender.trace restatement (LensedRender.swift:46-78), which the
  reference's own LensedRender properties pin (mirror symmetry, beaming), on
  the reference's SyntheticDisk fixture (LensedRenderTests.swift:7-53);
* the Kerr loop as the spin goes to zero is the Schwarzschild loop, pixel by
  pixel, with gas and gravitational redshift, under one shared step length;
  the direction each loop takes for the Doppler factor is pinned against
  the metric on its own, since the reference's loops use different ones;
* closed forms of radiative transfer through a uniform slab far from the
  hole (f -> 1, straight rays): Beer-Lambert, C = e (1 - exp(-k L)) / k with
  k = opacity x density; the transmittance floor that stops the gather
  (C = e (1 - T_stop) / k); the luma-weighted depth
  1/k - s* exp(-k s*) / (1 - exp(-k s*)); the relativistic Doppler factor
  D = sqrt((1 + beta) / (1 - beta)) of gas approaching along the line of
  sight (1/D receding), cubed; and the density blend of that factor below
  the beaming floor. A dropped transmittance, a wrong sign of the opacity
  exponent or of the Doppler dot product, a transposed blend or a depth
  weighted by something other than the contribution's luma fails one of them.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from tests.test_oracle_geodesic import synthetic_camera, synthetic_disk

OPACITY = 5e-4       # VOLUME_OPACITY
FLOOR = 0.004        # TRANSMITTANCE_FLOOR
R_ESCAPE = 60.0      # depth is reported as min(depth / (2 R_ESCAPE), 1)
LUMA = np.array([0.2126, 0.7152, 0.0722])


def _settings(h, escape, capture=2.02, cap=10_000_000):
    """One constant step length h everywhere: in and off the slab for the
    Schwarzschild loop, and for the Kerr loop wherever 0.2 (r - r+) and the
    pole factor do not cut it."""
    s = O.Settings()
    s.step_factor, s.step_min, s.step_max = 1.0, h, h
    s.capture_radius, s.escape_radius, s.step_cap = capture, escape, cap
    s.empty_step_factor, s.empty_step_max = 1.0, h
    return s


def _params(position, forward, right, up, tan_half, width, height, settings, spin, half, half_z, redshift):
    p = O.MarchParams()
    p.camera_position[:] = list(position)
    p.camera_forward[:] = list(forward)
    p.camera_right[:] = list(right)
    p.camera_up[:] = list(up)
    p.tan_half_fov[:] = list(tan_half)
    p.width, p.height = width, height
    p.volume_half_extent, p.volume_half_extent_z = half, half_z
    p.opacity_scale, p.star_brightness, p.drift_budget = OPACITY, 0.0, 1.0
    p.redshift = int(redshift)
    p.spin = spin
    p.sphere_radius = 0.0
    p.geodesic = settings
    p._bb = np.ascontiguousarray(O.blackbody_table())
    p.blackbody = O.ptr(p._bb)
    return p


def _pixels(p, field, w, h):
    xs, ys = np.meshgrid(np.arange(w), np.arange(h))
    xy = np.ascontiguousarray(np.stack([xs.ravel(), ys.ravel()], 1), dtype=np.int32)
    out = (O.Pixel * len(xy))()
    O.lib.orc_march_pixels(C.byref(p), C.byref(field), xy.ctypes.data_as(C.POINTER(C.c_int32)), len(xy), 8, out)
    return out


# ---- (1) the restatement is LensedRender.trace at spin 0 --------------------


@pytest.mark.parametrize("moving", [False, True], ids=["static", "moving"])
def test_restatement_equals_lensed_render_at_spin_zero(moving):
    w, h = 64, 36
    e, v = synthetic_disk(60, moving)
    field = O.make_field(e, v)
    cam = synthetic_camera()
    settings = O.settings("still")
    img = np.zeros((h, w, 3))
    O.lib.orc_render(w, h, C.byref(cam), C.byref(field), C.byref(settings), OPACITY, 1, 8, O.ptr(img))
    p = _params(cam.position, cam.forward, cam.right, cam.up, cam.tan_half_fov, w, h, settings, 0.0, 30.0, 30.0, True)
    pix = _pixels(p, field, w, h)
    got = np.array([px.color[:] for px in pix]).reshape(h, w, 3)
    assert img.sum() > 0
    # The same FP64 operations but for the blend's algebraic form:
    # gravity^3 (1 - bulk) + g^3 bulk against gravity^3 + (g^3 - gravity^3) bulk.
    np.testing.assert_allclose(got, img, rtol=1e-9, atol=1e-12 * img.max())


# ---- (2) Kerr -> Schwarzschild as the spin goes to zero ---------------------


@pytest.mark.parametrize("moving,redshift", [(False, True), (True, False)], ids=["static-redshift", "moving-no-redshift"])
def test_kerr_loop_reduces_to_the_schwarzschild_loop_with_gas(moving, redshift):
    """Geometry, gravitational redshift, opacity, transmittance and depth of
    the Kerr loop at a = 1e-6 against the Schwarzschild loop, per pixel. With
    moving gas and redshift the two loops differ by design: the reference's
    Doppler direction is the static observer's in Schwarzschild and the
    coordinate velocity's in Kerr (pinned separately below)."""
    w, h = 32, 18
    e, v = synthetic_disk(48, moving)
    field = O.make_field(e, v)
    cam = synthetic_camera()
    # Kerr captures at r+ + 0.05, which is 2.05 as the spin goes to zero.
    s = _settings(0.02, 60.0, capture=2.05, cap=200_000)
    out = {}
    for spin in (0.0, 1e-6):
        p = _params(cam.position, cam.forward, cam.right, cam.up, cam.tan_half_fov, w, h, s, spin, 30.0, 30.0, redshift)
        out[spin] = _pixels(p, field, w, h)
    schw, kerr = out[0.0], out[1e-6]
    cs = np.array([px.color[:] for px in schw])
    ck = np.array([px.color[:] for px in kerr])
    ls, lk = cs @ LUMA, ck @ LUMA
    assert ls.sum() > 0 and (ls > 0).sum() > w * h // 10
    assert [px.outcome for px in schw] == [px.outcome for px in kerr]
    steps = np.array([px.steps for px in schw]) - np.array([px.steps for px in kerr])
    assert np.abs(steps).max() <= 1
    rel = np.abs(ls - lk).sum() / ls.sum()
    worst = np.abs(ls - lk).max() / ls.max()
    print(f"Kerr(a=1e-6) vs Schwarzschild gather: total {rel:.2e}, worst pixel {worst:.2e}")
    assert rel < 1e-6 and worst < 1e-5
    np.testing.assert_allclose([px.transmittance for px in kerr], [px.transmittance for px in schw], atol=1e-6)
    # Depth is the distance travelled where a ray saw no gas, and the Kerr
    # loop may capture one step of 0.02 later (r+ sits a hair below 2).
    ds, dk = np.array([px.depth for px in schw]), np.array([px.depth for px in kerr])
    np.testing.assert_allclose(dk[steps == 0], ds[steps == 0], atol=1e-6)
    np.testing.assert_allclose(dk, ds, atol=0.02 / (2.0 * R_ESCAPE) + 1e-6)


def test_doppler_direction_conventions_of_the_two_loops():
    """The direction each loop hands the Doppler factor, from the metric.
    Schwarzschild (Geodesic.h:94-100): the unit vector of the static
    observer, whose proper radial component is rdot / sqrt(f) (g_rr = 1/f in
    Schwarzschild coordinates; the same after projection in Eddington-
    Finkelstein ones). Kerr (Kerr.h:158-173): the unit coordinate velocity,
    which at a = 0 is rdot = f p_r (p_r = g_rr rdot), thetadot = p_theta / r^2,
    phidot = L / (r^2 sin^2 theta), scaled by r and r sin theta."""
    r, phi, rdot, phidot = 5.0, 0.3, 0.4, 0.05
    f = 1.0 - 2.0 / r
    ray = O.Ray()
    ray.s.r, ray.s.phi, ray.s.rdot, ray.s.phidot = r, phi, rdot, phidot
    ray.e1[:], ray.e2[:] = [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]
    radial = np.array([math.cos(phi), math.sin(phi), 0.0])
    tangential = np.array([-math.sin(phi), math.cos(phi), 0.0])
    want = rdot / math.sqrt(f) * radial + r * phidot * tangential
    np.testing.assert_allclose(O.schw_direction(ray), want / np.linalg.norm(want), atol=1e-14)

    s = O.KerrState()
    s.r, s.theta, s.phi, s.pr, s.ptheta, s.energy, s.angular_momentum = 5.0, 1.1, 0.3, 0.7, 0.9, 1.0, 2.5
    st, ct, sp, cp = math.sin(s.theta), math.cos(s.theta), math.sin(s.phi), math.cos(s.phi)
    r_hat = np.array([st * cp, st * sp, ct])
    theta_hat = np.array([ct * cp, ct * sp, -st])
    phi_hat = np.array([-sp, cp, 0.0])
    fk = 1.0 - 2.0 / s.r
    want = fk * s.pr * r_hat + s.r * (s.ptheta / s.r ** 2) * theta_hat + \
        s.r * st * (s.angular_momentum / (s.r ** 2 * st * st)) * phi_hat
    np.testing.assert_allclose(O.kerr_direction(0.0, s), want / np.linalg.norm(want), atol=1e-14)
    # The two conventions differ by the radial factor sqrt(f).
    assert abs(np.dot(O.kerr_direction(0.0, s), r_hat) - fk * s.pr / np.linalg.norm(want)) < 1e-14


# ---- (3) closed forms through a uniform slab far from the hole --------------

# The camera sits at (14000, 10000, 0) looking along +x, in the plane z = 0
# (rays launched in it stay in it, for Kerr too: p_theta = 0 on the
# equator). The ray leaves at the escape radius after L = 600 units of path,
# always inside the uniform part of an 8^3 field of half extent 20000 (texels
# of 5000; the trilinear fade reaches in 2500 from each face) and |z| <= 8.
# Far out f = 1 - 2/r differs from 1 by 1e-4 and the ray bends by 4/b ~ 4e-4
# rad over its whole path, so straight-ray radiative transfer holds to 1e-4.
X0, Y0, L = 14000.0, 10000.0, 600.0
ESCAPE = math.hypot(X0 + L, Y0)
HALF, HALF_Z, SIZE = 20000.0, 8.0, 8
EMISSION = np.array([0.3, 0.5, 0.2])


def _slab(density, velocity=(0.0, 0.0, 0.0), colours=None):
    """Uniform density and velocity; the emission colour is EMISSION, or
    colours[0] in the texel columns x <= 12500 and colours[1] beyond."""
    e = np.zeros((SIZE, SIZE, SIZE, 4))
    e[..., :3] = EMISSION
    if colours is not None:
        e[..., :SIZE - 1, :3] = colours[0]
        e[..., SIZE - 1, :3] = colours[1]
    e[..., 3] = density
    v = np.zeros((SIZE, SIZE, SIZE, 3))
    v[...] = velocity
    return O.make_field(e, v, HALF, HALF_Z)


def _slab_pixel(field, spin, step, redshift=False):
    p = _params((X0, Y0, 0.0), (1.0, 0.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), (0.1, 0.1), 1, 1,
                _settings(step, ESCAPE), spin, HALF, HALF_Z, redshift)
    return _pixels(p, field, 1, 1)[0]


SPINS = pytest.mark.parametrize("spin", [0.0, 0.5], ids=["schwarzschild", "kerr"])


@SPINS
def test_slab_beer_lambert(spin):
    """k L = 1: C = e (1 - e^-kL) / k, transmittance e^-kL; k h = 1.7e-4."""
    k = 1.0 / L
    px = _slab_pixel(_slab(k / OPACITY), spin, 0.1)
    assert px.outcome == 1
    want = EMISSION * (1.0 - math.exp(-k * L)) / k
    np.testing.assert_allclose(px.color[:], want, rtol=5e-4)
    assert abs(px.transmittance - math.exp(-k * L)) < 5e-4


@SPINS
def test_slab_optically_thin_limit(spin):
    """k L = 0.01: C = e L (1 - kL/2 + ...), within 1e-4."""
    k = 0.01 / L
    px = _slab_pixel(_slab(k / OPACITY), spin, 0.05)
    want = EMISSION * (1.0 - math.exp(-k * L)) / k
    np.testing.assert_allclose(px.color[:], want, rtol=3e-4)
    np.testing.assert_allclose(px.color[:], EMISSION * L * (1 - 0.005), rtol=5e-4)


@SPINS
def test_slab_stops_at_the_transmittance_floor(spin):
    """k L = 20: the gather stops once T <= 0.004, so C = e (1 - T_stop) / k
    with T_stop in (0.004 e^-kh, 0.004]; k h = 6.7e-4."""
    k = 20.0 / L
    step = 0.02
    px = _slab_pixel(_slab(k / OPACITY), spin, step)
    q = math.exp(-k * step)
    assert FLOOR * q * 0.999 < px.transmittance <= FLOOR
    want = EMISSION * (1.0 - px.transmittance) / k
    np.testing.assert_allclose(px.color[:], want, rtol=1e-3)
    # Without the floor the gather would reach 1/k.
    assert np.all(np.array(px.color[:]) < EMISSION / k * (1 - 0.5 * FLOOR))


def _incomplete(n, k, s):
    """Integral of x^n e^(-k x) over [0, s]."""
    tail = sum((k * s) ** m / math.factorial(m) for m in range(n + 1))
    return math.factorial(n) / k ** (n + 1) * (1.0 - math.exp(-k * s) * tail)


@SPINS
def test_slab_depth_is_the_luma_weighted_mean_distance(spin):
    """The colour ramps along the ray from green to blue (trilinear between
    the texel columns centred at x = 12500 and 17500), so the luma of the
    emission is linear in the distance s: lum(s) = a + b s. Contributions
    lum(s) T(s) ds weigh the distance travelled to their end, so
    depth = int s lum e^-ks / int lum e^-ks over [0, s*], s* where T reaches
    the floor, plus h/2 for the end of each step (k h^2 / 12 beyond);
    reported as depth / (2 R_ESCAPE)."""
    k, step = 0.02, 1.0
    green, blue = np.array([0.1, 1.0, 0.1]), np.array([0.1, 0.1, 1.0])
    px = _slab_pixel(_slab(k / OPACITY, colours=(green, blue)), spin, step)
    lum_a, lum_b = LUMA @ green, LUMA @ blue
    slope = (lum_b - lum_a) / 5000.0
    a = lum_a + slope * (X0 - 12500.0)
    s_star = math.log(1.0 / FLOOR) / k
    i0, i1, i2 = (_incomplete(n, k, s_star) for n in range(3))
    depth = (a * i1 + slope * i2) / (a * i0 + slope * i1) + 0.5 * step
    assert abs(px.depth - depth / (2.0 * R_ESCAPE)) < 1.5e-3 * depth / (2.0 * R_ESCAPE)
    # The ramp moves the depth by 0.8% from the uniform colour's.
    uniform = i1 / i0 + 0.5 * step
    assert abs(depth - uniform) > 6e-3 * depth


def _speed_for_beta(beta):
    """|v| for which 0.85 tanh(|v| / 0.85) = beta (GAS_SPEED_CEILING)."""
    return 0.85 * math.atanh(beta / 0.85)


@SPINS
@pytest.mark.parametrize("sign", [-1.0, 1.0], ids=["approaching", "receding"])
def test_slab_doppler_factor_cubed(spin, sign):
    """Dense gas (density >= 40: full Doppler) moving at beta = 0.5 along the
    line of sight: g^3 = f^(3/2) D^3 with D = sqrt(3) approaching (toward the
    camera, against the marched direction) and 1/sqrt(3) receding."""
    beta = 0.5
    k = 0.02
    px = _slab_pixel(_slab(k / OPACITY, (sign * _speed_for_beta(beta), 0.0, 0.0)), spin, 0.05, redshift=True)
    doppler = math.sqrt((1 + beta) / (1 - beta)) if sign < 0 else math.sqrt((1 - beta) / (1 + beta))
    f = 1.0 - 2.0 / math.hypot(X0, Y0)
    want = EMISSION * f ** 1.5 * doppler ** 3 * (1.0 - px.transmittance) / k
    np.testing.assert_allclose(px.color[:], want, rtol=1e-3)


@SPINS
def test_slab_doppler_blends_in_with_density(spin):
    """Below the beaming floor the factor is mix(f^(3/2), f^(3/2) D^3,
    density / 40): at density 10, f^(3/2) (0.75 + 0.25 D^3)."""
    beta = 0.5
    k = 10.0 * OPACITY
    px = _slab_pixel(_slab(10.0, (-_speed_for_beta(beta), 0.0, 0.0)), spin, 0.2, redshift=True)
    d3 = ((1 + beta) / (1 - beta)) ** 1.5
    f = 1.0 - 2.0 / math.hypot(X0, Y0)
    want = EMISSION * f ** 1.5 * (0.75 + 0.25 * d3) * (1.0 - px.transmittance) / k
    np.testing.assert_allclose(px.color[:], want, rtol=1e-3)
