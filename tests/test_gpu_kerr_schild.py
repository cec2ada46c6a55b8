"""GPU parity of the forms regular on the spin axis: the Kerr-Schild form
(GARG_INTEGRATOR_KERR_SCHILD) and the polar form (GARG_INTEGRATOR_POLAR:
Boyer-Lindquist fast steps until the ray first comes within sin(theta) <
0.25 of the axis or plunges into the horizon, Kerr-Schild steps from
there), against the FP64
Kerr-Schild oracle (oracle/kerr_schild.cpp), which is pinned to the
Boyer-Lindquist metric and integrator of KerrGeodesic.swift: probes over the
metric's view including the column over the pole, the drift budget on the
whole image, and spin zero falling back to the Schwarzschild fast form."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from gargantua_amd import api
from tests import scenes
from tests.test_gpu_march import same_bits

pytestmark = pytest.mark.gpu

SPIN = 0.9
REGULAR = pytest.mark.parametrize("form", [api.INTEGRATOR_KERR_SCHILD, api.INTEGRATOR_POLAR],
                                  ids=["kerr-schild", "polar"])


def view(cols, rows, kind="still", cap=4096, spin=SPIN):
    cam = scenes.look_at(50.0, 60.0, 60.0)
    g = api.settings(kind, step_cap=cap, escape_radius=60.0)
    u = api.uniforms(cam, 1, 1, g, drift_budget=1.0, redshift=False, stars=False, spin=spin,
                     compensated=kind == "still")
    return cam, g, u, scenes.camera_grid(cam, cols, rows, 16.0 / 9.0)


def oracle_ks(origin, direction, g, spin=SPIN):
    s = O.kerr_settings(int(g.step_cap))
    s.step_factor, s.step_max, s.escape_radius = float(g.step_factor), float(g.step_max), float(g.escape_radius)
    ks = O.KsState()
    O.lib.orc_ks_from_bl(spin, C.byref(O.kerr_launch(spin, origin, direction)), C.byref(ks))
    res = O.KerrResult()
    O.lib.orc_ks_integrate(spin, C.byref(ks), C.byref(s), C.byref(res))
    d = (C.c_double * 3)()
    O.lib.orc_kerr_direction(spin, C.byref(res.state), d)
    return res, np.array(d[:])


def angle(x, y):
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    return math.atan2(np.linalg.norm(np.cross(x, y)), float(x @ y))


@REGULAR
@pytest.mark.parametrize("spin,kind,budget", [(0.9, "still", 1e-5), (0.5, "still", 1e-5), (0.99, "still", 1e-5),
                                               (0.9, "interactive", 1e-4)])
def test_kerr_schild_probes_match_the_fp64_oracle(ctx, spin, kind, budget, form):
    """A 16 x 9 grid of the metric's view (r = 50, inclination 60 deg) plus
    the frame's centre column, which passes over the pole, at spins 0.5,
    0.9 and 0.99 (horizons 1.87, 1.44, 1.14) with the still policy and at
    0.9 with the 256-step interactive policy: the same fate as the FP64
    Kerr-Schild oracle (all but rays within float reach of the shadow
    edge), step counts within 1%, escape directions within 2e-5 rad (FP32
    against FP64 at the same steps; compensated for stills), and every
    escaping ray inside the drift budget of its policy (1e-5 still, 1e-4
    interactive). The polar form integrates the Boyer-Lindquist equations
    with the same steps away from the axis, where both converge to the same
    geodesics, and is held to the same bar."""
    cam, g, u, grid = view(16, 9, kind, 4096 if kind == "still" else 256, spin)
    column = [p for p in scenes.camera_grid(cam, 33, 33, 16.0 / 9.0)[16::33]]
    launches = grid + column
    res = ctx.probe(launches, u, integrator=form)
    worst_angle, worst_drift, mismatches, compared = 0.0, 0.0, 0, 0
    for (o, d), r in zip(launches, res):
        ref, dref = oracle_ks(o, d, g, spin)
        if r.outcome != ref.outcome:
            mismatches += 1
            continue
        if ref.outcome == api.RAY_ESCAPED:
            assert abs(r.steps - ref.steps) <= max(2, 0.01 * ref.steps)
            worst_angle = max(worst_angle, angle(r.direction[:], dref))
            worst_drift = max(worst_drift, r.drift)
            compared += 1
    print(f"form {form} probes a={spin} {kind}: {compared} escaped compared, {mismatches} fate mismatches, "
          f"max angle {worst_angle:.2e} rad, max drift {worst_drift:.2e}")
    assert compared > 0.5 * len(launches) and mismatches <= 2
    assert worst_angle < 2e-5 and worst_drift < budget


@REGULAR
def test_kerr_schild_fixes_the_polar_column(ctx, form):
    """Rays of the centre column pass near the pole, where the
    Boyer-Lindquist forms clamp 1/sin^2 theta and shrink their steps: on
    them the Kerr-Schild form lands closer to the converged FP64 direction
    (the Kerr-Schild oracle at steps twenty times finer) than the
    Boyer-Lindquist fast form on every ray the latter misses by more than
    1e-5 rad, and keeps its drift inside the budget on all of them."""
    cam, g, u, _ = view(1, 1)
    column = scenes.camera_grid(cam, 1, 41, 16.0 / 9.0)
    ks = ctx.probe(column, u, integrator=form)
    bl = ctx.probe(column, u, integrator=api.INTEGRATOR_FAST)
    fine = api.settings("still", step_cap=400_000, escape_radius=60.0, step_factor=0.001)
    worse, compared = 0, 0
    for (o, d), a, b in zip(column, ks, bl):
        ref, dref = oracle_ks(o, d, fine)
        if ref.outcome != api.RAY_ESCAPED or a.outcome != api.RAY_ESCAPED or b.outcome != api.RAY_ESCAPED:
            continue
        compared += 1
        ek, eb = angle(a.direction[:], dref), angle(b.direction[:], dref)
        assert a.drift < 1e-5
        if eb > 1e-5:
            worse += 1
            assert ek < eb, (ek, eb)
    print(f"polar column: {compared} escaping rays, Boyer-Lindquist off by more than 1e-5 rad on {worse}")
    assert compared > 20


def test_kerr_schild_image_keeps_the_drift_budget(ctx):
    """The metric's view at 480 x 270 with the 4096-step still budget: the
    Boyer-Lindquist forms take a few percent of the escaping rays over the
    1e-5 drift budget (near-polar rays, in FP64 too); the Kerr-Schild form,
    whose metric is regular on the axis, and the polar form, which switches
    to it near the axis, keep at most 0.1% over it."""
    w, h = 480, 270
    cam = scenes.look_at(50.0, 60.0, 60.0)
    u = api.uniforms(cam, w, h, api.settings("still", escape_radius=60.0), spin=SPIN, compensated=True,
                     star_seed=1, drift_budget=1e-5)
    shares = {}
    for form in (api.INTEGRATOR_FAST, api.INTEGRATOR_KERR_SCHILD, api.INTEGRATOR_POLAR):
        t = api.Targets(ctx, w, h)
        t.reset_counters()
        ctx.march_image(u, t, None, None, form)
        _, debug, _ = t.read()
        t.close()
        kept = debug[..., 2] != api.RAY_CAPTURED
        shares[form] = float((debug[..., 0][kept] > 1e-5).mean())
    print(f"over the drift budget: Boyer-Lindquist fast {shares[api.INTEGRATOR_FAST]:.2%}, "
          f"Kerr-Schild {shares[api.INTEGRATOR_KERR_SCHILD]:.2%}, polar {shares[api.INTEGRATOR_POLAR]:.2%}")
    for form in (api.INTEGRATOR_KERR_SCHILD, api.INTEGRATOR_POLAR):
        assert shares[form] <= 0.001
        assert shares[api.INTEGRATOR_FAST] > 10 * max(shares[form], 1e-4)


@REGULAR
def test_kerr_schild_image_agrees_with_boyer_lindquist_off_the_axis(ctx, form):
    """With the gas: where the two coordinate systems both integrate the
    same geodesics accurately (rays that stay away from the pole), the
    pictures agree. Compared outside the middle quarter of the columns: the
    centre column lies in the plane of the spin axis, and the farther a
    pixel is from it, the larger the ray's axial angular momentum and the
    farther its turning point in theta stays from the pole. Fates equal on
    99.5% of those pixels and the colour of the matching ones within 1%
    (relative to the frame's brightest pixel) on 99% of them."""
    from tests.test_gpu_march import disk_volume
    w, h = 320, 180
    vol = disk_volume(ctx)
    cam = scenes.look_at(50.0, 60.0, 60.0)
    u = api.uniforms(cam, w, h, api.settings("still", escape_radius=60.0), spin=SPIN, compensated=True,
                     star_seed=1)
    out = {}
    for f in (api.INTEGRATOR_FAST, form):
        t = api.Targets(ctx, w, h)
        ctx.march_image(u, t, vol, None, f)
        out[f] = t.read()
        t.close()
    (ca, da, _), (cb, db, _) = out[api.INTEGRATOR_FAST], out[form]
    side = np.abs(np.arange(w) + 0.5 - 0.5 * w) > w / 8
    fate = da[:, side, 2] == db[:, side, 2]
    peak = float(np.abs(ca[..., :3]).max())
    diff = np.abs(ca[:, side, :3].astype(np.float64) - cb[:, side, :3]).max(-1) / peak
    close = float((diff[fate] < 0.01).mean())
    print(f"off the axis plane: fates equal on {fate.mean():.2%}, colour within 1% on {close:.2%}, "
          f"median difference {np.median(diff[fate]):.2e}")
    assert fate.mean() >= 0.995 and close >= 0.99


@REGULAR
def test_kerr_schild_at_zero_spin_is_the_schwarzschild_fast_form(ctx, form):
    from tests.test_gpu_march import disk_volume
    vol = disk_volume(ctx)
    cam = api.orbit_camera()
    u = api.uniforms(cam, 96, 64, api.settings("still"), compensated=True, star_seed=1)
    images = []
    for f in (api.INTEGRATOR_FAST, form):
        t = api.Targets(ctx, 96, 64)
        ctx.march_image(u, t, vol, None, f)
        images.append(t.read())
        t.close()
    assert same_bits(images[0], images[1])


@REGULAR
@pytest.mark.parametrize("threshold", [0, 40])
def test_batched_kerr_schild_matches_plain_bitwise(ctx, threshold, form):
    """The batched persistent march with garg_batch_config.kerr_schild runs
    the same per-ray state machine as the plain Kerr-Schild launch: every
    target bit-identical, for whole-tile claims and for lane refill, alone
    and with a second request of another spin, size and budget in the mix."""
    from tests.test_gpu_march import disk_volume, halves, kerr_request, render_plain
    vol = disk_volume(ctx)
    a = kerr_request(ctx, 256, 144, 0.9, 60.0, 50.0, 4096, vol)
    b = kerr_request(ctx, 120, 72, 0.5, 80.0, 40.0, 256, vol, compensated=False)
    plain_a = render_plain(ctx, a, form)[0]
    plain_b = render_plain(ctx, b, form)[0]
    cfg = ctx.batch_config(compaction_period=8, refill_threshold=threshold,
                           kerr_schild=1 if form == api.INTEGRATOR_KERR_SCHILD else 2)
    ctx.march_batch(halves(a), cfg)
    assert same_bits(plain_a, a["targets"].read())
    ctx.march_batch([b, a], cfg)
    assert same_bits(plain_a, a["targets"].read()) and same_bits(plain_b, b["targets"].read())
    for r in (a, b):
        r["targets"].close()


def test_spinning_rays_beyond_the_radius_bound_are_refused(ctx):
    """The fast Kerr forms take one reciprocal of products growing like r^6;
    spinning rays are therefore marched within r = 1e5, and the entry points
    refuse an escape radius, camera or probe origin beyond it with an
    argument error rather than march garbage."""
    import math
    cam, g, u, grid = view(2, 2)
    far = api.settings("still", escape_radius=2.0e5)
    uf = api.uniforms(cam, 1, 1, far, drift_budget=1.0, redshift=False, stars=False, spin=SPIN, compensated=True)
    with pytest.raises(api.GargError, match="within r = 100000"):
        ctx.probe(grid, uf, integrator=api.INTEGRATOR_FAST)
    origin, direction = grid[0]
    with pytest.raises(api.GargError, match="probe 0 starts at r"):
        ctx.probe([(np.asarray(origin) * 5e3, direction)], u, integrator=api.INTEGRATOR_KERR_SCHILD)
    t = api.Targets(ctx, 16, 16)
    um = api.uniforms(cam, 16, 16, far, spin=SPIN)
    with pytest.raises(api.GargError, match="within r = 100000"):
        ctx.march_image(um, t, None, None, api.INTEGRATOR_FAST)
    with pytest.raises(api.GargError, match="within r = 100000"):
        ctx.march_batch([{"uniforms": um, "targets": t}])
    # Spin 0 has no such bound (the deflection golden launches at r = 50000).
    u0 = api.uniforms(cam, 16, 16, far)
    ctx.march_image(u0, t, None, None, api.INTEGRATOR_FAST)
    assert math.isfinite(float(t.read()[1][..., 0].max()))
    t.close()


def pixel_ray(u, x, y):
    """The ray of pixel (x, y) as the march kernels form it: through the
    pixel centre, in the camera basis scaled by tan(fov/2)."""
    w, h = u.resolution
    nx = (2.0 * (x + 0.5) / w - 1.0) * u.tan_half_fov[0]
    ny = (1.0 - 2.0 * (y + 0.5) / h) * u.tan_half_fov[1]
    d = np.array(u.camera_forward[:]) + nx * np.array(u.camera_right[:]) + ny * np.array(u.camera_up[:])
    return np.array(u.camera_position[:], np.float64), d / np.linalg.norm(d)


def test_polar_form_captures_what_boyer_lindquist_lets_escape_near_the_horizon(ctx):
    """One request of the 64-request mix (960 x 540, a = 0.9837, inclination
    66.3 deg, r = 51.9, still budget): a few retrograde shadow-edge rays of
    the Boyer-Lindquist fast march fall fast enough (dr/dlambda < -5) for
    the near-horizon step 0.2 (r - r+) to put RK4 substeps onto Delta ~ 0;
    their FP32 state turns non-finite and leaves as "escaped". The polar
    form hands a ray inside r+ + 0.5 falling faster than dr/dlambda = -2 to
    the Kerr-Schild march (regular across the horizon): on the whole frame
    no ray it keeps has a non-finite drift and every escaping one keeps the
    1e-5 budget (the rays that graze the horizon and turn round keep their
    Boyer-Lindquist steps), and on the blown-up pixels its fates are the
    Kerr-Schild form's and, on all but one, the FP64 oracle's."""
    from gargantua_amd import workloads
    u = workloads.view(960, 540, 0.9837, 66.3, 51.89, 4096)
    debug = {}
    for form in (api.INTEGRATOR_FAST, api.INTEGRATOR_KERR_SCHILD, api.INTEGRATOR_POLAR):
        t = api.Targets(ctx, 960, 540)
        ctx.march_image(u, t, None, None, form)
        debug[form] = t.read()[1].astype(np.float64)
        t.close()
    fast = debug[api.INTEGRATOR_FAST]
    blown = (fast[..., 2] != api.RAY_CAPTURED) & ~np.isfinite(fast[..., 0])
    ys, xs = np.nonzero(blown)
    polar = debug[api.INTEGRATOR_POLAR]
    kept = polar[..., 2] != api.RAY_CAPTURED
    escaped = polar[..., 0][polar[..., 2] == api.RAY_ESCAPED]
    print(f"fast form: {len(xs)} kept rays with a non-finite drift; polar form: "
          f"{int((kept & ~np.isfinite(polar[..., 0])).sum())}, max drift of its escaping rays {escaped.max():.3g}")
    assert len(xs) >= 5
    assert np.isfinite(polar[..., 0][kept]).all()
    assert escaped.max() < 1e-5
    ks = debug[api.INTEGRATOR_KERR_SCHILD]
    assert (polar[ys, xs, 2] == ks[ys, xs, 2]).all()
    g = api.settings("still", step_cap=4096, escape_radius=max(60.0, 51.89 + 10.0))
    agree = sum(int(oracle_ks(*pixel_ray(u, x, y), g, 0.9837)[0].outcome == polar[y, x, 2]) for x, y in zip(xs, ys))
    assert agree >= len(xs) - 1, (agree, len(xs))
