"""GPU parity through the probe kernel: the probe-based goldens of
Validation/Validate.swift (shadow, deflection, conservation_interactive,
conservation_still, parity, sky, shadow_kerr) for both integrator forms,
and the FP64 table and shadow-boundary kernels against the FP64 oracle."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O
from gargantua_amd import api
from tests import scenes
from tests.conftest import golden_passes, load_golden

pytestmark = pytest.mark.gpu

FORMS = [(api.INTEGRATOR_REFERENCE, False), (api.INTEGRATOR_FAST, False), (api.INTEGRATOR_FAST, True)]
FORM_IDS = ["reference", "fast", "fast-compensated"]


def probe_uniforms(geodesic, spin=0.0, compensated=False, tables=None, camera=None):
    u = api.uniforms(camera or api.orbit_camera(), 1, 1, geodesic, drift_budget=1.0, redshift=False, stars=False,
                     spin=spin, compensated=compensated)
    if tables is not None:
        tables.apply(u)
    return u


@pytest.mark.parametrize("form,comp", FORMS, ids=FORM_IDS)
def test_shadow_golden(ctx, form, comp):
    """Validate.swift:125-150: parallel bisection of the capture boundary."""
    g = load_golden("shadow")
    p = g["parameters"]
    low, high, n = p["bracketLow"], p["bracketHigh"], int(p["raysPerRound"])
    u = probe_uniforms(api.settings("still"), compensated=comp)
    measured = None
    for _ in range(int(p["rounds"])):
        values = [low + (high - low) * (i + 0.5) / n for i in range(n)]
        res = ctx.probe([scenes.aimed(p["launchRadius"], b) for b in values], u, integrator=form)
        outcomes = [r.outcome for r in res]
        if api.RAY_EXHAUSTED in outcomes:
            measured = values[outcomes.index(api.RAY_EXHAUSTED)]
            break
        first = outcomes.index(api.RAY_ESCAPED) if api.RAY_ESCAPED in outcomes else n
        low, high = (values[first - 1] if first > 0 else low), (values[first] if first < n else high)
    critical = measured if measured is not None else 0.5 * (low + high)
    print(f"shadow b = {critical:.5f}")
    assert golden_passes(g, critical)


@pytest.mark.parametrize("form,comp", FORMS, ids=FORM_IDS)
def test_deflection_golden(ctx, form, comp):
    g = load_golden("deflection")
    radius, b = g["parameters"]["launchRadius"], g["parameters"]["impactParameter"]
    s = api.settings("still", step_max=1e9, escape_radius=radius)
    origin, direction = scenes.aimed(radius, b)
    r = ctx.probe([(origin, direction)], probe_uniforms(s, compensated=comp), integrator=form)[0]
    final = np.array(r.direction[:], dtype=np.float64)
    angle = math.atan2(np.linalg.norm(np.cross(direction, final)), direction @ final)
    print(f"deflection {angle:.6f} rad, outcome {r.outcome}, steps {r.steps}")
    assert r.outcome == api.RAY_ESCAPED and golden_passes(g, angle)


@pytest.mark.parametrize("kind", ["interactive", "still"])
@pytest.mark.parametrize("form,comp", FORMS, ids=FORM_IDS)
def test_conservation_goldens(ctx, kind, form, comp):
    g = load_golden(f"conservation_{kind}")
    p = g["parameters"]
    launches = scenes.camera_grid(api.orbit_camera(), int(p["columns"]), int(p["rows"]), 960 / 540,
                                  as_float32=False)
    res = ctx.probe(launches, probe_uniforms(api.settings(kind), compensated=comp), integrator=form)
    worst = max(r.drift for r in res)
    exhausted = sum(r.outcome == api.RAY_EXHAUSTED for r in res)
    print(f"{kind}: max drift {worst:.2e}, {exhausted} of {len(res)} exhausted")
    assert golden_passes(g, worst)


@pytest.mark.parametrize("form,comp", FORMS, ids=FORM_IDS)
def test_parity_golden(ctx, form, comp):
    """GPU endpoints vs the FP64 oracle from identical float32 launch values."""
    g = load_golden("parity")
    p = g["parameters"]
    launches = scenes.camera_grid(api.orbit_camera(), int(p["columns"]), int(p["rows"]), 960 / 540)
    res = ctx.probe(launches, probe_uniforms(api.settings("still"), compensated=comp), integrator=form)
    s = O.settings("still")
    worst, mismatches = 0.0, 0
    for (o, d), r in zip(launches, res):
        ref = O.schw_integrate(O.schw_launch(o, d), s)
        expected = O.schw_position(ref.ray)
        actual = np.array(r.position[:], dtype=np.float64)
        worst = max(worst, np.linalg.norm(expected - actual) / max(np.linalg.norm(expected), 1.0))
        mismatches += ref.outcome != r.outcome
    print(f"parity: max relative endpoint error {worst:.2e}, {mismatches} outcome mismatches")
    assert golden_passes(g, worst) and mismatches == 0


@pytest.mark.parametrize("form,comp", FORMS, ids=FORM_IDS)
def test_sky_golden(ctx, form, comp):
    """Sweep-table sky directions vs the oracle integrated to r = 1e6."""
    g = load_golden("sky")
    p = g["parameters"]
    cam = api.orbit_camera()
    tables = api.Tables(ctx)
    tables.update(cam["distance"])
    launches = scenes.camera_grid(cam, int(p["columns"]), int(p["rows"]), 960 / 540)
    res = ctx.probe(launches, probe_uniforms(api.settings("still"), compensated=comp, tables=tables), tables,
                    integrator=form)
    s = O.settings("still")
    s.step_factor, s.step_max, s.escape_radius, s.step_cap = 0.01, 20.0, p["oracleRadius"], 400_000
    s.empty_step_factor, s.empty_step_max = 0.02, 20.0
    worst, mismatches, compared = 0.0, 0, 0
    for (o, d), r in zip(launches, res):
        ref = O.schw_integrate(O.schw_launch(o, d), s)
        gpu_escaped = r.outcome == api.RAY_ESCAPED
        if gpu_escaped != (ref.outcome == 1):
            mismatches += 1
            continue
        if gpu_escaped:
            e = O.schw_direction(ref.ray)
            a = np.array(r.direction[:], dtype=np.float64)
            worst = max(worst, math.acos(min(max(e @ a, -1.0), 1.0)))
            compared += 1
    tables.close()
    print(f"sky: max angle {worst:.2e} rad over {compared} rays, {mismatches} fate mismatches")
    assert golden_passes(g, worst) and mismatches == 0 and compared > 0


@pytest.mark.parametrize("form,comp", FORMS, ids=FORM_IDS)
def test_shadow_kerr_golden(ctx, form, comp):
    """Validate.swift:350-374: the spin 0.9 shadow asymmetry, bisected through
    the probe kernel, against the oracle's bisection."""
    g = load_golden("shadow_kerr")
    p = g["parameters"]
    spin, radius, iterations = p["spin"], p["launchRadius"], int(p["iterations"])
    u = probe_uniforms(api.settings("still"), spin=spin, compensated=comp)

    def gpu_xi(side):
        low, high = 0.0, 12.0 / radius
        for _ in range(iterations):
            aim = 0.5 * (low + high)
            r = ctx.probe([((radius, 0.0, 0.0), (-1.0, side * aim, 0.0))], u, integrator=form)[0]
            if r.outcome == api.RAY_CAPTURED:
                low = aim
            else:
                high = aim
        k = O.kerr_launch(spin, [radius, 0, 0], [-1.0, side * 0.5 * (low + high), 0.0])
        return k.angular_momentum / k.energy

    gpu = abs(gpu_xi(-1.0) / gpu_xi(1.0))
    s = O.kerr_settings(4096)
    pro = O.lib.orc_kerr_bisect_critical_xi(spin, radius, 1.0, C.byref(s), iterations)
    retro = O.lib.orc_kerr_bisect_critical_xi(spin, radius, -1.0, C.byref(s), iterations)
    oracle = abs(retro / pro)
    print(f"shadow_kerr: gpu {gpu:.4f} vs oracle {oracle:.4f}, ratio {gpu / oracle:.4f}")
    assert golden_passes(g, gpu / oracle)


def test_sweep_tables_match_oracle(ctx):
    tables = api.Tables(ctx)
    tables.update(52.0)
    sphere, camera, sky = tables.read()
    tables.close()
    ref = scenes.oracle_tables(np.float32(52.0).item())
    for got, want, name in zip((sphere, camera, sky), ref, ("sphere", "camera", "sky")):
        err = np.max(np.abs(got.astype(np.float64) - want))
        print(f"{name} table max error {err:.2e}")
        assert err < 1e-6, name


def test_far_field_table_fp64_matches_oracle(ctx):
    """4096 impact parameters log-spaced on [5.2, 1000], 2048 intervals."""
    b, alpha, sweep = ctx.deflection_table(50.0, 4096, 5.2, 1000.0, 2048)
    rb, ralpha, rsweep = O.far_field_table(50.0, 4096, 5.2, 1000.0, 2048)
    assert np.max(np.abs(b - rb) / rb) < 1e-14
    assert np.max(np.abs(alpha - ralpha)) < 1e-10
    assert np.max(np.abs(sweep - rsweep)) < 1e-10
    assert abs(alpha[0] - O.lib.orc_strong_deflection(5.2)) < 0.7 * (5.2 - 3 * math.sqrt(3))


def test_shadow_boundary_fp64_matches_oracle_and_bardeen(ctx):
    """GPU FP64 multi-section along 360 screen azimuths; at 90 degrees the
    azimuths 0 and 180 are the prograde and retrograde equatorial edges."""
    edge, xi = ctx.shadow_boundary(0.9, 52.0, 90.0, azimuths=360, rounds=6)
    s = O.kerr_settings(4096)
    s.escape_radius = 62.0
    pro = O.lib.orc_kerr_bisect_critical_xi(0.9, 52.0, 1.0, C.byref(s), 48)
    retro = O.lib.orc_kerr_bisect_critical_xi(0.9, 52.0, -1.0, C.byref(s), 48)
    print(f"xi(0) {xi[0]:.6f} vs {pro:.6f}; xi(180) {xi[180]:.6f} vs {retro:.6f}")
    assert abs(xi[0] - pro) / abs(pro) < 1e-5
    assert abs(xi[180] - retro) / abs(retro) < 1e-5
    assert abs(abs(xi[180] / xi[0]) - 2.4020) < 2e-3
    assert np.all(edge > 0) and np.all(np.isfinite(edge))
    # The shadow is symmetric about the equatorial plane at 90 degrees.
    assert np.max(np.abs(edge[1:180] - edge[359:180:-1])) < 1e-6


def sky_settings():
    """The sky golden's far integration (Validate.swift:293-305): step
    factor 0.01 up to 20, escape at oracleRadius, 400,000 steps, the coarse
    empty-leg factor left at the default step factor."""
    g = api.settings("still", step_cap=400_000, escape_radius=load_golden("sky")["parameters"]["oracleRadius"])
    g.step_factor, g.step_max, g.empty_step_max = 0.01, 20.0, 20.0
    g.empty_step_factor = 0.02
    return g


@pytest.mark.parametrize("kind", ["parity", "sky"])
def test_fp64_probes_match_the_cpu_oracle(ctx, kind):
    """garg_probe_fp64, the reference's FP64 Schwarzschild integrator on the
    GPU that `gargantua validate` judges the parity and sky goldens against,
    against the CPU oracle's restatement of the same integrator with the
    same (float-valued) settings, on those goldens' 8 x 8 camera grids:
    equal outcomes and step counts, end points and end directions within
    1e-9 (the two differ only in the rounding of sqrt, sin, cos and the
    rate expressions), and drifts within 1e-9 of each other."""
    g = api.settings("still") if kind == "parity" else sky_settings()
    p = load_golden(kind)["parameters"]
    launches = scenes.camera_grid(api.orbit_camera(), int(p["columns"]), int(p["rows"]), 960 / 540)
    res = ctx.probe_fp64(launches, g)
    s = O.settings("still")
    for name in ("step_factor", "step_min", "step_max", "capture_radius", "escape_radius", "empty_step_factor",
                 "empty_step_max"):
        setattr(s, name, float(getattr(g, name)))
    s.step_cap = int(g.step_cap)
    worst_p = worst_d = worst_drift = 0.0
    escaped = 0
    for (o, d), r in zip(launches, res):
        o32 = np.array(o, np.float32).astype(np.float64)
        d32 = np.array(d, np.float32).astype(np.float64)
        ref = O.schw_integrate(O.schw_launch(o32, d32), s)
        assert r.outcome == ref.outcome and r.steps == ref.steps, (r.outcome, ref.outcome, r.steps, ref.steps)
        pos = O.schw_position(ref.ray)
        worst_p = max(worst_p, np.linalg.norm(np.array(r.position[:]) - pos) / max(np.linalg.norm(pos), 1.0))
        worst_d = max(worst_d, np.linalg.norm(np.array(r.direction[:]) - O.schw_direction(ref.ray)))
        worst_drift = max(worst_drift, abs(r.drift - ref.drift))
        escaped += r.outcome == api.RAY_ESCAPED
    print(f"{kind}: FP64 GPU vs CPU oracle: end point {worst_p:.2e}, direction {worst_d:.2e}, drift {worst_drift:.2e}, "
          f"{escaped} of {len(launches)} escaped")
    assert escaped > 10 and worst_p < 1e-9 and worst_d < 1e-9 and worst_drift < 1e-9


def test_fp64_probe_refuses_bad_launches(ctx):
    g = api.settings("still")
    with pytest.raises(api.GargError, match="inside the capture radius"):
        ctx.probe_fp64([((1.0, 0.0, 0.0), (0.0, 1.0, 0.0))], g)
    with pytest.raises(api.GargError, match="has no direction"):
        ctx.probe_fp64([((50.0, 0.0, 0.0), (0.0, 0.0, 0.0))], g)
