"""The batched mixed workload at full size: 64 render requests drawn from a
seed (resolutions 256x144 to 3840x2160, spin U[0, 0.99], inclination
U[10, 89] degrees, observer r U[30, 100], budgets 256 or 4096) on one
1M-particle / 128^3 disk, marched in one garg_march_batch call. Every
request must equal its own one-launch-per-frame march (counters, and the
colour, debug and depth planes bit for bit), and pixels sampled from four
requests spread over spin and observer radius must match the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

from gargantua_amd import api, workloads
from tests import scenes

pytestmark = pytest.mark.gpu


def _raw(targets):
    n = targets.width * targets.height
    color = np.zeros(4 * n, np.uint16)
    debug = np.zeros(4 * n, np.uint16)
    depth = np.zeros(n, np.uint32)
    api._check(targets.ctx.lib.garg_targets_read(targets.h, api._ptr(color), api._ptr(debug), api._ptr(depth)))
    return color, debug, depth


@pytest.fixture(scope="module")
def mixed_disk(ctx):
    p, v = workloads.disk(ctx)
    yield v
    v.close()
    p.close()


def test_mixed_batch_of_64_requests_equals_each_request_alone(ctx, mixed_disk):
    reqs = workloads.mixed(ctx, mixed_disk)
    assert len(reqs) == 64
    sizes = {(r["targets"].width, r["targets"].height) for r in reqs}
    spins = [float(r["uniforms"].spin) for r in reqs]
    print(f"{len(reqs)} requests, sizes {sorted(sizes)}, spins {min(spins):.3f}..{max(spins):.3f}", flush=True)
    total = ctx.march_batch(reqs, ctx.batch_config())
    steps = 0
    for i, r in enumerate(reqs):
        u = r["uniforms"]
        alone = api.Targets(ctx, r["targets"].width, r["targets"].height)
        ctx.march_image(u, alone, mixed_disk, None, api.INTEGRATOR_FAST)
        a, b = r["targets"].counters(), alone.counters()
        assert (a.rays, a.steps, a.captured, a.exhausted, a.over_drift_budget) == \
            (b.rays, b.steps, b.captured, b.exhausted, b.over_drift_budget), i
        for x, y in zip(_raw(r["targets"]), _raw(alone)):
            assert np.array_equal(x, y), f"request {i} differs from its plain march"
        steps += a.steps
        alone.close()
    assert total.steps == steps and total.rays == sum(r["targets"].width * r["targets"].height for r in reqs)
    print(f"{total.rays} rays, {total.steps} steps, all 64 equal alone", flush=True)

    # Four requests at the extremes of spin and observer radius against FP64.
    radii = [float(np.linalg.norm(np.array(r["uniforms"].camera_position[:]))) for r in reqs]
    picks = sorted({int(np.argmin(spins)), int(np.argmax(spins)), int(np.argmin(radii)), int(np.argmax(radii))})
    field = scenes.field_from_volume(mixed_disk)
    rng = np.random.default_rng(7)
    for i in picks:
        r = reqs[i]
        u = r["uniforms"]
        w, h = int(u.resolution[0]), int(u.resolution[1])
        color, debug, _ = r["targets"].read()
        xy = np.stack([rng.integers(0, w, 50), rng.integers(0, h, 50)], 1)
        pix, _ = scenes.oracle_pixels(scenes.oracle_params(u), xy, field)
        outcome_ok = steps_ok = color_ok = 0
        for (x, y), o in zip(xy, pix):
            if int(round(debug[y, x, 2])) != o.outcome:
                continue
            outcome_ok += 1
            steps_ok += abs(debug[y, x, 1] - o.steps) <= max(2, 0.01 * o.steps)
            want = np.array(o.color[:])
            color_ok += np.all(np.abs(color[y, x, :3] - want) <= 2e-3 + 2e-2 * np.abs(want))
        print(f"request {i}: {w}x{h} spin {u.spin:.3f} r_obs {radii[i]:.1f} budget {u.geodesic.step_cap}: "
              f"outcome {outcome_ok}/50, steps {steps_ok}/50, colour {color_ok}/50", flush=True)
        assert outcome_ok >= 49 and steps_ok >= 47 and color_ok >= 46
    workloads.close(reqs)
