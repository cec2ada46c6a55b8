"""garg_probe_noalloc, the probe entry point that allocates nothing: its
results equal garg_probe's bit for bit on two calls in a row, the second
smaller than the first (so stale staging from the first call cannot leak into
the second), the context's allocation count does not move, and a call above
GARG_PROBE_NOALLOC_MAX is refused with a sentence before anything runs."""
import ctypes as C

import numpy as np
import pytest

from gargantua_amd import api
from tests import scenes

pytestmark = pytest.mark.gpu


def _fields(results):
    raw = np.frombuffer(bytes(results), dtype=np.uint32)
    return raw.reshape(len(results), C.sizeof(api.RayProbeResult) // 4)


def _rays(count, seed):
    rng = np.random.default_rng(seed)
    cam = scenes.look_at(50.0, 60.0, 60.0, azimuth=float(rng.uniform(0, 6.28)))
    cols = int(np.ceil(np.sqrt(count)))
    return scenes.camera_grid(cam, cols, cols, 1.0)[:count], cam


@pytest.mark.parametrize("spin,integrator", [(0.9, api.INTEGRATOR_FAST), (0.9, api.INTEGRATOR_KERR_SCHILD),
                                             (0.0, api.INTEGRATOR_REFERENCE)],
                         ids=["kerr-fast", "kerr-schild", "schwarzschild-reference"])
def test_noalloc_probes_equal_the_allocating_entry_point(ctx, spin, integrator):
    for count, seed in ((3000, 1), (700, 2)):
        rays, cam = _rays(count, seed)
        print(f"noalloc probes: spin {spin}, integrator {integrator}, {count} rays", flush=True)
        u = api.uniforms(cam, 1, 1, api.settings("still"), spin=spin, compensated=True, redshift=False, stars=False)
        before = ctx.allocated()
        got = ctx.probe(rays, u, integrator=integrator, noalloc=True)
        assert ctx.allocated() == before
        want = ctx.probe(rays, u, integrator=integrator)
        assert len(got) == count
        np.testing.assert_array_equal(_fields(got), _fields(want))
        outcomes = {r.outcome for r in got}
        assert api.RAY_ESCAPED in outcomes


def test_noalloc_refuses_more_than_its_array_holds(ctx):
    n = api.PROBE_NOALLOC_MAX + 1
    rays, cam = _rays(n, 3)
    u = api.uniforms(cam, 1, 1, api.settings("interactive"), spin=0.5, redshift=False, stars=False)
    with pytest.raises(api.GargError) as err:
        ctx.probe(rays, u, noalloc=True)
    assert err.value.status == api.GARG_E_ARGUMENT
    assert "at most" in str(err.value)
