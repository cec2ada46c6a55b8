"""HIP graph capture of a disk frame: one sim dispatch, the stochastic splat
with its frame taken from the dispatch count, the resolve, and a march of a
fixed camera, recorded once and replayed. Every replay must equal the same
launches run eagerly, bit for bit, with eager work and synchronizations
between the replays."""
from __future__ import annotations

import numpy as np
import pytest

from gargantua_amd import api

pytestmark = pytest.mark.gpu


def _disk(ctx, n):
    p = api.Particles(ctx, n, seed=7)
    p.seed(api.GENERATOR_ANNULUS, sigma_z=0.3)
    p.step(20)
    v = api.Volume(ctx, 64, 30.0, 3.0)
    v.clear()
    return p, v


def test_captured_disk_frame_replays_like_eager_launches(ctx):
    n = 200_000
    eager_p, eager_v = _disk(ctx, n)
    graph_p, graph_v = _disk(ctx, n)
    cam = api.orbit_camera()
    u = api.uniforms(cam, 96, 54, api.settings("interactive"), spin=0.6, star_seed=3, volume_half_extent=30.0)
    eager_t, graph_t, other_t = (api.Targets(ctx, 96, 54) for _ in range(3))
    s = ctx.stream()
    ctx.sync()
    start = graph_p.dispatches()
    assert start == 20

    ctx.graph_begin(s)
    try:
        graph_p.step(1, stream=s)
        graph_v.splat(graph_p, exact=False, frame=api.FRAME_AUTO, stream=s)
        graph_v.resolve(clear=True, stream=s)
        ctx.march_image(u, graph_t, graph_v, None, api.INTEGRATOR_FAST, stream=s)
    finally:
        g = ctx.graph_end(s)
    assert graph_p.dispatches() == start  # capturing ran nothing

    for k in range(1, 5):
        eager_p.step(1)
        eager_v.splat(eager_p, exact=False, frame=start + k)
        eager_v.resolve(clear=True)
        ctx.march_image(u, eager_t, eager_v, None, api.INTEGRATOR_FAST)
        g.launch(s)
        s.sync()
        # Unrelated eager work between replays.
        ctx.march_image(u, other_t, eager_v, None, api.INTEGRATOR_REFERENCE)
        ctx.sync()
        assert graph_p.dispatches() == start + k
        assert graph_p.checksum() == eager_p.checksum()
        ge, gv = graph_v.read_fields()
        ee, ev = eager_v.read_fields()
        assert np.array_equal(ge.view(np.uint32), ee.view(np.uint32))
        assert np.array_equal(gv.view(np.uint32), ev.view(np.uint32))
        gc, gd, _ = graph_t.read()
        ec, ed, _ = eager_t.read()
        assert np.array_equal(gc.view(np.uint32), ec.view(np.uint32))
        assert np.array_equal(gd.view(np.uint32), ed.view(np.uint32))
    print(f"graph replays equal eager launches over 4 frames ({n} particles, 64^3, 96x54 march)")
    g.close()
    s.close()
    for o in (eager_t, graph_t, other_t, eager_v, graph_v, eager_p, graph_p):
        o.close()


def test_host_synchronous_calls_refuse_while_capturing(ctx):
    """A call that waits for the device or copies to the host refuses with
    GARG_E_STATE when its stream (or the context's stream) is capturing;
    work on other streams is unaffected."""
    p, v = _disk(ctx, 10_000)
    s = ctx.stream()
    ctx.sync()
    ctx.graph_begin(s)
    try:
        p.step(1, stream=s)
        with pytest.raises(api.GargError) as e:
            ctx.sync(s)
        assert e.value.status == api.GARG_E_STATE
        before = p.checksum()  # on the context's stream, which is not capturing
    finally:
        g = ctx.graph_end(s)
    ctx.graph_begin()  # now the context's own stream captures
    try:
        p.step(1)
        with pytest.raises(api.GargError) as e:
            p.checksum()
        assert e.value.status == api.GARG_E_STATE
    finally:
        g2 = ctx.graph_end()
    assert p.dispatches() == 20 and p.checksum() == before
    g.launch(s)
    s.sync()
    g2.launch()
    ctx.sync()
    assert p.dispatches() == 22
    for o in (g, g2, s, v, p):
        o.close()
