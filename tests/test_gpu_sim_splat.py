"""GPU parity of the sim and splat passes: the determinism and isco goldens,
one sim dispatch bit-exact against a CPU replay of the same float
operations, the integer deposit bit-exact against a CPU replay (exact and
stochastic corners), and the resolve against a float64 restatement."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from gargantua_amd import api
from tests.conftest import golden_passes, load_golden

pytestmark = pytest.mark.gpu


def settled(ctx, count, seed, dispatches, generator=api.GENERATOR_DISK):
    p = api.Particles(ctx, count, seed)
    p.seed(generator)
    if dispatches:
        p.step(dispatches)
    return p


def test_determinism_golden(ctx):
    g = load_golden("determinism")
    steps = int(g["parameters"]["steps"]) // 2
    sums = []
    for _ in range(2):
        p = settled(ctx, 250_000, 1, steps)
        sums.append(p.checksum())
        p.close()
    print(f"{sums[0]:016x} vs {sums[1]:016x}")
    assert golden_passes(g, float(sums[0] - sums[1]))


def test_checksum_is_fnv1a_of_the_records(ctx):
    p = settled(ctx, 10_000, 5, 3)
    rec = p.read()
    assert p.checksum() == O.lib.orc_fnv1a(rec.ctypes.data_as(C.c_void_p), rec.nbytes)
    p.close()


def test_isco_golden(ctx):
    """1M particles, seed 1, 1500 substeps, edge from Sigma r^1.5 (Validate.swift:102-110)."""
    g = load_golden("isco")
    p = settled(ctx, 1_000_000, 1, int(g["parameters"]["settlingSteps"]) // 2)
    rec = p.read()
    radii = np.ascontiguousarray(np.hypot(rec["position"][:, 0], rec["position"][:, 1]).astype(np.float64))
    q = g["parameters"]
    edge = O.lib.orc_inner_edge(O.ptr(radii), len(radii), q["binWidth"], q["fitInner"], q["fitOuter"])
    print(f"isco edge r = {edge:.3f}; respawns {int(rec['respawns'].sum())}")
    assert golden_passes(g, edge) and rec["respawns"].sum() > 0
    p.close()


@pytest.mark.parametrize("generator", [api.GENERATOR_DISK, api.GENERATOR_ANNULUS])
def test_sim_dispatch_bit_exact_against_cpu_replay(ctx, generator):
    p = settled(ctx, 200_000, 7, 40, generator)
    before = p.read()
    p.step(1)
    after = p.read()
    replay = before.copy()
    respawned = np.zeros(len(replay), np.uint8)
    O.lib.orc_sim_replay_step(replay.ctypes.data_as(C.c_void_p), len(replay), 1, 0.1, 0.003, _temperature_scale(),
                              respawned.ctypes.data_as(C.POINTER(C.c_uint8)))
    keep = respawned == 0
    assert keep.sum() > 0.9 * len(keep)
    # Positions and velocities round identically; respawned particles differ.
    assert np.array_equal(after["position"][keep].view(np.uint32), replay["position"][keep].view(np.uint32))
    assert np.array_equal(after["velocity"][keep].view(np.uint32), replay["velocity"][keep].view(np.uint32))
    assert np.array_equal(after["respawns"][keep], before["respawns"][keep])
    assert np.all(after["respawns"][~keep] == before["respawns"][~keep] + 1)
    # Temperature goes through pow (ocml vs libm): equal to a few ulp.
    rel = np.abs(after["temperature"][keep] - replay["temperature"][keep]) / replay["temperature"][keep]
    assert rel.max() < 4e-7
    p.close()


@pytest.mark.parametrize("generator", [api.GENERATOR_DISK, api.GENERATOR_ANNULUS])
def test_fused_dispatches_equal_one_launch_each(ctx, generator):
    """Consecutive dispatches run in one launch with the particle in
    registers (300 = a launch of 256 and one of 44): the store must equal
    one launch per dispatch bit for bit, respawns included, and the
    dispatch count must advance the same."""
    fused = settled(ctx, 150_000, 9, 0, generator)
    single = settled(ctx, 150_000, 9, 0, generator)
    fused.step(300)
    for _ in range(300):
        single.step(1)
    a, b = fused.read(), single.read()
    assert a.tobytes() == b.tobytes()
    assert fused.dispatches() == single.dispatches() == 300
    assert fused.checksum() == single.checksum()
    if generator == api.GENERATOR_DISK:
        # Its inner particles plunge, so respawns happen inside the fused
        # launches (the annulus settles without any in this span).
        assert int((a["respawns"] > 0).sum()) > 0
    fused.close()
    single.close()


def _temperature_scale():
    peak_r = 6.0 * 49.0 / 36.0
    return 8000.0 / O.lib.orc_disk_shape(peak_r, 0.01)


def test_seed_matches_oracle_spawn(ctx):
    """Seeded particles against the FP64 spawn of the same hashes."""
    p = settled(ctx, 4096, 11, 0)
    rec = p.read()
    cdf = O.inverse_cdf(1024)
    s = (C.c_double * 6)()
    worst = 0.0
    for i in range(0, 4096, 7):
        O.lib.orc_spawn(i, 0, 11, O.ptr(cdf), 1024, 1, 0, 0.0, s)
        got = np.concatenate([rec["position"][i], rec["velocity"][i]]).astype(np.float64)
        worst = max(worst, np.max(np.abs(got - np.array(s[:]))))
    assert worst < 2e-4
    p.close()


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("size,hz", [(96, 30.0), (128, 3.0), (256, 3.0), (12, 30.0)])
def test_splat_bit_exact_against_cpu_replay(ctx, exact, size, hz):
    """The exact deposit sorts particles by cell and deposits one voxel
    brick per workgroup; a 12^3 grid leaves partial 16 x 8 x 4 bricks on two
    axes."""
    p = settled(ctx, 250_000, 3, 100)
    v = api.Volume(ctx, size, 30.0, hz)
    v.splat(p, exact=exact, frame=17)
    e, m, a = v.read_sums()
    rec = p.read()
    re = np.zeros_like(e)
    rm = np.zeros_like(m)
    ra = np.zeros_like(a)
    O.lib.orc_splat_replay(rec.ctypes.data_as(C.c_void_p), len(rec), size, 30.0, hz, 8000.0, 0.0005,
                           1 if exact else 0, 17, 3, re.ctypes.data_as(C.POINTER(C.c_uint32)),
                           rm.ctypes.data_as(C.POINTER(C.c_int32)), ra.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert a.sum() > 0
    assert np.array_equal(e, re) and np.array_equal(m, rm) and np.array_equal(a, ra)
    v.close()
    p.close()


def _replay(rec, size, hz, exact, frame, seed):
    n = size ** 3
    re = np.zeros(2 * n, np.uint32)
    rm = np.zeros(3 * n, np.int32)
    ra = np.zeros(n, np.uint32)
    O.lib.orc_splat_replay(rec.ctypes.data_as(C.c_void_p), len(rec), size, 30.0, hz, 8000.0, 0.0005,
                           1 if exact else 0, frame, seed, re.ctypes.data_as(C.POINTER(C.c_uint32)),
                           rm.ctypes.data_as(C.POINTER(C.c_int32)), ra.ctypes.data_as(C.POINTER(C.c_uint32)))
    return re, rm, ra


@pytest.mark.parametrize("count,case", [(1, "inside"), (7, "inside"), (1000, "inside"), (1025, "inside"),
                                        (33_333, "inside"), (5000, "outside"), (5000, "edges")])
def test_exact_deposit_edge_cases_against_cpu_replay(ctx, count, case):
    """Exact deposits of one particle, counts that take the sort's
    single-block path (up to 1024) and counts that are not a multiple of a
    workgroup, every particle outside the cube (nothing deposited), and
    particles on the cube's faces and corners (clamped corners), each into
    a volume that already holds an earlier deposit: the sums must equal the
    CPU replay of both deposits."""
    size, hz = 64, 30.0
    rng = np.random.default_rng(count)
    p = settled(ctx, count, 13, 0)
    rec = p.read()
    if case == "outside":
        rec["position"] = rng.uniform(31.0, 40.0, (count, 3)).astype(np.float32) * rng.choice([-1, 1], (count, 3))
    elif case == "edges":
        # On and next to the faces: grid coordinates -0.5 .. 0 and n - 1 .. n - 0.5 clamp a corner.
        cell = 60.0 / size
        pick = rng.choice([-30.0, -30.0 + 0.25 * cell, 30.0 - 0.25 * cell, 30.0 - 1e-4, 0.0], (count, 3))
        rec["position"] = pick.astype(np.float32)
    else:
        rec["position"] = rng.uniform(-29.0, 29.0, (count, 3)).astype(np.float32)
    rec["velocity"] = rng.normal(0.0, 0.4, (count, 3)).astype(np.float32)
    rec["temperature"] = rng.uniform(2000.0, 9000.0, count).astype(np.float32)
    p.write(rec)
    first = settled(ctx, 20_000, 5, 10)
    v = api.Volume(ctx, size, 30.0, hz)
    v.splat(first, exact=True)
    v.splat(p, exact=True)
    e, m, a = v.read_sums()
    e0, m0, a0 = _replay(first.read(), size, hz, True, 0, 5)
    e1, m1, a1 = _replay(rec, size, hz, True, 0, 13)
    assert np.array_equal(e, e0 + e1) and np.array_equal(m, m0 + m1) and np.array_equal(a, a0 + a1)
    if case == "outside":
        assert np.array_equal(a1, np.zeros_like(a1))
    else:
        assert a1.sum() > 0
    for o in (v, p, first):
        o.close()


def test_resolve_against_float64_restatement(ctx):
    p = settled(ctx, 250_000, 3, 100)
    v = api.Volume(ctx, 96)
    v.splat(p, exact=True)
    e, m, a = v.read_sums()
    v.resolve(clear=True)
    emission, velocity = v.read_fields()
    e2, m2, a2 = v.read_sums()
    assert not e2.any() and not m2.any() and not a2.any()
    voxel = 60.0 / 96
    count = a.astype(np.float64) / 256.0
    energy = e[0::2].astype(np.float64) / 2 ** 24
    energy_t = e[1::2].astype(np.float64) / 2 ** 20
    density = count / voxel ** 3
    table = O.blackbody_table()
    with np.errstate(divide="ignore", invalid="ignore"):
        mean_t = np.where(energy > 0, energy_t / energy * 8000.0, 1000.0)
    coord = (np.log(np.clip(mean_t, 1000, 40000)) - np.log(1000)) / (np.log(40000) - np.log(1000))
    x = coord * 256 - 0.5
    i0 = np.clip(np.floor(x).astype(int), 0, 255)
    i1 = np.clip(np.floor(x).astype(int) + 1, 0, 255)
    f = (x - np.floor(x))[:, None]
    tint = table[i0] * (1 - f) + table[i1] * f
    want_e = tint * (energy / voxel ** 3)[:, None]
    got = emission.reshape(-1, 4)
    ok = density > 0
    assert np.allclose(got[:, 3], density, rtol=2e-3, atol=1e-6)
    assert np.allclose(got[ok, :3], want_e[ok], rtol=4e-3, atol=1e-6)
    mom = m.reshape(-1, 3).astype(np.float64) / 4096.0
    want_v = np.where(count[:, None] > 0, mom / np.maximum(count, 1e-30)[:, None], 0.0)
    assert np.allclose(velocity.reshape(-1, 4)[:, :3], want_v, rtol=2e-3, atol=2e-3)
    v.close()
    p.close()


def test_sorted_exact_deposit_equals_direct_deposit(ctx):
    """Inside a graph capture the exact deposit takes the direct path (no
    library sort in graphs); eagerly it sorts and deposits per brick. Both
    must give the same integer sums."""
    p = settled(ctx, 300_000, 5, 60)
    eager = api.Volume(ctx, 128, 30.0, 3.0)
    graphed = api.Volume(ctx, 128, 30.0, 3.0)
    s = ctx.stream()
    ctx.sync()
    ctx.graph_begin(s)
    try:
        graphed.splat(p, exact=True, stream=s)
    finally:
        g = ctx.graph_end(s)
    for _ in range(2):  # each replay adds the same deposit again
        g.launch(s)
        s.sync()
        eager.splat(p, exact=True)
        ctx.sync()
    # A second, smaller and denser particle set into the same volumes: the
    # cell tables still hold the first set's starts, which must be ignored.
    q = settled(ctx, 70_000, 6, 20, api.GENERATOR_ANNULUS)
    ctx.graph_begin(s)
    try:
        graphed.splat(q, exact=True, stream=s)
    finally:
        g2 = ctx.graph_end(s)
    g2.launch(s)
    s.sync()
    eager.splat(q, exact=True)
    ctx.sync()
    e1, m1, a1 = eager.read_sums()
    e2, m2, a2 = graphed.read_sums()
    assert a1.sum() > 0
    assert np.array_equal(e1, e2) and np.array_equal(m1, m2) and np.array_equal(a1, a2)
    for o in (g, g2, s, eager, graphed, p, q):
        o.close()



@pytest.mark.parametrize("size", [128, 96])
def test_every_lds_brick_shape_gives_the_same_sums(ctx, size):
    """The exact deposit's brick shape (the autotuner's LDS dimension) moves
    work between workgroups and their LDS boxes but not the integer sums:
    every shape deposits the direct (atomic, graph-captured) path's bits,
    also on a size that is not a multiple of the larger bricks, and after
    another shape's deposit into the same volume, whose brick stamps the
    new shape must ignore."""
    p = settled(ctx, 300_000, 7, 40)
    ref = api.Volume(ctx, size, 30.0, 3.0)
    s = ctx.stream()
    ctx.sync()
    ctx.graph_begin(s)
    try:
        ref.splat(p, exact=True, stream=s)
    finally:
        g = ctx.graph_end(s)
    g.launch(s)
    s.sync()
    want = ref.read_sums()
    assert want[2].sum() > 0
    v = api.Volume(ctx, size, 30.0, 3.0)
    assert v.brick == api.Volume.BRICKS[0]
    for shape in api.Volume.BRICKS:
        v.brick = shape
        assert v.brick == shape
        v.clear()
        v.splat(p, exact=True)
        ctx.sync()
        got = v.read_sums()
        assert all(np.array_equal(a, b) for a, b in zip(got, want)), shape
    with pytest.raises(api.GargError, match="not one of the exact deposit's shapes"):
        v.brick = (4, 4, 4)
    for o in (g, s, ref, v, p):
        o.close()


@pytest.mark.parametrize("size,hz,count", [(12, 30.0, 7), (64, 30.0, 4097), (96, 30.0, 250_000),
                                           (128, 3.0, 1_200_000), (256, 3.0, 300_000), (320, 3.0, 20_000),
                                           (336, 3.0, 20_000)])
def test_binned_one_corner_deposit_equals_atomic_deposit(ctx, size, hz, count):
    """The one-corner deposit sorted by 16x16x8 brick (the default method)
    against the reference's pass of atomics: the same integer sums, bit for
    bit, added onto an earlier deposit. The counts run from one count
    workgroup (7, 4097: a partial second) to all 256 (1.2M); the sides leave
    partial bricks (12, 320, the largest the brick tables hold) or exceed the
    tables (336, which deposits with atomics); a 3-unit half height puts
    most of the disk outside the cube."""
    p = settled(ctx, count, 9, 30)
    first = settled(ctx, 5000, 4, 5)
    binned, atomic = (api.Volume(ctx, size, 30.0, hz) for _ in range(2))
    assert binned.method == api.SPLAT_SORTED
    atomic.method = api.SPLAT_ATOMIC
    assert atomic.method == api.SPLAT_ATOMIC
    for v in (binned, atomic):
        v.splat(first, exact=False, frame=3)
        v.splat(p, exact=False, frame=11)
    got, want = binned.read_sums(), atomic.read_sums()
    assert want[2].sum() > 0
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    with pytest.raises(api.GargError, match="unknown splat method"):
        binned.method = 2
    for o in (binned, atomic, p, first):
        o.close()


def test_captured_binned_deposit_replays_like_the_atomic_deposit(ctx):
    """The binned one-corner deposit captured once with its frame from the
    dispatch count and replayed after each sim dispatch, against the atomic
    pass run eagerly with that frame: the sums agree after every replay
    (both volumes keep adding), with the counting sort's tables rewritten by
    each replay and no memset in the graph."""
    p = settled(ctx, 400_000, 8, 25)
    graphed, atomic = (api.Volume(ctx, 128, 30.0, 3.0) for _ in range(2))
    atomic.method = api.SPLAT_ATOMIC
    s = ctx.stream()
    ctx.sync()
    ctx.graph_begin(s)
    try:
        graphed.splat(p, exact=False, frame=api.FRAME_AUTO, stream=s)
    finally:
        g = ctx.graph_end(s)
    for _ in range(3):
        p.step(1)
        ctx.sync()
        g.launch(s)
        s.sync()
        atomic.splat(p, exact=False, frame=p.dispatches())
        ctx.sync()
        got, want = graphed.read_sums(), atomic.read_sums()
        assert want[2].sum() > 0
        assert all(np.array_equal(a, b) for a, b in zip(got, want))
    for o in (g, s, graphed, atomic, p):
        o.close()


@pytest.mark.parametrize("case", ["one_point", "brick_faces", "few"])
def test_binned_deposit_edge_cases_equal_the_atomic_pass(ctx, case):
    """The brick-sorted one-corner deposit where its sort and LDS gather are
    stressed, against the atomic pass bit for bit: every particle at one
    point (their corners are one cell's eight voxels: one hot brick, LDS
    atomics to the same words from every lane, sums that wrap modulo 2^32),
    particles within a voxel of the 16x16x8 bricks' faces on every axis (a
    particle's chosen voxel on either side of a face), and a few particles
    scattered over many bricks of a 64^3 cube."""
    size, n = 64, 60_000
    rng = np.random.default_rng(11)
    count = 37 if case == "few" else n
    p = settled(ctx, count, 17, 3)
    rec = p.read()
    cell = 60.0 / size
    if case == "one_point":
        rec["position"] = np.array([0.37, -0.21, 0.05], np.float32)
        rec["temperature"] = np.float32(30000.0)  # large sums, so the u32 words wrap
    elif case == "brick_faces":
        faces = np.stack([rng.integers(1, size // 16, count) * 16, rng.integers(1, size // 16, count) * 16,
                          rng.integers(1, size // 8, count) * 8], axis=1)
        grid = faces + rng.uniform(-0.75, 0.75, (count, 3))  # voxel-centre coordinates around a face
        rec["position"] = ((grid + 0.5) * cell - 30.0).astype(np.float32)
    else:
        rec["position"] = rng.uniform(-29.5, 29.5, (count, 3)).astype(np.float32)
    p.write(rec)
    binned, atomic = (api.Volume(ctx, size, 30.0, 30.0) for _ in range(2))
    atomic.method = api.SPLAT_ATOMIC
    for v in (binned, atomic):
        v.splat(p, exact=False, frame=29)
    got, want = binned.read_sums(), atomic.read_sums()
    assert want[2].sum() > 0
    if case == "one_point":
        assert np.count_nonzero(want[2]) <= 8
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    for o in (binned, atomic, p):
        o.close()
