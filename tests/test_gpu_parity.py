"""Parity of the HIP step (chipmunk2d_amd/libcpgpu.so, through its C ABI)
with the oracle -- needs an MI355X.

The device step and the oracle perform the same IEEE double operations in
the same order (the oracle solves in the device's colour order), so every
comparison below is bit-exact: body state after every step, the arbiter set
with its contacts, impulses, feature hashes, states and colours, and the
joints' accumulated impulses.
"""
import numpy as np
import pytest

from chipmunk2d_amd.cabi import concat
from chipmunk2d_amd.gpu import GPUWorld
from chipmunk2d_amd.scenes import pyramid_210, random_world, random_pile, chain_world, Rng
from chipmunk2d_amd.space import Space, Body, Circle, Segment, Box, moment_for_circle
from checkers import OracleWorld
import test_oracle_pins as pins

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def assert_bodies_equal(g, o, label=""):
    gb, ob = g.bodies(), o.bodies()
    for f in ob.FIELDS:
        x, y = getattr(gb, f), getattr(ob, f)
        if not np.array_equal(x, y):
            diff = np.abs(x - y)
            raise AssertionError("%s: body field %s differs (max %g at %s)" % (label, f, np.nanmax(diff),
                                                                            np.unravel_index(np.nanargmax(diff), diff.shape)))


def _key_order(shapes):
    """Order of arbiters by their unordered shape pair, and the sorted pairs."""
    lo, hi = np.minimum(shapes[:, 0], shapes[:, 1]), np.maximum(shapes[:, 0], shapes[:, 1])
    idx = np.lexsort((hi, lo))
    return idx, np.stack([lo[idx], hi[idx]], axis=1)


def assert_arbiters_equal(g, o):
    ga, oa = g.arbiters(), o.arbiters()
    assert ga.n == oa.n
    idx_g, keys_g = _key_order(ga.shapes)
    idx_o, keys_o = _key_order(oa.shapes)
    assert np.array_equal(keys_g, keys_o)
    assert np.array_equal(ga.shapes[idx_g], oa.shapes[idx_o])
    assert np.array_equal(ga.colour[idx_g], oa.colour[idx_o])
    assert np.array_equal(ga.state[idx_g], oa.state[idx_o])
    assert np.array_equal(ga.count[idx_g], oa.count[idx_o])
    assert np.array_equal(ga.normal[idx_g], oa.normal[idx_o])
    assert np.array_equal(ga.contact[idx_g], oa.contact[idx_o])
    assert np.array_equal(ga.hash[idx_g], oa.hash[idx_o])


def assert_pairs_equal(g, o, label=""):
    """The device's candidate pairs of the last step equal the oracle's as a set
    (both sides report (lower, higher) shape indices, sorted)."""
    gp, op = g.pairs(), o.candidate_pairs()
    if not np.array_equal(gp, op):
        gs, os_ = set(map(tuple, gp.tolist())), set(map(tuple, op.tolist()))
        raise AssertionError("%s: candidate pairs differ: %d on the device, %d in the oracle, %d only on the device "
                             "(e.g. %s), %d only in the oracle (e.g. %s)"
                             % (label, len(gp), len(op), len(gs - os_), sorted(gs - os_)[:3], len(os_ - gs),
                                sorted(os_ - gs)[:3]))


def run_lockstep(pw, steps, check_every=1, arbiters_every=0, g=None):
    g = g if g is not None else GPUWorld(pw)
    o = OracleWorld(pw)
    for s in range(steps):
        g.step(DT)
        o.step(DT)
        if (s + 1) % check_every == 0 or s == steps - 1:
            assert_bodies_equal(g, o, "step %d" % s)
        if arbiters_every and (s + 1) % arbiters_every == 0:
            assert_arbiters_equal(g, o)
            assert_pairs_equal(g, o, "step %d" % s)
    assert_pairs_equal(g, o, "last step")
    gs, os_ = g.stats(), o.stats()
    for k in ("candidate_pairs", "arbiters", "contacts", "cached_arbiters", "arbiter_colours"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    return g, o


def test_pyramid_bit_exact():
    g, o = run_lockstep(pyramid_210().pack(), 300, check_every=1, arbiters_every=50)
    assert_arbiters_equal(g, o)
    # the stack settles and stays up (the reference scene's purpose)
    b = g.bodies()
    assert b.p[1:, 1].max() > 300.0


@pytest.mark.parametrize("seed", [7, 8])
def test_random_world_bit_exact(seed):
    pw = random_world(n_circles=500, n_polys=500, n_segments=200, extent=120.0, seed=seed).pack()
    g, o = run_lockstep(pw, 150, check_every=1, arbiters_every=25)
    assert g.stats()["arbiters"] > 500


def test_chain_with_joints_bit_exact():
    pw = chain_world(n_links=80, n_circles=120, seed=42).pack()
    g, o = run_lockstep(pw, 150, check_every=5, arbiters_every=50)
    jg, cg = g.constraints()
    jo, co = o.constraints()
    assert np.array_equal(cg, co)
    assert np.array_equal(jg, jo)
    assert np.count_nonzero(jg) > 50


def test_batch_of_worlds_bit_exact():
    rng = Rng(0xBA7C4)
    spaces = [pyramid_210(rows=r).pack() for r in (3, 6, 9)]
    spaces += [random_pile(n, rng).pack() for n in (16, 100, 300)]
    spaces += [chain_world(n_links=20, n_circles=10, seed=5).pack()]
    pw = concat(spaces)
    assert pw.n_worlds == 7
    run_lockstep(pw, 120, check_every=10, arbiters_every=40)


def _strips_seen(o):
    """Strips holding arbiters, and arbiters of seams or the wide set, in the oracle's last step."""
    u = o.units()
    return len(np.unique(u[(u > 0) & (u <= 64)])), int(np.count_nonzero(u > 64))


def test_grid_solve_mid_size_world_bit_exact():
    """More dynamic bodies than one workgroup takes, in small scattered
    clusters: the grid solve's strips hold many independent clusters."""
    pw = random_world(n_circles=2600, n_polys=2600, n_segments=400, extent=260.0, seed=12).pack()
    assert pw.dynamic_body_count() > 4096
    g, o = run_lockstep(pw, 60, check_every=2, arbiters_every=30)
    assert g.stats()["arbiters"] > 300


def test_grid_strips_dense_world_bit_exact():
    """A dense large space of ~10k arbiters, solved on the grid in 64 strips,
    the seams between them and the wide set."""
    pw = random_world(n_circles=2600, n_polys=2600, n_segments=400, extent=140.0, seed=12).pack()
    g, o = run_lockstep(pw, 60, check_every=1, arbiters_every=20)
    strips, boundary = _strips_seen(o)
    assert strips == 64 and boundary > 1000, (strips, boundary)


def test_grid_strips_crowded_world_bit_exact():
    """A crowded large space (12k bodies, ~70k arbiters): strips of 400-700
    arbiters, so the grid solve keeps some strips whole in LDS with their
    bodies and of others only the smallest colours."""
    pw = random_world(n_circles=6000, n_polys=6000, n_segments=400, extent=110.0, seed=21).pack()
    g, o = run_lockstep(pw, 24, check_every=3, arbiters_every=12)
    u = o.units()
    per_strip = np.bincount(u[(u > 0) & (u <= 64)], minlength=65)[1:]
    print("strip arbiters: min %d max %d" % (per_strip.min(), per_strip.max()), flush=True)
    assert per_strip.max() > 600 and per_strip.min() < 500, per_strip


def test_grid_strips_wide_world_runtime_strip_count_bit_exact():
    """30k dynamic bodies spread wide: the grid set is cut into 79 strips
    (about 384 bodies each), one workgroup per strip."""
    pw = random_world(n_circles=15000, n_polys=15000, n_segments=1000, extent=800.0, seed=13).pack()
    g, o = run_lockstep(pw, 30, check_every=3, arbiters_every=15)
    u = o.units()
    assert len(np.unique(u[(u > 0) & (u <= 79)])) == 79 and np.count_nonzero(u > 79) > 0


def test_grid_strips_chain_with_joints_bit_exact():
    """A large space with a chain (pivots, springs, slides, gears) on the grid:
    strips of the links and the circles, then the joints."""
    pw = chain_world(n_links=300, n_circles=3900, seed=9).pack()
    g, o = run_lockstep(pw, 100, check_every=2, arbiters_every=50)
    strips, boundary = _strips_seen(o)
    assert strips >= 16 and boundary > 0, (strips, boundary)
    jg, cg = g.constraints()
    jo, co = o.constraints()
    assert np.array_equal(cg, co)
    assert np.array_equal(jg, jo)


@pytest.mark.parametrize("schedule", ["flow", "phases"])
def test_grid_strips_flat_pile_schedules_bit_exact(schedule):
    """A wide flat pile (no joints, no arbiter spanning non-neighbouring
    strips): the dependency-driven schedule of strips and seams, and the
    phase-barrier one, both bit-exact with the oracle."""
    from chipmunk2d_amd.gpu import GRID_FLOW, GRID_PHASES
    pw = random_world(n_circles=2600, n_polys=2600, n_segments=400, extent=400.0, seed=15).pack()
    g = GPUWorld(pw)
    g.set_grid_schedule(GRID_FLOW if schedule == "flow" else GRID_PHASES)
    o = OracleWorld(pw)
    seen = []
    for s in range(160):
        g.step(DT)
        o.step(DT)
        if s % 4 == 3 or s == 159:
            assert_bodies_equal(g, o, "step %d" % s)
        if s >= 100 and s % 20 == 19:
            u = o.units()
            seen.append((len(np.unique(u[(u > 0) & (u <= 64)])), int(np.count_nonzero((u > 64) & (u < 128))),
                         int(np.count_nonzero(u == 128))))
    assert_arbiters_equal(g, o)
    assert_pairs_equal(g, o)
    print(seen)
    assert all(st == 64 and seams > 500 and wide == 0 for st, seams, wide in seen), seen
    gs, os_ = g.stats(), o.stats()
    for k in ("candidate_pairs", "arbiters", "contacts", "cached_arbiters", "arbiter_colours"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])


def test_solve_only_timing_records_only_solver_launches():
    g = GPUWorld(concat([pyramid_210(rows=6).pack(), random_world(n_circles=2600, n_polys=2600, n_segments=400,
                                                                  extent=140.0, seed=12).pack()]))
    g.set_timing(True, solve_only=True)
    g.step(DT, 3)
    t = g.timing()
    assert t and all(k.startswith("solver") for k in t), t
    assert t["solver"][1] == 3
    g.set_timing(True)
    g.step(DT, 2)
    t = g.timing()
    assert "narrowphase" in t and t["solver"][1] == 2
    from chipmunk2d_amd.gpu import CpgpuError, _check
    with pytest.raises(CpgpuError):
        _check(g.L.cpgpuWorldSetTiming(g.h, 3))


def test_grid_flow_schedule_two_large_spaces_in_a_batch_bit_exact():
    """Two wide flat piles and small spaces in one step: the completion counters
    of both large spaces' strips and seams (no wide arbiters, no joints)."""
    rng = Rng(0xF10)
    a = random_world(n_circles=2600, n_polys=2600, n_segments=400, extent=400.0, seed=15).pack()
    b = random_world(n_circles=2500, n_polys=2500, n_segments=300, extent=380.0, seed=17).pack()
    pw = concat([pyramid_210(rows=6).pack(), a, random_pile(150, rng).pack(), b])
    g = GPUWorld(pw)
    o = OracleWorld(pw)
    for s in range(125):
        g.step(DT)
        o.step(DT)
        if s % 5 == 4:
            assert_bodies_equal(g, o, "step %d" % s)
    assert_arbiters_equal(g, o)
    assert_pairs_equal(g, o)
    u = o.units()
    assert np.count_nonzero(u == 128) == 0
    assert len(np.unique(u[(u > 0) & (u <= 64)])) == 64 and np.count_nonzero((u > 64) & (u < 128)) > 500


def test_grid_schedule_rejects_unknown_values():
    from chipmunk2d_amd.gpu import CpgpuError
    g = GPUWorld(pyramid_210(rows=3).pack())
    with pytest.raises(CpgpuError):
        g.set_grid_schedule(5)


def test_grid_strips_with_kinematic_bodies_bit_exact():
    """Kinematic paddles sweeping through a dense large space: they join no
    strip and constrain no colour, their contacts go with the dynamic body's
    strip."""
    sp = random_world(n_circles=2600, n_polys=2600, n_segments=400, extent=140.0, seed=16)
    for k in range(4):
        kin = sp.add(Body.kinematic())
        kin.position = (-60.0 + 30.0 * k, -40.0 + 20.0 * k)
        kin.velocity = (25.0 if k % 2 == 0 else -25.0, 0.0)
        kin.angular_velocity = 0.5
        sp.add(Box(kin, 24.0, 3.0, 0.0))
    pw = sp.pack()
    g, o = run_lockstep(pw, 40, check_every=2, arbiters_every=20)
    strips, others = _strips_seen(o)
    assert strips == 64 and others > 1000, (strips, others)


def test_mixed_batch_grid_and_workgroup_solves_bit_exact():
    """Two large spaces (grid sets in strips) and small spaces (a workgroup each) in one step."""
    rng = Rng(0x1A46E)
    big = random_world(n_circles=2400, n_polys=2400, n_segments=300, extent=130.0, seed=13).pack()
    big2 = random_world(n_circles=2300, n_polys=2300, n_segments=200, extent=125.0, seed=14).pack()
    spaces = [pyramid_210(rows=5).pack(), big, random_pile(120, rng).pack(), big2, pyramid_210(rows=8).pack()]
    pw = concat(spaces)
    g, o = run_lockstep(pw, 40, check_every=5, arbiters_every=20)
    strips, boundary = _strips_seen(o)
    assert strips == 64 and boundary > 1000, (strips, boundary)


def test_worlds_in_a_batch_step_as_when_alone():
    """A space stepped inside a batch ends bit-identical to the same space stepped alone."""
    rng = Rng(77)
    a = random_pile(150, rng).pack()
    b = pyramid_210(rows=7).pack()
    g_batch = GPUWorld(concat([a, b]))
    g_alone = GPUWorld(b)
    for _ in range(60):
        g_batch.step(DT)
        g_alone.step(DT)
    nb_a = a.n_bodies
    x, y = g_batch.bodies(), g_alone.bodies()
    assert np.array_equal(x.p[nb_a:], y.p) and np.array_equal(x.v[nb_a:], y.v) and np.array_equal(x.a[nb_a:], y.a)


def test_repeatable_run_to_run():
    pw = random_world(n_circles=800, n_polys=800, n_segments=200, extent=150.0, seed=19).pack()
    g1, g2 = GPUWorld(pw), GPUWorld(pw)
    for _ in range(80):
        g1.step(DT)
        g2.step(DT)
    b1, b2 = g1.bodies(), g2.bodies()
    assert np.array_equal(b1.p, b2.p) and np.array_equal(b1.v, b2.v) and np.array_equal(b1.a, b2.a)


def test_filters_groups_masks_and_no_collide_joints():
    sp = random_world(n_circles=300, n_polys=300, n_segments=80, extent=60.0, seed=23)
    for k, s in enumerate(sp.shapes):
        if k % 5 == 0:
            s.set_filter(group=7)
        if k % 9 == 0:
            s.set_filter(categories=1, mask=2)
    run_lockstep(sp.pack(), 80, check_every=4, arbiters_every=40)


def test_shapes_collide_bit_exact():
    """cpShapesCollide (cpShape.c:260-283) on the device for every candidate pair in both
    argument orders, plus far-apart pairs, at creation and after 30 steps: the contact point
    sets equal the oracle's (itself pinned against the reference's cpShapesCollide)."""
    pw = random_world(n_circles=300, n_polys=300, n_segments=80, extent=60.0, seed=23).pack()
    g, o = GPUWorld(pw), OracleWorld(pw)
    o.set_transforms(pw.a["body_transform"])
    rng = np.random.default_rng(11)
    for steps in (0, 30):
        if steps:
            g.step(DT, steps)
            o.step(DT, steps)
        cand = o.candidate_pairs()
        far = rng.integers(0, pw.n_shapes, size=(200, 2)).astype(np.int32)
        pairs = np.concatenate([cand, cand[:, ::-1], far]).astype(np.int32)
        count, sets = g.shapes_collide(pairs)
        touching = 0
        for i, (a, b) in enumerate(pairs):
            co, so = o.shapes_collide(int(a), int(b))
            assert count[i] == co, (steps, a, b)
            assert np.array_equal(sets[i, :2 + 5 * co], so[:2 + 5 * co]), (steps, a, b, sets[i], so)
            assert not np.any(sets[i, 2 + 5 * co:]), (steps, a, b)
            touching += co > 0
        assert touching > 200, touching
    with pytest.raises(Exception, match="out of range"):
        g.shapes_collide([[0, pw.n_shapes]])


# ---- known answers of the reference's own tests, on the device ----

def test_time_stepping_known_answer_gpu():
    pins.check_time_stepping(GPUWorld)


def test_basic_simulation_gpu():
    pins.check_basic_simulation(GPUWorld)


def test_ball_rests_on_bar_gpu():
    """CallbacksTest.m:257-292 without the removed bar: a ball (r 1) on a bar of
    radius 1 at y 0 rests at y 2 within 1.1 slop."""
    space = Space()
    space.gravity = (0.0, -100.0)
    space.add(Segment(space.static_body, (-10.0, 0.0), (10.0, 0.0), 1.0))
    ball = space.add(Body(1.0, moment_for_circle(1.0, 0.0, 1.0)))
    ball.position = (0.0, 10.0)
    space.add(Circle(ball, 1.0))
    g = GPUWorld(space.pack())
    for _ in range(100):
        g.step(0.01)
    assert abs(g.bodies().p[1][1] - 2.0) < 1.1 * space.collision_slop


# ---- edge cases ----

def test_empty_and_shapeless_worlds():
    space = Space()
    space.gravity = (0.0, -10.0)
    b = space.add(Body(2.0, 3.0))
    b.velocity = (1.0, 2.0)
    run_lockstep(space.pack(), 10)
    space = Space()
    space.add(Segment(space.static_body, (0.0, 0.0), (1.0, 0.0), 0.0))
    run_lockstep(space.pack(), 5)


def test_zero_dt_is_a_no_op():
    pw = pyramid_210(rows=4).pack()
    g = GPUWorld(pw)
    before = g.bodies()
    g.step(0.0)
    after = g.bodies()
    assert np.array_equal(before.p, after.p) and g.stats()["steps"] == 0


def test_coincident_circles():
    """Zero distance: cpCollide's fallback normal (1, 0) (cpCollision.c:533)."""
    space = Space()
    for _ in range(2):
        body = space.add(Body(1.0, 1.0))
        body.position = (3.0, 4.0)
        space.add(Circle(body, 1.0))
    g, o = run_lockstep(space.pack(), 3)
    assert_arbiters_equal(g, o)


def test_kinematic_body_pushes_dynamic():
    space = Space()
    kin = space.add(Body.kinematic())
    kin.velocity = (5.0, 0.0)
    space.add(Box(kin, 4.0, 4.0, 0.0))
    for k in range(5):
        body = space.add(Body(1.0, 1.0))
        body.position = (5.0 + 2.5 * k, 0.0)
        space.add(Circle(body, 1.0))
    g, o = run_lockstep(space.pack(), 120, check_every=1)
    assert g.bodies().p[1][0] > 10.0


def test_variable_dt_cached_impulse_scaling():
    """dt_coef = dt / prev_dt scales the warm start (cpSpaceStep.c:407)."""
    pw = pyramid_210(rows=6).pack()
    g, o = GPUWorld(pw), OracleWorld(pw)
    for k in range(60):
        dt = DT if k % 3 else DT * 0.5
        g.step(dt)
        o.step(dt)
    assert_bodies_equal(g, o)


def test_velocity_and_force_writes():
    pw = pyramid_210(rows=5).pack()
    g, o = GPUWorld(pw), OracleWorld(pw)
    for k in range(30):
        g.step(DT)
        o.step(DT)
    n = pw.n_bodies
    rng = np.random.default_rng(3)
    v = rng.normal(size=(n, 2))
    f = rng.normal(size=(n, 2)) * 10
    v[0] = 0.0
    f[0] = 0.0
    g.write_bodies(v=v, f=f)
    o.write_bodies(v=np.ascontiguousarray(v), f=np.ascontiguousarray(f))
    for k in range(30):
        g.step(DT)
        o.step(DT)
    assert_bodies_equal(g, o)


# ---- the full-size workloads: size-independent properties ----

def test_random_world_100k_first_steps_bit_exact():
    """The 100k-shape world (50k circles, 40k polys, 10k static segments):
    the first steps are small enough for the oracle."""
    pw = random_world().pack()
    assert pw.n_shapes == 100000
    g, o = GPUWorld(pw), OracleWorld(pw)
    for s in range(3):
        g.step(DT)
        o.step(DT)
        assert_bodies_equal(g, o, "step %d" % s)
        assert_pairs_equal(g, o, "step %d" % s)
    assert_arbiters_equal(g, o)
    assert g.stats()["candidate_pairs"] == o.stats()["candidate_pairs"]


def test_batch_tiling_and_autotuner_keep_results_bit_identical():
    """The workgroup size of a size bucket, LDS staging and the autotuner change
    only the schedule: the colour order, and so every result, is the same."""
    rng = Rng(0x7113)
    spaces = [random_pile(n, rng).pack() for n in (20, 70, 300, 900)]
    spaces += [pyramid_210(rows=r).pack() for r in (4, 12)]
    pw = concat(spaces)
    dt = DT
    g_default = GPUWorld(pw)
    g_single = GPUWorld(pw)
    g_single.set_buckets([1 << 30], [1024])
    assert g_single.buckets()[0][2] == pw.n_worlds
    g_tuned = GPUWorld(pw)
    threads = g_tuned.autotune(dt, 2)          # 10 steps
    assert all(t in (0, 64, 128, 256, 512, 1024) for t in threads)
    for g in (g_default, g_single):
        g.step(dt, 10)
    for g in (g_default, g_single, g_tuned):
        g.step(dt, 30)
    a, b, c = g_default.bodies(), g_single.bodies(), g_tuned.bodies()
    for f in a.FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
        assert np.array_equal(getattr(a, f), getattr(c, f)), f


def test_sensors_bit_exact():
    """A fifth of the shapes are sensors (cpShape.sensor): their pairs are never solved
    (cpSpaceStep.c:268-270) and are not candidate pairs; the rest steps bit for bit."""
    sp = random_world(n_circles=400, n_polys=400, n_segments=60, extent=60.0, seed=55)
    rng = Rng(56)
    for shape in sp.shapes:
        if rng.uniform(0.0, 1.0) < 0.2:
            shape.sensor = True
    pw = sp.pack()
    assert pw.a["shape_sensor"].sum() > 100
    g, o = run_lockstep(pw, 120, check_every=4, arbiters_every=40)
    sensor = pw.a["shape_sensor"].astype(bool)
    pairs = g.pairs()
    assert not np.any(sensor[pairs[:, 0]] | sensor[pairs[:, 1]])
