"""The flows of the cpGPUSpace / cpGPUBatch checks (include/cpGPUSpace.h),
run on either build of the binding: over the device library
(tests/test_gpuspace.py, needs an MI355X) and over the oracle's CPU step
(tests/test_binding_cpu.py).

The binding packs the spaces from Chipmunk's own structs; every flow compares
what it produced with the oracle stepping the binding's own description
(cpGPUSpaceGetDesc), bit for bit:
  * body state after cpGPUSpaceSyncHost, read from the cpBody structs;
  * the arbiters threaded onto the bodies (cpBodyEachArbiter, the contact
    graph of cpSpaceProcessComponents), with cpArbiterTotalImpulse and the
    contact point sets;
  * setters between steps (velocity, force, position, gravity, iterations);
  * adding and removing bodies and shapes mid-run: the new pack keeps the
    cached arbiters and warm starts of every surviving pair
    (cpgpuWorldCreateFrom against the oracle's ow_new_from);
  * a batch of plain cpSpaces, and a space in a batch against the same space
    alone;
  * every joint class: the accumulated impulses (and ratchet angles) written
    back into the constraint structs by cpGPUSpaceSyncHost, a motor's rate
    changed between steps.
"""
import ctypes

import numpy as np

from checkers import OracleWorld, packed_from_desc, ref_lib
from chipmunk2d_amd.cabi import ARB_NORMAL, BODY_DYNAMIC
from chipmunk2d_amd.gpu import GPUWorld
from chipmunk2d_amd.gpu import lib as gpu_lib
from chipmunk2d_amd.scenes import chain_world, machine_world, pyramid_210, random_world
from gpuspace import RAW, body_arbiters, body_raw, cp, gs, v

DT = 1.0 / 60.0


def errors_raise(kind):
    """Binding errors come back as messages (not cpAssertHard's abort) and fail the test."""
    gs(kind).cpGPUSetAbortOnError(0)
    gs(kind).cpGPUClearError()


def errors_abort(kind):
    gs(kind).cpGPUSetAbortOnError(1)


def _no_error(kind):
    err = gs(kind).cpGPUGetError().decode()
    assert not err, err


class WorldView(GPUWorld):
    """The world behind a cpGPUSpace, read through the cpgpuWorld readers of the
    library the binding was built over (not owned)."""

    def __init__(self, handle, packed, L):
        self.L = L
        self.packed = packed
        self.h = ctypes.c_void_p(handle)

    def close(self):
        self.h = None


class GPUSpace:
    """A cpGPUSpace of the binding build `kind` filled from a packed description
    through the public API (ref_bridge's rb_new_in); `bodies` and `shapes` hold
    its objects in the description's order (None once a test removes one)."""

    def __init__(self, pw, kind, precision=0):
        self.kind = kind
        self.B = gs(kind)
        self.space = self.B.cpGPUSpaceNewWithPrecision(precision)
        self.desc_keep = pw.desc()
        self.rb = ref_lib().rb_new_in(self.space, ctypes.byref(self.desc_keep), 0)
        self.bodies = [ref_lib().rb_body(self.rb, k) for k in range(pw.n_bodies)]
        self.shapes = [ref_lib().rb_shape(self.rb, k) for k in range(pw.n_shapes)]
        self.edited = False

    def step(self, dt=DT):
        self.B.cpGPUSpaceStep(self.space, dt)
        _no_error(self.kind)

    def sync(self):
        self.B.cpGPUSpaceSyncHost(self.space)
        _no_error(self.kind)

    def desc(self):
        return packed_from_desc(self.B.cpGPUSpaceGetDesc(self.space))

    def shape_map(self, n):
        m = self.B.cpGPUSpaceGetShapeMap(self.space)
        return np.array([m[k] for k in range(n)], np.int32)

    def world(self):
        return WorldView(self.B.cpGPUSpaceGetWorld(self.space), self.desc(), gpu_lib() if self.kind == "gpu" else self.B)

    def index(self, body):
        return self.B.cpGPUSpaceBodyIndex(self.space, body)

    def free(self):
        # after adds and removes ref_bridge's lists are stale: the space alone is freed
        if not self.edited:
            ref_lib().rb_free(self.rb)
        self.B.cpGPUSpaceFree(self.space)


def _assert_host_equals_oracle(g, o, what):
    """Every live body's state read from its cpBody struct equals the oracle's."""
    ob = o.bodies()
    bodies = [b for b in g.bodies if b is not None]
    assert len(bodies) == o.packed.n_bodies
    for b in bodies:
        j = g.index(b)
        raw = body_raw(b)
        for f in ("p", "v", "a", "w", "v_bias", "w_bias", "transform"):
            assert np.array_equal(raw[RAW[f]], getattr(ob, f)[j]), (what, f, j)


def steps_bit_exact_and_threads_arbiters(kind):
    pw = pyramid_210(rows=10).pack()
    g = GPUSpace(pw, kind)
    B = g.B
    g.step()
    o = OracleWorld(g.desc())
    o.step(DT)
    for s in range(59):
        g.step()
        o.step(DT)
    g.sync()
    _assert_host_equals_oracle(g, o, "after 60 steps")
    assert B.cpGPUSpaceGetPackCount(g.space) == 1
    # the contact graph: every body's arbiters through cpBodyEachArbiter
    A = o.arbiters()
    shape_of = {B.cpGPUSpaceShapeIndex(g.space, s): s for s in g.shapes}
    by_pair = {}
    for k in range(A.n):
        by_pair[(shape_of[int(A.shapes[k, 0])], shape_of[int(A.shapes[k, 1])])] = k
    seen = 0
    for body in g.bodies:
        j = B.cpGPUSpaceBodyIndex(g.space, body)
        mine = [k for k in range(A.n) if j in (o.packed.a["shape_body"][A.shapes[k, 0]],
                                                 o.packed.a["shape_body"][A.shapes[k, 1]])]
        got = body_arbiters(body)
        assert len(got) == len(mine), (j, len(got), len(mine))
        for arb in got:
            key = (arb["a"], arb["b"]) if (arb["a"], arb["b"]) in by_pair else (arb["b"], arb["a"])
            k = by_pair[key]
            c = A.contact[k]
            n = A.count[k]
            assert arb["count"] == n
            assert arb["jn"] == sum(c[i, 7] for i in range(n)) and arb["jt"] == sum(c[i, 8] for i in range(n))
            # cpArbiterTotalImpulse: sum of rotate(n, (jn, jt)), negated unless swapped (the body is body_b)
            sx = sy = 0.0
            nx, ny = A.normal[k]
            for i in range(n):
                jn, jt = c[i, 7], c[i, 8]
                sx, sy = sx + (nx * jn - ny * jt), sy + (nx * jt + ny * jn)
            expect = (sx, sy) if arb["swapped"] else (-sx, -sy)
            assert tuple(arb["impulse"]) == expect
            # cpArbiterGetContactPointSet: absolute points from the bodies' positions after the sync
            ob = o.bodies()
            ba = o.packed.a["shape_body"][A.shapes[k, 0]]
            bb = o.packed.a["shape_body"][A.shapes[k, 1]]
            q = arb["points"]
            sw = arb["swapped"]
            assert (q[0], q[1]) == ((-nx, -ny) if sw else (nx, ny))
            for i in range(n):
                p1 = ob.p[ba] + c[i, 0:2]
                p2 = ob.p[bb] + c[i, 2:4]
                pa, pb = (p2, p1) if sw else (p1, p2)
                assert tuple(q[2 + 5 * i:4 + 5 * i]) == tuple(pa) and tuple(q[4 + 5 * i:6 + 5 * i]) == tuple(pb)
                d = p2 - p1
                assert q[6 + 5 * i] == d[0] * nx + d[1] * ny
            seen += 1
    # every arbiter is threaded onto both of its bodies, static ones included (cpSpaceComponent.c:264-265)
    assert seen == 2 * A.n and A.n > 50
    g.free()


def setters_between_steps_reach_the_world(kind):
    pw = random_world(n_circles=150, n_polys=150, n_segments=40, extent=60.0, seed=31).pack()
    g = GPUSpace(pw, kind)
    B = g.B
    g.step()
    o = OracleWorld(g.desc())
    o.step(DT)
    for _ in range(9):
        g.step()
        o.step(DT)
    C = cp()
    idx = lambda body: B.cpGPUSpaceBodyIndex(g.space, body)
    dyn = [b for b, t in zip(g.bodies, pw.a["body_type"]) if t == BODY_DYNAMIC]
    g.sync()
    # a velocity, a force at a point (force and torque), a new position and angle
    C.cpBodySetVelocity(dyn[3], v(4.5, -2.25))
    C.cpBodyApplyForceAtWorldPoint(dyn[7], v(300.0, 50.0), C.cpBodyGetPosition(dyn[7]))
    C.cpBodyApplyForceAtWorldPoint(dyn[8], v(-10.0, 75.0), v(C.cpBodyGetPosition(dyn[8]).x + 0.5, 0.0))
    C.cpBodySetPosition(dyn[11], v(1.0, 55.0))
    C.cpBodySetAngle(dyn[11], 0.75)
    C.cpBodySetAngularVelocity(dyn[12], 3.0)
    for b in (dyn[3], dyn[7], dyn[8], dyn[11], dyn[12]):
        raw = body_raw(b)
        o.write_bodies_at([idx(b)], p=raw[RAW["p"]], v=raw[RAW["v"]], f=raw[RAW["f"]], a=[raw[RAW["a"]]],
                          w=[raw[RAW["w"]]], t=[raw[RAW["t"]]], transform=raw[RAW["transform"]])
    # space parameters
    C.cpSpaceSetGravity(g.space, v(10.0, -50.0))
    C.cpSpaceSetIterations(g.space, 7)
    sp = dict(o.packed.a["space"][0])
    # the space's stamp and dt after the ten steps stay as they are
    sp.update(gravity_x=10.0, gravity_y=-50.0, iterations=7, stamp=int(sp["stamp"]) + 10, curr_dt=DT)
    o.set_space_params(0, sp)
    for _ in range(10):
        g.step()
        o.step(DT)
    g.sync()
    _assert_host_equals_oracle(g, o, "after setters and 10 steps")
    assert B.cpGPUSpaceGetPackCount(g.space) == 1, "setters need no new pack"
    g.free()


def add_and_remove_keep_warm_starts(kind):
    pw = random_world(n_circles=200, n_polys=200, n_segments=40, extent=50.0, seed=37).pack()
    g = GPUSpace(pw, kind)
    B = g.B
    g.step()
    o = OracleWorld(g.desc())
    o.step(DT)
    for _ in range(39):
        g.step()
        o.step(DT)
    C = cp()
    dyn_shapes = [k for k in range(pw.n_shapes) if pw.a["body_type"][pw.a["shape_body"][k]] == BODY_DYNAMIC]
    # remove three bodies with their shapes
    removed = dyn_shapes[5:8]
    for k in removed:
        body = pw.a["shape_body"][k]
        C.cpSpaceRemoveShape(g.space, g.shapes[k])
        C.cpSpaceRemoveBody(g.space, g.bodies[body])
        g.shapes[k] = None
        g.bodies[body] = None
    g.edited = True
    # add two circles on new bodies
    added = []
    for x in (-5.0, 5.0):
        body = C.cpBodyNew(1.5, C.cpMomentForCircle(1.5, 0.0, 0.8, v(0, 0)))
        C.cpSpaceAddBody(g.space, body)
        C.cpBodySetPosition(body, v(x, 60.0))
        shape = C.cpCircleShapeNew(body, 0.8, v(0.0, 0.0))
        C.cpShapeSetFriction(shape, 0.7)
        C.cpSpaceAddShape(g.space, shape)
        added.append((body, shape))
    g.step()
    assert B.cpGPUSpaceGetPackCount(g.space) == 2
    d2 = g.desc()
    assert d2.n_shapes == pw.n_shapes - 3 + 2
    o2 = OracleWorld.from_old(d2, o, g.shape_map(d2.n_shapes))
    o2.step(DT)
    world = g.world()
    A, B = world.arbiters(), o2.arbiters()
    assert A.n == B.n and np.array_equal(A.shapes, B.shapes) and np.array_equal(A.state, B.state)
    assert np.array_equal(A.contact, B.contact) and np.array_equal(A.hash, B.hash)
    # warm starts survived the new pack: pairs solved last step are NORMAL with their cached impulses
    assert np.count_nonzero(A.state == ARB_NORMAL) > 0.8 * A.n
    for _ in range(20):
        g.step()
        o2.step(DT)
    g.bodies += [b for b, _ in added]
    g.sync()
    _assert_host_equals_oracle(g, o2, "20 steps after the new pack")
    g.free()


def batch_of_plain_spaces_and_alone(kind):
    B = gs(kind)
    packs = [pyramid_210(rows=6).pack(), random_world(n_circles=60, n_polys=60, n_segments=20, extent=30.0,
                                                       seed=41).pack(), chain_world(10, 30, seed=3).pack()]
    rbs = [ref_lib().rb_new(ctypes.byref(p.desc()), 0, 0) for p in packs]
    spaces = (ctypes.c_void_p * len(rbs))(*[ref_lib().rb_space(r) for r in rbs])
    batch = B.cpGPUBatchNew(spaces, len(rbs))
    B.cpGPUBatchStep(batch, DT)
    _no_error(kind)
    d = packed_from_desc(B.cpGPUBatchGetDesc(batch))
    assert d.n_worlds == 3
    o = OracleWorld(d)
    o.step(DT)
    for _ in range(29):
        B.cpGPUBatchStep(batch, DT)
        o.step(DT)
    B.cpGPUBatchSyncHost(batch)
    ob = o.bodies()
    for w, (p, r) in enumerate(zip(packs, rbs)):
        # the binding's order within a space: bodies of space->dynamicBodies (dynamic and
        # kinematic), then of space->staticBodies, each in the order they were added
        t = p.a["body_type"]
        order = np.concatenate([np.nonzero(t != 2)[0], np.nonzero(t == 2)[0]])
        b0 = int(d.a["world_body_start"][w])
        for j, k in enumerate(order):
            raw = body_raw(ref_lib().rb_body(r, int(k)))
            for f in ("p", "v", "a", "w", "transform"):
                assert np.array_equal(raw[RAW[f]], getattr(ob, f)[b0 + j]), (w, int(k), f)
    # the third space alone, as a cpGPUSpace, steps identically
    alone = GPUSpace(packs[2], kind)
    for _ in range(30):
        alone.step()
    alone.sync()
    r = rbs[2]
    for k in range(packs[2].n_bodies):
        assert np.array_equal(body_raw(alone.bodies[k]), body_raw(ref_lib().rb_body(r, k))), k
    alone.free()
    B.cpGPUBatchFree(batch)
    for r in rbs:
        ref_lib().rb_free(r)


def joints_sync_to_host(kind):
    """machine_world (all ten joint classes): after steps and cpGPUSpaceSyncHost each
    constraint struct holds the world's accumulated impulse (and a ratchet its angle),
    read back through the reference's own struct layout; a motor's new rate reaches the
    world with the next step (a new pack), and the states carried through it."""
    pw = machine_world(n_units=4, n_circles=30).pack()
    g = GPUSpace(pw, kind)
    B = g.B
    g.step()
    o = OracleWorld(g.desc())
    o.step(DT)
    for _ in range(59):
        g.step()
        o.step(DT)
    g.sync()
    _assert_host_equals_oracle(g, o, "machines after 60 steps")
    cons = [ref_lib().rb_constraint(g.rb, k) for k in range(pw.n_constraints)]

    def host_jacc():
        out = np.zeros((pw.n_constraints, 2))
        for k, c in enumerate(cons):
            ref_lib().rb_constraint_jacc(c, out[k].ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    jo, _ = o.constraints()
    at = np.array([B.cpGPUSpaceConstraintIndex(g.space, c) for c in cons])
    assert sorted(at) == list(range(pw.n_constraints))
    assert np.array_equal(host_jacc(), jo[at])
    assert np.count_nonzero(jo) > pw.n_constraints // 2
    # a simple motor's rate changed by the caller: the constant changed, so the step packs again
    motor = [c for c, t in zip(cons, pw.a["cons_type"]) if t == 8][0]
    cp().cpSimpleMotorSetRate(motor, 2.5)
    g.step()
    assert B.cpGPUSpaceGetPackCount(g.space) == 2
    d2 = g.desc()
    assert d2.a["cons_params"][B.cpGPUSpaceConstraintIndex(g.space, motor), 0] == 2.5
    o2 = OracleWorld.from_old(d2, o, g.shape_map(d2.n_shapes))
    o2.step(DT)
    for _ in range(19):
        g.step()
        o2.step(DT)
    g.sync()
    _assert_host_equals_oracle(g, o2, "20 steps after a new motor rate")
    jo, _ = o2.constraints()
    assert np.array_equal(host_jacc(), jo[at])
    g.free()


def handlers_run_the_callers_callbacks(kind):
    """Collision handlers set through the reference's API (ref_bridge installs
    them from the description: a pair handler with every callback, pair
    handlers with preSolve, a wildcard handler for a few sensors) run as the
    reference runs them: the binding calls their begin and preSolve between the
    device's two half-steps, then separate and postSolve.  The same
    deterministic answers (handler_rules.Rule) from the oracle stepping the
    binding's description: each step the same pairs get each callback, and the
    bodies stay bit-exact."""
    from chipmunk2d_amd.cabi import EVENT_BEGIN, EVENT_POST_SOLVE, EVENT_PRE_SOLVE, EVENT_SEPARATE, HANDLER_FN
    from handler_rules import Rule, handler_space
    pw = handler_space(n_circles=90, n_polys=90, n_segments=20, extent=35.0, seed=41).pack()
    g = GPUSpace(pw, kind)
    rg = Rule()
    hfn = HANDLER_FN(lambda ctx, ev, vv, dd: rg(ev, vv.contents, dd.contents))
    ref_lib().rb_set_handler_fn(g.rb, hfn, None)
    g.step()
    # the rule sees the description's shape indices on both sides
    assert [g.B.cpGPUSpaceShapeIndex(g.space, sh) for sh in g.shapes] == list(range(pw.n_shapes))
    o = OracleWorld(g.desc())
    assert len(o.packed.a["handlers"]) == 4
    ro = Rule()
    o.set_handler_fn(ro)
    o.step(DT)
    kinds = (EVENT_BEGIN, EVENT_PRE_SOLVE, EVENT_POST_SOLVE, EVENT_SEPARATE)
    for s in range(1, 90):
        rg.step = ro.step = s
        g.step()
        o.step(DT)
        for k in kinds:
            eg = {e for e in rg.events(k) if e[0] == s}
            eo = {e for e in ro.events(k) if e[0] == s}
            assert eg == eo, (s, k, sorted(eg ^ eo)[:6])
        if s % 15 == 14:
            g.sync()
            _assert_host_equals_oracle(g, o, "step %d" % s)
    counts = {k: len(rg.events(k)) for k in kinds}
    assert all(c > 10 for c in counts.values()), counts
    g.free()


def faults_report_like_a_debug_build(kind):
    """A body whose velocity is NaN: cpGPUSpaceStep runs the step in full and
    reports the check a debug build of the reference fails there (aborting,
    by default, as that build does); the report comes once, and the space
    steps and syncs on as a release build would."""
    from fault_scenes import build
    pw, bad, fault, step, message = build("nan_velocity")
    g = GPUSpace(pw, kind)
    g.B.cpGPUSpaceStep(g.space, DT)
    err = g.B.cpGPUGetError().decode()
    assert "returned -5" in err and "Body's state is invalid" in err, err
    g.B.cpGPUClearError()
    o = OracleWorld(g.desc())   # the binding's own description, in its body order
    o.step(DT)
    g.step()
    o.step(DT)
    g.sync()
    ob = o.bodies()
    for k, b in enumerate(g.bodies):
        raw, j = body_raw(b), g.index(b)
        if k == bad:
            assert np.isnan(raw[RAW["p"]][0]) and np.isnan(ob.p[j][0])
        else:
            assert np.array_equal(raw[RAW["p"]], ob.p[j]) and np.array_equal(raw[RAW["v"]], ob.v[j]), k
    g.free()
