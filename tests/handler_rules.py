"""Collision handler scenes and deterministic callbacks for the tests: the same
answers given by the reference's handlers (oracle/ref_bridge.c), by the oracle
(ow_set_handler_fn) and by the host between cpgpuWorldStepBegin and
cpgpuWorldStepEnd (GPUWorld.step_with_callbacks).

The answers depend only on the pair of shapes and the step, never on the order
in which the callbacks come, so each side may visit the pairs in its own order.
"""
from chipmunk2d_amd.cabi import (EVENT_BEGIN, EVENT_POST_SOLVE, EVENT_PRE_SOLVE, EVENT_SEPARATE)
from chipmunk2d_amd.scenes import Rng, random_world

CIRCLE_T, POLY_T, GROUND_T, SENSOR_T = 1, 2, 3, 4
ALL_EVENTS = EVENT_BEGIN | EVENT_PRE_SOLVE | EVENT_POST_SOLVE | EVENT_SEPARATE


def handler_space(n_circles=150, n_polys=150, n_segments=30, extent=45.0, seed=21, sensors=0.05):
    """A pile whose shapes carry collision types (circles 1, polygons 2, static
    segments 3, a few sensors 4) and handlers: circle-polygon with every event,
    polygon-polygon with preSolve, circle-ground with preSolve and separate, and
    a wildcard handler of the sensors with begin and separate."""
    sp = random_world(n_circles=n_circles, n_polys=n_polys, n_segments=n_segments, extent=extent, seed=seed)
    rng = Rng(seed + 1)
    for s in sp.shapes:
        if s.body.type == 2:
            s.collision_type = GROUND_T
        elif rng.uniform(0.0, 1.0) < sensors:
            s.sensor = True
            s.collision_type = SENSOR_T
        else:
            s.collision_type = CIRCLE_T if s.type == 0 else POLY_T
    sp.add_collision_handler(CIRCLE_T, POLY_T, ALL_EVENTS)
    sp.add_collision_handler(POLY_T, POLY_T, EVENT_PRE_SOLVE)
    sp.add_collision_handler(CIRCLE_T, GROUND_T, EVENT_PRE_SOLVE | EVENT_SEPARATE)
    sp.add_wildcard_handler(SENSOR_T, EVENT_BEGIN | EVENT_SEPARATE)
    return sp


def _mix(a, b):
    lo, hi = min(a, b), max(a, b)
    return (lo * 7919 + hi * 104729 + (lo * hi) % 65521) & 0x7FFFFFFF


class Rule:
    """begin: refuse one pair in 11 (ignored until it separates), ignore one in
    13 by cpArbiterIgnore; preSolve: skip one pair-step in 7, quarter the
    friction of one pair in 5, set the restitution of another, a surface
    velocity for one in 9, ignore one pair-step in 97.  Every call is logged as
    (step, event, space, lower shape, higher shape)."""

    def __init__(self):
        self.step = 0
        self.log = set()

    def __call__(self, event, v, d):
        h = _mix(v.shape_a, v.shape_b)
        self.log.add((self.step, int(event), int(v.space), min(v.shape_a, v.shape_b), max(v.shape_a, v.shape_b)))
        if event == EVENT_BEGIN:
            if h % 11 == 0:
                d.solve = 0
            elif h % 13 == 0:
                d.ignore = 1
        elif event == EVENT_PRE_SOLVE:
            if (h + self.step) % 7 == 0:
                d.solve = 0
            if h % 5 == 1:
                d.u = v.u * 0.25
            if h % 5 == 2:
                d.e = 0.5
            if h % 9 == 4:
                d.surface_vr[0], d.surface_vr[1] = 1.5, -0.5
            if (h + self.step) % 97 == 3:
                d.ignore = 1

    def begin(self, v, d):
        self(EVENT_BEGIN, v, d)

    def pre_solve(self, v, d):
        self(EVENT_PRE_SOLVE, v, d)

    def events(self, kind):
        return {e for e in self.log if e[1] == kind}
