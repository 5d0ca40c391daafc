"""Spaces that trip the checks a debug build of the reference makes during
cpSpaceStep, each beside a small pyramid that steps normally.

    unsolvable    k_scalar / k_tensor of zero (include/chipmunk/chipmunk_private.h:223,255)
    body mass     cpBodyUpdateVelocity's positive mass and moment (src/cpBody.c:498)
    body state    cpBodySanityCheck after an integration (src/cpBody.c:110-125,508,521)

Each scene names the fault it raises first, the step it raises it at, the
body that fails (for the body checks) and the message the reference's debug
build aborts with.  The bodies that take no part in the fault (the pyramid
and its ground) keep stepping as before: the joints join two kinematic bodies
with no shapes, so the NaN impulses the reference then applies to both
(v += j * 0) reach nothing the pyramid touches.
"""
import math

from chipmunk2d_amd.cabi import FAULT_BODY_MASS, FAULT_BODY_STATE, FAULT_UNSOLVABLE
from chipmunk2d_amd.scenes import pyramid_210
from chipmunk2d_amd.space import Body, Circle, DampedSpring, PinJoint, PivotJoint, moment_for_circle

DT = 1.0 / 60.0


def _base():
    sp = pyramid_210(rows=4)
    return sp


def _ball(sp, pos, mass=1.0, radius=2.0, shape=True):
    b = sp.add(Body(mass, moment_for_circle(mass, 0.0, radius)))
    b.position = pos
    if shape:
        sp.add(Circle(b, radius))
    return b


def _kinematic_pair(sp):
    a = sp.add(Body.kinematic())
    a.position = (900.0, 300.0)
    b = sp.add(Body.kinematic())
    b.position = (940.0, 300.0)
    return a, b


def kinematic_pin():
    """A pin joint between two kinematic bodies: both masses infinite,
    k_scalar is zero at the first step."""
    sp = _base()
    a, b = _kinematic_pair(sp)
    sp.add(PinJoint(a, b, (0.0, 0.0), (0.0, 0.0)))
    return sp, None


def kinematic_pivot():
    """A pivot joint between two kinematic bodies: the 2x2 effective mass is
    singular (k_tensor's determinant is zero)."""
    sp = _base()
    a, b = _kinematic_pair(sp)
    sp.add(PivotJoint(a, b, (920.0, 300.0)))
    return sp, None


def kinematic_spring():
    """A damped spring between two kinematic bodies, k_scalar zero."""
    sp = _base()
    a, b = _kinematic_pair(sp)
    sp.add(DampedSpring(a, b, (0.0, 0.0), (0.0, 0.0), 30.0, 50.0, 1.0))
    return sp, None


def zero_mass():
    """A dynamic body of mass zero: the velocity update's check fails first;
    the update then makes its velocity NaN (gravity plus 0 * inf)."""
    sp = _base()
    b = _ball(sp, (900.0, 300.0), shape=False)
    b.mass = 0.0
    return sp, b


def huge_force():
    """A light body under a force whose acceleration overflows: the velocity is
    infinite after the velocity update."""
    sp = _base()
    b = _ball(sp, (900.0, 300.0), mass=1e-300, shape=False)
    b.force = (1e300, 0.0)
    return sp, b


def nan_velocity():
    """A body whose velocity is NaN before the first step: its position is NaN
    after the first position update."""
    sp = _base()
    b = _ball(sp, (900.0, 300.0))
    b.v = (math.nan, 0.0)
    return sp, b


def runaway(steps_before=3):
    """A body flying along x near the largest double: its position overflows
    to infinity at the position update of step `steps_before`.  It carries a
    circle, so its shape's box goes infinite in the broadphase too."""
    sp = _base()
    b = _ball(sp, (0.0, 2000.0), radius=1.0)
    vx = 6e307
    step = vx * DT   # the position update's increment, exact enough to place the overflow
    b.position = (1.7976931348623157e308 - (steps_before + 0.5) * step, 2000.0)
    b.v = (vx, 0.0)
    return sp, b


# name: (builder, first fault, its step, the reference's abort message)
SCENES = {
    "kinematic_pin": (kinematic_pin, FAULT_UNSOLVABLE, 0, "Unsolvable collision or constraint."),
    "kinematic_pivot": (kinematic_pivot, FAULT_UNSOLVABLE, 0, "Unsolvable constraint."),
    "kinematic_spring": (kinematic_spring, FAULT_UNSOLVABLE, 0, "Unsolvable collision or constraint."),
    "zero_mass": (zero_mass, FAULT_BODY_MASS, 0, "Body's mass and moment must be positive to simulate."),
    "huge_force": (huge_force, FAULT_BODY_STATE, 0, "Body's velocity is invalid."),
    "nan_velocity": (nan_velocity, FAULT_BODY_STATE, 0, "Body's position is invalid."),
    "runaway": (runaway, FAULT_BODY_STATE, 3, "Body's position is invalid."),
}


def build(name):
    """(packed world, index of the failing body or -1, first fault, its step, message)."""
    fn, fault, step, message = SCENES[name]
    sp, bad = fn()
    pw = sp.pack()
    return pw, (bad.index if bad is not None else -1), fault, step, message
