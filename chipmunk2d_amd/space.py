"""Host-side mirror of Chipmunk2D's object model for building spaces.

The classes follow the reference's C API (include/chipmunk/cpSpace.h,
cpBody.h, cpShape.h, cpPolyShape.h, the joint headers) closely enough that
scene code reads the same: a `Space` holds `Body`, shape and joint objects,
the setters do what the reference's setters do (including the mass
accumulation of cpBodyAccumulateMassFromShapes and the transform rebuild of
cpBodySetPosition / cpBodySetAngle with libm cos/sin), and `Space.pack()`
turns it into the packed description of include/cpgpu.h.  Stepping happens
in the HIP library (chipmunk2d_amd.gpu).

Every arithmetic expression below keeps the reference's operation order, so
a space built here and the same space built with the reference's C API hold
bit-identical doubles (pinned in tests/test_host_mirror.py).
"""
import math

import numpy as np

from .cabi import (BODY_DYNAMIC, BODY_KINEMATIC, BODY_STATIC, SHAPE_CIRCLE, SHAPE_SEGMENT, SHAPE_POLY,
                   JOINT_PIVOT, JOINT_DAMPED_SPRING, JOINT_SLIDE, JOINT_GEAR, JOINT_PIN, JOINT_GROOVE,
                   JOINT_ROTARY_LIMIT, JOINT_RATCHET, JOINT_SIMPLE_MOTOR, JOINT_DAMPED_ROTARY_SPRING, PackedWorld,
                   HANDLER_DTYPE, WILDCARD)

INF = float("inf")
CP_PI = 3.14159265358979323846264338327950288
DBL_MIN = 2.2250738585072014e-308
ALL_CATEGORIES = 0xFFFFFFFF


# ---------------------------------------------------------------------------
# vector helpers (cpVect.h)

def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1])


def _add(a, b):
    return (a[0] + b[0], a[1] + b[1])


def _mult(a, s):
    return (a[0] * s, a[1] * s)


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1]


def _cross(a, b):
    return a[0] * b[1] - a[1] * b[0]


def _length(a):
    return math.sqrt(_dot(a, a))


def _normalize(a):
    return _mult(a, 1.0 / (_length(a) + DBL_MIN))


def _rperp(a):
    return (a[1], -a[0])


def _lerp(a, b, t):
    return _add(_mult(a, 1.0 - t), _mult(b, t))


def _xpoint(t, p):
    return (t[0] * p[0] + t[2] * p[1] + t[4], t[1] * p[0] + t[3] * p[1] + t[5])


def _xvect(t, v):
    return (t[0] * v[0] + t[2] * v[1], t[1] * v[0] + t[3] * v[1])


# ---------------------------------------------------------------------------
# mass properties (src/chipmunk.c:63-171)

def moment_for_circle(m, r1, r2, offset=(0.0, 0.0)):
    """cpMomentForCircle, chipmunk.c:63-67."""
    return m * (0.5 * (r1 * r1 + r2 * r2) + _dot(offset, offset))


def area_for_circle(r1, r2):
    """cpAreaForCircle, chipmunk.c:69-73."""
    return CP_PI * abs(r1 * r1 - r2 * r2)


def moment_for_segment(m, a, b, r):
    """cpMomentForSegment, chipmunk.c:75-83."""
    offset = _lerp(a, b, 0.5)
    length = _length(_sub(b, a)) + 2.0 * r
    return m * ((length * length + 4.0 * r * r) / 12.0 + _dot(offset, offset))


def area_for_segment(a, b, r):
    """cpAreaForSegment, chipmunk.c:85-89."""
    return r * (CP_PI * r + 2.0 * _length(_sub(a, b)))


def moment_for_poly(m, verts, offset=(0.0, 0.0), r=0.0):
    """cpMomentForPoly, chipmunk.c:91-110."""
    count = len(verts)
    if count == 2:
        return moment_for_segment(m, verts[0], verts[1], 0.0)
    sum1 = 0.0
    sum2 = 0.0
    for i in range(count):
        v1 = _add(verts[i], offset)
        v2 = _add(verts[(i + 1) % count], offset)
        a = _cross(v2, v1)
        b = _dot(v1, v1) + _dot(v1, v2) + _dot(v2, v2)
        sum1 += a * b
        sum2 += a
    return (m * sum1) / (6.0 * sum2)


def area_for_poly(verts, r=0.0):
    """cpAreaForPoly, chipmunk.c:112-126."""
    area = 0.0
    perimeter = 0.0
    count = len(verts)
    for i in range(count):
        v1 = verts[i]
        v2 = verts[(i + 1) % count]
        area += _cross(v1, v2)
        perimeter += _length(_sub(v1, v2))
    return r * (CP_PI * abs(r) + perimeter) + area / 2.0


def centroid_for_poly(verts):
    """cpCentroidForPoly, chipmunk.c:128-143."""
    s = 0.0
    vsum = (0.0, 0.0)
    count = len(verts)
    for i in range(count):
        v1 = verts[i]
        v2 = verts[(i + 1) % count]
        cross = _cross(v1, v2)
        s += cross
        vsum = _add(vsum, _mult(_add(v1, v2), cross))
    return _mult(vsum, 1.0 / (3.0 * s))


def moment_for_box(m, width, height):
    """cpMomentForBox, chipmunk.c:154-158."""
    return m * (width * width + height * height) / 12.0


# ---------------------------------------------------------------------------
# convex hull (cpConvexHull, chipmunk.c:176-274).  The vertex order decides
# the vertex indices that contact feature hashes are built from, so this
# follows the reference's in-place QuickHull step for step.

def _loop_indexes(verts):
    start = end = 0
    mn = mx = verts[0]
    for i in range(1, len(verts)):
        v = verts[i]
        if v[0] < mn[0] or (v[0] == mn[0] and v[1] < mn[1]):
            mn = v
            start = i
        elif v[0] > mx[0] or (v[0] == mx[0] and v[1] > mx[1]):
            mx = v
            end = i
    return start, end


def _qhull_partition(buf, lo, count, a, b, tol):
    if count == 0:
        return 0
    mx = 0.0
    pivot = 0
    delta = _sub(b, a)
    value_tol = tol * _length(delta)
    head = 0
    tail = count - 1
    while head <= tail:
        value = _cross(_sub(buf[lo + head], a), delta)
        if value > value_tol:
            if value > mx:
                mx = value
                pivot = head
            head += 1
        else:
            buf[lo + head], buf[lo + tail] = buf[lo + tail], buf[lo + head]
            tail -= 1
    if pivot != 0:
        buf[lo], buf[lo + pivot] = buf[lo + pivot], buf[lo]
    return head


def _qhull_reduce(tol, buf, lo, count, a, pivot, b, res, ro):
    if count < 0:
        return 0
    if count == 0:
        res[ro] = pivot
        return 1
    left_count = _qhull_partition(buf, lo, count, a, pivot, tol)
    index = _qhull_reduce(tol, buf, lo + 1, left_count - 1, a, buf[lo], pivot, res, ro)
    res[ro + index] = pivot
    index += 1
    right_count = _qhull_partition(buf, lo + left_count, count - left_count, pivot, b, tol)
    # the reference passes verts[left_count] even when right_count == 0 (then unused)
    nxt = buf[lo + left_count] if right_count > 0 else None
    return index + _qhull_reduce(tol, buf, lo + left_count + 1, right_count - 1, pivot, nxt, b, res, ro + index)


def convex_hull(verts, tol=0.0):
    """Counter-clockwise hull in the reference's vertex order; returns (hull, first)."""
    verts = [(float(v[0]), float(v[1])) for v in verts]
    count = len(verts)
    res = list(verts)
    start, end = _loop_indexes(verts)
    if start == end:
        return [res[0]], 0
    res[0], res[start] = res[start], res[0]
    j = start if end == 0 else end
    res[1], res[j] = res[j], res[1]
    a = res[0]
    b = res[1]
    n = _qhull_reduce(tol, res, 2, count - 2, a, b, a, res, 1) + 1
    return res[:n], start


# ---------------------------------------------------------------------------
# object model

class Body:
    """cpBody (src/cpBody.c)."""

    def __init__(self, mass=0.0, moment=0.0, body_type=BODY_DYNAMIC):
        self.type = BODY_DYNAMIC
        self.space = None
        self.index = -1
        self.shapes = []
        self.p = (0.0, 0.0)
        self.v = (0.0, 0.0)
        self.f = (0.0, 0.0)
        self.a = 0.0
        self.w = 0.0
        self.t = 0.0
        self.v_bias = (0.0, 0.0)
        self.w_bias = 0.0
        self.cog = (0.0, 0.0)
        self.transform = (1.0, 0.0, 0.0, 1.0, 0.0, 0.0)
        self.m = self.m_inv = self.i = self.i_inv = 0.0
        # cpBodyInit (cpBody.c:34-71): mass and moment through the setters, then
        # SetTransform at angle 0.
        self._set_mass(mass)
        self._set_moment(moment)
        self._set_transform(self.p, 0.0)
        if body_type != BODY_DYNAMIC:
            self._set_type(body_type)

    @classmethod
    def static(cls):
        return cls(0.0, 0.0, BODY_STATIC)

    @classmethod
    def kinematic(cls):
        return cls(0.0, 0.0, BODY_KINEMATIC)

    def _set_mass(self, mass):
        self.m = mass
        self.m_inv = INF if mass == 0.0 else 1.0 / mass

    def _set_moment(self, moment):
        self.i = moment
        self.i_inv = INF if moment == 0.0 else 1.0 / moment

    def _set_type(self, t):
        # cpBodySetType (cpBody.c:148-201) for a body not yet in a space
        self.type = t
        if t == BODY_DYNAMIC:
            self.m = self.i = 0.0
            self.m_inv = self.i_inv = INF
            self.accumulate_mass()
        else:
            self.m = self.i = INF
            self.m_inv = self.i_inv = 0.0
            self.v = (0.0, 0.0)
            self.w = 0.0

    def _set_transform(self, p, a):
        # SetTransform (cpBody.c:347-357), libm cos/sin like the reference
        rot = (math.cos(a), math.sin(a))
        c = self.cog
        self.transform = (rot[0], rot[1], -rot[1], rot[0],
                          p[0] - (c[0] * rot[0] - c[1] * rot[1]),
                          p[1] - (c[0] * rot[1] + c[1] * rot[0]))

    # -- setters with the reference's side effects
    @property
    def mass(self):
        return self.m

    @mass.setter
    def mass(self, m):
        assert self.type == BODY_DYNAMIC and 0.0 <= m < INF
        self._set_mass(m)

    @property
    def moment(self):
        return self.i

    @moment.setter
    def moment(self, i):
        assert i >= 0.0
        self._set_moment(i)

    @property
    def position(self):
        """cpBodyGetPosition: the transform applied to the origin."""
        return _xpoint(self.transform, (0.0, 0.0))

    @position.setter
    def position(self, pos):
        # cpBodySetPosition (cpBody.c:375-382)
        self.p = _add(_xvect(self.transform, self.cog), (float(pos[0]), float(pos[1])))
        self._set_transform(self.p, self.a)

    @property
    def angle(self):
        return self.a

    @angle.setter
    def angle(self, a):
        # cpBodySetAngle (cpBody.c:433-439)
        self.a = float(a)
        self._set_transform(self.p, self.a)

    @property
    def velocity(self):
        return self.v

    @velocity.setter
    def velocity(self, v):
        self.v = (float(v[0]), float(v[1]))

    @property
    def angular_velocity(self):
        return self.w

    @angular_velocity.setter
    def angular_velocity(self, w):
        self.w = float(w)

    @property
    def force(self):
        return self.f

    @force.setter
    def force(self, f):
        self.f = (float(f[0]), float(f[1]))

    @property
    def torque(self):
        return self.t

    @torque.setter
    def torque(self, t):
        self.t = float(t)

    @property
    def center_of_gravity(self):
        return self.cog

    @center_of_gravity.setter
    def center_of_gravity(self, cog):
        self.cog = (float(cog[0]), float(cog[1]))

    def accumulate_mass(self):
        """cpBodyAccumulateMassFromShapes (cpBody.c:207-239)."""
        if self.type != BODY_DYNAMIC:
            return
        self.m = self.i = 0.0
        self.cog = (0.0, 0.0)
        pos = self.position
        for s in self.shapes:
            m = s.mass_m
            if m > 0.0:
                msum = self.m + m
                d = _sub(self.cog, s.mass_cog)
                self.i += m * s.mass_i + _dot(d, d) * (m * self.m) / msum
                self.cog = _lerp(self.cog, s.mass_cog, m / msum)
                self.m = msum
        self.m_inv = 1.0 / self.m
        self.i_inv = 1.0 / self.i
        self.position = pos


class Shape:
    """cpShape (src/cpShape.c:22-130)."""

    type = -1

    def __init__(self, body, mass_i, mass_cog, area):
        self.body = body
        self.space = None
        self.index = -1
        self.hashid = 0
        self.e = 0.0
        self.u = 0.0
        self.surface_v = (0.0, 0.0)
        self.group = 0
        self.categories = ALL_CATEGORIES
        self.mask = ALL_CATEGORIES
        self.sensor = False
        self.collision_type = 0
        self.mass_m = 0.0
        self.mass_i = mass_i
        self.mass_cog = mass_cog
        self.area = area
        body.shapes.insert(0, self)

    elasticity = property(lambda s: s.e, lambda s, v: setattr(s, "e", float(v)))
    friction = property(lambda s: s.u, lambda s, v: setattr(s, "u", float(v)))

    @property
    def surface_velocity(self):
        return self.surface_v

    @surface_velocity.setter
    def surface_velocity(self, v):
        self.surface_v = (float(v[0]), float(v[1]))

    def set_filter(self, group=0, categories=ALL_CATEGORIES, mask=ALL_CATEGORIES):
        self.group, self.categories, self.mask = int(group), int(categories), int(mask)

    @property
    def mass(self):
        return self.mass_m

    @mass.setter
    def mass(self, m):
        # cpShapeSetMass (cpShape.c:96-102)
        self.mass_m = float(m)
        self.body.accumulate_mass()

    @property
    def density(self):
        return self.mass_m / self.area

    @density.setter
    def density(self, d):
        # cpShapeSetDensity (cpShape.c:105)
        self.mass = float(d) * self.area


class Circle(Shape):
    """cpCircleShape (cpShape.c:285-356)."""
    type = SHAPE_CIRCLE

    def __init__(self, body, radius, offset=(0.0, 0.0)):
        offset = (float(offset[0]), float(offset[1]))
        radius = float(radius)
        super().__init__(body, moment_for_circle(1.0, 0.0, radius, (0.0, 0.0)), offset, area_for_circle(0.0, radius))
        self.c = offset
        self.r = radius


class Segment(Shape):
    """cpSegmentShape (cpShape.c:370-508)."""
    type = SHAPE_SEGMENT

    def __init__(self, body, a, b, radius):
        a = (float(a[0]), float(a[1]))
        b = (float(b[0]), float(b[1]))
        radius = float(radius)
        super().__init__(body, moment_for_box(1.0, _length(_sub(a, b)) + 2.0 * radius, 2.0 * radius),
                         _lerp(a, b, 0.5), area_for_segment(a, b, radius))
        self.a = a
        self.b = b
        self.n = _rperp(_normalize(_sub(b, a)))
        self.r = radius
        self.a_tangent = (0.0, 0.0)
        self.b_tangent = (0.0, 0.0)

    def set_neighbors(self, prev, nxt):
        """cpSegmentShapeSetNeighbors (cpShape.c:540-547)."""
        self.a_tangent = _sub(prev, self.a)
        self.b_tangent = _sub(nxt, self.b)


class Poly(Shape):
    """cpPolyShape (cpPolyShape.c:147-250): hull of the transformed vertices."""
    type = SHAPE_POLY

    def __init__(self, body, verts, transform=None, radius=0.0, raw=False):
        verts = [(float(v[0]), float(v[1])) for v in verts]
        if not raw:
            if transform is not None:
                verts = [_xpoint(transform, v) for v in verts]
            verts, _ = convex_hull(verts, 0.0)
        radius = float(radius)
        centroid = centroid_for_poly(verts)
        super().__init__(body, moment_for_poly(1.0, verts, (-centroid[0], -centroid[1]), radius), centroid,
                         area_for_poly(verts, radius))
        self.r = radius
        self.verts = verts
        count = len(verts)
        self.normals = []
        for i in range(count):
            a = verts[(i - 1 + count) % count]
            b = verts[i]
            self.normals.append(_normalize(_rperp(_sub(b, a))))


def Box(body, width, height, radius=0.0):
    """cpBoxShapeNew (cpPolyShape.c:226-251)."""
    hw = width / 2.0
    hh = height / 2.0
    l, b, r, t = -hw, -hh, hw, hh
    return Poly(body, [(r, b), (r, t), (l, t), (l, b)], radius=radius, raw=True)


class Constraint:
    """cpConstraint defaults (cpConstraint.c:37-57)."""
    type = -1

    def __init__(self, a, b):
        self.a = a
        self.b = b
        self.space = None
        self.index = -1
        self.max_force = INF
        self.error_bias = math.pow(1.0 - 0.1, 60.0)
        self.max_bias = INF
        self.collide_bodies = True

    def params(self):
        raise NotImplementedError

    def jacc(self):
        return (0.0, 0.0)


def _world_to_local(body, p):
    """cpBodyWorldToLocal (cpBody.c:531-534) through cpTransformRigidInverse (cpTransform.h:80-88)."""
    t = body.transform
    inv = (t[3], -t[1], -t[2], t[0], t[2] * t[5] - t[4] * t[3], t[4] * t[1] - t[0] * t[5])
    return _xpoint(inv, p)


class PivotJoint(Constraint):
    """cpPivotJoint (src/cpPivotJoint.c)."""
    type = JOINT_PIVOT

    def __init__(self, a, b, anchor_a, anchor_b=None):
        super().__init__(a, b)
        if anchor_b is None:
            # cpPivotJointNew (cpPivotJoint.c:101-107): one world-space pivot
            pivot = anchor_a
            anchor_a = _world_to_local(a, pivot)
            anchor_b = _world_to_local(b, pivot)
        self.anchor_a = (float(anchor_a[0]), float(anchor_a[1]))
        self.anchor_b = (float(anchor_b[0]), float(anchor_b[1]))
        self.j_acc = (0.0, 0.0)

    def params(self):
        return [*self.anchor_a, *self.anchor_b]

    def jacc(self):
        return self.j_acc


class DampedSpring(Constraint):
    """cpDampedSpring (src/cpDampedSpring.c) with the default spring force."""
    type = JOINT_DAMPED_SPRING

    def __init__(self, a, b, anchor_a, anchor_b, rest_length, stiffness, damping):
        super().__init__(a, b)
        self.anchor_a = (float(anchor_a[0]), float(anchor_a[1]))
        self.anchor_b = (float(anchor_b[0]), float(anchor_b[1]))
        self.rest_length, self.stiffness, self.damping = float(rest_length), float(stiffness), float(damping)

    def params(self):
        return [*self.anchor_a, *self.anchor_b, self.rest_length, self.stiffness, self.damping]


class SlideJoint(Constraint):
    """cpSlideJoint (src/cpSlideJoint.c)."""
    type = JOINT_SLIDE

    def __init__(self, a, b, anchor_a, anchor_b, mn, mx):
        super().__init__(a, b)
        self.anchor_a = (float(anchor_a[0]), float(anchor_a[1]))
        self.anchor_b = (float(anchor_b[0]), float(anchor_b[1]))
        self.min, self.max = float(mn), float(mx)

    def params(self):
        return [*self.anchor_a, *self.anchor_b, self.min, self.max]


class GearJoint(Constraint):
    """cpGearJoint (src/cpGearJoint.c)."""
    type = JOINT_GEAR

    def __init__(self, a, b, phase, ratio):
        super().__init__(a, b)
        self.phase, self.ratio = float(phase), float(ratio)
        self.ratio_inv = 1.0 / self.ratio

    def params(self):
        return [self.phase, self.ratio, self.ratio_inv]


class PinJoint(Constraint):
    """cpPinJoint (src/cpPinJoint.c): the distance between the anchors is held at the
    one they have when the joint is made (cpPinJointInit, cpPinJoint.c:98-116)."""
    type = JOINT_PIN

    def __init__(self, a, b, anchor_a, anchor_b):
        super().__init__(a, b)
        self.anchor_a = (float(anchor_a[0]), float(anchor_a[1]))
        self.anchor_b = (float(anchor_b[0]), float(anchor_b[1]))
        self.dist = _length(_sub(_xpoint(b.transform, self.anchor_b), _xpoint(a.transform, self.anchor_a)))
        self.jn_acc = 0.0

    def params(self):
        return [*self.anchor_a, *self.anchor_b, self.dist]

    def jacc(self):
        return (self.jn_acc, 0.0)


class GrooveJoint(Constraint):
    """cpGrooveJoint (src/cpGrooveJoint.c): body B's anchor slides along the groove
    from groove_a to groove_b on body A."""
    type = JOINT_GROOVE

    def __init__(self, a, b, groove_a, groove_b, anchor_b):
        super().__init__(a, b)
        self.groove_a = (float(groove_a[0]), float(groove_a[1]))
        self.groove_b = (float(groove_b[0]), float(groove_b[1]))
        self.anchor_b = (float(anchor_b[0]), float(anchor_b[1]))
        self.j_acc = (0.0, 0.0)

    @property
    def groove_n(self):
        # cpvperp(cpvnormalize(groove_b - groove_a)) (cpGrooveJoint.c:126, cpVect.h:159-163)
        n = _normalize((self.groove_b[0] - self.groove_a[0], self.groove_b[1] - self.groove_a[1]))
        return (-n[1], n[0])

    def params(self):
        return [*self.groove_a, *self.groove_b, *self.anchor_b, *self.groove_n]

    def jacc(self):
        return self.j_acc


class RotaryLimitJoint(Constraint):
    """cpRotaryLimitJoint (src/cpRotaryLimitJoint.c): b.angle - a.angle kept in [min, max]."""
    type = JOINT_ROTARY_LIMIT

    def __init__(self, a, b, mn, mx):
        super().__init__(a, b)
        self.min, self.max = float(mn), float(mx)
        self.j_acc = 0.0

    def params(self):
        return [self.min, self.max]

    def jacc(self):
        return (self.j_acc, 0.0)


class RatchetJoint(Constraint):
    """cpRatchetJoint (src/cpRatchetJoint.c): a rotary ratchet with notches every
    `ratchet` radians from `phase`; the angle starts at b.angle - a.angle
    (cpRatchetJointInit, cpRatchetJoint.c:111-124) and is carried between steps."""
    type = JOINT_RATCHET

    def __init__(self, a, b, phase, ratchet):
        super().__init__(a, b)
        self.phase, self.ratchet = float(phase), float(ratchet)
        self.angle = b.a - a.a
        self.j_acc = 0.0

    def params(self):
        return [self.phase, self.ratchet]

    def jacc(self):
        return (self.j_acc, self.angle)


class SimpleMotor(Constraint):
    """cpSimpleMotor (src/cpSimpleMotor.c): holds b.w - a.w at -rate."""
    type = JOINT_SIMPLE_MOTOR

    def __init__(self, a, b, rate):
        super().__init__(a, b)
        self.rate = float(rate)
        self.j_acc = 0.0

    def params(self):
        return [self.rate]

    def jacc(self):
        return (self.j_acc, 0.0)


class DampedRotarySpring(Constraint):
    """cpDampedRotarySpring (src/cpDampedRotarySpring.c) with the default spring torque."""
    type = JOINT_DAMPED_ROTARY_SPRING

    def __init__(self, a, b, rest_angle, stiffness, damping):
        super().__init__(a, b)
        self.rest_angle, self.stiffness, self.damping = float(rest_angle), float(stiffness), float(damping)

    def params(self):
        return [self.rest_angle, self.stiffness, self.damping]


class Space:
    """cpSpace (src/cpSpace.c:119-178 defaults)."""

    def __init__(self):
        self.iterations = 10
        self.gravity = (0.0, 0.0)
        self.damping = 1.0
        self.collision_slop = 0.1
        self.collision_bias = math.pow(1.0 - 0.1, 60.0)
        self.collision_persistence = 3
        self.bodies = []
        self.shapes = []
        self.constraints = []
        self.static_body = Body.static()
        self._static_added = False
        # collision handlers: (type_a, type_b, events); WILDCARD as type_b makes a
        # wildcard handler of type_a, as both the default handler
        self.handlers = []

    def add_collision_handler(self, type_a, type_b, events):
        """cpSpaceAddCollisionHandler (cpSpace.c:382-399) with the CPGPU_EVENT_* callbacks it has."""
        self.handlers.append((int(type_a), int(type_b), int(events)))

    def add_wildcard_handler(self, type_a, events):
        """cpSpaceAddWildcardHandler (cpSpace.c:401-413)."""
        self.handlers.append((int(type_a), WILDCARD, int(events)))

    def add_default_handler(self, events):
        """cpSpaceAddDefaultCollisionHandler (cpSpace.c:373-380)."""
        self.handlers.append((WILDCARD, WILDCARD, int(events)))

    def add(self, obj):
        """cpSpaceAddBody / AddShape / AddConstraint (cpSpace.c:417-474)."""
        if isinstance(obj, Body):
            assert obj.space is None, "body already added to a space"
            obj.space = self
            self.bodies.append(obj)
        elif isinstance(obj, Shape):
            assert obj.space is None, "shape already added to a space"
            body = obj.body
            if body is self.static_body and not self._static_added:
                self.add(body)
                self._static_added = True
            assert body.space is self, "the shape's body must be added to the space before the shape"
            obj.hashid = len(self.shapes)
            obj.space = self
            self.shapes.append(obj)
        elif isinstance(obj, Constraint):
            assert obj.space is None
            for b in (obj.a, obj.b):
                if b is self.static_body and not self._static_added:
                    self.add(b)
                    self._static_added = True
            obj.space = self
            self.constraints.append(obj)
        else:
            raise TypeError(obj)
        return obj

    def pack(self):
        """The packed description (include/cpgpu.h) of this one space."""
        nb, ns, nc = len(self.bodies), len(self.shapes), len(self.constraints)
        for k, b in enumerate(self.bodies):
            b.index = k
        a = {}
        a["space"] = [dict(iterations=int(self.iterations), collision_persistence=int(self.collision_persistence),
                           gravity_x=float(self.gravity[0]), gravity_y=float(self.gravity[1]),
                           damping=float(self.damping), collision_slop=float(self.collision_slop),
                           collision_bias=float(self.collision_bias), curr_dt=0.0, stamp=0)]
        a["world_body_start"] = np.array([0, nb], dtype=np.int32)
        a["world_shape_start"] = np.array([0, ns], dtype=np.int32)
        a["world_constraint_start"] = np.array([0, nc], dtype=np.int32)
        a["body_type"] = np.array([b.type for b in self.bodies], dtype=np.int32).reshape(nb)
        for name, attr in [("m", "m"), ("m_inv", "m_inv"), ("i", "i"), ("i_inv", "i_inv"), ("a", "a"),
                           ("w", "w"), ("t", "t"), ("w_bias", "w_bias")]:
            a["body_" + name] = np.array([getattr(b, attr) for b in self.bodies], dtype=np.float64).reshape(nb)
        for name in ["cog", "p", "v", "f", "v_bias"]:
            a["body_" + name] = np.array([getattr(b, name) for b in self.bodies], dtype=np.float64).reshape(nb, 2)
        a["body_transform"] = np.array([b.transform for b in self.bodies], dtype=np.float64).reshape(nb, 6)
        a["shape_type"] = np.array([s.type for s in self.shapes], dtype=np.int32).reshape(ns)
        a["shape_body"] = np.array([s.body.index for s in self.shapes], dtype=np.int32).reshape(ns)
        a["shape_hashid"] = np.array([s.hashid for s in self.shapes], dtype=np.uint64).reshape(ns)
        a["shape_e"] = np.array([s.e for s in self.shapes], dtype=np.float64).reshape(ns)
        a["shape_u"] = np.array([s.u for s in self.shapes], dtype=np.float64).reshape(ns)
        a["shape_surface_v"] = np.array([s.surface_v for s in self.shapes], dtype=np.float64).reshape(ns, 2)
        a["shape_group"] = np.array([s.group for s in self.shapes], dtype=np.uint64).reshape(ns)
        a["shape_categories"] = np.array([s.categories for s in self.shapes], dtype=np.uint32).reshape(ns)
        a["shape_mask"] = np.array([s.mask for s in self.shapes], dtype=np.uint32).reshape(ns)
        a["shape_sensor"] = np.array([1 if s.sensor else 0 for s in self.shapes], dtype=np.uint8).reshape(ns)
        a["shape_ctype"] = np.array([s.collision_type for s in self.shapes], dtype=np.uint32).reshape(ns)
        a["handlers"] = np.array([(-1, ta, tb, ev) for ta, tb, ev in self.handlers], dtype=HANDLER_DTYPE)
        a["shape_radius"] = np.array([s.r for s in self.shapes], dtype=np.float64).reshape(ns)
        geom = np.zeros((ns, 10), dtype=np.float64)
        pstart = np.zeros(ns, dtype=np.int32)
        pcount = np.zeros(ns, dtype=np.int32)
        pv, pn = [], []
        for k, s in enumerate(self.shapes):
            if s.type == SHAPE_CIRCLE:
                geom[k, 0:2] = s.c
            elif s.type == SHAPE_SEGMENT:
                geom[k] = [*s.a, *s.b, *s.n, *s.a_tangent, *s.b_tangent]
            else:
                pstart[k] = len(pv)
                pcount[k] = len(s.verts)
                pv.extend(s.verts)
                pn.extend(s.normals)
        a["shape_geom"] = geom
        a["shape_poly_start"] = pstart
        a["shape_poly_count"] = pcount
        a["poly_verts"] = np.array(pv, dtype=np.float64).reshape(len(pv), 2)
        a["poly_normals"] = np.array(pn, dtype=np.float64).reshape(len(pn), 2)
        a["cons_type"] = np.array([c.type for c in self.constraints], dtype=np.int32).reshape(nc)
        a["cons_a"] = np.array([c.a.index for c in self.constraints], dtype=np.int32).reshape(nc)
        a["cons_b"] = np.array([c.b.index for c in self.constraints], dtype=np.int32).reshape(nc)
        a["cons_max_force"] = np.array([c.max_force for c in self.constraints], dtype=np.float64).reshape(nc)
        a["cons_error_bias"] = np.array([c.error_bias for c in self.constraints], dtype=np.float64).reshape(nc)
        a["cons_max_bias"] = np.array([c.max_bias for c in self.constraints], dtype=np.float64).reshape(nc)
        a["cons_collide_bodies"] = np.array([1 if c.collide_bodies else 0 for c in self.constraints],
                                            dtype=np.int32).reshape(nc)
        params = np.zeros((nc, 8), dtype=np.float64)
        jacc = np.zeros((nc, 2), dtype=np.float64)
        for k, c in enumerate(self.constraints):
            p = c.params()
            params[k, :len(p)] = p
            jacc[k] = c.jacc()
        a["cons_params"] = params
        a["cons_jacc"] = jacc
        return PackedWorld(a)
