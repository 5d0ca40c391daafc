"""ctypes binding of the gfx950 step library (include/cpgpu.h).

The library is built in-tree (chipmunk2d_amd/libcpgpu.so, see build.py).  If
it is missing this module raises on import of `lib()`: there is no CPU
fallback for the step.
"""
import atexit
import ctypes
import os
import weakref

import numpy as np

from .cabi import (SpaceParams, Stats, WorldDesc, ptr, FP64, CallbackArbiter, CallbackDecision, ARB_FIRST_COLLISION,
                   ARB_IGNORE, EVENT_BEGIN, EVENT_PRE_SOLVE)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcpgpu.so")

_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p

EXPORTS = [
    "cpgpuWorldCreate", "cpgpuWorldCreateFrom", "cpgpuWorldDestroy", "cpgpuWorldWriteBodiesAt",
    "cpgpuWorldStepBegin", "cpgpuWorldReadCallbacks", "cpgpuWorldWriteDecisions", "cpgpuWorldStepEnd",
    "cpgpuWorldReadSeparations",
    "cpgpuWorldSetSpaceParams", "cpgpuWorldStep", "cpgpuWorldStepN", "cpgpuWorldSync",
    "cpgpuWorldGetFaults",
    "cpgpuWorldReadBodies", "cpgpuWorldWriteBodies", "cpgpuWorldGetStats", "cpgpuWorldReadArbiters",
    "cpgpuWorldReadConstraints", "cpgpuWorldReadPairs", "cpgpuWorldShapesCollide",
    "cpgpuWorldSetTiming", "cpgpuWorldGetTiming", "cpgpuWorldGetSolveCounts",
    "cpgpuWorldGetBuckets", "cpgpuWorldSetBuckets", "cpgpuWorldAutotune", "cpgpuWorldSetGridSchedule", "cpgpuWorldSetGraph",
    "cpgpuWorldSaveState", "cpgpuWorldLoadState",
    "cpgpuWorldQueryPointNearest", "cpgpuWorldQuerySegmentFirst", "cpgpuWorldQueryPointAll", "cpgpuWorldQuerySegmentAll",
    "cpgpuWorldQueryBBAll", "cpgpuWorldQueryShapes",
    "cpgpuWorldTapeBegin", "cpgpuWorldTapeEnd",
    "cpgpuWorldBackward", "cpgpuWorldGetJointGrads", "cpgpuGetLastError", "cpgpuVersion",
]


GRID_FLOW, GRID_PHASES = 0, 1


class CpgpuError(RuntimeError):
    pass


class ShapeFilter(ctypes.Structure):
    """cpgpuShapeFilter (cpShapeFilter: group, categories, mask)."""
    _fields_ = [("group", ctypes.c_uint64), ("categories", ctypes.c_uint32), ("mask", ctypes.c_uint32)]


def _filters(filters, n):
    """None (every shape) or one (group, categories, mask) per query."""
    if filters is None:
        return None
    arr = (ShapeFilter * max(n, 1))()
    for k, (g, c, m) in enumerate(filters):
        arr[k] = ShapeFilter(g, c, m)
    return arr


def _f64(x, shape):
    return np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), shape))


def _space(space, n):
    return None if space is None else np.ascontiguousarray(np.broadcast_to(np.asarray(space, np.int32), (n,)))


def lib():
    """The loaded step library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CpgpuError("%s is missing: build it with `python -m chipmunk2d_amd.build`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.cpgpuWorldCreate.argtypes = [ctypes.POINTER(WorldDesc), ctypes.c_int32, ctypes.POINTER(_vp)]
        L.cpgpuWorldCreateFrom.argtypes = [ctypes.POINTER(WorldDesc), ctypes.c_int32, _vp, _ip, ctypes.POINTER(_vp)]
        L.cpgpuWorldWriteBodiesAt.argtypes = [_vp, ctypes.c_int32, _ip] + [_dp] * 7
        L.cpgpuWorldSetSpaceParams.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(SpaceParams)]
        L.cpgpuWorldDestroy.argtypes = [_vp]
        L.cpgpuWorldDestroy.restype = None
        L.cpgpuWorldStep.argtypes = [_vp, ctypes.c_double]
        L.cpgpuWorldStepN.argtypes = [_vp, ctypes.c_double, ctypes.c_int32]
        L.cpgpuWorldSync.argtypes = [_vp]
        L.cpgpuWorldGetFaults.argtypes = [_vp, _ip, _ip]
        L.cpgpuWorldReadBodies.argtypes = [_vp] + [_dp] * 7
        L.cpgpuWorldWriteBodies.argtypes = [_vp] + [_dp] * 4
        L.cpgpuWorldGetStats.argtypes = [_vp, ctypes.POINTER(Stats)]
        L.cpgpuWorldReadArbiters.argtypes = [_vp, ctypes.c_int64, _i64p, _ip, _ip, _ip, _ip, _dp, _dp, _u64p]
        L.cpgpuWorldReadConstraints.argtypes = [_vp, _dp, _ip]
        L.cpgpuWorldReadPairs.argtypes = [_vp, ctypes.c_int64, _i64p, _ip]
        L.cpgpuWorldShapesCollide.argtypes = [_vp, ctypes.c_int64, _ip, _ip, _dp]
        L.cpgpuWorldSetTiming.argtypes = [_vp, ctypes.c_int32]
        L.cpgpuWorldSetGridSchedule.argtypes = [_vp, ctypes.c_int32]
        L.cpgpuWorldSetGraph.argtypes = [_vp, ctypes.c_int32]
        L.cpgpuWorldSaveState.argtypes = [_vp, _vp, ctypes.c_int64, _i64p]
        L.cpgpuWorldLoadState.argtypes = [_vp, _vp, ctypes.c_int64]
        _fp = ctypes.POINTER(ShapeFilter)
        L.cpgpuWorldQueryPointNearest.argtypes = [_vp, ctypes.c_int32, _ip, _dp, _dp, _fp, _ip, _dp]
        L.cpgpuWorldQuerySegmentFirst.argtypes = [_vp, ctypes.c_int32, _ip, _dp, _dp, _dp, _fp, _ip, _dp]
        L.cpgpuWorldQueryPointAll.argtypes = [_vp, ctypes.c_int32, _ip, _dp, _dp, _fp, ctypes.c_int64, _i64p, _ip, _dp]
        L.cpgpuWorldQuerySegmentAll.argtypes = [_vp, ctypes.c_int32, _ip, _dp, _dp, _dp, _fp, ctypes.c_int64, _i64p,
                                                _ip, _dp]
        L.cpgpuWorldQueryBBAll.argtypes = [_vp, ctypes.c_int32, _ip, _dp, _fp, ctypes.c_int64, _i64p, _ip]
        L.cpgpuWorldQueryShapes.argtypes = [_vp, ctypes.c_int32, _ip, ctypes.c_int64, _i64p, _ip, _ip, _dp]
        L.cpgpuWorldStepBegin.argtypes = [_vp, ctypes.c_double, _i64p]
        L.cpgpuWorldReadCallbacks.argtypes = [_vp, ctypes.c_int64, _i64p, ctypes.POINTER(CallbackArbiter)]
        L.cpgpuWorldWriteDecisions.argtypes = [_vp, ctypes.c_int64, ctypes.POINTER(CallbackDecision)]
        L.cpgpuWorldStepEnd.argtypes = [_vp]
        L.cpgpuWorldReadSeparations.argtypes = [_vp, ctypes.c_int64, _i64p, _ip]
        L.cpgpuWorldGetTiming.argtypes = [_vp, ctypes.c_int32, _ip, _dp, _i64p, ctypes.c_char_p, ctypes.c_int32]
        L.cpgpuWorldGetSolveCounts.argtypes = [_vp, _i64p, _i64p]
        L.cpgpuWorldGetBuckets.argtypes = [_vp, _ip, _ip, _ip, _ip]
        L.cpgpuWorldSetBuckets.argtypes = [_vp, ctypes.c_int32, _ip, _ip]
        L.cpgpuWorldAutotune.argtypes = [_vp, ctypes.c_double, ctypes.c_int32, _ip]
        L.cpgpuWorldTapeBegin.argtypes = [_vp, ctypes.c_int32]
        L.cpgpuWorldTapeEnd.argtypes = [_vp]
        L.cpgpuWorldBackward.argtypes = [_vp] + [_dp] * 11
        L.cpgpuWorldGetJointGrads.argtypes = [_vp, _dp]
        L.cpgpuGetLastError.restype = ctypes.c_char_p
        L.cpgpuVersion.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CpgpuError("cpgpu error %d: %s" % (rc, lib().cpgpuGetLastError().decode()))


# Worlds still open when the interpreter exits are destroyed by an atexit
# handler, while the HIP runtime is certainly alive, rather than by __del__
# during module teardown, whose order against the runtime's (and a
# profiler's) own exit handlers is not fixed.
_live = weakref.WeakSet()


@atexit.register
def _close_live_worlds():
    for w in list(_live):
        w.close()


class GPUWorld:
    """One or more packed spaces resident on the current HIP device."""

    def __init__(self, packed, precision=FP64):
        self.L = lib()
        self.packed = packed
        self.desc = packed.desc()
        h = _vp()
        _check(self.L.cpgpuWorldCreate(ctypes.byref(self.desc), precision, ctypes.byref(h)))
        self.h = h
        _live.add(self)

    @classmethod
    def from_old(cls, packed, old, shape_from, precision=FP64):
        """cpgpuWorldCreateFrom: a world of `packed` that keeps the collision state of the
        GPUWorld `old` (shape_from[s]: old index of shape s, -1 for a new shape)."""
        self = cls.__new__(cls)
        self.L = lib()
        self.packed = packed
        self.desc = packed.desc()
        sf = np.ascontiguousarray(shape_from, dtype=np.int32)
        h = _vp()
        _check(self.L.cpgpuWorldCreateFrom(ctypes.byref(self.desc), precision, old.h, ptr(sf, ctypes.c_int32),
                                           ctypes.byref(h)))
        self.h = h
        _live.add(self)
        return self

    def close(self):
        if getattr(self, "h", None):
            self.L.cpgpuWorldDestroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, dt, n=1):
        if n == 1:
            _check(self.L.cpgpuWorldStep(self.h, dt))
        else:
            _check(self.L.cpgpuWorldStepN(self.h, dt, n))

    def sync(self):
        _check(self.L.cpgpuWorldSync(self.h))

    def faults(self):
        """(CPGPU_FAULT_* bits, lowest failing body or -1) as of the last
        synchronising call: the checks a debug build of the reference makes
        during cpSpaceStep."""
        f, b = ctypes.c_int32(0), ctypes.c_int32(0)
        _check(self.L.cpgpuWorldGetFaults(self.h, ctypes.byref(f), ctypes.byref(b)))
        return f.value, b.value

    # ---- collision handler callbacks ----

    def step_begin(self, dt):
        """The first half of a step: returns the arbiters whose begin or preSolve is due
        (a ctypes array of CallbackArbiter)."""
        n = ctypes.c_int64()
        _check(self.L.cpgpuWorldStepBegin(self.h, dt, ctypes.byref(n)))
        out = (CallbackArbiter * max(n.value, 1))()
        m = ctypes.c_int64()
        if n.value:
            _check(self.L.cpgpuWorldReadCallbacks(self.h, n.value, ctypes.byref(m), out))
        return out[:n.value]

    def step_end(self, decisions=()):
        """Hand back the decisions (CallbackDecision) and finish the step."""
        if decisions:
            arr = (CallbackDecision * len(decisions))(*decisions)
            _check(self.L.cpgpuWorldWriteDecisions(self.h, len(decisions), arr))
        _check(self.L.cpgpuWorldStepEnd(self.h))

    def step_with_callbacks(self, dt, begin=None, pre_solve=None):
        """One step, its due arbiters answered as cpSpaceCollideShapes calls a handler
        (cpSpaceStep.c:255-266): begin on a first collision (false, or an ignore, ignores
        the pair until it separates), then preSolve unless ignored, which sees what begin
        set.  begin / pre_solve(view, decision) edit the decision (solve, ignore, e, u,
        surface_vr) in place; None answers with the defaults."""
        decisions = []
        for v in self.step_begin(dt):
            d = CallbackDecision(slot=v.slot, solve=1, ignore=0, e=v.e, u=v.u)
            d.surface_vr[0], d.surface_vr[1] = v.surface_vr[0], v.surface_vr[1]
            ignored = False
            if v.state == ARB_FIRST_COLLISION and (v.events & EVENT_BEGIN) and begin:
                begin(v, d)
                if not d.solve or d.ignore:
                    d.solve, d.ignore, ignored = 0, 1, True
            if not ignored and (v.events & EVENT_PRE_SOLVE) and pre_solve:
                v.e, v.u = d.e, d.u
                v.surface_vr[0], v.surface_vr[1] = d.surface_vr[0], d.surface_vr[1]
                d.solve = 1
                pre_solve(v, d)
            decisions.append(d)
        self.step_end(decisions)

    def separations(self):
        """Pairs whose arbiter separated in the last step, among handler pairs with a
        separate event: (n, 3) int32 {space, shape_a, shape_b}."""
        n = ctypes.c_int64()
        _check(self.L.cpgpuWorldReadSeparations(self.h, 0, ctypes.byref(n), None))
        out = np.zeros((n.value, 3), np.int32)
        m = ctypes.c_int64()
        if n.value:
            _check(self.L.cpgpuWorldReadSeparations(self.h, n.value, ctypes.byref(m), ptr(out, ctypes.c_int32)))
        return out

    def bodies(self):
        from types import SimpleNamespace
        n = self.packed.n_bodies
        s = SimpleNamespace(p=np.zeros((n, 2)), v=np.zeros((n, 2)), a=np.zeros(n), w=np.zeros(n),
                            v_bias=np.zeros((n, 2)), w_bias=np.zeros(n), transform=np.zeros((n, 6)))
        s.FIELDS = ("p", "v", "a", "w", "v_bias", "w_bias", "transform")
        _check(self.L.cpgpuWorldReadBodies(self.h, *[ptr(getattr(s, f), ctypes.c_double) for f in s.FIELDS]))
        return s

    def write_bodies(self, v=None, w=None, f=None, t=None):
        args = [ptr(np.ascontiguousarray(x, dtype=np.float64), ctypes.c_double) if x is not None else None
                for x in (v, w, f, t)]
        _check(self.L.cpgpuWorldWriteBodies(self.h, *args))

    def write_bodies_at(self, index, p=None, v=None, f=None, a=None, w=None, t=None, transform=None):
        """cpgpuWorldWriteBodiesAt: the listed bodies' fields (rows in index order)."""
        index = np.ascontiguousarray(index, dtype=np.int32)
        args = [None if x is None else ptr(np.ascontiguousarray(x, dtype=np.float64), ctypes.c_double)
                for x in (p, v, f, a, w, t, transform)]
        _check(self.L.cpgpuWorldWriteBodiesAt(self.h, len(index), ptr(index, ctypes.c_int32), *args))

    def set_space_params(self, k, params):
        """cpgpuWorldSetSpaceParams for space k from a dict of cpgpuSpaceParams fields."""
        sp = SpaceParams(**params)
        _check(self.L.cpgpuWorldSetSpaceParams(self.h, k, ctypes.byref(sp)))

    def stats(self):
        st = Stats()
        _check(self.L.cpgpuWorldGetStats(self.h, ctypes.byref(st)))
        return st.as_dict()

    def arbiters(self):
        from types import SimpleNamespace
        n = ctypes.c_int64()
        _check(self.L.cpgpuWorldReadArbiters(self.h, 0, ctypes.byref(n), None, None, None, None, None, None, None))
        n = n.value
        A = SimpleNamespace(n=n, shapes=np.zeros((n, 2), np.int32), colour=np.zeros(n, np.int32),
                            state=np.zeros(n, np.int32), count=np.zeros(n, np.int32), normal=np.zeros((n, 2)),
                            contact=np.zeros((n, 2, 12)), hash=np.zeros((n, 2), np.uint64))
        m = ctypes.c_int64()
        _check(self.L.cpgpuWorldReadArbiters(self.h, n, ctypes.byref(m), ptr(A.shapes, ctypes.c_int32),
                                             ptr(A.colour, ctypes.c_int32), ptr(A.state, ctypes.c_int32),
                                             ptr(A.count, ctypes.c_int32), ptr(A.normal, ctypes.c_double),
                                             ptr(A.contact, ctypes.c_double), ptr(A.hash, ctypes.c_uint64)))
        return A

    def pairs(self):
        """Candidate pairs of the last step as an (n, 2) int32 array of (lower, higher) global
        shape indices, sorted."""
        n = ctypes.c_int64()
        _check(self.L.cpgpuWorldReadPairs(self.h, 0, ctypes.byref(n), None))
        out = np.zeros((n.value, 2), np.int32)
        m = ctypes.c_int64()
        _check(self.L.cpgpuWorldReadPairs(self.h, n.value, ctypes.byref(m), ptr(out, ctypes.c_int32)))
        return out

    def shapes_collide(self, pairs):
        """cpShapesCollide for each (a, b) row of `pairs` on the shapes' cached geometry: returns
        (count [n] int32, sets [n, 12]: normal, then per point pointA, pointB, distance)."""
        pairs = np.ascontiguousarray(pairs, dtype=np.int32).reshape(-1, 2)
        n = len(pairs)
        count = np.zeros(n, np.int32)
        sets = np.zeros((n, 12))
        _check(self.L.cpgpuWorldShapesCollide(self.h, n, ptr(pairs, ctypes.c_int32), ptr(count, ctypes.c_int32), ptr(sets, ctypes.c_double)))
        return count, sets

    def save_state(self):
        """cpgpuWorldSaveState: the state the next step starts from, as bytes."""
        n = ctypes.c_int64()
        _check(self.L.cpgpuWorldSaveState(self.h, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        _check(self.L.cpgpuWorldSaveState(self.h, buf, n.value, ctypes.byref(n)))
        return buf.raw

    def load_state(self, data):
        """cpgpuWorldLoadState of bytes from save_state (a world of the same description)."""
        buf = ctypes.create_string_buffer(bytes(data), len(data))
        _check(self.L.cpgpuWorldLoadState(self.h, buf, len(data)))

    def constraints(self):
        n = self.packed.n_constraints
        jacc = np.zeros((n, 2))
        col = np.zeros(n, np.int32)
        _check(self.L.cpgpuWorldReadConstraints(self.h, ptr(jacc, ctypes.c_double), ptr(col, ctypes.c_int32)))
        return jacc, col

    def set_timing(self, on=True, solve_only=False):
        """Time kernel launches with HIP events: all of them, or (solve_only) only the solver's and
        the adjoint's."""
        _check(self.L.cpgpuWorldSetTiming(self.h, (2 if solve_only else 1) if on else 0))

    def timing(self):
        """dict kernel name -> (total ms, launches) since timing was enabled."""
        cap = 64
        n = ctypes.c_int32()
        ms = np.zeros(cap)
        launches = np.zeros(cap, np.int64)
        names = ctypes.create_string_buffer(8192)
        _check(self.L.cpgpuWorldGetTiming(self.h, cap, ctypes.byref(n), ptr(ms, ctypes.c_double),
                                          ptr(launches, ctypes.c_int64), names, 8192))
        keys = names.value.decode().split("\n")[:n.value]
        return {k: (float(ms[i]), int(launches[i])) for i, k in enumerate(keys)}

    def buckets(self):
        """list of (max_bodies, threads, n_worlds, lds_bodies) per size bucket."""
        a = [np.zeros(4, np.int32) for _ in range(4)]
        _check(self.L.cpgpuWorldGetBuckets(self.h, *[ptr(x, ctypes.c_int32) for x in a]))
        return [tuple(int(x[q]) for x in a) for q in range(4)]

    def set_buckets(self, max_bodies, threads):
        mb = np.ascontiguousarray(max_bodies, np.int32)
        th = np.ascontiguousarray(threads, np.int32)
        _check(self.L.cpgpuWorldSetBuckets(self.h, len(mb), ptr(mb, ctypes.c_int32), ptr(th, ctypes.c_int32)))

    def set_grid_schedule(self, schedule):
        """GRID_FLOW (dependency-driven, the default) or GRID_PHASES (grid barriers between phases)."""
        _check(self.L.cpgpuWorldSetGridSchedule(self.h, int(schedule)))

    def set_graph(self, on=True):
        """Capture each step in a HIP graph (cpgpuWorldSetGraph)."""
        _check(self.L.cpgpuWorldSetGraph(self.h, 1 if on else 0))

    # ---- spatial queries (cpSpaceQuery.c) on the last step's shapes ----

    def query_point_nearest(self, points, max_distance, space=None, filters=None):
        """cpSpacePointQueryNearest per point: (shape [n], -1 for none; info [n, 5]: point, distance,
        gradient)."""
        p = _f64(points, (len(points), 2))
        n = len(p)
        d, sp, f = _f64(max_distance, (n,)), _space(space, n), _filters(filters, n)
        shape, info = np.zeros(n, np.int32), np.zeros((n, 5))
        _check(self.L.cpgpuWorldQueryPointNearest(self.h, n, ptr(sp, ctypes.c_int32), ptr(p, ctypes.c_double),
                                                  ptr(d, ctypes.c_double), f, ptr(shape, ctypes.c_int32),
                                                  ptr(info, ctypes.c_double)))
        return shape, info

    def query_segment_first(self, starts, ends, radius, space=None, filters=None):
        """cpSpaceSegmentQueryFirst per segment: (shape [n]; info [n, 5]: point, normal, alpha)."""
        a = _f64(starts, (len(starts), 2))
        n = len(a)
        b, r, sp, f = _f64(ends, (n, 2)), _f64(radius, (n,)), _space(space, n), _filters(filters, n)
        shape, info = np.zeros(n, np.int32), np.zeros((n, 5))
        _check(self.L.cpgpuWorldQuerySegmentFirst(self.h, n, ptr(sp, ctypes.c_int32), ptr(a, ctypes.c_double),
                                                  ptr(b, ctypes.c_double), ptr(r, ctypes.c_double), f,
                                                  ptr(shape, ctypes.c_int32), ptr(info, ctypes.c_double)))
        return shape, info

    def _query_all(self, fn, args, with_info=True):
        total = ctypes.c_int64()
        _check(fn(self.h, *args, 0, ctypes.byref(total), None, None) if with_info else
               fn(self.h, *args, 0, ctypes.byref(total), None))
        cap = total.value
        hits, info = np.zeros((cap, 2), np.int32), np.zeros((cap, 5))
        if with_info:
            _check(fn(self.h, *args, cap, ctypes.byref(total), ptr(hits, ctypes.c_int32), ptr(info, ctypes.c_double)))
            return hits, info
        _check(fn(self.h, *args, cap, ctypes.byref(total), ptr(hits, ctypes.c_int32)))
        return hits

    def query_point_all(self, points, max_distance, space=None, filters=None):
        """cpSpacePointQuery: hits [k, 2] (query, shape) sorted, info [k, 5] (point, distance, gradient)."""
        p = _f64(points, (len(points), 2))
        n = len(p)
        d, sp = _f64(max_distance, (n,)), _space(space, n)
        return self._query_all(self.L.cpgpuWorldQueryPointAll, [n, ptr(sp, ctypes.c_int32), ptr(p, ctypes.c_double),
                                                                ptr(d, ctypes.c_double), _filters(filters, n)])

    def query_segment_all(self, starts, ends, radius, space=None, filters=None):
        """cpSpaceSegmentQuery: hits [k, 2] (query, shape) sorted, info [k, 5] (point, normal, alpha)."""
        a = _f64(starts, (len(starts), 2))
        n = len(a)
        b, r, sp = _f64(ends, (n, 2)), _f64(radius, (n,)), _space(space, n)
        return self._query_all(self.L.cpgpuWorldQuerySegmentAll,
                               [n, ptr(sp, ctypes.c_int32), ptr(a, ctypes.c_double), ptr(b, ctypes.c_double),
                                ptr(r, ctypes.c_double), _filters(filters, n)])

    def query_bb_all(self, boxes, space=None, filters=None):
        """cpSpaceBBQuery: hits [k, 2] (query, shape) sorted; boxes [n, 4] l, b, r, t."""
        bb = _f64(boxes, (len(boxes), 4))
        n = len(bb)
        return self._query_all(self.L.cpgpuWorldQueryBBAll, [n, ptr(_space(space, n), ctypes.c_int32),
                                                             ptr(bb, ctypes.c_double), _filters(filters, n)],
                               with_info=False)

    def query_shapes(self, shapes):
        """cpSpaceShapeQuery of each shape of the world at its body's present
        transform: hits [k, 2] (query, shape) sorted, contact counts [k] and
        contact point sets [k, 12] (normal, then per point pointA, pointB,
        distance)."""
        q = np.ascontiguousarray(shapes, dtype=np.int32).reshape(-1)
        n = len(q)
        total = ctypes.c_int64()
        fn = self.L.cpgpuWorldQueryShapes
        _check(fn(self.h, n, ptr(q, ctypes.c_int32), 0, ctypes.byref(total), None, None, None))
        cap = total.value
        hits, counts, sets = np.zeros((cap, 2), np.int32), np.zeros(cap, np.int32), np.zeros((cap, 12))
        _check(fn(self.h, n, ptr(q, ctypes.c_int32), cap, ctypes.byref(total), ptr(hits, ctypes.c_int32),
                  ptr(counts, ctypes.c_int32), ptr(sets, ctypes.c_double)))
        return hits, counts, sets

    def autotune(self, dt, trial_steps=3):
        """Pick each size bucket's workgroup size by timing (advances the world 5 x trial_steps steps)."""
        out = np.zeros(4, np.int32)
        _check(self.L.cpgpuWorldAutotune(self.h, dt, trial_steps, ptr(out, ctypes.c_int32)))
        return [int(x) for x in out]

    def tape_begin(self, checkpoint_every=10):
        """Record the rollout from here on for backward()."""
        _check(self.L.cpgpuWorldTapeBegin(self.h, checkpoint_every))

    def tape_end(self):
        _check(self.L.cpgpuWorldTapeEnd(self.h))

    def backward(self, g_p=None, g_a=None, g_v=None, g_w=None):
        """Adjoint of the taped rollout.  g_*: dL/d of the final positions [n,2], angles [n],
        velocities [n,2], angular velocities [n] (None: zero).  Returns a dict of dL/d of the
        initial state ('p', 'a', 'v', 'w') and of the parameters ('m_inv', 'i_inv' per body,
        'u' per shape, 'joint' [n_constraints, 8] in the layout of cons_params)."""
        nb, ns = self.packed.n_bodies, self.packed.n_shapes

        def inp(x, shape):
            return None if x is None else np.ascontiguousarray(np.broadcast_to(x, shape), np.float64)

        seeds = [inp(g_p, (nb, 2)), inp(g_a, (nb,)), inp(g_v, (nb, 2)), inp(g_w, (nb,))]
        out = {"p": np.zeros((nb, 2)), "a": np.zeros(nb), "v": np.zeros((nb, 2)), "w": np.zeros(nb),
               "m_inv": np.zeros(nb), "i_inv": np.zeros(nb), "u": np.zeros(ns)}
        args = [None if x is None else ptr(x, ctypes.c_double) for x in seeds]
        args += [ptr(out[k], ctypes.c_double) for k in ("p", "a", "v", "w", "m_inv", "i_inv", "u")]
        _check(self.L.cpgpuWorldBackward(self.h, *args))
        out["joint"] = np.zeros((self.packed.n_constraints, 8))
        _check(self.L.cpgpuWorldGetJointGrads(self.h, ptr(out["joint"], ctypes.c_double)))
        return out

    SOLVE_UNITS = ("solver_worlds_b0", "solver_worlds_b1", "solver_worlds_b2", "solver_worlds_b3", "solver")

    def solve_counts(self):
        """dict solver kernel name -> (contacts solved, bodies integrated) since creation."""
        c = np.zeros(len(self.SOLVE_UNITS), np.int64)
        b = np.zeros(len(self.SOLVE_UNITS), np.int64)
        _check(self.L.cpgpuWorldGetSolveCounts(self.h, ptr(c, ctypes.c_int64), ptr(b, ctypes.c_int64)))
        return {k: (int(c[i]), int(b[i])) for i, k in enumerate(self.SOLVE_UNITS)}


def version():
    return lib().cpgpuVersion().decode()
