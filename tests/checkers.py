"""ctypes handles on the checkers: the oracle (oracle/liboracle.so) and the
compiled reference with its bridge (oracle/_ref/libcpref.so).

Test infrastructure only.  Both libraries are built by `make -C oracle`
(called from __graft_entry__.build()).
"""
import ctypes
import os

import numpy as np

from chipmunk2d_amd.cabi import HANDLER_DTYPE, HANDLER_FN, PackedWorld, SpaceParams, Stats, WorldDesc, ptr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
GPUSPACE_SO = os.path.join(ROOT, "oracle", "_ref", "libcpgpuspace.so")
GPUSPACE_CPU_SO = os.path.join(ROOT, "oracle", "_ref", "libcpgpuspace_cpu.so")
DUAL_SO = os.path.join(ROOT, "oracle", "libdualstep.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libcpref.so")
REF_RELEASE_SO = os.path.join(ROOT, "oracle", "_ref", "libcpref_release.so")
# the strict build with the reference's assertions on: it aborts where a debug build would
REF_DEBUG_SO = os.path.join(ROOT, "oracle", "_ref", "libcpref_debug.so")

_d = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_up = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p

_oracle = None
_refs = {}


def oracle_lib():
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(ORACLE_SO)
        lib.ow_new.restype = _vp
        lib.ow_new.argtypes = [ctypes.POINTER(WorldDesc)]
        lib.ow_free.argtypes = [_vp]
        lib.ow_step.argtypes = [_vp, _d]
        lib.ow_set_libm_trig.argtypes = [_vp, ctypes.c_int]
        lib.ow_set_joint_list_order.argtypes = [_vp, ctypes.c_int]
        lib.ow_set_next_order.argtypes = [_vp, ctypes.c_int64, _ip]
        lib.ow_set_next_collisions.argtypes = [_vp, ctypes.c_int64, _ip]
        lib.ow_read_bodies.argtypes = [_vp] + [_dp] * 7
        lib.ow_write_bodies.argtypes = [_vp] + [_dp] * 4
        lib.ow_read_arbiters.restype = ctypes.c_int64
        lib.ow_read_arbiters.argtypes = [_vp, ctypes.c_int64, _ip, _ip, _ip, _ip, _dp, _dp, _u64p]
        lib.ow_read_constraints.argtypes = [_vp, _dp, _ip]
        lib.ow_read_units.restype = ctypes.c_int64
        lib.ow_read_units.argtypes = [_vp, ctypes.c_int64, _ip]
        lib.ow_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
        lib.ow_cache_all_shapes.argtypes = [_vp]
        lib.ow_read_shape_bbs.argtypes = [_vp, _dp]
        lib.ow_set_transforms.argtypes = [_vp, _dp]
        lib.ow_collide.restype = ctypes.c_int
        lib.ow_collide.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, _ip, _up, _dp, _dp, _u64p]
        lib.ow_shapes_collide.restype = ctypes.c_int
        lib.ow_shapes_collide.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp]
        lib.ow_candidate_pairs.restype = ctypes.c_int64
        lib.ow_candidate_pairs.argtypes = [_vp, ctypes.c_int64, _ip]
        lib.ow_write_bodies_at.argtypes = [_vp, ctypes.c_int32, _ip] + [_dp] * 7
        _q = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
        lib.ow_point_query_nearest.restype = ctypes.c_int32
        lib.ow_point_query_nearest.argtypes = [_vp, ctypes.c_int32, _dp, _d] + _q + [_dp]
        lib.ow_segment_query_first.restype = ctypes.c_int32
        lib.ow_segment_query_first.argtypes = [_vp, ctypes.c_int32, _dp, _dp, _d] + _q + [_dp]
        lib.ow_point_query.restype = ctypes.c_int64
        lib.ow_point_query.argtypes = [_vp, ctypes.c_int32, _dp, _d] + _q + [ctypes.c_int64, _ip, _dp]
        lib.ow_segment_query.restype = ctypes.c_int64
        lib.ow_segment_query.argtypes = [_vp, ctypes.c_int32, _dp, _dp, _d] + _q + [ctypes.c_int64, _ip, _dp]
        lib.ow_bb_query.restype = ctypes.c_int64
        lib.ow_bb_query.argtypes = [_vp, ctypes.c_int32, _dp] + _q + [ctypes.c_int64, _ip]
        lib.ow_set_handler_fn.argtypes = [_vp, HANDLER_FN, _vp]
        lib.ow_read_separations.restype = ctypes.c_int64
        lib.ow_read_separations.argtypes = [_vp, ctypes.c_int64, _ip]
        lib.ow_faults.restype = ctypes.c_int32
        lib.ow_faults.argtypes = [_vp, _ip]
        lib.ow_shape_query.restype = ctypes.c_int64
        lib.ow_shape_query.argtypes = [_vp, ctypes.c_int32, ctypes.c_int64, _ip, _ip, _dp]
        lib.ow_set_space_params.argtypes = [_vp, ctypes.c_int32, ctypes.POINTER(SpaceParams)]
        lib.ow_new_from.restype = _vp
        lib.ow_new_from.argtypes = [ctypes.POINTER(WorldDesc), _vp, _ip]
        lib.cp_sincos_oracle.argtypes = [_d, _dp, _dp]
        lib.cp_exp_oracle.restype = _d
        lib.cp_exp_oracle.argtypes = [_d]
        _oracle = lib
    return _oracle


def ref_lib(release=False, debug=False):
    path = REF_RELEASE_SO if release else REF_DEBUG_SO if debug else REF_SO
    if path not in _refs:
        lib = ctypes.CDLL(path)
        lib.rb_new.restype = _vp
        lib.rb_new.argtypes = [ctypes.POINTER(WorldDesc), ctypes.c_int, ctypes.c_int]
        lib.rb_free.argtypes = [_vp]
        lib.rb_step.argtypes = [_vp, _d, ctypes.c_int]
        lib.rb_read_bodies.argtypes = [_vp] + [_dp] * 7
        lib.rb_read_arbiters.restype = ctypes.c_int64
        lib.rb_read_arbiters.argtypes = [_vp, ctypes.c_int64, _ip, _ip, _ip, _dp, _dp, _u64p]
        lib.rb_num_cached_arbiters.restype = ctypes.c_int64
        lib.rb_num_cached_arbiters.argtypes = [_vp]
        lib.rb_read_constraints.argtypes = [_vp, _dp]
        lib.rb_set_transforms.argtypes = [_vp, _dp]
        lib.rb_read_shape_bbs.argtypes = [_vp, _dp]
        lib.rb_collide.restype = ctypes.c_int
        lib.rb_collide.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, _ip, _up, _dp, _dp, _u64p]
        lib.rb_shapes_collide.restype = ctypes.c_int
        lib.rb_shapes_collide.argtypes = [_vp, ctypes.c_int32, ctypes.c_int32, _dp, _dp]
        lib.rb_candidate_pairs.restype = ctypes.c_int64
        lib.rb_candidate_pairs.argtypes = [_vp, ctypes.c_int64, _ip]
        lib.rb_body_update_position.argtypes = [_vp, ctypes.c_int, _d]
        lib.rb_body_update_velocity.argtypes = [_vp, ctypes.c_int, _d, _d, _d, _d]
        lib.rb_log_collisions.argtypes = [ctypes.c_int]
        lib.rb_read_collision_log.restype = ctypes.c_int64
        lib.rb_read_collision_log.argtypes = [ctypes.c_int64, _ip]
        lib.rb_space_stamp.restype = ctypes.c_int64
        lib.rb_space_stamp.argtypes = [_vp]
        lib.rb_remove_shape.argtypes = lib.rb_remove_body.argtypes = [_vp, ctypes.c_int]
        _q = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32]
        lib.rb_point_query_nearest.restype = ctypes.c_int32
        lib.rb_point_query_nearest.argtypes = [_vp, _dp, _d] + _q + [_dp]
        lib.rb_segment_query_first.restype = ctypes.c_int32
        lib.rb_segment_query_first.argtypes = [_vp, _dp, _dp, _d] + _q + [_dp]
        lib.rb_point_query.restype = ctypes.c_int64
        lib.rb_point_query.argtypes = [_vp, _dp, _d] + _q + [ctypes.c_int64, _ip, _dp]
        lib.rb_segment_query.restype = ctypes.c_int64
        lib.rb_segment_query.argtypes = [_vp, _dp, _dp, _d] + _q + [ctypes.c_int64, _ip, _dp]
        lib.rb_bb_query.restype = ctypes.c_int64
        lib.rb_bb_query.argtypes = [_vp, _dp] + _q + [ctypes.c_int64, _ip]
        lib.rb_set_handler_fn.argtypes = [_vp, HANDLER_FN, _vp]
        lib.rb_shape_query.restype = ctypes.c_int64
        lib.rb_shape_query.argtypes = [_vp, ctypes.c_int32, ctypes.c_int64, _ip, _ip, _dp, _ip]
        lib.rb_new_in.restype = _vp
        lib.rb_new_in.argtypes = [_vp, ctypes.POINTER(WorldDesc), ctypes.c_int]
        for name in ("rb_space", "rb_body", "rb_shape", "rb_constraint"):
            getattr(lib, name).restype = _vp
        lib.rb_space.argtypes = [_vp]
        lib.rb_body.argtypes = lib.rb_shape.argtypes = lib.rb_constraint.argtypes = [_vp, ctypes.c_int]
        lib.rb_body_raw.argtypes = [_vp, _dp]
        lib.rb_constraint_jacc.argtypes = [_vp, _dp]
        lib.rb_body_arbiters.restype = ctypes.c_int
        lib.rb_body_arbiters.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _ip, _ip, _dp,
                                         _dp, _dp, _dp]
        # a few reference entry points called directly
        lib.cpConvexHull.restype = ctypes.c_int
        lib.cpConvexHull.argtypes = [ctypes.c_int, _dp, _dp, _ip, _d]
        lib.cpMomentForPoly.restype = _d
        lib.cpAreaForPoly.restype = _d
        _refs[path] = lib
    return _refs[path]


def ref_available():
    return os.path.exists(REF_SO)


def _arr(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype)


class BodyState:
    FIELDS = ("p", "v", "a", "w", "v_bias", "w_bias", "transform")

    def __init__(self, n):
        self.p = _arr((n, 2))
        self.v = _arr((n, 2))
        self.a = _arr(n)
        self.w = _arr(n)
        self.v_bias = _arr((n, 2))
        self.w_bias = _arr(n)
        self.transform = _arr((n, 6))

    def ptrs(self):
        return [ptr(getattr(self, f), ctypes.c_double) for f in self.FIELDS]


class Arbiters:
    def __init__(self, n):
        self.n = n
        self.shapes = _arr((n, 2), np.int32)
        self.colour = _arr(n, np.int32)
        self.state = _arr(n, np.int32)
        self.count = _arr(n, np.int32)
        self.normal = _arr((n, 2))
        self.contact = _arr((n, 2, 12))
        self.hash = _arr((n, 2), np.uint64)

    def by_key(self):
        """dict (min shape, max shape) -> index."""
        out = {}
        for k in range(self.n):
            a, b = int(self.shapes[k, 0]), int(self.shapes[k, 1])
            out[(min(a, b), max(a, b))] = k
        return out


class OracleWorld:
    def __init__(self, packed):
        self.lib = oracle_lib()
        self.packed = packed
        self.desc = packed.desc()
        self.h = self.lib.ow_new(ctypes.byref(self.desc))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ow_free(self.h)
            self.h = None

    def step(self, dt, n=1):
        for _ in range(n):
            self.lib.ow_step(self.h, dt)

    def set_handler_fn(self, fn):
        """Answer the description's collision handlers with fn(event, view, decision)
        (None: the defaults)."""
        self._hfn = HANDLER_FN(lambda ctx, ev, v, d: fn(ev, v.contents, d.contents)) if fn else HANDLER_FN()
        self.lib.ow_set_handler_fn(self.h, self._hfn, None)

    def separations(self):
        """The last step's separations of pairs with a separate event: (n, 3) {space, a, b}."""
        n = self.lib.ow_read_separations(self.h, 0, None)
        out = np.zeros((n, 3), np.int32)
        if n:
            self.lib.ow_read_separations(self.h, n, ptr(out, ctypes.c_int32))
        return out

    def faults(self):
        """(CPGPU_FAULT_* bits, lowest failing body or -1) raised so far."""
        b = ctypes.c_int32(0)
        return int(self.lib.ow_faults(self.h, ctypes.byref(b))), b.value

    def libm_trig(self, on=True):
        self.lib.ow_set_libm_trig(self.h, 1 if on else 0)

    def joint_list_order(self, on=True):
        self.lib.ow_set_joint_list_order(self.h, 1 if on else 0)

    def set_next_order(self, pairs):
        pairs = np.ascontiguousarray(pairs, dtype=np.int32)
        self.lib.ow_set_next_order(self.h, len(pairs), ptr(pairs, ctypes.c_int32))

    def set_next_collisions(self, coll):
        """coll: (n, >=3) int array of {a, b, id_in, ...} from RefWorld.step_logged."""
        c = np.ascontiguousarray(np.asarray(coll)[:, :3], dtype=np.int32)
        self.lib.ow_set_next_collisions(self.h, len(c), ptr(c, ctypes.c_int32))

    def replay(self, ref_arbiters, collisions, dt):
        """One step in the reference's own pair orientations, ids and solve order."""
        self.set_next_order(ref_arbiters.shapes)
        self.set_next_collisions(collisions)
        self.step(dt)

    def bodies(self):
        s = BodyState(self.packed.n_bodies)
        self.lib.ow_read_bodies(self.h, *s.ptrs())
        return s

    def write_bodies(self, v=None, w=None, f=None, t=None):
        self.lib.ow_write_bodies(self.h, *[ptr(x, ctypes.c_double) if x is not None else None for x in (v, w, f, t)])

    def write_bodies_at(self, index, p=None, v=None, f=None, a=None, w=None, t=None, transform=None):
        index = np.ascontiguousarray(index, dtype=np.int32)
        args = [None if x is None else ptr(np.ascontiguousarray(x, dtype=np.float64), ctypes.c_double)
                for x in (p, v, f, a, w, t, transform)]
        assert self.lib.ow_write_bodies_at(self.h, len(index), ptr(index, ctypes.c_int32), *args) == 0

    def set_space_params(self, k, params):
        sp = SpaceParams(**params)
        assert self.lib.ow_set_space_params(self.h, k, ctypes.byref(sp)) == 0

    @classmethod
    def from_old(cls, packed, old, shape_from):
        """The oracle's counterpart of cpgpuWorldCreateFrom: a world of `packed` that keeps
        the collision state of `old` under the new shape indices."""
        self = cls.__new__(cls)
        self.lib = oracle_lib()
        self.packed = packed
        self.desc = packed.desc()
        sf = np.ascontiguousarray(shape_from, dtype=np.int32)
        self.h = self.lib.ow_new_from(ctypes.byref(self.desc), old.h, ptr(sf, ctypes.c_int32))
        return self

    def stats(self):
        st = Stats()
        self.lib.ow_stats(self.h, ctypes.byref(st))
        return st.as_dict()

    def arbiters(self):
        n = self.lib.ow_read_arbiters(self.h, 0, None, None, None, None, None, None, None)
        A = Arbiters(n)
        self.lib.ow_read_arbiters(self.h, n, ptr(A.shapes, ctypes.c_int32), ptr(A.colour, ctypes.c_int32),
                                  ptr(A.state, ctypes.c_int32), ptr(A.count, ctypes.c_int32),
                                  ptr(A.normal, ctypes.c_double), ptr(A.contact, ctypes.c_double),
                                  ptr(A.hash, ctypes.c_uint64))
        return A

    def units(self):
        """Solve unit of every arbiter of the last step (arbiters() order)."""
        n = self.lib.ow_read_units(self.h, 0, None)
        u = _arr(n, np.int32)
        self.lib.ow_read_units(self.h, n, ptr(u, ctypes.c_int32))
        return u

    def constraints(self):
        n = self.packed.n_constraints
        jacc = _arr((n, 2))
        col = _arr(n, np.int32)
        self.lib.ow_read_constraints(self.h, ptr(jacc, ctypes.c_double), ptr(col, ctypes.c_int32))
        return jacc, col

    def set_transforms(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        self.lib.ow_set_transforms(self.h, ptr(t, ctypes.c_double))
        self.lib.ow_cache_all_shapes(self.h)

    def shape_bbs(self):
        bb = _arr((self.packed.n_shapes, 4))
        self.lib.ow_read_shape_bbs(self.h, ptr(bb, ctypes.c_double))
        return bb

    def collide(self, a, b, cid=0):
        return _collide(self.lib.ow_collide, self.h, a, b, cid)

    def shapes_collide(self, a, b):
        return _shapes_collide(self.lib.ow_shapes_collide, self.h, a, b)

    def candidate_pairs(self):
        n = self.lib.ow_candidate_pairs(self.h, 0, None)
        out = _arr((n, 2), np.int32)
        self.lib.ow_candidate_pairs(self.h, n, ptr(out, ctypes.c_int32))
        return out


def _shapes_collide(fn, h, a, b):
    """(count, set [12]) of cpShapesCollide in the library's layout: normal, then per point
    pointA, pointB, distance; zeros past the count."""
    n = _arr(2)
    pts = _arr((2, 5))
    count = fn(h, a, b, ptr(n, ctypes.c_double), ptr(pts, ctypes.c_double))
    out = np.zeros(12)
    out[:2] = n
    out[2:2 + 5 * count] = pts[:count].ravel()
    return count, out


def _collide(fn, h, a, b, cid):
    ab = _arr(2, np.int32)
    idout = _arr(1, np.uint32)
    n = _arr(2)
    pts = _arr((2, 4))
    hashes = _arr(2, np.uint64)
    count = fn(h, a, b, cid, ptr(ab, ctypes.c_int32), ptr(idout, ctypes.c_uint32), ptr(n, ctypes.c_double),
               ptr(pts, ctypes.c_double), ptr(hashes, ctypes.c_uint64))
    return dict(count=count, ab=tuple(int(x) for x in ab), id=int(idout[0]), n=n.copy(),
                points=pts[:count].copy(), hash=hashes[:count].copy())


class RefWorld:
    """World `w` of a packed description built as a real cpSpace."""

    def __init__(self, packed, w=0, hasty_threads=0, release=False, debug=False):
        self.lib = ref_lib(release, debug)
        self.packed = packed
        self.desc = packed.desc()
        self.w = w
        (self.b0, self.b1), (self.s0, self.s1), (self.c0, self.c1) = packed.world_slice(w)
        self.h = self.lib.rb_new(ctypes.byref(self.desc), w, hasty_threads)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.rb_free(self.h)
            self.h = None

    def set_handler_fn(self, fn):
        """Answer the collision handlers installed from the description with
        fn(event, view, decision) (None: the defaults)."""
        self._hfn = HANDLER_FN(lambda ctx, ev, v, d: fn(ev, v.contents, d.contents)) if fn else HANDLER_FN()
        self.lib.rb_set_handler_fn(self.h, self._hfn, None)

    def step(self, dt, n=1):
        self.lib.rb_step(self.h, dt, n)

    def step_logged(self, dt):
        """One step; returns the {a, b, id_in, id_out} of every cpCollide it made."""
        self.lib.rb_log_collisions(1)
        self.lib.rb_step(self.h, dt, 1)
        n = self.lib.rb_read_collision_log(0, None)
        out = _arr((n, 4), np.int32)
        self.lib.rb_read_collision_log(n, ptr(out, ctypes.c_int32))
        self.lib.rb_log_collisions(0)
        out[:, 0] += self.s0
        out[:, 1] += self.s0
        return out

    def bodies(self):
        s = BodyState(self.b1 - self.b0)
        self.lib.rb_read_bodies(self.h, *s.ptrs())
        return s

    def arbiters(self):
        n = self.lib.rb_read_arbiters(self.h, 0, None, None, None, None, None, None)
        A = Arbiters(n)
        self.lib.rb_read_arbiters(self.h, n, ptr(A.shapes, ctypes.c_int32), ptr(A.state, ctypes.c_int32),
                                  ptr(A.count, ctypes.c_int32), ptr(A.normal, ctypes.c_double),
                                  ptr(A.contact, ctypes.c_double), ptr(A.hash, ctypes.c_uint64))
        return A

    def cached_arbiters(self):
        return self.lib.rb_num_cached_arbiters(self.h)

    def constraints(self):
        jacc = _arr((self.c1 - self.c0, 2))
        self.lib.rb_read_constraints(self.h, ptr(jacc, ctypes.c_double))
        return jacc

    def set_transforms(self, t):
        t = np.ascontiguousarray(t, dtype=np.float64)
        self.lib.rb_set_transforms(self.h, ptr(t, ctypes.c_double))

    def shape_bbs(self):
        bb = _arr((self.s1 - self.s0, 4))
        self.lib.rb_read_shape_bbs(self.h, ptr(bb, ctypes.c_double))
        return bb

    def collide(self, a, b, cid=0):
        r = _collide(self.lib.rb_collide, self.h, a - self.s0, b - self.s0, cid)
        return r

    def shapes_collide(self, a, b):
        return _shapes_collide(self.lib.rb_shapes_collide, self.h, a - self.s0, b - self.s0)

    def candidate_pairs(self):
        n = self.lib.rb_candidate_pairs(self.h, 0, None)
        return n

    def candidate_pair_list(self, cap):
        out = _arr((cap, 2), np.int32)
        n = self.lib.rb_candidate_pairs(self.h, cap, ptr(out, ctypes.c_int32))
        return out[:min(n, cap)]


def sincos(x):
    lib = oracle_lib()
    s = ctypes.c_double()
    c = ctypes.c_double()
    lib.cp_sincos_oracle(x, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def packed_from_desc(desc_ptr):
    """A PackedWorld holding copies of the arrays a cpgpuWorldDesc points to."""
    d = ctypes.cast(desc_ptr, ctypes.POINTER(WorldDesc)).contents

    def arr(p, n, dtype, width=1):
        shape = (n, width) if width > 1 else (n,)
        if n == 0 or not p:
            return np.zeros(shape, dtype)
        return np.ctypeslib.as_array(p, shape=(n * width,)).astype(dtype).reshape(shape).copy()

    nw, nb, ns, nc, nv = d.n_worlds, d.n_bodies, d.n_shapes, d.n_constraints, d.n_poly_verts
    a = {"space": [{name: getattr(d.space[k], name) for name, _ in SpaceParams._fields_ if name != "_pad"}
                   for k in range(nw)]}
    a["world_body_start"] = arr(d.world_body_start, nw + 1, np.int32)
    a["world_shape_start"] = arr(d.world_shape_start, nw + 1, np.int32)
    a["world_constraint_start"] = arr(d.world_constraint_start, nw + 1, np.int32)
    a["body_type"] = arr(d.body_type, nb, np.int32)
    for n in ("m", "m_inv", "i", "i_inv", "a", "w", "t", "w_bias"):
        a["body_" + n] = arr(getattr(d, "body_" + n), nb, np.float64)
    for n in ("cog", "p", "v", "f", "v_bias"):
        a["body_" + n] = arr(getattr(d, "body_" + n), nb, np.float64, 2)
    a["body_transform"] = arr(d.body_transform, nb, np.float64, 6)
    a["shape_type"] = arr(d.shape_type, ns, np.int32)
    a["shape_body"] = arr(d.shape_body, ns, np.int32)
    a["shape_hashid"] = arr(d.shape_hashid, ns, np.uint64)
    a["shape_e"] = arr(d.shape_e, ns, np.float64)
    a["shape_u"] = arr(d.shape_u, ns, np.float64)
    a["shape_surface_v"] = arr(d.shape_surface_v, ns, np.float64, 2)
    a["shape_group"] = arr(d.shape_group, ns, np.uint64)
    a["shape_categories"] = arr(d.shape_categories, ns, np.uint32)
    a["shape_mask"] = arr(d.shape_mask, ns, np.uint32)
    a["shape_sensor"] = arr(d.shape_sensor, ns, np.uint8)
    a["shape_ctype"] = arr(d.shape_ctype, ns, np.uint32)
    nh = d.n_handlers
    a["handlers"] = (np.ctypeslib.as_array(ctypes.cast(d.handlers, ctypes.POINTER(ctypes.c_uint32)), shape=(4 * nh,))
                     .copy().view(HANDLER_DTYPE) if nh > 0 else np.zeros(0, HANDLER_DTYPE))
    a["shape_radius"] = arr(d.shape_radius, ns, np.float64)
    a["shape_geom"] = arr(d.shape_geom, ns, np.float64, 10)
    a["shape_poly_start"] = arr(d.shape_poly_start, ns, np.int32)
    a["shape_poly_count"] = arr(d.shape_poly_count, ns, np.int32)
    a["poly_verts"] = arr(d.poly_verts, nv, np.float64, 2)
    a["poly_normals"] = arr(d.poly_normals, nv, np.float64, 2)
    a["cons_type"] = arr(d.cons_type, nc, np.int32)
    a["cons_a"] = arr(d.cons_a, nc, np.int32)
    a["cons_b"] = arr(d.cons_b, nc, np.int32)
    for n in ("max_force", "error_bias", "max_bias"):
        a["cons_" + n] = arr(getattr(d, "cons_" + n), nc, np.float64)
    a["cons_collide_bodies"] = arr(d.cons_collide_bodies, nc, np.int32)
    a["cons_params"] = arr(d.cons_params, nc, np.float64, 8)
    a["cons_jacc"] = arr(d.cons_jacc, nc, np.float64, 2)
    return PackedWorld(a)


def without(pw, shapes=(), bodies=()):
    """The one-space description pw with the listed shapes and bodies taken out
    (the others keep their order, indices renumbered); constraints must not use
    a removed body.  Returns (description, kept shape indices, kept body indices)."""
    a = pw.a
    keep_s = np.array([k for k in range(pw.n_shapes) if k not in set(shapes)], np.int64)
    keep_b = np.array([k for k in range(pw.n_bodies) if k not in set(bodies)], np.int64)
    new_b = -np.ones(pw.n_bodies, np.int64)
    new_b[keep_b] = np.arange(len(keep_b))
    assert np.all(new_b[a["shape_body"][keep_s]] >= 0)
    out = {"space": [dict(sp) for sp in a["space"]]}
    for k, val in a.items():
        if k == "space" or k.startswith("world_"):
            continue
        if k.startswith("body_"):
            out[k] = np.ascontiguousarray(val[keep_b])
        elif k.startswith("shape_"):
            out[k] = np.ascontiguousarray(val[keep_s])
        else:
            out[k] = val.copy()
    out["shape_body"] = new_b[a["shape_body"][keep_s]].astype(np.int32)
    if pw.n_constraints:
        out["cons_a"] = new_b[a["cons_a"]].astype(np.int32)
        out["cons_b"] = new_b[a["cons_b"]].astype(np.int32)
        assert np.all(out["cons_a"] >= 0) and np.all(out["cons_b"] >= 0)
    out["world_body_start"] = np.array([0, len(keep_b)], np.int32)
    out["world_shape_start"] = np.array([0, len(keep_s)], np.int32)
    out["world_constraint_start"] = a["world_constraint_start"].copy()
    return PackedWorld(out), keep_s, keep_b


_dual = None


def dual_lib():
    global _dual
    if _dual is None:
        lib = ctypes.CDLL(DUAL_SO)
        lib.dual_new.restype = _vp
        lib.dual_new.argtypes = [ctypes.POINTER(WorldDesc)]
        lib.dual_free.argtypes = [_vp]
        lib.dual_step.argtypes = [_vp, _d]
        lib.dual_set_libm_trig.argtypes = [_vp, ctypes.c_int]
        lib.dual_seed.argtypes = [_vp, _dp, _dp, _dp, _dp]
        lib.dual_read_bodies.argtypes = [_vp] + [_dp] * 8
        _dual = lib
    return _dual


class DualWorld:
    """The oracle's step on dual numbers (oracle/libdualstep.so): the rollout
    and its directional derivative along a seeded direction of the initial
    velocities, angular velocities, masses and friction coefficients."""

    def __init__(self, packed, dv=None, dw=None, dm=None, du=None):
        self.lib = dual_lib()
        self.packed = packed
        self.desc = packed.desc()
        self.h = self.lib.dual_new(ctypes.byref(self.desc))
        seeds = [None if x is None else np.ascontiguousarray(x, dtype=np.float64) for x in (dv, dw, dm, du)]
        self._seeds = seeds
        self.lib.dual_seed(self.h, *[None if x is None else ptr(x, ctypes.c_double) for x in seeds])

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.dual_free(self.h)
            self.h = None

    def step(self, dt, n=1):
        for _ in range(n):
            self.lib.dual_step(self.h, dt)

    def bodies(self):
        """(values, tangents): dicts of p, v [n,2], a, w [n]."""
        n = self.packed.n_bodies
        val = {"p": _arr((n, 2)), "v": _arr((n, 2)), "a": _arr(n), "w": _arr(n)}
        tan = {k: np.zeros_like(x) for k, x in val.items()}
        args = []
        for k in ("p", "v", "a", "w"):
            args += [ptr(val[k], ctypes.c_double), ptr(tan[k], ctypes.c_double)]
        self.lib.dual_read_bodies(self.h, *args)
        return val, tan


ALL = (0, 0xFFFFFFFF, 0xFFFFFFFF)   # CP_SHAPE_FILTER_ALL: group, categories, mask


def _v(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class Queries:
    """Spatial queries (cpSpaceQuery.c) with one calling convention for the oracle
    (`ow_*`, a space index first) and the reference (`rb_*`, one space); shape
    indices are the description's, hits sorted by shape."""

    def __init__(self, fn_prefix, lib, h, world=None, cap=4096):
        self.lib, self.h, self.world, self.cap, self.pre = lib, h, world, cap, fn_prefix

    def _args(self):
        return [self.h] if self.world is None else [self.h, self.world]

    def point_nearest(self, p, max_distance, filt=ALL):
        info = np.zeros(5)
        s = getattr(self.lib, self.pre + "point_query_nearest")(*self._args(), ptr(_v(p), ctypes.c_double),
                                                                 max_distance, *filt, ptr(info, ctypes.c_double))
        return int(s), info

    def segment_first(self, a, b, radius, filt=ALL):
        info = np.zeros(5)
        s = getattr(self.lib, self.pre + "segment_query_first")(*self._args(), ptr(_v(a), ctypes.c_double),
                                                                 ptr(_v(b), ctypes.c_double), radius, *filt,
                                                                 ptr(info, ctypes.c_double))
        return int(s), info

    def _hits(self, name, args, with_info=True):
        shapes = np.zeros(self.cap, np.int32)
        info = np.zeros((self.cap, 5))
        extra = [ptr(shapes, ctypes.c_int32)] + ([ptr(info, ctypes.c_double)] if with_info else [])
        n = getattr(self.lib, self.pre + name)(*self._args(), *args, self.cap, *extra)
        assert n <= self.cap
        order = np.argsort(shapes[:n], kind="stable")
        return shapes[:n][order], info[:n][order]

    def point_all(self, p, max_distance, filt=ALL):
        return self._hits("point_query", [ptr(_v(p), ctypes.c_double), max_distance, *filt])

    def segment_all(self, a, b, radius, filt=ALL):
        return self._hits("segment_query", [ptr(_v(a), ctypes.c_double), ptr(_v(b), ctypes.c_double), radius, *filt])

    def bb_all(self, bb, filt=ALL):
        return self._hits("bb_query", [ptr(_v(bb), ctypes.c_double), *filt], with_info=False)[0]

    def shape(self, s):
        """cpSpaceShapeQuery of shape s: the touching shapes, their contact
        counts and point sets [n, 12], sorted by shape; and, for the reference,
        the function's return value (None for the oracle)."""
        shapes = np.zeros(self.cap, np.int32)
        counts = np.zeros(self.cap, np.int32)
        sets = np.zeros((self.cap, 12))
        any_ = np.zeros(1, np.int32)
        args = [self.h, s, self.cap, ptr(shapes, ctypes.c_int32), ptr(counts, ctypes.c_int32), ptr(sets, ctypes.c_double)]
        if self.pre == "rb_":
            args.append(ptr(any_, ctypes.c_int32))
        n = getattr(self.lib, self.pre + "shape_query")(*args)
        assert n <= self.cap
        order = np.argsort(shapes[:n], kind="stable")
        return shapes[:n][order], counts[:n][order], sets[:n][order], (bool(any_[0]) if self.pre == "rb_" else None)


def oracle_queries(o, world=0):
    return Queries("ow_", o.lib, o.h, world)


def ref_queries(r):
    return Queries("rb_", r.lib, r.h)
