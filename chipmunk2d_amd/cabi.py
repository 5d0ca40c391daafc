"""ctypes mirror of include/cpgpu.h: the packed world description and the
step library's entry points.

The structures here are plain C structs of doubles, integers and pointers;
`PackedWorld` owns the numpy arrays a description points into.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

BODY_DYNAMIC, BODY_KINEMATIC, BODY_STATIC = 0, 1, 2
SHAPE_CIRCLE, SHAPE_SEGMENT, SHAPE_POLY = 0, 1, 2
JOINT_PIVOT, JOINT_DAMPED_SPRING, JOINT_SLIDE, JOINT_GEAR = 0, 1, 2, 3
JOINT_PIN, JOINT_GROOVE, JOINT_ROTARY_LIMIT, JOINT_RATCHET, JOINT_SIMPLE_MOTOR, JOINT_DAMPED_ROTARY_SPRING = \
    4, 5, 6, 7, 8, 9
ARB_FIRST_COLLISION, ARB_NORMAL, ARB_IGNORE, ARB_CACHED = 0, 1, 2, 3
FP64, FP32 = 0, 1
EVENT_BEGIN, EVENT_PRE_SOLVE, EVENT_POST_SOLVE, EVENT_SEPARATE = 1, 2, 4, 8
WILDCARD = 0xFFFFFFFF
# the checks a debug build of the reference makes during cpSpaceStep (cpgpuWorldGetFaults)
FAULT_UNSOLVABLE, FAULT_BODY_STATE, FAULT_BODY_MASS = 128, 512, 1024
# a description's collision handlers: {space (-1: every space), type_a, type_b, events}
HANDLER_DTYPE = np.dtype([("space", np.int32), ("type_a", np.uint32), ("type_b", np.uint32), ("events", np.uint32)])

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_up = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)


class SpaceParams(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("collision_persistence", ctypes.c_uint32),
        ("gravity_x", ctypes.c_double),
        ("gravity_y", ctypes.c_double),
        ("damping", ctypes.c_double),
        ("collision_slop", ctypes.c_double),
        ("collision_bias", ctypes.c_double),
        ("curr_dt", ctypes.c_double),
        ("stamp", ctypes.c_uint32),
        ("_pad", ctypes.c_uint32),
    ]


class WorldDesc(ctypes.Structure):
    _fields_ = [
        ("n_worlds", ctypes.c_int32),
        ("space", ctypes.POINTER(SpaceParams)),
        ("world_body_start", _ip),
        ("world_shape_start", _ip),
        ("world_constraint_start", _ip),
        ("n_bodies", ctypes.c_int32),
        ("body_type", _ip),
        ("body_m", _dp), ("body_m_inv", _dp),
        ("body_i", _dp), ("body_i_inv", _dp),
        ("body_cog", _dp), ("body_p", _dp), ("body_v", _dp), ("body_f", _dp),
        ("body_a", _dp), ("body_w", _dp), ("body_t", _dp),
        ("body_v_bias", _dp), ("body_w_bias", _dp),
        ("body_transform", _dp),
        ("n_shapes", ctypes.c_int32),
        ("shape_type", _ip), ("shape_body", _ip),
        ("shape_hashid", _u64p),
        ("shape_e", _dp), ("shape_u", _dp), ("shape_surface_v", _dp),
        ("shape_group", _u64p), ("shape_categories", _up), ("shape_mask", _up),
        ("shape_radius", _dp), ("shape_geom", _dp),
        ("shape_poly_start", _ip), ("shape_poly_count", _ip),
        ("n_poly_verts", ctypes.c_int32),
        ("poly_verts", _dp), ("poly_normals", _dp),
        ("n_constraints", ctypes.c_int32),
        ("cons_type", _ip), ("cons_a", _ip), ("cons_b", _ip),
        ("cons_max_force", _dp), ("cons_error_bias", _dp), ("cons_max_bias", _dp),
        ("cons_collide_bodies", _ip),
        ("cons_params", _dp), ("cons_jacc", _dp),
        ("shape_sensor", ctypes.POINTER(ctypes.c_uint8)),
        ("shape_ctype", _up),
        ("n_handlers", ctypes.c_int32),
        ("handlers", ctypes.c_void_p),
    ]


class CallbackArbiter(ctypes.Structure):
    """cpgpuCallbackArbiter: a touching arbiter whose begin or preSolve is due."""
    _fields_ = [("slot", ctypes.c_int32), ("space", ctypes.c_int32), ("shape_a", ctypes.c_int32),
                ("shape_b", ctypes.c_int32), ("state", ctypes.c_int32), ("events", ctypes.c_int32),
                ("count", ctypes.c_int32), ("_pad", ctypes.c_int32), ("normal", ctypes.c_double * 2),
                ("e", ctypes.c_double), ("u", ctypes.c_double), ("surface_vr", ctypes.c_double * 2),
                ("contact", (ctypes.c_double * 6) * 2), ("hash", ctypes.c_uint64 * 2)]


class CallbackDecision(ctypes.Structure):
    """cpgpuCallbackDecision: what begin / preSolve decided for one arbiter."""
    _fields_ = [("slot", ctypes.c_int32), ("solve", ctypes.c_int32), ("ignore", ctypes.c_int32),
                ("_pad", ctypes.c_int32), ("e", ctypes.c_double), ("u", ctypes.c_double),
                ("surface_vr", ctypes.c_double * 2)]


HANDLER_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(CallbackArbiter),
                              ctypes.POINTER(CallbackDecision))


class Stats(ctypes.Structure):
    _fields_ = [
        ("candidate_pairs", ctypes.c_int64),
        ("arbiters", ctypes.c_int64),
        ("contacts", ctypes.c_int64),
        ("cached_arbiters", ctypes.c_int64),
        ("arbiter_colours", ctypes.c_int32),
        ("constraint_colours", ctypes.c_int32),
        ("steps", ctypes.c_int64),
        ("total_arbiters", ctypes.c_int64),
        ("total_contacts", ctypes.c_int64),
        ("total_pairs", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def ptr(arr, ctype):
    """Pointer to a contiguous numpy array (None for an empty/None array)."""
    if arr is None:
        return None
    assert arr.flags["C_CONTIGUOUS"], "array must be contiguous"
    return arr.ctypes.data_as(ctypes.POINTER(ctype))


_BODY_F64 = ["m", "m_inv", "i", "i_inv", "a", "w", "t", "w_bias"]
_BODY_V2 = ["cog", "p", "v", "f", "v_bias"]


class PackedWorld:
    """Numpy arrays of a packed description (one or more worlds).

    Arrays are named after the cpgpuWorldDesc fields without their prefix
    group: bodies `body_*`, shapes `shape_*`, constraints `cons_*`.
    """

    def __init__(self, arrays):
        self.a = arrays
        self._desc = None

    @property
    def n_worlds(self):
        return len(self.a["space"])

    @property
    def n_bodies(self):
        return len(self.a["body_type"])

    @property
    def n_shapes(self):
        return len(self.a["shape_type"])

    @property
    def n_constraints(self):
        return len(self.a["cons_type"])

    def desc(self):
        """Build (and keep alive) the ctypes cpgpuWorldDesc."""
        a = self.a
        d = WorldDesc()
        self._space = (SpaceParams * max(1, len(a["space"])))()
        for k, sp in enumerate(a["space"]):
            for name, val in sp.items():
                setattr(self._space[k], name, val)
        d.n_worlds = len(a["space"])
        d.space = ctypes.cast(self._space, ctypes.POINTER(SpaceParams))
        d.world_body_start = ptr(a["world_body_start"], ctypes.c_int32)
        d.world_shape_start = ptr(a["world_shape_start"], ctypes.c_int32)
        d.world_constraint_start = ptr(a["world_constraint_start"], ctypes.c_int32)
        d.n_bodies = self.n_bodies
        d.body_type = ptr(a["body_type"], ctypes.c_int32)
        for n in _BODY_F64 + _BODY_V2 + ["transform"]:
            setattr(d, "body_" + n, ptr(a["body_" + n], ctypes.c_double))
        d.n_shapes = self.n_shapes
        d.shape_type = ptr(a["shape_type"], ctypes.c_int32)
        d.shape_body = ptr(a["shape_body"], ctypes.c_int32)
        d.shape_hashid = ptr(a["shape_hashid"], ctypes.c_uint64)
        for n in ["e", "u", "surface_v", "radius", "geom"]:
            setattr(d, "shape_" + n, ptr(a["shape_" + n], ctypes.c_double))
        d.shape_group = ptr(a["shape_group"], ctypes.c_uint64)
        d.shape_categories = ptr(a["shape_categories"], ctypes.c_uint32)
        d.shape_mask = ptr(a["shape_mask"], ctypes.c_uint32)
        d.shape_poly_start = ptr(a["shape_poly_start"], ctypes.c_int32)
        d.shape_poly_count = ptr(a["shape_poly_count"], ctypes.c_int32)
        d.n_poly_verts = len(a["poly_verts"])
        d.poly_verts = ptr(a["poly_verts"], ctypes.c_double)
        d.poly_normals = ptr(a["poly_normals"], ctypes.c_double)
        d.n_constraints = self.n_constraints
        d.cons_type = ptr(a["cons_type"], ctypes.c_int32)
        d.cons_a = ptr(a["cons_a"], ctypes.c_int32)
        d.cons_b = ptr(a["cons_b"], ctypes.c_int32)
        for n in ["max_force", "error_bias", "max_bias", "params", "jacc"]:
            setattr(d, "cons_" + n, ptr(a["cons_" + n], ctypes.c_double))
        d.cons_collide_bodies = ptr(a["cons_collide_bodies"], ctypes.c_int32)
        if "shape_sensor" in a and np.any(a["shape_sensor"]):
            d.shape_sensor = ptr(a["shape_sensor"], ctypes.c_uint8)
        if "shape_ctype" in a and np.any(a["shape_ctype"]):
            d.shape_ctype = ptr(a["shape_ctype"], ctypes.c_uint32)
        if "handlers" in a and len(a["handlers"]):
            self._handlers = np.ascontiguousarray(a["handlers"], dtype=HANDLER_DTYPE)
            d.n_handlers = len(self._handlers)
            d.handlers = self._handlers.ctypes.data
        self._desc = d
        return d

    def world_slice(self, w):
        """(body range, shape range, constraint range) of world w."""
        a = self.a
        return ((int(a["world_body_start"][w]), int(a["world_body_start"][w + 1])),
                (int(a["world_shape_start"][w]), int(a["world_shape_start"][w + 1])),
                (int(a["world_constraint_start"][w]), int(a["world_constraint_start"][w + 1])))

    def dynamic_body_count(self):
        return int(np.count_nonzero(self.a["body_type"] == BODY_DYNAMIC))


def concat(worlds):
    """Pack several PackedWorld objects (each possibly multi-world) into one."""
    out = {"space": []}
    keys = [k for k in worlds[0].a if k not in ("space", "handlers") and not k.startswith("world_")]
    handlers = []
    for k in keys:
        out[k] = []
    wb, ws, wc = [0], [0], [0]
    boff = soff = coff = voff = 0
    for pw in worlds:
        a = pw.a
        out["space"].extend(a["space"])
        for k in keys:
            v = a[k]
            if k == "shape_body":
                v = v + boff
            elif k in ("cons_a", "cons_b"):
                v = v + boff
            elif k == "shape_poly_start":
                v = v + voff
            out[k].append(v)
        # a world's handlers for every one of its spaces become handlers of each, by index
        for h in a.get("handlers", np.zeros(0, HANDLER_DTYPE)):
            spaces = range(len(a["space"])) if h["space"] < 0 else [int(h["space"])]
            for sp in spaces:
                handlers.append((sp + len(wb) - 1, h["type_a"], h["type_b"], h["events"]))
        for w in range(len(a["space"])):
            wb.append(int(a["world_body_start"][w + 1]) + boff)
            ws.append(int(a["world_shape_start"][w + 1]) + soff)
            wc.append(int(a["world_constraint_start"][w + 1]) + coff)
        boff += len(a["body_type"])
        soff += len(a["shape_type"])
        coff += len(a["cons_type"])
        voff += len(a["poly_verts"])
    for k in keys:
        out[k] = np.ascontiguousarray(np.concatenate(out[k]))
    out["world_body_start"] = np.array(wb, dtype=np.int32)
    out["world_shape_start"] = np.array(ws, dtype=np.int32)
    out["world_constraint_start"] = np.array(wc, dtype=np.int32)
    out["handlers"] = np.array(handlers, dtype=HANDLER_DTYPE)
    return PackedWorld(out)
