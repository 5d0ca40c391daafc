"""The C-ABI library loads and exports every entry point include/cpgpu.h
declares; argument validation runs before any device work -- CPU only."""
import ctypes
import os
import re

import numpy as np
import pytest

from chipmunk2d_amd import gpu
from chipmunk2d_amd.cabi import WorldDesc, concat
from chipmunk2d_amd.scenes import pyramid_210

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "cpgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cpgpu[A-Za-z]+)\s*\(", text)))


def test_header_declares_the_step_api():
    names = declared_functions()
    for must in ("cpgpuWorldCreate", "cpgpuWorldStep", "cpgpuWorldDestroy", "cpgpuWorldReadBodies"):
        assert must in names


def test_library_exports_every_declared_symbol():
    L = gpu.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    assert sorted(gpu.EXPORTS) == declared_functions()


def test_version_string():
    assert "gfx950" in gpu.version()


def _create(pw):
    L = gpu.lib()
    d = pw.desc()
    h = ctypes.c_void_p()
    rc = L.cpgpuWorldCreate(ctypes.byref(d), 0, ctypes.byref(h))
    return rc, L.cpgpuGetLastError().decode(), h


def test_invalid_descriptions_are_rejected():
    pw = pyramid_210().pack()
    # shape pointing at a body out of range
    bad = pyramid_210().pack()
    bad.a["shape_body"] = bad.a["shape_body"].copy()
    bad.a["shape_body"][5] = 10_000
    rc, msg, _ = _create(bad)
    assert rc == -1 and "body index" in msg
    # polygon with too many vertices
    bad = pyramid_210().pack()
    bad.a["shape_poly_count"] = bad.a["shape_poly_count"].copy()
    bad.a["shape_poly_count"][3] = 1000
    rc, msg, _ = _create(bad)
    assert rc == -1 and "vertex count" in msg
    # world ranges that do not cover the arrays
    bad = pyramid_210().pack()
    bad.a["world_body_start"] = np.array([0, 5], np.int32)
    rc, msg, _ = _create(bad)
    assert rc == -1 and "world_body_start" in msg
    # unsupported precision
    L = gpu.lib()
    d = pw.desc()
    h = ctypes.c_void_p()
    assert L.cpgpuWorldCreate(ctypes.byref(d), 7, ctypes.byref(h)) == -1


def test_null_handles():
    L = gpu.lib()
    assert L.cpgpuWorldStep(None, 0.1) == -1
    assert L.cpgpuWorldSync(None) == -1
    L.cpgpuWorldDestroy(None)


def test_batch_packing_offsets():
    a, b = pyramid_210(rows=3).pack(), pyramid_210(rows=4).pack()
    pw = concat([a, b])
    assert pw.n_worlds == 2
    assert list(pw.a["world_body_start"]) == [0, a.n_bodies, a.n_bodies + b.n_bodies]
    assert pw.a["shape_body"][a.n_shapes] == a.n_bodies + b.a["shape_body"][0]
    # hashids stay local to each space, as cpSpaceAddShape assigns them
    assert pw.a["shape_hashid"][a.n_shapes] == 0


def test_binding_exports_every_symbol_of_its_header():
    """oracle/_ref/libcpgpuspace.so (the cpGPUSpace binding, built against the
    reference's headers) exports every function include/cpGPUSpace.h declares."""
    import gpuspace
    if not gpuspace.available():
        pytest.skip("the binding is built only where the reference sources are present")
    text = open(os.path.join(ROOT, "include", "cpGPUSpace.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = sorted(set(re.findall(r"\b(cpGPU[A-Za-z]+)\s*\(", text)))
    assert len(names) >= 20
    gpuspace.cp()
    L = ctypes.CDLL(gpuspace.GPUSPACE_SO)
    for name in names:
        assert hasattr(L, name), name
