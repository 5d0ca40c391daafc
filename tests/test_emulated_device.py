"""The step library's device code run on the CPU (tests/emu: the csrc sources
built with g++ against a HIP stand-in), checked against the oracle in this
container, where there is no GPU.  A short sample of the GPU tests, so that
every run of the CPU suite also exercises the device kernels' logic: the
workgroup solve, the cooperative grid solve with its strips and LDS-kept
colours, collision handler callbacks and shape queries.  The whole GPU suite
runs the same way with `CPGPU_EMULATE=1 python -m pytest tests -m gpu`.
"""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_SO = os.path.join(HERE, "emu", "_build", "libcpgpu_emu.so")
DT = 1.0 / 60.0


def _sources_newer_than(path):
    csrc = os.path.join(os.path.dirname(HERE), "chipmunk2d_amd", "csrc")
    t = os.path.getmtime(path)
    files = [os.path.join(csrc, f) for f in os.listdir(csrc)] + \
        [os.path.join(HERE, "emu", f) for f in ("emu_runtime.cpp", "build_emu.py")] + \
        [os.path.join(HERE, "emu", "hip", "hip_runtime.h")]
    return any(os.path.getmtime(f) > t for f in files if os.path.isfile(f))


@pytest.fixture(scope="module")
def emulated():
    """chipmunk2d_amd.gpu over the emulated library for this module's tests."""
    if not (shutil.which("g++") and shutil.which("hipcc")):
        pytest.skip("the emulation is built with g++ and takes one figure from hipcc")
    if not os.path.exists(EMU_SO) or _sources_newer_than(EMU_SO):
        r = subprocess.run([sys.executable, os.path.join(HERE, "emu", "build_emu.py")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
    import chipmunk2d_amd.gpu as G
    saved_path, saved_lib = G.LIB_PATH, G._lib
    G.LIB_PATH, G._lib = EMU_SO, None
    yield G
    G.LIB_PATH, G._lib = saved_path, saved_lib


def test_workgroup_solve_bit_exact(emulated):
    from chipmunk2d_amd.scenes import pyramid_210, random_world
    from chipmunk2d_amd.cabi import concat
    from test_gpu_parity import run_lockstep
    pw = concat([pyramid_210(rows=8).pack(),
                 random_world(n_circles=120, n_polys=120, n_segments=30, extent=40.0, seed=3).pack()])
    run_lockstep(pw, 40, check_every=5, arbiters_every=20)


def test_grid_solve_bit_exact(emulated):
    """A crowded space of more than 4096 dynamic bodies: the cooperative grid
    solve, some strips kept whole in LDS and of others the smallest colours."""
    from chipmunk2d_amd.scenes import random_world
    from test_gpu_parity import run_lockstep
    pw = random_world(n_circles=2600, n_polys=2600, n_segments=300, extent=70.0, seed=23).pack()
    run_lockstep(pw, 6, check_every=3, arbiters_every=6)


def test_collision_handler_callbacks_bit_exact(emulated):
    from test_handlers_gpu import _lockstep
    from handler_rules import handler_space
    _lockstep(handler_space(n_circles=80, n_polys=80, n_segments=20, seed=24).pack(), 30, check_every=10)


def test_shape_queries_bit_exact(emulated):
    from chipmunk2d_amd.scenes import random_world
    from checkers import OracleWorld
    from test_queries import _check_shape_queries
    pw = random_world(n_circles=150, n_polys=150, n_segments=30, extent=40.0, seed=25).pack()
    g, o = emulated.GPUWorld(pw), OracleWorld(pw)
    for _ in range(10):
        g.step(DT)
        o.step(DT)
    assert np.array_equal(g.bodies().p, o.bodies().p)
    assert _check_shape_queries(g, o, pw) > 50


def test_device_allocation_failures_are_reported(emulated):
    """Every device allocation of a world's creation and first steps made to fail
    in turn (the emulation's hipMalloc fails on demand): each one comes back as
    an error from the C ABI -- no null or stale buffer is ever launched on -- and
    a world made afterwards steps normally."""
    import ctypes
    from chipmunk2d_amd.scenes import pyramid_210, random_world
    from chipmunk2d_amd.cabi import concat
    L = ctypes.CDLL(EMU_SO)
    L.emu_fail_malloc_after.argtypes = [ctypes.c_long]
    L.emu_malloc_countdown.restype = ctypes.c_long
    pw = concat([pyramid_210(rows=5).pack(), random_world(n_circles=60, n_polys=60, n_segments=10, extent=25.0,
                                                          seed=26).pack()])
    failed, k = 0, 1
    while True:
        L.emu_fail_malloc_after(k)
        try:
            g = emulated.GPUWorld(pw)
            g.step(DT, 4)
            g.bodies()
            g.close()
        except emulated.CpgpuError:
            failed += 1
        left = L.emu_malloc_countdown()
        L.emu_fail_malloc_after(0)
        if left > 0:   # fewer than k allocations: every one has been made to fail
            break
        k += 1
    assert failed == k - 1 and k > 20, (failed, k)
    g = emulated.GPUWorld(pw)
    g.step(DT, 4)
    assert np.all(np.isfinite(g.bodies().p))


def test_faults_reported_as_the_oracle(emulated):
    """The reference's debug-build checks: an unsolvable joint, a NaN state, a
    position overflowing some steps in; the same faults, step and first body
    as the oracle, and the bodies the fault does not touch still bit-exact."""
    from test_faults import test_device_reports_the_oracles_faults as check
    for name in ("kinematic_pivot", "nan_velocity", "runaway"):
        check(name)
