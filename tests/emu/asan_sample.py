"""A sample of the device tests over the AddressSanitizer build of the
emulation (tests/emu/_build_asan, made by EMU_ASAN=1 python tests/emu/build_emu.py),
where every device allocation has its exact size, so that any access past one
is reported:

    ASAN_OPTIONS=detect_leaks=0:detect_stack_use_after_return=0:halt_on_error=1 \\
    LD_PRELOAD=$(gcc -print-file-name=libasan.so) python tests/emu/asan_sample.py
"""
import os
import sys
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(os.path.dirname(HERE)), os.path.dirname(HERE)]
import chipmunk2d_amd.gpu as G
G.LIB_PATH = os.path.join(HERE, "_build_asan", "libcpgpu_emu.so")
from chipmunk2d_amd.scenes import pyramid_210, random_world, chain_world
from chipmunk2d_amd.cabi import concat
from test_gpu_parity import run_lockstep
print("workgroup", flush=True)
run_lockstep(concat([pyramid_210(rows=6).pack(), random_world(n_circles=80, n_polys=80, n_segments=20, extent=30.0, seed=3).pack()]), 20, check_every=5, arbiters_every=10)
print("chains", flush=True)
run_lockstep(chain_world(40, 60).pack(), 15, check_every=5, arbiters_every=15)
print("grid", flush=True)
run_lockstep(random_world(n_circles=2600, n_polys=2600, n_segments=300, extent=70.0, seed=23).pack(), 3, check_every=3, arbiters_every=3)
print("handlers", flush=True)
from test_handlers_gpu import _lockstep
from handler_rules import handler_space
_lockstep(handler_space(n_circles=60, n_polys=60, n_segments=15, seed=24).pack(), 15, check_every=5)
print("faults", flush=True)
from test_faults import test_device_reports_the_oracles_faults as check
for n in ("kinematic_pivot", "nan_velocity", "runaway", "zero_mass"):
    check(n)
print("queries", flush=True)
import numpy as np
from checkers import OracleWorld
from test_queries import _check_shape_queries
pw = random_world(n_circles=100, n_polys=100, n_segments=20, extent=30.0, seed=25).pack()
g, o = G.GPUWorld(pw), OracleWorld(pw)
for _ in range(5):
    g.step(1/60); o.step(1/60)
_check_shape_queries(g, o, pw)
print("tapes", flush=True)
import test_adjoint
test_adjoint.GPUWorld = G.GPUWorld
test_adjoint.test_a_second_tape_reuses_the_first_ones_buffers()
test_adjoint.test_backward_leaves_the_world_unchanged()
print("fp32", flush=True)
from chipmunk2d_amd.cabi import FP32
from chipmunk2d_amd.scenes import chains_packed
for pw in (chains_packed(n_worlds=64, processes=4),
           random_world(n_circles=2600, n_polys=2600, n_segments=300, extent=70.0, seed=23).pack()):
    g = G.GPUWorld(pw, FP32)
    g.step(1 / 60, 3)
    g.bodies()
    g.close()
print("ALL OK", flush=True)
