"""Every joint class on the device -- needs an MI355X.

machine_world puts all ten constraint classes of the reference to work: pin,
groove, rotary limit, ratchet, simple motor and damped rotary spring beside
pivot, damped spring, slide and gear (the oracle's restatement of the six is
pinned step by step against the compiled reference in
tests/test_oracle_pins.py).  The device steps it bit for bit like the oracle:
  * one space of eight mechanisms, bodies every few steps, the accumulated
    impulses and ratchet angles at the end;
  * a batch of machine spaces of different seeds with a chain of pivots,
    springs, slides and gears, solved one workgroup each;
  * a machine space of more than 4096 dynamic bodies, solved on the grid;
  * FP32 stays finite and near FP64 over the first steps;
  * the adjoint, which has no reverse for these six classes, refuses them.
"""
import numpy as np
import pytest

from chipmunk2d_amd.cabi import FP32, concat
from chipmunk2d_amd.gpu import CpgpuError, GPUWorld
from chipmunk2d_amd.scenes import chain_world, machine_world
from test_gpu_parity import run_lockstep

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _assert_constraints_equal(g, o):
    jg, cg = g.constraints()
    jo, co = o.constraints()
    assert np.array_equal(cg, co)
    assert np.array_equal(jg, jo)
    return jg


def test_machine_space_bit_exact():
    pw = machine_world(n_units=8, n_circles=60).pack()
    g, o = run_lockstep(pw, 240, check_every=3, arbiters_every=60)
    j = _assert_constraints_equal(g, o)
    t = pw.a["cons_type"]
    # every class carried an impulse at the end or (a ratchet) an angle off its start
    active = np.bincount(t[np.any(j != 0.0, axis=1)], minlength=10)
    print("classes carrying state after 240 steps:", active.tolist(), flush=True)
    assert np.count_nonzero(active) >= 8, active
    assert not np.array_equal(j[t == 7, 1], pw.a["cons_jacc"][t == 7, 1])


def test_machine_batch_bit_exact():
    spaces = [machine_world(n_units=2 + k % 3, n_circles=10 + 5 * (k % 4), seed=100 + k).pack() for k in range(24)]
    spaces.append(chain_world(n_links=20, n_circles=10, seed=5).pack())
    pw = concat(spaces)
    g, o = run_lockstep(pw, 120, check_every=10, arbiters_every=60)
    _assert_constraints_equal(g, o)


def test_machine_space_on_the_grid_bit_exact():
    pw = machine_world(n_units=520, n_circles=1400).pack()
    n_dyn = int(np.count_nonzero(pw.a["body_type"] == 0))
    assert n_dyn > 4096, n_dyn
    g, o = run_lockstep(pw, 30, check_every=5, arbiters_every=15)
    _assert_constraints_equal(g, o)
    print("%d dynamic bodies, %d constraints: 30 grid steps bit-exact" % (n_dyn, pw.n_constraints), flush=True)


def test_machine_space_fp32():
    pw = machine_world(n_units=8, n_circles=60).pack()
    g64, g32 = GPUWorld(pw), GPUWorld(pw, FP32)
    g64.step(DT, 10)
    g32.step(DT, 10)
    b64, b32 = g64.bodies(), g32.bodies()
    assert np.allclose(b32.p, b64.p, atol=5e-2), np.abs(b32.p - b64.p).max()
    g32.step(DT, 190)
    b32 = g32.bodies()
    assert np.all(np.isfinite(b32.p)) and np.all(np.isfinite(b32.v))


def test_adjoint_refuses_the_other_joint_classes():
    g = GPUWorld(machine_world(n_units=1, n_circles=0).pack())
    with pytest.raises(CpgpuError, match="pivot, damped-spring, slide and gear joints only"):
        g.tape_begin(10)
    g.step(DT, 5)   # the world still steps
    assert np.all(np.isfinite(g.bodies().p))
