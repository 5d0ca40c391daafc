"""The checks a debug build of the reference makes during cpSpaceStep --
unsolvable collisions and constraints (chipmunk_private.h:223,255), the
positive mass and moment of a simulated body (cpBody.c:498) and the body
sanity check after each integration (cpBody.c:110-125,508,521) -- reported by
the oracle and the device as CPGPU_FAULT_* (cpgpuWorldGetFaults).

The oracle's reports are pinned against the reference itself, compiled
without -DNDEBUG (oracle/_ref/libcpref_debug.so): stepped in a child process,
it aborts at the step where the oracle first raises a fault, with the message
of that fault, and runs ordinary scenes to the end with no fault in either.
The device tests (GPU) check that the device raises the same faults at the
same step for the same first body, returns the reference's message once, and
keeps stepping the bodies the fault does not touch bit-exactly.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from chipmunk2d_amd.cabi import FAULT_BODY_MASS, FAULT_BODY_STATE, FAULT_UNSOLVABLE
from checkers import REF_DEBUG_SO, OracleWorld, RefWorld, ref_available
from fault_scenes import DT, SCENES, build

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

needs_debug_ref = pytest.mark.skipif(not os.path.exists(REF_DEBUG_SO),
                                     reason="oracle/_ref/libcpref_debug.so is built from the reference sources")

# the reference's abort messages, by the fault that reports the same condition
MESSAGES = {
    "Unsolvable collision or constraint.": FAULT_UNSOLVABLE,
    "Unsolvable constraint.": FAULT_UNSOLVABLE,
    "Body's mass and moment must be positive to simulate.": FAULT_BODY_MASS,
    "Body's mass is NaN.": FAULT_BODY_STATE,
    "Body's moment is NaN.": FAULT_BODY_STATE,
    "Body's mass is negative.": FAULT_BODY_STATE,
    "Body's moment is negative.": FAULT_BODY_STATE,
    "Body's position is invalid.": FAULT_BODY_STATE,
    "Body's velocity is invalid.": FAULT_BODY_STATE,
    "Body's force is invalid.": FAULT_BODY_STATE,
    "Body's angle is invalid.": FAULT_BODY_STATE,
    "Body's angular velocity is invalid.": FAULT_BODY_STATE,
    "Body's torque is invalid.": FAULT_BODY_STATE,
}

_CHILD = r"""
import sys
sys.path[:0] = [{root!r}, {tests!r}]
from checkers import RefWorld
{make}
r = RefWorld(pw, debug=True)
for k in range({steps}):
    print("step", k, flush=True)
    r.step({dt!r})
print("done", flush=True)
"""


def _debug_reference(make, steps):
    """Step a world made by the `make` source (which binds pw) in the debug
    build of the reference, in a child process: (last step begun, its abort
    message or None when all steps ran)."""
    code = _CHILD.format(root=ROOT, tests=HERE, make=make, steps=steps, dt=DT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    lines = r.stdout.split()
    steps_begun = [int(lines[i + 1]) for i, t in enumerate(lines) if t == "step"]
    if "done" in lines:
        assert r.returncode == 0, r.stderr[-2000:]
        return steps_begun[-1], None
    tag = "Aborting due to Chipmunk error: "
    assert r.returncode != 0 and tag in r.stderr, (r.returncode, r.stderr[-2000:])
    message = r.stderr.split(tag, 1)[1].splitlines()[0]
    return steps_begun[-1], message


def _first_fault(o, steps):
    for k in range(steps):
        o.step(DT)
        f, b = o.faults()
        if f:
            return k, f, b
    return None, 0, -1


@needs_debug_ref
@pytest.mark.parametrize("name", sorted(SCENES))
def test_oracle_faults_match_the_debug_reference(name):
    pw, bad, fault, step, message = build(name)
    k, f, b = _first_fault(OracleWorld(pw), step + 3)
    assert (k, b) == (step, bad), (k, b)
    assert f & fault, (f, fault)
    ref_step, ref_message = _debug_reference("from fault_scenes import build\npw = build(%r)[0]" % name, step + 3)
    assert ref_step == step and ref_message is not None
    assert ref_message.startswith(message), ref_message
    # the oracle raises the reference's condition, and at most the one its update makes next
    assert MESSAGES[message] & f
    assert f in (fault, FAULT_BODY_MASS | FAULT_BODY_STATE), f


@pytest.mark.skipif(not ref_available(), reason="oracle/_ref is built from the reference sources")
@pytest.mark.parametrize("name", sorted(SCENES))
def test_after_a_fault_the_oracle_steps_as_the_release_reference(name):
    """A release build of the reference checks nothing and steps on: replayed
    in its order, the oracle finds the same pairs every step (a shape whose box
    has a NaN side overlaps nothing, an infinite one only what it reaches) and
    leaves every body as the reference does, the failed ones included."""
    pw, bad, fault, step, message = build(name)
    o, r = OracleWorld(pw), RefWorld(pw)
    o.libm_trig(True)
    for k in range(step + 4):
        coll = r.step_logged(DT)
        o.replay(r.arbiters(), coll, DT)
        assert {(int(min(a, b)), int(max(a, b))) for a, b in coll[:, :2]} == \
            {tuple(map(int, p)) for p in o.candidate_pairs()}, k
        bo, br = o.bodies(), r.bodies()
        for f in ("p", "v", "a", "w"):
            assert np.array_equal(getattr(bo, f), getattr(br, f), equal_nan=True), (k, f)


@needs_debug_ref
@pytest.mark.parametrize("scene", ["pyramid", "chains", "random"])
def test_ordinary_scenes_raise_no_fault(scene):
    """Ordinary spaces -- a pyramid, chains of pivot joints and springs hung
    from static bodies, a random pile -- run in the debug reference without
    an abort and raise no fault in the oracle."""
    make = {
        "pyramid": "from chipmunk2d_amd.scenes import pyramid_210\npw = pyramid_210(rows=8).pack()",
        "chains": "from chipmunk2d_amd.scenes import chain_world\npw = chain_world(40, 60).pack()",
        "random": "from chipmunk2d_amd.scenes import random_world\n"
                  "pw = random_world(n_circles=120, n_polys=120, n_segments=30, extent=40.0, seed=7).pack()",
    }[scene]
    steps = 60
    ns = {}
    exec(make, ns)
    k, f, b = _first_fault(OracleWorld(ns["pw"]), steps)
    assert (k, f, b) == (None, 0, -1)
    last, message = _debug_reference(make, steps)
    assert (last, message) == (steps - 1, None)


def _healthy(pw, bad):
    """Bodies no fault of the scene touches: the pyramid and its ground.  Not
    the kinematic bodies of the joint scenes: the reference (and the oracle)
    add the NaN impulses of an unsolvable joint to them (v += j * 0), the
    device writes no velocity of a body that is not dynamic."""
    keep = np.ones(pw.n_bodies, bool)
    keep[pw.a["body_type"] == 1] = False      # kinematic bodies of the joint scenes
    if bad >= 0:
        keep[bad] = False
    return keep


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(SCENES))
def test_device_reports_the_oracles_faults(name):
    from chipmunk2d_amd.gpu import CpgpuError, GPUWorld
    pw, bad, fault, step, message = build(name)
    g, o = GPUWorld(pw), OracleWorld(pw)
    keep = _healthy(pw, bad)
    raised = None
    for k in range(step + 4):
        for call in (lambda: g.step(DT), g.sync):   # the step, or the synchronising call after it, reports
            try:
                call()
            except CpgpuError as e:
                assert raised is None, "a fault is reported once: %s" % e
                raised = (k, str(e))
        o.step(DT)
        if k < step:
            assert g.faults() == (0, -1) and o.faults() == (0, -1), k
        elif k == step:
            assert g.faults() == o.faults(), (g.faults(), o.faults())
        bg, bo = g.bodies(), o.bodies()
        assert np.array_equal(bg.p[keep], bo.p[keep]) and np.array_equal(bg.v[keep], bo.v[keep]), k
        if bad >= 0:   # the failed body itself, NaN where the oracle has NaN
            assert np.array_equal(bg.p[bad], bo.p[bad], equal_nan=True), (k, bg.p[bad], bo.p[bad])
            assert np.array_equal(bg.v[bad], bo.v[bad], equal_nan=True), (k, bg.v[bad], bo.v[bad])
    assert raised is not None and raised[0] == step, raised
    assert "cpgpu error -5: the step failed the reference's checks" in raised[1]
    text = {FAULT_UNSOLVABLE: "Unsolvable collision or constraint", FAULT_BODY_MASS: "mass and moment must be positive",
            FAULT_BODY_STATE: "state is invalid"}[fault]
    assert text in raised[1], raised[1]
    if bad >= 0:
        assert "first body: %d;" % bad in raised[1], raised[1]
