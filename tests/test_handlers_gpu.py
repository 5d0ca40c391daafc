"""Collision handler callbacks around the device step (cpgpuWorldStepBegin /
ReadCallbacks / WriteDecisions / StepEnd / ReadSeparations) -- needs an MI355X.

The same deterministic callbacks (handler_rules.Rule) answer the device's
due arbiters and the oracle's, whose handler semantics are pinned against the
compiled reference (test_collision_handlers_replay_matches_reference).  Each
step the same pairs get begin and preSolve and the same pairs separate; the
bodies and the solved arbiters stay bit-exact.
"""
import numpy as np
import pytest

from chipmunk2d_amd.cabi import EVENT_BEGIN, EVENT_PRE_SOLVE, concat
from chipmunk2d_amd.gpu import CpgpuError, GPUWorld
from chipmunk2d_amd.scenes import pyramid_210
from checkers import OracleWorld
from handler_rules import Rule, handler_space
from test_gpu_parity import assert_arbiters_equal, assert_bodies_equal

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _lockstep(pw, steps, check_every=5):
    g, o = GPUWorld(pw), OracleWorld(pw)
    rg, ro = Rule(), Rule()
    o.set_handler_fn(ro)
    n_sep = 0
    for s in range(steps):
        rg.step = ro.step = s
        g.step_with_callbacks(DT, rg.begin, rg.pre_solve)
        o.step(DT)
        for kind in (EVENT_BEGIN, EVENT_PRE_SOLVE):
            eg = {e for e in rg.events(kind) if e[0] == s}
            eo = {e for e in ro.events(kind) if e[0] == s}
            assert eg == eo, (s, kind, sorted(eg ^ eo)[:6])
        sg = {(int(x[0]), int(min(x[1], x[2])), int(max(x[1], x[2]))) for x in g.separations()}
        so = {(int(x[0]), int(min(x[1], x[2])), int(max(x[1], x[2]))) for x in o.separations()}
        assert sg == so, (s, sorted(sg ^ so)[:6])
        n_sep += len(sg)
        if (s + 1) % check_every == 0 or s == steps - 1:
            assert_bodies_equal(g, o, "step %d" % s)
    assert_arbiters_equal(g, o)
    return g, o, rg, n_sep


def test_callbacks_match_the_oracle():
    pw = handler_space().pack()
    g, o, rule, n_sep = _lockstep(pw, 120)
    counts = {k: len(rule.events(k)) for k in (EVENT_BEGIN, EVENT_PRE_SOLVE)}
    print("device callbacks over 120 steps:", counts, "separations:", n_sep, flush=True)
    assert counts[EVENT_BEGIN] > 100 and counts[EVENT_PRE_SOLVE] > 1000 and n_sep > 50


def test_callbacks_in_a_batch():
    """Two handler spaces beside a plain pyramid: each space's handlers are its own."""
    pw = concat([handler_space(seed=31).pack(), pyramid_210(rows=6).pack(),
                 handler_space(n_circles=80, n_polys=80, seed=32).pack()])
    assert len(pw.a["handlers"]) == 8
    _lockstep(pw, 60, check_every=10)


def test_default_answers_and_refusals():
    """cpgpuWorldStep answers every callback with the defaults: solved, values
    unchanged, as the oracle with no handler function."""
    pw = handler_space(seed=33).pack()
    g, o = GPUWorld(pw), OracleWorld(pw)
    for s in range(40):
        g.step(DT)
        o.step(DT)
    assert_bodies_equal(g, o, "default answers")
    assert_arbiters_equal(g, o)
    with pytest.raises(CpgpuError, match="no step begun"):
        g.step_end()
    g.step_begin(DT)
    with pytest.raises(CpgpuError, match="not ended"):
        g.step(DT)
    g.step_end()
    with pytest.raises(CpgpuError, match="without graphs"):
        g.set_graph(True)
    with pytest.raises(CpgpuError, match="collision handlers"):
        g.tape_begin(10)
