"""The adjoint of taped rollouts with joints -- needs an MI355X.

Pivot joints, damped springs, slide joints and gear joints are differentiated
by hand-written reverse kernels.  The reference has no derivative, so the
checks are mathematical:
  * finite differences of the device's own FP64 rollout for a loss linear in
    the final positions, to 1e-6 relative: spring stiffness, damping and
    rest length, slide limits, pivot anchors, gear phase, initial velocities
    and masses.  Without contacts the rollout is smooth, and a 4-point
    central stencil (truncation O(h^4)) with h = 1e-5 times the parameter's
    magnitude keeps the rounding noise of the differences (~1e-12 in the
    loss) far below the tolerance;
  * central differences (h = 1e-6) with falling circles colliding with a
    jointed chain, to 1e-5 (rollouts whose contact set changes under the
    perturbation are skipped);
  * gradients do not depend on the checkpoint interval (bit for bit), and the
    backward pass leaves the world -- joints' accumulated impulses included --
    exactly as it was.
"""
import numpy as np
import pytest

from chipmunk2d_amd.cabi import concat
from chipmunk2d_amd.gpu import GPUWorld
from chipmunk2d_amd import grad as G
from chipmunk2d_amd.scenes import Rng, chain_world
from chipmunk2d_amd.space import (Space, Body, Box, Circle, Segment, PivotJoint, DampedSpring, SlideJoint, GearJoint,
                                  moment_for_box)

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
PIVOT, SPRING, SLIDE, GEAR = 0, 1, 2, 3


def joint_zoo(seed, n_links=6):
    """A chain of boxes on pivots from a static anchor, with springs (i, i+2),
    slides (i, i+3) held beyond their maximum, two gears, a spring from the
    anchor to the last link; the links do not collide (one shape group)."""
    rng = np.random.default_rng(seed)
    space = Space()
    space.iterations = 10
    space.gravity = (0.0, -100.0)
    links, prev = [], space.static_body
    for i in range(n_links):
        body = space.add(Body(1.0 + 0.2 * i, moment_for_box(1.0 + 0.2 * i, 4.0, 1.0)))
        body.position = (4.0 * i + 2.0, 0.0)
        body.velocity = tuple(rng.normal(0.0, 2.0, 2))
        body.angular_velocity = float(rng.normal(0.0, 0.5))
        shape = space.add(Box(body, 4.0, 1.0, 0.0))
        shape.set_filter(group=1)
        space.add(PivotJoint(prev, body, (4.0 * i, 0.0)))
        links.append(body)
        prev = body
    for i in range(n_links - 2):
        space.add(DampedSpring(links[i], links[i + 2], (0.5, 0.0), (-0.5, 0.2), 6.5, rng.uniform(50.0, 500.0),
                               rng.uniform(1.0, 10.0)))
    for i in range(n_links - 3):
        space.add(SlideJoint(links[i], links[i + 3], (0.0, 0.0), (0.0, 0.0), 4.0, 10.0))
    space.add(GearJoint(links[0], links[1], 0.1, 1.0))
    space.add(GearJoint(links[3], links[4], 0.0, 2.0))
    space.add(DampedSpring(space.static_body, links[-1], (0.0, 5.0), (1.0, 0.0), 20.0, 100.0, 5.0))
    return space


def _rollout_loss(pw, weights, steps, counts=False):
    g = GPUWorld(pw)
    g.step(DT, steps)
    L = float(np.sum(weights * g.bodies().p))
    st = g.stats()
    g.close()
    return (L, (st["total_arbiters"], st["total_contacts"])) if counts else L


def _fd_check(pw, steps, cases, rel_tol, min_cases, seed=3, h_rel=1e-6, four_point=False):
    rng = np.random.default_rng(seed)
    dyn = (pw.a["body_type"] == G.DYNAMIC)[:, None]
    weights = rng.normal(size=(pw.n_bodies, 2)) * dyn
    g = G.rollout(pw, steps, DT, 10)
    grads = G.param_grads(pw, g.backward(g_p=weights))
    g.close()
    _, counts0 = _rollout_loss(pw, weights, steps, counts=True)
    report = []
    for kind, idx in cases:
        base = {"v0": pw.a["body_v"], "m": pw.a["body_m"], "joint": pw.a["cons_params"]}[kind][idx]
        eps = h_rel * max(1.0, abs(float(base)))
        vals = []
        for sgn in ((1.0, -1.0, 2.0, -2.0) if four_point else (1.0, -1.0)):
            v0 = pw.a["body_v"].copy()
            m = pw.a["body_m"].copy()
            jp = pw.a["cons_params"].copy()
            if kind == "v0":
                v0[idx] += sgn * eps
            elif kind == "m":
                m[idx] += sgn * eps
            else:
                jp[idx] += sgn * eps
            vals.append(_rollout_loss(G.with_params(pw, v0, m, None, jp), weights, steps, counts=True))
        if any(v[1] != counts0 for v in vals):
            continue   # the perturbation changed the contact set
        if four_point:
            fd = (8.0 * (vals[0][0] - vals[1][0]) - (vals[2][0] - vals[3][0])) / (12.0 * eps)
        else:
            fd = (vals[0][0] - vals[1][0]) / (2.0 * eps)
        ad = grads[kind][idx]
        report.append((kind, idx, ad, fd, abs(ad - fd) / max(abs(fd), 1e-2)))
    print("\n".join("%s %s adjoint %.12g fd %.12g rel %.2e" % r for r in report))
    assert len(report) >= min_cases, len(report)
    assert max(r[4] for r in report) < rel_tol, max(report, key=lambda r: r[4])
    return grads


def _joint_cases(pw):
    t = pw.a["cons_type"]
    cases = []
    for c in np.nonzero(t == SPRING)[0]:
        cases += [("joint", (int(c), 5)), ("joint", (int(c), 6)), ("joint", (int(c), 4))]
    for c in np.nonzero(t == SLIDE)[0]:
        cases += [("joint", (int(c), 5))]
    for c in np.nonzero(t == PIVOT)[0][:3]:
        cases += [("joint", (int(c), 0)), ("joint", (int(c), 3))]
    for c in np.nonzero(t == GEAR)[0]:
        cases += [("joint", (int(c), 0))]
    return cases


def test_joint_parameter_gradients_match_central_differences():
    pw = joint_zoo(11).pack()
    cases = _joint_cases(pw)
    dyn = np.nonzero(pw.a["body_type"] == G.DYNAMIC)[0]
    cases += [("v0", (int(b), k)) for b in dyn[:3] for k in (0, 1)]
    cases += [("m", int(b)) for b in dyn[1:4]]
    grads = _fd_check(pw, 40, cases, 1e-6, len(cases), h_rel=1e-5, four_point=True)
    assert np.count_nonzero(grads["spring_k"]) == np.count_nonzero(pw.a["cons_type"] == SPRING)


def test_chain_with_falling_circles_spring_gradients():
    """A short chain_world (pivots, springs, slides, gears) hit by circles."""
    pw = chain_world(n_links=12, n_circles=8, seed=5).pack()
    t = pw.a["cons_type"]
    springs = np.nonzero(t == SPRING)[0]
    cases = [("joint", (int(c), k)) for c in springs[::2] for k in (5, 6)]
    _fd_check(pw, 30, cases, 1e-5, len(cases) // 2)


def test_joint_gradients_independent_of_checkpoint_interval():
    pw = concat([joint_zoo(21).pack(), joint_zoo(22, n_links=5).pack()])
    rng = np.random.default_rng(1)
    targets = rng.normal(size=(pw.n_bodies, 2))
    ref = None
    for every in (1, 3, 7, 100):
        L, grads = G.loss_and_grads(pw, targets, 25, DT, checkpoint_every=every)
        if ref is None:
            ref = (L, grads)
            continue
        assert L == ref[0]
        for k in grads:
            assert np.array_equal(grads[k], ref[1][k]), (every, k)


def test_backward_restores_joints_and_bodies():
    pw = joint_zoo(31).pack()
    a, b = GPUWorld(pw), GPUWorld(pw)
    a.tape_begin(4)
    for g in (a, b):
        g.step(DT, 18)
    a.backward(g_p=np.ones((pw.n_bodies, 2)))
    a.tape_end()
    for g in (a, b):
        g.step(DT, 12)
    x, y = a.bodies(), b.bodies()
    for f in x.FIELDS:
        assert np.array_equal(getattr(x, f), getattr(y, f)), f
    assert np.array_equal(a.constraints()[0], b.constraints()[0])
