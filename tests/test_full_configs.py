"""The three batched workloads at their full sizes on the device -- needs an MI355X.

  * 4096 spaces of log-uniform size in [16, 2048] (pyramids and random piles,
    1.67M bodies): the size buckets autotuned, then stepped in lockstep with
    the oracle, bodies, arbiters and candidate pairs bit-exact;
  * 2048 jointed chains (200 links with pivot, damped-spring, slide and gear
    joints, 300 falling circles each; 1.03M bodies, 1.26M joints): FP64 in
    lockstep with the oracle, joints' accumulated impulses bit-exact, and the
    FP32 run's drift from it;
  * 1024 spaces of 64 circles and boxes launched onto ground: the adjoint of
    the 120-step rollout (checkpoints every 10 steps) against central finite
    differences with eps = 1e-6, one perturbed parameter per space so one
    pair of rollouts differentiates every space at once.

Each test prints what it checked.  The oracle steps a few steps of the two
forward workloads (about 4 s a step for the 4096 spaces).
"""
import time

import numpy as np
import pytest

from chipmunk2d_amd import grad as G
from chipmunk2d_amd.cabi import FP32
from chipmunk2d_amd.drift import compare
from chipmunk2d_amd.gpu import GPUWorld
from chipmunk2d_amd.scenes import batch_packed, chains_packed, grad_batch
from checkers import OracleWorld
from test_gpu_parity import assert_arbiters_equal, assert_bodies_equal, assert_pairs_equal

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _log(msg, t0):
    print("[%6.1f s] %s" % (time.time() - t0, msg), flush=True)


def _assert_stats_equal(g, o):
    gs, os_ = g.stats(), o.stats()
    for k in ("candidate_pairs", "arbiters", "contacts", "cached_arbiters", "arbiter_colours"):
        assert gs[k] == os_[k], (k, gs[k], os_[k])
    return gs


def test_batch_4096_autotuned_lockstep_bit_exact():
    t0 = time.time()
    pw = batch_packed()
    assert pw.n_worlds == 4096 and pw.n_bodies > 1_500_000
    _log("packed 4096 spaces, %d bodies" % pw.n_bodies, t0)
    g = GPUWorld(pw)
    threads = g.autotune(DT, 1)          # five steps, one at each workgroup size
    _log("autotuned bucket threads %s" % threads, t0)
    o = OracleWorld(pw)
    o.step(DT, 5)
    assert_bodies_equal(g, o, "after the autotuner's 5 steps")
    _log("bodies bit-exact after 5 steps", t0)
    for s in range(2):
        g.step(DT)
        o.step(DT)
        assert_bodies_equal(g, o, "step %d" % (6 + s))
    assert_pairs_equal(g, o, "step 7")
    assert_arbiters_equal(g, o)
    st = _assert_stats_equal(g, o)
    _log("step 7: %d candidate pairs, %d arbiters, %d contacts bit-exact" %
         (st["candidate_pairs"], st["arbiters"], st["contacts"]), t0)
    assert st["arbiters"] > 1_000_000


def test_chains_2048_lockstep_bit_exact_and_fp32_drift():
    t0 = time.time()
    pw = chains_packed()
    assert pw.n_worlds == 2048 and pw.n_constraints > 1_000_000
    _log("packed 2048 chains: %d bodies, %d joints" % (pw.n_bodies, pw.n_constraints), t0)
    g, o = GPUWorld(pw), OracleWorld(pw)
    for s in range(4):
        g.step(DT)
        o.step(DT)
        assert_bodies_equal(g, o, "step %d" % s)
    assert_pairs_equal(g, o, "step 3")
    assert_arbiters_equal(g, o)
    jg, cg = g.constraints()
    jo, co = o.constraints()
    assert np.array_equal(cg, co)
    assert np.array_equal(jg, jo)
    st = _assert_stats_equal(g, o)
    _log("4 steps bit-exact: %d arbiters, %d joints' impulses" % (st["arbiters"], pw.n_constraints), t0)
    # FP32 beside the FP64 run (which is the oracle's, bit for bit)
    g32 = GPUWorld(pw, FP32)
    g32.step(DT, 4)
    rec = compare(pw, g.bodies(), g32.bodies(), st, g32.stats())
    _log("FP32 after 4 steps: max |dp| %.3e (rel %.3e), max |dv| %.3e, contacts %d / %d" %
         (rec["max_dp"], rec["max_dp_rel"], rec["max_dv"], rec["contacts_fp32"], rec["contacts_fp64"]), t0)
    assert rec["finite"]
    # positions are O(100): one float ulp there is 7.6e-6; four steps keep a few ulps
    assert rec["max_dp_rel"] < 2e-5, rec
    assert abs(rec["contacts_fp32"] - rec["contacts_fp64"]) <= 0.01 * rec["contacts_fp64"] + 4, rec


def _contact_signature(g, world_of_shape, n_worlds):
    """Per space: arbiters, contacts, states and shape pairs of the last step."""
    A = g.arbiters()
    sig = np.zeros((n_worlds, 4), np.int64)
    if A.n:
        w = world_of_shape[A.shapes[:, 0]]
        lo = np.minimum(A.shapes[:, 0], A.shapes[:, 1]).astype(np.int64)
        hi = np.maximum(A.shapes[:, 0], A.shapes[:, 1]).astype(np.int64)
        np.add.at(sig[:, 0], w, 1)
        np.add.at(sig[:, 1], w, A.count.astype(np.int64))
        np.add.at(sig[:, 2], w, A.state.astype(np.int64))
        np.add.at(sig[:, 3], w, lo * 1000003 + hi)
    return sig


def _rollout(pw, steps, world_of_shape, tape_every=0):
    g = GPUWorld(pw)
    if tape_every:
        g.tape_begin(tape_every)
    sig = np.zeros((pw.n_worlds, 4), np.int64)
    for _ in range(steps):
        g.step(DT)
        sig += _contact_signature(g, world_of_shape, pw.n_worlds)
    return g, sig


def _world_losses(p, targets, dyn, world_of_body, n_worlds):
    d = (p - targets) * dyn[:, None]
    return np.bincount(world_of_body, weights=np.sum(d * d, axis=1), minlength=n_worlds)


def _world_loss_diffs(p_plus, p_minus, targets, dyn, world_of_body, n_worlds):
    """L(+) - L(-) per space, as the sum of (p+ - p-) . (p+ + p- - 2 target).
    Subtracting the two losses instead would round the difference at the
    loss's own magnitude: with L ~ 1e4 that is ~1e-12, as large as the whole
    difference of a gradient of 1e-3 at eps 1e-6."""
    d = np.sum((p_plus - p_minus) * (p_plus + p_minus - 2.0 * targets), axis=1) * dyn
    return np.bincount(world_of_body, weights=d, minlength=n_worlds)


def test_grad_1024_adjoint_matches_central_differences():
    """The loss is the sum over bodies of |p_T - target|^2.  Space k perturbs
    one parameter (an initial velocity component in half of the spaces, a
    mass or a friction coefficient in a quarter each), so the k-th space's loss difference is the central
    difference of that parameter.  Spaces whose contact set changes under any
    perturbation are left out (the contact set is held fixed by the adjoint),
    and so are those whose differences at eps and 2 eps disagree (a clamp of
    the solver switching inside the stencil)."""
    t0 = time.time()
    steps, every, eps = 120, 10, 1e-6
    pw, targets = grad_batch(1024, 64)
    a = pw.a
    n_worlds = pw.n_worlds
    world_of_body = np.repeat(np.arange(n_worlds), np.diff(a["world_body_start"]))
    world_of_shape = np.repeat(np.arange(n_worlds), np.diff(a["world_shape_start"]))
    dyn = a["body_type"] == G.DYNAMIC
    _log("packed 1024 x 64 bodies", t0)

    g, sig0 = _rollout(pw, steps, world_of_shape, tape_every=every)
    p = g.bodies().p
    L0 = _world_losses(p, targets, dyn.astype(np.float64), world_of_body, n_worlds)
    grads = G.param_grads(pw, g.backward(g_p=2.0 * (p - targets) * dyn[:, None]))
    g.close()
    _log("taped rollout and backward", t0)

    rng = np.random.default_rng(0xC4)
    kinds = ("v0", "m", "u")
    cases = []
    for k in range(n_worlds):
        bodies = np.nonzero(dyn & (world_of_body == k))[0]
        # velocities, whose central differences are the noisiest, in half of the spaces
        kind = ("v0", "m", "v0", "u")[k % 4]
        if kind == "v0":
            cases.append((kind, (int(rng.choice(bodies)), int(rng.integers(2)))))
        elif kind == "m":
            cases.append((kind, int(rng.choice(bodies))))
        else:
            shapes = np.nonzero((world_of_shape == k) & dyn[a["shape_body"]])[0]
            cases.append((kind, int(rng.choice(shapes))))

    def perturbed(h):
        v0, m, u = a["body_v"].copy(), a["body_m"].copy(), a["shape_u"].copy()
        for kind, idx in cases:
            if kind == "v0":
                v0[idx] += h
            elif kind == "m":
                m[idx] += h
            else:
                u[idx] += h
        return G.with_params(pw, v0, m, u)

    P, same = {}, np.ones(n_worlds, bool)
    for h in (eps, -eps, 2 * eps, -2 * eps):
        gh, sig = _rollout(perturbed(h), steps, world_of_shape)
        P[h] = gh.bodies().p
        gh.close()
        same &= np.all(sig == sig0, axis=1)
    _log("four perturbed rollouts; %d of %d spaces keep their contact set" % (same.sum(), n_worlds), t0)

    dynf = dyn.astype(np.float64)
    fd1 = _world_loss_diffs(P[eps], P[-eps], targets, dynf, world_of_body, n_worlds) / (2 * eps)
    fd2 = _world_loss_diffs(P[2 * eps], P[-2 * eps], targets, dynf, world_of_body, n_worlds) / (4 * eps)
    ad = np.array([grads[kind][idx] for kind, idx in cases])
    report = {kind: [] for kind in kinds}
    rejected = {kind: 0 for kind in kinds}
    for k, (kind, idx) in enumerate(cases):
        if not same[k]:
            continue
        scale = max(abs(fd1[k]), 1e-3)
        if abs(fd1[k] - fd2[k]) > 1e-7 * scale:
            rejected[kind] += 1
            continue
        report[kind].append((k, ad[k], fd1[k], abs(ad[k] - fd1[k]) / scale))
    for kind in kinds:
        r = report[kind]
        errs = np.array([x[3] for x in r]) if r else np.zeros(1)
        worst = max(r, key=lambda x: x[3]) if r else None
        _log("%s: %d spaces checked (%d left out: a clamp switching inside the stencil); max rel err %.2e, "
             "median %.2e; worst space %s" % (kind, len(r), rejected[kind], errs.max(), np.median(errs), worst), t0)
    # about a third of the spaces keep their contact set under all four perturbations, and a
    # clamp switches inside the stencil in some of those: 20-34 spaces per kind are left (the
    # CPU emulation of this rollout, bit-exact with the oracle, checks 34, 24 and 20)
    for kind in kinds:
        assert len(report[kind]) >= 20, (kind, len(report[kind]), rejected[kind])
        assert max(x[3] for x in report[kind]) < 1e-6, (kind, sorted(report[kind], key=lambda x: -x[3])[:5])
    assert np.all(np.isfinite(L0))


def test_grad_1024_adjoint_matches_dual_numbers():
    """The same 120-step rollout of 1024 x 64 bodies, differentiated forward on
    the CPU by the oracle's step on dual numbers (oracle/dual_step.cpp): its
    value parts are the device's positions bit for bit, and for random
    directions e of the initial velocities, of the masses and of the friction
    coefficients, each space's directional derivative sum 2 (p_T - target) .
    dp_T/de equals the adjoint's e . grad -- no finite differences, every
    space (the contact set is the rollout's in both)."""
    from checkers import DualWorld
    t0 = time.time()
    steps, every = 120, 10
    pw, targets = grad_batch(1024, 64)
    a = pw.a
    n_worlds = pw.n_worlds
    world_of_body = np.repeat(np.arange(n_worlds), np.diff(a["world_body_start"]))
    world_of_shape = np.repeat(np.arange(n_worlds), np.diff(a["world_shape_start"]))
    dyn = (a["body_type"] == G.DYNAMIC).astype(np.float64)
    g = GPUWorld(pw)
    g.tape_begin(every)
    g.step(DT, steps)
    p = g.bodies().p
    grads = G.param_grads(pw, g.backward(g_p=2.0 * (p - targets) * dyn[:, None]))
    g.close()
    _log("taped rollout and backward", t0)
    rng = np.random.default_rng(0xD0A1)
    directions = {"v0": dict(dv=rng.normal(size=(pw.n_bodies, 2)) * dyn[:, None]),
                  "m": dict(dm=rng.normal(size=pw.n_bodies) * dyn),
                  "u": dict(du=rng.normal(size=pw.n_shapes) * dyn[a["shape_body"]])}
    for kind, seed in directions.items():
        d = DualWorld(pw, **seed)
        d.step(DT, steps)
        val, tan = d.bodies()
        assert np.array_equal(val["p"], p), kind
        dual = np.bincount(world_of_body, weights=np.sum(2.0 * (p - targets) * dyn[:, None] * tan["p"], axis=1),
                           minlength=n_worlds)
        if kind == "v0":
            e_grad = np.sum(seed["dv"] * grads["v0"], axis=1)
            adj = np.bincount(world_of_body, weights=e_grad, minlength=n_worlds)
        elif kind == "m":
            adj = np.bincount(world_of_body, weights=seed["dm"] * grads["m"], minlength=n_worlds)
        else:
            adj = np.bincount(world_of_shape, weights=seed["du"] * grads["u"], minlength=n_worlds)
        # relative to the space's own derivative (1e-3 floor)
        rel = np.abs(adj - dual) / np.maximum(np.abs(dual), 1e-3)
        _log("%s: %d spaces, max rel diff %.2e, median %.2e, 99th pct %.2e" %
             (kind, n_worlds, rel.max(), np.median(rel), np.percentile(rel, 99)), t0)
        assert rel.max() < 1e-6, (kind, int(np.argmax(rel)), rel.max(), adj[np.argmax(rel)], dual[np.argmax(rel)])
