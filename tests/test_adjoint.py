"""The adjoint of a taped rollout (cpgpuWorldTapeBegin / cpgpuWorldBackward)
-- needs an MI355X.

The reference has no derivative, so the checks are mathematical:
  * a contact-free projectile, whose gradient has a closed form;
  * central finite differences (eps 1e-6) of the device's own FP64 rollout,
    on rollouts whose contact set and counts do not change under the
    perturbation, for initial velocities, masses and friction coefficients;
  * the checkpoint interval changes only how the backward pass replays, so
    gradients are bit-identical for every interval;
  * the backward pass leaves the world exactly as it was;
  * the torch.autograd wrapper returns the library's gradients.
"""
import numpy as np
import pytest

from chipmunk2d_amd.cabi import concat
from chipmunk2d_amd.gpu import GPUWorld
from chipmunk2d_amd import grad as G
from chipmunk2d_amd.scenes import Rng, grad_world, grad_batch
from chipmunk2d_amd.space import Space, Body, Circle, moment_for_circle

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def small_batch(n_worlds=3, n_bodies=6, seed=0xADA1):
    rng = Rng(seed)
    packs, targets = [], []
    for _ in range(n_worlds):
        sp, tg = grad_world(rng, n_bodies, vel_sigma=10.0)
        packs.append(sp.pack())
        targets.append(np.vstack([np.zeros((1, 2)), tg]))
    return concat(packs), np.vstack(targets)


def test_projectile_gradient_closed_form():
    """No contacts: p_T = p_0 + T dt v_0 + g dt^2 T(T-1)/2 (positions are
    integrated before velocities), so dL/dv_0 = 2 (p_T - target) T dt."""
    space = Space()
    space.gravity = (0.0, -100.0)
    body = space.add(Body(2.0, moment_for_circle(2.0, 0.0, 1.0)))
    body.position = (1.0, 500.0)
    body.velocity = (3.0, 7.0)
    space.add(Circle(body, 1.0))
    pw = space.pack()
    k = int(np.nonzero(pw.a["body_type"] == G.DYNAMIC)[0][0])
    T = 25
    g = G.rollout(pw, T, DT)
    p = g.bodies().p
    target = np.zeros((pw.n_bodies, 2))
    target[k] = (4.0, 480.0)
    dyn = (pw.a["body_type"] == G.DYNAMIC).astype(np.float64)
    L, gp = G.position_loss(p, target, dyn)
    adj = g.backward(g_p=gp)
    expect = 2.0 * (p[k] - target[k]) * T * DT
    assert np.allclose(adj["v"][k], expect, rtol=1e-13, atol=0.0)
    assert np.allclose(adj["p"][k], 2.0 * (p[k] - target[k]), rtol=1e-13, atol=0.0)
    # the mass does not enter a force-free trajectory
    assert adj["m_inv"][k] == 0.0


def test_checkpoint_interval_gives_identical_gradients():
    pw, targets = small_batch()
    ref = None
    for every in (1, 4, 7, 200):
        L, grads = G.loss_and_grads(pw, targets, 30, DT, checkpoint_every=every)
        if ref is None:
            ref = (L, grads)
        else:
            assert L == ref[0]
            for k in grads:
                assert np.array_equal(grads[k], ref[1][k]), (every, k)


def test_backward_leaves_the_world_unchanged():
    pw, targets = small_batch(seed=0xADA2)
    a = GPUWorld(pw)
    b = GPUWorld(pw)
    a.tape_begin(5)
    for g in (a, b):
        g.step(DT, 20)
    a.backward(g_p=np.ones((pw.n_bodies, 2)))
    a.tape_end()
    for g in (a, b):
        g.step(DT, 15)
    x, y = a.bodies(), b.bodies()
    for f in x.FIELDS:
        assert np.array_equal(getattr(x, f), getattr(y, f)), f


def test_a_second_tape_reuses_the_first_ones_buffers():
    """A world taped twice (the second tape on the checkpoint and record
    buffers the first one left, with a different interval) gives the
    gradients of a fresh world taped once from the same state."""
    pw, _ = small_batch(seed=0xADA4)
    g_p = np.ones((pw.n_bodies, 2))
    a = GPUWorld(pw)
    a.tape_begin(3)
    a.step(DT, 12)
    a.backward(g_p=g_p)
    a.tape_end()
    a.tape_begin(5)
    a.step(DT, 20)
    second = a.backward(g_p=g_p)
    a.tape_end()
    b = GPUWorld(pw)
    b.step(DT, 12)
    b.tape_begin(5)
    b.step(DT, 20)
    fresh = b.backward(g_p=g_p)
    b.tape_end()
    for k in fresh:
        assert np.array_equal(second[k], fresh[k]), k


def _linear_loss(pw, weights, steps):
    """L = sum of weights . p_T, and the rollout's contact counts."""
    g = GPUWorld(pw)
    g.step(DT, steps)
    L = float(np.sum(weights * g.bodies().p))
    st = g.stats()
    g.close()
    return L, (st["total_arbiters"], st["total_contacts"], st["total_pairs"])


def test_adjoint_matches_central_differences():
    """Initial velocities, masses and friction coefficients, eps = 1e-6.  The
    loss is linear in the final positions (random weights): its rounding noise
    is small next to the gradients, so central differences resolve them to
    about 1e-8; a quadratic loss of positions ~30 carries ~1e-12 of noise in
    L, i.e. ~1e-6 in the differences."""
    pw, _ = small_batch(seed=0xADA3)
    steps = 40
    rng = np.random.default_rng(5)
    weights = rng.normal(size=(pw.n_bodies, 2)) * (pw.a["body_type"] == G.DYNAMIC)[:, None]
    g = G.rollout(pw, steps, DT, 10)
    grads = G.param_grads(pw, g.backward(g_p=weights))
    g.close()
    _, counts0 = _linear_loss(pw, weights, steps)
    dyn = np.nonzero(pw.a["body_type"] == G.DYNAMIC)[0]
    cases = []
    for b in rng.choice(dyn, 8, replace=False):
        cases.append(("v0", (int(b), int(rng.integers(2)))))
    for b in rng.choice(dyn, 6, replace=False):
        cases.append(("m", int(b)))
    shapes = np.nonzero(pw.a["body_type"][pw.a["shape_body"]] == G.DYNAMIC)[0]
    for s in rng.choice(shapes, 6, replace=False):
        cases.append(("u", int(s)))
    eps = 1e-6
    report = []
    for kind, idx in cases:
        vals = []
        for sgn in (1.0, -1.0):
            v0 = pw.a["body_v"].copy()
            m = pw.a["body_m"].copy()
            u = pw.a["shape_u"].copy()
            if kind == "v0":
                v0[idx] += sgn * eps
            elif kind == "m":
                m[idx] += sgn * eps
            else:
                u[idx] += sgn * eps
            vals.append(_linear_loss(G.with_params(pw, v0, m, u), weights, steps))
        if vals[0][1] != counts0 or vals[1][1] != counts0:
            continue   # the perturbation changed the contact set
        fd = (vals[0][0] - vals[1][0]) / (2.0 * eps)
        ad = grads[kind][idx]
        report.append((kind, idx, ad, fd, abs(ad - fd) / max(abs(fd), 1e-2)))
    print("\n".join("%s %s adjoint %.12g fd %.12g rel %.2e" % r for r in report))
    assert len(report) >= 14, len(report)
    assert max(r[4] for r in report) < 1e-6, report


def test_torch_autograd_wrapper():
    torch = pytest.importorskip("torch")
    pw, targets = small_batch(seed=0xADA4)
    steps = 25
    v0 = torch.tensor(pw.a["body_v"], dtype=torch.float64, requires_grad=True)
    m = torch.tensor(pw.a["body_m"], dtype=torch.float64, requires_grad=True)
    u = torch.tensor(pw.a["shape_u"], dtype=torch.float64, requires_grad=True)
    p = G.torch_rollout(v0, m, u, pw, steps, DT, 5)
    dyn = torch.tensor((pw.a["body_type"] == G.DYNAMIC).astype(np.float64))
    loss = (((p - torch.tensor(targets)) ** 2).sum(dim=1) * dyn).sum()
    loss.backward()
    L, grads = G.loss_and_grads(pw, targets, steps, DT, 5)
    assert abs(loss.item() - L) <= 1e-9 * abs(L)
    assert np.array_equal(v0.grad.numpy(), grads["v0"])
    dynb = pw.a["body_type"] == G.DYNAMIC
    assert np.array_equal(m.grad.numpy()[dynb], grads["m"][dynb])
    assert np.array_equal(u.grad.numpy(), grads["u"])


def test_grad_batch_workload_runs():
    """A slice of the 1024 x 64 differentiation workload: every gradient finite."""
    pw, targets = grad_batch(n_worlds=16, n_bodies=64)
    L, grads = G.loss_and_grads(pw, targets, 60, DT, checkpoint_every=10)
    assert np.isfinite(L)
    for k, v in grads.items():
        assert np.all(np.isfinite(v)), k
    assert np.count_nonzero(grads["u"]) > 0
