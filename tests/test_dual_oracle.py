"""The oracle's step on dual numbers (oracle/dual_step.cpp) -- CPU only.

The dual build is the second check of the device adjoint (with finite
differences); it is pinned here before it checks anything:
  * its value parts equal the double oracle's bit for bit, step after step,
    on every kind of space (polygons, segments, joints, a batch);
  * its tangents are the derivative: a contact-free projectile's closed form
    (p_T = p_0 + T dt v_0 + ...: dp_T / dv_0 = T dt exactly), and central
    finite differences of the double oracle on rollouts whose contact set does
    not change, for velocities, masses and friction coefficients.
"""
import numpy as np
import pytest

from chipmunk2d_amd.scenes import Rng, chain_world, grad_world, machine_world, pyramid_210, random_world
from chipmunk2d_amd.cabi import concat
from chipmunk2d_amd import grad as G
from chipmunk2d_amd.space import Space, Body, Circle, moment_for_circle
from checkers import DualWorld, OracleWorld

DT = 1.0 / 60.0


def _grad_batch(n_worlds, n_bodies, seed):
    rng = Rng(seed)
    packs, targets = [], []
    for _ in range(n_worlds):
        sp, tg = grad_world(rng, n_bodies, vel_sigma=10.0)
        packs.append(sp.pack())
        targets.append(np.vstack([np.zeros((1, 2)), tg]))
    return concat(packs), np.vstack(targets)


@pytest.mark.parametrize("scene", ["pyramid", "random", "chain", "grad", "machines"])
def test_value_parts_equal_the_double_oracle(scene):
    pw = {"pyramid": lambda: pyramid_210(rows=10).pack(),
          "random": lambda: random_world(n_circles=150, n_polys=150, n_segments=30, extent=50.0, seed=81).pack(),
          "chain": lambda: chain_world(20, 40, seed=9).pack(),
          "grad": lambda: _grad_batch(4, 16, 0xD1)[0],
          "machines": lambda: machine_world(n_units=4, n_circles=30).pack()}[scene]()
    rng = np.random.default_rng(1)
    o = OracleWorld(pw)
    d = DualWorld(pw, dv=rng.normal(size=(pw.n_bodies, 2)), dm=rng.normal(size=pw.n_bodies),
                  du=rng.normal(size=pw.n_shapes))
    for s in range(60):
        o.step(DT)
        d.step(DT)
        if s % 15 == 14:
            ob = o.bodies()
            val, tan = d.bodies()
            for f in ("p", "v", "a", "w"):
                assert np.array_equal(val[f], getattr(ob, f)), (scene, s, f)
    assert np.any(tan["p"] != 0.0)


def test_projectile_tangent_closed_form():
    space = Space()
    space.gravity = (0.0, -100.0)
    body = space.add(Body(2.0, moment_for_circle(2.0, 0.0, 1.0)))
    body.position = (1.0, 500.0)
    body.velocity = (3.0, 7.0)
    space.add(Circle(body, 1.0))
    pw = space.pack()
    k = int(np.nonzero(pw.a["body_type"] == G.DYNAMIC)[0][0])
    dv = np.zeros((pw.n_bodies, 2))
    dv[k] = (1.0, -2.0)
    d = DualWorld(pw, dv=dv)
    T = 25
    d.step(DT, T)
    val, tan = d.bodies()
    # positions are integrated before velocities: p_T = p_0 + dt sum_{t<T} v_t, each v_t = v_0 + t g dt
    assert np.allclose(tan["p"][k], np.array([1.0, -2.0]) * T * DT, rtol=1e-14, atol=0)
    assert np.array_equal(tan["v"][k], np.array([1.0, -2.0]))


def _fd(pw, steps, v0=None, m=None, u=None, eps=1e-6):
    """Central differences of the final positions along a direction (v0, m, u), and whether
    the two perturbed rollouts kept the unperturbed arbiters (pairs and counts)."""
    def run(h):
        a = pw.a
        pv = G.with_params(pw, a["body_v"] + (h * v0 if v0 is not None else 0.0),
                           a["body_m"] + (h * m if m is not None else 0.0),
                           a["shape_u"] + (h * u if u is not None else 0.0))
        o = OracleWorld(pv)
        sig = []
        for _ in range(steps):
            o.step(DT)
            A = o.arbiters()
            sig.append((A.n, A.shapes.tobytes(), A.count.tobytes()))
        return o.bodies().p, sig
    (pp, sp), (pm, sm) = run(eps), run(-eps)
    (p0, s0) = run(0.0)
    return (pp - pm) / (2 * eps), sp == s0 and sm == s0


@pytest.mark.parametrize("kind", ["v0", "m", "u"])
def test_tangents_match_central_differences(kind):
    """A small pile of circles and boxes on ground: the dual tangent of every final
    position along a random direction equals the central difference (eps 1e-6) of
    the double oracle, on seeds whose contact set stays fixed."""
    checked = 0
    for seed in range(20):
        pw, _ = _grad_batch(1, 8, 0xD200 + seed)
        rng = np.random.default_rng(seed)
        dyn = (pw.a["body_type"] == G.DYNAMIC).astype(float)
        v0 = m = u = None
        if kind == "v0":
            v0 = rng.normal(size=(pw.n_bodies, 2)) * dyn[:, None]
        elif kind == "m":
            m = rng.normal(size=pw.n_bodies) * dyn * 0.1
        else:
            u = rng.normal(size=pw.n_shapes) * dyn[pw.a["shape_body"]] * 0.1
        fd, fixed = _fd(pw, 40, v0, m, u)
        if not fixed:
            continue
        d = DualWorld(pw, dv=v0, dm=m, du=u)
        d.step(DT, 40)
        _, tan = d.bodies()
        scale = max(np.abs(fd).max(), 1e-3)
        err = np.abs(tan["p"] - fd).max() / scale
        assert err < 1e-6, (kind, seed, err)
        checked += 1
        if checked == 4:
            break
    assert checked >= 2, (kind, checked)
