"""Seeded workload generators for the step benchmarks and tests.

One project PRNG (xoshiro256** seeded through SplitMix64) so every host
produces the same scene from the same seed.  Scene builders use the host
mirror (chipmunk2d_amd.space) exactly as a Chipmunk program would use the C
API, and return a `Space` (or a list of them) ready for `pack()`.

Workloads (sizes as named):
  pyramid_210       20-row pyramid of 210 boxes on a ground segment, slop 0.5
  random_world      circles + convex polys falling on a static container with
                    sinusoidal terrain; default 50k circles, 40k polys, 10k segments
  batch_worlds      many independent spaces of log-uniform body counts, each a
                    pyramid or a random pile on its own ground
  grad_batch        many spaces of 64 circles / boxes launched onto a ground
                    segment, with target positions (differentiation)
  chain_world       a pivot-joint chain with springs, slides and gears, and
                    falling circles
  machine_world     mechanisms with every joint class (pin, groove, rotary
                    limit, ratchet, simple motor, damped rotary spring beside
                    pivot, spring, slide, gear), and falling circles
"""
import math

import numpy as np

from .space import Space, Body, Circle, Segment, Poly, Box, moment_for_box, moment_for_circle, PivotJoint, \
    DampedSpring, SlideJoint, GearJoint, PinJoint, GrooveJoint, RotaryLimitJoint, RatchetJoint, SimpleMotor, \
    DampedRotarySpring

MASK64 = (1 << 64) - 1

SEED_PYRAMID = 0xC4100001
SEED_RANDOM_WORLD = 0xC4100002
SEED_BATCH = 0xC4100003
SEED_GRAD = 0xC4100004
SEED_CHAINS = 0xC4100005
SEED_MACHINES = 0xC4100006


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & MASK64


class Rng:
    """xoshiro256** (Blackman & Vigna), seeded through SplitMix64."""

    def __init__(self, seed):
        s = seed & MASK64
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & MASK64
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
            self.s.append(z ^ (z >> 31))

    def next_u64(self):
        s = self.s
        result = (_rotl((s[1] * 5) & MASK64, 7) * 9) & MASK64
        t = (s[1] << 17) & MASK64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = _rotl(s[3], 45)
        return result

    def uniform(self, lo=0.0, hi=1.0):
        u = (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def randint(self, lo, hi):
        """Integer in [lo, hi] inclusive."""
        lo, hi = int(lo), int(hi)
        return lo + self.next_u64() % (hi - lo + 1)

    def normal(self, mu=0.0, sigma=1.0):
        u1 = self.uniform()
        u2 = self.uniform()
        r = math.sqrt(-2.0 * math.log(1.0 - u1))
        return mu + sigma * r * math.cos(2.0 * math.pi * u2)


def pyramid_210(rows=20, ground_friction=0.6, x_offset=0.0, space=None):
    """20-row pyramid of boxes (side 30, bevel 0.5, mass 1, friction 0.6) on a
    static ground segment from (-500,-240) to (500,-240); gravity (0,-100),
    10 iterations, collision slop 0.5.  Row r holds rows-r boxes spaced 32."""
    if space is None:
        space = Space()
        space.iterations = 10
        space.gravity = (0.0, -100.0)
        space.collision_slop = 0.5
    ground = space.add(Segment(space.static_body, (x_offset - 500.0, -240.0), (x_offset + 500.0, -240.0), 0.0))
    ground.elasticity = 0.0
    ground.friction = ground_friction
    for r in range(rows):
        for j in range(rows - r):
            body = space.add(Body(1.0, moment_for_box(1.0, 30.0, 30.0)))
            body.position = (x_offset + (j - (rows - 1 - r) / 2.0) * 32.0, -240.0 + 15.5 + 31.0 * r)
            shape = space.add(Box(body, 30.0, 30.0, 0.5))
            shape.elasticity = 0.0
            shape.friction = 0.6
    return space


def _random_poly_verts(rng, k, radius):
    verts = []
    for _ in range(k):
        ang = rng.uniform(0.0, 2.0 * math.pi)
        verts.append((radius * math.cos(ang), radius * math.sin(ang)))
    return verts


def random_world(n_circles=50000, n_polys=40000, n_segments=10000, extent=2000.0, seed=SEED_RANDOM_WORLD,
                 terrain_amplitude=20.0, terrain_period=200.0, segment_radius=0.5, vel_sigma=5.0):
    """Circles (radius U[0.5,1.5]) and convex polys (hull of 3-8 points at a
    radius U[1,2], bevel 0.05), density 1, friction U[0.3,0.9], elasticity 0,
    positions uniform over an extent x extent box, velocities N(0, vel_sigma);
    a static container: 10% of the segments form the two side walls, the rest
    a sinusoidal terrain along the bottom.  Gravity (0,-100), 10 iterations."""
    rng = Rng(seed)
    space = Space()
    space.iterations = 10
    space.gravity = (0.0, -100.0)
    half = extent / 2.0
    n_wall = max(2, n_segments // 10) if n_segments else 0
    n_terrain = n_segments - n_wall
    sb = space.static_body
    if n_wall:
        per_side = n_wall // 2
        for side, x in ((0, -half - 10.0), (1, half + 10.0)):
            cnt = per_side if side == 0 else n_wall - per_side
            y0, y1 = -half - 50.0, half + 50.0
            for i in range(cnt):
                a = (x, y0 + (y1 - y0) * i / cnt)
                b = (x, y0 + (y1 - y0) * (i + 1) / cnt)
                s = space.add(Segment(sb, a, b, segment_radius))
                s.friction = 0.7
    if n_terrain:
        x0, x1 = -half - 10.0, half + 10.0
        ybase = -half - 30.0

        def ty(x):
            return ybase + terrain_amplitude * math.sin(2.0 * math.pi * x / terrain_period)

        for i in range(n_terrain):
            xa = x0 + (x1 - x0) * i / n_terrain
            xb = x0 + (x1 - x0) * (i + 1) / n_terrain
            s = space.add(Segment(sb, (xa, ty(xa)), (xb, ty(xb)), segment_radius))
            s.friction = 0.7
    # dynamic bodies interleaved so circles and polys share the index range
    kinds = [0] * n_circles + [1] * n_polys
    for i in range(len(kinds) - 1, 0, -1):
        j = rng.randint(0, i)
        kinds[i], kinds[j] = kinds[j], kinds[i]
    for kind in kinds:
        body = space.add(Body(0.0, 0.0))
        if kind == 0:
            shape = space.add(Circle(body, rng.uniform(0.5, 1.5)))
        else:
            k = rng.randint(3, 8)
            verts = _random_poly_verts(rng, k, rng.uniform(1.0, 2.0))
            shape = space.add(Poly(body, verts, radius=0.05))
        shape.density = 1.0
        shape.friction = rng.uniform(0.3, 0.9)
        shape.elasticity = 0.0
        body.position = (rng.uniform(-half, half), rng.uniform(-half, half))
        body.velocity = (rng.normal(0.0, vel_sigma), rng.normal(0.0, vel_sigma))
    return space


def random_pile(n_bodies, rng, width=None, x_offset=0.0):
    """A random circle/poly pile of n_bodies falling on its own ground segment."""
    space = Space()
    space.iterations = 10
    space.gravity = (0.0, -100.0)
    if width is None:
        width = max(40.0, 4.0 * math.sqrt(n_bodies) * 2.0)
    half = width / 2.0
    ground = space.add(Segment(space.static_body, (x_offset - half - 20.0, 0.0), (x_offset + half + 20.0, 0.0), 0.5))
    ground.friction = 0.7
    height = 4.0 * n_bodies * 4.0 / width + 20.0
    for _ in range(n_bodies):
        body = space.add(Body(0.0, 0.0))
        if rng.uniform() < 0.5:
            shape = space.add(Circle(body, rng.uniform(0.5, 1.5)))
        else:
            k = rng.randint(3, 8)
            shape = space.add(Poly(body, _random_poly_verts(rng, k, rng.uniform(1.0, 2.0)), radius=0.05))
        shape.density = 1.0
        shape.friction = rng.uniform(0.3, 0.9)
        body.position = (x_offset + rng.uniform(-half, half), rng.uniform(3.0, 3.0 + height))
        body.velocity = (rng.normal(0.0, 5.0), rng.normal(0.0, 5.0))
    return space


def grad_world(rng, n_bodies=64, vel_sigma=20.0):
    """One space of the differentiation workload: n_bodies circles or boxes
    (radius or half-side U[0.5,2], mass U[0.5,2], friction U[0.2,0.8],
    elasticity 0) on an 8-wide grid 6 apart above a ground segment (friction
    0.7), launched with velocities N(0, vel_sigma); gravity (0,-100), 10
    iterations.  Returns (space, targets [n_bodies, 2]) with seeded target
    positions U[-40,40] x U[0,30].  The moment follows the mass (i = k m for
    the body's geometry)."""
    space = Space()
    space.iterations = 10
    space.gravity = (0.0, -100.0)
    ground = space.add(Segment(space.static_body, (-80.0, 0.0), (80.0, 0.0), 0.5))
    ground.friction = 0.7
    for k in range(n_bodies):
        size = rng.uniform(0.5, 2.0)
        m = rng.uniform(0.5, 2.0)
        if rng.uniform() < 0.5:
            body = space.add(Body(m, moment_for_circle(m, 0.0, size)))
            shape = space.add(Circle(body, size))
        else:
            body = space.add(Body(m, moment_for_box(m, 2.0 * size, 2.0 * size)))
            shape = space.add(Box(body, 2.0 * size, 2.0 * size, 0.0))
        shape.friction = rng.uniform(0.2, 0.8)
        shape.elasticity = 0.0
        body.position = (-21.0 + 6.0 * (k % 8), 5.0 + 6.0 * (k // 8))
        body.velocity = (rng.normal(0.0, vel_sigma), rng.normal(0.0, vel_sigma))
    targets = np.array([(rng.uniform(-40.0, 40.0), rng.uniform(0.0, 30.0)) for _ in range(n_bodies)])
    return space, targets


def grad_batch(n_worlds=1024, n_bodies=64, seed=SEED_GRAD):
    """The differentiation batch: n_worlds spaces of grad_world.  Returns
    (packed description, targets [n_bodies_total, 2] with zeros for static
    bodies)."""
    from .cabi import concat
    rng = Rng(seed)
    packs, targets = [], []
    for _ in range(n_worlds):
        sp, tg = grad_world(rng, n_bodies)
        packs.append(sp.pack())
        targets.append(np.vstack([np.zeros((1, 2)), tg]))   # the static body comes first
    return concat(packs), np.vstack(targets)


def chain_world(n_links=200, n_circles=300, seed=SEED_CHAINS, rng=None):
    """A chain of n_links boxes (4 x 1, mass 1) joined by pivot joints
    (collideBodies off) and hung from a static body; every link i >= 2 also
    has a damped spring to link i-2 (rest length 8, stiffness U[50,500],
    damping U[1,10]) and a slide joint to it (limits [6, 8]); a gear joint
    (ratio 1) joins every 10th link to the next; n_circles circles (radius
    U[0.5,1.5], density 1) fall on the chain.  Gravity (0,-100), 10
    iterations."""
    if rng is None:
        rng = Rng(seed)
    space = Space()
    space.iterations = 10
    space.gravity = (0.0, -100.0)
    x0, y0 = -2.0 * n_links, 0.0
    links = []
    prev = space.static_body
    for i in range(n_links):
        body = space.add(Body(1.0, moment_for_box(1.0, 4.0, 1.0)))
        body.position = (x0 + 4.0 * i + 2.0, y0)
        shape = space.add(Box(body, 4.0, 1.0, 0.0))
        shape.friction = 0.6
        pivot = (x0 + 4.0 * i, y0)
        j = space.add(PivotJoint(prev, body, pivot))
        j.collide_bodies = False
        links.append(body)
        prev = body
    for i in range(2, n_links):
        a, b = links[i - 2], links[i]
        space.add(DampedSpring(a, b, (0.0, 0.0), (0.0, 0.0), 8.0, rng.uniform(50.0, 500.0), rng.uniform(1.0, 10.0)))
        space.add(SlideJoint(a, b, (0.0, 0.0), (0.0, 0.0), 6.0, 8.0))
    for i in range(0, n_links - 1, 10):
        space.add(GearJoint(links[i], links[i + 1], 0.0, 1.0))
    for _ in range(n_circles):
        body = space.add(Body(0.0, 0.0))
        shape = space.add(Circle(body, rng.uniform(0.5, 1.5)))
        shape.density = 1.0
        shape.friction = 0.5
        body.position = (rng.uniform(x0, -x0), rng.uniform(10.0, 10.0 + 0.5 * n_links))
    return space


def machine_world(n_units=8, n_circles=60, seed=SEED_MACHINES, rng=None):
    """n_units mechanisms side by side (12 apart) on a ground segment, each using
    every joint class, and n_circles circles (radius U[0.4,1.2]) falling on them.
    Gravity (0,-100), 10 iterations.  Unit k at x = 12 k:
      * a two-link pendulum of circles hung from the static body by pin joints,
        the links joined by a damped rotary spring (rest angle U[-1,1]);
      * a wheel on a pivot driven by a simple motor (rate U[-4,4], max force
        U[2e3, 2e4]) and an arm on the same pivot held to the wheel by a rotary
        limit joint ([-0.4, 0.4]);
      * a slider on a tilted groove of the static body (it slides into the
        groove's lower end) and a second slider on a groove along the arm;
      * a paddle on a pivot with a ratchet (notch pi/8 or -pi/8, phase U[0,1])
        against the static body, wound by a damped rotary spring;
      * a gear joint (ratio U[0.5,2]) between the paddle and the wheel every
        other unit, a damped spring and a slide joint between the pendulum's
        tip and the slider.
    Max bias and error bias vary between units."""
    if rng is None:
        rng = Rng(seed)
    space = Space()
    space.iterations = 10
    space.gravity = (0.0, -100.0)
    st = space.static_body
    width = 12.0 * n_units + 20.0
    ground = space.add(Segment(st, (-20.0, -2.0), (width, -2.0), 0.5))
    ground.friction = 0.8

    def circle_body(pos, r, mass=1.0):
        body = space.add(Body(mass, moment_for_circle(mass, 0.0, r)))
        body.position = pos
        sh = space.add(Circle(body, r))
        sh.friction = 0.5
        return body

    def box_body(pos, w, h, mass=1.0, angle=0.0):
        body = space.add(Body(mass, moment_for_box(mass, w, h)))
        body.position = pos
        body.angle = angle
        sh = space.add(Box(body, w, h, 0.0))
        sh.friction = 0.6
        return body

    def tune(j, k):
        if k % 3 == 1:
            j.max_bias = 40.0
        if k % 4 == 2:
            j.error_bias = math.pow(1.0 - 0.2, 60.0)
        return space.add(j)

    for k in range(n_units):
        x = 12.0 * k
        # pendulum: pin joints, a damped rotary spring between the links
        p1 = circle_body((x + 3.0, 22.0), 0.8)
        p2 = circle_body((x + 5.5, 19.0), 0.6, mass=0.5)
        tune(PinJoint(st, p1, (x, 24.0), (0.0, 0.0)), k)
        tune(PinJoint(p1, p2, (0.3, 0.0), (0.0, 0.4)), k)
        tune(DampedRotarySpring(p1, p2, rng.uniform(-1.0, 1.0), rng.uniform(20.0, 200.0), rng.uniform(1.0, 20.0)), k)
        # wheel, motor, arm held by a rotary limit
        wheel = circle_body((x, 6.0), 2.0, mass=2.0)
        tune(PivotJoint(st, wheel, (x, 6.0)), k)
        motor = tune(SimpleMotor(st, wheel, rng.uniform(-4.0, 4.0)), k)
        motor.max_force = rng.uniform(2.0e3, 2.0e4)
        arm = box_body((x + 3.0, 6.0), 6.0, 0.5, mass=0.8)
        piv = tune(PivotJoint(wheel, arm, (x, 6.0)), k)
        piv.collide_bodies = False
        tune(RotaryLimitJoint(wheel, arm, -0.4, 0.4), k)
        # sliders on grooves: one in the static body, one along the arm
        s1 = box_body((x + 2.0, 14.0), 1.5, 1.0, mass=0.6)
        tune(GrooveJoint(st, s1, (x - 3.0, 16.0), (x + 4.0, 12.5), (0.0, 0.0)), k)
        s2 = circle_body((x + 4.0, 6.8), 0.4, mass=0.3)
        g2 = tune(GrooveJoint(arm, s2, (-2.5, 0.8), (2.5, 0.8), (0.0, 0.0)), k)
        g2.collide_bodies = False
        # paddle: pivot, ratchet against the static body, wound by a rotary spring
        paddle = box_body((x + 8.0, 10.0), 3.0, 0.6, mass=1.0, angle=rng.uniform(-0.5, 0.5))
        tune(PivotJoint(st, paddle, (x + 8.0, 10.0)), k)
        notch = math.pi / 8 if k % 2 == 0 else -math.pi / 8
        tune(RatchetJoint(st, paddle, rng.uniform(0.0, 1.0), notch), k)
        tune(DampedRotarySpring(st, paddle, rng.uniform(-3.0, 3.0), rng.uniform(50.0, 400.0), rng.uniform(2.0, 30.0)), k)
        if k % 2 == 0:
            tune(GearJoint(paddle, wheel, 0.0, rng.uniform(0.5, 2.0)), k)
        tune(DampedSpring(p2, s1, (0.0, 0.0), (0.0, 0.0), rng.uniform(3.0, 6.0), rng.uniform(10.0, 100.0),
                          rng.uniform(0.5, 5.0)), k)
        tune(SlideJoint(p2, s1, (0.0, 0.0), (0.0, 0.0), 2.0, 7.0), k)
    for _ in range(n_circles):
        r = rng.uniform(0.4, 1.2)
        body = space.add(Body(0.0, 0.0))
        sh = space.add(Circle(body, r))
        sh.density = 1.0
        sh.friction = 0.5
        body.position = (rng.uniform(-5.0, width - 20.0), rng.uniform(26.0, 60.0))
    return space


def _pack_chain(seed):
    return chain_world(rng=Rng(seed)).pack()


def chains_packed(n_worlds=2048, seed=SEED_CHAINS, processes=8):
    """The constraints-and-contacts batch: n_worlds chain_world spaces (200
    links, 300 circles), each seeded from the batch seed, built in worker
    processes and packed into one description."""
    from .cabi import concat
    rng = Rng(seed)
    seeds = [rng.next_u64() for _ in range(n_worlds)]
    if processes and processes > 1:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(processes) as pool:
            packed = pool.map(_pack_chain, seeds, chunksize=16)
    else:
        packed = [_pack_chain(s) for s in seeds]
    return concat(packed)


def batch_specs(n_worlds=4096, lo=16, hi=2048, seed=SEED_BATCH, pyramid_fraction=0.5):
    """(kind, body count, world seed) of each space of a batch: counts
    log-uniform in [lo, hi], pyramid or random pile, each with its own seed
    drawn from the batch seed (so worlds can be built in any order)."""
    rng = Rng(seed)
    specs = []
    for _ in range(n_worlds):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        kind = "pyramid" if rng.uniform() < pyramid_fraction else "pile"
        specs.append((kind, n, rng.next_u64()))
    return specs


def build_batch_world(spec):
    """One space of a batch: a pyramid stack (rows chosen so rows*(rows+1)/2
    is near the count) or a random pile, on its own ground."""
    kind, n, wseed = spec
    if kind == "pyramid":
        rows = max(1, int((math.sqrt(8 * n + 1) - 1) / 2))
        return pyramid_210(rows=rows)
    return random_pile(n, Rng(wseed))


def batch_worlds(n_worlds=4096, lo=16, hi=2048, seed=SEED_BATCH, pyramid_fraction=0.5):
    """Independent spaces with body counts log-uniform in [lo, hi]."""
    return [build_batch_world(s) for s in batch_specs(n_worlds, lo, hi, seed, pyramid_fraction)]


def _pack_spec(spec):
    return build_batch_world(spec).pack()


def batch_packed(n_worlds=4096, lo=16, hi=2048, seed=SEED_BATCH, pyramid_fraction=0.5, processes=8):
    """The batch packed into one description, worlds built in worker processes."""
    from .cabi import concat
    specs = batch_specs(n_worlds, lo, hi, seed, pyramid_fraction)
    if processes and processes > 1:
        import multiprocessing as mp
        with mp.get_context("fork").Pool(processes) as pool:
            packed = pool.map(_pack_spec, specs, chunksize=16)
    else:
        packed = [_pack_spec(s) for s in specs]
    return concat(packed)
