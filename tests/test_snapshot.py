"""Snapshots of a world's state (cpgpuWorldSaveState / cpgpuWorldLoadState):
checkpoint and resume -- needs an MI355X.

A world saved after some steps and loaded into a new world made from the
same description steps on bit for bit like the world it was saved from, and
like the oracle stepping the whole rollout without a pause: bodies, the
arbiters with their cached impulses, the pair set, the joints' impulses and
ratchet angles, the running totals.  Checked on a mixed batch (pyramids,
piles, a chain), on the machine scene (every joint class), on a space solved
on the grid, and for FP32.  Snapshots of another description, precision or
length are refused.
"""
import numpy as np
import pytest

from chipmunk2d_amd.cabi import FP32, concat
from chipmunk2d_amd.gpu import CpgpuError, GPUWorld
from chipmunk2d_amd.scenes import chain_world, machine_world, pyramid_210, random_world
from checkers import OracleWorld
from test_gpu_parity import assert_arbiters_equal, assert_bodies_equal, assert_pairs_equal

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _same_worlds(a, b, what):
    ba, bb = a.bodies(), b.bodies()
    for f in ba.FIELDS:
        assert np.array_equal(getattr(ba, f), getattr(bb, f)), (what, f)
    A, B = a.arbiters(), b.arbiters()
    assert A.n == B.n, what
    for f in ("shapes", "state", "count", "normal", "contact", "hash"):
        assert np.array_equal(getattr(A, f), getattr(B, f)), (what, f)
    assert np.array_equal(a.pairs(), b.pairs()), what
    if a.packed.n_constraints:
        assert np.array_equal(a.constraints()[0], b.constraints()[0]), what
    sa, sb = a.stats(), b.stats()
    for k in ("steps", "total_pairs", "total_arbiters", "total_contacts", "cached_arbiters"):
        assert sa[k] == sb[k], (what, k, sa[k], sb[k])


def _resume_check(pw, before, after, precision=0, oracle=True):
    g = GPUWorld(pw, precision)
    g.step(DT, before)
    snap = g.save_state()
    g.step(DT, after)
    h = GPUWorld(pw, precision)
    h.load_state(snap)
    h.step(DT, after)
    _same_worlds(g, h, "resumed after %d + %d steps" % (before, after))
    if oracle:
        o = OracleWorld(pw)
        o.step(DT, before + after)
        assert_bodies_equal(h, o, "resumed world against the oracle")
        assert_arbiters_equal(h, o)
        assert_pairs_equal(h, o, "resumed world against the oracle")
    return len(snap)


def test_mixed_batch_resumes_bit_exact():
    pw = concat([pyramid_210(rows=10).pack(), random_world(n_circles=400, n_polys=400, n_segments=60, extent=50.0,
                                                           seed=61).pack(), chain_world(25, 40, seed=6).pack()])
    size = _resume_check(pw, 45, 30)
    print("snapshot of %d bodies, %d shapes: %d bytes" % (pw.n_bodies, pw.n_shapes, size), flush=True)


def test_machine_scene_resumes_bit_exact():
    _resume_check(machine_world(n_units=6, n_circles=40).pack(), 80, 60)


def test_grid_space_resumes_bit_exact():
    pw = random_world(n_circles=2600, n_polys=2600, n_segments=400, extent=140.0, seed=12).pack()
    _resume_check(pw, 20, 10)


def test_fp32_resumes_bit_identically():
    pw = random_world(n_circles=300, n_polys=300, n_segments=40, extent=50.0, seed=62).pack()
    _resume_check(pw, 30, 30, precision=FP32, oracle=False)


def test_loading_before_the_first_step_and_refusals():
    pw = pyramid_210(rows=6).pack()
    g = GPUWorld(pw)
    snap0 = g.save_state()        # a world not yet stepped
    g.step(DT, 10)
    h = GPUWorld(pw)
    h.load_state(snap0)
    h.step(DT, 10)
    _same_worlds(g, h, "from the unstepped state")
    other = GPUWorld(pyramid_210(rows=7).pack())
    with pytest.raises(CpgpuError, match="another description"):
        other.load_state(snap0)
    with pytest.raises(CpgpuError, match="other precision"):
        GPUWorld(pw, FP32).load_state(snap0)
    with pytest.raises(CpgpuError, match="bytes"):
        h.load_state(snap0[:-8])
    bad = bytearray(snap0)
    bad[0:8] = b"notasnap"
    with pytest.raises(CpgpuError, match="not a cpgpu snapshot"):
        h.load_state(bytes(bad))
