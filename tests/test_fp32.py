"""FP32 worlds (cpgpuWorldCreate with CPGPU_FP32) -- needs an MI355X.

The FP32 step runs the FP64 step's kernels compiled with float arithmetic,
so it is checked for drift against the double-precision oracle, not for bit
equality:
  * after one step the state is within a few hundred float ulps of the
    oracle's (float rounding of positions ~1e3 is ~6e-5);
  * a settled scene (the 210-box pyramid, the 80-link chain) stays within a
    small fraction of a body size of the FP64 run after 300 steps;
  * schedule independence holds in FP32 too: a space stepped inside a batch
    is bit-identical to the same space alone, and runs repeat bit for bit;
  * the adjoint refuses FP32 worlds.
"""
import numpy as np
import pytest

from chipmunk2d_amd.cabi import concat, FP32, FP64
from chipmunk2d_amd.drift import drift_report, compare
from chipmunk2d_amd.gpu import GPUWorld, CpgpuError
from chipmunk2d_amd.scenes import pyramid_210, random_pile, random_world, chain_world, Rng
from checkers import OracleWorld

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _show(rep):
    for r in rep:
        print("step %(step)d max_dp %(max_dp).3e rms_dp %(rms_dp).3e max_da %(max_da).3e max_dv %(max_dv).3e "
              "ke %(kinetic_fp64).4g/%(kinetic_fp32).4g mean_y %(mean_y_fp64).6g/%(mean_y_fp32).6g "
              "contacts %(contacts_fp64)d/%(contacts_fp32)d" % r)


def test_fp32_first_step_within_float_rounding_of_oracle():
    rng = Rng(0xF32)
    pw = concat([random_pile(400, rng).pack(), pyramid_210(rows=10).pack()])
    o = OracleWorld(pw)
    g = GPUWorld(pw, FP32)
    o.step(DT)
    g.step(DT)
    rec = compare(pw, o.bodies(), g.bodies(), o.stats(), g.stats())
    _show([dict(rec, step=1)])
    assert rec["finite"]
    # positions are O(100): one float ulp there is 7.6e-6
    assert rec["max_dp"] < 2e-3, rec
    assert rec["max_dp_rel"] < 2e-5, rec
    assert rec["kinetic_rel"] < 1e-4, rec
    assert g.stats()["candidate_pairs"] == o.stats()["candidate_pairs"]


def test_fp32_pyramid_stays_up_next_to_fp64():
    pw = pyramid_210().pack()
    rep = drift_report(pw, DT, checkpoints=(1, 10, 100, 300))
    _show(rep)
    for r in rep:
        assert r["finite"]
    last = rep[-1]
    # boxes are 30 units: the FP32 pyramid settles where the FP64 one does
    assert last["max_dp"] < 0.5, last
    assert abs(last["mean_y_fp32"] - last["mean_y_fp64"]) < 0.05, last
    assert abs(last["contacts_fp32"] - last["contacts_fp64"]) <= 4, last
    # at rest: the kinetic energy left is round-off, well under one box moving at 1 unit/s (0.5)
    assert last["kinetic_fp32"] < 0.5 and last["kinetic_fp64"] < 0.5, last


def test_fp32_chain_with_joints_hangs_like_fp64():
    pw = chain_world(n_links=80, n_circles=120, seed=42).pack()
    rep = drift_report(pw, DT, checkpoints=(1, 30, 150))
    _show(rep)
    assert all(r["finite"] for r in rep)
    assert rep[0]["max_dp"] < 1e-3, rep[0]
    # half a second in, links and circles (sizes 1..4) are still within a few percent of a link of the
    # FP64 run; later the circles bounce off the chain chaotically and the runs separate
    assert rep[1]["max_dp"] < 0.2, rep[1]
    assert abs(rep[1]["mean_y_fp32"] - rep[1]["mean_y_fp64"]) < 0.01, rep[1]


def test_fp32_batch_member_bit_identical_to_alone_and_repeatable():
    rng = Rng(177)
    a = random_pile(150, rng).pack()
    b = pyramid_210(rows=7).pack()
    g_batch = GPUWorld(concat([a, b]), FP32)
    g_alone = GPUWorld(b, FP32)
    g_again = GPUWorld(b, FP32)
    for g in (g_batch, g_alone, g_again):
        g.step(DT, 60)
    x, y, z = g_batch.bodies(), g_alone.bodies(), g_again.bodies()
    nb_a = a.n_bodies
    for f in ("p", "v", "a", "w"):
        assert np.array_equal(getattr(x, f)[nb_a:], getattr(y, f)), f
        assert np.array_equal(getattr(y, f), getattr(z, f)), f
    # FP32 state is float: every value read back is exactly representable in float
    assert np.array_equal(y.p.astype(np.float32).astype(np.float64), y.p)


def test_fp32_random_world_100k_first_steps():
    pw = random_world().pack()
    rep = drift_report(pw, DT, checkpoints=(1, 5))
    _show(rep)
    for r in rep:
        assert r["finite"]
        assert abs(r["contacts_fp32"] - r["contacts_fp64"]) <= 0.01 * r["contacts_fp64"], r
    assert rep[0]["max_dp_rel"] < 2e-5, rep[0]


def test_fp32_dense_large_world_on_the_grid():
    """A dense large space (its grid set in 64 strips, seams and a wide set) in
    FP32 beside FP64; both grid schedules give the same FP32 run."""
    from chipmunk2d_amd.gpu import GRID_FLOW, GRID_PHASES
    pw = random_world(n_circles=2600, n_polys=2600, n_segments=400, extent=140.0, seed=12).pack()
    rep = drift_report(pw, DT, checkpoints=(1, 20))
    _show(rep)
    assert all(r["finite"] for r in rep)
    assert rep[0]["max_dp_rel"] < 2e-5, rep[0]
    assert abs(rep[-1]["contacts_fp32"] - rep[-1]["contacts_fp64"]) <= 0.02 * rep[-1]["contacts_fp64"], rep[-1]
    flat = random_world(n_circles=2600, n_polys=2600, n_segments=400, extent=400.0, seed=15).pack()
    runs = []
    for sched in (GRID_FLOW, GRID_PHASES):
        g = GPUWorld(flat, FP32)
        g.set_grid_schedule(sched)
        g.step(DT, 130)
        runs.append(g.bodies())
        g.close()
    for f in ("p", "v", "a", "w"):
        assert np.array_equal(getattr(runs[0], f), getattr(runs[1], f)), f


def test_adjoint_refuses_fp32_worlds():
    g = GPUWorld(pyramid_210(rows=3).pack(), FP32)
    with pytest.raises(CpgpuError, match="FP64"):
        g.tape_begin(5)


def test_unknown_precision_is_rejected():
    with pytest.raises(CpgpuError):
        GPUWorld(pyramid_210(rows=3).pack(), 7)
    assert FP64 == 0 and FP32 == 1
