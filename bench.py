#!/usr/bin/env python
"""Benchmark of the gfx950 cpSpaceStep on one MI355X.

Workload (default): the single 100k-shape world -- 50k circles, 40k convex
polygons, 10k static segments (container + sinusoidal terrain), seeded,
10 solver iterations, dt 1/60 -- stepped for the configured 300-step rollout
from its initial state.  A "step" is one cpSpaceStep of the whole world.

    python bench.py [--steps K] [--warmup W] [--workload random_world_100k|batch_4096|chains_2048|grad_1024]
                    [--precision fp64|fp32]

W warm-up steps run on a separate copy of the world (first-launch costs),
then exactly K steps are timed on a fresh copy, so the timed region is
steps 0..K-1 of the rollout.  The step library synchronises its own stream
at both ends of the timed region.  Prints one JSON line.

Other workloads: batch_4096 (4096 spaces, log-uniform 16..2048 bodies, 240
steps), chains_2048 (2048 spaces of a 200-link jointed chain and 300
circles, 300 steps), grad_1024 (1024 spaces of 64 circles / boxes: a taped
120-step rollout and its adjoint; a step is one forward step and its share
of the backward pass, so value is forward+adjoint body-steps/s),
grad_chains_2048 (the chains_2048 spaces: a taped 300-step rollout and its
adjoint for a loss on the final link positions, gradients w.r.t. every
spring's stiffness and damping).

roofline: the dominant kernel's algorithmic bytes per launch (streaming
model: every phase reads and writes the state the reference's dataflow
touches, FP64) over its average launch time from HIP events recorded on
the library's stream during the timed region, against 8 TB/s HBM.
cpu_baseline: the reference's own cpSpaceStep (compiled from its sources,
strict IEEE build, one thread) on the first steps of the same rollout.
plain_per_space_gpu (batch_4096): a sample of the batch's spaces stepped one
at a time by the plain single-space GPU step, beside the batched rate.

--precision fp32 steps CPGPU_FP32 worlds (float state and arithmetic,
4-byte scalars in the byte model) and adds a "drift" object: the FP32
state against the FP64 step (bit-for-bit the oracle) after 1, 10 and 60
steps of the same workload, measured outside the timed region.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "body-steps/s @10 iters: 100k-shape world & 4096-world batch; fwd+adjoint steps/s"
HBM_PEAK_GBS = 8000.0
S = 8  # bytes per scalar: 8 for FP64 worlds, 4 for FP32 (set from --precision)


def build_workload(name):
    from chipmunk2d_amd import scenes
    from chipmunk2d_amd.cabi import concat
    if name == "random_world_100k":
        sp = scenes.random_world()
        return sp.pack(), {"workload": "random_world_100k", "circles": 50000, "polys": 40000,
                           "static_segments": 10000, "extent": 2000.0, "iterations": 10, "dt": 1.0 / 60.0,
                           "seed": hex(scenes.SEED_RANDOM_WORLD), "rollout_steps": 300}
    if name == "chains_2048":
        pw = scenes.chains_packed(n_worlds=2048, processes=min(16, os.cpu_count() or 1))
        return pw, {"workload": "chains_2048", "worlds": pw.n_worlds, "links": 200, "circles": 300,
                    "constraints_per_world": pw.n_constraints // pw.n_worlds,
                    "joints": "pivot chain (collideBodies off), damped spring and slide joint to link i-2, "
                              "gear every 10 links", "iterations": 10, "dt": 1.0 / 60.0,
                    "seed": hex(scenes.SEED_CHAINS), "rollout_steps": 300}
    if name == "grad_1024":
        pw, targets = scenes.grad_batch(n_worlds=1024, n_bodies=64)
        pw.targets = targets
        return pw, {"workload": "grad_1024", "worlds": pw.n_worlds, "bodies_per_world": 64,
                    "loss": "sum |p_T - target|^2", "grads": "initial velocities, masses, friction",
                    "checkpoint_every": 10, "iterations": 10, "dt": 1.0 / 60.0, "seed": hex(scenes.SEED_GRAD),
                    "rollout_steps": 120}
    if name == "grad_chains_2048":
        pw = scenes.chains_packed(n_worlds=2048, processes=min(16, os.cpu_count() or 1))
        links = np.zeros(pw.n_bodies, bool)
        links[pw.a["cons_a"]] = True
        links[pw.a["cons_b"]] = True
        links &= pw.a["body_type"] == 0
        # system identification: final link positions against those of the chains at rest, the
        # gradients taken w.r.t. every spring's stiffness and damping
        pw.targets = pw.a["body_p"].copy()
        pw.loss_mask = links.astype(float)
        return pw, {"workload": "grad_chains_2048", "worlds": pw.n_worlds, "links": 200, "circles": 300,
                    "loss": "sum over links |p_T - p_0|^2", "grads": "spring stiffness and damping (and every "
                    "joint parameter, initial velocities, masses, friction)", "checkpoint_every": 10,
                    "iterations": 10, "dt": 1.0 / 60.0, "seed": hex(scenes.SEED_CHAINS), "rollout_steps": 300}
    if name == "batch_4096":
        pw = scenes.batch_packed(n_worlds=4096, processes=min(16, os.cpu_count() or 1))
        return pw, {"workload": "batch_4096", "worlds": pw.n_worlds, "bodies_log_uniform": [16, 2048],
                    "kinds": "pyramid stack or random circle/poly pile (1:1), each on its own ground",
                    "iterations": 10, "dt": 1.0 / 60.0, "seed": hex(scenes.SEED_BATCH), "rollout_steps": 240}
    raise SystemExit("unknown workload " + name)


def solver_bytes(contacts, bodies, iterations):
    """Streaming-model bytes of solving `contacts` contacts and integrating `bodies` bodies:
    contact pre-step 26s + cached impulse 24s, then 47s per contact and iteration;
    velocity integration reads 8s and writes 6s per dynamic body."""
    return contacts * (26 + 24) * S + contacts * iterations * 47 * S + bodies * 14 * S


def algorithmic_bytes(kernel, per_step, n_dyn_bodies, n_dyn_shapes, iterations, launches, solve_counts):
    """Streaming-model bytes of one launch of `kernel`, from the per-step averages (solver kernels: from the
    contacts and bodies each solve unit handled, counted on the device)."""
    c, a, p = per_step["contacts"], per_step["arbiters"], per_step["pairs"]
    if kernel in solve_counts:
        uc, ub = solve_counts[kernel]
        return solver_bytes(uc / max(launches, 1), ub / max(launches, 1), iterations)
    if kernel == "reverse_solve":
        # reads the forward solve's state and writes as many adjoints (twice the forward's bytes), plus
        # the recorded pre-state of every arbiter application (18 scalars)
        return 2 * solver_bytes(c, n_dyn_bodies, iterations) + a * iterations * 18 * S
    if kernel == "reverse_geometry":
        # per arbiter: both shapes' geometry and poses, the arbiter and its adjoints (read), the bodies'
        # position adjoints (written): the narrowphase's bytes twice over
        return 2 * (a * (2 * 8 * S + 2 * (4 * S + 8) + 24) + a * (6 * S + 8) + c * (4 * S + 8))
    if kernel == "narrowphase":
        # per pair: two shapes' world geometry (~8s each for the mix), material 4s + 8 B each,
        # warm-start record 24 B; per colliding pair its arbiter header 6s + 8 B, per contact 4s + 8 B
        return p * (2 * 8 * S + 2 * (4 * S + 8) + 24) + a * (6 * S + 8) + c * (4 * S + 8)
    if kernel == "bvh_pairs":
        return n_dyn_shapes * (4 * S + 20) + p * 8
    if kernel == "integrate_positions":
        return n_dyn_bodies * (11 + 12) * S
    if kernel == "cache_shapes":
        return n_dyn_shapes * 15 * S
    return None


def measured_traffic(workload, kernel, precision="fp64"):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    (profiles/traffic_<workload>.txt, or traffic_<workload>_fp32.txt for FP32 worlds; calibrated on a device
    copy), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_%s%s.txt" % (workload, "" if precision == "fp64" else "_fp32"))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bench_kernel=%s " % kernel):
            return int(line.split("hbm_bytes_per_launch=")[1].split()[0])
    return None


def run_gpu(pw, steps, warmup, dt, autotune_steps=0, autotune_at=60, precision=0, bucket_threads=None):
    """Warm-up (and, for batches, the autotuner, once the copy has run `autotune_at` steps so that piles have
    formed) on a throwaway copy of the world, then `steps` timed steps on a fresh copy with the tuned tiling.
    bucket_threads (four workgroup sizes) fixes the buckets' sizes in place of the autotuner."""
    from chipmunk2d_amd.gpu import GPUWorld

    def fix_buckets(g):
        if bucket_threads:
            g.set_buckets([b[0] for b in g.buckets()], bucket_threads)

    if bucket_threads:
        autotune_steps = 0
    tiling = None
    if warmup > 0 or autotune_steps > 0:
        g = GPUWorld(pw, precision)
        if warmup > 0:
            g.step(dt, warmup)
        if autotune_steps > 0:
            g.step(dt, max(autotune_at - warmup, 0))
            g.autotune(dt, autotune_steps)
        g.sync()
        tiling = g.buckets()
        g.close()
    g = GPUWorld(pw, precision)
    if autotune_steps > 0:
        g.set_buckets([b[0] for b in tiling], [b[1] for b in tiling])
    fix_buckets(g)
    tiling = g.buckets()
    g.sync()
    # inside the measured run only the solver's launches carry HIP events
    g.set_timing(True, solve_only=True)
    t_step = []
    g.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.step(dt)
        t_step.append(time.perf_counter())
    g.sync()
    t1 = time.perf_counter()
    timing = g.timing()
    stats = g.stats()
    counts = g.solve_counts()
    g.close()
    # the per-kernel table: the same run again, every launch timed
    g = GPUWorld(pw, precision)
    if autotune_steps > 0:
        g.set_buckets([b[0] for b in tiling], [b[1] for b in tiling])
    fix_buckets(g)
    g.set_timing(True)
    g.step(dt, steps)
    g.sync()
    profile = g.timing()
    g.close()
    return t1 - t0, [t - t0 for t in t_step], timing, profile, stats, counts, tiling


def run_grad(pw, steps, warmup, dt, checkpoint_every=10):
    """A taped rollout of `steps` steps and its backward pass: (forward seconds, backward seconds, timing,
    stats, loss)."""
    from chipmunk2d_amd.gpu import GPUWorld
    from chipmunk2d_amd import grad as G
    dyn = getattr(pw, "loss_mask", (pw.a["body_type"] == G.DYNAMIC).astype(float))
    if warmup > 0:
        g = GPUWorld(pw)
        g.tape_begin(checkpoint_every)
        g.step(dt, warmup)
        g.backward(g_p=np.ones((pw.n_bodies, 2)))
        g.close()
    def rollout(solve_only):
        g = GPUWorld(pw)
        g.sync()
        g.set_timing(True, solve_only=solve_only)
        g.tape_begin(checkpoint_every)
        t0 = time.perf_counter()
        for _ in range(steps):
            g.step(dt)
        g.sync()
        t1 = time.perf_counter()
        L, gp = G.position_loss(g.bodies().p, pw.targets, dyn)
        t2 = time.perf_counter()
        g.backward(g_p=gp)
        t3 = time.perf_counter()
        out = (t1 - t0, t3 - t2, g.timing(), g.stats(), g.solve_counts(), L)
        g.close()
        return out
    # measured with HIP events on the solver's and the adjoint's launches only; then the same run again,
    # every launch timed, for the per-kernel table
    t_fwd, t_bwd, timing, stats, counts, L = rollout(True)
    profile = rollout(False)[2]
    return t_fwd, t_bwd, timing, profile, stats, counts, L


def run_cpu_reference(pw, sample_steps, dt, hasty_threads=0, release=False):
    """Seconds for the reference's own cpSpaceStep (or cpHastySpaceStep with hasty_threads) to step world 0
    of pw sample_steps times."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from checkers import RefWorld, REF_SO
    if not os.path.exists(REF_SO):
        return None
    r = RefWorld(pw, hasty_threads=hasty_threads, release=release)
    t0 = time.perf_counter()
    r.step(dt, sample_steps)
    return time.perf_counter() - t0


def run_plain_per_space(pw, n_sample, steps, dt):
    """The batch's spaces stepped one at a time by the plain single-space GPU step: every
    (n_worlds / n_sample)-th space of the seeded batch rebuilt alone, `steps` steps each after one
    untimed step.  Returns (body-steps/s, bodies in the sample, seconds)."""
    from chipmunk2d_amd import scenes
    from chipmunk2d_amd.gpu import GPUWorld
    specs = scenes.batch_specs(n_worlds=pw.n_worlds)
    stride = max(1, pw.n_worlds // n_sample)
    spaces = [scenes.build_batch_world(specs[k]).pack() for k in range(0, pw.n_worlds, stride)][:n_sample]
    worlds = [GPUWorld(sp) for sp in spaces]
    for g in worlds:
        g.step(dt)
        g.sync()
    t0 = time.perf_counter()
    for g in worlds:
        g.step(dt, steps)
        g.sync()
    t = time.perf_counter() - t0
    bodies = sum(sp.dynamic_body_count() for sp in spaces)
    for g in worlds:
        g.close()
    return bodies * steps / t, bodies, t


def run_cpu_reference_batch(pw, sample_steps, dt, threads):
    """Seconds for the reference's cpSpaceStep to step every world of pw sample_steps times, the worlds
    split into contiguous slices over `threads` host threads (each slice stepped one step at a time,
    all slices joined after every step)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from checkers import RefWorld, REF_SO
    if not os.path.exists(REF_SO):
        return None
    from concurrent.futures import ThreadPoolExecutor
    worlds = [RefWorld(pw, w) for w in range(pw.n_worlds)]
    slices = [worlds[k::threads] for k in range(threads)]

    def step_slice(sl):
        for r in sl:
            r.step(dt)

    with ThreadPoolExecutor(threads) as ex:
        t0 = time.perf_counter()
        for _ in range(sample_steps):
            list(ex.map(step_slice, slices))
        return time.perf_counter() - t0


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default: the workload's rollout length)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="random_world_100k",
                    choices=["random_world_100k", "batch_4096", "chains_2048", "grad_1024", "grad_chains_2048"])
    ap.add_argument("--cpu-sample-steps", type=int, default=120,
                    help="steps of the same rollout the reference CPU step is timed on")
    ap.add_argument("--cpu-batch-sample-steps", type=int, default=20,
                    help="steps of a batch rollout the reference CPU step is timed on (all worlds)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain-baseline", action="store_true",
                    help="skip the batch's comparison with its spaces stepped one at a time")
    ap.add_argument("--bucket-threads", default=None,
                    help="four comma-separated workgroup sizes for the small-space buckets, in place of the "
                         "autotuner (to profile the sizes a measured run chose)")
    ap.add_argument("--autotune-steps", type=int, default=6,
                    help="steps per trial of the batch tiling autotuner (on the warm-up copy; 0: default tiling)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32"],
                    help="world precision (CPGPU_FP64: bit-for-bit the reference's double build; CPGPU_FP32)")
    args = ap.parse_args()
    global S
    S = 8 if args.precision == "fp64" else 4
    from chipmunk2d_amd.cabi import FP32, FP64
    precision = FP64 if args.precision == "fp64" else FP32
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would step its own "
              "slice of worlds independently); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    bucket_threads = None
    if args.bucket_threads:
        try:
            bucket_threads = [int(x) for x in args.bucket_threads.split(",")]
        except ValueError:
            bucket_threads = []
        if len(bucket_threads) != 4 or any(t not in (64, 128, 256, 512, 1024) for t in bucket_threads):
            print("--bucket-threads takes four workgroup sizes, each 64, 128, 256, 512 or 1024", file=sys.stderr)
            sys.exit(2)

    dt = 1.0 / 60.0
    pw, config = build_workload(args.workload)
    if args.steps is None:
        args.steps = config["rollout_steps"]
    n_dyn = pw.dynamic_body_count()
    body_type = pw.a["body_type"]
    n_dyn_shapes = int(np.count_nonzero(body_type[pw.a["shape_body"]] != 2))
    iterations = max(sp["iterations"] for sp in pw.a["space"])
    config.update({"bodies": n_dyn, "shapes": pw.n_shapes, "worlds": pw.n_worlds})

    autotune = args.autotune_steps if args.workload in ("batch_4096", "chains_2048") else 0
    grad_extra = None
    if args.workload in ("grad_1024", "grad_chains_2048"):
        if args.precision != "fp64":
            print("the adjoint runs on FP64 worlds only", file=sys.stderr)
            sys.exit(2)
        t_fwd, t_bwd, timing, profile, stats, solve_counts, loss = run_grad(pw, args.steps, args.warmup, dt)
        wall = t_fwd + t_bwd
        t_steps = None
        tiling = None
        grad_extra = {"forward_s": round(t_fwd, 4), "backward_s": round(t_bwd, 4), "loss": loss,
                      "fwd_adjoint_steps_per_s": round(args.steps / wall, 2),
                      "forward_only_body_steps_per_s": round(n_dyn * args.steps / t_fwd, 1)}
    else:
        if bucket_threads:
            autotune = 0
        wall, t_steps, timing, profile, stats, solve_counts, tiling = run_gpu(pw, args.steps, args.warmup, dt,
                                                                               autotune, precision=precision,
                                                                               bucket_threads=bucket_threads)
    if tiling:
        config["tiling"] = {"autotuned": bool(autotune), "buckets": [
            {"bucket": i, "max_bodies": b[0], "threads": b[1], "worlds": b[2], "lds_bodies": b[3]}
            for i, b in enumerate(tiling) if b[2] > 0]}
        if autotune:
            config["tiling"].update({"trial_steps": autotune, "tuned_after_steps": 60})
    value = n_dyn * args.steps / wall
    per_step = {"contacts": stats["total_contacts"] / args.steps, "arbiters": stats["total_arbiters"] / args.steps,
                "pairs": stats["total_pairs"] / args.steps}

    def kernel_table(tm):
        out = {}
        for name, (ms, launches) in tm.items():
            avg_ms = ms / max(launches, 1)
            b = algorithmic_bytes(name, per_step, n_dyn, n_dyn_shapes, iterations, launches, solve_counts)
            out[name] = {"total_ms": round(ms, 3), "launches": launches, "avg_us": round(avg_ms * 1e3, 2)}
            if b is not None:
                out[name]["bytes_per_launch"] = int(b)
                out[name]["GBps"] = round(b / (avg_ms * 1e-3) / 1e9, 1)
        return out
    # the measured run's events (solver and adjoint launches), and every kernel of an identical second run
    measured = kernel_table(timing)
    kernels = kernel_table(profile)
    dominant = max(timing.items(), key=lambda kv: kv[1][0])[0] if timing else None
    roofline = None
    if dominant and "GBps" in measured.get(dominant, {}):
        ach = measured[dominant]["GBps"]
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 5), "traffic": measured_traffic(args.workload, dominant, args.precision),
                    "bytes_per_launch": measured[dominant]["bytes_per_launch"],
                    "avg_launch_us": measured[dominant]["avg_us"]}

    cpu = None
    if not args.no_cpu_baseline and args.workload == "random_world_100k":
        k = min(args.cpu_sample_steps, args.steps)
        t_cpu = run_cpu_reference(pw, k, dt)
        if t_cpu is not None:
            gpu_same = n_dyn * k / t_steps[k - 1]
            cpu = {"value": round(n_dyn * k / t_cpu, 1), "unit": "body-steps/s", "cores": 1, "kind": "reference",
                   "sample": "steps 0-%d of the same 300-step rollout (the reference's cpSpaceStep, strict build "
                             "-O2 -ffp-contract=off, one thread; later steps cost more as piles form)" % (k - 1),
                   "seconds": round(t_cpu, 2), "gpu_same_sample": round(gpu_same, 1),
                   "gpu_over_cpu_same_sample": round(gpu_same / (n_dyn * k / t_cpu), 1),
                   "host": platform.processor() or platform.machine(), "nproc": os.cpu_count()}
            # the reference's own Release flags (-O3 -ffast-math) and cpHastySpace, same sample
            t_rel = run_cpu_reference(pw, k, dt, release=True)
            t_hasty = run_cpu_reference(pw, k, dt, hasty_threads=os.cpu_count() or 1)
            cpu["also"] = {
                "reference_release_1_core": round(n_dyn * k / t_rel, 1),
                "hasty_space": {"value": round(n_dyn * k / t_hasty, 1), "threads_requested": os.cpu_count(),
                                "note": "cpHastySpace clamps its solver to 2 threads (cpHastySpace.c:387,534)"},
            }
    plain = None
    if args.workload == "batch_4096" and not args.no_plain_baseline:
        # the same spaces stepped one at a time by the plain per-space GPU step (each its own world)
        rate, bodies, secs = run_plain_per_space(pw, 64, 20, dt)
        plain = {"value": round(rate, 1), "unit": "body-steps/s",
                 "sample": "every 64th of the 4096 spaces (%d dynamic bodies), each alone, 20 steps after one "
                           "untimed step, stepped one space after another" % bodies,
                 "seconds": round(secs, 3), "batched_over_plain": round(value / rate, 1)}
    if not args.no_cpu_baseline and args.workload in ("batch_4096", "chains_2048"):
        k = min(args.cpu_batch_sample_steps, args.steps)
        threads = min(16, os.cpu_count() or 1)
        t_cpu = run_cpu_reference_batch(pw, k, dt, threads)
        if t_cpu is not None:
            gpu_same = n_dyn * k / t_steps[k - 1]
            cpu = {"value": round(n_dyn * k / t_cpu, 1), "unit": "body-steps/s", "cores": threads,
                   "kind": "reference",
                   "sample": "steps 0-%d of the same rollout, every world, the reference's cpSpaceStep (strict "
                             "build) with the worlds split over %d host threads, joined every step" % (k - 1, threads),
                   "seconds": round(t_cpu, 2), "gpu_same_sample": round(gpu_same, 1),
                   "gpu_over_cpu_same_sample": round(gpu_same / (n_dyn * k / t_cpu), 1),
                   "host": platform.processor() or platform.machine(), "nproc": os.cpu_count()}

    if cpu is not None:
        cpu["precision"] = "f64 (the reference's default CP_USE_DOUBLES build)"
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded scene generator, xoshiro256**)",
        "config": config,
        "contacts_per_s": round(stats["total_contacts"] / wall, 1),
        "avg_contacts_per_step": round(per_step["contacts"], 1),
        "avg_arbiters_per_step": round(per_step["arbiters"], 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_per_space_gpu": plain,
        "kernels_measured_run": measured,
        "kernels": kernels,
        "kernels_note": "kernels_measured_run: HIP events on the solver's (and adjoint's) launches inside the "
                        "measured steps; kernels: every launch timed, in a second identical run outside them",
    }
    if grad_extra:
        out["metric_part"] = "fwd+adjoint"
        out.update(grad_extra)
    if args.precision == "fp32":
        from chipmunk2d_amd.drift import drift_report
        out["drift"] = {"against": "FP64 step of the same workload (bit-for-bit the oracle)",
                        "at_steps": drift_report(pw, dt, checkpoints=(1, 10, 60))}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
