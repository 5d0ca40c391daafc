#!/usr/bin/env python
"""Benchmark of the gfx950 cpSpaceStep on one MI355X.

    python bench.py [--steps K] [--warmup W] [--workload metric|random_world_100k|batch_4096|chains_2048|
                                                         grad_1024|grad_chains_2048]
                    [--precision fp64|fp32]

The metric has three parts, and the default workload ("metric") measures all
three in one run:
  * the single 100k-shape world -- 50k circles, 40k convex polygons, 10k
    static segments (container and sinusoidal terrain), seeded, 10 solver
    iterations, dt 1/60, a 300-step rollout.  `value` is its body-steps/s;
  * the 4096-space batch (log-uniform 16..2048 bodies, pyramids and piles,
    240-step rollout), beside the same spaces stepped one at a time by the
    plain single-space step (key "batch_4096");
  * forward+adjoint: 1024 spaces of 64 circles and boxes, a taped rollout of
    the 120-step workload and its backward pass (key "grad_1024").

A "step" is one cpSpaceStep of every space of the workload.  The cost of a
step grows as piles form (the 100k world has 14k contacts at step 10 and
140k at step 290), so the K timed steps are a fixed window of the rollout:
its last K steps, steps R-K .. R-1 of an R-step rollout (steps 0 .. K-1 when
K >= R).  The steps before the window run untimed on the same world (W of
them are the warm-up; more when the window starts later), so the timed steps
are exactly K full steps from the rollout's own state: nothing is skipped or
cached.  The library synchronises its stream at both ends of the window, and
inputs are resident in device memory throughout.

roofline: the dominant kernel's algorithmic bytes per launch (the streaming
model: every phase reads and writes the state the reference's dataflow
touches) over its average launch time from HIP events recorded on the
library's stream during the window, against 8 TB/s HBM; `traffic` is the
HBM bytes per launch that rocprofv3's counters read over the same window
(profiles/traffic_<workload>_w<K>.txt), or null when no counter run of this
window exists.  step_roofline: the whole step's algorithmic bytes (every
phase of the model) over the measured time per step.

Under rocprofv3 --selected-regions only the measured windows are collected:
the timed steps are bracketed by roctxProfilerResume(0) / roctxProfilerPause(0)
(no-ops without a profiler), so a kernel-trace or counter run of this command
covers exactly the steps the line reports.

cpu_baseline: the reference's own cpSpaceStep (compiled from its sources,
strict IEEE build), timed on the host in child processes pinned with
taskset, on a bounded sample that starts from the window's first state
(the device's bodies at that step, the reference's arbiter cache cold): the
median of three runs.  The batch is timed on all of the host's allowed cores
(each thread steps a contiguous slice of the spaces, joined every step).

--precision fp32 steps CPGPU_FP32 worlds (float state and arithmetic,
4-byte scalars in the byte model) and adds a "drift" object: the FP32 state
against the FP64 step (bit-for-bit the oracle) after 1, 10 and 60 steps.
"""
import argparse
import ctypes
import json
import os
import platform
import statistics
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "body-steps/s @10 iters: 100k-shape world & 4096-world batch; fwd+adjoint steps/s"
HBM_PEAK_GBS = 8000.0
S = 8  # bytes per scalar: 8 for FP64 worlds, 4 for FP32 (set from --precision)
DT = 1.0 / 60.0
ROLLOUT = {"random_world_100k": 300, "batch_4096": 240, "chains_2048": 300, "grad_1024": 120,
           "grad_chains_2048": 300}


# ---- profiler regions (rocprofv3 --selected-regions) ----

class _Regions:
    def __init__(self):
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so"):
            try:
                self.lib = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if self.lib is not None:
            self.lib.roctxProfilerResume.argtypes = [ctypes.c_uint64]
            self.lib.roctxProfilerPause.argtypes = [ctypes.c_uint64]

    def resume(self):
        if self.lib is not None:
            self.lib.roctxProfilerResume(0)

    def pause(self):
        if self.lib is not None:
            self.lib.roctxProfilerPause(0)


REGIONS = _Regions()


def window_of(workload, steps):
    """First step of the timed window: the rollout's last `steps` steps."""
    return max(0, ROLLOUT[workload] - steps)


# ---- workloads ----

def build_workload(name):
    from chipmunk2d_amd import scenes
    procs = min(16, os.cpu_count() or 1)
    if name == "random_world_100k":
        sp = scenes.random_world()
        return sp.pack(), {"workload": "random_world_100k", "circles": 50000, "polys": 40000,
                           "static_segments": 10000, "extent": 2000.0, "iterations": 10, "dt": DT,
                           "seed": hex(scenes.SEED_RANDOM_WORLD), "rollout_steps": 300}
    if name == "chains_2048":
        pw = scenes.chains_packed(n_worlds=2048, processes=procs)
        return pw, {"workload": "chains_2048", "worlds": pw.n_worlds, "links": 200, "circles": 300,
                    "constraints_per_world": pw.n_constraints // pw.n_worlds,
                    "joints": "pivot chain (collideBodies off), damped spring and slide joint to link i-2, "
                              "gear every 10 links", "iterations": 10, "dt": DT,
                    "seed": hex(scenes.SEED_CHAINS), "rollout_steps": 300}
    if name == "grad_1024":
        pw, targets = scenes.grad_batch(n_worlds=1024, n_bodies=64)
        pw.targets = targets
        return pw, {"workload": "grad_1024", "worlds": pw.n_worlds, "bodies_per_world": 64,
                    "loss": "sum |p_T - target|^2", "grads": "initial velocities, masses, friction",
                    "checkpoint_every": 10, "iterations": 10, "dt": DT, "seed": hex(scenes.SEED_GRAD),
                    "rollout_steps": 120}
    if name == "grad_chains_2048":
        pw = scenes.chains_packed(n_worlds=2048, processes=procs)
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
                    "iterations": 10, "dt": DT, "seed": hex(scenes.SEED_CHAINS), "rollout_steps": 300}
    if name == "batch_4096":
        pw = scenes.batch_packed(n_worlds=4096, processes=procs)
        return pw, {"workload": "batch_4096", "worlds": pw.n_worlds, "bodies_log_uniform": [16, 2048],
                    "kinds": "pyramid stack or random circle/poly pile (1:1), each on its own ground",
                    "iterations": 10, "dt": DT, "seed": hex(scenes.SEED_BATCH), "rollout_steps": 240}
    raise SystemExit("unknown workload " + name)


# ---- the streaming byte model ----

def solver_bytes(contacts, bodies, iterations):
    """Streaming-model bytes of solving `contacts` contacts and integrating `bodies` bodies:
    contact pre-step 26s + cached impulse 24s, then 47s per contact and iteration;
    velocity integration reads 8s and writes 6s per dynamic body."""
    return contacts * (26 + 24) * S + contacts * iterations * 47 * S + bodies * 14 * S


JOINT_ITER_SCALARS = {0: 31, 1: 24, 2: 27, 3: 13}   # pivot, damped spring, slide, gear


def narrowphase_bytes(pairs, arbiters, contacts):
    """Per pair: two shapes' world geometry (~8s each for the mix), material 4s + 8 B each, the pair
    record's id and warm-start 24 B; per colliding pair its arbiter header 6s + 8 B, per contact 4s + 8 B."""
    return pairs * (2 * 8 * S + 2 * (4 * S + 8) + 24) + arbiters * (6 * S + 8) + contacts * (4 * S + 8)


def shape_model(pw):
    """Per-step cacheData and broadphase bytes of the dynamic shapes (circle 15s, segment 23s,
    polygon (11 + 8k)s; broadphase 4s + 20 B each)."""
    a = pw.a
    dyn = a["body_type"][a["shape_body"]] != 2
    t = a["shape_type"][dyn]
    k = a["shape_poly_count"][dyn]
    cache = np.where(t == 0, 15, np.where(t == 1, 23, 11 + 8 * k)).sum() * S
    return int(cache), int(np.count_nonzero(dyn)) * (4 * S + 20)


def step_bytes(pw, n_dyn, per_step, iterations):
    """Algorithmic bytes of one whole step (every phase of the streaming model) at the window's
    average counts."""
    cache, broad = shape_model(pw)
    joints = 0
    if pw.n_constraints:
        ct = pw.a["cons_type"]
        per_iter = sum(JOINT_ITER_SCALARS[j] * int(np.count_nonzero(ct == j)) for j in JOINT_ITER_SCALARS) * S
        joints = per_iter * (iterations + 2)
    return int(n_dyn * 37 * S + cache + broad + per_step["pairs"] * 8
               + narrowphase_bytes(per_step["pairs"], per_step["arbiters"], per_step["contacts"])
               + per_step["contacts"] * ((26 + 24) + iterations * 47) * S + joints)


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
        return 2 * narrowphase_bytes(a, a, c)
    if kernel == "narrowphase":
        return narrowphase_bytes(p, a, c)
    if kernel == "bvh_pairs":
        return n_dyn_shapes * (4 * S + 20) + p * 8
    if kernel == "integrate_positions":
        return n_dyn_bodies * (11 + 12) * S
    if kernel == "cache_shapes":
        return n_dyn_shapes * 15 * S
    return None


def measured_traffic(workload, kernel, steps, precision="fp64"):
    """HBM bytes per launch of `kernel` read by rocprofv3's FETCH_SIZE / WRITE_SIZE counters over the same
    window (profiles/traffic_<workload>_w<steps>.txt, _fp32 for FP32 worlds; calibrated on a device copy), or
    None when no counter run of this window exists."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_w%d%s.txt" % (workload, steps,
                                                                    "" if precision == "fp64" else "_fp32"))
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("bench_kernel=%s " % kernel):
            return int(line.split("hbm_bytes_per_launch=")[1].split()[0])
    return None


def progress(msg):
    """A progress line on stderr (the JSON line alone goes to stdout)."""
    print("[bench %7.1f s] %s" % (time.time() - _T0, msg), file=sys.stderr, flush=True)


_T0 = time.time()


# ---- device runs ----

def _advance(g, dt, n, chunk=50):
    """n untimed steps (synchronised every `chunk` steps so queues stay short)."""
    done = 0
    while done < n:
        k = min(chunk, n - done)
        g.step(dt, k)
        g.sync()
        done += k
        if n > chunk:
            progress("  %d of %d untimed steps" % (done, n))


def run_gpu(pw, steps, warmup, start, dt, autotune_steps=0, autotune_at=60, precision=0, bucket_threads=None):
    """Warm-up (and, for batches, the autotuner, once the copy has run `autotune_at` steps so that piles have
    formed) on a throwaway copy of the world; then a fresh copy runs `start` steps untimed and the `steps`
    steps of the window timed, with the tuned tiling.  bucket_threads (four workgroup sizes) fixes the
    buckets' sizes in place of the autotuner.  Returns (seconds, per-step times, solver timing, every
    kernel's timing from an identical second run, stats over the window, solve counts over the window,
    tiling, the world's bodies at the window's start)."""
    from chipmunk2d_amd.gpu import GPUWorld

    if bucket_threads:
        autotune_steps = 0
    tiling = None
    progress("warm-up (%d steps)%s" % (warmup, ", autotuner" if autotune_steps else ""))
    if warmup > 0 or autotune_steps > 0:
        g = GPUWorld(pw, precision)
        if warmup > 0:
            g.step(dt, warmup)
        if autotune_steps > 0:
            _advance(g, dt, max(autotune_at - warmup, 0))
            g.autotune(dt, autotune_steps)
        g.sync()
        tiling = g.buckets()
        g.close()

    def fresh():
        g = GPUWorld(pw, precision)
        if autotune_steps > 0:
            g.set_buckets([b[0] for b in tiling], [b[1] for b in tiling])
        if bucket_threads:
            g.set_buckets([b[0] for b in g.buckets()], bucket_threads)
        _advance(g, dt, start)
        return g

    def totals(g):
        st = g.stats()
        return st, g.solve_counts()

    progress("the measured run: %d untimed steps, then %d timed" % (start, steps))
    g = fresh()
    tiling = g.buckets()
    state0 = g.bodies() if start > 0 else None
    st0, sc0 = totals(g)
    # inside the measured steps only the solver's launches carry HIP events
    g.set_timing(True, solve_only=True)
    t_step = []
    g.sync()
    REGIONS.resume()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.step(dt)
        t_step.append(time.perf_counter())
    g.sync()
    t1 = time.perf_counter()
    REGIONS.pause()
    timing = g.timing()
    st1, sc1 = totals(g)
    g.close()
    stats = dict(st1)
    for k in ("total_arbiters", "total_contacts", "total_pairs"):
        stats[k] = st1[k] - st0[k]
    counts = {k: (sc1[k][0] - sc0[k][0], sc1[k][1] - sc0[k][1]) for k in sc1}
    progress("timed window: %.3f s; the per-kernel table run" % (t1 - t0))
    # the per-kernel table: the same window again, every launch timed
    g = fresh()
    g.set_timing(True)
    g.step(dt, steps)
    g.sync()
    profile = g.timing()
    g.close()
    return t1 - t0, [t - t0 for t in t_step], timing, profile, stats, counts, tiling, state0


def run_grad(pw, steps, warmup, start, dt, checkpoint_every=10):
    """A taped rollout of the window's `steps` steps (after `start` untimed steps) and its backward pass:
    (forward seconds, backward seconds, solver/adjoint timing, every kernel's timing, stats, solve counts,
    loss)."""
    from chipmunk2d_amd.gpu import GPUWorld
    from chipmunk2d_amd import grad as G
    dyn = getattr(pw, "loss_mask", (pw.a["body_type"] == G.DYNAMIC).astype(float))
    if warmup > 0:
        g = GPUWorld(pw)
        g.tape_begin(checkpoint_every)
        g.step(dt, warmup)
        g.backward(g_p=np.ones((pw.n_bodies, 2)))
        g.close()

    def rollout(solve_only, measured):
        g = GPUWorld(pw)
        _advance(g, dt, start)
        st0 = g.stats()
        sc0 = g.solve_counts()
        g.set_timing(True, solve_only=solve_only)
        g.tape_begin(checkpoint_every)
        g.sync()
        if measured:
            REGIONS.resume()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.step(dt)
        g.sync()
        t1 = time.perf_counter()
        if measured:
            REGIONS.pause()
        L, gp = G.position_loss(g.bodies().p, pw.targets, dyn)
        if measured:
            REGIONS.resume()
        t2 = time.perf_counter()
        g.backward(g_p=gp)
        t3 = time.perf_counter()
        if measured:
            REGIONS.pause()
        st1 = g.stats()
        sc1 = g.solve_counts()
        stats = dict(st1)
        for k in ("total_arbiters", "total_contacts", "total_pairs"):
            stats[k] = st1[k] - st0[k]
        counts = {k: (sc1[k][0] - sc0[k][0], sc1[k][1] - sc0[k][1]) for k in sc1}
        out = (t1 - t0, t3 - t2, g.timing(), stats, counts, L)
        g.close()
        return out
    # measured with HIP events on the solver's and the adjoint's launches only; then the same run again,
    # every launch timed, for the per-kernel table
    t_fwd, t_bwd, timing, stats, counts, L = rollout(True, True)
    profile = rollout(False, False)[2]
    return t_fwd, t_bwd, timing, profile, stats, counts, L


def run_plain_per_space(pw, n_sample, steps, start, dt):
    """The batch's spaces stepped one at a time by the plain single-space GPU step: every
    (n_worlds / n_sample)-th space of the seeded batch rebuilt alone, `start` steps untimed, then the
    window's `steps` steps each.  Returns (body-steps/s, bodies in the sample, seconds)."""
    from chipmunk2d_amd import scenes
    from chipmunk2d_amd.gpu import GPUWorld
    progress("the batch's spaces one at a time (plain per-space step)")
    specs = scenes.batch_specs(n_worlds=pw.n_worlds)
    stride = max(1, pw.n_worlds // n_sample)
    spaces = [scenes.build_batch_world(specs[k]).pack() for k in range(0, pw.n_worlds, stride)][:n_sample]
    worlds = [GPUWorld(sp) for sp in spaces]
    for g in worlds:
        _advance(g, dt, max(start, 1))
    t0 = time.perf_counter()
    for g in worlds:
        g.step(dt, steps)
        g.sync()
    t = time.perf_counter() - t0
    bodies = sum(sp.dynamic_body_count() for sp in spaces)
    for g in worlds:
        g.close()
    return bodies * steps / t, bodies, t


# ---- the CPU baseline: the reference, in pinned child processes ----

def _with_state(pw, state, start, dt):
    """The description with the device's bodies at the window's first step (forces are zero after a
    step; the spaces' stamp and previous dt as the device had them)."""
    from chipmunk2d_amd.cabi import PackedWorld
    a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in pw.a.items()}
    if state is not None:
        for f, key in (("p", "body_p"), ("v", "body_v"), ("a", "body_a"), ("w", "body_w"),
                       ("v_bias", "body_v_bias"), ("w_bias", "body_w_bias"), ("transform", "body_transform")):
            a[key] = np.ascontiguousarray(getattr(state, f), dtype=np.float64)
        a["space"] = [dict(sp, stamp=int(sp["stamp"]) + start, curr_dt=dt) for sp in a["space"]]
    return PackedWorld(a)


def _save_packed(pw, path):
    arrays = {k: v for k, v in pw.a.items() if isinstance(v, np.ndarray)}
    np.savez(path, __space__=np.frombuffer(json.dumps(pw.a["space"]).encode(), np.uint8), **arrays)


def _load_packed(path):
    from chipmunk2d_amd.cabi import PackedWorld
    z = np.load(path, allow_pickle=False)
    a = {k: z[k] for k in z.files if k != "__space__"}
    a["space"] = json.loads(bytes(z["__space__"]).decode())
    return PackedWorld(a)


def host_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        glibc = os.confstr("CS_GNU_LIBC_VERSION")
    except (ValueError, OSError):
        glibc = " ".join(platform.libc_ver())
    return {"cpu_model": model or platform.processor() or platform.machine(), "glibc": glibc,
            "allowed_cpus": len(os.sched_getaffinity(0)), "nproc": os.cpu_count(),
            "compiler": "gcc -std=gnu99 -O2 -DNDEBUG -ffp-contract=off -fno-fast-math (strict); "
                        "-O3 -DNDEBUG -ffast-math (Release)"}


def cpu_child(spec_path):
    """Runs in a child process pinned by taskset: times the reference on the packed description in the
    spec, three runs from the same state, the first calibrating the step count; prints one JSON line."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from checkers import RefWorld
    spec = json.load(open(spec_path))
    pw = _load_packed(spec["npz"])
    dt, target, runs = spec["dt"], spec["target_seconds"], spec["runs"]
    steps = spec.get("steps")
    out = []
    for _ in range(runs):
        if spec["threads"] > 1 and spec["batch"]:
            from concurrent.futures import ThreadPoolExecutor
            threads = spec["threads"]
            worlds = [RefWorld(pw, w) for w in range(pw.n_worlds)]
            slices = [worlds[q * len(worlds) // threads:(q + 1) * len(worlds) // threads] for q in range(threads)]

            def step_slice(sl):
                for r in sl:
                    r.step(dt)

            with ThreadPoolExecutor(threads) as ex:
                t0 = time.perf_counter()
                k = 0
                while (steps is None and (k == 0 or time.perf_counter() - t0 < target)) or \
                        (steps is not None and k < steps):
                    list(ex.map(step_slice, slices))
                    k += 1
                t = time.perf_counter() - t0
            del worlds
        else:
            r = RefWorld(pw, 0, hasty_threads=spec["hasty"], release=spec["release"])
            t0 = time.perf_counter()
            k = 0
            while (steps is None and (k == 0 or time.perf_counter() - t0 < target)) or \
                    (steps is not None and k < steps):
                r.step(dt)
                k += 1
            t = time.perf_counter() - t0
            del r
        steps = k
        out.append({"steps": k, "seconds": t})
    print(json.dumps(out))


def run_cpu_pinned(pw, start, n_dyn, cores, threads=1, batch=False, hasty=0, release=False, target=4.0, runs=3,
                   tmpdir=None):
    """Rates (body-steps/s) of `runs` runs of the reference on pw in a child pinned to `cores`; the first
    run steps until `target` seconds have passed, the others as many steps."""
    if not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libcpref.so")):
        return None
    npz = os.path.join(tmpdir, "state.npz")
    if not os.path.exists(npz):
        _save_packed(pw, npz)
    spec = os.path.join(tmpdir, "spec.json")
    json.dump({"npz": npz, "dt": DT, "target_seconds": target, "runs": runs, "threads": threads, "batch": batch,
               "hasty": hasty, "release": release}, open(spec, "w"))
    cmd = [sys.executable, os.path.join(ROOT, "bench.py"), "--cpu-child", spec]
    progress("CPU baseline on cores %s (%s%s%s)" % (",".join(str(c) for c in cores), "batch " if batch else "",
                                                     "hasty " if hasty else "", "release" if release else "strict"))
    pinned = "taskset"
    try:
        subprocess.run(["taskset", "-V"], capture_output=True, check=True)
        cmd = ["taskset", "-c", ",".join(str(c) for c in cores)] + cmd
    except (OSError, subprocess.CalledProcessError):
        pinned = "none (taskset missing)"
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=ROOT)
    if r.returncode != 0:
        raise RuntimeError("CPU baseline child failed:\n" + r.stderr[-2000:])
    res = json.loads(r.stdout.strip().splitlines()[-1])
    rates = [n_dyn * x["steps"] / x["seconds"] for x in res]
    return {"median": statistics.median(rates), "runs": [round(x, 1) for x in rates],
            "steps_per_run": res[0]["steps"], "seconds_per_run": [round(x["seconds"], 3) for x in res],
            "pinned": pinned, "cores": list(cores)}


def cpu_baselines(workload, pw, state, start, n_dyn, gpu_rate, k_steps):
    allowed = sorted(os.sched_getaffinity(0))
    with tempfile.TemporaryDirectory() as tmp:
        pws = _with_state(pw, state, start, DT)
        if workload == "random_world_100k":
            core = [allowed[-1]]
            main = run_cpu_pinned(pws, start, n_dyn, core, tmpdir=tmp)
            if main is None:
                return None
            rel = run_cpu_pinned(pws, start, n_dyn, core, release=True, tmpdir=tmp)
            hasty_cores = allowed[-2:] if len(allowed) >= 2 else allowed
            hasty = run_cpu_pinned(pws, start, n_dyn, hasty_cores, hasty=len(hasty_cores), tmpdir=tmp)
            out = {"value": round(main["median"], 1), "unit": "body-steps/s", "cores": 1, "kind": "reference",
                   "sample": "the reference's cpSpaceStep (strict build, one thread pinned with taskset) from the "
                             "window's first state, step %d of the 300-step rollout (the device's bodies there; the "
                             "reference's arbiter cache starts cold): %d steps a run, median of 3 runs"
                             % (start, main["steps_per_run"]),
                   "runs": main["runs"], "seconds_per_run": main["seconds_per_run"],
                   "gpu_window_rate": round(gpu_rate, 1), "gpu_over_cpu": round(gpu_rate / main["median"], 1),
                   "also": {"reference_release_1_core": {"value": round(rel["median"], 1), "runs": rel["runs"]},
                            "hasty_space": {"value": round(hasty["median"], 1), "runs": hasty["runs"],
                                            "threads": len(hasty_cores),
                                            "note": "cpHastySpace clamps its solver to 2 threads "
                                                    "(cpHastySpace.c:387,534)"}}}
        else:
            threads = min(16, len(allowed))
            cores = allowed[:threads]
            main = run_cpu_pinned(pws, start, n_dyn, cores, threads=threads, batch=True, target=3.0, tmpdir=tmp)
            if main is None:
                return None
            out = {"value": round(main["median"], 1), "unit": "body-steps/s", "cores": threads, "kind": "reference",
                   "sample": "every space, the reference's cpSpaceStep (strict build) with the spaces split over %d "
                             "threads pinned with taskset, joined every step, from the window's first state (step "
                             "%d; arbiter caches cold): %d steps a run, median of 3 runs"
                             % (threads, start, main["steps_per_run"]),
                   "runs": main["runs"], "seconds_per_run": main["seconds_per_run"],
                   "gpu_window_rate": round(gpu_rate, 1), "gpu_over_cpu": round(gpu_rate / main["median"], 1)}
    out["host"] = host_info()
    out["precision"] = "f64 (the reference's default CP_USE_DOUBLES build)"
    return out


# ---- one workload ----

def measure(workload, args, precision, bucket_threads, cpu=True, plain=True, pw_config=None):
    progress("building %s" % workload)
    pw, config = pw_config or build_workload(workload)
    steps = args.steps if args.steps is not None else config["rollout_steps"]
    start = window_of(workload, steps)
    n_dyn = pw.dynamic_body_count()
    body_type = pw.a["body_type"]
    n_dyn_shapes = int(np.count_nonzero(body_type[pw.a["shape_body"]] != 2))
    iterations = max(sp["iterations"] for sp in pw.a["space"])
    config.update({"bodies": n_dyn, "shapes": pw.n_shapes, "worlds": pw.n_worlds,
                   "window": {"first_step": start, "steps": steps,
                              "of": "the last %d steps of the %d-step rollout" % (steps, config["rollout_steps"])
                              if start > 0 else "steps 0-%d" % (steps - 1)}})
    autotune = args.autotune_steps if workload in ("batch_4096", "chains_2048") and not bucket_threads else 0
    state0 = None
    out = {}
    if workload in ("grad_1024", "grad_chains_2048"):
        t_fwd, t_bwd, timing, profile, stats, solve_counts, loss = run_grad(pw, steps, args.warmup, start, DT)
        wall = t_fwd + t_bwd
        tiling = None
        out.update({"forward_s": round(t_fwd, 4), "backward_s": round(t_bwd, 4), "loss": loss,
                    "fwd_adjoint_steps_per_s": round(steps / wall, 2),
                    "forward_only_body_steps_per_s": round(n_dyn * steps / t_fwd, 1)})
    else:
        wall, t_steps, timing, profile, stats, solve_counts, tiling, state0 = run_gpu(
            pw, steps, args.warmup, start, DT, autotune, precision=precision, bucket_threads=bucket_threads)
    if tiling:
        config["tiling"] = {"autotuned": bool(autotune), "buckets": [
            {"bucket": i, "max_bodies": b[0], "threads": b[1], "worlds": b[2], "lds_bodies": b[3]}
            for i, b in enumerate(tiling) if b[2] > 0]}
        if autotune:
            config["tiling"].update({"trial_steps": autotune, "tuned_after_steps": 60})
    value = n_dyn * steps / wall
    per_step = {"contacts": stats["total_contacts"] / steps, "arbiters": stats["total_arbiters"] / steps,
                "pairs": stats["total_pairs"] / steps}

    def kernel_table(tm):
        tab = {}
        for name, (ms, launches) in tm.items():
            avg_ms = ms / max(launches, 1)
            b = algorithmic_bytes(name, per_step, n_dyn, n_dyn_shapes, iterations, launches, solve_counts)
            tab[name] = {"total_ms": round(ms, 3), "launches": launches, "avg_us": round(avg_ms * 1e3, 2)}
            if b is not None:
                tab[name]["bytes_per_launch"] = int(b)
                tab[name]["GBps"] = round(b / (avg_ms * 1e-3) / 1e9, 1)
        return tab
    measured = kernel_table(timing)
    kernels = kernel_table(profile)
    dominant = max(timing.items(), key=lambda kv: kv[1][0])[0] if timing else None
    roofline = None
    if dominant and "GBps" in measured.get(dominant, {}):
        ach = measured[dominant]["GBps"]
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 5),
                    "traffic": measured_traffic(workload, dominant, steps, args.precision),
                    "bytes_per_launch": measured[dominant]["bytes_per_launch"],
                    "avg_launch_us": measured[dominant]["avg_us"],
                    "window": "the same %d steps as value (HIP events on the library's stream)" % steps}
    step_roofline = None
    if workload not in ("grad_1024", "grad_chains_2048"):
        sb = step_bytes(pw, n_dyn, per_step, iterations)
        ach = sb / (wall / steps) / 1e9
        step_roofline = {"bytes_per_step": sb, "ms_per_step": round(wall / steps * 1e3, 4),
                         "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(ach / HBM_PEAK_GBS, 5),
                         "model": "bodies 37s, shape cache and broadphase per dynamic shape, 8 B per pair, the "
                                  "narrowphase per pair and arbiter, per contact 50s + 47s per iteration, joints "
                                  "(s = %d B), at the window's average counts" % S}
    out.update({
        "value": value, "steps": steps, "wall": wall, "config": config,
        "ms_per_step": round(wall / steps * 1e3, 4),
        "contacts_per_s": round(stats["total_contacts"] / wall, 1),
        "avg_contacts_per_step": round(per_step["contacts"], 1),
        "avg_arbiters_per_step": round(per_step["arbiters"], 1),
        "avg_pairs_per_step": round(per_step["pairs"], 1),
        "roofline": roofline, "step_roofline": step_roofline,
        "kernels_measured_run": measured, "kernels": kernels,
    })
    if cpu and not args.no_cpu_baseline and workload in ("random_world_100k", "batch_4096", "chains_2048"):
        out["cpu_baseline"] = cpu_baselines(workload, pw, state0, start, n_dyn, value, steps)
    if plain and workload == "batch_4096" and not args.no_plain_baseline:
        rate, bodies, secs = run_plain_per_space(pw, 64, steps, start, DT)
        out["plain_per_space_gpu"] = {
            "value": round(rate, 1), "unit": "body-steps/s",
            "sample": "every 64th of the 4096 spaces (%d dynamic bodies), each alone, the same window (%d steps "
                      "after %d untimed), one space after another" % (bodies, steps, max(start, 1)),
            "seconds": round(secs, 3), "batched_over_plain": round(value / rate, 1)}
    return pw, out


def part_summary(name, r):
    """The headline numbers of one part of the metric, as an extra key of the line."""
    keep = ("value", "ms_per_step", "contacts_per_s", "avg_contacts_per_step", "roofline", "step_roofline",
            "cpu_baseline", "plain_per_space_gpu", "forward_s", "backward_s", "fwd_adjoint_steps_per_s",
            "forward_only_body_steps_per_s", "loss")
    out = {k: r[k] for k in keep if k in r}
    out["value"] = round(out["value"], 1)
    out["unit"] = "body-steps/s" if not name.startswith("grad") else "body-steps/s (forward+adjoint)"
    out["config"] = r["config"]
    out["dominant_kernels"] = {k: v for k, v in sorted(r["kernels_measured_run"].items(),
                                                        key=lambda kv: -kv[1]["total_ms"])[:3]}
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps: the last K steps of each workload's rollout (default: the whole rollout)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="metric",
                    choices=["metric", "random_world_100k", "batch_4096", "chains_2048", "grad_1024",
                             "grad_chains_2048"])
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
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args.cpu_child)
        return
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
    if args.precision != "fp64" and args.workload in ("grad_1024", "grad_chains_2048", "metric"):
        print("the adjoint runs on FP64 worlds only: --precision fp32 takes a forward workload", file=sys.stderr)
        sys.exit(2)
    if args.steps is not None and args.steps < 1:
        print("--steps must be positive", file=sys.stderr)
        sys.exit(2)

    primary = "random_world_100k" if args.workload == "metric" else args.workload
    pw, r = measure(primary, args, precision, bucket_threads)
    parts = {}
    if args.workload == "metric":
        for name in ("batch_4096", "grad_1024"):
            print("[bench] %s: %.3g body-steps/s; measuring %s" % (primary, r["value"], name), file=sys.stderr,
                  flush=True)
            parts[name] = part_summary(name, measure(name, args, precision, bucket_threads)[1])
    out = {
        "metric": METRIC,
        "value": round(r["value"], 1),
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic (seeded scene generator, xoshiro256**)",
        "config": r["config"],
        "contacts_per_s": r["contacts_per_s"],
        "avg_contacts_per_step": r["avg_contacts_per_step"],
        "avg_arbiters_per_step": r["avg_arbiters_per_step"],
        "avg_pairs_per_step": r["avg_pairs_per_step"],
        "roofline": r["roofline"],
        "step_roofline": r["step_roofline"],
        "cpu_baseline": r.get("cpu_baseline"),
        "plain_per_space_gpu": r.get("plain_per_space_gpu"),
    }
    for k in ("forward_s", "backward_s", "loss", "fwd_adjoint_steps_per_s", "forward_only_body_steps_per_s"):
        if k in r:
            out[k] = r[k]
    if primary.startswith("grad"):
        out["metric_part"] = "fwd+adjoint"
    out.update(parts)
    out["kernels_measured_run"] = r["kernels_measured_run"]
    out["kernels"] = r["kernels"]
    out["kernels_note"] = ("kernels_measured_run: HIP events on the solver's (and adjoint's) launches inside the "
                           "measured steps; kernels: every launch timed, in a second identical run of the window")
    if args.precision == "fp32":
        from chipmunk2d_amd.drift import drift_report
        out["drift"] = {"against": "FP64 step of the same workload (bit-for-bit the oracle)",
                        "at_steps": drift_report(pw, DT, checkpoints=(1, 10, 60))}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
