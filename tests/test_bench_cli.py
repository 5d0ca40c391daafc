"""bench.py's command line refuses what it cannot run before it builds a
workload -- CPU only (no GPU is touched on these paths)."""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("sizes", ["64,64", "64,64,256,512,1024", "64,64,100,512", "a,b,c,d"])
def test_bucket_threads_need_four_valid_sizes(sizes):
    r = _bench("--bucket-threads", sizes)
    assert r.returncode == 2
    assert "four workgroup sizes" in r.stderr


def test_fp32_needs_a_forward_workload():
    r = _bench("--precision", "fp32")
    assert r.returncode == 2
    assert "FP64" in r.stderr


def test_window_is_the_rollouts_last_steps():
    sys.path.insert(0, ROOT)
    import bench
    assert bench.window_of("random_world_100k", 20) == 280
    assert bench.window_of("batch_4096", 20) == 220
    assert bench.window_of("grad_1024", 20) == 100
    assert bench.window_of("random_world_100k", 300) == 0
    assert bench.window_of("random_world_100k", 500) == 0


def test_cpu_baseline_child_times_the_reference_pinned(tmp_path):
    """The CPU baseline's child process: three runs of the reference from the same state, the first
    calibrating the step count, pinned with taskset when the host has it."""
    sys.path.insert(0, ROOT)
    import bench
    from checkers import ref_available
    from chipmunk2d_amd.scenes import pyramid_210
    if not ref_available():
        pytest.skip("the compiled reference (oracle/_ref) is not built")
    pw = pyramid_210(rows=8).pack()
    n_dyn = pw.dynamic_body_count()
    core = [sorted(os.sched_getaffinity(0))[-1]]
    r = bench.run_cpu_pinned(pw, 0, n_dyn, core, target=0.2, tmpdir=str(tmp_path))
    assert len(r["runs"]) == 3 and r["steps_per_run"] >= 1
    assert all(x > 0 for x in r["runs"]) and abs(r["median"] - sorted(r["runs"])[1]) < 0.1
    # the packed description survives the trip through the file the child reads
    back = bench._load_packed(os.path.join(str(tmp_path), "state.npz"))
    assert back.a["space"] == pw.a["space"]
    for k, v in pw.a.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(back.a[k], v), k
    info = bench.host_info()
    assert info["glibc"] and info["cpu_model"]
