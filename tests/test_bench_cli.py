"""bench.py's command line refuses what it cannot run before it builds a
workload -- CPU only (no GPU is touched on these paths)."""
import os
import subprocess
import sys

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
