"""Build the gfx950 step library chipmunk2d_amd/libcpgpu.so with hipcc.

Every translation unit is compiled with -ffp-contract=off (one IEEE
operation per + - * /, as in the reference's strict build) and with packed
FP32 instructions disabled.  Run as `python -m chipmunk2d_amd.build` or via
__graft_entry__.build().
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcpgpu.so")
SOURCES = ["world.hip", "broadphase.hip", "narrowphase.hip", "solver.hip", "adjoint.hip"]
ARCH = os.environ.get("CPGPU_ARCH", "gfx950")

HIPCC_FLAGS = [
    "--offload-arch=" + ARCH,
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _obj(src):
    return os.path.join(CSRC, "_build", os.path.splitext(src)[0] + ".o")


def _compile(src):
    obj = _obj(src)
    cmd = ["hipcc"] + HIPCC_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(line for line in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in line)
    return src, r.returncode, msgs


def build(verbose=True):
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        results = list(ex.map(_compile, SOURCES))
    failed = False
    for src, rc, msgs in results:
        if msgs.strip() and verbose:
            print("[%s]\n%s" % (src, msgs), file=sys.stderr)
        if rc != 0:
            failed = True
    if failed:
        raise RuntimeError("hipcc failed; see messages above")
    tmp = OUT + ".tmp"
    cmd = ["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [_obj(s) for s in SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
