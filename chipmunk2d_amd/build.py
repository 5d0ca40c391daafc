"""Build the gfx950 step library chipmunk2d_amd/libcpgpu.so with hipcc.

Every translation unit is compiled with -ffp-contract=off (one IEEE
operation per + - * /, as in the reference's strict build) and with packed
FP32 instructions disabled.  The step sources are compiled twice: once with
double arithmetic into namespace cpgpu64 (CPGPU_FP64 worlds, bit-for-bit the
reference's double build) and once with float into cpgpu32 (CPGPU_FP32
worlds); dispatch.hip holds the extern "C" entry points, which forward each
call to the build the world was created with.  Run as
`python -m chipmunk2d_amd.build` or via __graft_entry__.build().
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcpgpu.so")
SOURCES = ["world.hip", "broadphase.hip", "narrowphase.hip", "solver.hip", "adjoint.hip", "query.hip"]
VARIANTS = {"64": ["-DCPGPU_NS=cpgpu64", "-DCPGPU_REAL=double"],
            "32": ["-DCPGPU_NS=cpgpu32", "-DCPGPU_REAL=float"]}
UNITS = [(src, v) for v in VARIANTS for src in SOURCES] + [("dispatch.hip", None)]
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


def _obj(unit):
    src, v = unit
    base = os.path.splitext(src)[0]
    return os.path.join(CSRC, "_build", base + ("_fp" + v if v else "") + ".o")


def _compile(unit):
    src, v = unit
    cmd = ["hipcc"] + HIPCC_FLAGS + (VARIANTS[v] if v else []) + ["-c", os.path.join(CSRC, src), "-o", _obj(unit)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(line for line in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in line)
    return "%s%s" % (src, " (fp%s)" % v if v else ""), r.returncode, msgs


def build(verbose=True):
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(UNITS), os.cpu_count() or 4)) as ex:
        results = list(ex.map(_compile, UNITS))
    failed = False
    for src, rc, msgs in results:
        if msgs.strip() and verbose:
            print("[%s]\n%s" % (src, msgs), file=sys.stderr)
        if rc != 0:
            failed = True
    if failed:
        raise RuntimeError("hipcc failed; see messages above")
    tmp = OUT + ".tmp"
    cmd = ["hipcc", "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [_obj(u) for u in UNITS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
