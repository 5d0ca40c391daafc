"""Build tests/emu/_build/libcpgpu_emu.so: the step library's own device and
host sources (chipmunk2d_amd/csrc) compiled with g++ for the CPU against the
HIP stand-in of tests/emu/hip/hip_runtime.h.  TEST INFRASTRUCTURE: the library
it makes exports the same C ABI as chipmunk2d_amd/libcpgpu.so and lets the
GPU tests run in a container without a GPU (tests/conftest.py, CPGPU_EMULATE);
it is never loaded by the package itself.

The sources are copied with these mechanical rewrites, each counted so that a
change of the sources that it no longer fits stops the build:
  * `kernel<<<grid, block, shm, stream>>>(args)` -> `emu::launch(kernel, grid,
    block, shm, stream)(args)`; the one cooperative launch likewise;
  * `__shared__` -> `static thread_local` (one copy per host thread, and a
    workgroup runs on one host thread); `extern __shared__ T x[];` -> a pointer
    to the launch's dynamic shared memory;
  * where the device code takes the lanes that reach a point of divergent code
    with `__ballot(1)` (append_index, emit_pair), each lane alone; the
    leader's broadcast among exactly those lanes (`emu_shfl_in`); and at the
    end of k_pairs, where every lane of the wave meets again before reading the
    wave's staged count, an explicit meeting of the wave (`emu_wave_sync`).
Run as `python tests/emu/build_emu.py`.
"""
import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "chipmunk2d_amd", "csrc")
# EMU_ASAN=1: an AddressSanitizer build into _build_asan, with allocations of
# their exact size, run with libasan preloaded (python itself is not built with it):
#   ASAN_OPTIONS=detect_leaks=0 LD_PRELOAD=$(gcc -print-file-name=libasan.so) python ...
ASAN = os.environ.get("EMU_ASAN") == "1"
SAN = ["-fsanitize=address", "-fno-omit-frame-pointer", "-DEMU_EXACT_ALLOC"] if ASAN else []
BUILD = os.path.join(HERE, "_build_asan" if ASAN else "_build")
OUT = os.path.join(BUILD, "libcpgpu_emu.so")
SOURCES = ["world.hip", "broadphase.hip", "narrowphase.hip", "solver.hip", "adjoint.hip", "query.hip"]
VARIANTS = {"64": ("cpgpu64", "double"), "32": ("cpgpu32", "float")}


def _body_span(text, signature):
    """(start, end) of the braces of the first function whose text starts with signature."""
    i = text.index(signature)
    j = text.index("{", i)
    depth = 0
    for k in range(j, len(text)):
        if text[k] == "{":
            depth += 1
        elif text[k] == "}":
            depth -= 1
            if depth == 0:
                return j, k + 1
    raise ValueError("unbalanced braces after " + signature)


def _in_body(text, signature, old, new, count):
    a, b = _body_span(text, signature)
    body = text[a:b]
    if body.count(old) != count:
        raise RuntimeError("%s: expected %d of %r, found %d" % (signature, count, old, body.count(old)))
    return text[:a] + body.replace(old, new) + text[b:]


LAUNCH = re.compile(r"([A-Za-z_][\w:]*(?:<[^<>;()]*>)?)\s*<<<")
COOP = re.compile(r"void \*args\[\] = \{([^}]*)\};(.*?)hipLaunchCooperativeKernel\(reinterpret_cast<const void \*>"
                  r"\((\w+)\),\s*([^,]+),\s*([^,]+),\s*args,\s*([^,]+),\s*(\w+)\)", re.S)
EXTERN_SHARED = re.compile(r"extern\s+__shared__\s+(?:__align__\(\d+\)\s+)?([^;\[]+?)\s+(\w+)\[\];")


def rewrite(name, text):
    n_launch = len(LAUNCH.findall(text))
    text = LAUNCH.sub(lambda m: "emu::launch(%s, " % m.group(1), text)
    if text.count(">>>") != n_launch:
        raise RuntimeError("%s: %d launches but %d '>>>'" % (name, n_launch, text.count(">>>")))
    text = re.sub(r">>>\s*\(", ")(", text)

    def coop(m):
        args = [a.strip() for a in m.group(1).split(",") if a.strip()]
        names = []
        for a in args:
            mm = re.fullmatch(r"\(void \*\)&(\w+)", a)
            if not mm:
                raise RuntimeError("%s: cooperative argument %r" % (name, a))
            names.append(mm.group(1))
        return ("void *args[] = {%s};%semu::launch_cooperative(%s, %s, %s, %s, %s)(%s)"
                % (m.group(1), m.group(2), m.group(3), m.group(4), m.group(5), m.group(6), m.group(7),
                   ", ".join(names)))
    text, n_coop = COOP.subn(coop, text)
    if "hipLaunchCooperativeKernel" in text:
        raise RuntimeError("%s: a cooperative launch the rewrite does not fit" % name)
    text = EXTERN_SHARED.sub(lambda m: "%s *%s = reinterpret_cast<%s *>(emu::dyn_shared());"
                             % (m.group(1), m.group(2), m.group(1)), text)
    if "extern __shared__" in text:
        raise RuntimeError("%s: an extern __shared__ declaration the rewrite does not fit" % name)
    text = re.sub(r"\b__shared__\b", "static thread_local", text)
    if name == "narrowphase.hip":
        text = _in_body(text, "__device__ __forceinline__ int append_index(", "__ballot(1)", "emu_self_mask()", 1)
        text = _in_body(text, "__device__ __forceinline__ int append_index(", "base = __shfl(base, leader);",
                        "base = emu_shfl_in(base, leader, active);", 1)
    if name == "broadphase.hip":
        text = _in_body(text, "__device__ __forceinline__ void emit_pair(", "__ballot(1)", "emu_self_mask()", 1)
        text = _in_body(text, "__device__ __forceinline__ void stage_flush(", "base = __shfl(base, leader);",
                        "base = emu_shfl_in(base, leader, active);", 1)
        text = _in_body(text, "__global__ void __launch_bounds__(256) k_pairs(", "const int staged = *st.n;",
                        "emu_wave_sync(); const int staged = *st.n;", 1)
    return text, n_launch, n_coop


def grid_static_lds(ns, real):
    """Static LDS of k_solve_grid in the gfx950 build (hipcc's resource usage)."""
    cmd = ["hipcc", "--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-Xclang", "-target-feature",
           "-Xclang", "-packed-fp32-ops", "-DCPGPU_NS=" + ns, "-DCPGPU_REAL=" + real,
           "-Rpass-analysis=kernel-resource-usage", "--cuda-device-only", "-c", os.path.join(CSRC, "solver.hip"),
           "-o", os.devnull]
    r = subprocess.run(cmd, capture_output=True, text=True)
    lines = r.stderr.splitlines()
    for i, line in enumerate(lines):
        if "Function Name:" in line and "k_solve_grid" in line:
            for follow in lines[i:i + 12]:
                m = re.search(r"LDS Size \[bytes/block\]: (\d+)", follow)
                if m:
                    return int(m.group(1))
    raise RuntimeError("no LDS figure for k_solve_grid from hipcc:\n" + r.stderr[-2000:])


def build(verbose=True):
    src = os.path.join(BUILD, "src", "csrc")
    inc = os.path.join(BUILD, "include")
    os.makedirs(src, exist_ok=True)
    os.makedirs(os.path.join(BUILD, "obj"), exist_ok=True)
    shutil.copytree(os.path.join(ROOT, "include"), inc, dirs_exist_ok=True)
    counts = {}
    for f in sorted(os.listdir(CSRC)):
        if not (f.endswith(".hip") or f.endswith(".h")):
            continue
        text, n_launch, n_coop = rewrite(f, open(os.path.join(CSRC, f)).read())
        counts[f] = (n_launch, n_coop)
        with open(os.path.join(src, f), "w") as fh:
            fh.write(text)
    if counts.get("solver.hip", (0, 0))[1] != 1:
        raise RuntimeError("solver.hip: expected one cooperative launch")
    lds = {v: grid_static_lds(*VARIANTS[v]) for v in VARIANTS}
    units = [(s, v) for v in VARIANTS for s in SOURCES] + [("dispatch.hip", None)]

    def compile_unit(unit):
        s, v = unit
        obj = os.path.join(BUILD, "obj", os.path.splitext(s)[0] + ("_fp" + v if v else "") + ".o")
        cmd = ["g++", "-x", "c++", "-std=c++17", "-O2", "-g", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
               "-Wno-attributes", "-Wno-unused-result", "-I", HERE] + SAN
        if v:
            ns, real = VARIANTS[v]
            cmd += ["-DCPGPU_NS=" + ns, "-DCPGPU_REAL=" + real, "-DEMU_GRID_STATIC_LDS=%d" % lds[v]]
        cmd += ["-c", os.path.join(src, s), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        return s, v, obj, r.returncode, r.stdout + r.stderr

    with ThreadPoolExecutor(max_workers=min(len(units), os.cpu_count() or 4)) as ex:
        results = list(ex.map(compile_unit, units))
    failed = [x for x in results if x[3] != 0]
    for s, v, obj, rc, msg in results:
        if rc != 0 and verbose:
            print("[%s fp%s]\n%s" % (s, v, msg[-6000:]), file=sys.stderr)
    if failed:
        raise RuntimeError("g++ failed on %s" % ", ".join("%s (fp%s)" % (s, v) for s, v, *_ in failed))
    rt = os.path.join(BUILD, "obj", "emu_runtime.o")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-g", "-fPIC", "-I", HERE] + SAN +
                       ["-c", os.path.join(HERE, "emu_runtime.cpp"), "-o", rt], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("emu_runtime.cpp:\n" + r.stderr)
    tmp = OUT + ".tmp"
    r = subprocess.run(["g++", "-shared"] + SAN[:1] + ["-o", tmp] + [x[2] for x in results] + [rt, "-lpthread"],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, OUT)
    if not ASAN:
        build_binding()
    return OUT


REFERENCE = "/root/reference"
BINDING_OUT = os.path.join(BUILD, "libcpgpuspace_emu.so")


def build_binding():
    """The cpGPUSpace binding (chipmunk2d_amd/binding/cpGPUSpace.c) over the
    emulated library, as oracle/Makefile links it over libcpgpu.so; only where
    the reference's headers and oracle/_ref/libcpref.so are at hand."""
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    if not (os.path.isdir(os.path.join(REFERENCE, "include")) and os.path.exists(os.path.join(ref_dir, "libcpref.so"))):
        return None
    cmd = ["gcc", "-std=gnu99", "-O2", "-DNDEBUG", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-Wall", "-shared",
           "-I", os.path.join(REFERENCE, "include"), "-I", os.path.join(ROOT, "include"), "-o", BINDING_OUT,
           os.path.join(ROOT, "chipmunk2d_amd", "binding", "cpGPUSpace.c"),
           "-L", ref_dir, "-l:libcpref.so", "-L", BUILD, "-l:libcpgpu_emu.so",
           "-Wl,-rpath," + ref_dir, "-Wl,-rpath," + BUILD, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("binding over the emulation:\n" + r.stderr)
    return BINDING_OUT


if __name__ == "__main__":
    print(build())
