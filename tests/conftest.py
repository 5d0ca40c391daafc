import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, TESTS):
    if p not in sys.path:
        sys.path.insert(0, p)

# CPGPU_EMULATE=1 runs the GPU tests in a container without a GPU: the step
# library's own sources built for the CPU by tests/emu/build_emu.py (device
# threads as fibers, device memory in host memory).  Never set where a GPU runs them.
if os.environ.get("CPGPU_EMULATE") == "1":
    import chipmunk2d_amd.gpu as _gpu
    import checkers as _checkers
    _gpu.LIB_PATH = os.path.join(TESTS, "emu", "_build", "libcpgpu_emu.so")
    _checkers.GPUSPACE_SO = os.path.join(TESTS, "emu", "_build", "libcpgpuspace_emu.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: runs the HIP step library on an MI355X (gfx950)")
