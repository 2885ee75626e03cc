"""Kernels run on the CPU: every csrc/*.hip translation unit compiled as host
C++ against the emulated HIP runtime in tests/cpu_emu (fibers per work-item,
wave ops and MFMA with the gfx950 lane layouts) under AddressSanitizer, then
driven through the C ABI by tests/cpu_emu/scenarios.py.  An out-of-bounds
access aborts with an ASan report; a barrier that not every work-item
reaches aborts as a deadlock; results are checked against the oracle.

The default CPU run executes the `quick`, `complex_ops` and `batched_tiled`
groups (a few minutes including the build).  DM_EMU_GROUPS="tiles tiled mixed batched
serving capped" selects the long groups (tens of minutes)."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EMU = os.path.join(HERE, "cpu_emu")


def _clang_available():
    return os.path.exists("/opt/rocm/llvm/bin/clang++")


@pytest.mark.skipif(not _clang_available(), reason="ROCm clang++ (host compiler) not found")
def test_kernels_under_cpu_emulation():
    sys.path.insert(0, EMU)
    try:
        import build_emu
    finally:
        sys.path.pop(0)
    lib = build_emu.build()
    env = dict(os.environ)
    env["DM_TEST_EMU_LIB"] = lib
    env["LD_PRELOAD"] = build_emu.asan_runtime()
    env["ASAN_OPTIONS"] = "detect_leaks=0"
    groups = os.environ.get("DM_EMU_GROUPS", "quick complex_ops batched_tiled").split()
    r = subprocess.run([sys.executable, os.path.join(EMU, "scenarios.py"), *groups], env=env,
                       capture_output=True, text=True, timeout=3000)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "ALL OK" in r.stdout, out[-4000:]
