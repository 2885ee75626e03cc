"""The native planner (dag.hip, tiled.hip, trsm_plan.hip) compiled for the host with HIP and
the device launches stubbed out, run under AddressSanitizer and UBSan over
every executor option set (sequential, batched multi-stream look-ahead,
graph capture, profiling).  CPU only: catches host memory errors in
planning/issue code without a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
H = os.path.join(ROOT, "tests", "host_harness")
CSRC = os.path.join(ROOT, "dagger.jl_amd", "csrc")


@pytest.mark.skipif(not shutil.which("g++"), reason="g++ not available")
def test_planner_host_under_sanitizers(tmp_path):
    exe = tmp_path / "plan_harness"
    cmd = ["g++", "-std=c++17", "-g", "-O1", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-x", "c++",
           "-I", H, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
           os.path.join(CSRC, "dag.hip"), os.path.join(CSRC, "tiled.hip"),
           os.path.join(CSRC, "trsm_plan.hip"),
           os.path.join(H, "stubs.cpp"), os.path.join(H, "driver.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    p = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    assert p.stdout.strip().endswith("ok")
    # look-ahead plans batch each k-step into a handful of launches
    lines = [l for l in p.stdout.splitlines() if l.startswith("opts 21100 U")]
    assert lines and "tasks 56" in lines[0]
