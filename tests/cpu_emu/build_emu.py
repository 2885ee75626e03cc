"""Build the CPU-emulated library (test infrastructure): every
dagger.jl_amd/csrc/*.hip translation unit compiled as host C++ against the
emulated HIP header in this folder, with AddressSanitizer, into
tests/cpu_emu/_build/libdagger_mi355x_emu.so.

Usage: python tests/cpu_emu/build_emu.py  (prints the library path and the
ASan runtime to preload)."""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "dagger.jl_amd", "csrc")
OUT = os.path.join(HERE, "_build")
CXX = "/opt/rocm/llvm/bin/clang++"
FLAGS = ["-std=c++17", "-g", "-O1", "-fPIC", "-fno-omit-frame-pointer", "-fsanitize=address",
         "-fsanitize-address-use-after-scope", "-I", HERE, "-I", os.path.join(ROOT, "include"),
         "-I", CSRC, "-Wno-unused-command-line-argument"]
LIB = os.path.join(OUT, "libdagger_mi355x_emu.so")


def asan_runtime():
    out = subprocess.run([CXX, "-print-file-name=libclang_rt.asan-x86_64.so"],
                         capture_output=True, text=True, check=True).stdout.strip()
    return out


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + ".o")
    if os.path.exists(obj) and os.path.getmtime(obj) >= max(
            os.path.getmtime(src), *(os.path.getmtime(os.path.join(d, f))
                                     for d in (CSRC, os.path.join(HERE, "hip"))
                                     for f in os.listdir(d) if f.endswith(".h"))):
        return obj
    cmd = [CXX, *FLAGS, "-x", "c++", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"compile failed: {src}\n{r.stderr}")
    return obj


def build():
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    srcs.append(os.path.join(HERE, "emu.cpp"))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    cmd = [CXX, "-shared", "-fsanitize=address", "-shared-libasan", "-o", LIB, *objs, "-lpthread"]
    subprocess.run(cmd, check=True)
    return LIB


if __name__ == "__main__":
    lib = build()
    print(lib)
    print(asan_runtime())
    sys.exit(0)
