"""Builds tests/sanitize/_build/host_check: the library's sources with their
host code compiled under AddressSanitizer and UndefinedBehaviorSanitizer
(hipcc's -Xarch_host, so device code is built as usual and not instrumented),
linked with the driver host_check.cpp.

Usage: python tests/sanitize/build.py [-j N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "acoustic-field-calculator_amd", "csrc")
OBJ = os.path.join(HERE, "_build")
EXE = os.path.join(OBJ, "host_check")
ARCH = os.environ.get("AFH_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CLANG = "/opt/rocm/llvm/bin/clang++"

SAN = ["-fsanitize=address", "-fsanitize=undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
HOST_ONLY = [x for f in SAN for x in ("-Xarch_host", f)]
HIP_FLAGS = [
    # -O3 as the library's own build, so the device code is the product's
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Xarch_host", "-g",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", *HOST_ONLY,
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"failed: {cmd[0]} ... {cmd[-1]}")


def _compile_hip(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", obj])
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile_hip, srcs))
    # the driver is plain C++ against the C ABI; it and the link carry the
    # sanitizer runtimes (host code only: nothing here targets the GPU)
    drv = os.path.join(OBJ, "host_check.o")
    _run([CLANG, "-O1", "-g", "-std=c++17", *SAN, "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
          "-I" + os.path.join(ROOT, "include"), "-c", os.path.join(HERE, "host_check.cpp"), "-o", drv])
    tmp = EXE + ".tmp"
    _run([CLANG, *SAN, "-o", tmp, drv, *objs, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib",
          "-lpthread"])
    os.replace(tmp, EXE)
    return EXE


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    print(build(ap.parse_args().j))
