"""Builds the package's native libraries in-tree (xkaapi_amd/lib/)."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("XKRT_HIP_ARCH", "gfx950")

# hipcc flags for every HIP source: gfx950 only, no packed-FP32 instruction forms
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-shared", "-pthread",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]

HIP_SOURCES = ["resolver.hip", "coherency.hip", "release.hip", "handles.hip", "taskdom.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    return out


def build_host():
    os.makedirs(LIB, exist_ok=True)
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-Wextra", "-o",
          os.path.join(LIB, "libxkstreams.so"), os.path.join(CSRC, "streams.c"), "-lm"])


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    _run([HIPCC, *HIP_FLAGS, "-o", os.path.join(LIB, "libxkrt_hip.so"), *srcs])


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-B", "all"])


ABI_DIR = os.path.join(ROOT, "tests", "abi")


def build_abi_driver():
    """The compiled C++ caller of the C ABI (tests/abi/xkrt_abi_driver.cpp), linked
    against lib/libxkrt_hip.so (found at run time through its rpath)."""
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "lib")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-I", os.path.join(ROOT, "include"),
          "-o", os.path.join(ABI_DIR, "xkrt_abi_driver"), os.path.join(ABI_DIR, "xkrt_abi_driver.cpp"),
          "-L", LIB, "-lxkrt_hip", "-Wl,-rpath,$ORIGIN/../../xkaapi_amd/lib", "-Wl,-rpath-link," + rocm_lib])


def build_all():
    build_host()
    build_hip()
    build_abi_driver()
    build_oracle()


if __name__ == "__main__":
    build_all()
