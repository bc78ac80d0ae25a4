"""Builds libspecfempp_amd.so in-tree with hipcc for gfx950.

The host-side assembly is compiled with -ffp-contract=off so its FP32
arithmetic keeps the reference host build's rounding; the device code keeps
hipcc's default contraction.  Packed-FP32 instructions are disabled for all
device code.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(HERE, "libspecfempp_amd.so")
BUILD_DIR = os.path.join(HERE, "_build")

ARCH = os.environ.get("SMX_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

DEVICE_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-munsafe-fp-atomics", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall",
                "-Wno-unused-result"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", f"-I{INCLUDE}",
              f"-I{CSRC}", "-Wall"]

SOURCES = [
    ("solver2d.hip", "device"),
    ("acoustic2d.hip", "device"),
    ("coupled2d.hip", "device"),
    ("solver3d.hip", "device"),
    ("assembly2d.cpp", "host"),
    ("assembly3d.cpp", "host"),
    ("coupled2d.cpp", "host"),
    ("capi.cpp", "host"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []
    for src, kind in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(BUILD_DIR, src + ".o")
        if kind == "device":
            cmd = [HIPCC, *DEVICE_FLAGS, "-c", path, "-o", obj]
        else:
            cmd = ["g++", *HOST_FLAGS, "-I/opt/rocm/include", "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
