"""Builds the native libraries in-tree (gfx950 only).

  qtorch_amd/lib/libqtorch_hip.so   HIP kernels + the engine's C ABI (include/qtorch_hip.h)
  qtorch_amd/lib/libqtorch_host.so  qTorch's C++ host surface + its C ABI (include/qtorch_host.h)
  qtorch_amd/bin/qtorch             the measure-file CLI (src/main.cpp counterpart)

Run ``python qtorch_amd/build.py`` (hipcc cross-compiles; no GPU needed).
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall",
    # never emit packed FP32 instructions (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O2", "-g", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed: {cmd[0]} (rc={res.returncode})")


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "hip", "*.hip")))
    out = os.path.join(LIB, "libqtorch_hip.so")
    _run([HIPCC, *HIP_FLAGS, "-shared", "-I", os.path.join(REPO, "include"), *srcs, "-lhiprtc", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_host():
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(PKG, "csrc", "host", "*.cpp")))
    lib_srcs = [s for s in srcs if not os.path.basename(s).startswith("main_")]
    inc = ["-I", os.path.join(REPO, "include"), "-I", os.path.join(PKG, "csrc", "host")]
    out = os.path.join(LIB, "libqtorch_host.so")
    _run(["g++", *HOST_FLAGS, "-shared", *inc, *lib_srcs, "-L", LIB, "-lqtorch_hip",
          "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    for main in [s for s in srcs if os.path.basename(s).startswith("main_")]:
        exe = os.path.join(BIN, os.path.basename(main)[len("main_"):-len(".cpp")])
        _run(["g++", *HOST_FLAGS, *inc, main, "-L", LIB, "-lqtorch_host", "-lqtorch_hip",
              "-Wl,-rpath,$ORIGIN/../lib", "-o", exe])
    return out


def build_oracle():
    """The checker: oracle/liboracle.so always; oracle/_ref when the reference sources exist."""
    _run(["make", "-s", "-C", os.path.join(REPO, "oracle")])
    if os.path.isdir(os.environ.get("QTORCH_REFERENCE", "/root/reference")):
        _run([os.path.join(REPO, "oracle", "build_ref.sh")])


def build_all(oracle: bool = True):
    build_hip()
    build_host()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build_all(oracle="--no-oracle" not in sys.argv)
