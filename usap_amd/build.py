"""Builds usap_amd/lib/libusap.so in-tree.

Device code: hipcc --offload-arch=gfx950 (packed-FP32 ops disabled).
Host code (driver, utilities): g++ with -ffp-contract=off, so the double
arithmetic of the driver is plain IEEE (no FMA contraction).

Usage: python -m usap_amd.build   (or usap_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OUT_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(OUT_DIR, "libusap.so")
OBJ_DIR = os.path.join(ROOT, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("USAP_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    # hipcc's host pass warns about the device-only target feature; keep
    # everything else visible
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l)
    if err.strip():
        sys.stderr.write(err + "\n")


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    inc = ["-I", INC, "-I", CSRC]
    objs = []
    o = os.path.join(OBJ_DIR, "engine.o")
    _run([HIPCC, *HIP_FLAGS, *inc, "-c", os.path.join(CSRC, "engine.hip"), "-o", o])
    objs.append(o)
    for src in ("host.cpp", "driver.cpp", "batch.cpp", "serve.cpp"):
        o = os.path.join(OBJ_DIR, src.replace(".cpp", ".o"))
        _run(["g++", *CXX_FLAGS, *inc, "-c", os.path.join(CSRC, src), "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
