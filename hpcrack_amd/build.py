"""Builds the gfx950 HIP library (libhpcrack_amd.so) and the host driver in-tree.

    python -m hpcrack_amd.build
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libhpcrack_amd.so")
DRIVER = os.path.join(HERE, "hpcrack_run")
SERVER = os.path.join(HERE, "hpcrack_serve")

HIP_SOURCES = ["csrc/hpcr_kernels.hip", "csrc/hpcr_mixed.hip", "csrc/hpcr_solver.hip"]
HOST_SOURCES = ["csrc/hpcr_io.cpp"]
DRIVER_SOURCES = ["host/hpcrack_run.cpp"]
SERVER_SOURCES = ["host/hpcrack_serve.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# FP32 packed-math instructions are kept out of the generated code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, cwd=HERE, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "unknown target feature" not in l
                    and "packed-fp32-ops" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    srcs = [s for s in HIP_SOURCES + HOST_SOURCES if os.path.exists(os.path.join(HERE, s))]
    tmp = LIB + ".tmp"
    _run([HIPCC] + COMMON + DEVICE_FLAGS + ["-shared", "-o", tmp] + srcs)
    os.replace(tmp, LIB)
    drv = [s for s in DRIVER_SOURCES if os.path.exists(os.path.join(HERE, s))]
    if drv:
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o", DRIVER] + drv +
             ["-L", HERE, "-lhpcrack_amd", "-Wl,-rpath,$ORIGIN"])
    srv = [s for s in SERVER_SOURCES if os.path.exists(os.path.join(HERE, s))]
    if srv:
        _run(["g++", "-O2", "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "include"), "-o", SERVER] + srv +
             ["-L", HERE, "-lhpcrack_amd", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
