"""Build the gfx950 HIP library libns3d_hip.so in-tree.

    python -m ns3dperiodic_amd.build            # compile (always recompiles)

The library is one translation unit (csrc/ns3d_api.hip) exposing the C ABI of
include/ns3d.h.  It is compiled for gfx950 only.
"""
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libns3d_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    "-Wno-unused-function",
    # packed FP32 instructions stay off (only ordinary float code is written)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc():
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return exe


def build(verbose=True, extra_flags=()):
    src = os.path.join(CSRC, "ns3d_api.hip")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *extra_flags, src, "-o", tmp]
    t0 = time.time()
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if verbose:
        print(f"[build] {LIB} ({time.time() - t0:.1f} s)", flush=True)
    build_pattern_ceiling(verbose)
    return LIB


# The access-pattern bandwidth benchmark (a standalone program bench.py runs).
PATTERN_BIN = os.path.join(HERE, "pattern_ceiling")


def build_pattern_ceiling(verbose=True):
    flags = [f for f in HIPCC_FLAGS if f not in ("-shared", "-fPIC")]
    tmp = PATTERN_BIN + ".tmp"
    cmd = [hipcc(), *flags, os.path.join(CSRC, "pattern_ceiling.hip"), "-o", tmp]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, PATTERN_BIN)
    return PATTERN_BIN


if __name__ == "__main__":
    build(extra_flags=sys.argv[1:])
