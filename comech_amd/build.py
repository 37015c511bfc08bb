"""Build of the HIP library ``libcomech_hip.so`` (gfx950) in-tree.

``python -m comech_amd.build`` or ``__graft_entry__.build()``.  hipcc cross-compiles without a GPU.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libcomech_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no packed FP32 instructions in any kernel (ordinary float code only).
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared", "-munsafe-fp-atomics",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include")]

SOURCES = [os.path.join(CSRC, "system.hip"), os.path.join(CSRC, "learned.hip")]
HOST_LIB = os.path.join(HERE, "libcomech_host.so")
HOST_SOURCES = [os.path.join(CSRC, "host_element.cpp")]


def build(verbose=False):
    """Compile every HIP source into comech_amd/libcomech_hip.so (always recompiles)."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + SOURCES
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + err)
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    build_host()
    return LIB


def build_host():
    """CPU build of the element math (g++), for tests without a GPU."""
    tmp = HOST_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", tmp] + HOST_SOURCES
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed:\n" + res.stderr)
    os.replace(tmp, HOST_LIB)
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose=True))
