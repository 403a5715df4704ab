"""Build recipe for the gfx950 shared library (and the test-only oracle).

The library is compiled in-tree with hipcc into
``spec_spars_amd/libspecspars_hip.so``; nothing is installed elsewhere.
"""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libspecspars_hip.so")

HIP_SOURCES = ["kernels.hip", "core_kernels.hip", "batch_kernels.hip", "device.hip", "noalloc.hip"]
CPP_SOURCES = ["factor.cpp", "graph.cpp", "capi.cpp", "sparsifier.cpp", "mtx.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 ops are disabled for all device code (their wrong-result
# failure modes are not reliably caught by tests on this hardware).
DEVICE_FLAGS = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r.stdout


def build_library(verbose=False):
    """Compile every HIP/C++ source for gfx950 and link the C-ABI library."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, cmds = [], []
    for src in HIP_SOURCES + CPP_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmds.append([HIPCC, "-c"] + lang + [os.path.join(CSRC, src), "-o", obj] + COMMON + DEVICE_FLAGS)
        objs.append(obj)
    # one hipcc per source, at most 8 at a time
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        for out in ex.map(_run, cmds):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-o", tmp] + objs + DEVICE_FLAGS + ["-fPIC"])
    os.replace(tmp, LIB)
    return LIB


ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")


def build_oracle():
    """Compile the CPU oracle (test infrastructure, not part of the library)."""
    src = os.path.join(ORACLE_DIR, "oracle.c")
    tmp = ORACLE_LIB + ".tmp"
    _run(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-std=c99",
          "-o", tmp, src, "-lm"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_oracle())
