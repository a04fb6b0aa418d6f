"""Build libgcdt.so (HIP kernels for gfx950 + C++ host driver + C ABI) in-tree.

    python -m gpu_cdt_refine_amd.build        # or: python gpu_cdt_refine_amd/build.py

Compiles every source under csrc/ with hipcc in one shared library next to
this file.  FP contraction is disabled so that the device arithmetic rounds
exactly like the host's (the refined mesh is reproduced bit for bit by the
CPU oracle), and packed-FP32 instruction selection is turned off.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libgcdt.so")
SOURCES = ["gcdt_refine.hip", "gcdt_capi.cpp", "gcdt_batch.cpp", "gcdt_tune.cpp", "cdt_host.cpp",
           "gcdt_mlp.hip", "gcdt_check.hip", "cdt_io.cpp", "gcdt_cdt.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
    "-I" + os.path.join(HERE, "csrc"),
]


def build(verbose=True):
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES
            if os.path.exists(os.path.join(HERE, "csrc", s))]
    tmp = LIB + ".tmp"
    cmd = [HIPCC] + FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libgcdt.so")
    os.replace(tmp, LIB)
    return LIB


DROPIN_SRC = os.path.join(ROOT, "examples", "dropin_main.cpp")
DROPIN_BIN = os.path.join(ROOT, "examples", "dropin_main")


def build_dropin_example(verbose=True):
    """Compile the reference-style caller against include/ and libgcdt.so
    (plain g++: the drop-in needs no HIP headers)."""
    cmd = ["g++", "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"), DROPIN_SRC,
           "-L" + HERE, "-lgcdt", "-Wl,-rpath," + HERE, "-o", DROPIN_BIN]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the drop-in example")
    return DROPIN_BIN


if __name__ == "__main__":
    build()
    build_dropin_example()
