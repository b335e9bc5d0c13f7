"""Build librqcopt_hip.so in-tree with hipcc for gfx950.

    python -m rqcopt_hpc_amd.build        (or __graft_entry__.build())

Compiles every HIP/C++ source under rqcopt_hpc_amd/csrc into one shared
library, rqcopt_hpc_amd/librqcopt_hip.so, exporting the C ABI declared in
include/rqcopt/*.h.  Packed FP32 instructions are disabled for the device
code (-packed-fp32-ops).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OUT = os.path.join(HERE, "librqcopt_hip.so")
ARCH = os.environ.get("RQC_OFFLOAD_ARCH", "gfx950")

SOURCES = ["rqc_engine.hip", "rqc_batch.hip", "rqc_server.hip", "rqc_host.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{INC}", f"-I{CSRC}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build(verbose=True):
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] + DEVICE
        _run([HIPCC] + COMMON + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs +
         ["-L/opt/rocm/lib", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
