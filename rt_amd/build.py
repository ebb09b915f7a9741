"""Build the in-tree native libraries.

librt_hip.so         HIP kernels + the C ABI of include/rt_hip.h (hipcc, gfx950)
librt_engine_hip.so  the rt_engine_func.h C++ plugin API over that ABI

Run as `python -m rt_amd.build` or through __graft_entry__.build().  Every
HIP translation unit is built with packed FP32 math disabled and hardware
FP64 atomics enabled; FP contraction is left on (FMA), parity is to 1e-12.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", f"-I{INCLUDE}",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if p.returncode != 0:
        raise RuntimeError(f"build step failed ({p.returncode}): {' '.join(cmd)}")


def build_hip():
    out = os.path.join(HERE, "librt_hip.so")
    srcs = [os.path.join(CSRC, f) for f in ("rt_trace.hip",)]
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-o", out] + srcs)
    return out


def build_all():
    build_hip()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
