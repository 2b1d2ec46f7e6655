"""Builds the native library in-tree: bevy_hair_amd/libhairmi.so (gfx950).

The flags matter for correctness, not only speed:
  -ffp-contract=off                          no FMA contraction: the integer
                                             G-buffer/DOM keys must round like
                                             the CPU restatement;
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE divide and sqrt;
  -target-feature -packed-fp32-ops           no packed FP32 instructions.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libhairmi.so")
SOURCES = ["hm_kernels.hip", "hm_binned.hip", "hm_mlp.hip", "hm_noalloc.hip", "hm_compose.hip", "hm_host.cpp", "hm_bake.cpp", "hm_mlp_host.cpp",
           "hm_batch.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources() -> list[str]:
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build_library(verbose: bool = False) -> str:
    """Compiles every HIP source into libhairmi.so; raises on failure."""
    tmp = LIB + ".tmp"
    cmd = [HIPCC, *FLAGS, "-x", "hip", *sources(), "-o", tmp]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    noise = "packed-fp32-ops"
    err = "\n".join(l for l in proc.stderr.splitlines() if noise not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed ({proc.returncode}):\n{' '.join(cmd)}\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
