"""Builds libvecsz_hip.so (gfx950) in-tree with hipcc.

    python -m vecsz_amd.build          # or: from vecsz_amd.build import build; build()
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libvecsz_hip.so")
ARCH = os.environ.get("VECSZ_OFFLOAD_ARCH", "gfx950")

SOURCES = ["api.hip", "kernels_compress.hip", "kernels_decompress.hip", "codebook.cpp", "lossless.cpp", "kernels_verify.hip",
           "kernels_pad.hip", "kernels_codebook.hip"]
# zlib (the same 1.2.11 release the reference vendors) for the gzip pass;
# libdl to load the system zstd runtime on first use
LIBS = ["-lz", "-ldl"]

# -ffp-contract=off: no fused multiply-add may change a rounding the
#   reference performs separately (prequant product, reconstruction scale).
# -fno-gpu-flush-denormals-to-zero: IEEE denormals, like the reference's SSE/AVX
#   arithmetic (no FTZ/DAZ).
# -packed-fp32-ops: keep the compiler from forming packed fp32 instructions.
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *FLAGS, "-o", tmp] + [os.path.join(CSRC, s) for s in SOURCES] + LIBS
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, cwd=CSRC, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed")
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
