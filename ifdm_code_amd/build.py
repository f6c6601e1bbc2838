"""Builds libifdm.so (gfx950) in-tree with hipcc.

    python -m ifdm_code_amd.build        # or ifdm_code_amd.build.build()
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libifdm.so")
ARCH = os.environ.get("IFDM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["ifdm_capi.hip", "ifdm_io.hip", "cosmo.cpp"]
HEADERS = ["fft_core.h", "kdk_passes.h", "constants.h", "ctx.h", "fastmath.h", "ic_kernels.h"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def command(out: str = LIB) -> list[str]:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            # packed FP32 instructions stay off: some of their forms have given wrong results on this hardware
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-Wno-unused-result",
            "-o", out] + srcs


CLI = os.path.join(HERE, "ifdm")


def cli_command(out: str = CLI) -> list[str]:
    """The command-line driver: plain C++ over the C ABI."""
    return ["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "ifdm_main.cpp"), "-o", out,
            "-L" + HERE, "-lifdm", "-Wl,-rpath,$ORIGIN"]


def build(verbose: bool = True) -> str:
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libifdm.so")
    os.replace(tmp, LIB)
    cmd = cli_command()
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("g++ failed building the ifdm driver")
    return LIB


if __name__ == "__main__":
    build()
