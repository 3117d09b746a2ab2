"""Builds libmpqp_hip.so for gfx950 with hipcc (in-tree, next to this file)."""
from __future__ import annotations

import os
import shutil
import subprocess

_DIR = os.path.dirname(os.path.abspath(__file__))
SOURCES = [os.path.join(_DIR, "csrc", "mpqp_hip.hip")]
DEPS = [os.path.join(_DIR, "csrc", f) for f in ("lp_engine.hpp", "kernels.hpp", "opt_kernel.hpp", "regions.hpp")] + [
    os.path.join(os.path.dirname(_DIR), "include", "mpqp_hip.h")]
OUT = os.path.join(_DIR, "libmpqp_hip.so")

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # the kernels' arithmetic is specified operation by operation: no fusion
    # beyond the explicit fma() calls
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False, extra_flags: list[str] | None = None) -> str:
    """Compiles the HIP sources into libmpqp_hip.so; returns its path."""
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, *(extra_flags or []), "-o", tmp, *SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{r.stderr[-8000:]}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
