"""Builds libsgs.so for gfx950 with hipcc (no torch involved).

Flags: IEEE evaluation in source order (-ffp-contract=off, no fast math, correctly
rounded fp32 division and sqrt), and packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

CSRC = Path(__file__).resolve().parent
PKG = CSRC.parent
REPO = PKG.parent
OUT = PKG / "StochasticGSCudaBackend" / "libsgs.so"
# measurement kernels (copy ceiling, random atomicMin rate) for bench.py, a separate library
MICROBENCH_OUT = PKG / "StochasticGSCudaBackend" / "libsgs_microbench.so"
MICROBENCH_SOURCES = ["sgs_microbench.hip"]
SOURCES = ["sgs_api.hip", "sgs_accumulate.hip", "sgs_batch.hip", "sgs_hybrid.hip", "sgs_resolve.hip", "sgs_scene.hip"]
ARCH = os.environ.get("SGS_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    return "hipcc"


def flags() -> list[str]:
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-result",
        f"-I{REPO / 'include'}", f"-I{CSRC}",
    ]


def _compile(sources: list[str], out: Path, verbose: bool) -> Path:
    cmd = [hipcc(), *flags(), *[str(CSRC / s) for s in sources], "-o", str(out) + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(str(out) + ".tmp", out)
    return out


def build(verbose: bool = False) -> Path:
    _compile(MICROBENCH_SOURCES, MICROBENCH_OUT, verbose)
    return _compile(SOURCES, OUT, verbose)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
