"""Build recipe of the HIP library: hipcc for gfx950 into diffractom_amd/libdiffractom_hip.so (in-tree, so it travels
with the source tree). `python -m diffractom_amd.build` or __graft_entry__.build()."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libdiffractom_hip.so"
SOURCES = ("operator.hip", "solver.hip", "dense_pf.hip", "spmm_t.hip", "radon_band.hip", "radon_sweep.hip",
           "pf_blocks.hip", "pf_gen.hip", "pf_grad.hip", "fourier.hip",
           "local_basis.hip")
ARCH = os.environ.get("DFM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    raise RuntimeError("hipcc not found")


def build(verbose: bool = True) -> Path:
    srcs = [str(CSRC / s) for s in SOURCES if (CSRC / s).exists()]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           # packed FP32 instructions stay off (their wrong results on this hardware are not caught by tests)
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", str(PKG.parent / "include"), *srcs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
