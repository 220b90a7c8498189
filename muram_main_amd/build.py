"""Build the gfx950 HIP library in-tree (muram_main_amd/libmuram_mi355x.so)."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmuram_mi355x.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fno-fast-math",
    # packed FP32 instructions stay off (see the design notes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# translation units and their contraction mode: the drop-in operations keep
# plain IEEE arithmetic in the reference's operation order (no FMA); the
# fused-step kernels are also built with contraction in mrm_fast.hip; the
# FP32-mixed kernels (mrm_mixed.hip) are contracted; the adjoint kernels
# (mrm_adjoint.hip) are not
UNITS = [("mrm_step.hip", "-ffp-contract=off"), ("mrm_fast.hip", "-ffp-contract=fast"),
         ("mrm_mixed.hip", "-ffp-contract=fast"), ("mrm_adjoint.hip", "-ffp-contract=off"),
         ("mrm_io.cpp", "-ffp-contract=off")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libmuram_mi355x.so")


def build(verbose: bool = False) -> str:
    objs = []
    for src, contract in UNITS:
        obj = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        _run([HIPCC] + FLAGS + [contract, "-I", CSRC, "-c", "-o", obj, os.path.join(CSRC, src)], verbose)
        objs.append(obj)
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, OUT)
    for o in objs:
        os.remove(o)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
