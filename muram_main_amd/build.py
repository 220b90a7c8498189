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
# (mrm_adjoint.hip) and the potential-field FFTs (mrm_potential.hip) are not
UNITS = [("mrm_step.hip", "-ffp-contract=off"), ("mrm_fast.hip", "-ffp-contract=fast"),
         ("mrm_mixed.hip", "-ffp-contract=fast"), ("mrm_adjoint.hip", "-ffp-contract=off"),
         ("mrm_potential.hip", "-ffp-contract=off"), ("mrm_io.cpp", "-ffp-contract=off")]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libmuram_mi355x.so")


def build(verbose: bool = False) -> str:
    """compile the translation units side by side (at most MAX_JOBS, default
    the CPU count) and link the library"""
    from concurrent.futures import ThreadPoolExecutor
    objs = [os.path.join(CSRC, os.path.splitext(src)[0] + ".o") for src, _ in UNITS]
    cmds = [[HIPCC] + FLAGS + [contract, "-I", CSRC, "-c", "-o", obj, os.path.join(CSRC, src)]
            for (src, contract), obj in zip(UNITS, objs)]
    jobs = max(1, min(len(cmds), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        list(ex.map(lambda cmd: _run(cmd, verbose), cmds))
    tmp = OUT + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, OUT)
    for o in objs:
        os.remove(o)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
