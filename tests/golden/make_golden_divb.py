"""Write tests/golden/ref_divb.npz: the compiled reference's Clean_divB
(clean_divB.C:35-369) on a box whose field has a divergent part, as single
calls and inside full steps (test infrastructure; needs oracle/_ref built by
``make -C oracle ref`` and the reference tree):

    python tests/golden/make_golden_divb.py

Arrays (data only): the input state; after each of four successive
Clean_divB calls div B and phi, and the state after the first and the last;
the state and scalars after 1 and 5 full steps with the cleaning on, with
va_adjust 0 and 1.  The EOS table is
tests/golden/eos_small.dat.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from muram_main_amd import config, eos_table, initial  # noqa: E402
from orc import ref_available, run_reference  # noqa: E402

GSIZE = (12, 10, 8)
DX = (3.2e6, 3.2e6, 4.8e6)


def divergent_state(tab, seed):
    """stratified box + a curl-free field perturbation grad(chi), so that the
    4th-order div B is O(1) of |B|/dx and the cleaning has work to do"""
    box = initial.BoxSpec(gsize=GSIZE, dx=DX, seed=seed, v_rms=2.0e5)
    U = initial.make_state(tab, box)
    rng = np.random.default_rng(seed + 100)
    chi = rng.standard_normal(U.shape[1:]) * 5.0 * DX[0]
    g = np.gradient(chi, DX[2], DX[1], DX[0])     # array axes (k, j, i) -> dims 2, 1, 0
    U[5] += g[2]
    U[6] += g[1]
    U[7] += g[0]
    return U


def main():
    assert ref_available(), "build the reference first: make -C oracle ref"
    raw = open(os.path.join(HERE, "eos_small.dat"), "rb").read()
    tab = eos_table.load_tables(raw)
    grid = config.grid_desc(GSIZE, DX)
    phys = config.stratified_phys()
    phys.divB[:] = [1, 5, 0.2]
    out = {}
    U = divergent_state(tab, 61)
    run = config.run_desc()
    ops = [("STARTUP", 0, False)] + [("DIVB", 0, True)] * 4
    D = run_reference(grid, phys, run, raw, U, ops)
    out["op_in"] = U
    for n, d in enumerate(D):
        if n in (0, len(D) - 1):
            out[f"op{n}_U"] = d["U"]
        out[f"op{n}_divB"] = d["divB"]
        out[f"op{n}_phi"] = d["phi"]
    for name, va_max, vadj in (("steps", 3.0e7, 0.0), ("steps_vadj", 1.0e6, 1.0)):
        p = config.stratified_phys(va_max=va_max, va_adjust=vadj)
        p.divB[:] = [1, 5, 0.2]
        U = divergent_state(tab, 62)
        run = config.run_desc(CFL=0.8)
        D = run_reference(grid, p, run, raw, U, [("STARTUP", 0, False), ("STEP", 1, True), ("STEP", 4, True)])
        out[f"{name}_in"] = U
        out[f"{name}_step1_U"] = D[0]["U"]
        out[f"{name}_step1_sc"] = D[0]["sc"]
        out[f"{name}_step5_U"] = D[1]["U"]
        out[f"{name}_step5_sc"] = D[1]["sc"]
        out[f"{name}_step5_phi"] = D[1]["phi"]
    np.savez_compressed(os.path.join(HERE, "ref_divb.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
