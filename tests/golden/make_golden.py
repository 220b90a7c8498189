"""Write the golden fixtures of the compiled reference (test infrastructure).

Run in a container that has the reference tree and oracle/_ref built
(``make -C oracle ref``):

    python tests/golden/make_golden.py

Outputs (committed):
  tests/golden/eos_small.dat   the EOS table the fixtures were made with
                               (this build's generator, reference binary format)
  tests/golden/ref_small.npz   inputs and reference outputs on a 10x8x6 box

Every array is data: the seeded input state and what the reference's own
functions returned for it.
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
from orc import run_reference, ref_available  # noqa: E402

GSIZE = (10, 8, 6)
DX = (3.2e6, 3.2e6, 3.2e6)


def edge_state(tab, U):
    """Cells that drive ConsToPrim through each fallback (cons_to_prim.C:180-204)."""
    E = U.copy()
    eps_max = np.exp(tab.xeps[-1]) - tab.eps_off
    rho_low = np.exp(tab.xlr[0]) * 0.5
    k, j = 3, 4
    # ln rho below the table: ideal gas with pV = nRT electrons
    E[0, k, j, 3] = rho_low
    E[4, k, j, 3] = rho_low * 1.0e12 + 0.5 * (E[1, k, j, 3] ** 2 + E[2, k, j, 3] ** 2 + E[3, k, j, 3] ** 2) / rho_low
    # eps above 0.8 eps_max: blend with the ideal gas
    r = E[0, k, j, 5]
    ek = 0.5 * (E[1, k, j, 5] ** 2 + E[2, k, j, 5] ** 2 + E[3, k, j, 5] ** 2) / r
    E[4, k, j, 5] = r * 0.9 * eps_max + ek
    E[4, k, j, 6] = r * 1.2 * eps_max + ek
    # negative internal energy: invalid p, T -> ideal gas; n_e floor
    E[4, k, j, 7] = ek - 3.0e10 * r
    # eps + eps_off below the table, ln rho above the table
    E[4, k, j, 8] = ek + r * 1.0e9
    E[0, k, j, 9] = 5.0e-3
    return E


def main():
    assert ref_available(), "build the reference first: make -C oracle ref"
    raw = eos_table.generate(eos_table.TableSpec())
    with open(os.path.join(HERE, "eos_small.dat"), "wb") as f:
        f.write(raw)
    tab = eos_table.load_tables(raw)
    grid = config.grid_desc(GSIZE, DX)
    out = {}
    cases = {
        "base": (initial.BoxSpec(gsize=GSIZE, dx=DX, seed=7), config.stratified_phys(), 0.8),
        "split": (initial.BoxSpec(gsize=GSIZE, dx=DX, seed=8, v_rms=3.0e5), config.stratified_phys(), 0.99),
        "vadj": (initial.BoxSpec(gsize=GSIZE, dx=DX, seed=9, v_rms=3.0e5),
                 config.stratified_phys(va_max=3.0e7, va_adjust=1.0), 0.8),
    }
    for name, (box, phys, cfl) in cases.items():
        U = initial.make_state(tab, box)
        run = config.run_desc(CFL=cfl)
        if name == "base":
            ops = [("STARTUP", 0, True), ("SAVECONS", 0, False), ("C2P", 0, True), ("VALF", 0, False),
                   ("RESID", 0, True), ("SYNC", 0, False), ("SRCINT", 1, True),
                   ("EXCHANGE", 1, False), ("BC", 1, True)]
            D = run_reference(grid, phys, run, raw, U, ops)
            out["base_in"] = U
            out["base_startup_U"] = D[0]["U"]
            out["base_c2p_U"] = D[1]["U"]
            out["base_c2p_F"] = D[1]["fields"]
            out["base_resid_R"] = D[2]["Res"]
            out["base_resid_sc"] = D[2]["sc"]
            out["base_srcint_U"] = D[3]["U"]
            out["base_srcint_sc"] = D[3]["sc"]
            out["base_bc1_U"] = D[4]["U"]
            # TVD alone on the state after one integrate + BC (pres from the stage's ConsToPrim)
            ops_tvd = ops + [("TVD", 0, True)]
            D = run_reference(grid, phys, run, raw, U, ops_tvd)
            out["base_tvd_U"] = D[-1]["U"]
            # ConsToPrim fallbacks
            E = edge_state(tab, D[0]["U"])
            D = run_reference(grid, phys, run, raw, E, [("C2P", 0, True)])
            out["edge_in"] = E
            out["edge_c2p_U"] = D[0]["U"]
            out["edge_c2p_F"] = D[0]["fields"]
        ops = [("STARTUP", 0, False), ("STEP", 1, True), ("STEP", 4, True)]
        D = run_reference(grid, phys, run, raw, U, ops)
        out[f"{name}_in"] = U
        out[f"{name}_step1_U"] = D[0]["U"]
        out[f"{name}_step1_sc"] = D[0]["sc"]
        out[f"{name}_step5_U"] = D[1]["U"]
        out[f"{name}_step5_sc"] = D[1]["sc"]
        out[f"{name}_cfl"] = np.array(cfl)
    np.savez_compressed(os.path.join(HERE, "ref_small.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
