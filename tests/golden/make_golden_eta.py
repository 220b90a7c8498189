"""Write tests/golden/ref_eta.npz: the compiled reference with resistivity
(param_eta > 0; mhdres_SR.C:422-462, 742-770, 815-819): one MHD_Residual on
a stratified box and full steps (test infrastructure; needs oracle/_ref built
by ``make -C oracle ref`` and the reference tree):

    python tests/golden/make_golden_eta.py

Arrays (data only): per case the input state, the residual and its dt_cfl
after one ConsToPrim + Adjust_Valf_Max + MHD_Residual, and the state and
scalars after 1 and 5 full steps.  eta = 2e11 cm^2/s leaves the CFL limit in
charge; 1e12 makes the diffusion limit dt.  The EOS table is
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
CASES = {"eta_cfl": (2.0e11, 71), "eta_diff": (1.0e12, 72)}


def main():
    assert ref_available(), "build the reference first: make -C oracle ref"
    raw = open(os.path.join(HERE, "eos_small.dat"), "rb").read()
    tab = eos_table.load_tables(raw)
    grid = config.grid_desc(GSIZE, DX)
    out = {}
    for name, (eta, seed) in CASES.items():
        phys = config.stratified_phys()
        phys.params[config.I_PARAM_ETA] = eta
        U = initial.make_state(tab, initial.BoxSpec(gsize=GSIZE, dx=DX, seed=seed, v_rms=2.0e5))
        run = config.run_desc(CFL=0.8)
        D = run_reference(grid, phys, run, raw, U,
                          [("STARTUP", 0, False), ("C2P", 0, False), ("VALF", 0, False), ("RESID", 0, True)])
        out[f"{name}_in"] = U
        out[f"{name}_resid_R"] = D[0]["Res"]
        out[f"{name}_resid_sc"] = D[0]["sc"]
        D = run_reference(grid, phys, run, raw, U, [("STARTUP", 0, False), ("STEP", 1, True), ("STEP", 4, True)])
        out[f"{name}_step1_U"] = D[0]["U"]
        out[f"{name}_step1_sc"] = D[0]["sc"]
        out[f"{name}_step5_U"] = D[1]["U"]
        out[f"{name}_step5_sc"] = D[1]["sc"]
        out[f"{name}_eta"] = np.array(eta)
    np.savez_compressed(os.path.join(HERE, "ref_eta.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
