"""Write tests/golden/ref_damp.npz: the compiled reference with the damping
on (dmp_switch 1; get_damping, Add_Sources_Integrate_SR.C:92-108, 628-761)
on a stratified box with a mean vertical flow (test infrastructure; needs
oracle/_ref built by ``make -C oracle ref`` and the reference tree):

    python tests/golden/make_golden_damp.py

Arrays (data only): per case the input state, the state after one
ConsToPrim + Adjust_Valf_Max + MHD_Residual + Sync_timestep +
Source_Integrate_Tcheck(stage 1), and the state and scalars after 1 and 5
full steps.  Case "damp" has vel_ref 1e4 (the lower layers on the
1/tau_min cap, the upper ones on the power branch), "damp_weak" vel_ref
1e6 (every layer on the power branch).  The EOS table is
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
from damp_cases import damp_phys, upflow_state  # noqa: E402

GSIZE = (12, 10, 8)
DX = (3.2e6, 3.2e6, 4.8e6)
CASES = {"damp": (1.0e4, 81), "damp_weak": (1.0e6, 82)}


def main():
    assert ref_available(), "build the reference first: make -C oracle ref"
    raw = open(os.path.join(HERE, "eos_small.dat"), "rb").read()
    tab = eos_table.load_tables(raw)
    grid = config.grid_desc(GSIZE, DX)
    out = {}
    for name, (vel_ref, seed) in CASES.items():
        phys = damp_phys(vel_ref)
        U = upflow_state(tab, initial.BoxSpec(gsize=GSIZE, dx=DX, seed=seed, v_rms=2.0e5))
        run = config.run_desc(CFL=0.8)
        D = run_reference(grid, phys, run, raw, U,
                          [("STARTUP", 0, False), ("SAVECONS", 0, False), ("C2P", 0, False), ("VALF", 0, False),
                           ("RESID", 0, False), ("SYNC", 0, False), ("SRCINT", 1, True)])
        out[f"{name}_in"] = U
        out[f"{name}_srcint_U"] = D[0]["U"]
        out[f"{name}_srcint_sc"] = D[0]["sc"]
        D = run_reference(grid, phys, run, raw, U, [("STARTUP", 0, False), ("STEP", 1, True), ("STEP", 4, True)])
        out[f"{name}_step1_U"] = D[0]["U"]
        out[f"{name}_step1_sc"] = D[0]["sc"]
        out[f"{name}_step5_U"] = D[1]["U"]
        out[f"{name}_step5_sc"] = D[1]["sc"]
        out[f"{name}_vel_ref"] = np.array(vel_ref)
    np.savez_compressed(os.path.join(HERE, "ref_damp.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
