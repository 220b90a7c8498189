"""Write tests/golden/ref_diag.npz: the diagnostics arrays tvar1..8, Qres and
Qvis (grid.H:69-100) the compiled reference collects over the last step of
a run (solver.C:264-296: zeroed at stage 4, then MHD_Residual,
Source_Integrate_Tcheck, TVDlimit, Clean_divB and TCheck add to them) --
test infrastructure; needs oracle/_ref built by ``make -C oracle ref`` and
the reference tree:

    python tests/golden/make_golden_diag.py

Cases (data only: the input state, and the diagnostics over the interior
and the scalars after two plain steps and one collecting step):
  diag_plain  stratified 12x10x8 box, default physics, one TVD pass
  diag_full   24x10x8 box with a 1e6 K corona, Spitzer conduction,
              resistivity, Clean_divB, hot-plate top, and CFL_tvd = 0.1 so
              that TVDlimit runs as two half steps (dt_fac = 0.5)
The EOS table is tests/golden/eos_small.dat.
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


def cases():
    """name -> (box, phys, CFL_tvd)"""
    plain = config.stratified_phys()
    full = config.stratified_phys(va_max=1.0e8)
    full.params[config.I_PARAM_SPITZER] = 1e-6
    full.params[config.I_PARAM_ETA] = 3e9
    full.bnd[config.I_BND_EPS_TOP] = 2.0e14
    full.divB[0] = 1.0
    return {
        "diag_plain": (initial.BoxSpec(gsize=(12, 10, 8), dx=(3.2e6, 3.2e6, 4.8e6), seed=91, v_rms=2.0e5),
                       plain, 0.95),
        "diag_full": (initial.BoxSpec(gsize=(24, 10, 8), dx=(1.6e7, 3.2e6, 3.2e6), seed=92, v_rms=2.0e5,
                                      corona=True, B_vert_G=100.0, dB_rms_G=30.0), full, 0.1),
    }


def main():
    assert ref_available(), "build the reference first: make -C oracle ref"
    raw = open(os.path.join(HERE, "eos_small.dat"), "rb").read()
    tab = eos_table.load_tables(raw)
    out = {}
    inner = (slice(None),) + (slice(2, -2),) * 3
    for name, (box, phys, cfl_tvd) in cases().items():
        U = initial.make_state(tab, box)
        grid = config.grid_desc(box.gsize, box.dx)
        run = config.run_desc(CFL=0.8, CFL_tvd=cfl_tvd)
        D = run_reference(grid, phys, run, raw, U, [("STARTUP", 0, False), ("STEP", 2, False),
                                                    ("STEPDIAG", 1, True)])
        out[f"{name}_in"] = U
        out[f"{name}_diag"] = D[0]["diag"][inner]
        out[f"{name}_sc"] = D[0]["sc"]
        print(name, "dt", D[0]["sc"][0], "max|diag|", np.abs(out[f"{name}_diag"]).reshape(10, -1).max(axis=1))
    np.savez_compressed(os.path.join(HERE, "ref_diag.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
