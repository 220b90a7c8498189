"""Write tests/golden/ref_spitzer.npz: the compiled reference with Spitzer
heat conduction (param_spitzer > 0; mhdres_SR.C:350-359, 486-521, 680-690;
Add_Sources_Integrate_SR.C:410-419, 504-517; boundary_pdmp_1_fftw3.C:306-310,
455-466) over full steps from sflx = 0 (test infrastructure; needs
oracle/_ref built by ``make -C oracle ref`` and the reference tree):

    python tests/golden/make_golden_spitzer.py

Cases (data only: inputs, the state and sflx after 1 and 5 steps, scalars):
  spt_cool    stratified 12x10x8 box at 6000 K, zero-flux top (bnd_eps_top 0)
  spt_corona  24x10x8 box whose upper half is a 1e6 K corona (dz 160 km),
              hot plate at the top (bnd_eps_top 2e14: symmetric sflx ghosts)
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
    cool = config.stratified_phys()
    cool.params[config.I_PARAM_SPITZER] = 1e-6
    hot = config.stratified_phys(va_max=1.0e8)
    hot.params[config.I_PARAM_SPITZER] = 1e-6
    hot.bnd[config.I_BND_EPS_TOP] = 2.0e14
    return {
        "spt_cool": (initial.BoxSpec(gsize=(12, 10, 8), dx=(3.2e6, 3.2e6, 4.8e6), seed=81, v_rms=2.0e5), cool),
        "spt_corona": (initial.BoxSpec(gsize=(24, 10, 8), dx=(1.6e7, 3.2e6, 3.2e6), seed=82, v_rms=2.0e5,
                                       corona=True, B_vert_G=100.0, dB_rms_G=30.0), hot),
    }


def main():
    assert ref_available(), "build the reference first: make -C oracle ref"
    raw = open(os.path.join(HERE, "eos_small.dat"), "rb").read()
    tab = eos_table.load_tables(raw)
    out = {}
    for name, (box, phys) in cases().items():
        U = initial.make_state(tab, box)
        grid = config.grid_desc(box.gsize, box.dx)
        run = config.run_desc(CFL=0.8)
        D = run_reference(grid, phys, run, raw, U, [("STARTUP", 0, False), ("STEP", 1, True), ("STEP", 4, True)])
        out[f"{name}_in"] = U
        for n, d in zip((1, 5), D):
            out[f"{name}_step{n}_U"] = d["U"]
            out[f"{name}_step{n}_sflx"] = d["sflx"]
            out[f"{name}_step{n}_sc"] = d["sc"]
        print(name, "dt", D[1]["sc"][0], "max|sflx|", np.abs(D[1]["sflx"]).max())
    np.savez_compressed(os.path.join(HERE, "ref_spitzer.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
