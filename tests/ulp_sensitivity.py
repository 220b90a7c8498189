"""Sensitivity of the FP64 step to rounding, measured with the oracle alone.

The oracle runs a workload for N steps twice: from the seeded state, and from
the same state with half of its values (chosen at random) moved by one ulp.
It prints the per-field differences between the two runs (max|d| / max|ref|
and relative L2) and the limiter branch accounting of both: decisive counts,
near ties and exact ties (the rule of mrm_branch_count).  Two FP64
implementations whose transcendental functions differ in the last bit cannot
be expected to agree more closely than this control does.

    python tests/ulp_sensitivity.py box512 10          # box512, 10 steps
    python tests/ulp_sensitivity.py box64 10 - 1e-10   # near-tie margin 1e-10

Arguments: workload name (config.workloads()), steps, optional interior size
"n0xn1xn2" (or -), optional near-tie margin.  CPU only; box512 needs about
35 GB of memory and a few minutes per step on 8 cores.
"""
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import numpy as np  # noqa: E402

from muram_main_amd import abi, config, eos_table, initial  # noqa: E402
from orc import Oracle  # noqa: E402


def main():
    wlname, nsteps = sys.argv[1], int(sys.argv[2])
    gs = tuple(int(x) for x in sys.argv[3].split("x")) if len(sys.argv) > 3 and sys.argv[3] != "-" else None
    tie = float(sys.argv[4]) if len(sys.argv) > 4 else abi.BRANCH_TIE
    spec = eos_table.TableSpec(n_eps=512, n_lr=512, n_lp3=200, n_s3=150)
    table = eos_table.generate(spec)
    tab = eos_table.load_tables(table)
    wl = config.workloads()[wlname]
    box = wl.box
    if gs:
        box = initial.BoxSpec(gsize=gs, dx=box.dx, v_rms=box.v_rms, seed=box.seed, corona=box.corona,
                              B_vert_G=box.B_vert_G, dB_rms_G=box.dB_rms_G)
    U = initial.make_state(tab, box)
    g = config.grid_desc(box.gsize, box.dx)

    def run(U0, tag):
        o = Oracle(g, wl.phys, table)
        r = config.run_desc()
        o.set_state(U0)
        o.exchange_grid()
        o.set_boundary_conditions(r, 0)
        o.branch_count(1, tie)
        for s in range(nsteps):
            t = time.time()
            o.step(r, config.DT_RAD_OFF, 1)
            print(tag, "step", s + 1, "dt", r.dt, f"{time.time() - t:.1f}s", flush=True)
        return o.state(0).copy(), o.branch_counts().copy(), r.dt

    A, ca, dta = run(U, "ref")
    rng = np.random.default_rng(7)
    V = U
    del U
    mask = rng.random(V.shape) < 0.5
    V[mask] = np.nextafter(V[mask], np.where(rng.random(int(mask.sum())) < 0.5, np.inf, -np.inf))
    B, cb, dtb = run(V, "perturbed")
    linf = [float(np.max(np.abs(A[v] - B[v])) / np.max(np.abs(A[v]))) for v in range(8)]
    l2 = [float(np.linalg.norm(A[v] - B[v]) / np.linalg.norm(A[v])) for v in range(8)]
    print(f"after {nsteps} steps, per field (rho, M0, M1, M2, e, B0, B1, B2):")
    print("  max|d|/max|ref|:", ["%.2e" % x for x in linf])
    print("  relative L2:    ", ["%.2e" % x for x in l2])
    v = int(np.argmax(linf))
    d = np.abs(A[v] - B[v])
    print("  largest difference: field", v, "at (k, j, i)", np.unravel_index(int(np.argmax(d)), d.shape))
    print("dt", dta, dtb)
    print("branch accounting:", abi.branch_mismatch(ca, cb))


if __name__ == "__main__":
    main()
