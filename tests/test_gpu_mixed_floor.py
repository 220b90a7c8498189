"""FP32-mixed step against the FP32-storage floor of the box512_v3
distributions, on a 256 x 64 x 64 crop (the full depth of the stratified box,
a horizontal cut), 100 steps.

The FP32-mixed form stores rho, M and B in FP32 after every RK stage and
after TCheck.  Its floor is the same storage rounding with FP64 arithmetic:
the reference's stage loop (solver.C:109-365) run through the drop-in
operations in FP64, with rho, M and B rounded to FP32 after every stage's
update and after TCheck (control_each_stage).  Both are compared with the
FP64 step from the same initial state; relative L2 (||a - b|| / ||b||) per
field over the interior's lower half (rho >~ 1e-11 g/cm^3) and upper half.

Bar: the mixed form's largest per-field error within 2 times the floor's,
in each half.  Measured: lower half 1.21e-5 against 1.02e-5 (1.2x), upper
half 1.08e-2 against 7.1e-3 (1.5x).  With the FP32 library logarithm in the
TVD limiter's ln rho and ln eps (up to 2.3 ulps of error, biased low by 0.35
to 0.65 ulp on average) the lower half was 2.8e-5 (2.75x: M1).
"""
import numpy as np
import pytest

from muram_main_amd import abi, config, eos_table, initial
from muram_main_amd.mhd import MhdGrid

pytestmark = pytest.mark.gpu
NSTEPS = 100
GSIZE = (256, 64, 64)
FIELDS = ("rho", "M0", "M1", "M2", "e", "B0", "B1", "B2")


def rounded(V):
    V = V.copy()
    for v in (0, 1, 2, 3, 5, 6, 7):
        V[v] = V[v].astype(np.float32).astype(np.float64)
    return V


def run_step(table, g, phys, U, prec):
    m = MhdGrid(g, phys, prec)
    m.eos_init(table)
    r = config.run_desc()
    m.set_state(U)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    for _ in range(NSTEPS):
        m.step(r, config.DT_RAD_OFF, 1)
    out = m.get_state()
    m.close()
    return out


def run_stage_loop_rounded(table, g, phys, U):
    """solver.C:109-365 through the drop-in operations (FP64), rho, M and B
    rounded to FP32 after each stage's update and after TCheck"""
    m = MhdGrid(g, phys, abi.PREC_FP64)
    m.eos_init(table)
    r = config.run_desc()
    m.set_state(U)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    for _ in range(NSTEPS):
        m.SaveCons()
        cfl = 0.0
        for stage in range(1, 5):
            m.ConsToPrim()
            if stage == 1:
                m.Adjust_Valf_Max(r)
            dt_mhd = m.MHD_Residual(r)
            if stage == 1:
                cfl = m.Sync_timestep(r, config.DT_RAD_OFF, dt_mhd)
            m.Source_Integrate_Tcheck(r, stage)
            m.set_state(rounded(m.get_state()))
            m.exchange_grid()
            m.SetBoundaryConditions(r, stage)
        if cfl <= r.CFL_tvd:
            m.TVDlimit(r, r.dt)
        else:
            m.TVDlimit(r, 0.5 * r.dt)
            m.exchange_grid()
            m.TVDlimit(r, 0.5 * r.dt)
        m.TCheck(r)
        m.set_state(rounded(m.get_state()))
        m.exchange_grid()
        m.SetBoundaryConditions(r, 5)
        r.iteration += 1
        r.globiter += 1
        r.time += r.dt
    out = m.get_state()
    m.close()
    return out


def max_l2(a, b, sl):
    return max(float(np.linalg.norm(a[v][sl] - b[v][sl]) / np.linalg.norm(b[v][sl])) for v in range(len(FIELDS)))


def test_mixed_within_storage_floor():
    print(f"case: FP32-mixed vs the FP32-storage floor, box512_v3 on {GSIZE}, {NSTEPS} steps", flush=True)
    table = eos_table.generate(eos_table.TableSpec(n_eps=512, n_lr=512, n_lp3=200, n_s3=150))
    tab = eos_table.load_tables(table)
    wl = config.workloads()["box512_v3"]
    box = initial.BoxSpec(gsize=GSIZE, dx=wl.box.dx, v_rms=wl.box.v_rms, seed=wl.box.seed)
    U = initial.make_state(tab, box)
    g = config.grid_desc(box.gsize, box.dx)
    ref = run_step(table, g, wl.phys, U, abi.PREC_FP64)
    mixed = run_step(table, g, wl.phys, U, abi.PREC_FP32_MIXED)
    floor = run_stage_loop_rounded(table, g, wl.phys, rounded(U))
    n0 = GSIZE[0]
    inner = (slice(2, -2),) * 2
    for name, k in (("lower", slice(2, 2 + n0 // 2)), ("upper", slice(2 + n0 // 2, 2 + n0))):
        sl = inner + (k,)
        em, ef = max_l2(mixed, ref, sl), max_l2(floor, ref, sl)
        print(f"{name} half: mixed max rel L2 {em:.3e}, storage floor {ef:.3e} ({em / ef:.2f}x)", flush=True)
        assert em <= 2.0 * ef, (name, em, ef)
