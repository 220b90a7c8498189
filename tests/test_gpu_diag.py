"""Diagnostics (tvar1..8, Qres, Qvis; grid.H:69-100) on the GPU: mrm_step
with need_diagnostics collects them over its last step through the
per-operation kernels (solver.C:264-296).  Checked against the compiled
reference's arrays in tests/golden/ref_diag.npz, against the oracle from the
same input, and against the energy budget the arrays close."""
import os
import sys

import numpy as np
import pytest

from muram_main_amd import abi, config
from muram_main_amd.mhd import MhdGrid
from orc import Oracle, relerr

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden_diag  # noqa: E402

INNER = (slice(2, -2),) * 3
# per array, max |gpu - ref| / max |ref| over the interior
TOL = 1e-12


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "ref_diag.npz"))


def gpu_case(table, gold, case, phys=None, plain_steps=2):
    box, phys0, cfl_tvd = make_golden_diag.cases()[case]
    m = MhdGrid(config.grid_desc(box.gsize, box.dx), phys or phys0)
    m.eos_init(table)
    r = config.run_desc(CFL=0.8, CFL_tvd=cfl_tvd)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    if plain_steps:
        m.step(r, 1e30, plain_steps)
    return m, r


def diag_array(m):
    d = m.diagnostics()
    return np.stack([d[n][INNER] for n in abi.DIAG_NAMES])


@pytest.mark.parametrize("case", ["diag_plain", "diag_full"])
def test_diagnostics_match_reference(table, gold, case):
    print(f"case: diagnostics vs compiled reference, {case}, 2 plain steps + 1 collecting step", flush=True)
    m, r = gpu_case(table, gold, case)
    r.need_diagnostics = 1
    m.step(r, 1e30, 1)
    assert r.need_diagnostics == 1
    assert r.dt == pytest.approx(gold[f"{case}_sc"][0], rel=1e-12)
    G = diag_array(m)
    errs = {n: relerr(G[t], gold[f"{case}_diag"][t]) for t, n in enumerate(abi.DIAG_NAMES)}
    print(" ".join(f"{n} {e:.2e}" for n, e in errs.items()))
    assert max(errs.values()) <= TOL, errs


@pytest.mark.parametrize("case", ["diag_plain", "diag_full"])
def test_diagnostics_match_oracle_from_same_state(table, gold, case):
    print(f"case: diagnostics vs oracle, {case}, one collecting step", flush=True)
    box, phys, cfl_tvd = make_golden_diag.cases()[case]
    m, r = gpu_case(table, gold, case, plain_steps=0)
    o = Oracle(config.grid_desc(box.gsize, box.dx), phys, table)
    ro = config.run_desc(CFL=0.8, CFL_tvd=cfl_tvd)
    o.set_state(gold[f"{case}_in"])
    o.exchange_grid()
    o.set_boundary_conditions(ro, 0)
    r.need_diagnostics = ro.need_diagnostics = 1
    m.step(r, 1e30, 1)
    o.step(ro, 1e30, 1)
    G, Go = diag_array(m), o.diag()[(slice(None),) + INNER]
    errs = {n: relerr(G[t], Go[t]) for t, n in enumerate(abi.DIAG_NAMES)}
    print(" ".join(f"{n} {e:.2e}" for n, e in errs.items()))
    assert max(errs.values()) <= TOL, errs
    assert relerr(m.get_state()[4], o.state(0)[4]) <= 1e-12


@pytest.mark.parametrize("case", ["diag_plain", "diag_full"])
def test_gpu_diagnostics_close_the_energy_budget(table, gold, case):
    """e(after) - e(before) = dt (tvar1 + tvar2 + tvar3 + tvar5 + tvar6 +
    tvar7 + tvar8 + Qres) per interior cell, with Qtot = 0 and no
    resistivity (whose eta J^2 has no diagnostic)"""
    print(f"case: diagnostics energy budget on the GPU, {case}", flush=True)
    box, phys, _ = make_golden_diag.cases()[case]
    phys.params[config.I_PARAM_ETA] = 0.0
    m, r = gpu_case(table, gold, case, phys=phys, plain_steps=1)
    e0 = m.get_state()[4][INNER]
    r.need_diagnostics = 1
    m.step(r, 1e30, 1)
    de = m.get_state()[4][INNER] - e0
    G = diag_array(m)
    rate = G[0] + G[1] + G[2] + G[4] + G[5] + G[6] + G[7] + G[8]
    err = np.abs(de - r.dt * rate).max() / np.abs(de).max()
    print("budget imbalance", err)
    assert err < 1e-12


def test_diagnostics_api(table, gold):
    print("case: diagnostics API (zero before use, zeroing, per-operation calls, refusals)", flush=True)
    case = "diag_plain"
    m, r = gpu_case(table, gold, case, plain_steps=1)
    assert not any(np.any(v) for v in m.diagnostics().values())
    # the per-operation sequence of a collecting stage 4 adds to the arrays
    m.zero_diagnostics()
    r.need_diagnostics = 1
    m.SaveCons()
    m.ConsToPrim()
    m.MHD_Residual(r)
    d1 = m.diagnostics()
    assert np.any(d1["tvar1"]) and np.any(d1["tvar3"]) and not np.any(d1["tvar4"])
    m.Source_Integrate_Tcheck(r, 4)
    d2 = m.diagnostics()
    assert np.array_equal(d2["tvar1"], d1["tvar1"])
    assert np.all(d2["tvar4"][INNER] == 1.0)   # Qtot = 0: efaci = 1
    # without need_diagnostics nothing changes
    r.need_diagnostics = 0
    m.TVDlimit(r, r.dt)
    assert all(np.array_equal(v, d2[k]) for k, v in m.diagnostics().items())
    m.zero_diagnostics()
    assert not any(np.any(v) for v in m.diagnostics().values())
    # a call with the flag set after one with it clear zeroes the arrays
    # first (solver.C:264-296); consecutive flagged calls accumulate
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.ConsToPrim()
    r.need_diagnostics = 1
    m.MHD_Residual(r)
    t1 = m.diagnostics()["tvar1"]
    m.MHD_Residual(r)
    assert np.array_equal(m.diagnostics()["tvar1"], 2 * t1)
    r.need_diagnostics = 0
    m.Sync_timestep(r, 1e30, 1.0)   # any call taking the run descriptor
    r.need_diagnostics = 1
    m.MHD_Residual(r)
    assert np.array_equal(m.diagnostics()["tvar1"], t1)
    # FP32-mixed contexts do not collect diagnostics
    box, phys, _ = make_golden_diag.cases()[case]
    mx = MhdGrid(config.grid_desc(box.gsize, box.dx), phys, abi.PREC_FP32_MIXED)
    mx.eos_init(table)
    mx.set_state(gold[f"{case}_in"])
    rr = config.run_desc()
    rr.need_diagnostics = 1
    with pytest.raises(abi.MrmError, match="UNSUPPORTED"):
        mx.step(rr, 1e30, 1)
