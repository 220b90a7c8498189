"""Resistivity (param_eta > 0) on the GPU against the compiled reference's
outputs in tests/golden/ref_eta.npz: MHD_Residual with its diffusion-limited
dt, and full steps on the fused and the drop-in sequence paths, within 1e-12
per field."""
import os

import numpy as np
import pytest

from muram_main_amd import abi, config
from muram_main_amd.mhd import MhdGrid
from orc import relerr

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12
GSIZE, DX = (12, 10, 8), (3.2e6, 3.2e6, 4.8e6)


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "ref_eta.npz"))


def eta_phys(eta):
    p = config.stratified_phys()
    p.params[config.I_PARAM_ETA] = eta
    return p


def make(table, phys, path="fused"):
    saved = os.environ.get("MRM_STEP_PATH")
    os.environ["MRM_STEP_PATH"] = path
    try:
        m = MhdGrid(config.grid_desc(GSIZE, DX), phys)
    finally:
        if saved is None:
            os.environ.pop("MRM_STEP_PATH", None)
        else:
            os.environ["MRM_STEP_PATH"] = saved
    m.eos_init(table)
    return m


def maxrel(a, b):
    return max(relerr(a[v], b[v]) for v in range(a.shape[0]))


@pytest.mark.parametrize("case", ["eta_cfl", "eta_diff"])
def test_resistive_residual_matches_reference(table, gold, case):
    print(f"case: resistive residual, {case}", flush=True)
    m = make(table, eta_phys(float(gold[f"{case}_eta"])))
    r = config.run_desc(CFL=0.8)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.ConsToPrim()
    m.Adjust_Valf_Max(r)
    dt = m.MHD_Residual(r)
    e = maxrel(m.get_state(abi.RES), gold[f"{case}_resid_R"])
    print("residual rel err", e, "dt", dt, gold[f"{case}_resid_sc"][2])
    assert e <= TOL
    assert dt == pytest.approx(gold[f"{case}_resid_sc"][2], rel=TOL)


@pytest.mark.parametrize("path", ["fused", "sequence"])
@pytest.mark.parametrize("case", ["eta_cfl", "eta_diff"])
def test_resistive_steps_match_reference(table, gold, case, path):
    print(f"case: resistive steps, {case}, {path}", flush=True)
    m = make(table, eta_phys(float(gold[f"{case}_eta"])), path)
    r = config.run_desc(CFL=0.8)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.step(r, 1e30, 1)
    e1 = maxrel(m.get_state(), gold[f"{case}_step1_U"])
    assert r.dt == pytest.approx(gold[f"{case}_step1_sc"][0], rel=TOL)
    m.step(r, 1e30, 4)
    e5 = maxrel(m.get_state(), gold[f"{case}_step5_U"])
    print("step1", e1, "step5", e5, "dt", r.dt)
    assert e1 <= TOL and e5 <= TOL
    assert r.dt == pytest.approx(gold[f"{case}_step5_sc"][0], rel=TOL)
    assert r.time == pytest.approx(gold[f"{case}_step5_sc"][1], rel=TOL)


@pytest.mark.parametrize("case", ["eta_cfl", "eta_diff"])
def test_mixed_resistive_steps(table, gold, case):
    """the FP32-mixed form with resistivity (FP32 stencils, eta terms
    included) against the reference's FP64 states after 5 steps; dt, which
    the diffusion limit fixes in eta_diff, to FP32 accuracy"""
    print(f"case: FP32-mixed resistive steps, {case}", flush=True)
    m = MhdGrid(config.grid_desc(GSIZE, DX), eta_phys(float(gold[f"{case}_eta"])), abi.PREC_FP32_MIXED)
    m.eos_init(table)
    r = config.run_desc(CFL=0.8)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.step(r, 1e30, 5)
    U = m.get_state()
    inner = (slice(None),) + (slice(2, -2),) * 3
    errs = [relerr(U[inner][v], gold[f"{case}_step5_U"][inner][v]) for v in range(8)]
    print("per-field rel err", ["%.2e" % e for e in errs], "dt", r.dt, gold[f"{case}_step5_sc"][0])
    assert max(errs) <= 1e-5
    assert r.dt == pytest.approx(gold[f"{case}_step5_sc"][0], rel=1e-5)
