"""Clean_divB on the GPU (mrm_clean_divB, and inside mrm_step between
TVDlimit and TCheck) against the compiled reference's outputs in
tests/golden/ref_divb.npz: single calls bit for bit in the interior (the
cleaning is additions and products of B and phi only), full steps within
1e-12 per field."""
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
INNER = (slice(2, -2),) * 3


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "ref_divb.npz"))


def divb_phys(va_max=3.0e8, va_adjust=0.0):
    p = config.stratified_phys(va_max=va_max, va_adjust=va_adjust)
    p.divB[:] = [1, 5, 0.2]
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


def test_clean_divb_calls_match_reference(table, gold):
    print("case: Clean_divB, four calls, 12x10x8", flush=True)
    m = make(table, divb_phys())
    r = config.run_desc()
    m.set_state(gold["op_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    for n in range(4):
        err = m.Clean_divB(r)
        print(n, "err", err)
        assert err[1] < err[0]
        divb = m.get_field(abi.DIVB)
        phi = m.get_field(abi.PHI)
        assert np.array_equal(divb[INNER], gold[f"op{n}_divB"][INNER]), n
        assert np.array_equal(phi, gold[f"op{n}_phi"]), n
        if f"op{n}_U" in gold.files:
            U = m.get_state()
            assert np.array_equal(U[(slice(None),) + INNER], gold[f"op{n}_U"][(slice(None),) + INNER]), n
            assert max(relerr(U[v], gold[f"op{n}_U"][v]) for v in range(8)) <= TOL


@pytest.mark.parametrize("path", ["fused", "sequence"])
@pytest.mark.parametrize("case,va_max,vadj", [("steps", 3.0e7, 0.0), ("steps_vadj", 1.0e6, 1.0)])
def test_steps_with_clean_divb_match_reference(table, gold, path, case, va_max, vadj):
    print(f"case: steps with Clean_divB, {case}, {path}", flush=True)
    m = make(table, divb_phys(va_max, vadj), path)
    r = config.run_desc(CFL=0.8)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.step(r, 1e30, 1)
    e1 = max(relerr(m.get_state()[v], gold[f"{case}_step1_U"][v]) for v in range(8))
    assert r.dt == pytest.approx(gold[f"{case}_step1_sc"][0], rel=TOL)
    m.step(r, 1e30, 4)
    U = m.get_state()
    e5 = max(relerr(U[v], gold[f"{case}_step5_U"][v]) for v in range(8))
    ephi = relerr(m.get_field(abi.PHI), gold[f"{case}_step5_phi"])
    print("step1", e1, "step5", e5, "phi", ephi, "dt", r.dt)
    assert e1 <= TOL and e5 <= TOL and ephi <= TOL
    assert r.time == pytest.approx(gold[f"{case}_step5_sc"][1], rel=TOL)
    assert m.limits()[0] == pytest.approx(gold[f"{case}_step5_sc"][6], rel=TOL)


@pytest.mark.parametrize("case,va_max,vadj", [("steps", 3.0e7, 0.0), ("steps_vadj", 1.0e6, 1.0)])
def test_mixed_steps_with_clean_divb(table, gold, case, va_max, vadj):
    """the FP32-mixed form (FP32 B, FP64 phi, divergence and corrections)
    against the reference's FP64 states and potential after 5 steps"""
    print(f"case: FP32-mixed steps with Clean_divB, {case}", flush=True)
    m = MhdGrid(config.grid_desc(GSIZE, DX), divb_phys(va_max, vadj), abi.PREC_FP32_MIXED)
    m.eos_init(table)
    r = config.run_desc(CFL=0.8)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.step(r, 1e30, 5)
    U = m.get_state()
    errs = [relerr(U[v][INNER], gold[f"{case}_step5_U"][v][INNER]) for v in range(8)]
    ephi = relerr(m.get_field(abi.PHI)[INNER], gold[f"{case}_step5_phi"][INNER])
    print("per-field rel err", ["%.2e" % e for e in errs], "phi", "%.2e" % ephi, "dt", r.dt)
    assert max(errs) <= 1e-5 and ephi <= 1e-4
    assert r.time == pytest.approx(gold[f"{case}_step5_sc"][1], rel=1e-5)


def test_clean_divb_refused_where_unsupported(table):
    p = divb_phys()
    p.divB[1] = 0   # divB_itmax < 1
    for prec in (abi.PREC_FP64, abi.PREC_FP32_MIXED):
        with pytest.raises(abi.MrmError):
            MhdGrid(config.grid_desc(GSIZE, DX), p, prec)
