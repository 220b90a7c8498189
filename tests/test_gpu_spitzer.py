"""Spitzer heat conduction (param_spitzer > 0) on the GPU against the
compiled reference's outputs in tests/golden/ref_spitzer.npz: the state and
the heat flux sflx after 1 and 5 steps from sflx = 0, within 1e-12 per field,
on the fused step and on the per-operation sequence."""
import os
import sys

import numpy as np
import pytest

from muram_main_amd import abi, config
from muram_main_amd.mhd import MhdGrid
from orc import relerr

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_golden_spitzer  # noqa: E402


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "ref_spitzer.npz"))


def maxrel(a, b):
    return max(relerr(a[v], b[v]) for v in range(a.shape[0]))


def make(box, phys, path):
    saved = os.environ.get("MRM_STEP_PATH")
    os.environ["MRM_STEP_PATH"] = path
    try:
        return MhdGrid(config.grid_desc(box.gsize, box.dx), phys)
    finally:
        if saved is None:
            os.environ.pop("MRM_STEP_PATH", None)
        else:
            os.environ["MRM_STEP_PATH"] = saved


@pytest.mark.parametrize("path", ["fused", "sequence"])
@pytest.mark.parametrize("case", ["spt_cool", "spt_corona"])
def test_spitzer_steps_match_reference(table, gold, case, path):
    print(f"case: Spitzer conduction, {case}, {path} step", flush=True)
    box, phys = make_golden_spitzer.cases()[case]
    m = make(box, phys, path)
    m.eos_init(table)
    r = config.run_desc(CFL=0.8)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.step(r, 1e30, 1)
    e1 = maxrel(m.get_state(), gold[f"{case}_step1_U"])
    s1 = relerr(m.get_field(abi.SFLX), gold[f"{case}_step1_sflx"])
    assert r.dt == pytest.approx(gold[f"{case}_step1_sc"][0], rel=TOL)
    m.step(r, 1e30, 4)
    e5 = maxrel(m.get_state(), gold[f"{case}_step5_U"])
    s5 = relerr(m.get_field(abi.SFLX), gold[f"{case}_step5_sflx"])
    print("U step1", e1, "step5", e5, "sflx step1", s1, "step5", s5, "dt", r.dt)
    assert max(e1, e5, s1, s5) <= TOL
    assert r.dt == pytest.approx(gold[f"{case}_step5_sc"][0], rel=TOL)


@pytest.mark.parametrize("case", ["spt_cool", "spt_corona"])
def test_mixed_spitzer_steps(table, gold, case):
    """the FP32-mixed form with conduction (FP32 temperature lookup, FP64
    heat flux) against the reference's FP64 state and heat flux after 5 steps"""
    print(f"case: FP32-mixed Spitzer conduction, {case}", flush=True)
    box, phys = make_golden_spitzer.cases()[case]
    m = MhdGrid(config.grid_desc(box.gsize, box.dx), phys, abi.PREC_FP32_MIXED)
    m.eos_init(table)
    r = config.run_desc(CFL=0.8)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.step(r, 1e30, 5)
    inner = (slice(2, -2),) * 3
    U = m.get_state()
    errs = [relerr(U[v][inner], gold[f"{case}_step5_U"][v][inner]) for v in range(8)]
    es = relerr(m.get_field(abi.SFLX)[inner], gold[f"{case}_step5_sflx"][inner])
    print("per-field rel err", ["%.2e" % e for e in errs], "sflx", "%.2e" % es, "dt", r.dt)
    assert max(errs) <= 1e-5 and es <= 1e-4
    assert r.dt == pytest.approx(gold[f"{case}_step5_sc"][0], rel=1e-5)
