"""The damping (dmp_switch; get_damping, Add_Sources_Integrate_SR.C:92-108,
628-761) on the GPU:

* against the compiled reference's outputs (tests/golden/ref_damp.npz):
  Source_Integrate_Tcheck at stage 1 through the drop-in operations, and 1
  and 5 full steps on the fused and the drop-in sequence paths, within 1e-12
  per field, dt equal;
* against the oracle on a 40 x 36 x 28 box with a mean vertical flow, three
  steps on both paths: decisive limiter choices equal, fields within 1e-12;
* an FP32-mixed context refuses the damping.
"""
import os

import numpy as np
import pytest

from muram_main_amd import abi, config, eos_table, initial
from muram_main_amd.mhd import MhdGrid, MrmError
from orc import relerr
from gpu_pairs import check_pair, env, start_gpu, start_oracle
from damp_cases import damp_phys, upflow_state

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12
GSIZE, DX = (12, 10, 8), (3.2e6, 3.2e6, 4.8e6)


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "ref_damp.npz"))


def make(table, phys, path="fused"):
    with env(MRM_STEP_PATH=path):
        m = MhdGrid(config.grid_desc(GSIZE, DX), phys)
    m.eos_init(table)
    return m


def maxrel(a, b):
    return max(relerr(a[v], b[v]) for v in range(a.shape[0]))


@pytest.mark.parametrize("case", ["damp", "damp_weak"])
def test_damped_integration_matches_reference(table, gold, case):
    print(f"case: damped Source_Integrate_Tcheck, {case}", flush=True)
    m = make(table, damp_phys(float(gold[f"{case}_vel_ref"])))
    r = config.run_desc(CFL=0.8)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.SaveCons()
    m.ConsToPrim()
    m.Adjust_Valf_Max(r)
    dt = m.MHD_Residual(r)
    m.Sync_timestep(r, 1e30, dt)
    m.Source_Integrate_Tcheck(r, 1)
    e = maxrel(m.get_state(), gold[f"{case}_srcint_U"])
    print("state rel err", e, "dt", r.dt, gold[f"{case}_srcint_sc"][0], flush=True)
    assert e <= TOL
    assert r.dt == pytest.approx(gold[f"{case}_srcint_sc"][0], rel=TOL)


@pytest.mark.parametrize("path", ["fused", "sequence"])
@pytest.mark.parametrize("case", ["damp", "damp_weak"])
def test_damped_steps_match_reference(table, gold, case, path):
    print(f"case: damped steps, {case}, {path}", flush=True)
    m = make(table, damp_phys(float(gold[f"{case}_vel_ref"])), path)
    r = config.run_desc(CFL=0.8)
    m.set_state(gold[f"{case}_in"])
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.step(r, 1e30, 1)
    e1 = maxrel(m.get_state(), gold[f"{case}_step1_U"])
    m.step(r, 1e30, 4)
    e5 = maxrel(m.get_state(), gold[f"{case}_step5_U"])
    print("1 step", e1, "5 steps", e5, "dt", r.dt, gold[f"{case}_step5_sc"][0], flush=True)
    assert e1 <= TOL and e5 <= TOL
    assert r.dt == pytest.approx(gold[f"{case}_step5_sc"][0], rel=TOL)


@pytest.mark.parametrize("path", ["fused", "sequence"])
def test_damped_box_matches_oracle(table, path):
    print(f"case: damped 40x36x28 box, three steps, {path}", flush=True)
    box = initial.BoxSpec(gsize=(40, 36, 28), dx=(3.2e6,) * 3, seed=9, v_rms=2.0e5)
    U = upflow_state(eos_table.load_tables(table), box)
    phys = damp_phys()
    o, ro = start_oracle(table, box, phys, U)
    m, r = start_gpu(table, box, phys, U, path=path)
    for s in range(3):
        o.step(ro, config.DT_RAD_OFF, 1)
        m.step(r, config.DT_RAD_OFF, 1)
        print(f"step {s + 1}: dt gpu {r.dt:.15e} oracle {ro.dt:.15e}", flush=True)
    check_pair(m, r, o, ro)


def test_mixed_refuses_damping():
    with pytest.raises(MrmError):
        MhdGrid(config.grid_desc(GSIZE, DX), damp_phys(), abi.PREC_FP32_MIXED)
