"""CPU tests of the oracle's damping (get_damping and its use in
Source_Integrate_Tcheck, Add_Sources_Integrate_SR.C:92-108, 164, 340,
628-761).

1. Golden vectors from the compiled reference (tests/golden/ref_damp.npz,
   tests/golden/make_golden_damp.py): Source_Integrate_Tcheck at stage 1 and
   1 and 5 full steps, bit for bit, with layers on the 1/tau_min cap and on
   the power branch.
2. What the definition fixes, independently of the reference: with no
   vertical flow the damping is exactly zero and the step is bit for bit the
   undamped one (the next one is not); with a weak field (Boris factor ~1e-14) the stage-1 update
   differs from the undamped one by a force uniform over each layer that
   opposes the layer's mean mass flux <rho v_x>, equal to <rho v_x>/tau_min
   in the layers on the cap, with the work -dmp v_x in the energy.
"""
import ctypes as C
import os

import numpy as np
import pytest

from muram_main_amd import config, eos_table, initial
from orc import Oracle
from damp_cases import TAU_MIN, TAU_REF, damp_phys, upflow_state

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
GSIZE = (12, 10, 8)
DX = (3.2e6, 3.2e6, 4.8e6)


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(GOLD, "eos_small.dat"), "rb").read()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "ref_damp.npz"))


def start(table, phys, U):
    o = Oracle(config.grid_desc(GSIZE, DX), phys, table)
    r = config.run_desc(CFL=0.8)
    o.set_state(U)
    o.exchange_grid()
    o.set_boundary_conditions(r, 0)
    return o, r


def stage1(o, r):
    """SaveCons, ConsToPrim, Adjust_Valf_Max, MHD_Residual, Sync_timestep and
    Source_Integrate_Tcheck(1); returns the primitive state the integration
    saw (velocities in slots 1-3)"""
    o.save_cons()
    o.cons_to_prim()
    o.adjust_valf_max(r, None)
    dt = o.mhd_residual(r)
    cfl = C.c_double()
    assert o.sync_timestep(r, 1e30, dt, C.byref(cfl)) == 0
    prim = o.state(0).copy()
    o.source_integrate_tcheck(r, 1)
    return prim


@pytest.mark.parametrize("case", ["damp", "damp_weak"])
def test_golden_damped_integration_and_steps(gold, table, case):
    p = damp_phys(float(gold[f"{case}_vel_ref"]))
    o, r = start(table, p, gold[f"{case}_in"])
    stage1(o, r)
    assert r.dt == gold[f"{case}_srcint_sc"][0]
    assert np.array_equal(o.state(0), gold[f"{case}_srcint_U"])
    o, r = start(table, p, gold[f"{case}_in"])
    o.step(r, 1e30, 1)
    assert np.array_equal(o.state(0), gold[f"{case}_step1_U"])
    o.step(r, 1e30, 4)
    assert np.array_equal(o.state(0), gold[f"{case}_step5_U"])
    assert r.dt == gold[f"{case}_step5_sc"][0]


def test_no_vertical_flow_no_damping(table):
    tab = eos_table.load_tables(table)
    U = initial.make_state(tab, initial.BoxSpec(gsize=GSIZE, dx=DX, seed=5))
    U[4] -= 0.5 * U[1] ** 2 / U[0]
    U[1] = 0.0
    a, ra = start(table, config.stratified_phys(), U)
    b, rb = start(table, damp_phys(), U)
    # the profile of stage 1 serves the whole step
    a.step(ra, 1e30, 1)
    b.step(rb, 1e30, 1)
    assert np.array_equal(a.state(0), b.state(0))
    # the next step starts from a flow the first one made, and damps it
    a.step(ra, 1e30, 1)
    b.step(rb, 1e30, 1)
    assert not np.array_equal(a.state(0), b.state(0))


def test_damping_force_opposes_mean_flux(table):
    tab = eos_table.load_tables(table)
    U = upflow_state(tab, initial.BoxSpec(gsize=GSIZE, dx=DX, seed=6, v_rms=2.0e5))
    base = config.stratified_phys(va_max=1.0e12)      # Boris factor ~1e-14
    a, ra = start(table, base, U)
    b, rb = start(table, damp_phys(base=config.stratified_phys(va_max=1.0e12)), U)
    prim = stage1(a, ra)
    stage1(b, rb)
    assert ra.dt == rb.dt
    wdt = ra.dt / 4.0
    inner = (slice(2, -2), slice(2, -2), slice(2, -2))
    dM = (b.state(0)[1] - a.state(0)[1])[inner] / wdt     # = -dmp(i)
    de = (b.state(0)[4] - a.state(0)[4])[inner] / wdt     # = -dmp(i) v_x
    rho, vx = prim[0][inner], prim[1][inner]
    flux = np.mean(rho * vx, axis=(0, 1))                  # <rho v_x> per layer
    rho_max = rho.max()
    force = dM.mean(axis=(0, 1))
    scale = np.abs(force).max()
    assert scale > 0.0
    # uniform over each layer, opposite to the mean flux
    assert np.abs(dM - force).max() <= 1e-6 * scale
    assert np.all(force * flux < 0.0)
    # the cap: rate 1/tau_min where (flux / (vel_ref rho_max))^4 / tau_ref
    # exceeds it; below it the rate is that power
    rate = -force / flux
    capped = (1.0 / TAU_REF) * (flux / (1.0e4 * rho_max)) ** 4 >= 1.0 / TAU_MIN
    assert capped.any() and (~capped).any()
    np.testing.assert_allclose(rate[capped], 1.0 / TAU_MIN, rtol=1e-6)
    assert np.all(rate[~capped] < 1.0 / TAU_MIN)
    # the energy takes the work of the force
    np.testing.assert_allclose(de, dM * vx, rtol=1e-5, atol=1e-6 * np.abs(dM * vx).max())
