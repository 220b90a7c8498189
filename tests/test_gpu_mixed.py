"""FP32-mixed step (MRM_PREC_FP32_MIXED) against the FP64 oracle: rho, M, B
and p in FP32, e in FP64.  Per-field relative L2 / Linf errors after several
steps, dt agreement, the split-TVD branch, Qtot/Tau, the state conversion, and
the per-operation calls that an FP32-mixed context refuses.
"""
import os

import numpy as np
import pytest

from muram_main_amd import abi, config, eos_table, initial
from muram_main_amd.mhd import MhdGrid, MrmError
from orc import Oracle, field_errors

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
# after the steps of each case: relative L2 and Linf error per field
TOL_L2 = 2e-6
TOL_LINF = 2e-5
TOL_DT = 1e-6


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


def setup_pair(table, gsize, seed, v_rms=2.0e5, cfl=0.8, fields=None, phys=None):
    tab = eos_table.load_tables(table)
    box = initial.BoxSpec(gsize=gsize, dx=(3.2e6,) * 3, seed=seed, v_rms=v_rms)
    U = initial.make_state(tab, box)
    g = config.grid_desc(box.gsize, box.dx)
    p = phys or config.stratified_phys()
    o = Oracle(g, p, table)
    ro = config.run_desc(CFL=cfl)
    o.set_state(U)
    m = MhdGrid(g, p, precision=abi.PREC_FP32_MIXED)
    m.eos_init(table)
    r = config.run_desc(CFL=cfl)
    m.set_state(U)
    for f, arr in (fields or {}).items():
        o.field(f)[...] = arr
        m.set_field(f, arr)
    o.exchange_grid()
    o.set_boundary_conditions(ro, 0)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    return o, ro, m, r


def check(o, ro, m, r, tol_l2=TOL_L2, tol_linf=TOL_LINF):
    errs = field_errors(m.get_state(), o.state(0))
    print("dt", r.dt, ro.dt)
    for k, v in errs.items():
        print(f"  {k:4s} L2 {v['l2']:.2e}  Linf {v['linf']:.2e}")
    assert abs(r.dt - ro.dt) <= TOL_DT * ro.dt
    for k, v in errs.items():
        assert v["l2"] <= tol_l2, (k, v)
        assert v["linf"] <= tol_linf, (k, v)


def test_mixed_state_conversion(table):
    print("case: mixed set/get state", flush=True)
    o, ro, m, r = setup_pair(table, (20, 16, 12), seed=41)
    U = o.state(0).copy()
    m.set_state(U)
    V = m.get_state()
    for v in (0, 1, 2, 3, 5, 6, 7):
        np.testing.assert_array_equal(V[v], U[v].astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(V[4], U[4])


@pytest.mark.parametrize("nsteps", [1, 5])
def test_mixed_steps_match_oracle(table, nsteps):
    print(f"case: mixed 40x36x28, {nsteps} steps", flush=True)
    o, ro, m, r = setup_pair(table, (40, 36, 28), seed=42)
    o.step(ro, 1e30, nsteps)
    m.step(r, 1e30, nsteps)
    check(o, ro, m, r)


def test_mixed_va_adjust(table):
    print("case: mixed 40x36x28, 5 steps, va_adjust 1", flush=True)
    p = config.stratified_phys(va_max=3.0e7, va_adjust=1.0)
    o, ro, m, r = setup_pair(table, (40, 36, 28), seed=45, phys=p)
    o.step(ro, 1e30, 5)
    m.step(r, 1e30, 5)
    check(o, ro, m, r)
    print("limits", m.limits(), o.limits())
    np.testing.assert_allclose(m.limits(), o.limits(), rtol=TOL_DT)


def test_mixed_split_tvd(table):
    print("case: mixed split TVD", flush=True)
    o, ro, m, r = setup_pair(table, (24, 20, 18), seed=43, v_rms=3.0e5, cfl=0.99)
    o.step(ro, 1e30, 3)
    m.step(r, 1e30, 3)
    check(o, ro, m, r)


def test_mixed_with_qtot_and_tau(table):
    print("case: mixed with Qtot, Tau", flush=True)
    rng = np.random.default_rng(6)
    shape = tuple(s + 4 for s in (20, 18, 16))[::-1]
    q = -1e-3 * np.abs(rng.standard_normal(shape)) * 1e6
    tau = np.where(rng.random(shape) < 0.5, 1.0, 0.0)
    o, ro, m, r = setup_pair(table, (20, 18, 16), seed=44, fields={abi.QTOT: q, abi.TAU: tau})
    o.step(ro, 1e30, 2)
    m.step(r, 1e30, 2)
    check(o, ro, m, r)


def test_mixed_pressure_is_eos_of_state(table):
    print("case: mixed carried pressure", flush=True)
    o, ro, m, r = setup_pair(table, (20, 16, 16), seed=45)
    m.step(r, 1e30, 2)
    p_mx = m.get_field(abi.PRES)
    o.set_state(m.get_state())
    o.cons_to_prim()
    p64 = o.field(abi.PRES)
    # interior cells: FP32 table interpolation against the FP64 one
    sl = (slice(2, -2), slice(2, -2), slice(2, -2))
    rel = np.max(np.abs(p_mx[sl] - p64[sl]) / np.abs(p64[sl]))
    print("pressure max rel diff", rel)
    assert rel <= 1e-5


def test_mixed_refuses_per_operation_calls(table):
    print("case: mixed refuses drop-in calls", flush=True)
    o, ro, m, r = setup_pair(table, (12, 8, 8), seed=46)
    with pytest.raises(MrmError, match="UNSUPPORTED|whole steps"):
        m.ConsToPrim()
    with pytest.raises(MrmError):
        m.TVDlimit(r, r.dt)
    with pytest.raises(MrmError):
        m.get_state(abi.U0)
