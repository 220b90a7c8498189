"""GPU parity of the fused, device-resident step (mrm_step's default path)
against the oracle: 1e-12 relative per field after several steps, equal dt.
Covers both stage-kernel variants, grids that are not multiples of the tile,
the split TVD branch, non-zero Qtot/Tau, and interleaving with drop-in calls.
"""
import os

import numpy as np
import pytest

from muram_main_amd import abi, config, eos_table, initial
from muram_main_amd.mhd import MhdGrid
from orc import Oracle, relerr

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


def setup_pair(table, gsize, seed, v_rms=2.0e5, cfl=0.8, phys=None, fields=None, zwin="1", fma="1"):
    tab = eos_table.load_tables(table)
    box = initial.BoxSpec(gsize=gsize, dx=(3.2e6,) * 3, seed=seed, v_rms=v_rms)
    U = initial.make_state(tab, box)
    g = config.grid_desc(box.gsize, box.dx)
    p = phys or config.stratified_phys()
    o = Oracle(g, p, table)
    ro = config.run_desc(CFL=cfl)
    o.set_state(U)
    saved = {k: os.environ.get(k) for k in ("MRM_STAGE_ZWIN", "MRM_FMA")}
    os.environ["MRM_STAGE_ZWIN"] = zwin
    os.environ["MRM_FMA"] = fma
    try:
        m = MhdGrid(g, p)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    m.eos_init(table)
    r = config.run_desc(CFL=cfl)
    m.set_state(U)
    for f, arr in (fields or {}).items():
        o.field(f)[...] = arr
        m.set_field(f, arr)
    for x, rr in ((o, ro), (m, r)):
        if x is o:
            o.exchange_grid()
            o.set_boundary_conditions(ro, 0)
        else:
            m.exchange_grid()
            m.SetBoundaryConditions(r, 0)
    return o, ro, m, r


def compare(o, ro, m, r):
    errs = [relerr(m.get_state()[v], o.state(0)[v]) for v in range(8)]
    print("per-field rel err:", ["%.2e" % e for e in errs], "dt", r.dt, ro.dt)
    assert max(errs) <= TOL
    assert r.dt == pytest.approx(ro.dt, rel=TOL)
    assert r.time == pytest.approx(ro.time, rel=TOL)


@pytest.mark.parametrize("zwin,fma", [("0", "0"), ("1", "0"), ("2", "0"), ("3", "0"), ("4", "0"), ("0", "1"), ("1", "1"),
                                      ("2", "1"), ("3", "1"), ("4", "1")])
def test_fused_steps_match_oracle(table, zwin, fma):
    print(f"case: fused 40x36x28 zwin={zwin} fma={fma}", flush=True)
    o, ro, m, r = setup_pair(table, (40, 36, 28), seed=31, zwin=zwin, fma=fma)
    o.step(ro, 1e30, 3)
    m.step(r, 1e30, 3)
    compare(o, ro, m, r)


@pytest.mark.parametrize("fma", ["0", "1"])
def test_stage_wave_groups_bit_identical(table, fma):
    """the three-group stage kernel (the EOS pressure in its own wave group)
    and the two-group one perform the same operations: equal states and
    pressures bit for bit after three steps"""
    print(f"case: fused 40x36x28, stage kernel with three vs two wave groups, fma={fma}", flush=True)
    out = []
    for zwin in ("3", "4"):
        o, ro, m, r = setup_pair(table, (40, 36, 28), seed=33, zwin=zwin, fma=fma)
        m.step(r, 1e30, 3)
        out.append((m.get_state(), m.get_field(abi.PRES), r.dt))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_fused_split_tvd(table):
    print("case: fused split TVD", flush=True)
    o, ro, m, r = setup_pair(table, (24, 20, 18), seed=32, v_rms=3.0e5, cfl=0.99)
    o.step(ro, 1e30, 3)
    m.step(r, 1e30, 3)
    compare(o, ro, m, r)


def test_fused_with_qtot_and_tau(table):
    print("case: fused with Qtot, Tau", flush=True)
    rng = np.random.default_rng(5)
    shape = tuple(s + 4 for s in (20, 18, 16))[::-1]
    U_e = None
    q = -1e-3 * np.abs(rng.standard_normal(shape)) * 1e6
    tau = np.where(rng.random(shape) < 0.5, 1.0, 0.0)
    o, ro, m, r = setup_pair(table, (20, 18, 16), seed=33, fields={abi.QTOT: q, abi.TAU: tau})
    o.step(ro, 1e30, 2)
    m.step(r, 1e30, 2)
    compare(o, ro, m, r)


def test_fused_interleaved_with_drop_in(table):
    print("case: fused interleaved with drop-in", flush=True)
    o, ro, m, r = setup_pair(table, (20, 16, 20), seed=34)
    o.step(ro, 1e30, 1)
    m.step(r, 1e30, 1)
    # a drop-in TCheck between steps invalidates the carried pressure/cmax
    o.tcheck(ro)
    m.TCheck(r)
    o.step(ro, 1e30, 2)
    m.step(r, 1e30, 2)
    compare(o, ro, m, r)


@pytest.mark.parametrize("fma,tol", [("0", 0.0), ("1", 1e-13)])
def test_fused_pressure_is_eos_of_new_state(table, fma, tol):
    o, ro, m, r = setup_pair(table, (20, 16, 16), seed=35, fma=fma)
    m.step(r, 1e30, 2)
    p_fused = m.get_field(abi.PRES)
    U = m.get_state()
    m.ConsToPrim()
    p_ref = m.get_field(abi.PRES)
    assert relerr(p_fused, p_ref) <= tol
    m.set_state(U)


def full_phys():
    p = config.stratified_phys(va_max=1.0e8, va_adjust=1.0)
    p.divB[:] = [1, 5, 0.2]
    p.params[config.I_PARAM_ETA] = 1e10
    p.params[config.I_PARAM_SPITZER] = 1e-6
    return p


@pytest.mark.parametrize("gsize", [(13, 7, 5), (33, 17, 37)])
@pytest.mark.parametrize("path", ["fused", "sequence", "mixed"])
def test_odd_sizes_every_option(table, gsize, path):
    """grids below one tile and with ragged tiles in every kernel (stage
    32x8 with 64-plane FP64 and 32-plane mixed chunks, Clean_divB and
    conduction 64x4), with va_adjust,
    Clean_divB, resistivity and Spitzer conduction on together: the FP64
    paths to 1e-12 of the oracle, the FP32-mixed form to 1e-5"""
    print(f"case: odd sizes {gsize}, every physics option, {path}", flush=True)
    saved = os.environ.get("MRM_STEP_PATH")
    os.environ["MRM_STEP_PATH"] = "sequence" if path == "sequence" else "fused"
    try:
        tab = eos_table.load_tables(table)
        box = initial.BoxSpec(gsize=gsize, dx=(3.2e6, 3.2e6, 4.8e6), seed=37, v_rms=2.0e5)
        U = initial.make_state(tab, box)
        g = config.grid_desc(box.gsize, box.dx)
        m = MhdGrid(g, full_phys(), abi.PREC_FP32_MIXED if path == "mixed" else abi.PREC_FP64)
    finally:
        if saved is None:
            os.environ.pop("MRM_STEP_PATH", None)
        else:
            os.environ["MRM_STEP_PATH"] = saved
    m.eos_init(table)
    o = Oracle(g, full_phys(), table)
    r, ro = config.run_desc(CFL=0.8), config.run_desc(CFL=0.8)
    o.set_state(U)
    m.set_state(U)
    o.exchange_grid()
    o.set_boundary_conditions(ro, 0)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    o.step(ro, 1e30, 3)
    m.step(r, 1e30, 3)
    inner = (slice(2, -2),) * 3
    Ug = m.get_state()
    errs = [relerr(Ug[v][inner], o.state(0)[v][inner]) for v in range(8)]
    errs.append(relerr(m.get_field(abi.SFLX)[inner], o.field(abi.SFLX)[inner]))
    print("per-field rel err (rho, M, e, B, sflx):", ["%.2e" % e for e in errs], "dt", r.dt, ro.dt)
    tol = 1e-5 if path == "mixed" else TOL
    assert max(errs[:8]) <= tol and errs[8] <= (1e-4 if path == "mixed" else TOL)
    assert r.dt == pytest.approx(ro.dt, rel=tol)
