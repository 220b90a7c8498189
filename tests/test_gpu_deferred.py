"""The drop-in path's deferred stage on the GPU: SaveCons, ConsToPrim and
MHD_Residual are recorded, and Source_Integrate_Tcheck runs the stage kernel
on the conservative state, when no other call comes between them; any other
call carries the recorded operations out first (MRM_DEFER=0 runs every
operation as it is called).

* the compiled reference's stage (tests/golden/ref_small.npz): SaveCons,
  ConsToPrim, Adjust_Valf_Max, MHD_Residual, Sync_timestep,
  Source_Integrate_Tcheck with no read in between, then the state (ghost
  cells in ConsToPrim's primitive form, as the in-place update leaves them),
  and after exchange and boundary, within 1e-12 of the reference's;
* reads in the middle: U after a recorded ConsToPrim is the primitive state,
  the EOS fields and Res after a recorded MHD_Residual are ConsToPrim's and
  the residual, U0 after a recorded SaveCons is U; ConsToPrim's five EOS
  outputs read after the deferred stage, after exchange and boundary, and
  after stage 2 (a recorded MHD_Residual forms the pressure alone, the
  other four are formed from the recorded state when read), and a
  temperature written by the caller in between; the stage output's ghost
  cells (recorded: the exchange and boundaries overwrite them) read right
  after the stage, after the exchange alone, with the boundaries before the
  exchange and after a ConsToPrim with no boundaries before it; TVDlimit's
  recorded dim-2
  sweep run by TCheck, or plain before a read or a change of the limits,
  whole and split -- each bit for bit the calls with MRM_DEFER=0;
* three whole steps on the drop-in sequence (mrm_step with MRM_STEP_PATH=
  sequence), deferred and not, plain, with va_adjust 1 (whose stage 1 runs
  ConsToPrim for Adjust_Valf_Max's statistics) and with resistivity, against
  the oracle: decisive limiter choices equal, fields within 1e-12, and the
  last ConsToPrim's five EOS outputs within 1e-12.
"""
import os

import numpy as np
import pytest

from muram_main_amd import abi, config, eos_table, initial
from muram_main_amd.mhd import MhdGrid
from orc import relerr
from gpu_pairs import check_pair, env, start_gpu, start_oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(HERE, "golden", "ref_small.npz"))


def small_grid():
    return config.grid_desc((10, 8, 6), (3.2e6,) * 3)


def make(table, defer="1", phys=None, grid=None):
    with env(MRM_DEFER=defer):
        m = MhdGrid(grid or small_grid(), phys or config.stratified_phys())
    m.eos_init(table)
    return m


def maxrel(a, b):
    return max(relerr(a[v], b[v]) for v in range(a.shape[0]))


def start(m, U):
    r = config.run_desc(CFL=0.8)
    m.set_state(U)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    return r


def test_deferred_stage_matches_reference(gold, table):
    print("case: deferred stage 1, 10x8x6, against the reference", flush=True)
    m = make(table)
    r = start(m, gold["base_in"])
    m.SaveCons()
    m.ConsToPrim()
    m.Adjust_Valf_Max(r)
    dt_mhd = m.MHD_Residual(r)
    assert dt_mhd == pytest.approx(gold["base_resid_sc"][2], rel=TOL)
    m.Sync_timestep(r, 1e30, dt_mhd)
    m.Source_Integrate_Tcheck(r, 1)
    e = maxrel(m.get_state(), gold["base_srcint_U"])
    print("after Source_Integrate_Tcheck:", e, flush=True)
    assert e <= TOL
    assert np.all(m.get_state(abi.RES)[:, 2:-2, 2:-2, 2:-2] == 0.0)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 1)
    e = maxrel(m.get_state(), gold["base_bc1_U"])
    print("after exchange and boundary:", e, flush=True)
    assert e <= TOL


def _ops(table, gold, defer, read):
    """stage 1 by the drop-in calls, with one read at the point `read`;
    returns what was read and the state at the end"""
    m = make(table, defer)
    r = start(m, gold["base_in"])
    got = {}
    m.SaveCons()
    if read == "u0":
        got["u0"] = m.get_state(abi.U0)
    m.ConsToPrim()
    if read == "prim":
        got["u"] = m.get_state()
    m.Adjust_Valf_Max(r)
    dt = m.MHD_Residual(r)
    if read == "res":
        got["res"] = m.get_state(abi.RES)
        got["f"] = np.stack([m.get_field(f) for f in range(5)])
    m.Sync_timestep(r, 1e30, dt)
    m.Source_Integrate_Tcheck(r, 1)
    if read == "fields_srcint":
        got["f"] = np.stack([m.get_field(f) for f in range(5)])
    # the ghost cells of the stage's output: recorded, read with every ghost
    # cell pending (u_srcint), with the dim-0 layers pending (u_exch), taken
    # over by boundaries before the exchange (bc_first), the dim-0 layers
    # converted by a ConsToPrim with no boundaries in between (skip_bc: the
    # primitive state; the EOS of such twice-converted ghost cells is NaN)
    if read == "u_srcint":
        got["u"] = m.get_state()
    if read == "bc_first":
        m.SetBoundaryConditions(r, 1)
    m.exchange_grid()
    if read == "u_exch":
        got["u"] = m.get_state()
    if read == "skip_bc":
        m.ConsToPrim()
        got["u"] = m.get_state()
    else:
        m.SetBoundaryConditions(r, 1)
    if read == "fields_bc":
        got["f"] = np.stack([m.get_field(f) for f in range(5)])
    if read in ("fields_stage2", "temp_set"):
        # stage 2 up to Source_Integrate_Tcheck, then the fields
        m.ConsToPrim()
        m.Adjust_Valf_Max(r)
        m.MHD_Residual(r)
        m.Source_Integrate_Tcheck(r, 2)
        if read == "temp_set":
            m.set_field(abi.TEMP, np.full(m.get_field(abi.PRES).shape, 1234.5))
        got["f"] = np.stack([m.get_field(f) for f in range(5)])
    return got, m.get_state(), m.get_state(abi.U0), r.dt


@pytest.mark.parametrize("read", ["u0", "prim", "res", "fields_srcint", "fields_bc", "fields_stage2", "temp_set",
                                  "u_srcint", "u_exch", "bc_first", "skip_bc"])
def test_reads_carry_out_recorded_operations(gold, table, read):
    print(f"case: read {read} in the middle of a deferred stage", flush=True)
    g1, U1, U01, dt1 = _ops(table, gold, "1", read)
    g0, U0, U00, dt0 = _ops(table, gold, "0", read)
    for k in g0:
        assert np.array_equal(g1[k], g0[k]), k
    assert dt1 == dt0
    # the read made the stage run as called: the end state is bit for bit
    assert np.array_equal(U1, U0)
    assert np.array_equal(U01, U00)


def _tvd_ops(table, gold, defer, read, split):
    """a whole stage by the drop-in calls, then TVDlimit (whole, or two
    halves with an exchange between) and TCheck, with one read at the point
    `read`; returns what was read and the state at the end"""
    m = make(table, defer)
    r = start(m, gold["base_in"])
    m.SaveCons()
    m.ConsToPrim()
    m.Adjust_Valf_Max(r)
    dt = m.MHD_Residual(r)
    m.Sync_timestep(r, 1e30, dt)
    m.Source_Integrate_Tcheck(r, 1)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 1)
    got = {}
    if split:
        m.TVDlimit(r, 0.5 * r.dt)
        m.exchange_grid()
        m.TVDlimit(r, 0.5 * r.dt)
    else:
        m.TVDlimit(r, r.dt)
    if read == "u":
        got["u"] = m.get_state()
    if read == "limits":
        m.set_limits(m.limits())
    m.TCheck(r)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 5)
    return got, m.get_state()


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("read", ["none", "u", "limits"])
def test_recorded_tvd_sweep(gold, table, read, split):
    """TVDlimit's dim-2 sweep recorded: run by the following TCheck with the
    clamp as its epilogue, or plain before a read or a change of the limits;
    bit for bit the calls with MRM_DEFER=0"""
    print(f"case: recorded dim-2 TVD sweep, read {read}, split {split}", flush=True)
    g1, U1 = _tvd_ops(table, gold, "1", read, split)
    g0, U0 = _tvd_ops(table, gold, "0", read, split)
    for k in g0:
        assert np.array_equal(g1[k], g0[k]), k
    assert np.array_equal(U1, U0)


@pytest.mark.parametrize("defer", ["1", "0"])
@pytest.mark.parametrize("case", ["plain", "va_adjust", "eta"])
def test_sequence_steps_match_oracle(table, case, defer):
    print(f"case: three drop-in steps, {case}, defer={defer}", flush=True)
    box = initial.BoxSpec(gsize=(40, 36, 28), dx=(3.2e6,) * 3, seed=23, v_rms=2.0e5)
    U = initial.make_state(eos_table.load_tables(table), box)
    phys = config.stratified_phys(va_adjust=1.0) if case == "va_adjust" else config.stratified_phys()
    if case == "eta":
        phys.params[config.I_PARAM_ETA] = 1.0e11
    o, ro = start_oracle(table, box, phys, U)
    with env(MRM_DEFER=defer):
        m, r = start_gpu(table, box, phys, U, path="sequence")
    for s in range(3):
        o.step(ro, config.DT_RAD_OFF, 1)
        m.step(r, config.DT_RAD_OFF, 1)
        print(f"step {s + 1}: dt gpu {r.dt:.15e} oracle {ro.dt:.15e}", flush=True)
    check_pair(m, r, o, ro)
    # the last ConsToPrim's outputs (stage 4's input), formed on this read
    # when the stage was deferred
    for f in (abi.PRES, abi.TEMP, abi.NE, abi.RHOI, abi.AMB):
        e = relerr(m.get_field(f), o.field(f))
        print("EOS field", f, "rel err", e, flush=True)
        assert e <= TOL, f
