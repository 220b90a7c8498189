"""The reference's Test_3D initial state (a 192 x 10 x 10 crop of
TEST/Test_3D/ini, solar-convection stratification from rho 9e-16 to 5e-6
g/cm^3) restored through the restart-file path and stepped on the GPU:
FP64 (fused and drop-in sequence, with and without va_adjust and divB
cleaning) against the oracle at 1e-12, a BackupSolution/RestoreSolution round trip, and the
FP32-mixed form's error against FP64 on real data."""
import os

import numpy as np
import pytest

from muram_main_amd import abi, config, mhd
from muram_main_amd.mhd import MhdGrid
from orc import Oracle, field_errors, relerr

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
TOL = 1e-12


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


@pytest.fixture(scope="module")
def crop():
    return np.load(os.path.join(HERE, "golden", "test3d_crop.npz"))


def restart_dir(tmp_path, crop):
    """the crop as the reference lays out a restart: result_prim_<v>, backup.dat"""
    d = str(tmp_path) + "/"
    for v in range(9):
        crop["prim"][v].tofile(d + f"result_prim_{v}.000000")
    open(d + "backup.dat", "w").write(str(crop["backup_ini"]))
    n2, n1, n0 = crop["prim"].shape[1:]
    return d, config.grid_desc((n0, n1, n2), tuple(crop["dx"]))


def restored(d, g, table, precision=abi.PREC_FP64, va_adjust=1.0, path="fused", divb=True, spitzer=True):
    saved = os.environ.get("MRM_STEP_PATH")
    os.environ["MRM_STEP_PATH"] = path
    try:
        m = MhdGrid(g, config.test3d_phys(va_adjust, divb, spitzer), precision)
    finally:
        if saved is None:
            os.environ.pop("MRM_STEP_PATH", None)
        else:
            os.environ["MRM_STEP_PATH"] = saved
    m.eos_init(table)
    r = config.run_desc()
    m.RestoreSolution(d, "result", d + "backup.dat", r)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    return m, r


@pytest.mark.parametrize("va_adjust,divb,spitzer,path", [(1.0, True, True, "fused"), (1.0, True, True, "sequence"),
                                                        (1.0, True, False, "fused"), (1.0, True, False, "sequence"),
                                                        (1.0, False, False, "fused"), (0.0, False, False, "fused")])
def test_restored_test3d_steps_match_oracle(tmp_path, crop, table, va_adjust, divb, spitzer, path):
    # va_adjust 1 with divB cleaning and Spitzer conduction is the case's own
    # setting
    print(f"case: Test_3D crop 192x10x10, restore + 4 FP64 {path} steps vs oracle, va_adjust {va_adjust},"
          f" divB {divb}, Spitzer {spitzer}", flush=True)
    d, g = restart_dir(tmp_path, crop)
    m, r = restored(d, g, table, va_adjust=va_adjust, path=path, divb=divb, spitzer=spitzer)
    assert (r.globiter, r.time) == (0, 0.0)
    o = Oracle(g, config.test3d_phys(va_adjust, divb, spitzer), table)
    ro = config.run_desc()
    o.set_state(mhd.read_prim(d, "result", 0, g, m.shape))
    if spitzer:
        o.field(abi.SFLX)[...] = mhd.read_var(d, "result", 8, 0, g, m.shape)
        # the GPU's ghosts are already filled; the interior came from the file
        inner = (slice(2, -2),) * 3
        assert np.array_equal(m.get_field(abi.SFLX)[inner], o.field(abi.SFLX)[inner])
    o.set_bc(-1.0, -1.0)
    o.exchange_grid()
    o.set_boundary_conditions(ro, 0)
    assert m.bc() == pytest.approx(o.bc(), rel=TOL)
    o.step(ro, config.DT_RAD_OFF, 4)
    m.step(r, config.DT_RAD_OFF, 4)
    errs = [relerr(m.get_state()[v], o.state(0)[v]) for v in range(8)]
    if spitzer:
        errs.append(relerr(m.get_field(abi.SFLX), o.field(abi.SFLX)))
    print("per-field rel err:", ["%.2e" % e for e in errs], "dt", r.dt, ro.dt)
    assert max(errs) <= TOL
    assert r.dt == pytest.approx(ro.dt, rel=TOL)
    assert r.time == pytest.approx(ro.time, rel=TOL)
    np.testing.assert_allclose(m.limits(), o.limits(), rtol=TOL)


def test_backup_restore_round_trip(tmp_path, crop, table):
    print("case: Test_3D crop backup/restore round trip", flush=True)
    d, g = restart_dir(tmp_path, crop)
    m, r = restored(d, g, table)
    m.step(r, config.DT_RAD_OFF, 3)
    r.globiter = 3
    out = tmp_path / "out"
    out.mkdir()
    od = str(out) + "/"
    m.BackupSolution(od, "result", od + "backup.dat", r)
    U = m.get_state()
    p_bc, s_bc = m.bc()
    gs, dx, t, dt, va = mhd.read_header(od, 3)
    assert gs == tuple(g.gsize) and t == pytest.approx(r.time, abs=1e-6)
    assert dt == pytest.approx(r.dt, abs=1e-6)
    m2, r2 = restored(od, g, table)
    # backup.dat keeps time with %.16lg, 16 significant digits
    assert (r2.globiter, r2.time) == (3, float("%.16g" % r.time))
    assert m2.bc() == pytest.approx((p_bc, s_bc), rel=1e-6)
    U2 = m2.get_state()
    inner = (slice(2, -2),) * 3
    for v in range(8):
        # the files hold floats: the restored state is the step's state rounded
        assert relerr(U2[v][inner], U[v][inner]) <= 1e-6, v


@pytest.mark.parametrize("va_adjust,divb,spitzer", [(0.0, False, False), (1.0, False, False), (1.0, True, False),
                                                   (1.0, True, True)])
def test_mixed_precision_error_on_test3d(tmp_path, crop, table, va_adjust, divb, spitzer):
    print(f"case: Test_3D crop FP32-mixed vs FP64, 20 fused steps, va_adjust {va_adjust}, divB {divb},"
          f" Spitzer {spitzer}", flush=True)
    d, g = restart_dir(tmp_path, crop)
    m64, r64 = restored(d, g, table, va_adjust=va_adjust, divb=divb, spitzer=spitzer)
    mx, rx = restored(d, g, table, abi.PREC_FP32_MIXED, va_adjust=va_adjust, divb=divb, spitzer=spitzer)
    m64.step(r64, config.DT_RAD_OFF, 20)
    mx.step(rx, config.DT_RAD_OFF, 20)
    errs = field_errors(mx.get_state(), m64.get_state(), interior=(slice(2, -2),) * 3)
    print("FP32-mixed vs FP64 after 20 steps (rel L2, rel Linf):")
    for k, e in errs.items():
        print(f"  {k:4s} {e['l2']:.2e} {e['linf']:.2e}")
    if spitzer:
        es = relerr(mx.get_field(abi.SFLX)[(slice(2, -2),) * 3], m64.get_field(abi.SFLX)[(slice(2, -2),) * 3])
        print(f"  sflx rel Linf {es:.2e}")
        assert es <= 1e-3
    print("dt", rx.dt, r64.dt, "limits", mx.limits(), m64.limits())
    assert max(e["l2"] for e in errs.values()) <= 1e-4
    assert rx.dt == pytest.approx(r64.dt, rel=1e-5)
    np.testing.assert_allclose(mx.limits(), m64.limits(), rtol=1e-5)
