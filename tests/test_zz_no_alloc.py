"""GPU test of mrm_mhd_residual_noalloc, the entry point that allocates
nothing: on caller-owned device arrays it must give the residual and dt_cfl
of mrm_mhd_residual (same kernels: bit for bit), on two calls in a row, the
second on a smaller grid, with and without resistivity; and it must refuse
grids beyond its limit and physics it does not cover before doing any work."""
import ctypes as C
import os

import numpy as np
import pytest

from muram_main_amd import abi, config, eos_table, initial
from muram_main_amd.mhd import MhdGrid

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


def drop_in_residual(table, gsize, phys, seed):
    """the primitive state, pressure, Res before and after, dt_cfl and
    inv_va2max of a drop-in MHD_Residual call (solver.C's order: ConsToPrim,
    Adjust_Valf_Max, MHD_Residual onto a non-zero Res)"""
    tab = eos_table.load_tables(table)
    box = initial.BoxSpec(gsize=gsize, dx=(3.2e6, 3.2e6, 4.8e6), seed=seed, v_rms=2.0e5)
    U = initial.make_state(tab, box)
    g = config.grid_desc(box.gsize, box.dx)
    m = MhdGrid(g, phys)
    m.eos_init(table)
    r = config.run_desc(CFL=0.8)
    m.set_state(U)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    m.ConsToPrim()
    m.Adjust_Valf_Max(r)
    Up = m.get_state()
    P = m.get_field(abi.PRES)
    R0 = np.random.default_rng(seed).standard_normal(Up.shape) * 1e-3
    m.set_state(R0, abi.RES)
    dt = m.MHD_Residual(r)
    R1 = m.get_state(abi.RES)
    inv_va2max = float(m.limits()[0])
    m.close()
    return g, r, Up, P, R0, R1, dt, inv_va2max


def noalloc(g, phys, r, inv_va2max, Up, P, R0):
    import torch
    dU = torch.from_numpy(Up).cuda()
    dP = torch.from_numpy(P).cuda()
    dR = torch.from_numpy(R0).cuda()
    dT = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    lib = abi.load()
    torch.cuda.synchronize()
    rc = lib.mrm_mhd_residual_noalloc(C.byref(g), C.byref(phys), C.byref(r), inv_va2max, dU.data_ptr(),
                                      dP.data_ptr(), dR.data_ptr(), dT.data_ptr(), C.c_void_p(stream))
    assert rc == abi.MRM_OK
    torch.cuda.synchronize()
    return dR.cpu().numpy(), float(dT.cpu().numpy()[0])


@pytest.mark.parametrize("eta", [0.0, 1.0e10])
def test_noalloc_matches_drop_in_residual(table, eta):
    import torch
    assert torch.cuda.is_available()
    phys = config.stratified_phys()
    phys.params[config.I_PARAM_ETA] = eta
    for gsize, seed in (((40, 36, 28), 3), ((20, 12, 10), 4)):     # the second call on a smaller grid
        print(f"case: no-alloc residual {gsize}, eta {eta}", flush=True)
        g, r, Up, P, R0, R1, dt, inv = drop_in_residual(table, gsize, phys, seed)
        Rn, dtn = noalloc(g, phys, r, inv, Up, P, R0)
        assert np.array_equal(Rn, R1), float(np.max(np.abs(Rn - R1)))
        assert dtn == dt, (dtn, dt)
        assert not np.array_equal(R1, R0)


def test_noalloc_refusals(table):
    import torch
    lib = abi.load()
    dummy = torch.zeros(8, dtype=torch.float64, device="cuda")
    p = dummy.data_ptr()
    r = config.run_desc()
    phys = config.stratified_phys()
    big = config.grid_desc((1024, 1024, 512), (3.2e6,) * 3)    # 2^29 cells and more
    assert lib.mrm_mhd_residual_noalloc(C.byref(big), C.byref(phys), C.byref(r), 0.0, p, p, p, p, None) \
        == abi.ERRORS_BY_NAME["MRM_ERR_ARG"]
    small = config.grid_desc((8, 8, 8), (3.2e6,) * 3)
    spt = config.stratified_phys()
    spt.params[config.I_PARAM_SPITZER] = 1e-6
    assert lib.mrm_mhd_residual_noalloc(C.byref(small), C.byref(spt), C.byref(r), 0.0, p, p, p, p, None) \
        == abi.ERRORS_BY_NAME["MRM_ERR_UNSUPPORTED"]
    assert lib.mrm_mhd_residual_noalloc(C.byref(small), C.byref(phys), C.byref(r), 0.0, None, p, p, p, None) \
        == abi.ERRORS_BY_NAME["MRM_ERR_ARG"]
    torch.cuda.synchronize()
