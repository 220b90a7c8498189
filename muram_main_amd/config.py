"""Parameter sets for the MHD step.

``default_phys()`` returns PhysicsData's constructor defaults
(physics.H:111-198).  ``stratified_phys()`` applies the settings of the
stratified-box workloads: RT, Spitzer conduction, resistivity, ambipolar
diffusion, damping, divB cleaning, potential-field BC and diagnostics off;
gravity 2.74e4 cm/s^2 (hydrostatic 150 km scale height at T = 6000 K,
mu = 1.21); the TVD settings of the reference's 3D test case
(TEST/Test_3D/parameters.dat:67-91); half-open top boundary.
"""
from __future__ import annotations

import dataclasses
import numpy as np

from .abi import GridDesc, PhysDesc, RunDesc
from .initial import BoxSpec

# index constants (physics.H:15-75)
I_PARAM_GRAV, I_PARAM_VA_MAX, I_PARAM_VA_ADJUST, I_PARAM_SPITZER, I_PARAM_ETA, \
    I_PARAM_MAX_FILL, I_PARAM_AMBIPOLAR, I_PARAM_AMBFAC_MAX, I_PARAM_AMBVEL_MAX = range(9)
I_BND_TOP, I_BND_POT, I_BND_BCRIT, I_BND_EPS_TOP, I_BND_FEM = range(5)
I_TVD_RHOLEV, I_TVD_RHOLOG, I_TVD_QRHO, I_TVD_BPAR, I_TVD_VHYP, I_TVD_QDIFF_BND, \
    I_TVD_PM_V, I_TVD_PM_B, I_TVD_VMAX_LIM, I_TVD_CME_THRESH = range(10)
I_TCHK_EPS_MIN, I_TCHK_RHO_MIN, I_TCHK_EPS_MAX, I_TCHK_VMAX = range(4)
I_RT_TSTEP, I_RT_CFL = 3, 4
I_DIVB_SWITCH, I_DIVB_ITMAX, I_DIVB_ERR = range(3)
DIVB_FIELD, PHI_FIELD, SFLX_FIELD = 7, 8, 9     # mrm field ids of div B, phi, Spitzer heat flux


def default_phys() -> PhysDesc:
    p = PhysDesc()
    p.params[:] = [0.0, 1e7, 0, 0.0, 0.0, 1000.0, 0.0, 1e12, 1e7]
    p.bnd[:] = [0.0, 0.0, 1.0e10, 0.0, 0.0]
    p.tvd_h[:] = [2.0, 2.0, 2.0, 2.0]
    p.tvd_cs[:] = [0.2, 0.2, 0.2, 0.2]
    p.tvd_coeff[:] = [1.0, 1.0, 1.0, 1.0]
    p.tvd[:] = [1.0e-11, 1.0, 10.0, 0.2, 1.0, 1.0, 1.0, 1.0, 0.75, 100.0]
    p.tvd_h_bnd[:] = [0.0, 0.0]
    p.tvd_visc_bnd[:] = [0.0, 0.0]
    p.tvd_eta_bnd[:] = [0.0, 0.0]
    p.tchk[:] = [1.0e11, 1.0e-16, 4.0e15, 100.0e5]
    p.dmp[:] = [0, 1.0e3, 1.0e3, 1.0e2]
    p.rt[:] = [1, 1.0e-8, 20.0e3, 0.0, 0.5, 0, 0.0, 0, 1e2, 1e2]
    p.divB[:] = [1, 5, 0.2]
    p.rt_ext_cor = 0
    return p


def stratified_phys(va_max: float = 3.0e8, va_adjust: float = 0.0) -> PhysDesc:
    p = default_phys()
    p.params[I_PARAM_GRAV] = 2.74e4
    p.params[I_PARAM_VA_MAX] = va_max
    p.params[I_PARAM_VA_ADJUST] = va_adjust
    p.bnd[I_BND_TOP] = 1.0
    p.bnd[I_BND_POT] = 0.0
    p.bnd[I_BND_EPS_TOP] = 0.0
    p.divB[0] = 0.0
    # TEST/Test_3D/parameters.dat:67-91
    p.tvd_h[:] = [2.0, 2.0, 1.0, 0.0]
    p.tvd_cs[:] = [0.2, 0.2, 0.2, 1.0]
    p.tvd[:] = [1.0e-11, 1.0, 10.0, 0.2, 1.0, 0.0, 1.0, 4.0, 0.75, 0.5]
    p.tvd_h_bnd[:] = [0.01, 0.01]
    p.tvd_visc_bnd[:] = [1.0, 1.0]
    p.tvd_eta_bnd[:] = [1.0, 1.0]
    p.tvd_coeff[:] = [1.0, 1.0, 1.0, 1.0]
    p.tchk[:] = [1.0e11, 1.0e-20, 4.0e15, 1.0e7]
    return p


def test3d_phys(va_adjust: float = 1.0, divb: bool = True, spitzer: bool = True) -> PhysDesc:
    """TEST/Test_3D/parameters.dat with damping, RT and the potential-field
    top boundary off (outside the path): va_max 1e7 with va_adjust 1
    (:51-52), Spitzer conduction 1e-6 (:54), half-open top (:60), B_crit
    2.5e30 (:63), hot plate eps_top 2e14 (:64), divB cleaning on with 5
    passes at most and a 0.2 error reduction target (:94-96), tchk eps_min
    1.75e12, rho_min 1e-18, eps_max 2e14, vmax 5e7 (:99-102)."""
    p = stratified_phys(va_max=1.0e7, va_adjust=va_adjust)
    if divb:
        p.divB[:] = [1, 5, 0.2]
    if spitzer:
        p.params[I_PARAM_SPITZER] = 1.0e-6
    p.bnd[I_BND_BCRIT] = 2.5e30
    p.bnd[I_BND_EPS_TOP] = 2.0e14
    p.tchk[:] = [1.75e12, 1.0e-18, 2.0e14, 5.0e7]
    return p


def grid_desc(gsize, dx, ghosts: int = 2, periods=(0, 1, 1)) -> GridDesc:
    g = GridDesc()
    for d in range(3):
        g.gsize[d] = int(gsize[d])
        g.ghosts[d] = ghosts
        g.periods[d] = int(periods[d])
        g.gxmin[d] = 0.0
        g.gxmax[d] = float(gsize[d]) * float(dx[d])
    return g


def run_desc(CFL: float = 0.8, CFL_tvd: float = 0.95) -> RunDesc:
    r = RunDesc()
    r.dt = 0.0
    r.time = 0.0
    r.CFL = CFL
    r.CFL_tvd = CFL_tvd
    r.outcad = 0.0
    r.autostep = 1          # dt = 0 in parameters.dat -> autostep (run.H:52)
    r.need_diagnostics = 0
    r.globiter = 0
    r.iteration = 0
    return r


DT_RAD_OFF = 1.0e30     # finite: +inf would abort in Sync_timestep (solver.C:497-501)


@dataclasses.dataclass
class Workload:
    name: str
    box: BoxSpec
    phys: PhysDesc
    steps: int


def workloads():
    """The stratified-box workloads, MURaM axis order (vertical first)."""
    return {
        "box64": Workload("box64: 64x64x64, dx=32 km, one RK step", BoxSpec(
            gsize=(64, 64, 64), dx=(3.2e6,) * 3), stratified_phys(), 1),
        "box512": Workload("box512: 256x512x512 (vertical 256), dx=16 km, FP64", BoxSpec(
            gsize=(256, 512, 512), dx=(1.6e6,) * 3), stratified_phys(va_max=3.0e8), 10),
        "box512_v3": Workload("box512_v3: 256x512x512, v_rms 3 km/s", BoxSpec(
            gsize=(256, 512, 512), dx=(1.6e6,) * 3, v_rms=3.0e5), stratified_phys(va_max=3.0e8), 100),
        "corona": Workload("corona: 768x384x384 (vertical 768), dz=32 km", BoxSpec(
            gsize=(768, 384, 384), dx=(3.2e6,) * 3, corona=True, B_vert_G=100.0, dB_rms_G=30.0),
            stratified_phys(va_max=1.0e9), 10),
        "adjoint128": Workload("adjoint128: 128^3, dx=32 km", BoxSpec(
            gsize=(128, 128, 128), dx=(3.2e6,) * 3), stratified_phys(), 50),
    }
