"""Seeded synthetic initial states for the MHD step.

Layout: one float64 array ``U[8, n2, n1, n0]`` per state in the field order of
MURaM's ``cState`` (rho, M0, M1, M2, e, B0, B1, B2; mhd3d.H:24-27), with dim 0
(vertical, MURaM "x") fastest, exactly the reference's node order
``node = i + j*stride[1] + k*stride[2]`` (grid.C:196-199).  ``n_d = gsize[d] + 2*ghosts``.

The stratified box: rho = rho0*exp(-h/H), isothermal T, e_int from the EOS table
(eps such that T_interp(eps, rho) = T, found by bisection along the table's eps
axis as T_interp evaluates it, eos_init.C:292-316), Gaussian velocities, and
B = B_vertical + curl(A) with a Gaussian vector potential A and the same
4th-order centred operator the residual uses (mhdres_SR.C:183-184), so the
perturbation is divergence-free to rounding.  Magnetic field is in code units
B_code = B_Gauss/sqrt(4 pi) (mhdres_SR.C:317 uses p_mag = B^2/2).
Energy is e = e_int + rho v^2/2 (no magnetic part, cons_to_prim.C:118).
"""
from __future__ import annotations

import dataclasses
import numpy as np

from .eos_table import Tables

SQRT4PI = np.sqrt(4.0 * np.pi)


@dataclasses.dataclass
class BoxSpec:
    gsize: tuple = (64, 64, 64)       # MURaM order: (vertical, horizontal, horizontal)
    dx: tuple = (3.2e6, 3.2e6, 3.2e6)  # cm
    ghosts: int = 2
    rho0: float = 3.0e-7              # g/cm^3 at h = 0
    scale_height: float = 1.5e7       # cm
    T: float = 6000.0                 # K
    v_rms: float = 1.0e5              # cm/s per component
    B_vert_G: float = 200.0
    dB_rms_G: float = 50.0
    seed: int = 0
    # optional coronal profile (tanh transition of ln rho and T)
    corona: bool = False
    rho_top: float = 1.0e-15
    T_top: float = 1.0e6
    z_tr: float = 2.0e8
    w_tr: float = 2.0e7


def t_interp(tab: Tables, ee, dd):
    """Vectorised T(eps, rho) with the reference's bilinear form (eos_init.C:292-316)."""
    ee1 = np.log(ee + tab.eps_off)
    logr = np.log(dd)
    i = np.clip(((ee1 - tab.xeps[0]) * (1.0 / tab.del_eps)).astype(np.int64), 0, tab.n_eps - 2)
    j = np.clip(((logr - tab.xlr[0]) * (1.0 / tab.del_lr)).astype(np.int64), 0, tab.n_lr - 2)
    T = tab.fwd[1].astype(np.float64)
    v = ((logr - tab.xlr[j]) * (ee1 - tab.xeps[i]) * T[i + 1, j + 1]
         + (logr - tab.xlr[j]) * (tab.xeps[i + 1] - ee1) * T[i, j + 1]
         + (tab.xlr[j + 1] - logr) * (ee1 - tab.xeps[i]) * T[i + 1, j]
         + (tab.xlr[j + 1] - logr) * (tab.xeps[i + 1] - ee1) * T[i, j])
    return np.exp(v)


def eps_for_T(tab: Tables, rho, T, iters=200):
    """Bisection for eps with T_interp(eps, rho) = T (T_interp is monotone in eps)."""
    rho = np.asarray(rho, dtype=np.float64)
    T = np.broadcast_to(np.asarray(T, dtype=np.float64), rho.shape)
    lo = np.full(rho.shape, np.exp(tab.xeps[0]) - tab.eps_off)
    hi = np.full(rho.shape, np.exp(tab.xeps[-1]) - tab.eps_off)
    lo = np.maximum(lo, 1.0)
    for _ in range(iters):
        mid = np.sqrt(lo * hi)
        f = t_interp(tab, mid, rho) - T
        lo = np.where(f < 0, mid, lo)
        hi = np.where(f < 0, hi, mid)
    return np.sqrt(lo * hi)


def _d4(a, axis, dx):
    """4th-order centred derivative with periodic wrap (mhdres_SR.C:183-184)."""
    w1 = 8.0 / (12.0 * dx)
    w2 = -1.0 / (12.0 * dx)
    return (w1 * (np.roll(a, -1, axis) - np.roll(a, 1, axis))
            + w2 * (np.roll(a, -2, axis) - np.roll(a, 2, axis)))


def make_state(tab: Tables, spec: BoxSpec) -> np.ndarray:
    g = spec.ghosts
    n0, n1, n2 = (s + 2 * g for s in spec.gsize)
    rng = np.random.default_rng(spec.seed)
    U = np.empty((8, n2, n1, n0), dtype=np.float64)

    h = (np.arange(n0, dtype=np.float64) - g + 0.5) * spec.dx[0]
    if spec.corona:
        f = 0.5 * (1.0 + np.tanh((h - spec.z_tr) / spec.w_tr))
        lnrho = (1 - f) * (np.log(spec.rho0) - h / spec.scale_height) + f * np.log(spec.rho_top)
        rho_h = np.exp(lnrho)
        T_h = np.exp((1 - f) * np.log(spec.T) + f * np.log(spec.T_top))
    else:
        rho_h = spec.rho0 * np.exp(-h / spec.scale_height)
        T_h = np.full_like(h, spec.T)
    eps_h = eps_for_T(tab, rho_h, T_h)

    rho = np.broadcast_to(rho_h, (n2, n1, n0))
    U[0] = rho
    for c in range(3):
        U[1 + c] = rng.standard_normal((n2, n1, n0)) * spec.v_rms
    eint = rho * eps_h
    U[4] = eint + 0.5 * rho * (U[1] ** 2 + U[2] ** 2 + U[3] ** 2)
    for c in range(3):
        U[1 + c] *= rho          # v -> momentum

    # dB = curl A on the buffer, axis order of arrays is (2, 1, 0)
    A = [rng.standard_normal((n2, n1, n0)) for _ in range(3)]
    ax = {0: 2, 1: 1, 2: 0}      # MURaM dim -> numpy axis
    def d(a, dim):
        return _d4(a, ax[dim], spec.dx[dim])
    dB0 = d(A[2], 1) - d(A[1], 2)
    dB1 = d(A[0], 2) - d(A[2], 0)
    dB2 = d(A[1], 0) - d(A[0], 1)
    del A
    sl = (slice(g, n2 - g), slice(g, n1 - g), slice(g, n0 - g))
    rms = np.sqrt(np.mean(dB0[sl] ** 2 + dB1[sl] ** 2 + dB2[sl] ** 2))
    fac = (spec.dB_rms_G / SQRT4PI) / rms if rms > 0 else 0.0
    U[5] = spec.B_vert_G / SQRT4PI + fac * dB0
    U[6] = fac * dB1
    U[7] = fac * dB2
    return U
