"""Shared pieces of the damping tests (get_damping,
Add_Sources_Integrate_SR.C:92-108, 628-761): physics with the damping on and
a seeded state carrying a mean vertical flow, so that the layer-mean mass
flux the damping acts on is far from zero."""
from __future__ import annotations

import numpy as np

from muram_main_amd import config, initial

# dmp_tau_ref, dmp_tau_min as TEST/Test_3D/parameters.dat:105-108; vel_ref
# 1e4 cm/s puts the lower layers of these boxes on the 1/tau_min cap and the
# upper ones on the (flux / (vel_ref rho_max))^4 branch
TAU_REF, VEL_REF, TAU_MIN = 1.0e3, 1.0e4, 1.0e2


def damp_phys(vel_ref: float = VEL_REF, base=None):
    p = base if base is not None else config.stratified_phys()
    p.dmp[0] = 1.0
    p.dmp[1] = TAU_REF
    p.dmp[2] = vel_ref
    p.dmp[3] = TAU_MIN
    return p


def upflow_state(tab, box, v0: float = 1.0e5) -> np.ndarray:
    """the seeded stratified state plus a vertical velocity v0 sin(pi h/H)
    per layer (internal energy kept)"""
    U = initial.make_state(tab, box)
    n0 = U.shape[3]
    dv = v0 * np.sin(np.pi * (np.arange(n0) - 1.5) / (n0 - 4))
    rho, m = U[0], U[1]
    U[4] += m * dv + 0.5 * rho * dv * dv
    U[1] = m + rho * dv
    return U
