"""Helpers of the GPU parity tests: one seeded state stepped by the oracle
and by the library (through the C ABI) side by side, with the limiter
branch accounting on in both, and the comparison to the tests' bar."""
from __future__ import annotations

import contextlib
import os

import numpy as np

from muram_main_amd import abi, config, eos_table, initial
from muram_main_amd.mhd import MhdGrid
from orc import Oracle, relerr

TOL = 1e-12


@contextlib.contextmanager
def env(**kw):
    """set MRM_* knobs (read at context creation) for the duration"""
    saved = {k: os.environ.get(k) for k in kw}
    for k, v in kw.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v
    try:
        yield
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def make_box_state(table, box):
    tab = eos_table.load_tables(table)
    return initial.make_state(tab, box)


def start_oracle(table, box, phys, U, cfl=0.8, count=True, kernels=None):
    """kernels: the potential-field kernels (8, gsize[2], gsize[1]) for bnd_pot > 0"""
    g = config.grid_desc(box.gsize, box.dx)
    o = Oracle(g, phys, table)
    ro = config.run_desc(CFL=cfl)
    if kernels is not None:
        o.potential_kernels(kernels)
    o.set_state(U)
    o.exchange_grid()
    o.set_boundary_conditions(ro, 0)
    if count:
        o.branch_count(1, abi.BRANCH_TIE)
    return o, ro


def start_gpu(table, box, phys, U, cfl=0.8, path="fused", fma="0", tvd_x=None, precision=abi.PREC_FP64,
              count=True, kernels=None):
    """kernels: the potential-field kernels as an array, or the directory of
    their PSF-kernel-<c>.dat files"""
    g = config.grid_desc(box.gsize, box.dx)
    with env(MRM_STEP_PATH=path, MRM_FMA=fma, MRM_TVD_X=tvd_x):
        m = MhdGrid(g, phys, precision)
    m.eos_init(table)
    if isinstance(kernels, str):
        m.potential_kernels_load(kernels)
    elif kernels is not None:
        m.potential_kernels(kernels)
    r = config.run_desc(CFL=cfl)
    m.set_state(U)
    m.exchange_grid()
    m.SetBoundaryConditions(r, 0)
    if count and precision == abi.PREC_FP64:
        m.branch_count(True, abi.BRANCH_TIE)
    return m, r


def field_errs(Ug, Uo):
    return [relerr(Ug[v], Uo[v]) for v in range(8)]


def rel_l2(a, b):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    nb = float(np.sqrt(np.sum(np.asarray(b, np.float64) ** 2)))
    return float(np.sqrt(np.sum(d * d))) / nb if nb > 0 else float(np.sqrt(np.sum(d * d)))


def check_branches(m, o):
    """every decisive limiter choice equal in count and position checksum;
    the near and exact ties (margin below 1e-10 of the operands' scale, or
    zero) may differ where the inputs differ in their last bits"""
    bg, bo = m.branch_counts(), o.branch_counts()
    mis = abi.branch_mismatch(bg, bo)
    print("branch mismatch:", mis, flush=True)
    if mis["slots_differing"] or mis["checksums_differing"]:
        print("gpu:", abi.branch_dict(bg), "\noracle:", abi.branch_dict(bo), flush=True)
    assert (mis["slots_differing"], mis["count_diff"], mis["checksums_differing"]) == (0, 0, 0)
    assert mis["decisive"] > 0
    return mis


def check_pair(m, r, o, ro, tol=TOL, counts=True, linf_tol=None):
    """decisive branch choices equal (first, so that a field failure still
    reports them); over the interior cells, every field within tol in
    relative L2 (||gpu - ref|| / ||ref||) and within linf_tol (tol unless
    given) in max|gpu - ref| / max|ref|; dt and time to tol.  The ghost
    layers, which the boundary rebuilds from the interior every stage, are
    reported beside them."""
    mis = check_branches(m, o) if counts else None
    Ug, Uo = m.get_state(), o.state(0)
    inner = (slice(2, -2),) * 3
    errs = [relerr(Ug[v][inner], Uo[v][inner]) for v in range(8)]
    l2 = [rel_l2(Ug[v][inner], Uo[v][inner]) for v in range(8)]
    ebuf = field_errs(Ug, Uo)
    print("interior, per-field rel err (max):", ["%.2e" % e for e in errs], flush=True)
    print("interior, per-field rel err (L2): ", ["%.2e" % e for e in l2], "dt", r.dt, ro.dt, flush=True)
    print("whole buffer, per-field rel err (max):", ["%.2e" % e for e in ebuf], flush=True)
    w = int(np.argmax(ebuf))
    d = np.abs(Ug[w] - Uo[w])
    print("largest difference in the buffer: field", w, "at (k, j, i)", np.unravel_index(int(np.argmax(d)), d.shape),
          flush=True)
    assert max(l2) <= tol, l2
    assert max(errs) <= (tol if linf_tol is None else linf_tol), errs
    assert abs(r.dt - ro.dt) <= tol * abs(ro.dt)
    assert abs(r.time - ro.time) <= tol * abs(ro.time)
    return errs
