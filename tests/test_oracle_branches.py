"""CPU tests of the oracle's limiter branch accounting (orc_branch_count,
categories MRM_BR_* of include/muram_step.h, branch points of
tvdlimit_SR.C:281-657).

The reference keeps no such counts, so the accounting is pinned by what the
arithmetic fixes:
1. every evaluation is counted once: the slope, minmod and cmax categories
   sum to the number of slopes, interfaces and interfaces of the counted lines
   (interior transverse indices, every point along the line), and a Tau above
   1e-5 in chosen cells is seen once per sweep whose counted lines hold them;
2. a state whose only non-constant fields are rho = 2^k and B1 = small
   integers along the vertical (M = 0, e = 2^40 rho, rho_log = 0) has exactly
   representable slopes, so the slope choice (which of sl, sr, (sl+sr)/4 the
   min or max takes), the density-jump switch (|ln rho_i - ln rho_i+1| >
   ln 10, i.e. a jump of 2^4 or more) and the minmod's choice of every point
   and interface are known in advance; here they are formed with numpy's
   argmin/argmax (first index on ties) rather than the oracle's comparisons;
3. with rho_min above every density each update of the first sweep takes
   the density floor.
"""
import os

import numpy as np
import pytest

from muram_main_amd import abi, config
from orc import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
NB = len(abi.BRANCH_COUNTS)
IDX = {k: i for i, k in enumerate(abi.BRANCH_COUNTS)}


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


def lines_and_lengths(o):
    """(number of counted lines, points per line) of the three sweeps"""
    n2, n1, n0 = o.shape
    g = o.grid
    ls = [g.gsize[0], g.gsize[1], g.gsize[2]]
    return [(ls[1] * ls[2], n0), (ls[0] * ls[2], n1), (ls[0] * ls[1], n2)]


def prepared(o, U):
    """pressure from ConsToPrim of U, U restored, dt from the residual"""
    o.set_state(U)
    r = config.run_desc()
    o.cons_to_prim()
    o.adjust_valf_max(r, None)
    dt = o.mhd_residual(r)
    o.state(0)[...] = U
    o.state(2)[...] = 0.0
    return r, 0.8 * dt


def counts_of(o):
    return [int(x) for x in o.branch_counts()[:NB]]


def test_every_evaluation_counted_once(table):
    from muram_main_amd import eos_table, initial
    tab = eos_table.load_tables(table)
    box = initial.BoxSpec(gsize=(14, 10, 8), dx=(3.2e6,) * 3, seed=21, v_rms=3.0e5)
    U = initial.make_state(tab, box)
    g = config.grid_desc(box.gsize, box.dx)
    o = Oracle(g, config.stratified_phys(), table)
    o.set_state(U)
    r = config.run_desc()
    o.exchange_grid()
    o.set_boundary_conditions(r, 0)
    U = o.state(0).copy()
    r, dt = prepared(o, U)
    o.branch_count(1, 0.0)
    o.tvdlimit(r, dt)
    c = counts_of(o)
    L = lines_and_lengths(o)
    slopes = sum(n * (s - 2) for n, s in L)
    ifaces = sum(n * (s - 3) for n, s in L)
    updates = sum(n * (s - 4) for n, s in L)
    cat = lambda names: sum(c[IDX[k]] for k in names)
    assert cat(["slope_pos_l", "slope_pos_r", "slope_pos_m", "slope_neg_l", "slope_neg_r", "slope_neg_m",
                "slope_zero"]) == 8 * slopes
    assert cat(["mm_pos_lim", "mm_pos_sl", "mm_neg_lim", "mm_neg_sl", "mm_zero"]) == 8 * ifaces
    assert cat(["cmax_limit", "cmax_left", "cmax_right"]) == ifaces
    assert c[IDX["cz"]] == 0                   # Tau = 0
    assert c[IDX["qrho"]] <= slopes and c[IDX["clip"]] <= 8 * ifaces
    assert c[IDX["rho_floor"]] <= updates and c[IDX["e_floor"]] <= updates
    # a non-trivial state takes many of the choices
    assert min(c[IDX[k]] for k in ("slope_pos_l", "slope_pos_r", "slope_pos_m", "slope_neg_l", "slope_neg_r",
                                   "slope_neg_m", "mm_pos_lim", "mm_neg_lim", "mm_zero", "clip", "cme")) > 0
    # the limited slopes at both ends of an interface have the sign of its
    # difference or are zero, so sl_lim lies between 0 and sl: the minmod
    # never takes sl itself (a sign error in a slope would make it do so)
    assert c[IDX["mm_pos_sl"]] == 0 and c[IDX["mm_neg_sl"]] == 0
    # the checksums are reproducible and off when the accounting is off
    first = o.branch_counts().copy()
    o.set_state(U)
    o.branch_count(1, 0.0)
    o.tvdlimit(r, dt)
    assert np.array_equal(o.branch_counts(), first)
    o.branch_count(0, 0.0)
    o.set_state(U)
    o.tvdlimit(r, dt)
    assert np.array_equal(o.branch_counts(), first)
    # with a near-tie margin, every evaluation is decisive, a near tie or an
    # exact tie: the categories and the ties of each branch point add up again
    o.set_state(U)
    o.branch_count(1, 1e-3)
    o.tvdlimit(r, dt)
    t = [int(x) for x in o.branch_counts()]
    k = len(abi.BRANCH_SUMS)
    near = {c: t[NB + k + i] + t[NB + 2 * k + i] for i, c in enumerate(abi.BRANCH_SUMS)}
    assert near["slope"] > 0 and near["minmod"] > 0
    assert sum(t[IDX[k]] for k in ("slope_pos_l", "slope_pos_r", "slope_pos_m", "slope_neg_l", "slope_neg_r",
                                   "slope_neg_m", "slope_zero")) + near["slope"] == 8 * slopes
    assert sum(t[IDX[k]] for k in ("mm_pos_lim", "mm_pos_sl", "mm_neg_lim", "mm_neg_sl", "mm_zero")) \
        + near["minmod"] == 8 * ifaces
    assert sum(t[IDX[k]] for k in ("cmax_limit", "cmax_left", "cmax_right")) + near["cmax"] == ifaces


def test_tau_counted_once_per_sweep(table):
    g = config.grid_desc((8, 6, 6), (3.2e6,) * 3)
    o = Oracle(g, config.stratified_phys(), table)
    from muram_main_amd import eos_table, initial
    tab = eos_table.load_tables(table)
    U = initial.make_state(tab, initial.BoxSpec(gsize=(8, 6, 6), dx=(3.2e6,) * 3, seed=2))
    tau = np.zeros(o.shape)
    # cells (k, j, i): interior; vertical ghost of an interior column; dim-1
    # ghost of interior (i, k); dim-2 ghost of interior (i, j); a corner ghost
    cells = {(3, 3, 4): 3, (4, 2, 0): 1, (2, 0, 5): 1, (0, 4, 3): 1, (0, 0, 4): 0}
    for cell in cells:
        tau[cell] = 1.0
    o.field(abi.TAU)[...] = tau
    r, dt = prepared(o, U)
    o.branch_count(1, 0.0)
    o.tvdlimit(r, dt)
    assert counts_of(o)[IDX["cz"]] == sum(cells.values())


def designed_state(o, k, b1):
    """rho = 2^k[i] and B1 = b1[i] along dim 0, the same on every (j, k)
    line; M = 0, e = 2^40 rho (so e/rho is constant), B0 and B2 constant"""
    n2, n1, n0 = o.shape
    rho = np.ldexp(1.0, np.asarray(k))
    U = np.zeros((8, n2, n1, n0))
    U[0] = rho
    U[4] = np.ldexp(rho, 40)
    U[5], U[7] = 32.0, 4.0
    U[6] = np.asarray(b1, dtype=np.float64)
    return U


def expected_line(x, jump=None):
    """slope and minmod choices of one line of exactly representable values
    x: slope category per point 1..n-2 (0..2 positive from sl, sr, slm; 3..5
    negative; 6 zero), minmod category per interface 1..n-3 (0, 1 positive
    from sl_lim, sl; 2, 3 negative; 4 zero); jump: points whose slope the
    density-jump switch zeroes"""
    n = len(x)
    sl = x[1:-1] - x[:-2]
    sr = x[2:] - x[1:-1]
    slm = 0.25 * (sl + sr)
    cand = np.stack([sl, sr, slm])
    lo, hi = cand.min(axis=0), cand.max(axis=0)
    cat = np.where(lo > 0, np.argmin(cand, axis=0), np.where(hi < 0, 3 + np.argmax(cand, axis=0), 6))
    S = np.where(lo > 0, lo, 0.0) + np.where(hi < 0, hi, 0.0)
    if jump is not None:
        S = np.where(jump, 0.0, S)
    Sfull = np.zeros(n)
    Sfull[1:-1] = S
    mm = []
    for i in range(1, n - 2):
        s = x[i + 1] - x[i]
        s_lim = s - (Sfull[i] + Sfull[i + 1])
        pair = np.array([s_lim, s])
        if pair.min() > 0:
            mm.append(int(np.argmin(pair)) if s != s_lim else 0)
        elif pair.max() < 0:
            mm.append(2 + (int(np.argmax(pair)) if s != s_lim else 0))
        else:
            mm.append(4)
    return cat, np.array(mm)


def test_designed_profile_choices(table):
    g = config.grid_desc((12, 6, 6), (3.2e6,) * 3)
    p = config.stratified_phys()
    p.tvd[config.I_TVD_RHOLOG] = 0.0     # slopes of rho itself: exact differences
    o = Oracle(g, p, table)
    k = [-24, -24, -23, -21, -21, -22, -26, -25, -25, -24, -20, -20, -21, -23, -22, -22]
    b1 = [64, 66, 62, 68, 66, 72, 73, 69, 68, 63, 61, 62, 68, 74, 70, 72]
    assert len(k) == len(b1) == o.shape[2]
    U = designed_state(o, k, b1)
    r, dt = prepared(o, U)
    o.branch_count(1, 0.0)
    o.tvdlimit(r, dt)
    c = counts_of(o)
    lg = np.asarray(k)
    jump = np.abs(lg[1:-1] - lg[2:]) >= 4        # ratio 2^4 > 10 > 2^3
    cat_r, mm_r = expected_line(np.ldexp(1.0, lg), jump)
    cat_b, mm_b = expected_line(np.asarray(b1, dtype=np.float64))
    (L0, s0), (L1, s1), (L2, s2) = lines_and_lengths(o)
    names = ["slope_pos_l", "slope_pos_r", "slope_pos_m", "slope_neg_l", "slope_neg_r", "slope_neg_m"]
    for ci, name in enumerate(names):
        assert c[IDX[name]] == L0 * int(np.sum(cat_r == ci) + np.sum(cat_b == ci)), name
    # the six other fields are constant along dim 0, and every field is
    # constant along dims 1 and 2 (the dim-0 update is the same on every line)
    zero = L0 * (int(np.sum(cat_r == 6) + np.sum(cat_b == 6)) + 6 * (s0 - 2)) + 8 * (L1 * (s1 - 2) + L2 * (s2 - 2))
    assert c[IDX["slope_zero"]] == zero
    assert c[IDX["qrho"]] == L0 * int(np.sum(jump))
    mnames = ["mm_pos_lim", "mm_pos_sl", "mm_neg_lim", "mm_neg_sl"]
    for ci, name in enumerate(mnames):
        assert c[IDX[name]] == L0 * int(np.sum(mm_r == ci) + np.sum(mm_b == ci)), name
    zero_mm = L0 * (int(np.sum(mm_r == 4) + np.sum(mm_b == 4)) + 6 * (s0 - 3)) + 8 * (L1 * (s1 - 3) + L2 * (s2 - 3))
    assert c[IDX["mm_zero"]] == zero_mm
    # the profiles exercise every slope choice, the jump switch and the
    # minmod's possible choices
    cats = np.concatenate([cat_r, cat_b])
    mms = np.concatenate([mm_r, mm_b])
    assert all(np.sum(cats == ci) > 0 for ci in range(7)) and np.sum(jump) > 0
    assert all(np.sum(mms == ci) > 0 for ci in (0, 2, 4))


def test_density_floor_counted_per_update(table):
    """rho_min above every density: each update of the dim-0 sweep floors
    rho; the later sweeps then see rho = rho_min on every point of their
    lines (all of them were dim-0 updates), so no density flux and no floor"""
    g = config.grid_desc((8, 6, 6), (3.2e6,) * 3)
    p = config.stratified_phys()
    p.tchk[config.I_TCHK_RHO_MIN] = 1.0
    o = Oracle(g, p, table)
    from muram_main_amd import eos_table, initial
    tab = eos_table.load_tables(table)
    U = initial.make_state(tab, initial.BoxSpec(gsize=(8, 6, 6), dx=(3.2e6,) * 3, seed=3))
    r, dt = prepared(o, U)
    o.branch_count(1, 0.0)
    o.tvdlimit(r, dt)
    (L0, s0), _, _ = lines_and_lengths(o)
    assert counts_of(o)[IDX["rho_floor"]] == L0 * (s0 - 4)
