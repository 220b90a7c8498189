"""Hand-written adjoint of the FP64 step (mrm_adjoint_forward/backward)
against central finite differences of the forward, through a
torch.autograd.Function wrapper, and against the autograd tape of the FP64
PyTorch restatement of the same steps (tests/tape_step.py).

Objective J = sum of v_vertical^2 = (M0/rho)^2 over one vertical layer of the
interior after N steps.  Inputs: the interior of the initial state (rho, M, e,
B); its ghosts are filled by exchange + boundary as the forward does.  dt is
held fixed (autostep off) and p_bc/s_bc are pinned, which are the parameters
the adjoint holds fixed.
"""
import os

import numpy as np
import pytest
import torch

from muram_main_amd import abi, config, eos_table, initial
from muram_main_amd.mhd import MhdGrid

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


class Problem:
    def __init__(self, gsize, seed, nsteps, every, v_rms=2.0e5, dt_frac=0.5, layer=None, precision=abi.PREC_FP64,
                 dt=None):
        self.table = open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()
        tab = eos_table.load_tables(self.table)
        box = initial.BoxSpec(gsize=gsize, dx=(3.2e6,) * 3, seed=seed, v_rms=v_rms)
        self.U = initial.make_state(tab, box)
        self.g = config.grid_desc(box.gsize, box.dx)
        self.phys = config.stratified_phys()
        self.m = MhdGrid(self.g, self.phys, precision)
        self.m.eos_init(self.table)
        self.nsteps, self.every = nsteps, every
        self.layer = 2 + (gsize[0] // 2 if layer is None else layer)
        # the autostep dt of the initial state sets the fixed dt
        self.m.set_state(self.U)
        r = config.run_desc()
        self.m.exchange_grid()
        self.m.SetBoundaryConditions(r, 0)
        self.p_bc, self.s_bc = self.m.bc()
        self.m.step(r, config.DT_RAD_OFF, 1)
        self.dt = dt_frac * r.dt if dt is None else dt
        self.inner = (slice(2, -2),) * 3

    def run(self, U0, record=False):
        """forward from U0 (interior used, ghosts refilled); returns U after nsteps"""
        m = self.m
        m.set_state(U0)
        m.set_bc(self.p_bc, self.s_bc)
        r = config.run_desc()
        r.autostep = 0
        r.dt = self.dt
        m.exchange_grid()
        m.SetBoundaryConditions(r, 0)
        if record:
            m.adjoint_forward(r, config.DT_RAD_OFF, self.nsteps, self.every)
        else:
            m.step(r, config.DT_RAD_OFF, self.nsteps)
        return m.get_state()

    def J(self, UN):
        L = self.layer
        v = UN[1, 2:-2, 2:-2, L] / UN[0, 2:-2, 2:-2, L]
        return float(np.sum(v * v))


class StepFn(torch.autograd.Function):
    """U0 -> U after the problem's steps; backward by the library's adjoint."""

    @staticmethod
    def forward(ctx, U0, prob):
        ctx.prob = prob
        UN = prob.run(U0.detach().numpy(), record=True)
        return torch.from_numpy(UN)

    @staticmethod
    def backward(ctx, gUN):
        g = ctx.prob.m.adjoint_backward(gUN.detach().contiguous().numpy(), fold_initial_ghosts=True)
        return torch.from_numpy(g), None


def gradient(prob):
    U0 = torch.from_numpy(prob.U.copy()).requires_grad_(True)
    UN = StepFn.apply(U0, prob)
    L = prob.layer
    v = UN[1, 2:-2, 2:-2, L] / UN[0, 2:-2, 2:-2, L]
    J = (v * v).sum()
    J.backward()
    return U0.grad.numpy(), float(J)


def directions(prob, k, seed):
    """random interior directions scaled by each cell's own magnitudes"""
    rng = np.random.default_rng(seed)
    U, inner = prob.U, prob.inner
    rho = U[0][inner]
    v_rms = np.sqrt(np.mean((U[1][inner] / rho) ** 2 + (U[2][inner] / rho) ** 2 + (U[3][inner] / rho) ** 2) / 3)
    b_rms = np.sqrt(np.mean(U[5][inner] ** 2 + U[6][inner] ** 2 + U[7][inner] ** 2) / 3)
    out = []
    for _ in range(k):
        d = np.zeros_like(U)
        d[0][inner] = rho * rng.standard_normal(rho.shape)
        for q in (1, 2, 3):
            d[q][inner] = rho * v_rms * rng.standard_normal(rho.shape)
        d[4][inner] = U[4][inner] * rng.standard_normal(rho.shape)
        for q in (5, 6, 7):
            d[q][inner] = b_rms * rng.standard_normal(rho.shape)
        out.append(d)
    return out


def fd_check(prob, grad, dirs, hs=(1e-8, 1e-9)):
    """relative difference between the adjoint's directional derivative and
    central differences; the limiter has many kinks (minmod, the
    hyperdiffusion switch), so the steps are small enough that a difference
    rarely straddles one, and the better of the two steps is kept"""
    errs = []
    for d in dirs:
        ad = float(np.sum(grad[(slice(None),) + prob.inner] * d[(slice(None),) + prob.inner]))
        best = np.inf
        for h in hs:
            jp = prob.J(prob.run(prob.U + h * d))
            jm = prob.J(prob.run(prob.U - h * d))
            fd = (jp - jm) / (2 * h)
            e = abs(fd - ad) / max(abs(fd), abs(ad), 1e-300)
            print(f"  h {h:.0e}  fd {fd:+.12e}  adjoint {ad:+.12e}  rel {e:.2e}")
            best = min(best, e)
        errs.append(best)
    return errs


def test_adjoint_matches_finite_differences():
    print("case: adjoint 16x12x12, 4 steps, checkpoints every 2", flush=True)
    prob = Problem((16, 12, 12), seed=51, nsteps=4, every=2)
    grad, J = gradient(prob)
    print("J", J, "dt", prob.dt)
    errs = fd_check(prob, grad, directions(prob, 8, 7))
    assert np.median(errs) <= 1e-6
    assert max(errs) <= 1e-5


def test_adjoint_split_tvd_branch():
    print("case: adjoint, split TVD", flush=True)
    prob = Problem((12, 12, 12), seed=52, nsteps=2, every=1, dt_frac=1.2)
    grad, J = gradient(prob)
    errs = fd_check(prob, grad, directions(prob, 4, 8))
    assert np.median(errs) <= 1e-6
    assert max(errs) <= 1e-5


def test_adjoint_checkpoint_spacing_gives_the_same_gradient():
    print("case: adjoint checkpoint spacing", flush=True)
    p1 = Problem((12, 12, 12), seed=53, nsteps=5, every=1)
    g1, _ = gradient(p1)
    p1.every = 3
    g3, _ = gradient(p1)
    # the recorded forward carries the pressure of the contracted build from
    # step to step, the recomputation evaluates it with the exact build: the
    # states agree to rounding, the gradients to far below the FD tolerance
    np.testing.assert_allclose(g1, g3, rtol=1e-9, atol=1e-9 * np.abs(g1).max())


def tape_gradient(prob):
    """dJ/dU0 from the autograd tape of the PyTorch FP64 restatement of the
    same steps (tests/tape_step.py, matched to the oracle's forward)"""
    from tape_step import TapeStep
    tape = TapeStep(prob.g, prob.phys, prob.table, prob.p_bc, prob.s_bc)
    U0 = torch.tensor(prob.U, dtype=torch.float64, requires_grad=True)
    X = tape.fill(U0)
    for _ in range(prob.nsteps):
        X = tape.step(X, prob.dt)
    L = prob.layer
    v = X[1, 2:-2, 2:-2, L] / X[0, 2:-2, 2:-2, L]
    Jt = (v * v).sum()
    Jt.backward()
    return U0.grad.numpy(), float(Jt)


@pytest.mark.parametrize("gsize,seed,nsteps,every", [((16, 12, 12), 51, 2, 1), ((12, 10, 14), 55, 3, 2),
                                                     ((40, 32, 24), 57, 4, 2)])
def test_adjoint_matches_fp64_tape_gradient(gsize, seed, nsteps, every):
    print(f"case: adjoint vs FP64 tape gradient {gsize}, {nsteps} steps", flush=True)
    prob = Problem(gsize, seed=seed, nsteps=nsteps, every=every)
    grad, J = gradient(prob)
    gt, Jt = tape_gradient(prob)
    assert J == pytest.approx(Jt, rel=1e-12)
    errs = []
    for q in range(8):
        a, b = grad[q][prob.inner], gt[q][prob.inner]
        den = np.abs(b).max()
        errs.append(float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a).max()))
    print("per-field max|adjoint - tape| / max|tape|:", ["%.1e" % e for e in errs])
    # the same derivative taken two ways: agreement to rounding
    assert max(errs) <= 1e-10
    # the initial ghosts are overwritten by exchange + BC: no sensitivity there
    assert not np.any(gt[:, 0:2]) and not np.any(gt[:, :, :, 0:2])
