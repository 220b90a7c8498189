"""Adjoint in an FP32-mixed context (mrm_adjoint_forward/backward): the
forward runs the mixed step and keeps mixed checkpoints, the backward replays
each segment with the mixed step and applies the FP64 transposes at the
widened mixed states.  It is the derivative of the FP64 step taken along the
mixed trajectory, so it is compared with the FP64 context's adjoint from the
same initial state and dt (which the FP64 tape and finite differences pin):
per field, max|mixed - FP64| / max|FP64| over the interior, and the
directional derivatives along random directions.  The bound is set by the
distance between the two trajectories (FP32 storage and stencils: about 1e-6
of the state after a few steps), which moves the linearisation point.
"""
import numpy as np
import pytest

from muram_main_amd import abi
from test_gpu_adjoint import Problem, directions, gradient

pytestmark = pytest.mark.gpu


def field_errs(a, b, inner):
    out = []
    for q in range(8):
        x, y = a[q][inner], b[q][inner]
        den = np.abs(y).max()
        out.append(float(np.abs(x - y).max() / den) if den > 0 else float(np.abs(x).max()))
    return out


@pytest.mark.parametrize("nsteps,every", [(4, 2), (3, 1)])
def test_mixed_adjoint_matches_fp64_adjoint(nsteps, every):
    print(f"case: mixed adjoint 24x16x16, {nsteps} steps, checkpoints every {every}", flush=True)
    p64 = Problem((24, 16, 16), seed=61, nsteps=nsteps, every=every)
    g64, J64 = gradient(p64)
    pmx = Problem((24, 16, 16), seed=61, nsteps=nsteps, every=every, precision=abi.PREC_FP32_MIXED, dt=p64.dt)
    gmx, Jmx = gradient(pmx)
    print("J", J64, Jmx, "rel", abs(Jmx - J64) / abs(J64))
    errs = field_errs(gmx, g64, p64.inner)
    print("per-field max|mixed - FP64| / max|FP64|:", ["%.1e" % e for e in errs])
    # directional derivatives: the difference relative to the sum of the
    # magnitudes of the terms (a derivative that nearly cancels is not
    # measured against its own small value)
    dd = []
    for d in directions(p64, 6, 9):
        t64 = g64[(slice(None),) + p64.inner] * d[(slice(None),) + p64.inner]
        tmx = gmx[(slice(None),) + p64.inner] * d[(slice(None),) + p64.inner]
        dd.append(abs(float(np.sum(tmx)) - float(np.sum(t64))) / float(np.sum(np.abs(t64))))
    print("directional derivatives, diff / sum|terms|:", ["%.1e" % e for e in dd])
    # measured: J 6e-7, fields <= 3.1e-6, directional <= 4.3e-8
    assert abs(Jmx - J64) <= 1e-5 * abs(J64)
    assert max(errs) <= 1e-5
    assert max(dd) <= 1e-6


def test_mixed_adjoint_checkpoint_spacing_and_state():
    """checkpoints every step (no replay) and every 3 steps (replayed mixed
    steps) give the same gradient; the backward leaves the context's state
    as the forward left it"""
    print("case: mixed adjoint checkpoint spacing", flush=True)
    p = Problem((16, 12, 12), seed=63, nsteps=5, every=1, precision=abi.PREC_FP32_MIXED)
    g1, _ = gradient(p)
    UN1 = p.m.get_state()
    p.every = 3
    g3, _ = gradient(p)
    UN3 = p.m.get_state()
    assert np.array_equal(UN1, UN3)
    # the replay forms the pressure of a restored state with k_eos_p_mx where
    # the recorded forward carried it from the previous step's last TVD sweep
    errs = field_errs(g3, g1, p.inner)
    print("per-field max|every 3 - every 1| / max:", ["%.1e" % e for e in errs])
    # measured: <= 8.3e-10
    assert max(errs) <= 1e-8
    # the context steps on from the forward's final state
    r = p.m.get_state()
    assert np.array_equal(r, UN3)
