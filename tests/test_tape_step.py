"""The PyTorch FP64 restatement of the step (tests/tape_step.py), whose
autograd tape is the reference gradient for the adjoint, against the
oracle's forward: 1e-12 relative per field after two steps at fixed dt."""
import os

import numpy as np
import pytest
import torch

from muram_main_amd import config, eos_table, initial
from orc import Oracle, relerr
from tape_step import TapeStep

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


def setup(table, gsize, seed, v_rms=2.0e5, dt_frac=0.5):
    tab = eos_table.load_tables(table)
    box = initial.BoxSpec(gsize=gsize, dx=(3.2e6,) * 3, seed=seed, v_rms=v_rms)
    U = initial.make_state(tab, box)
    g, p = config.grid_desc(box.gsize, box.dx), config.stratified_phys()
    o = Oracle(g, p, table)
    r = config.run_desc()
    o.set_state(U)
    o.exchange_grid()
    o.set_boundary_conditions(r, 0)
    p_bc, s_bc = o.bc()
    o.step(r, config.DT_RAD_OFF, 1)
    return U, g, p, p_bc, s_bc, dt_frac * r.dt


def oracle_steps(table, U, g, p, p_bc, s_bc, dt, n):
    o = Oracle(g, p, table)
    o.set_state(U)
    o.set_bc(p_bc, s_bc)
    r = config.run_desc()
    r.autostep = 0
    r.dt = dt
    o.exchange_grid()
    o.set_boundary_conditions(r, 0)
    o.step(r, config.DT_RAD_OFF, n)
    return o.state(0).copy()


@pytest.mark.parametrize("gsize,seed", [((16, 12, 12), 51), ((12, 10, 14), 54)])
def test_tape_step_matches_oracle(table, gsize, seed):
    U, g, p, p_bc, s_bc, dt = setup(table, gsize, seed)
    ref = oracle_steps(table, U, g, p, p_bc, s_bc, dt, 2)
    tape = TapeStep(g, p, table, p_bc, s_bc)
    with torch.no_grad():
        X = tape.fill(torch.tensor(U))
        for _ in range(2):
            X = tape.step(X, dt)
    got = X.numpy()
    errs = [relerr(got[v], ref[v]) for v in range(8)]
    print("tape vs oracle:", ["%.1e" % e for e in errs])
    assert max(errs) <= 1e-12
