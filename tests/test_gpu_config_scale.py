"""The precision and adjoint workloads at their own sizes (bench.py runs the
same studies and reports the numbers):

* box512_v3 (256 x 512 x 512, v_rms 3 km/s), 100 steps: the FP32-mixed step
  against the FP64 step, with the control -- the FP64 step from the
  FP32-rounded initial state, the configuration's own sensitivity to one
  rounding of the stored state.  Bars: largest per-field relative L2 over
  the interior within 2.5x the control's (measured 8.2e-4 against 4.7e-4:
  B in the upper half, where rho falls to 4e-19 g/cm^3 and v_A/v_A,max
  reaches ~300), over the lower half within 5x (1.5e-5 against 4.4e-6: the
  storage rounding of the mixed form's stages -- it rounds after every RK
  stage and TCheck -- which tests/test_gpu_mixed_floor.py separates), dt
  within 1e-6, e_int positive.
* corona (768 x 384 x 384, tanh transition to a 1e6 K corona, Alfven
  limiter 1e9 cm/s), 10 steps: largest per-field relative L2 within 4x the
  control's (4.5e-5 against 1.9e-5); the momentum's relative Linf over the
  upper half (the corona) within 1.5x the control's (0.23 against 0.20);
  e_int positive in both forms.
* adjoint128 (128^3), J = sum v_vertical^2 over layer 64: the adjoint
  gradient against central differences along 16 random interior directions
  (steps 1e-8 .. 3e-10 of each cell's magnitudes, each direction's estimate
  at the step where it agrees best with the next smaller one).  10 steps
  (checkpoints every 5): at least 14 of 16 within 1e-6, all within 5e-6
  (measured 14, largest 3.2e-6); 50 steps (every 10): at least 12 within
  1e-6, all within 3e-5 (measured 12, largest 1.2e-5).  The finite
  differences, not the adjoint, set these limits: the forward is smooth only
  between the limiter's switches, about 1e6 of which per step lie within
  1e-10 of their operands' scale; the adjoint matches an FP64 tape of the
  same step to 3e-14 (tests/test_gpu_adjoint.py).
"""
import numpy as np
import pytest

import bench
from muram_main_amd import config, eos_table

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables():
    table = eos_table.generate(eos_table.TableSpec(n_eps=512, n_lr=512, n_lp3=200, n_s3=150))
    return table, eos_table.load_tables(table)


def test_box512_v3_fp32_mixed_100_steps(tables):
    print("case: box512_v3 256x512x512, 100 steps, FP32-mixed vs FP64 with the rounded-state control", flush=True)
    table, tab = tables
    out = bench.precision_study(config.workloads()["box512_v3"], table, tab, 100, control=True, positivity=True)
    print({k: out[k] for k in ("max_l2", "control_max_l2", "max_l2_lower_half", "control_max_l2_lower_half",
                               "dt_rel_diff")}, flush=True)
    assert out["max_l2"] <= 2.5 * out["control_max_l2"]
    assert out["max_l2_lower_half"] <= 5.0 * out["control_max_l2_lower_half"]
    assert out["dt_rel_diff"] <= 1e-6
    assert out["eint_fp32_mixed"]["n_nonpositive"] == 0
    assert out["eint_fp64"]["n_nonpositive"] == 0


def test_corona_fp32_mixed_10_steps(tables):
    print("case: corona 768x384x384, 10 steps, FP32-mixed vs FP64 with the rounded-state control", flush=True)
    table, tab = tables
    out = bench.precision_study(config.workloads()["corona"], table, tab, 10, control=True, positivity=True)
    mx = max(out["errors_upper_half"][f]["linf"] for f in ("M0", "M1", "M2"))
    cx = max(out["control_upper_half"][f]["linf"] for f in ("M0", "M1", "M2"))
    print({"max_l2": out["max_l2"], "control_max_l2": out["control_max_l2"], "upper_M_linf": mx,
           "control_upper_M_linf": cx}, flush=True)
    assert out["max_l2"] <= 4.0 * out["control_max_l2"]
    assert mx <= 1.5 * cx
    assert out["eint_fp32_mixed"]["n_nonpositive"] == 0
    assert out["eint_fp64"]["n_nonpositive"] == 0


@pytest.mark.parametrize("steps,every,nmin,worst", [(10, 5, 14, 5e-6), (50, 10, 12, 3e-5)])
def test_adjoint128_against_finite_differences(tables, steps, every, nmin, worst):
    print(f"case: adjoint128 128^3, {steps} steps, checkpoints every {every}, 16 directions", flush=True)
    table, tab = tables
    out, _ = bench.adjoint_study(table, tab, steps=steps, every=every, ndir=16, hs=(1e-8, 3e-9, 1e-9, 3e-10))
    pl = np.array(out["fd_rel_diff_plateau"])
    print("plateau relative differences:", out["fd_rel_diff_plateau"], flush=True)
    assert len(pl) == 16
    assert np.count_nonzero(pl <= 1e-6) >= nmin
    assert pl.max() <= worst
