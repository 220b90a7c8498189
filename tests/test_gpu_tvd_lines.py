"""GPU parity of the dim-0 TVD sweep in each of its kernels, with the
limiter branch accounting compared count for count against the oracle:

* the block-per-line kernel (k_tvd_x_lines) on short columns (<= 64 KB of
  LDS per block) and on the 772-point columns of a coronal crop (130 KB of
  LDS, the larger dynamic LDS of gfx950);
* the marching kernel (k_tvd_sweep<0>), forced on an ordinary box and taken
  by itself on a 1004-point column (whose line arrays exceed 160 KB);

each in the exact (no FMA) and contracted builds of the fused step, plus the
drop-in sequence; fields to 1e-12 of the oracle, dt equal, every branch count
and position checksum equal.
"""
import os

import pytest

from muram_main_amd import abi, config, initial
from gpu_pairs import TOL, check_pair, make_box_state, start_gpu, start_oracle

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def table():
    return open(os.path.join(HERE, "golden", "eos_small.dat"), "rb").read()


def corona_box(n0, n1, n2, seed):
    return initial.BoxSpec(gsize=(n0, n1, n2), dx=(3.2e6,) * 3, seed=seed, corona=True, B_vert_G=100.0,
                           dB_rms_G=30.0)


def run_case(table, box, phys, steps, **gpu):
    U = make_box_state(table, box)
    o, ro = start_oracle(table, box, phys, U)
    m, r = start_gpu(table, box, phys, U, **gpu)
    o.step(ro, config.DT_RAD_OFF, steps)
    m.step(r, config.DT_RAD_OFF, steps)
    return check_pair(m, r, o, ro)


@pytest.mark.parametrize("fma", ["0", "1"])
@pytest.mark.parametrize("tvd_x", [None, "march"])
def test_small_box_both_dim0_kernels(table, fma, tvd_x):
    print(f"case: 40x36x28 fused fma={fma} dim-0 {tvd_x or 'lines'}", flush=True)
    box = initial.BoxSpec(gsize=(40, 36, 28), dx=(3.2e6,) * 3, seed=31, v_rms=3.0e5)
    run_case(table, box, config.stratified_phys(), 3, fma=fma, tvd_x=tvd_x)


@pytest.mark.parametrize("tvd_x", [None, "march"])
def test_small_box_drop_in_sequence(table, tvd_x):
    print(f"case: 40x36x28 drop-in sequence dim-0 {tvd_x or 'lines'}", flush=True)
    box = initial.BoxSpec(gsize=(40, 36, 28), dx=(3.2e6,) * 3, seed=32, v_rms=3.0e5)
    run_case(table, box, config.stratified_phys(), 2, path="sequence", tvd_x=tvd_x)


@pytest.mark.parametrize("fma", ["0", "1"])
def test_corona_columns_large_lds(table, fma):
    """772-point columns: 21 x 772 x 8 B = 130 KB of line arrays per block"""
    print(f"case: corona 768x8x8 fused fma={fma} (x-lines, 130 KB LDS)", flush=True)
    run_case(table, corona_box(768, 8, 8, 41), config.stratified_phys(va_max=1.0e9), 2, fma=fma)


@pytest.mark.parametrize("fma", ["0", "1"])
def test_tall_columns_marching(table, fma):
    """1004-point columns exceed one block's LDS: the marching kernel"""
    print(f"case: corona 1000x6x6 fused fma={fma} (marching dim 0)", flush=True)
    run_case(table, corona_box(1000, 6, 6, 42), config.stratified_phys(va_max=1.0e9), 2, fma=fma)


def test_corona_columns_drop_in(table):
    print("case: corona 768x8x8 drop-in sequence", flush=True)
    run_case(table, corona_box(768, 8, 8, 43), config.stratified_phys(va_max=1.0e9), 2, path="sequence")


def test_branch_counts_off_by_default(table):
    """the accounting is off unless asked for, and turning it off stops it"""
    print("case: branch accounting off", flush=True)
    box = initial.BoxSpec(gsize=(16, 12, 12), dx=(3.2e6,) * 3, seed=5)
    U = make_box_state(table, box)
    m, r = start_gpu(table, box, config.stratified_phys(), U, count=False)
    m.step(r, config.DT_RAD_OFF, 1)
    assert not m.branch_counts().any()
    m.branch_count(True, 0.0)
    m.step(r, config.DT_RAD_OFF, 1)
    first = m.branch_counts().copy()
    assert first.any()
    m.branch_count(False)
    m.step(r, config.DT_RAD_OFF, 1)
    assert (m.branch_counts() == first).all()
