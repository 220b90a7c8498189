"""GPU parity of the potential-field top boundary (bnd_pot > 0;
boundary_pdmp_1_fftw3.C:335-474, potential_sd_fftw3.C:27-257) against the
oracle, whose direct circular convolution the CPU tests pin to numpy's FFT
and to the potential field of single Fourier modes.

* b_ext after the first SetBoundaryConditions: the device's FFT convolution
  (mixed-radix Stockham transforms in LDS) within 1e-14 of max|B_x(lend)|
  sum|K_q| (the bound of the convolution's sum) of the oracle's and of
  numpy's FFT convolution, its ghosts exact periodic copies, with the Test_3D kernels read
  from their files (64 x 64) and with random kernels on 60 x 42 and 50 x 96
  planes (radices 4, 3, 5, 2, 7);
* three steps of a 24 x 64 x 64 box with the Test_3D kernels for bnd_pot = 1
  (b_ext every stage), 2 (every step) and 3 (every other step), on the fused
  step and on the drop-in sequence: every decisive limiter choice equal,
  every field within 1e-12 in relative L2 and max norm, dt equal to 1e-12
  (the top layer's new B_y, B_z enter the next step's cmax);
* the errors: no kernels, a grid the kernels do not fit, an FP32-mixed
  context.
"""
import numpy as np
import pytest

from muram_main_amd import abi, config, eos_table, initial
from muram_main_amd.mhd import MhdGrid, MrmError
from gpu_pairs import check_pair, start_gpu, start_oracle
from potential_cases import PSF_DIR, conv_fft, pot_box, pot_phys, psf_kernels, with_ghosts

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def table():
    import os
    return open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "eos_small.dat"), "rb").read()


def state(table, box):
    return initial.make_state(eos_table.load_tables(table), box)


def check_bext(m, o, K, tol=1e-14):
    """per plane q, max|gpu - oracle| within tol of max|B_x(lend)| sum|K_q|
    (the bound of the convolution's sum), and the ghosts exact periodic copies"""
    bg, bo = m.potential_bext(), o.potential_bext()
    bz = np.abs(o.state(0)[5][2:-2, 2:-2, -3]).max()
    for q in range(8):
        scale = bz * np.abs(np.asarray(K[q], np.float64)).sum()
        err = np.abs(bg[q] - bo[q]).max() / scale
        print(f"b_ext[{q}] max|gpu - oracle| / scale: {err:.2e}", flush=True)
        assert err <= tol
        assert np.array_equal(bg[q], with_ghosts(bg[q][2:-2, 2:-2]))
    return bg


def test_bext_test3d_kernels(table):
    print("case: b_ext, Test_3D kernels from their files, 24 x 64 x 64", flush=True)
    box = pot_box(24)
    U = state(table, box)
    K = psf_kernels()
    o, ro = start_oracle(table, box, pot_phys(1), U, count=False, kernels=K)
    m, r = start_gpu(table, box, pot_phys(1), U, count=False, kernels=PSF_DIR)
    bg = check_bext(m, o, K)
    # and the numpy FFT convolution of the device's own B_x(lend)
    bz = m.get_state()[5][2:-2, 2:-2, -3]
    for q in range(8):
        ref = conv_fft(bz, K[q])
        assert np.abs(bg[q][2:-2, 2:-2] - ref).max() <= 1e-14 * np.abs(bz).max() * np.abs(K[q]).sum()
    Ug, Uo = m.get_state(), o.state(0)
    d = max(np.abs(Ug[v] - Uo[v]).max() / np.abs(Uo[v]).max() for v in range(8))
    print("state after the boundary, max rel diff:", d, flush=True)
    assert d <= 1e-13


@pytest.mark.parametrize("n1,n2", [(60, 42), (96, 50)])
def test_bext_mixed_radix(table, n1, n2):
    print(f"case: b_ext, random kernels, 8 x {n1} x {n2}", flush=True)
    box = initial.BoxSpec(gsize=(8, n1, n2), dx=(3.2e6, 9.6e6, 9.6e6), seed=4)
    U = state(table, box)
    rng = np.random.default_rng(n1 * n2)
    K = rng.normal(0.0, 1.0, (8, n2, n1)).astype(np.float32)
    o, ro = start_oracle(table, box, pot_phys(1), U, count=False, kernels=K)
    m, r = start_gpu(table, box, pot_phys(1), U, count=False, kernels=K)
    bg = check_bext(m, o, K)
    bz = m.get_state()[5][2:-2, 2:-2, -3]
    for q in range(8):
        ref = conv_fft(bz, K[q])
        assert np.abs(bg[q][2:-2, 2:-2] - ref).max() <= 1e-14 * np.abs(bz).max() * np.abs(K[q]).sum()


@pytest.mark.parametrize("pot", [1, 2, 3])
@pytest.mark.parametrize("path", ["fused", "sequence"])
def test_steps_match_oracle(table, pot, path):
    print(f"case: three steps, bnd_pot = {pot}, {path}", flush=True)
    box = pot_box(24)
    U = state(table, box)
    K = psf_kernels()
    o, ro = start_oracle(table, box, pot_phys(pot), U, kernels=K)
    m, r = start_gpu(table, box, pot_phys(pot), U, path=path, kernels=K)
    for s in range(3):
        o.step(ro, config.DT_RAD_OFF, 1)
        m.step(r, config.DT_RAD_OFF, 1)
        print(f"step {s + 1}: dt gpu {r.dt:.15e} oracle {ro.dt:.15e}", flush=True)
        assert abs(r.dt - ro.dt) <= 1e-12 * ro.dt
    check_pair(m, r, o, ro)
    check_bext(m, o, K, tol=1e-12)


def test_errors(table):
    print("case: potential-field errors", flush=True)
    box = pot_box(8)
    U = state(table, box)
    g = config.grid_desc(box.gsize, box.dx)
    # no kernels: the first boundary call fails
    m = MhdGrid(g, pot_phys(2))
    m.eos_init(table)
    m.set_state(U)
    m.exchange_grid()
    with pytest.raises(MrmError, match="MRM_ERR_STATE"):
        m.SetBoundaryConditions(config.run_desc(), 0)
    # files whose header does not fit the grid
    g2 = config.grid_desc((8, 32, 32), box.dx)
    m2 = MhdGrid(g2, pot_phys(2))
    with pytest.raises(MrmError, match="MRM_ERR_ARG"):
        m2.potential_kernels_load(PSF_DIR)
    # a horizontal extent with a prime factor above 7
    g3 = config.grid_desc((8, 22, 22), box.dx)
    m3 = MhdGrid(g3, pot_phys(2))
    with pytest.raises(MrmError, match="MRM_ERR_UNSUPPORTED"):
        m3.potential_kernels(np.zeros((8, 22, 22), np.float32))
    # FP32-mixed contexts keep the vertical-field boundary
    with pytest.raises(MrmError):
        MhdGrid(g, pot_phys(1), abi.PREC_FP32_MIXED)
