"""Restart files (mrm_io_*): result_prim_<v>.<iter>, Header.<iter> and
backup.dat against the files the reference wrote for its Test_3D case
(io_xysl.C:39-48, 224-412; boundary_pdmp_1_fftw3.C:44-62).  Host code only:
these run without a GPU."""
import os

import numpy as np
import pytest

from muram_main_amd import abi, config, mhd

HERE = os.path.dirname(os.path.abspath(__file__))
REF_INI = "/root/reference/TEST/Test_3D/ini"

pytestmark = pytest.mark.skipif(not os.path.exists(abi.LIB_PATH), reason="library not built")


@pytest.fixture(scope="module")
def crop():
    return np.load(os.path.join(HERE, "golden", "test3d_crop.npz"))


def crop_grid(z):
    n2, n1, n0 = z["prim"].shape[1:]
    return config.grid_desc((n0, n1, n2), tuple(z["dx"]))


def write_raw(dirpath, prim, it):
    for v in range(8):
        prim[v].astype(np.float32).tofile(os.path.join(dirpath, f"result_prim_{v}.{it:06d}"))


def cons_of_prim(prim, shape, g=2):
    """what RestoreSolution leaves in U (io_xysl.C:401-406), ghosts zero"""
    U = np.zeros((8,) + shape)
    p = prim.astype(np.float64)
    inner = (slice(g, -g),) * 3
    d = p[0]
    for v in range(8):
        U[v][inner] = p[v]
    U[4][inner] = p[4] + 0.5 * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]) * d
    for q in (1, 2, 3):
        U[q][inner] = p[q] * d
    return U


def shape_of(g):
    return tuple(g.gsize[d] + 2 * g.ghosts[d] for d in range(3))[::-1]


def test_backup_file_text_matches_reference(tmp_path, crop):
    it, t, p_bc, s_bc = crop["bc_20"]
    f = str(tmp_path / "backup.dat")
    mhd.write_backup(f, int(it), t, p_bc, s_bc)
    assert open(f).read() == str(crop["backup_20"])
    assert mhd.read_backup(f) == (20, t, p_bc, s_bc)


def test_header_file_text_matches_reference(tmp_path, crop):
    g = config.grid_desc(tuple(int(v) for v in crop["gsize"]), tuple(crop["dx"]))
    time, dt = crop["run_10"]
    d = str(tmp_path) + "/"
    mhd.write_header(d, 10, g, time, dt, float(crop["va_10"]))
    assert open(d + "Header.000010").read() == str(crop["header_10"])
    gs, dx, t, dt_, va = mhd.read_header(d, 10)
    assert gs == (192, 64, 64) and dx == (3.2e6, 9.6e6, 9.6e6)
    assert (t, dt_, va) == (1.354290, 0.135432, 28726091.32118)


def test_initial_header_and_backup_of_reference_case(tmp_path, crop):
    d = str(tmp_path) + "/"
    open(d + "Header.000000", "w").write(str(crop["header_ini"]))
    open(d + "backup.dat", "w").write(str(crop["backup_ini"]))
    assert mhd.read_header(d, 0) == ((192, 64, 64), (3.2e6, 9.6e6, 9.6e6), 0.0, 0.0, 0.0)
    # p_bc = s_bc = -1: SetBoundaryConditions derives them from the state
    assert mhd.read_backup(d + "backup.dat") == (0, 0.0, -1.0, -1.0)
    open(d + "two.dat", "w").write("7  1.5")
    assert mhd.read_backup(d + "two.dat") == (7, 1.5, -1.0, -1.0)


def test_missing_files_raise(tmp_path, crop):
    g = crop_grid(crop)
    d = str(tmp_path) + "/"
    with pytest.raises(abi.MrmError):
        mhd.read_backup(d + "none.dat")
    with pytest.raises(abi.MrmError):
        mhd.read_header(d, 3)
    with pytest.raises(abi.MrmError):
        mhd.read_prim(d, "result", 0, g, shape_of(g))
    # a truncated variable file is an error, not a partial state
    write_raw(str(tmp_path), crop["prim"], 0)
    with open(d + "result_prim_5.000000", "r+b") as f:
        f.truncate(100)
    with pytest.raises(abi.MrmError):
        mhd.read_prim(d, "result", 0, g, shape_of(g))


def test_read_prim_is_restore_solution_conversion(tmp_path, crop):
    g = crop_grid(crop)
    write_raw(str(tmp_path), crop["prim"], 0)
    U = mhd.read_prim(str(tmp_path) + "/", "result", 0, g, shape_of(g))
    np.testing.assert_array_equal(U, cons_of_prim(crop["prim"], shape_of(g)))


def test_write_then_read_reproduces_the_files(tmp_path, crop):
    g = crop_grid(crop)
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    write_raw(str(a), crop["prim"], 0)
    U = mhd.read_prim(str(a) + "/", "result", 0, g, shape_of(g))
    mhd.write_prim(str(b) + "/", "result", 12, g, U)
    for v in range(8):
        x = np.fromfile(a / f"result_prim_{v}.000000", dtype=np.float32)
        y = np.fromfile(b / f"result_prim_{v}.000012", dtype=np.float32)
        # M/d and e - rho v^2/2 recover the stored floats up to double rounding
        np.testing.assert_allclose(y, x, rtol=2 ** -23, atol=0)


@pytest.mark.skipif(not os.path.isdir(REF_INI), reason="reference tree not present")
def test_read_full_reference_initial_state():
    g = config.grid_desc((192, 64, 64), (3.2e6, 9.6e6, 9.6e6))
    U = mhd.read_prim(REF_INI + "/", "result", 0, g, shape_of(g))
    prim = np.stack([np.fromfile(f"{REF_INI}/result_prim_{v}.000000", dtype=np.float32).reshape(64, 64, 192)
                     for v in range(8)])
    np.testing.assert_array_equal(U, cons_of_prim(prim, shape_of(g)))


def test_heat_flux_file_round_trip(tmp_path, crop):
    """result_prim_8 (Spitzer's heat flux) through mrm_io_read_var /
    mrm_io_write_var: the interior in the reference's order, ghosts zero"""
    g = crop_grid(crop)
    d = str(tmp_path) + "/"
    crop["prim"][8].tofile(d + "result_prim_8.000000")
    f = mhd.read_var(d, "result", 8, 0, g, shape_of(g))
    assert np.array_equal(f[2:-2, 2:-2, 2:-2], crop["prim"][8].astype(np.float64))
    assert not f[:2].any() and not f[:, :, :2].any()
    mhd.write_var(d, "result", 8, 3, g, f)
    assert open(d + "result_prim_8.000003", "rb").read() == open(d + "result_prim_8.000000", "rb").read()
