"""Host-side mirror of the reference's hot-path interface over the C ABI.

``MhdGrid`` owns one device context (one GridData) and exposes the reference's
operations under their own names -- eos_init, ConsToPrim, Adjust_Valf_Max,
MHD_Residual, Sync_timestep, Source_Integrate_Tcheck, SaveCons, TCheck,
TVDlimit, exchange_grid, SetBoundaryConditions -- plus ``step`` (the
ComputeSolution loop body).  Failures raise ``MrmError`` where the reference
would call MPI_Abort.  Arrays are numpy float64, SoA ``[8, n2, n1, n0]``.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi
from .abi import MrmError

_dp = C.POINTER(C.c_double)


class MhdGrid:
    def __init__(self, grid: abi.GridDesc, phys: abi.PhysDesc, precision: int = abi.PREC_FP64):
        self.lib = abi.load()
        self.grid, self.phys = grid, phys
        h = C.c_void_p()
        rc = self.lib.mrm_create(C.byref(grid), C.byref(phys), precision, C.byref(h))
        if rc != 0:
            raise MrmError(f"mrm_create failed: {abi.ERRORS.get(rc, rc)}")
        self.h = h
        bs = C.c_long()
        dims = (C.c_int * 3)()
        self.lib.mrm_bufsize(self.h, C.byref(bs), dims)
        self.bufsize = bs.value
        self.shape = (dims[2], dims[1], dims[0])

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.lib.mrm_last_error(self.h).decode(errors="replace")
            raise MrmError(f"{what}: {abi.ERRORS.get(rc, rc)}: {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.mrm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------ data
    def eos_init(self, table: bytes):
        self._tab = C.create_string_buffer(table, len(table))
        self._chk(self.lib.mrm_eos_load(self.h, self._tab, len(table)), "eos_init")

    def set_state(self, U: np.ndarray, which: int = abi.U, layout: int = abi.LAYOUT_SOA):
        a = np.ascontiguousarray(U, dtype=np.float64)
        assert a.size == 8 * self.bufsize
        self._chk(self.lib.mrm_set_state(self.h, which, a.ctypes.data_as(_dp), layout), "set_state")

    def get_state(self, which: int = abi.U, layout: int = abi.LAYOUT_SOA) -> np.ndarray:
        a = np.empty((8,) + self.shape if layout == abi.LAYOUT_SOA else self.shape + (8,), dtype=np.float64)
        self._chk(self.lib.mrm_get_state(self.h, which, a.ctypes.data_as(_dp), layout), "get_state")
        return a

    def set_field(self, f: int, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self._chk(self.lib.mrm_set_field(self.h, f, a.ctypes.data_as(_dp)), "set_field")

    def get_field(self, f: int) -> np.ndarray:
        a = np.empty(self.shape, dtype=np.float64)
        self._chk(self.lib.mrm_get_field(self.h, f, a.ctypes.data_as(_dp)), "get_field")
        return a

    def limits(self):
        a = (C.c_double * 4)()
        self.lib.mrm_get_limits(self.h, a)
        return np.array(a[:])

    def set_limits(self, lim):
        a = (C.c_double * 4)(*lim)
        self.lib.mrm_set_limits(self.h, a)

    def bc(self):
        p, s = C.c_double(), C.c_double()
        self.lib.mrm_get_bc(self.h, C.byref(p), C.byref(s))
        return p.value, s.value

    def set_bc(self, p_bc, s_bc):
        self.lib.mrm_set_bc(self.h, p_bc, s_bc)

    def synchronize(self):
        self._chk(self.lib.mrm_synchronize(self.h), "synchronize")

    # ------------------------------------------------------------ operations
    def ConsToPrim(self):
        self._chk(self.lib.mrm_cons_to_prim(self.h), "ConsToPrim")

    def Adjust_Valf_Max(self, run):
        st = (C.c_double * 5)()
        self._chk(self.lib.mrm_adjust_valf_max(self.h, C.byref(run), st), "Adjust_Valf_Max")
        return np.array(st[:])

    def MHD_Residual(self, run) -> float:
        dt = C.c_double()
        self._chk(self.lib.mrm_mhd_residual(self.h, C.byref(run), C.byref(dt)), "MHD_Residual")
        return dt.value

    def Sync_timestep(self, run, dt_rad, dt_mhd) -> float:
        cfl = C.c_double()
        self._chk(self.lib.mrm_sync_timestep(self.h, C.byref(run), dt_rad, dt_mhd, C.byref(cfl)),
                  "Sync_timestep")
        return cfl.value

    def Source_Integrate_Tcheck(self, run, stage: int):
        self._chk(self.lib.mrm_source_integrate_tcheck(self.h, C.byref(run), stage),
                  "Source_Integrate_Tcheck")

    def SaveCons(self):
        self._chk(self.lib.mrm_save_cons(self.h), "SaveCons")

    def TCheck(self, run):
        self._chk(self.lib.mrm_tcheck(self.h, C.byref(run)), "TCheck")

    def TVDlimit(self, run, dt_tvd: float):
        self._chk(self.lib.mrm_tvdlimit(self.h, C.byref(run), dt_tvd), "TVDlimit")

    def Clean_divB(self, run):
        """Clean_divB (clean_divB.C:35-369) -> max |div B| before and after"""
        err = (C.c_double * 2)()
        self._chk(self.lib.mrm_clean_divB(self.h, C.byref(run), err), "Clean_divB")
        return np.array(err[:])

    def exchange_grid(self):
        self._chk(self.lib.mrm_exchange_grid(self.h), "exchange_grid")

    # ------------------------------------------------------------ diagnostics
    def zero_diagnostics(self):
        """the zeroing of tvar1..8, Qres, Qvis at the start of the stage that
        collects them (solver.C:271-296)"""
        self._chk(self.lib.mrm_diag_zero(self.h), "zero_diagnostics")

    def diagnostics(self) -> dict:
        """tvar1..tvar8, Qres, Qvis (grid.H:69-100) by name, each [nz, ny, nx]"""
        return {n: self.get_field(abi.TVAR1 + t) for t, n in enumerate(abi.DIAG_NAMES)}

    def branch_count(self, on: bool = True, tie: float = 0.0):
        """limiter branch accounting on (zeroed) or off; choices within tie
        of their scale are counted as near ties"""
        self._chk(self.lib.mrm_branch_count(self.h, 1 if on else 0, tie), "mrm_branch_count")

    def branch_counts(self) -> np.ndarray:
        """the MRM_BR_NSLOT count and checksum slots (uint64)"""
        out = np.zeros(abi.BR_NSLOT, dtype=np.uint64)
        self._chk(self.lib.mrm_get_branch_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_ulonglong))),
                  "mrm_get_branch_counts")
        return out

    def SetBoundaryConditions(self, run, stage: int, pt_update: int | None = None):
        """SetBoundaryConditions(Run, Grid, Physics, stage, pt_update); pt_update
        None takes it from stage as solver.C passes it"""
        if pt_update is None:
            self._chk(self.lib.mrm_set_boundary_conditions(self.h, C.byref(run), stage), "SetBoundaryConditions")
        else:
            self._chk(self.lib.mrm_set_boundary_conditions_pt(self.h, C.byref(run), stage, int(pt_update)),
                      "SetBoundaryConditions")

    # ------------------------------------------------------------ potential field
    def potential_kernels(self, k8: np.ndarray):
        """the 8 potential-field kernels, shape (8, gsize[2], gsize[1]) float32,
        in the order abi.PSF_KERNELS"""
        k = np.ascontiguousarray(k8, dtype=np.float32)
        self._chk(self.lib.mrm_potential_kernels(self.h, k.ctypes.data_as(C.POINTER(C.c_float))),
                  "mrm_potential_kernels")

    def potential_kernels_load(self, directory: str):
        """the kernels from the PSF-kernel-<c>.dat files of a run directory"""
        self._chk(self.lib.mrm_potential_kernels_load(self.h, directory.encode()), "mrm_potential_kernels_load")

    def potential_bext(self) -> np.ndarray:
        """b_ext of the last potential update, shape (8, n2, n1)"""
        n2, n1, _ = self.shape
        out = np.empty((8, n2, n1), dtype=np.float64)
        self._chk(self.lib.mrm_get_potential_bext(self.h, out.ctypes.data_as(_dp)), "mrm_get_potential_bext")
        return out

    def step(self, run, dt_rad: float = 1.0e30, nsteps: int = 1):
        self._chk(self.lib.mrm_step(self.h, C.byref(run), dt_rad, nsteps), "step")

    # ------------------------------------------------------------ restart files
    def RestoreSolution(self, dir_3d: str, resfile: str, backfile: str, run):
        """RestoreSolution (io_xysl.C:272-412): read backup.dat (globiter, time,
        p_bc, s_bc), then result_prim_<v>.<globiter> (and v = 8, the heat
        flux, with Spitzer conduction); the state goes to the device."""
        it, t, p_bc, s_bc = read_backup(backfile)
        run.globiter = it
        run.time = t
        self.set_bc(p_bc, s_bc)
        self.set_state(read_prim(dir_3d, resfile, it, self.grid, self.shape))
        if self.phys.params[3] > 0.0:     # Spitzer's heat flux is variable 8 (io_xysl.C:389-392)
            self.set_field(abi.SFLX, read_var(dir_3d, resfile, 8, it, self.grid, self.shape))

    def BackupSolution(self, dir_3d: str, resfile: str, backfile: str, run):
        """BackupSolution (io_xysl.C:224-270): result_prim_<v>.<globiter>
        (v = 8 with Spitzer conduction), backup.dat and Header.<globiter> from
        the device state."""
        write_prim(dir_3d, resfile, run.globiter, self.grid, self.get_state())
        if self.phys.params[3] > 0.0:
            write_var(dir_3d, resfile, 8, run.globiter, self.grid, self.get_field(abi.SFLX))
        p_bc, s_bc = self.bc()
        write_backup(backfile, run.globiter, run.time, p_bc, s_bc)
        inv_va2max = self.limits()[0]
        va_max = 1.0 / np.sqrt(inv_va2max) if inv_va2max > 0 else 0.0
        write_header(dir_3d, run.globiter, self.grid, run.time, run.dt, va_max)

    def adjoint_forward(self, run, dt_rad: float = 1.0e30, nsteps: int = 1, every: int = 1):
        """nsteps steps as step(), recorded for adjoint_backward (checkpoint every `every` steps)."""
        self._chk(self.lib.mrm_adjoint_forward(self.h, C.byref(run), dt_rad, nsteps, every), "adjoint_forward")

    def adjoint_backward(self, seed: np.ndarray, fold_initial_ghosts: bool = True) -> np.ndarray:
        """dJ/dU before the recorded steps from seed = dJ/dU after them ([8, n2, n1, n0] float64)."""
        seed = np.ascontiguousarray(seed, dtype=np.float64)
        assert seed.size == 8 * self.bufsize
        grad = np.empty((8,) + self.shape, dtype=np.float64)
        self._chk(self.lib.mrm_adjoint_backward(self.h, seed.ctypes.data_as(C.POINTER(C.c_double)),
                                                grad.ctypes.data_as(C.POINTER(C.c_double)),
                                                1 if fold_initial_ghosts else 0), "adjoint_backward")
        return grad


# ---------------------------------------------------------------- restart files
def _chk_io(rc, what):
    if rc != 0:
        raise MrmError(f"{what}: {abi.ERRORS.get(rc, rc)}")


def read_prim(dir_3d: str, resfile: str, it: int, grid: abi.GridDesc, shape) -> np.ndarray:
    """result_prim_<v>.<it> -> conservative SoA state [8, n2, n1, n0] (ghosts zero)."""
    lib = abi.load()
    U = np.empty((8,) + tuple(shape), dtype=np.float64)
    _chk_io(lib.mrm_io_read_prim(dir_3d.encode(), resfile.encode(), it, C.byref(grid),
                                 U.ctypes.data_as(C.POINTER(C.c_double))), "read_prim")
    return U


def write_prim(dir_3d: str, resfile: str, it: int, grid: abi.GridDesc, U: np.ndarray):
    lib = abi.load()
    U = np.ascontiguousarray(U, dtype=np.float64)
    _chk_io(lib.mrm_io_write_prim(dir_3d.encode(), resfile.encode(), it, C.byref(grid),
                                  U.ctypes.data_as(C.POINTER(C.c_double))), "write_prim")


def read_var(dir_3d: str, resfile: str, v: int, it: int, grid: abi.GridDesc, shape) -> np.ndarray:
    """result_prim_<v>.<it> -> field [n2, n1, n0] (ghosts zero)"""
    f = np.empty(tuple(shape), dtype=np.float64)
    _chk_io(abi.load().mrm_io_read_var(dir_3d.encode(), resfile.encode(), v, it, C.byref(grid),
                                       f.ctypes.data_as(C.POINTER(C.c_double))), "read_var")
    return f


def write_var(dir_3d: str, resfile: str, v: int, it: int, grid: abi.GridDesc, f: np.ndarray):
    f = np.ascontiguousarray(f, dtype=np.float64)
    _chk_io(abi.load().mrm_io_write_var(dir_3d.encode(), resfile.encode(), v, it, C.byref(grid),
                                        f.ctypes.data_as(C.POINTER(C.c_double))), "write_var")


def write_header(dir_3d: str, it: int, grid: abi.GridDesc, time: float, dt: float, va_max: float):
    _chk_io(abi.load().mrm_io_write_header(dir_3d.encode(), it, C.byref(grid), time, dt, va_max), "write_header")


def read_header(dir_3d: str, it: int):
    """-> (gsize, dx, time, dt, va_max)"""
    g = (C.c_int * 3)()
    dx = (C.c_double * 3)()
    t, dt, va = C.c_double(), C.c_double(), C.c_double()
    _chk_io(abi.load().mrm_io_read_header(dir_3d.encode(), it, g, dx, C.byref(t), C.byref(dt), C.byref(va)),
            "read_header")
    return tuple(g), tuple(dx), t.value, dt.value, va.value


def write_backup(path: str, it: int, time: float, p_bc: float, s_bc: float):
    _chk_io(abi.load().mrm_io_write_backup(path.encode(), it, time, p_bc, s_bc), "write_backup")


def read_backup(path: str):
    """-> (globiter, time, p_bc, s_bc); p_bc, s_bc = -1 when the file holds only two numbers"""
    it = C.c_int()
    t, p, s_ = C.c_double(), C.c_double(), C.c_double()
    _chk_io(abi.load().mrm_io_read_backup(path.encode(), C.byref(it), C.byref(t), C.byref(p), C.byref(s_)),
            "read_backup")
    return it.value, t.value, p.value, s_.value
