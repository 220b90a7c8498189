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

    def exchange_grid(self):
        self._chk(self.lib.mrm_exchange_grid(self.h), "exchange_grid")

    def SetBoundaryConditions(self, run, stage: int):
        self._chk(self.lib.mrm_set_boundary_conditions(self.h, C.byref(run), stage),
                  "SetBoundaryConditions")

    def step(self, run, dt_rad: float = 1.0e30, nsteps: int = 1):
        self._chk(self.lib.mrm_step(self.h, C.byref(run), dt_rad, nsteps), "step")

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
