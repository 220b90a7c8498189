"""Test-side wrappers: the C oracle (oracle/liboracle.so) and the compiled
reference harness (oracle/_ref/muram_ref_run).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess
import tempfile

import numpy as np

from muram_main_amd.abi import BR_NSLOT, GridDesc, PhysDesc, RunDesc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_BIN = os.path.join(ROOT, "oracle", "_ref", "muram_ref_run")

dp = C.POINTER(C.c_double)


def _as_dp(a):
    return a.ctypes.data_as(dp)


class Oracle:
    """ctypes handle on one orc_ctx (SoA FP64 state)."""
    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(ORACLE_SO):
                subprocess.check_call(["make", "-C", os.path.join(ROOT, "oracle")])
            L = C.CDLL(ORACLE_SO)
            P = C.c_void_p
            L.orc_create.argtypes = [C.POINTER(GridDesc), C.POINTER(PhysDesc)]
            L.orc_create.restype = P
            L.orc_destroy.argtypes = [P]
            L.orc_bufsize.argtypes = [P]
            L.orc_bufsize.restype = C.c_long
            L.orc_eos_load.argtypes = [P, C.c_void_p, C.c_size_t]
            L.orc_state.argtypes = [P, C.c_int]
            L.orc_state.restype = dp
            L.orc_field.argtypes = [P, C.c_int]
            L.orc_field.restype = dp
            for n in ("orc_get_limits", "orc_set_limits"):
                getattr(L, n).argtypes = [P, dp]
            L.orc_get_bc.argtypes = [P, dp, dp]
            L.orc_set_bc.argtypes = [P, C.c_double, C.c_double]
            L.orc_set_threads.argtypes = [C.c_int]
            L.orc_cons_to_prim.argtypes = [P]
            L.orc_adjust_valf_max.argtypes = [P, C.POINTER(RunDesc), dp]
            L.orc_mhd_residual.argtypes = [P, C.POINTER(RunDesc)]
            L.orc_mhd_residual.restype = C.c_double
            L.orc_sync_timestep.argtypes = [P, C.POINTER(RunDesc), C.c_double, C.c_double, dp]
            L.orc_source_integrate_tcheck.argtypes = [P, C.POINTER(RunDesc), C.c_int]
            L.orc_save_cons.argtypes = [P]
            L.orc_tcheck.argtypes = [P, C.POINTER(RunDesc)]
            L.orc_tvdlimit.argtypes = [P, C.POINTER(RunDesc), C.c_double]
            L.orc_exchange_grid.argtypes = [P]
            L.orc_set_boundary_conditions.argtypes = [P, C.POINTER(RunDesc), C.c_int]
            L.orc_set_boundary_conditions_pt.argtypes = [P, C.POINTER(RunDesc), C.c_int, C.c_int]
            L.orc_potential_kernels.argtypes = [P, C.POINTER(C.c_float)]
            L.orc_potential_bext.argtypes = [P]
            L.orc_potential_bext.restype = dp
            L.orc_step.argtypes = [P, C.POINTER(RunDesc), C.c_double, C.c_int]
            L.orc_clean_divb.argtypes = [P, C.POINTER(RunDesc), dp]
            L.orc_diag_zero.argtypes = [P]
            L.orc_diag.argtypes = [P, C.c_int]
            L.orc_diag.restype = dp
            L.orc_branch_count.argtypes = [P, C.c_int, C.c_double]
            L.orc_get_branch_counts.argtypes = [P, C.POINTER(C.c_ulonglong)]
            for n in ("orc_p_interp", "orc_s_interp", "orc_T_interp", "orc_d3_interp", "orc_eps3_interp"):
                getattr(L, n).argtypes = [P, C.c_double, C.c_double]
                getattr(L, n).restype = C.c_double
            cls._lib = L
        return cls._lib

    def __init__(self, grid: GridDesc, phys: PhysDesc, table: bytes):
        L = self.lib()
        self.L = L
        self.grid = grid
        self.h = L.orc_create(C.byref(grid), C.byref(phys))
        self.bs = L.orc_bufsize(self.h)
        self.shape = tuple(int(grid.gsize[d] + 2 * grid.ghosts[d]) for d in (2, 1, 0))
        self._tab = C.create_string_buffer(table, len(table))
        rc = L.orc_eos_load(self.h, self._tab, len(table))
        assert rc == 0, rc

    def __del__(self):
        try:
            self.L.orc_destroy(self.h)
        except Exception:
            pass

    def state(self, which=0):
        p = self.L.orc_state(self.h, which)
        return np.ctypeslib.as_array(p, shape=(8 * self.bs,)).reshape((8,) + self.shape)

    def field(self, f):
        p = self.L.orc_field(self.h, f)
        return np.ctypeslib.as_array(p, shape=(self.bs,)).reshape(self.shape)

    def set_state(self, U, which=0):
        self.state(which)[...] = U

    def limits(self):
        a = np.zeros(4)
        self.L.orc_get_limits(self.h, _as_dp(a))
        return a

    def set_limits(self, lim):
        a = np.ascontiguousarray(lim, dtype=np.float64)
        self.L.orc_set_limits(self.h, _as_dp(a))

    def potential_kernels(self, k8):
        """the 8 kernels, (8, gsize[2], gsize[1]) float32"""
        self._k8 = np.ascontiguousarray(k8, dtype=np.float32)
        rc = self.L.orc_potential_kernels(self.h, self._k8.ctypes.data_as(C.POINTER(C.c_float)))
        assert rc == 0, rc

    def potential_bext(self):
        """b_ext, (8, n2, n1) copy"""
        n1, n2 = self.shape[1], self.shape[0]
        p = self.L.orc_potential_bext(self.h)
        return np.ctypeslib.as_array(p, shape=(8 * n1 * n2,)).reshape(8, n2, n1).copy()

    def set_boundary_conditions_pt(self, run, stage, pt_update):
        self.L.orc_set_boundary_conditions_pt(self.h, C.byref(run), stage, pt_update)

    def bc(self):
        p, s = C.c_double(), C.c_double()
        self.L.orc_get_bc(self.h, C.byref(p), C.byref(s))
        return p.value, s.value

    def set_bc(self, p, s):
        self.L.orc_set_bc(self.h, p, s)

    def diag(self):
        """tvar1..8, Qres, Qvis as a [10, nz, ny, nx] copy (zeros before any
        diagnostics were collected)"""
        out = np.zeros((10,) + self.shape)
        for t in range(10):
            p = self.L.orc_diag(self.h, t)
            if p:
                out[t] = np.ctypeslib.as_array(p, shape=(self.bs,)).reshape(self.shape)
        return out

    def branch_counts(self):
        """the MRM_BR_NSLOT count and checksum slots (uint64)"""
        out = np.zeros(BR_NSLOT, dtype=np.uint64)
        self.L.orc_get_branch_counts(self.h, out.ctypes.data_as(C.POINTER(C.c_ulonglong)))
        return out

    def clean_divb(self, run):
        err = np.zeros(2)
        self.L.orc_clean_divb(self.h, C.byref(run), _as_dp(err))
        return err

    def __getattr__(self, name):
        fn = getattr(self.L, "orc_" + name)
        return lambda *a: fn(self.h, *a)


# --------------------------------------------------------------------------
OP = dict(STARTUP=0, SAVECONS=1, C2P=2, VALF=3, RESID=4, SYNC=5, SRCINT=6, EXCHANGE=7,
          BC=8, TVD=9, TCHECK=10, STEP=11, ADVANCE=12, DIVB=13, DIAG=14, STEPDIAG=15)
DUMP = 0x100


def ref_available():
    return os.path.exists(REF_BIN)


def run_reference(grid: GridDesc, phys: PhysDesc, run: RunDesc, table: bytes, U: np.ndarray,
                  ops, p_bc=-1.0, s_bc=-1.0, dt_rad=1e30):
    """Run the compiled reference on (U, ops).  ops: list of (name, arg, dump).
    Returns a list of dumps: dict(op, sc, U, Res, fields)."""
    shape = U.shape
    with tempfile.TemporaryDirectory() as td:
        tpath = os.path.join(td, "eos.dat")
        with open(tpath, "wb") as f:
            f.write(table)
        case = os.path.join(td, "case.bin")
        with open(case, "wb") as f:
            f.write(struct.pack("<i", 0x4D524D31))
            f.write(bytes(grid))
            f.write(bytes(phys))
            f.write(bytes(run))
            f.write(struct.pack("<3d", p_bc, s_bc, dt_rad))
            f.write(tpath.encode().ljust(256, b"\0"))
            f.write(struct.pack("<i", len(ops)))
            for name, arg, dump in ops:
                f.write(struct.pack("<2i", OP[name] | (DUMP if dump else 0), arg))
            f.write(np.ascontiguousarray(U, dtype=np.float64).tobytes())
        out = os.path.join(td, "out.bin")
        r = subprocess.run([REF_BIN, case, out], cwd=td, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference harness failed: {r.stdout}\n{r.stderr}")
        raw = open(out, "rb").read()
    ndump, bs = struct.unpack_from("<2i", raw, 0)
    off = 8
    dumps = []
    for _ in range(ndump):
        (o,) = struct.unpack_from("<i", raw, off)
        off += 4
        sc = np.frombuffer(raw, np.float64, 16, off).copy()
        off += 128
        Uo = np.frombuffer(raw, np.float64, 8 * bs, off).reshape((8,) + shape[1:]).copy()
        off += 64 * bs
        Ro = np.frombuffer(raw, np.float64, 8 * bs, off).reshape((8,) + shape[1:]).copy()
        off += 64 * bs
        Fo = np.frombuffer(raw, np.float64, 5 * bs, off).reshape((5,) + shape[1:]).copy()
        off += 40 * bs
        Do = np.frombuffer(raw, np.float64, 3 * bs, off).reshape((3,) + shape[1:]).copy()
        off += 24 * bs
        Go = np.frombuffer(raw, np.float64, 10 * bs, off).reshape((10,) + shape[1:]).copy()
        off += 80 * bs
        dumps.append(dict(op=o, sc=sc, U=Uo, Res=Ro, fields=Fo, divB=Do[0], phi=Do[1], sflx=Do[2],
                          diag=Go))
    return dumps


def relerr(a, b):
    """max |a-b| / max|b| per field (the parity metric of the tests)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.max(np.abs(b))
    if den == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / den)


FIELD_NAMES = ("rho", "M0", "M1", "M2", "e", "B0", "B1", "B2")


def field_errors(a, b, interior=None):
    """Per-field relative L2 (||a-b||_2 / ||b||_2) and Linf (max|a-b| / max|b|)
    errors of state a [8, n2, n1, n0] against b, optionally over a slice."""
    out = {}
    for v, name in enumerate(FIELD_NAMES):
        x = np.asarray(a[v], dtype=np.float64)
        y = np.asarray(b[v], dtype=np.float64)
        if interior is not None:
            x, y = x[interior], y[interior]
        d = x - y
        ny2 = float(np.sqrt(np.sum(y * y)))
        nyi = float(np.max(np.abs(y)))
        out[name] = {"l2": float(np.sqrt(np.sum(d * d))) / ny2 if ny2 > 0 else float(np.sqrt(np.sum(d * d))),
                     "linf": float(np.max(np.abs(d))) / nyi if nyi > 0 else float(np.max(np.abs(d)))}
    return out
