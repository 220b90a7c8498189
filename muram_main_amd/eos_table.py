"""Tabulated equation of state in MURaM's binary table format.

The reference reads one binary file (eos_init.C:57-212):

* 14 float32 header values
  ``[N_eps, N_lr, N_lp3, N_s3, eps0, eps1, lr0, lr1, lp0, lp1, s0, s1, eps_off, ss_off]``
  (eos_init.C:80-102);
* six forward tables on the (ln(eps+eps_off), ln rho) grid, float32, stored with
  the eps index fastest (``ind = j*N_eps + i``, eos_init.C:161-177), in the order
  ln p, ln T, s+ss_off, ln n_e, ln rho_i, ln amb;
* two inverse tables on the (ln p, s+ss_off) grid, float32, ln p index fastest
  (eos_init.C:184-198), in the order ln(eps+eps_off), ln rho.

The physical table the reference ships is not in the tree, so this module
generates a smooth, thermodynamically plausible table of its own in exactly that
format: a hydrogen/helium mixture whose ionisation degree follows a smooth
function of (T, rho), with ionisation energy included in eps.  Parity between the
GPU path and the CPU restatement never depends on which table is used, only on
both reading the same bytes.

``load_tables`` performs the reference loader's transform (transpose to
``[i_eps][i_rho]`` and division by ``del_eps*del_lr`` before rounding back to
float32, eos_init.C:170-175 and :195-196), so every consumer sees identical
float32 tables.
"""
from __future__ import annotations

import dataclasses
import numpy as np

R_GAS = 8.314e7          # erg / (mol K), the constant cons_to_prim.C:57 uses
X_H = 0.7                # hydrogen mass fraction
MU_NEUTRAL = 1.21        # mean molecular weight of the neutral mixture
MU_IONISED = 0.60        # mean molecular weight when fully ionised
M_H = 1.67262158e-24     # g (cons_to_prim.C:24)
CHI_H = 2.18e-11         # erg, hydrogen ionisation energy
E_ION = X_H * CHI_H / M_H  # erg/g released per unit ionisation degree


def _ion_degree(T, rho):
    """Smooth ionisation degree in [0,1]; the threshold drifts with density."""
    t_ion = 9.0e3 * np.power(np.maximum(rho, 1e-300) / 1e-7, 0.03)
    return 0.5 * (1.0 + np.tanh((T - t_ion) / (0.15 * t_ion)))


def _mu(x):
    return 1.0 / ((1.0 - x) / MU_NEUTRAL + x / MU_IONISED)


def eps_of(T, rho):
    """Specific internal energy eps(T, rho) [erg/g]."""
    x = _ion_degree(T, rho)
    return 1.5 * R_GAS * T / _mu(x) + x * E_ION


def pres_of(T, rho):
    return rho * R_GAS * T / _mu(_ion_degree(T, rho))


def entropy_of(T, rho):
    """Specific entropy up to a constant (monotone in T at fixed rho and p)."""
    x = _ion_degree(T, rho)
    return (R_GAS / _mu(x)) * (1.5 * np.log(T) - np.log(rho)) + x * E_ION / T


def ne_of(T, rho):
    x = _ion_degree(T, rho)
    return (x * X_H + 1.0e-4) * rho / M_H


def rhoi_of(T, rho):
    return rho * (_ion_degree(T, rho) + 1.0e-4)


def amb_of(T, rho):
    x = _ion_degree(T, rho)
    return 1.0e14 * rho * (1.0 - x + 1.0e-6) * np.sqrt(T / 6000.0)


def _bisect_T_from_eps(eps, rho, iters=100):
    lo = np.full_like(eps, np.log(1.0))
    hi = np.full_like(eps, np.log(1.0e10))
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        f = eps_of(np.exp(mid), rho) - eps
        lo = np.where(f < 0, mid, lo)
        hi = np.where(f < 0, hi, mid)
    return np.exp(0.5 * (lo + hi))


def _invert_p_s(lnp, s, iters=80):
    """(ln p, s) -> (rho, T) by nested bisection: T outer, ln rho inner."""
    p = np.exp(lnp)
    lo_t = np.full_like(p, np.log(1.0))
    hi_t = np.full_like(p, np.log(1.0e10))

    def rho_at(T):
        lo = np.full_like(p, np.log(1e-40))
        hi = np.full_like(p, np.log(1e10))
        for _ in range(iters):
            mid = 0.5 * (lo + hi)
            f = pres_of(T, np.exp(mid)) - p
            lo = np.where(f < 0, mid, lo)
            hi = np.where(f < 0, hi, mid)
        return np.exp(0.5 * (lo + hi))

    for _ in range(iters):
        mid = 0.5 * (lo_t + hi_t)
        T = np.exp(mid)
        f = entropy_of(T, rho_at(T)) - s
        lo_t = np.where(f < 0, mid, lo_t)
        hi_t = np.where(f < 0, hi_t, mid)
    T = np.exp(0.5 * (lo_t + hi_t))
    return rho_at(T), T


@dataclasses.dataclass
class TableSpec:
    n_eps: int = 128
    n_lr: int = 128
    n_lp3: int = 64
    n_s3: int = 48
    eps_lo: float = 3.0e10      # eps + eps_off at the first eps node
    eps_hi: float = 5.0e16
    rho_lo: float = 1.0e-24
    rho_hi: float = 1.0e-3
    p_lo: float = 1.0e-6
    p_hi: float = 1.0e8
    s_lo: float = -2.0e9
    s_hi: float = 6.0e10
    eps_off: float = 2.0e10
    ss_off: float = 4.0e9


def generate(spec: TableSpec = TableSpec()) -> bytes:
    """Return the raw bytes of a table file in the reference's format."""
    hdr = np.array([spec.n_eps, spec.n_lr, spec.n_lp3, spec.n_s3,
                    np.log(spec.eps_lo), np.log(spec.eps_hi),
                    np.log(spec.rho_lo), np.log(spec.rho_hi),
                    np.log(spec.p_lo), np.log(spec.p_hi),
                    spec.s_lo + spec.ss_off, spec.s_hi + spec.ss_off,
                    spec.eps_off, spec.ss_off], dtype=np.float32)
    h = hdr.astype(np.float64)
    n_eps, n_lr, n_lp3, n_s3 = (int(v) for v in hdr[:4])
    eps_off = h[12]
    ss_off = h[13]
    # axes exactly as the loader builds them (eos_init.C:122-152)
    del_eps = (h[5] - h[4]) / float(n_eps - 1)
    del_lr = (h[7] - h[6]) / float(n_lr - 1)
    xeps = h[4] + del_eps * np.arange(n_eps, dtype=np.float64)
    xlr = h[6] + del_lr * np.arange(n_lr, dtype=np.float64)
    del_lp = (h[9] - h[8]) / float(n_lp3 - 1)
    del_s = (h[11] - h[10]) / float(n_s3 - 1)
    xlp = h[8] + del_lp * np.arange(n_lp3, dtype=np.float64)
    xs = h[10] + del_s * np.arange(n_s3, dtype=np.float64)

    # forward tables, array [j (rho), i (eps)] so that ravel() gives j*N_eps+i
    LE, LR = np.meshgrid(xeps, xlr)
    rho = np.exp(LR)
    eps = np.exp(LE) - eps_off
    T = _bisect_T_from_eps(eps, rho)
    fwd = [np.log(pres_of(T, rho)), np.log(T), entropy_of(T, rho) + ss_off,
           np.log(ne_of(T, rho)), np.log(rhoi_of(T, rho)), np.log(amb_of(T, rho))]

    # inverse tables, array [j (s), i (ln p)]
    LP, SS = np.meshgrid(xlp, xs)
    rho3, T3 = _invert_p_s(LP, SS - ss_off)
    inv = [np.log(eps_of(T3, rho3) + eps_off), np.log(rho3)]

    parts = [hdr] + [a.astype(np.float32).ravel() for a in fwd] + \
            [a.astype(np.float32).ravel() for a in inv]
    return b"".join(p.tobytes() for p in parts)


@dataclasses.dataclass
class Tables:
    """Tables after the reference loader's transform (eos_init.C:121-198)."""
    n_eps: int
    n_lr: int
    n_lp3: int
    n_s3: int
    eps_off: float
    ss_off: float
    del_eps: float
    del_lr: float
    del_lp3: float
    del_s3: float
    xeps: np.ndarray    # float64 [N_eps]
    xlr: np.ndarray     # float64 [N_lr]
    xlp3: np.ndarray    # float64 [N_lp3]
    xs3: np.ndarray     # float64 [N_s3]
    fwd: np.ndarray     # float32 [6, N_eps, N_lr]  (p, T, s, ne, rhoi, amb), pre-divided
    inv: np.ndarray     # float32 [2, N_lp3, N_s3]  (eps3, d3), pre-divided
    raw_header: np.ndarray  # float32 [14]
    raw_fwd: np.ndarray     # float32 [6*N_eps*N_lr] file order
    raw_inv: np.ndarray     # float32 [2*N_lp3*N_s3] file order


def load_tables(raw: bytes) -> Tables:
    buf = np.frombuffer(raw, dtype=np.float32)
    hdr = buf[:14].copy()
    n_eps, n_lr, n_lp3, n_s3 = (int(v) for v in hdr[:4])
    h = hdr.astype(np.float64)
    del_eps = (h[5] - h[4]) / float(n_eps - 1)
    del_lr = (h[7] - h[6]) / float(n_lr - 1)
    del_lp3 = (h[9] - h[8]) / float(n_lp3 - 1)
    del_s3 = (h[11] - h[10]) / float(n_s3 - 1)
    nf = 6 * n_eps * n_lr
    raw_fwd = buf[14:14 + nf].copy()
    raw_inv = buf[14 + nf:14 + nf + 2 * n_lp3 * n_s3].copy()
    # [t][j][i] -> [t][i][j], divided in double then rounded to float32
    f = raw_fwd.reshape(6, n_lr, n_eps).transpose(0, 2, 1).astype(np.float64)
    fwd = (f / (del_eps * del_lr)).astype(np.float32)
    g = raw_inv.reshape(2, n_s3, n_lp3).transpose(0, 2, 1).astype(np.float64)
    inv = (g / (del_lp3 * del_s3)).astype(np.float32)
    return Tables(n_eps, n_lr, n_lp3, n_s3, h[12], h[13], del_eps, del_lr, del_lp3, del_s3,
                  h[4] + del_eps * np.arange(n_eps, dtype=np.float64),
                  h[6] + del_lr * np.arange(n_lr, dtype=np.float64),
                  h[8] + del_lp3 * np.arange(n_lp3, dtype=np.float64),
                  h[10] + del_s3 * np.arange(n_s3, dtype=np.float64),
                  np.ascontiguousarray(fwd), np.ascontiguousarray(inv), hdr, raw_fwd, raw_inv)
