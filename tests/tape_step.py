"""FP64 restatement of one MURaM RK step in PyTorch -- test infrastructure.

Its autograd tape is the reference FP64 tape gradient the hand-written
adjoint (mrm_adjoint_forward/backward) is checked against.  The step is the
one the adjoint differentiates: va_adjust = 0, no divB cleaning, no
resistivity or conduction, RT off (Qtot = Tau = 0), dt fixed, one TVD call
(cfl <= CFL_tvd), p_bc and s_bc fixed.  Every piece follows the oracle's
function of the same name in oracle/muram_oracle.c, which is bit-identical
to the compiled reference; the forward matches the oracle to 1e-12 in
tests/test_tape_step.py:
  ConsToPrim             cons_to_prim.C:39-236
  EOS interpolations     eos_init.C:216-416
  MHD_Residual           mhdres_SR.C:27-854
  Source_Integrate_Tcheck Add_Sources_Integrate_SR.C:38-472
  exchange_grid_acc      exchange.C:532-782 (one rank)
  SetBoundaryConditions  boundary_pdmp_1_fftw3.C:81-493 (bnd_pot = 0)
  TVDlimit               tvdlimit_SR.C:33-695
  TCheck                 Add_Sources_Integrate_SR.C:534-625
State layout: U[v, k, j, i] (dim 0, the vertical, last), ghosts included.
"""
from __future__ import annotations

import math

import numpy as np
import torch

NV = 8
EOS_GAMMA, EOS_MU, EOS_MU_N = 1.65, 0.62, 1.3
MH, KB, RGAS = 1.67262158e-24, 1.380622e-16, 8.314e7


class Tables:
    """the EOS tables as orc_eos_load builds them (eos_init.C:57-212)"""

    def __init__(self, raw: bytes):
        buf = np.frombuffer(raw, dtype=np.float32)
        n_eps, n_lr, n_lp3, n_s3 = (int(v) for v in buf[:4])
        h = buf[:14].astype(np.float64)
        self.n_eps, self.n_lr, self.n_lp3, self.n_s3 = n_eps, n_lr, n_lp3, n_s3
        self.eps_off, self.ss_off = float(h[12]), float(h[13])
        del_eps = (h[5] - h[4]) / float(n_eps - 1)
        del_lr = (h[7] - h[6]) / float(n_lr - 1)
        del_lp3 = (h[9] - h[8]) / float(n_lp3 - 1)
        del_s3 = (h[11] - h[10]) / float(n_s3 - 1)
        self.inv_del_eps, self.inv_del_lr = 1.0 / del_eps, 1.0 / del_lr
        self.inv_del_lp3, self.inv_del_s3 = 1.0 / del_lp3, 1.0 / del_s3
        self.xeps = torch.tensor(h[4] + del_eps * np.arange(n_eps, dtype=np.float64))
        self.xlr = torch.tensor(h[6] + del_lr * np.arange(n_lr, dtype=np.float64))
        self.xlp3 = torch.tensor(h[8] + del_lp3 * np.arange(n_lp3, dtype=np.float64))
        self.xs3 = torch.tensor(h[10] + del_s3 * np.arange(n_s3, dtype=np.float64))
        nt = n_eps * n_lr
        fw = buf[14:14 + 6 * nt].astype(np.float64).reshape(6, n_lr, n_eps)
        sc = del_eps * del_lr
        # [t, i_eps, i_rho], divided by del_eps del_lr and rounded to float
        self.tab = torch.tensor((fw / sc).astype(np.float32).astype(np.float64).transpose(0, 2, 1).copy())
        ni = n_lp3 * n_s3
        iv = buf[14 + 6 * nt:14 + 6 * nt + 2 * ni].astype(np.float64).reshape(2, n_s3, n_lp3)
        sci = del_lp3 * del_s3
        self.itab = torch.tensor((iv / sci).astype(np.float32).astype(np.float64).transpose(0, 2, 1).copy())
        self.eps_max = math.exp(float(self.xeps[-1])) - self.eps_off


def _index(x, x0, inv, n):
    """(int)((x - x0) inv) clamped to [0, n - 2]; the cast truncates toward 0"""
    i = ((x.detach() - x0) * inv).trunc().to(torch.int64)
    return i.clamp(0, n - 2)


class TapeStep:
    def __init__(self, grid, phys, table: bytes, p_bc: float, s_bc: float):
        self.T = Tables(table)
        self.gsize = [int(grid.gsize[d]) for d in range(3)]
        self.gh = 2
        self.n = [g + 4 for g in self.gsize]                 # (n0, n1, n2)
        self.lb = [2, 2, 2]
        self.le = [g + 1 for g in self.gsize]
        self.dx = [(grid.gxmax[d] - grid.gxmin[d]) / grid.gsize[d] for d in range(3)]
        self.gxmin = [grid.gxmin[d] for d in range(3)]
        self.gxmax = [grid.gxmax[d] for d in range(3)]
        self.ph = phys
        self.p_bc, self.s_bc = p_bc, s_bc
        # Adjust_Valf_Max with va_adjust = 0 (Adjust_Valf_Max.C:136-144)
        self.inv_va2max = 1.0 / math.pow(phys.params[1], 2)
        self.v_lim, self.e_lim = phys.tchk[3], phys.tchk[2]
        self._tvd_profiles()

    # ------------------------------------------------------------ EOS
    def _fwd(self, t, ee1, logr):
        T = self.T
        i = _index(ee1, float(T.xeps[0]), T.inv_del_eps, T.n_eps)
        j = _index(logr, float(T.xlr[0]), T.inv_del_lr, T.n_lr)
        xe0, xe1 = T.xeps[i], T.xeps[i + 1]
        xr0, xr1 = T.xlr[j], T.xlr[j + 1]
        tab = T.tab[t]
        return ((logr - xr0) * (ee1 - xe0) * tab[i + 1, j + 1] + (logr - xr0) * (xe1 - ee1) * tab[i, j + 1]
                + (xr1 - logr) * (ee1 - xe0) * tab[i + 1, j] + (xr1 - logr) * (xe1 - ee1) * tab[i, j])

    def p_interp(self, ee, dd):
        return torch.exp(self._fwd(0, torch.log(ee + self.T.eps_off), torch.log(dd)))

    def s_interp(self, ee, dd):
        return self._fwd(2, torch.log(ee + self.T.eps_off), torch.log(dd)) - self.T.ss_off

    def _inv(self, t, pp, ss):
        T = self.T
        ss1 = ss + T.ss_off
        logp = torch.log(pp)
        i = _index(logp, float(T.xlp3[0]), T.inv_del_lp3, T.n_lp3)
        j = _index(ss1, float(T.xs3[0]), T.inv_del_s3, T.n_s3)
        x0, x1 = T.xlp3[i], T.xlp3[i + 1]
        s0, s1 = T.xs3[j], T.xs3[j + 1]
        tab = T.itab[t]
        return ((logp - x0) * (ss1 - s0) * tab[i + 1, j + 1] + (logp - x0) * (s1 - ss1) * tab[i + 1, j]
                + (x1 - logp) * (ss1 - s0) * tab[i, j + 1] + (x1 - logp) * (s1 - ss1) * tab[i, j])

    def d3_interp(self, pp, ss):
        return torch.exp(self._inv(1, pp, ss))

    def eps3_interp(self, pp, ss):
        return torch.exp(self._inv(0, pp, ss)) - self.T.eps_off

    def cons_to_prim(self, U):
        """primitive state W (rho, v, e_int, B) and p over the whole buffer"""
        T = self.T
        d = U[0]
        dn = 1.0 / d
        m1, m2, m3 = U[1] * dn, U[2] * dn, U[3] * dn
        vv = m1 * m1 + m2 * m2 + m3 * m3
        e = U[4] - 0.5 * vv * d
        eps = e * dn + T.eps_off
        lgr = torch.log(d)
        ep = torch.log(eps)
        i_d = _index(lgr, float(T.xlr[0]), T.inv_del_lr, T.n_lr)
        i_e = _index(ep, float(T.xeps[0]), T.inv_del_eps, T.n_eps)
        c0 = (ep - T.xeps[i_e]) * (lgr - T.xlr[i_d])
        c1 = (ep - T.xeps[i_e]) * (T.xlr[i_d + 1] - lgr)
        c2 = (T.xeps[i_e + 1] - ep) * (lgr - T.xlr[i_d])
        c3 = (T.xeps[i_e + 1] - ep) * (T.xlr[i_d + 1] - lgr)
        tab = T.tab
        r0 = c0 * tab[0][i_e + 1, i_d + 1]
        r0 = r0 + c1 * tab[0][i_e + 1, i_d]
        r0 = r0 + c2 * tab[0][i_e, i_d + 1]
        r0 = r0 + c3 * tab[0][i_e, i_d]
        pres = torch.exp(r0)
        c_pres = EOS_GAMMA - 1.0
        ideal = c_pres * e
        pres = torch.where(lgr < float(T.xlr[0]), ideal, pres)
        xx = torch.clamp(5.0 * (T.eps_max - eps) / T.eps_max, min=0.0)
        pres = torch.where(eps > 0.8 * T.eps_max, pres * xx + (1.0 - xx) * ideal, pres)
        bad = ~torch.isfinite(pres) | (pres <= 0.0)
        pres = torch.where(bad, ideal, pres)
        # (temperature and n_e do not reach the step)
        W = torch.stack([d, m1, m2, m3, e, U[5], U[6], U[7]])
        return W, pres

    # ------------------------------------------------------------ residual
    def residual(self, W, p):
        """MHD_Residual over the interior (mhdres_SR.C:27-854), sweeps 0, 1, 2"""
        lb, le = self.lb, self.le
        inner = (slice(lb[2], le[2] + 1), slice(lb[1], le[1] + 1), slice(lb[0], le[0] + 1))
        R = None
        for d1 in range(3):
            d2 = 1 if d1 == 0 else 0
            d3 = 1 if d1 == 2 else 2
            ax = 3 - d1                        # tensor axis of dim d1 in U[v, k, j, i]
            dx = self.dx[d1]
            w1, w2 = 8. / (12. * dx), -1. / (12. * dx)
            var0 = W[0]
            vel, mag = W[1 + d1], W[5 + d1]
            dn = 1.0 / var0
            vv = W[1] * W[1] + W[2] * W[2] + W[3] * W[3]
            bb = W[5] * W[5] + W[6] * W[6] + W[7] * W[7]
            ekin = 0.5 * var0 * vv
            pmag = 0.5 * bb
            mflx = vel * var0
            pd = [p if q == d1 else None for q in range(3)]
            flx = [mflx,
                   mflx * W[1] + (p if d1 == 0 else 0.0 * p),
                   mflx * W[2] + (p if d1 == 1 else 0.0 * p),
                   mflx * W[3] + (p if d1 == 2 else 0.0 * p),
                   vel * (W[4] + ekin + p),
                   vel * W[5] - W[1] * mag,
                   vel * W[6] - W[2] * mag,
                   vel * W[7] - W[3] * mag]
            del pd
            ms = [mag * W[5] - (pmag if d1 == 0 else 0.0 * pmag),
                  mag * W[6] - (pmag if d1 == 1 else 0.0 * pmag),
                  mag * W[7] - (pmag if d1 == 2 else 0.0 * pmag)]

            def sh(a, s):
                """a at offset s along dim d1, over the interior"""
                idx = list(inner)
                sl = idx[ax - 1]
                idx[ax - 1] = slice(sl.start + s, sl.stop + s)
                return a[tuple(idx)]

            def ddx(a):
                return w1 * (sh(a, 1) - sh(a, -1)) + w2 * (sh(a, 2) - sh(a, -2))

            res = [0.0 - ddx(f) for f in flx]
            dB2 = ddx(W[5 + d2])
            dB3 = ddx(W[5 + d3])
            B1c, B2c, B3c = W[5 + d1][inner], W[5 + d2][inner], W[5 + d3][inner]
            cb = [None] * 3
            cb[d1] = -B2c * dB2 - B3c * dB3
            cb[d2] = B1c * dB2
            cb[d3] = B1c * dB3
            lrt = [ddx(m) for m in ms]
            va2 = bb[inner] * dn[inner]
            x2 = va2 * self.inv_va2max
            x4 = x2 * x2
            s = 1.0 + x2
            lf = s / (s + x4)
            lrt = [lf * lrt[q] + (1.0 - lf) * cb[q] for q in range(3)]
            res[1] = res[1] + lrt[0]
            res[2] = res[2] + lrt[1]
            res[3] = res[3] + lrt[2]
            res[4] = res[4] + (W[1][inner] * lrt[0] + W[2][inner] * lrt[1] + W[3][inner] * lrt[2])
            r = torch.stack(res)
            R = r if R is None else R + r
        return R

    # ------------------------------------------------------------ sources
    def source_integrate(self, U, W, R, U0, stage, dt):
        """Source_Integrate_Tcheck (Add_Sources_Integrate_SR.C:331-465), RT off"""
        lb, le = self.lb, self.le
        inner = (slice(None), slice(lb[2], le[2] + 1), slice(lb[1], le[1] + 1), slice(lb[0], le[0] + 1))
        w = W[inner]
        wdt = dt / float(4 + 1 - stage)
        g = -self.ph.params[0]
        dn = g * w[0]
        R1 = R[1] + dn
        R4 = R[4] + dn * w[1] + 0.0
        rh, rd = w[0], R[0]
        vx, vy, vz = w[1], w[2], w[3]
        bx, by, bz = w[5], w[6], w[7]
        bb = bx * bx + by * by + bz * bz
        va2 = bb / rh
        x2 = va2 * self.inv_va2max
        x4 = x2 * x2
        bf = x4 / (1 + x2 + x4)
        fx = R1 - vx * rd
        fy = R[2] - vy * rd
        fz = R[3] - vz * rd
        dd = (fx * bx + fy * by + fz * bz) / torch.clamp(bb, min=1e-100)
        b0 = -bf * (fx - dd * bx)
        b1 = -bf * (fy - dd * by)
        b2 = -bf * (fz - dd * bz)
        b3 = b0 * vx + b1 * vy + b2 * vz
        Rn = torch.stack([R[0], R1 + b0, R[2] + b1, R[3] + b2, R4 + b3, R[5], R[6], R[7]])
        un = U0[inner] + wdt * Rn
        un = self._tcheck_cells(un)
        out = U.clone()
        out[inner] = un
        return out

    def _tcheck_cells(self, u):
        """Tcheck clamp of each cell (Add_Sources_Integrate_SR.C:447-465)"""
        rho_min, eps_min = self.ph.tchk[1], self.ph.tchk[0]
        vmax = self.v_lim if self.v_lim > 0.0 else self.ph.tchk[3]
        eps_max = self.e_lim if self.e_lim > 0.0 else self.ph.tchk[2]
        dn = torch.clamp(u[0], min=rho_min)
        vv = torch.sqrt(u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / dn
        s = vmax / torch.clamp(vv, min=vmax)
        ekin = 0.5 * dn * vv * vv
        ee_low = eps_min * dn + ekin * s * s
        ee_up = eps_max * dn + ekin * s * s
        e = torch.minimum(torch.maximum(u[4], ee_low), ee_up)
        return torch.stack([dn, u[1] * s, u[2] * s, u[3] * s, e, u[5], u[6], u[7]])

    def tcheck(self, U):
        lb, le = self.lb, self.le
        inner = (slice(None), slice(lb[2], le[2] + 1), slice(lb[1], le[1] + 1), slice(lb[0], le[0] + 1))
        out = U.clone()
        out[inner] = self._tcheck_cells(U[inner])
        return out

    # ------------------------------------------------------------ ghosts
    def exchange(self, U):
        """periodic dims 1 then 2 on one rank (exchange.C:532-782)"""
        lb, le, gh = self.lb, self.le, self.gh
        out = U.clone()
        i_in = slice(lb[0], le[0] + 1)
        k_in = slice(lb[2], le[2] + 1)
        # dim 1 at interior (i, k)
        out[:, k_in, le[1] + 1:le[1] + 1 + gh, i_in] = U[:, k_in, lb[1]:lb[1] + gh, i_in]
        out[:, k_in, lb[1] - gh:lb[1], i_in] = U[:, k_in, le[1] - gh + 1:le[1] + 1, i_in]
        out2 = out.clone()
        # dim 2 at interior i, all j
        out2[:, le[2] + 1:le[2] + 1 + gh, :, i_in] = out[:, lb[2]:lb[2] + gh, :, i_in]
        out2[:, lb[2] - gh:lb[2], :, i_in] = out[:, le[2] - gh + 1:le[2] + 1, :, i_in]
        return out2

    def bc(self, U):
        """SetBoundaryConditions with fixed p_bc, s_bc (boundary_pdmp_1_fftw3.C:172-474)"""
        lb, le = self.lb, self.le
        ph = self.ph
        open_top = int(ph.bnd[0])
        B_crit = ph.bnd[2] / math.sqrt(8.0 * math.asin(1.0))
        eps_top = ph.bnd[3]
        p1_dmp = 0.8
        # layer means of the interior (j, k)
        jk = (slice(lb[2], le[2] + 1), slice(lb[1], le[1] + 1))
        L1 = U[(slice(None),) + jk + (lb[0],)]
        L2 = U[(slice(None),) + jk + (lb[0] + 1,)]
        e1 = L1[4] - 0.5 * ((L1[1] * L1[1] + L1[2] * L1[2] + L1[3] * L1[3]) / L1[0])
        e2 = L2[4] - 0.5 * ((L2[1] * L2[1] + L2[2] * L2[2] + L2[3] * L2[3]) / L2[0])
        cc = 1.0 / float(self.gsize[1] * self.gsize[2])
        pmean1 = self.p_interp(e1 / L1[0], L1[0]).sum() * cc
        pmean2 = self.p_interp(e2 / L2[0], L2[0]).sum() * cc
        # bottom ghosts, every (j, k)
        u1 = U[:, :, :, lb[0]]
        u2 = U[:, :, :, lb[0] + 1]
        ei1 = u1[4] - 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3]) / u1[0]
        ei2 = u2[4] - 0.5 * (u2[1] * u2[1] + u2[2] * u2[2] + u2[3] * u2[3]) / u2[0]
        eps1, eps2 = ei1 / u1[0], ei2 / u2[0]
        pmag = 0.5 * u1[5] * u1[5]
        pmin1 = self.p_interp(eps1, u1[0]) - pmean1
        smin1 = self.s_interp(eps1, u1[0])
        smin2 = self.s_interp(eps2, u2[0])
        inflow = u1[1] >= 0.0
        s1 = torch.where(inflow, torch.full_like(smin1, self.s_bc), smin1)
        s2 = torch.where(inflow, torch.full_like(smin2, self.s_bc), smin2)
        cfac = self.p_bc / torch.sqrt(pmean1 * pmean2)
        p1 = pmean1 * cfac
        p2 = pmean1 * cfac * cfac
        p1 = p1 + pmin1 * p1_dmp
        p2 = p2 + pmin1 * p1_dmp * p1_dmp
        p1 = p1 - pmag * (1.0 - p1_dmp)
        p2 = p2 - pmag * (1.0 - p1_dmp * p1_dmp)
        cs = (u1[5] / B_crit) ** 4
        cs = (1.0 - cs) / (1.0 + cs)
        m1 = [u1[q] * cs for q in (1, 2, 3)]
        m2 = [u2[q] * cs for q in (1, 2, 3)]
        rho1 = self.d3_interp(p1, s1)
        rho2 = self.d3_interp(p2, s2)
        ep1 = self.eps3_interp(p1, s1)
        ep2 = self.eps3_interp(p2, s2)
        en1 = ep1 * rho1 + 0.5 * (m1[0] * m1[0] + m1[1] * m1[1] + m1[2] * m1[2]) / rho1
        en2 = ep2 * rho2 + 0.5 * (m2[0] * m2[0] + m2[1] * m2[1] + m2[2] * m2[2]) / rho2
        g1 = torch.stack([rho1, m1[0], m1[1], m1[2], en1, u1[5], u1[6], u1[7]])
        g2 = torch.stack([rho2, m2[0], m2[1], m2[2], en2, u2[5], u2[6], u2[7]])
        out = U.clone()
        out[:, :, :, lb[0] - 1] = g1
        out[:, :, :, lb[0] - 2] = g2
        # top ghosts (:402-474), vertical field
        t1 = U[:, :, :, le[0]]
        t2 = U[:, :, :, le[0] - 1]
        if open_top == 1:
            flip = t1[1] < 0.0
            mx1 = torch.where(flip, -t1[1], t1[1])
            mx2 = torch.where(flip, -t2[1], t2[1])
        else:
            mx1, mx2 = -t1[1], -t2[1]
        e_t1, e_t2 = t1[4], t2[4]
        if eps_top > 0.0:
            e_t1 = eps_top * t1[0] + 0.5 * (mx1 * mx1 + t1[2] * t1[2] + t1[3] * t1[3]) / t1[0]
            e_t2 = eps_top * t2[0] + 0.5 * (mx2 * mx2 + t2[2] * t2[2] + t2[3] * t2[3]) / t2[0]
        out[:, :, :, le[0] + 1] = torch.stack([t1[0], mx1, t1[2], t1[3], e_t1, t1[5], -t1[6], -t1[7]])
        out[:, :, :, le[0] + 2] = torch.stack([t2[0], mx2, t2[2], t2[3], e_t2, t2[5], -t2[6], -t2[7]])
        return out

    # ------------------------------------------------------------ TVD
    def _tvd_profiles(self):
        """hfb, hft, qft of tvdlimit_SR.C:132-149 over dim 0"""
        P = self.ph
        h_bnd_bot = P.tvd_h_bnd[0] if P.tvd_h_bnd[0] < 1 else P.tvd_h_bnd[0] / float(self.gsize[0])
        h_bnd_top = P.tvd_h_bnd[1] if P.tvd_h_bnd[1] < 1 else P.tvd_h_bnd[1] / float(self.gsize[0])
        hfb, hft, qft = [], [], []
        for i in range(self.n[0]):
            coord = self.gxmin[0] + (0.5 + (i - self.lb[0])) * self.dx[0]
            hh = coord / self.gxmax[0]
            vv = max(0.0, 1.0 - 0.5 * (1.0 - hh) / h_bnd_top)
            hft.append(min(1.0, vv * vv))
            vv = max(0.0, 1.0 - 0.5 * hh / h_bnd_bot)
            hfb.append(min(1.0, vv * vv))
            qft.append(1.0 if hh < 1 - 2 * h_bnd_top else P.tvd[5])
        self.hfb = torch.tensor(hfb, dtype=torch.float64)
        self.hft = torch.tensor(hft, dtype=torch.float64)
        self.qft = torch.tensor(qft, dtype=torch.float64)

    def tvd(self, U, p, dt_tvd):
        """TVDlimit (tvdlimit_SR.C:33-695) with Tau = 0 and no ambipolar term"""
        P = self.ph
        rho_min, eps_min = P.tchk[1], P.tchk[0]
        tvd_h, tvd_cs = list(P.tvd_h), list(P.tvd_cs)
        rho_lev = math.pow(P.tvd[0], 2)
        rho_log = int(P.tvd[1])
        q_rho_max = math.log(P.tvd[2])
        visc_slope_bot = max(0.0, P.tvd_visc_bnd[0] - 1.0)
        visc_slope_top = max(0.0, P.tvd_visc_bnd[1] - 1.0)
        eta_slope_bot = max(0.0, P.tvd_eta_bnd[0] - 1.0)
        eta_slope_top = max(0.0, P.tvd_eta_bnd[1] - 1.0)
        visc_coeff_bot = min(1.0, P.tvd_visc_bnd[0])
        visc_coeff_top = min(1.0, P.tvd_visc_bnd[1])
        eta_coeff_bot = min(1.0, P.tvd_eta_bnd[0])
        eta_coeff_top = min(1.0, P.tvd_eta_bnd[1])
        B_par_diff, vhyp = P.tvd[3], P.tvd[4]
        tvd_pm_v, tvd_pm_B, vmax_lim, CME_thresh = P.tvd[6], P.tvd[7], P.tvd[8], P.tvd[9]
        tc = list(P.tvd_coeff)
        need_tvd_coeff = tc[0] * tc[1] * tc[2] * tc[3] != 1.0
        vsqr_diff = math.pow(vmax_lim * self.v_lim, 2.0) if vmax_lim < 1.0 else math.pow(vmax_lim, 2.0)
        lb, le, n, gh = self.lb, self.le, self.n, self.gh
        bounds = [[0, n[d] - 1] for d in range(3)]
        for d1 in range(3):
            d2 = 1 if d1 == 0 else 0
            d3 = 1 if d1 == 2 else 2
            cmax = 0.975 * self.dx[d1] / dt_tvd
            idxd1 = 1.0 / self.dx[d1]
            # lines: the transverse ranges of this sweep; the line axis last
            sel = [slice(bounds[d][0], bounds[d][1] + 1) for d in range(3)]
            Xs = U[:, sel[2], sel[1], sel[0]]
            Ps = p[sel[2], sel[1], sel[0]]
            perm = {0: (0, 1, 2, 3), 1: (0, 1, 3, 2), 2: (0, 2, 3, 1)}[d1]
            X = Xs.permute(*perm)
            Pl = Ps.permute(*[q - 1 for q in perm[1:]])
            sz = X.shape[-1]
            # position profiles: along the line for sweep 0, per line otherwise
            if d1 == 0:
                hfb1, hft1, qft1 = self.hfb, self.hft, self.qft
            else:
                ii = torch.arange(bounds[0][0], bounds[0][1] + 1)
                shape = [1] * (X.dim() - 1)
                ax_i = 2 if d1 == 1 else 2
                shape[ax_i - 1] = len(ii)
                shape = shape[:-1] + [1]
                hfb1 = self.hfb[ii].reshape(shape)
                hft1 = self.hft[ii].reshape(shape)
                qft1 = self.qft[ii].reshape(shape)
            var0 = torch.clamp(X[0], min=rho_min)
            dn = 1.0 / var0
            v = [X[1] * dn, X[2] * dn, X[3] * dn]
            vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2]
            bb = X[5] * X[5] + X[6] * X[6] + X[7] * X[7]
            e4 = X[4] * dn - 0.5 * vv
            vsqr, bsqr = vv, bb
            vabs = torch.sqrt(vv)
            va2 = bb * dn
            x2 = va2 * self.inv_va2max
            x4 = x2 * x2
            s = 1.0 + x2
            lf = s / (s + x4)
            cs2 = 5.0 / 3.0 * Pl * dn
            cfast = torch.sqrt(va2 + cs2)
            CME_mode = (vabs >= CME_thresh * cfast).to(torch.float64)     # CZ = 0 (Tau = 0)
            dsq = var0 * var0
            rf = torch.minimum(lf, dsq / (dsq + rho_lev))
            hh = 1.0 - hfb1 - hft1
            cf = hfb1 * tvd_cs[0] + (rf * tvd_cs[1] + (1.0 - rf) * tvd_cs[2]) * hh + hft1 * tvd_cs[3]
            cf = torch.maximum(cf, CME_mode)
            cs2 = cs2 * (cf * cf)
            cm = 0.0 + vabs + torch.sqrt(torch.maximum(cs2, lf * (cs2 + va2)))
            hyp_e = hfb1 * tvd_h[0] + (rf * tvd_h[1] + (1.0 - rf) * tvd_h[2]) * hh + hft1 * tvd_h[3]
            hyp_v = (hfb1 * tvd_h[0] + (rf * tvd_h[1] + (1.0 - rf) * tvd_h[2] * tvd_pm_v) * hh
                     + hft1 * tvd_h[3] * tvd_pm_v)
            hyp_B = (hfb1 * tvd_h[0] + (rf * tvd_h[1] + (1.0 - rf) * tvd_h[2] * tvd_pm_B) * hh
                     + hft1 * tvd_h[3] * tvd_pm_B)
            hyp_e = hyp_e * (1.0 - CME_mode)
            hyp_v = hyp_v * (1.0 - CME_mode)
            BC2 = 1.0 - lf
            hyperdiff = None
            if d1 == 0 and vhyp > 0.0:
                hd = vhyp * torch.sqrt(vsqr)
                hyperdiff = 0.5 * idxd1 * torch.clamp(torch.minimum(hd, cmax - cm), min=0.0)
            var = [var0, v[0], v[1], v[2], e4, X[5], X[6], X[7]]
            if rho_log:
                var[0] = torch.log(var[0])
                var[4] = torch.log(var[4])
                qrho = torch.abs(var[0][..., :-1] - var[0][..., 1:])
            else:
                qrho = torch.abs(torch.log(var[0][..., :-1]) - torch.log(var[0][..., 1:]))
            # slopes at 1 .. sz-2
            slp = []
            for q in range(NV):
                Xq = var[q]
                sl = Xq[..., 1:-1] - Xq[..., :-2]
                sr = Xq[..., 2:] - Xq[..., 1:-1]
                slm = 0.25 * (sl + sr)
                lower = torch.minimum(torch.minimum(sl, sr), slm)
                upper = torch.maximum(torch.maximum(sl, sr), slm)
                slp.append(torch.clamp(lower, min=0.0) + torch.clamp(upper, max=0.0))

            def mid(a):   # a[..., 1:-1] for per-point arrays (or the profile for sweep 0)
                return a[..., 1:-1] if (isinstance(a, torch.Tensor) and a.shape[-1] == sz) else a

            if visc_slope_bot > 0.0 or visc_slope_top > 0.0:
                c = (1.0 - visc_slope_bot * mid(hfb1)) * (1.0 - visc_slope_top * mid(hft1))
                for q in (1, 2, 3):
                    slp[q] = slp[q] * c
            if eta_slope_bot > 0.0 or eta_slope_top > 0.0:
                c = (1.0 - eta_slope_bot * mid(hfb1)) * (1.0 - eta_slope_top * mid(hft1))
                for q in (5, 6, 7):
                    slp[q] = slp[q] * c
            slp[0] = torch.where(qrho[..., 1:] > q_rho_max, torch.zeros_like(slp[0]), slp[0])
            if vmax_lim > 0.0:
                fast = vsqr[..., 1:-1] > vsqr_diff
                for q in (1, 2, 3):
                    slp[q] = torch.where(fast, torch.zeros_like(slp[q]), slp[q])
            # pad slopes to full line length (index i -> slp[i-1])
            zero = torch.zeros_like(var[0][..., :1])
            S = [torch.cat([zero, s_, zero], dim=-1) for s_ in slp]
            # interface fluxes at 1 .. sz-3 (between i and i+1)
            I = slice(1, sz - 2)
            Ip = slice(2, sz - 1)
            flx = []
            for q in range(NV):
                Xq, Sq = var[q], S[q]
                H = hyp_v if 1 <= q <= 3 else hyp_B if 5 <= q <= 7 else hyp_e
                sl = Xq[..., Ip] - Xq[..., I]
                sl_lim = Xq[..., Ip] - Xq[..., I] - (Sq[..., I] + Sq[..., Ip])
                dnh = H[..., I]
                rfq = torch.abs(sl_lim) / torch.clamp(torch.abs(sl), min=1e-100)
                cfq = (torch.clamp(torch.minimum(sl_lim, sl), min=0.0)
                       + torch.clamp(torch.maximum(sl_lim, sl), max=0.0))
                flx.append(torch.clamp(1.0 + dnh * (rfq - 1.0), min=0.0) * cfq)
            tvd_faci = 0.5 * idxd1 * torch.clamp(torch.maximum(cm[..., I], cm[..., Ip]), max=cmax)
            flx = [f * tvd_faci for f in flx]
            flx[5 + d1] = flx[5 + d1] * B_par_diff
            if need_tvd_coeff:
                hb, ht = (hfb1[I], hft1[I]) if d1 == 0 else (hfb1, hft1)
                flx[0] = flx[0] * tc[0]
                c = torch.clamp(visc_coeff_bot * hb + visc_coeff_top * ht, min=tc[1])
                for q in (1, 2, 3):
                    flx[q] = flx[q] * c
                flx[4] = flx[4] * tc[2]
                c = torch.clamp(eta_coeff_bot * hb + eta_coeff_top * ht, min=tc[3])
                for q in (5, 6, 7):
                    flx[q] = flx[q] * c
            if hyperdiff is not None:
                istart = 1 + gh
                Ih = slice(istart, sz - 2)
                dnh = torch.maximum(hyperdiff[..., Ih], hyperdiff[..., istart + 1:sz - 1])
                for q in (0, 1, 4):
                    Xq = var[q]
                    corr = dnh * (0.125 * (Xq[..., istart + 2:sz] - Xq[..., istart - 1:sz - 3])
                                  - 0.375 * (Xq[..., istart + 1:sz - 1] - Xq[..., Ih]))
                    head = flx[q][..., :istart - 1]
                    flx[q] = torch.cat([head, flx[q][..., istart - 1:] - corr], dim=-1)
            uif = [0.5 * (var[q][..., I] + S[q][..., I] + var[q][..., Ip] - S[q][..., Ip]) for q in range(5)]
            f0, f1, f2, f3, f4 = flx[0], flx[1], flx[2], flx[3], flx[4]
            u0_, u4_ = uif[0], uif[4]
            if rho_log:
                u0_ = torch.exp(u0_)
                f0 = f0 * u0_
                u4_ = torch.exp(u4_)
                f4 = f4 * u4_
            min_cz = f0 * 0.0
            f1 = f1 * u0_ + uif[1] * min_cz
            f2 = f2 * u0_ + uif[2] * min_cz
            f3 = f3 * u0_ + uif[3] * min_cz
            qf = qft1[I] if d1 == 0 else qft1
            f4 = f4 * u0_
            f4 = f4 + ((u4_ * min_cz) + qf * (uif[1] * f1 + uif[2] * f2 + uif[3] * f3))
            flx[0], flx[1], flx[2], flx[3], flx[4] = f0, f1, f2, f3, f4
            # full-length interface arrays: F[i] at i = 1 .. sz-3
            zero = torch.zeros_like(var[0][..., :1])
            F = [torch.cat([zero, f, zero, zero], dim=-1) for f in flx]
            if d1 == 0:
                F5 = F[5].clone()
                F5[..., sz - gh - 1] = 0.0
                F5[..., gh - 1] = 0.0
                F[5] = F5
            qr = ((var[5][..., 2:sz - 1] - var[5][..., 1:sz - 2]) * F[5][..., 1:sz - 2]
                  + (var[6][..., 2:sz - 1] - var[6][..., 1:sz - 2]) * F[6][..., 1:sz - 2]
                  + (var[7][..., 2:sz - 1] - var[7][..., 1:sz - 2]) * F[7][..., 1:sz - 2])
            QR = torch.cat([zero, qr, zero, zero], dim=-1)
            # updates of the points gh .. sz-gh-1
            J = slice(gh, sz - gh)
            Jm = slice(gh - 1, sz - gh - 1)
            r = [dt_tvd * (F[q][..., J] - F[q][..., Jm]) for q in range(NV)]
            v1, v2, v3 = var[1][..., J], var[2][..., J], var[3][..., J]
            b5, b6, b7 = var[5][..., J], var[6][..., J], var[7][..., J]
            qj = qft1[J] if d1 == 0 else qft1
            r[4] = r[4] + (1.0 - qj) * (r[1] * v1 + r[2] * v2 + r[3] * v3)
            dnb = (r[1] * b5 + r[2] * b6 + r[3] * b7) / torch.clamp(bsqr[..., J], min=1e-100)
            bc2 = BC2[..., J]
            bo0 = -bc2 * (r[1] - dnb * b5)
            bo1 = -bc2 * (r[2] - dnb * b6)
            bo2 = -bc2 * (r[3] - dnb * b7)
            bo3 = bo0 * v1 + bo1 * v2 + bo2 * v3
            r[1] = r[1] + bo0
            r[2] = r[2] + bo1
            r[3] = r[3] + bo2
            r[4] = r[4] + bo3
            qres = 0.5 * (QR[..., J] + QR[..., Jm]) * qj
            r[4] = r[4] + dt_tvd * qres
            Xn = [X[q][..., J] + r[q] for q in range(NV)]
            Xn[0] = torch.clamp(Xn[0], min=rho_min)
            Xn[4] = torch.maximum(Xn[4], eps_min * Xn[0] + 0.5 * (Xn[1] * Xn[1] + Xn[2] * Xn[2]
                                                                   + Xn[3] * Xn[3]) / Xn[0])
            Xnew = torch.cat([X[:, ..., :gh], torch.stack(Xn), X[:, ..., sz - gh:]], dim=-1)
            inv = [perm.index(q) for q in range(4)]
            out = U.clone()
            out[:, sel[2], sel[1], sel[0]] = Xnew.permute(*inv)
            U = out
            bounds[d1] = [lb[d1], le[d1]]
        return U

    # ------------------------------------------------------------ step
    def step(self, U, dt):
        U0 = U
        p4 = None
        for stage in range(1, 5):
            W, p = self.cons_to_prim(U)
            if stage == 4:
                p4 = p
            R = self.residual(W, p)
            U = self.source_integrate(U, W, R, U0, stage, dt)
            U = self.bc(self.exchange(U))
        U = self.tvd(U, p4, dt)
        U = self.tcheck(U)
        return self.bc(self.exchange(U))

    def fill(self, U):
        """the initial ghost fill the forward starts from"""
        return self.bc(self.exchange(U))
