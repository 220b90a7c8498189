// mrm_adjoint.h -- hand-written reverse mode (vector-Jacobian products) of the
// FP64 cell functions of one RK step, and the gather-form adjoint kernels.
//
// Every *_vjp function re-evaluates its forward function (the one in
// mrm_device.h / mrm_kernels.h, same operation order) and propagates output
// cotangents back to its inputs, accumulating (+=) into the input cotangents.
// Piecewise functions (min/max, clamps, limiter switches, table cells,
// fallback branches) are differentiated along the branch the forward takes.
// dt, Qtot, Tau, p_bc, s_bc and the limiter constants are parameters.
#pragma once
#include "mrm_fused.h"

namespace MRM_NS {

// ------------------------------------------------------------------ primitives
// load_prim / ConsToPrim (cons_to_prim.C:110-118): w = (rho, M/rho, e - rho v^2/2, B)
MRM_HD void prim_vjp(const double u[MRM_NV], const double lw[MRM_NV], double lu[MRM_NV]) {
  const double d = u[0], dn = 1.0 / d;
  const double m1 = u[1] * dn, m2 = u[2] * dn, m3 = u[3] * dn;
  const double vv = m1 * m1 + m2 * m2 + m3 * m3;
  const double lvv = -0.5 * d * lw[4];
  double ld = lw[0] - 0.5 * vv * lw[4];
  const double lm1 = lw[1] + 2.0 * m1 * lvv, lm2 = lw[2] + 2.0 * m2 * lvv, lm3 = lw[3] + 2.0 * m3 * lvv;
  lu[1] += lm1 * dn;
  lu[2] += lm2 * dn;
  lu[3] += lm3 * dn;
  const double ldn = lm1 * u[1] + lm2 * u[2] + lm3 * u[3];
  ld -= ldn * dn * dn;
  lu[0] += ld;
  lu[4] += lw[4];
  lu[5] += lw[5];
  lu[6] += lw[6];
  lu[7] += lw[7];
}

// ------------------------------------------------------------------ EOS pressure
// pressure of ConsToPrim on one conservative cell (cons_to_prim.C:108-198, the
// path of cons_to_prim_cell<2>), and its VJP w.r.t. (rho, M, e).  Returns p.
MRM_HD double eos_p_vjp(const EosDev& E, const double u[MRM_NV], double lp, double lu[MRM_NV]) {
  const double eos_gamma = 1.65, eos_mu = 0.62, Rgas = 8.314e7;
  const double c_temp = (eos_gamma - 1.) * eos_mu / Rgas;
  const double c_pres = (eos_gamma - 1.);
  const double d = u[0], dn = 1.0 / d;
  const double v1 = u[1] * dn, v2 = u[2] * dn, v3 = u[3] * dn;
  const double vv = v1 * v1 + v2 * v2 + v3 * v3;
  const double e = u[4] - 0.5 * vv * d;
  const double eps = e * dn + E.eps_off;
  const double lgr = log(d), ep = log(eps), lr = lgr;
  int i_d = cvt_int((lr - E.xlr0) * E.inv_del_lr);
  i_d = imax_(0, i_d);
  i_d = imin_(E.N_lr - 2, i_d);
  int i_e = cvt_int((ep - E.xeps0) * E.inv_del_eps);
  i_e = imax_(0, i_e);
  i_e = imin_(E.N_eps - 2, i_e);
  const double ae = ep - E.xeps[i_e], be = E.xeps[i_e + 1] - ep;
  const double ar = lr - E.xlr[i_d], br = E.xlr[i_d + 1] - lr;
  const long L = E.N_lr, NT = (long)E.N_eps * E.N_lr;
  const long o11 = (i_e + 1) * L + i_d + 1, o10 = (i_e + 1) * L + i_d, o01 = i_e * L + i_d + 1, o00 = i_e * L + i_d;
  const double P11 = E.tab[o11], P10 = E.tab[o10], P01 = E.tab[o01], P00 = E.tab[o00];
  double v = ae * ar * P11;
  v += ae * br * P10;
  v += be * ar * P01;
  v += be * br * P00;
  double tt = ae * ar * (double)E.tab[NT + o11];
  tt += ae * br * (double)E.tab[NT + o10];
  tt += be * ar * (double)E.tab[NT + o01];
  tt += be * br * (double)E.tab[NT + o00];
  const double p0 = exp(v);
  const bool lowr = lgr < E.xlr0;
  const bool hot = eps > 0.8 * E.eps_max;
  double pres = p0, temp;
  if (!lowr && !hot && tt > -700.0 && tt < 700.0) temp = 1.0;
  else temp = exp(tt);
  if (lowr) {
    pres = c_pres * e;
    temp = c_temp * eps;
  }
  const double p_pre_hot = pres;
  double xx = 1.0, xarg = 0.0;
  if (hot) {
    xarg = 5.0 * (E.eps_max - eps) / E.eps_max;
    xx = smax(0.0, xarg);
    pres = pres * xx + (1.0 - xx) * c_pres * e;
    temp = temp * xx + (1.0 - xx) * c_temp * eps;
  }
  int bad = 0;
  if ((pres != pres) || (pres <= 0.) || (pres == __builtin_inf())) bad = 1;
  if ((temp != temp) || (temp <= 0.) || (temp == __builtin_inf())) bad = 1;
  // ---- reverse
  double le = 0.0, leps = 0.0, lv = 0.0;
  if (bad) {
    le += c_pres * lp;
    pres = c_pres * e;
  } else {
    double lpre = lp;
    if (hot) {
      lpre = lp * xx;
      le += (1.0 - xx) * c_pres * lp;
      if (0.0 < xarg) leps += lp * (p_pre_hot - c_pres * e) * (-5.0 / E.eps_max);
    }
    if (lowr) le += c_pres * lpre;
    else lv += p0 * lpre;
  }
  const double dv_dep = ar * (P11 - P01) + br * (P10 - P00);
  const double dv_dlr = ae * (P11 - P10) + be * (P01 - P00);
  leps += lv * dv_dep / eps;
  double ld = lv * dv_dlr / d;
  le += leps * dn;
  double ldn = leps * e;
  lu[4] += le;
  const double lvv = -0.5 * d * le;
  ld += -0.5 * vv * le;
  const double l1 = 2.0 * v1 * lvv, l2 = 2.0 * v2 * lvv, l3 = 2.0 * v3 * lvv;
  lu[1] += l1 * dn;
  lu[2] += l2 * dn;
  lu[3] += l3 * dn;
  ldn += l1 * u[1] + l2 * u[2] + l3 * u[3];
  ld -= ldn * dn * dn;
  lu[0] += ld;
  return pres;
}

// ------------------------------------------------------------------ fluxes
// point_flux<D> (mhdres_SR.C:296-360) as the 13 quantities the residual
// differences: 8 fluxes, 3 Maxwell-stress components, B_D2, B_D3.
template <int D>
MRM_HD void flux13(const double var[ST_NF], double f[13]) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  PtFlux o;
  point_flux<D>(var, var[8], o);
#pragma unroll
  for (int q = 0; q < 8; q++) f[q] = o.f[q];
  f[8] = o.ms[0]; f[9] = o.ms[1]; f[10] = o.ms[2];
  f[11] = var[5 + D2];
  f[12] = var[5 + D3];
}

template <int D>
MRM_HD void flux13_vjp(const double var[ST_NF], const double lf[13], double lvar[ST_NF]) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  const double p1 = D == 0 ? 1.0 : 0.0, p2 = D == 1 ? 1.0 : 0.0, p3 = D == 2 ? 1.0 : 0.0;
  const double var0 = var[0], var1 = var[1], var2 = var[2], var3 = var[3];
  const double var4 = var[4], var5 = var[5], var6 = var[6], var7 = var[7], pres = var[8];
  const double vel = var[1 + D], mag = var[5 + D];
  const double vv = var1 * var1 + var2 * var2 + var3 * var3;
  const double ekin = 0.5 * var0 * vv;
  const double mflx = vel * var0;
  const double lmflx = lf[0] + lf[1] * var1 + lf[2] * var2 + lf[3] * var3;
  double l0 = lmflx * vel;
  double l1 = lf[1] * mflx - lf[5] * mag;
  double l2 = lf[2] * mflx - lf[6] * mag;
  double l3 = lf[3] * mflx - lf[7] * mag;
  double lpres = p1 * lf[1] + p2 * lf[2] + p3 * lf[3] + vel * lf[4];
  double lvel = lmflx * var0 + lf[4] * (var4 + ekin + pres) + lf[5] * var5 + lf[6] * var6 + lf[7] * var7;
  const double lekin = vel * lf[4];
  const double l4 = vel * lf[4];
  l0 += 0.5 * vv * lekin;
  const double lvv = 0.5 * var0 * lekin;
  l1 += 2.0 * var1 * lvv;
  l2 += 2.0 * var2 * lvv;
  l3 += 2.0 * var3 * lvv;
  double l5 = lf[5] * vel + lf[8] * mag;
  double l6 = lf[6] * vel + lf[9] * mag;
  double l7 = lf[7] * vel + lf[10] * mag;
  const double lmag = -(lf[5] * var1 + lf[6] * var2 + lf[7] * var3) + lf[8] * var5 + lf[9] * var6 + lf[10] * var7;
  const double lbb = 0.5 * (-(p1 * lf[8] + p2 * lf[9] + p3 * lf[10]));
  l5 += 2.0 * var5 * lbb;
  l6 += 2.0 * var6 * lbb;
  l7 += 2.0 * var7 * lbb;
  double l[8] = {l0, l1, l2, l3, l4, l5, l6, l7};
  l[1 + D] += lvel;
  l[5 + D] += lmag;
  l[5 + D2] += lf[11];
  l[5 + D3] += lf[12];
#pragma unroll
  for (int q = 0; q < 8; q++) lvar[q] += l[q];
  lvar[8] += lpres;
}

// lfac_of (mhdres_SR.C:319-323): lf = (1+x2)/(1+x2+x2^2), x2 = B^2/(rho va_max^2)
MRM_HD void lfac_vjp(const double var[ST_NF], double inv_va2max, double llf, double lvar[ST_NF]) {
  const double dn = 1.0 / var[0];
  const double bb = var[5] * var[5] + var[6] * var[6] + var[7] * var[7];
  const double va2 = bb * dn;
  const double x2 = va2 * inv_va2max;
  const double x4 = x2 * x2;
  const double s = 1.0 + x2;
  const double den = s + x4;
  const double lx2 = llf * (-x2 * (2.0 + x2)) / (den * den);
  const double lva2 = lx2 * inv_va2max;
  const double lbb = lva2 * dn;
  const double ldn = lva2 * bb;
  lvar[0] += -ldn * dn * dn;
  lvar[5] += 2.0 * var[5] * lbb;
  lvar[6] += 2.0 * var[6] * lbb;
  lvar[7] += 2.0 * var[7] * lbb;
}

// Centre part of one sweep's residual (mhdres_SR.C:361-412): given the
// differences Delta_q = w1 d1_q + w2 d2_q of the 13 flux quantities, the
// sweep adds -Delta_q (q < 8) and the Boris-blended Lorentz force
// L = lf lrt + (1-lf) cb with its work c.v . L.  VJP: from the residual
// cotangent lres, G[q] = d/dDelta_q, plus the centre and lf cotangents.
template <int D>
MRM_HD void sweep_centre_vjp(const double c[ST_NF], const double Dl[13], double lf, const double lres[MRM_NV],
                             double G[13], double lc[ST_NF], double& llf) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  const double dB2 = Dl[11], dB3 = Dl[12];
  double cb[3];
  cb[D] = -c[5 + D2] * dB2 - c[5 + D3] * dB3;
  cb[D2] = c[5 + D] * dB2;
  cb[D3] = c[5 + D] * dB3;
  double Lk[3], lL[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    Lk[k] = lf * Dl[8 + k] + (1.0 - lf) * cb[k];
    lL[k] = lres[1 + k] + lres[4] * c[1 + k];
    lc[1 + k] += lres[4] * Lk[k];
  }
#pragma unroll
  for (int q = 0; q < 8; q++) G[q] = -lres[q];
  double lcb[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    G[8 + k] = lf * lL[k];
    lcb[k] = (1.0 - lf) * lL[k];
    llf += lL[k] * (Dl[8 + k] - cb[k]);
  }
  G[11] = -c[5 + D2] * lcb[D] + c[5 + D] * lcb[D2];
  G[12] = -c[5 + D3] * lcb[D] + c[5 + D] * lcb[D3];
  lc[5 + D2] += -dB2 * lcb[D];
  lc[5 + D3] += -dB3 * lcb[D];
  lc[5 + D] += dB2 * lcb[D2] + dB3 * lcb[D3];
}

// ------------------------------------------------------------------ Tcheck
// tcheck_vals (Add_Sources_Integrate_SR.C:447-465): u -> clamped u.
MRM_HD void tcheck_vjp(const double u[MRM_NV], double rho_min, double eps_min, double eps_max, double vmax,
                       const double lo[MRM_NV], double lu[MRM_NV]) {
  const double dn = fmax(rho_min, u[0]);
  const double msq = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
  const double am = sqrt(msq);
  const double vv = am / dn;
  const bool vlim = vmax < vv;
  const double s = vmax / fmax(vmax, vv);
  const double ekin = 0.5 * dn * vv * vv;
  const double lo_e = eps_min * dn + ekin * s * s;
  const double hi_e = eps_max * dn + ekin * s * s;
  const double t = fmax(u[4], lo_e);
  double ldn = lo[0], ls = 0.0, llo = 0.0, lhi = 0.0;
  double lm[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    lm[k] = lo[1 + k] * s;
    ls += lo[1 + k] * u[1 + k];
  }
  if (t > hi_e) lhi = lo[4];
  else if (u[4] >= lo_e) lu[4] += lo[4];
  else llo = lo[4];
  ldn += eps_min * llo + eps_max * lhi;
  const double lek = (llo + lhi) * s * s;
  ls += (llo + lhi) * 2.0 * ekin * s;
  ldn += 0.5 * vv * vv * lek;
  double lvv = dn * vv * lek;
  if (vlim) lvv += -ls * vmax / (vv * vv);
  const double lam_ = lvv / dn;
  ldn += -lvv * am / (dn * dn);
  if (am > 0.0) {
#pragma unroll
    for (int k = 0; k < 3; k++) lm[k] += lam_ * u[1 + k] / am;
  }
  if (u[0] > rho_min) lu[0] += ldn;
  lu[1] += lm[0]; lu[2] += lm[1]; lu[3] += lm[2];
  lu[5] += lo[5]; lu[6] += lo[6]; lu[7] += lo[7];
}

// ------------------------------------------------------------------ sources, Boris, RK
// source_integrate_cell (Add_Sources_Integrate_SR.C:331-465).  Inputs: up =
// primitive centre (8), R = residual (8), u0 = U0 (8).  Given the output
// cotangent lo, adds to lup, lR, lu0.
MRM_HD void source_integrate_vjp(const SrcParams& P, const double up[MRM_NV], const double Rin[MRM_NV],
                                 const double u0[MRM_NV], double Qrad, const double lo[MRM_NV],
                                 double lup[MRM_NV], double lR[MRM_NV], double lu0[MRM_NV]) {
  // forward
  double R[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) R[q] = Rin[q];
  const double dn1 = P.invRdt * up[4];
  const double efaci = dn1 / fmax(dn1, fabs(Qrad));
  const double dng = P.g * up[0] - 0.0;
  const double e_dmpi = dng * up[1];
  R[1] += dng;
  R[4] += e_dmpi + Qrad * efaci;
  const double rh = up[0], rd = R[0];
  const double vx = up[1], vy = up[2], vz = up[3];
  const double bx = up[5], by = up[6], bz = up[7];
  double fx = R[1], fy = R[2], fz = R[3];
  const double bb = bx * bx + by * by + bz * bz;
  const double va2 = bb / rh;
  const double x2 = va2 * P.inv_va2max;
  const double x4 = x2 * x2;
  const double sden = 1 + x2 + x4;
  const double bf = x4 / sden;
  fx -= vx * rd;
  fy -= vy * rd;
  fz -= vz * rd;
  const double den = fmax(1e-100, bb);
  const double fb = fx * bx + fy * by + fz * bz;
  const double dnb = fb / den;
  const double pxv = fx - dnb * bx, pyv = fy - dnb * by, pzv = fz - dnb * bz;
  const double b0 = -bf * pxv, b1 = -bf * pyv, b2 = -bf * pzv;
  double out[MRM_NV];
  out[0] = u0[0] + P.wdt * R[0];
  out[1] = u0[1] + P.wdt * (R[1] + b0);
  out[2] = u0[2] + P.wdt * (R[2] + b1);
  out[3] = u0[3] + P.wdt * (R[3] + b2);
  out[4] = u0[4] + P.wdt * (R[4] + (b0 * vx + b1 * vy + b2 * vz));
  out[5] = u0[5] + P.wdt * R[5];
  out[6] = u0[6] + P.wdt * R[6];
  out[7] = u0[7] + P.wdt * R[7];
  // reverse: Tcheck, RK
  double lpre[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  tcheck_vjp(out, P.rho_min, P.eps_min, P.eps_max, P.vmax, lo, lpre);
  double lRf[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    lu0[q] += lpre[q];
    lRf[q] = P.wdt * lpre[q];
  }
  // Boris
  const double lb0 = lRf[1] + lRf[4] * vx, lb1 = lRf[2] + lRf[4] * vy, lb2 = lRf[3] + lRf[4] * vz;
  double lvx = lRf[4] * b0, lvy = lRf[4] * b1, lvz = lRf[4] * b2;
  const double lbf = -(lb0 * pxv + lb1 * pyv + lb2 * pzv);
  double lfx = -bf * lb0, lfy = -bf * lb1, lfz = -bf * lb2;
  const double ldnb = bf * (lb0 * bx + lb1 * by + lb2 * bz);
  double lbx = bf * dnb * lb0, lby = bf * dnb * lb1, lbz = bf * dnb * lb2;
  lfx += ldnb * bx / den;
  lfy += ldnb * by / den;
  lfz += ldnb * bz / den;
  lbx += ldnb * fx / den;
  lby += ldnb * fy / den;
  lbz += ldnb * fz / den;
  double lbb = 0.0;
  if (bb > 1e-100) lbb += -ldnb * fb / (den * den);
  const double lx2 = lbf * x2 * (2.0 + x2) / (sden * sden);
  const double lva2 = lx2 * P.inv_va2max;
  lbb += lva2 / rh;
  double lrh = -lva2 * bb / (rh * rh);
  lbx += 2.0 * bx * lbb;
  lby += 2.0 * by * lbb;
  lbz += 2.0 * bz * lbb;
  // f = R - v rd
  double lR1[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) lR1[q] = lRf[q];
  lR1[1] += lfx;
  lR1[2] += lfy;
  lR1[3] += lfz;
  lvx += -rd * lfx;
  lvy += -rd * lfy;
  lvz += -rd * lfz;
  lR1[0] += -(vx * lfx + vy * lfy + vz * lfz);
  // gravity and radiative source
  const double ldng = lR1[1] + lR1[4] * up[1];
  lvx += lR1[4] * dng;
  lrh += P.g * ldng;
  const double lefaci = lR1[4] * Qrad;
  if (fabs(Qrad) > dn1) lup[4] += P.invRdt * lefaci / fabs(Qrad);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) lR[q] += lR1[q];
  lup[0] += lrh;
  lup[1] += lvx;
  lup[2] += lvy;
  lup[3] += lvz;
  lup[5] += lbx;
  lup[6] += lby;
  lup[7] += lbz;
}

// ------------------------------------------------------------------ TVD
// TvdPt fields in the adjoint's per-point buffers
enum { TW_V = 0, TW_LNR = 8, TW_VSQR, TW_BSQR, TW_CM, TW_HE, TW_HV, TW_HB, TW_BC2, TW_CZ, TW_HD, TW_N };

MRM_HD void tvdpt_store(const TvdPt& p, double* W, long bs, long n) {
#pragma unroll
  for (int q = 0; q < 8; q++) W[(TW_V + q) * bs + n] = p.v[q];
  W[TW_LNR * bs + n] = p.lnr; W[TW_VSQR * bs + n] = p.vsqr; W[TW_BSQR * bs + n] = p.bsqr;
  W[TW_CM * bs + n] = p.cm; W[TW_HE * bs + n] = p.hyp_e; W[TW_HV * bs + n] = p.hyp_v;
  W[TW_HB * bs + n] = p.hyp_B; W[TW_BC2 * bs + n] = p.BC2; W[TW_CZ * bs + n] = p.CZ; W[TW_HD * bs + n] = p.hd;
}
MRM_HD void tvdpt_load(TvdPt& p, const double* W, long bs, long n) {
#pragma unroll
  for (int q = 0; q < 8; q++) p.v[q] = W[(TW_V + q) * bs + n];
  p.lnr = W[TW_LNR * bs + n]; p.vsqr = W[TW_VSQR * bs + n]; p.bsqr = W[TW_BSQR * bs + n];
  p.cm = W[TW_CM * bs + n]; p.hyp_e = W[TW_HE * bs + n]; p.hyp_v = W[TW_HV * bs + n];
  p.hyp_B = W[TW_HB * bs + n]; p.BC2 = W[TW_BC2 * bs + n]; p.CZ = W[TW_CZ * bs + n]; p.hd = W[TW_HD * bs + n];
}

// tvd_point (tvdlimit_SR.C:269-385).  lo: cotangents of the TvdPt fields
// (indexed by TW_*, TW_CZ ignored).  Adds to lu[8] and lpres.
MRM_HD void tvd_point_vjp(const TvdParams& P, const double u[MRM_NV], double pres, double tau, double hfb1,
                          double hft1, bool xsweep, const double lo[TW_N], double lu[MRM_NV], double& lpres) {
  const double v0 = smax(P.rho_min, u[0]);
  const double CZ = (double)(tau > 1.0e-5);
  const double dn = 1.0 / v0;
  const double v1 = u[1] * dn, v2 = u[2] * dn, v3 = u[3] * dn;
  const double vv = v1 * v1 + v2 * v2 + v3 * v3;
  const double bb = u[5] * u[5] + u[6] * u[6] + u[7] * u[7];
  const double v4 = u[4] * dn - 0.5 * vv;
  const double vvs = sqrt(vv);
  const double va2 = bb * dn;
  const double x2 = va2 * P.inv_va2max;
  const double x4 = x2 * x2;
  const double s = 1.0 + x2;
  const double lf = s / (s + x4);
  const double cs2a = 5.0 / 3.0 * pres * dn;
  const double cfast = sqrt(va2 + cs2a);
  const double CME = (1.0 - CZ) * ((double)(vvs >= P.CME_thresh * cfast));
  const double dn2 = v0 * v0;
  const double rfa = dn2 / (dn2 + P.rho_lev);
  const double rf = smin(lf, rfa);
  const double hh = 1.0 - hfb1 - hft1;
  const double cf0 = hfb1 * P.tvd_cs[0] + (rf * P.tvd_cs[1] + (1.0 - rf) * P.tvd_cs[2]) * hh + hft1 * P.tvd_cs[3];
  const double cf = smax(cf0, CME);
  const double cs2 = cs2a * cf * cf;
  const double lfw = lf * (cs2 + va2);
  const double inner = smax(cs2, lfw);
  const double cm = vvs + sqrt(inner);
  // reverse
  double lv0 = 0.0, lv4 = 0.0;
  if (P.rho_log) {
    lv0 += (lo[TW_V + 0] + lo[TW_LNR]) / v0;
    lv4 += lo[TW_V + 4] / v4;
  } else {
    lv0 += lo[TW_V + 0] + lo[TW_LNR] / v0;
    lv4 += lo[TW_V + 4];
  }
  double lcm = lo[TW_CM], lvv = lo[TW_VSQR], lbb = lo[TW_BSQR];
  if (xsweep && (P.vhyp > 0.0)) {
    const double h = P.vhyp * sqrt(vv);
    const double t = smin(h, P.cmax - cm);
    if (0.0 < t) {
      const double lt = 0.5 * P.idxd1 * lo[TW_HD];
      if (P.cmax - cm < h) lcm += -lt;
      else if (vv > 0.0) lvv += lt * P.vhyp * 0.5 / sqrt(vv);
    }
  }
  double llf = -lo[TW_BC2];
  const double lhe0 = lo[TW_HE] * (1.0 - CME), lhv0 = lo[TW_HV] * (1.0 - CME);
  double lrf = lhe0 * (P.tvd_h[1] - P.tvd_h[2]) * hh + lhv0 * (P.tvd_h[1] - P.tvd_h[2] * P.tvd_pm_v) * hh +
               lo[TW_HB] * (P.tvd_h[1] - P.tvd_h[2] * P.tvd_pm_B) * hh;
  double lvvs = lcm;
  const double linner = lcm * 0.5 / sqrt(inner);
  double lcs2 = 0.0, lva2 = 0.0;
  if (cs2 < lfw) {
    llf += linner * (cs2 + va2);
    lcs2 += linner * lf;
    lva2 += linner * lf;
  } else {
    lcs2 += linner;
  }
  double lcs2a = lcs2 * cf * cf;
  const double lcf = lcs2 * cs2a * 2.0 * cf;
  if (!(cf0 < CME)) lrf += lcf * (P.tvd_cs[1] - P.tvd_cs[2]) * hh;
  if (rfa < lf) {
    const double lrfa = lrf;
    const double q = dn2 + P.rho_lev;
    lv0 += lrfa * P.rho_lev / (q * q) * 2.0 * v0;
  } else {
    llf += lrf;
  }
  double ldn = 0.0;
  lpres += lcs2a * 5.0 / 3.0 * dn;
  ldn += lcs2a * 5.0 / 3.0 * pres;
  {
    const double den = s + x4;
    const double lx2 = llf * (-x2 * (2.0 + x2)) / (den * den);
    lva2 += lx2 * P.inv_va2max;
  }
  lbb += lva2 * dn;
  ldn += lva2 * bb;
  if (vv > 0.0) lvv += lvvs * 0.5 / vvs;
  lu[4] += lv4 * dn;
  ldn += lv4 * u[4];
  lvv += -0.5 * lv4;
  const double l1 = lo[TW_V + 1] + 2.0 * v1 * lvv, l2 = lo[TW_V + 2] + 2.0 * v2 * lvv, l3 = lo[TW_V + 3] + 2.0 * v3 * lvv;
  lu[1] += l1 * dn;
  lu[2] += l2 * dn;
  lu[3] += l3 * dn;
  ldn += l1 * u[1] + l2 * u[2] + l3 * u[3];
  lu[5] += lo[TW_V + 5] + 2.0 * u[5] * lbb;
  lu[6] += lo[TW_V + 6] + 2.0 * u[6] * lbb;
  lu[7] += lo[TW_V + 7] + 2.0 * u[7] * lbb;
  lv0 += -ldn * dn * dn;
  if (P.rho_min < u[0]) lu[0] += lv0;
}

// tvd_slope (tvdlimit_SR.C:388-434): cotangent lS -> la.v, lb.v, lc.v
MRM_HD void tvd_slope_vjp(const TvdParams& P, const TvdPt& a, const TvdPt& b, const TvdPt& c, double hfb1,
                          double hft1, const double lS[MRM_NV], double la[MRM_NV], double lb[MRM_NV],
                          double lc[MRM_NV]) {
  double m[MRM_NV] = {1, 1, 1, 1, 1, 1, 1, 1};
  if ((P.visc_slope_bot > 0.0) || (P.visc_slope_top > 0.0)) {
    const double cf = (1.0 - P.visc_slope_bot * hfb1) * (1.0 - P.visc_slope_top * hft1);
    m[1] = m[2] = m[3] = cf;
  }
  if ((P.eta_slope_bot > 0.0) || (P.eta_slope_top > 0.0)) {
    const double cf = (1.0 - P.eta_slope_bot * hfb1) * (1.0 - P.eta_slope_top * hft1);
    m[5] = m[6] = m[7] = cf;
  }
  if (fabs(b.lnr - c.lnr) > P.q_rho_max) m[0] = 0.0;
  if (P.vmax_lim > 0.0 && b.vsqr > P.vsqr_diff) m[1] = m[2] = m[3] = 0.0;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    const double g = lS[q] * m[q];
    const double sl = b.v[q] - a.v[q];
    const double sr = c.v[q] - b.v[q];
    const double slm = 0.25 * (sl + sr);
    const bool in_sr = sr < sl;                 // smin(sl, sr) picks sr
    const double inner = in_sr ? sr : sl;
    const bool lo_m = slm < inner;
    const double lower = lo_m ? slm : inner;
    const bool in2_sr = sl < sr;                // smax(sl, sr) picks sr
    const double inner2 = in2_sr ? sr : sl;
    const bool up_m = inner2 < slm;
    const double upper = up_m ? slm : inner2;
    double lsl = 0.0, lsr = 0.0, lslm = 0.0;
    if (0.0 < lower) {
      if (lo_m) lslm += g;
      else if (in_sr) lsr += g;
      else lsl += g;
    }
    if (upper < 0.0) {
      if (up_m) lslm += g;
      else if (in2_sr) lsr += g;
      else lsl += g;
    }
    lsl += 0.25 * lslm;
    lsr += 0.25 * lslm;
    la[q] -= lsl;
    lb[q] += lsl - lsr;
    lc[q] += lsr;
  }
}

// cotangents of the TvdPt fields an interface reads
struct TvdIfaceAdj {
  double la[3];    // a.v[0], a.v[1], a.v[4]
  double lb[TW_N]; // b: v, cm, hyp_e, hyp_v, hyp_B, hd
  double lc[TW_N]; // c: v, cm, hd
  double ld[3];    // d.v[0], d.v[1], d.v[4]
  double lSb[MRM_NV], lSc[MRM_NV];
};

// tvd_iface<D> (tvdlimit_SR.C:436-563, 609-614): cotangents lF (8) of the
// interface flux and lq of the returned resistive term.
template <int D>
MRM_HD void tvd_iface_vjp(const TvdParams& P, const TvdPt& a, const TvdPt& b, const TvdPt& c, const TvdPt& d,
                          const double Sb[MRM_NV], const double Sc[MRM_NV], double hfb1, double hft1, double qft1,
                          bool hyper_on, bool zero_bflux, const double lFo[MRM_NV], double lq, TvdIfaceAdj& A) {
  // forward with intermediates
  double F0[MRM_NV], g[MRM_NV], cfq[MRM_NV], rfq[MRM_NV], hq[MRM_NV], sl[MRM_NV], sll[MRM_NV], den[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    sl[q] = c.v[q] - b.v[q];
    sll[q] = c.v[q] - b.v[q] - (Sb[q] + Sc[q]);
    hq[q] = (q >= 1 && q <= 3) ? b.hyp_v : (q >= 5 ? b.hyp_B : b.hyp_e);
    den[q] = smax(1e-100, fabs(sl[q]));
    rfq[q] = fabs(sll[q]) / den[q];
    cfq[q] = smax(0.0, smin(sll[q], sl[q])) + smin(0.0, smax(sll[q], sl[q]));
    g[q] = smax(0.0, 1.0 + hq[q] * (rfq[q] - 1.0));
    F0[q] = g[q] * cfq[q];
  }
  const double mm = smax(b.cm, c.cm);
  const double faci = 0.5 * P.idxd1 * smin(P.cmax, mm);
  double mult[MRM_NV] = {1, 1, 1, 1, 1, 1, 1, 1};
  mult[5 + D] = P.B_par_diff;
  if (P.need_tvd_coeff) {
    mult[0] *= P.tc[0];
    const double cf1 = smax(P.tc[1], P.visc_coeff_bot * hfb1 + P.visc_coeff_top * hft1);
    mult[1] *= cf1; mult[2] *= cf1; mult[3] *= cf1;
    mult[4] *= P.tc[2];
    const double cf2 = smax(P.tc[3], P.eta_coeff_bot * hfb1 + P.eta_coeff_top * hft1);
    mult[5] *= cf2; mult[6] *= cf2; mult[7] *= cf2;
  }
  double F2[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) F2[q] = F0[q] * faci * mult[q];
  const bool hyp = (D == 0 && hyper_on);
  const double dnh = smax(b.hd, c.hd);
  double H[3] = {0, 0, 0};
  const int hk[3] = {0, 1, 4};
  double F3[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) F3[q] = F2[q];
  if (hyp) {
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const int k = hk[t];
      H[t] = 0.125 * (d.v[k] - a.v[k]) - 0.375 * (c.v[k] - b.v[k]);
      F3[k] -= dnh * H[t];
    }
  }
  double uifr[5];
#pragma unroll
  for (int k = 0; k < 5; k++) uifr[k] = 0.5 * (b.v[k] + Sb[k] + c.v[k] - Sc[k]);
  double uif0 = uifr[0], uif4 = uifr[4];
  double flx0 = F3[0], flx4 = F3[4];
  if (P.rho_log) {
    uif0 = exp(uifr[0]);
    flx0 = F3[0] * uif0;
    uif4 = exp(uifr[4]);
    flx4 = F3[4] * uif4;
  }
  const double Z = smin(b.CZ, c.CZ);
  const double mincz = flx0 * Z;
  double flxp[4];
  flxp[1] = F3[1] * uif0 + uifr[1] * mincz;
  flxp[2] = F3[2] * uif0 + uifr[2] * mincz;
  flxp[3] = F3[3] * uif0 + uifr[3] * mincz;
  double Fo[MRM_NV];
  Fo[0] = flx0; Fo[1] = flxp[1]; Fo[2] = flxp[2]; Fo[3] = flxp[3];
  Fo[4] = flx4 * uif0 + (uif4 * mincz) + qft1 * (uifr[1] * flxp[1] + uifr[2] * flxp[2] + uifr[3] * flxp[3]);
  Fo[5] = zero_bflux ? 0.0 : F3[5];
  Fo[6] = F3[6]; Fo[7] = F3[7];
  // reverse
  for (int q = 0; q < TW_N; q++) { A.lb[q] = 0.0; A.lc[q] = 0.0; }
  for (int q = 0; q < 3; q++) { A.la[q] = 0.0; A.ld[q] = 0.0; }
  for (int q = 0; q < MRM_NV; q++) { A.lSb[q] = 0.0; A.lSc[q] = 0.0; }
  double lF[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) lF[q] = lFo[q];
#pragma unroll
  for (int k = 5; k < 8; k++) {
    lF[k] += lq * (c.v[k] - b.v[k]);
    A.lc[TW_V + k] += lq * Fo[k];
    A.lb[TW_V + k] -= lq * Fo[k];
  }
  double lF3[MRM_NV];
  lF3[5] = zero_bflux ? 0.0 : lF[5];
  lF3[6] = lF[6];
  lF3[7] = lF[7];
  double luif[5] = {0, 0, 0, 0, 0};
  double luif0 = 0.0, luif4 = 0.0, lmincz = 0.0;
  // Fo[4]
  const double lflx4p = lF[4] * uif0;
  luif0 += lF[4] * flx4;
  luif4 += lF[4] * mincz;
  lmincz += lF[4] * uif4;
  const double lqs = lF[4] * qft1;
  double lflxp[4];
#pragma unroll
  for (int k = 1; k < 4; k++) {
    luif[k] += lqs * flxp[k];
    lflxp[k] = lF[k] + lqs * uifr[k];
  }
#pragma unroll
  for (int k = 1; k < 4; k++) {
    lF3[k] = lflxp[k] * uif0;
    luif0 += lflxp[k] * F3[k];
    luif[k] += lflxp[k] * mincz;
    lmincz += lflxp[k] * uifr[k];
  }
  const double lflx0 = lF[0] + lmincz * Z;
  if (P.rho_log) {
    lF3[0] = lflx0 * uif0;
    luif[0] += (lflx0 * F3[0] + luif0) * uif0;
    lF3[4] = lflx4p * uif4;
    luif[4] += (lflx4p * F3[4] + luif4) * uif4;
  } else {
    lF3[0] = lflx0;
    luif[0] += luif0;
    lF3[4] = lflx4p;
    luif[4] += luif4;
  }
#pragma unroll
  for (int k = 0; k < 5; k++) {
    A.lb[TW_V + k] += 0.5 * luif[k];
    A.lSb[k] += 0.5 * luif[k];
    A.lc[TW_V + k] += 0.5 * luif[k];
    A.lSc[k] -= 0.5 * luif[k];
  }
  // hyperdiffusion
  if (hyp) {
    double ldnh = 0.0;
#pragma unroll
    for (int t = 0; t < 3; t++) {
      const int k = hk[t];
      ldnh += -lF3[k] * H[t];
      const double lH = -dnh * lF3[k];
      A.ld[t] += 0.125 * lH;
      A.la[t] -= 0.125 * lH;
      A.lc[TW_V + k] -= 0.375 * lH;
      A.lb[TW_V + k] += 0.375 * lH;
    }
    if (b.hd < c.hd) A.lc[TW_HD] += ldnh;
    else A.lb[TW_HD] += ldnh;
  }
  // F2 = F0 * faci * mult
  double lfaci = 0.0;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    const double lF2 = lF3[q];
    lfaci += lF2 * F0[q] * mult[q];
    const double lF0 = lF2 * faci * mult[q];
    const double lg = lF0 * cfq[q];
    const double lcf = lF0 * g[q];
    double lsl = 0.0, lsll = 0.0;
    if (0.0 < 1.0 + hq[q] * (rfq[q] - 1.0)) {
      const double lhq = lg * (rfq[q] - 1.0);
      if (q >= 1 && q <= 3) A.lb[TW_HV] += lhq;
      else if (q >= 5) A.lb[TW_HB] += lhq;
      else A.lb[TW_HE] += lhq;
      const double lrf = lg * hq[q];
      const double sgl = sll[q] > 0.0 ? 1.0 : (sll[q] < 0.0 ? -1.0 : 0.0);
      lsll += lrf * sgl / den[q];
      if (1e-100 < fabs(sl[q])) {
        const double sg = sl[q] > 0.0 ? 1.0 : -1.0;
        lsl += -lrf * fabs(sll[q]) / (den[q] * den[q]) * sg;
      }
    }
    // cf = smax(0, smin(sll, sl)) + smin(0, smax(sll, sl))
    const bool mn_sl = sl[q] < sll[q];   // smin(sll, sl) picks sl
    const double mn = mn_sl ? sl[q] : sll[q];
    if (0.0 < mn) { if (mn_sl) lsl += lcf; else lsll += lcf; }
    const bool mx_sl = sll[q] < sl[q];   // smax(sll, sl) picks sl
    const double mx = mx_sl ? sl[q] : sll[q];
    if (mx < 0.0) { if (mx_sl) lsl += lcf; else lsll += lcf; }
    lsl += lsll;
    A.lSb[q] -= lsll;
    A.lSc[q] -= lsll;
    A.lc[TW_V + q] += lsl;
    A.lb[TW_V + q] -= lsl;
  }
  if (mm < P.cmax) {
    const double lmm = lfaci * 0.5 * P.idxd1;
    if (b.cm < c.cm) A.lc[TW_CM] += lmm;
    else A.lb[TW_CM] += lmm;
  }
}

// tvd_update + its clamps (tvdlimit_SR.C:565-657).  Given lo (8) on the
// updated point: lu (direct), lF, lFm, lq, lqm and the cotangents of the
// point fields the update reads (v1..3, v5..7, BC2, bsqr) in lb (TW_* index).
MRM_HD void tvd_update_vjp(const TvdParams& P, const TvdPt& b, const double Fm[MRM_NV], const double F[MRM_NV],
                           double qm, double q, double qft1, const double u[MRM_NV], const double lo[MRM_NV],
                           double lu[MRM_NV], double lF[MRM_NV], double lFm[MRM_NV], double& lq, double& lqm,
                           double lb[TW_N]) {
  double r[MRM_NV];
#pragma unroll
  for (int k = 0; k < MRM_NV; k++) r[k] = P.dt_tvd * (F[k] - Fm[k]);
  const double v1 = b.v[1], v2 = b.v[2], v3 = b.v[3], b5 = b.v[5], b6 = b.v[6], b7 = b.v[7];
  const double rv = r[1] * v1 + r[2] * v2 + r[3] * v3;
  const double den = smax(1e-100, b.bsqr);
  const double rb = r[1] * b5 + r[2] * b6 + r[3] * b7;
  const double dnb = rb / den;
  const double q1 = r[1] - dnb * b5, q2 = r[2] - dnb * b6, q3 = r[3] - dnb * b7;
  const double bo0 = -b.BC2 * q1, bo1 = -b.BC2 * q2, bo2 = -b.BC2 * q3;
  double up[MRM_NV];
  up[0] = u[0] + r[0];
  up[1] = u[1] + (r[1] + bo0);
  up[2] = u[2] + (r[2] + bo1);
  up[3] = u[3] + (r[3] + bo2);
  up[4] = u[4] + (r[4] + (1.0 - qft1) * rv + (bo0 * v1 + bo1 * v2 + bo2 * v3) +
                  P.dt_tvd * (0.5 * (q + qm) * qft1));
  up[5] = u[5] + r[5]; up[6] = u[6] + r[6]; up[7] = u[7] + r[7];
  const double u0c = smax(P.rho_min, up[0]);
  const double msq = up[1] * up[1] + up[2] * up[2] + up[3] * up[3];
  const double emin = P.eps_min * u0c + 0.5 * msq / u0c;
  // reverse clamps
  double lup[MRM_NV];
#pragma unroll
  for (int k = 0; k < MRM_NV; k++) lup[k] = lo[k];
  double lemin = 0.0;
  if (up[4] < emin) { lemin = lo[4]; lup[4] = 0.0; }
  double lu0c = lo[0] + lemin * (P.eps_min - 0.5 * msq / (u0c * u0c));
  lup[1] += lemin * up[1] / u0c;
  lup[2] += lemin * up[2] / u0c;
  lup[3] += lemin * up[3] / u0c;
  lup[0] = (P.rho_min < up[0]) ? lu0c : 0.0;
#pragma unroll
  for (int k = 0; k < MRM_NV; k++) lu[k] += lup[k];
  // r cotangents
  double lr[MRM_NV];
#pragma unroll
  for (int k = 0; k < MRM_NV; k++) lr[k] = lup[k];
  const double lqres = P.dt_tvd * lup[4];
  lq += lqres * 0.5 * qft1;
  lqm += lqres * 0.5 * qft1;
  const double lbo0 = lup[1] + lup[4] * v1, lbo1 = lup[2] + lup[4] * v2, lbo2 = lup[3] + lup[4] * v3;
  double lv1 = lup[4] * bo0, lv2 = lup[4] * bo1, lv3 = lup[4] * bo2;
  lb[TW_BC2] += -(lbo0 * q1 + lbo1 * q2 + lbo2 * q3);
  lr[1] += -b.BC2 * lbo0;
  lr[2] += -b.BC2 * lbo1;
  lr[3] += -b.BC2 * lbo2;
  const double ldnb = b.BC2 * (lbo0 * b5 + lbo1 * b6 + lbo2 * b7);
  double lb5 = b.BC2 * dnb * lbo0, lb6 = b.BC2 * dnb * lbo1, lb7 = b.BC2 * dnb * lbo2;
  lr[1] += ldnb * b5 / den;
  lr[2] += ldnb * b6 / den;
  lr[3] += ldnb * b7 / den;
  lb5 += ldnb * r[1] / den;
  lb6 += ldnb * r[2] / den;
  lb7 += ldnb * r[3] / den;
  if (1e-100 < b.bsqr) lb[TW_BSQR] += -ldnb * rb / (den * den);
  const double lrv = lup[4] * (1.0 - qft1);
  lr[1] += lrv * v1;
  lr[2] += lrv * v2;
  lr[3] += lrv * v3;
  lv1 += lrv * r[1];
  lv2 += lrv * r[2];
  lv3 += lrv * r[3];
  lb[TW_V + 1] += lv1; lb[TW_V + 2] += lv2; lb[TW_V + 3] += lv3;
  lb[TW_V + 5] += lb5; lb[TW_V + 6] += lb6; lb[TW_V + 7] += lb7;
#pragma unroll
  for (int k = 0; k < MRM_NV; k++) {
    lF[k] += P.dt_tvd * lr[k];
    lFm[k] -= P.dt_tvd * lr[k];
  }
}

// ------------------------------------------------------------------ EOS scalar interps
// derivative of eos_fwd_raw(t, ee, dd) (eos_init.C:268-416) w.r.t. (ee, dd)
MRM_HD double eos_fwd_raw_d(const EosDev& E, int t, double ee, double dd, double& dee, double& ddd) {
  const double ee1 = log(ee + E.eps_off);
  const double logr = log(dd);
  int i = cvt_int((ee1 - E.xeps0) * E.inv_del_eps);
  if (i < 0) i = 0;
  if (i > E.N_eps - 2) i = E.N_eps - 2;
  int j = cvt_int((logr - E.xlr0) * E.inv_del_lr);
  if (j < 0) j = 0;
  if (j > E.N_lr - 2) j = E.N_lr - 2;
  const float* T = E.tab + (size_t)t * E.N_eps * E.N_lr;
  const long L = E.N_lr;
  const double T11 = T[(i + 1) * L + j + 1], T01 = T[i * L + j + 1], T10 = T[(i + 1) * L + j], T00 = T[i * L + j];
  const double v = (logr - E.xlr[j]) * (ee1 - E.xeps[i]) * T11 + (logr - E.xlr[j]) * (E.xeps[i + 1] - ee1) * T01 +
                   (E.xlr[j + 1] - logr) * (ee1 - E.xeps[i]) * T10 + (E.xlr[j + 1] - logr) * (E.xeps[i + 1] - ee1) * T00;
  const double dv_dee1 = (logr - E.xlr[j]) * (T11 - T01) + (E.xlr[j + 1] - logr) * (T10 - T00);
  const double dv_dlogr = (ee1 - E.xeps[i]) * (T11 - T10) + (E.xeps[i + 1] - ee1) * (T01 - T00);
  dee = dv_dee1 / (ee + E.eps_off);
  ddd = dv_dlogr / dd;
  return v;
}
// derivative of inv_raw(t, pp, ss) (eos_init.C:216-265) w.r.t. (pp, ss)
MRM_HD double inv_raw_d(const EosDev& E, int t, double pp, double ss, double& dpp, double& dss) {
  const double ss1 = ss + E.ss_off;
  const double logp = log(pp);
  int i = cvt_int((logp - E.xlp30) * E.inv_del_lp3);
  if (i < 0) i = 0;
  if (i > E.N_lp3 - 2) i = E.N_lp3 - 2;
  int j = cvt_int((ss1 - E.xs30) * E.inv_del_s3);
  if (j < 0) j = 0;
  if (j > E.N_s3 - 2) j = E.N_s3 - 2;
  const float* T = E.itab + (size_t)t * E.N_lp3 * E.N_s3;
  const long L = E.N_s3;
  const double T11 = T[(i + 1) * L + j + 1], T10 = T[(i + 1) * L + j], T01 = T[i * L + j + 1], T00 = T[i * L + j];
  const double v = (logp - E.xlp3[i]) * (ss1 - E.xs3[j]) * T11 + (logp - E.xlp3[i]) * (E.xs3[j + 1] - ss1) * T10 +
                   (E.xlp3[i + 1] - logp) * (ss1 - E.xs3[j]) * T01 + (E.xlp3[i + 1] - logp) * (E.xs3[j + 1] - ss1) * T00;
  const double dv_dlogp = (ss1 - E.xs3[j]) * (T11 - T01) + (E.xs3[j + 1] - ss1) * (T10 - T00);
  const double dv_dss1 = (logp - E.xlp3[i]) * (T11 - T10) + (E.xlp3[i + 1] - logp) * (T01 - T00);
  dpp = dv_dlogp / pp;
  dss = dv_dss1;
  return v;
}

// p of the bottom-layer sums (k_bc_sums<0>, boundary_pdmp_1_fftw3.C:172-202):
// p_interp(e_int/rho, rho) with e_int = e - M^2/(2 rho); VJP into lu
MRM_HD void bc_sum_p_vjp(const EosDev& E, const double u[MRM_NV], double lp, double lu[MRM_NV]) {
  const double d = u[0];
  const double msq = u[1] * u[1] + u[2] * u[2] + u[3] * u[3];
  const double e = u[4] - 0.5 * (msq / d);
  const double ee = e / d;
  double dee, ddd;
  const double v = eos_fwd_raw_d(E, 0, ee, d, dee, ddd);
  const double p = exp(v);
  const double lee = lp * p * dee;
  double ld = lp * p * ddd;
  const double le = lee / d;
  ld += -lee * e / (d * d);
  lu[4] += le;
  ld += le * 0.5 * msq / (d * d);
  lu[1] += -le * u[1] / d;
  lu[2] += -le * u[2] / d;
  lu[3] += -le * u[3] / d;
  lu[0] += ld;
}

// bottom boundary column (k_bc_bottom, boundary_pdmp_1_fftw3.C:240-318):
// inputs the two lowest interior cells un1 (lbeg), un2 (lbeg+1) and the layer
// means pm1, pm2; outputs the ghost cells g1 (lbeg-1), g2 (lbeg-2).  Given
// their cotangents lg1, lg2: adds to lu1, lu2, lpm1, lpm2.
MRM_HD void bc_bottom_vjp(const EosDev& E, const BcParams& B, const double un1[MRM_NV], const double un2[MRM_NV],
                          double pm1, double pm2, const double lg1[MRM_NV], const double lg2[MRM_NV],
                          double lu1[MRM_NV], double lu2[MRM_NV], double& lpm1, double& lpm2) {
  const double p1_dmp = 0.8;
  const double d1 = un1[0], d2 = un2[0];
  const double ms1 = un1[1] * un1[1] + un1[2] * un1[2] + un1[3] * un1[3];
  const double ms2 = un2[1] * un2[1] + un2[2] * un2[2] + un2[3] * un2[3];
  const double ei1 = un1[4] - 0.5 * ms1 / d1;
  const double ei2 = un2[4] - 0.5 * ms2 / d2;
  const double eps1 = ei1 / d1, eps2 = ei2 / d2;
  const double pmag = 0.5 * un1[5] * un1[5];
  double pi_e, pi_d, s1_e, s1_d, s2_e, s2_d;
  const double pint = exp(eos_fwd_raw_d(E, 0, eps1, d1, pi_e, pi_d));
  eos_fwd_raw_d(E, 2, eps1, d1, s1_e, s1_d);
  eos_fwd_raw_d(E, 2, eps2, d2, s2_e, s2_d);
  const double smin1 = eos_fwd_raw(E, 2, eps1, d1) - E.ss_off;
  const double smin2 = eos_fwd_raw(E, 2, eps2, d2) - E.ss_off;
  const bool inflow = un1[1] >= 0.e0;
  const double s1 = inflow ? B.s_bc : smin1, s2 = inflow ? B.s_bc : smin2;
  const double sq = sqrt(pm1 * pm2);
  const double c = B.p_bc / sq;
  const double pmin1 = pint - pm1;
  const double p1 = pm1 * c + pmin1 * p1_dmp - pmag * (1.0 - p1_dmp);
  const double p2 = pm1 * c * c + pmin1 * p1_dmp * p1_dmp - pmag * (1.0 - p1_dmp * p1_dmp);
  const double cb = un1[5] / B.B_crit;
  const double cb4 = pow(cb, 4.0);
  const double fac = (1.0 - cb4) / (1.0 + cb4);
  double m1[3], m2[3];
  for (int k = 0; k < 3; k++) { m1[k] = un1[1 + k] * fac; m2[k] = un2[1 + k] * fac; }
  double r1p, r1s, r2p, r2s, e1p, e1s, e2p, e2s;
  const double rho1 = exp(inv_raw_d(E, 1, p1, s1, r1p, r1s));
  const double rho2 = exp(inv_raw_d(E, 1, p2, s2, r2p, r2s));
  const double ex1 = exp(inv_raw_d(E, 0, p1, s1, e1p, e1s));
  const double ex2 = exp(inv_raw_d(E, 0, p2, s2, e2p, e2s));
  const double eps1n = ex1 - E.eps_off, eps2n = ex2 - E.eps_off;
  const double mm1 = m1[0] * m1[0] + m1[1] * m1[1] + m1[2] * m1[2];
  const double mm2 = m2[0] * m2[0] + m2[1] * m2[1] + m2[2] * m2[2];
  // g1 = (rho1, m1, eps1n*rho1 + mm1/(2 rho1), B of un1); g2 likewise with un2's B
  // reverse
  double lrho1 = lg1[0], lrho2 = lg2[0];
  double lm1[3] = {lg1[1], lg1[2], lg1[3]}, lm2[3] = {lg2[1], lg2[2], lg2[3]};
  lrho1 += lg1[4] * (eps1n - 0.5 * mm1 / (rho1 * rho1));
  lrho2 += lg2[4] * (eps2n - 0.5 * mm2 / (rho2 * rho2));
  const double leps1n = lg1[4] * rho1, leps2n = lg2[4] * rho2;
  for (int k = 0; k < 3; k++) {
    lm1[k] += lg1[4] * m1[k] / rho1;
    lm2[k] += lg2[4] * m2[k] / rho2;
  }
  lu1[5] += lg1[5]; lu1[6] += lg1[6]; lu1[7] += lg1[7];
  lu2[5] += lg2[5]; lu2[6] += lg2[6]; lu2[7] += lg2[7];
  // rho = exp(inv_raw(1, p, s)), eps = exp(inv_raw(0, p, s)) - eps_off
  const double lp1 = lrho1 * rho1 * r1p + leps1n * ex1 * e1p;
  const double ls1 = lrho1 * rho1 * r1s + leps1n * ex1 * e1s;
  const double lp2 = lrho2 * rho2 * r2p + leps2n * ex2 * e2p;
  const double ls2 = lrho2 * rho2 * r2s + leps2n * ex2 * e2s;
  // momentum scaling by the spot factor
  double lfac = 0.0;
  for (int k = 0; k < 3; k++) {
    lu1[1 + k] += lm1[k] * fac;
    lu2[1 + k] += lm2[k] * fac;
    lfac += lm1[k] * un1[1 + k] + lm2[k] * un2[1 + k];
  }
  // fac = (1 - cb4)/(1 + cb4), cb4 = (B0/B_crit)^4
  const double lcb4 = lfac * (-2.0 / ((1.0 + cb4) * (1.0 + cb4)));
  lu1[5] += lcb4 * 4.0 * cb * cb * cb / B.B_crit;
  // p1, p2
  const double lpmin1 = lp1 * p1_dmp + lp2 * p1_dmp * p1_dmp;
  const double lpmag = -lp1 * (1.0 - p1_dmp) - lp2 * (1.0 - p1_dmp * p1_dmp);
  const double lc = lp1 * pm1 + lp2 * pm1 * 2.0 * c;
  lpm1 += lp1 * c + lp2 * c * c - lpmin1;
  // c = p_bc / sqrt(pm1 pm2)
  lpm1 += lc * (-0.5 * c / pm1);
  lpm2 += lc * (-0.5 * c / pm2);
  lu1[5] += lpmag * un1[5];
  // entropies of the outflow branch
  double leps1 = 0.0, ld1 = 0.0, leps2 = 0.0, ld2 = 0.0;
  if (!inflow) {
    leps1 += ls1 * s1_e; ld1 += ls1 * s1_d;
    leps2 += ls2 * s2_e; ld2 += ls2 * s2_d;
  }
  // pint = p_interp(eps1, d1)
  leps1 += lpmin1 * pint * pi_e;
  ld1 += lpmin1 * pint * pi_d;
  // eps = ei/d, ei = e - M^2/(2d)
  auto back = [&](const double* un, double ei, double d, double ms, double leps, double ldd, double* lu) {
    const double lei = leps / d;
    ldd += -leps * ei / (d * d);
    lu[4] += lei;
    ldd += lei * 0.5 * ms / (d * d);
    lu[1] += -lei * un[1] / d;
    lu[2] += -lei * un[2] / d;
    lu[3] += -lei * un[3] / d;
    lu[0] += ldd;
  };
  back(un1, ei1, d1, ms1, leps1, ld1, lu1);
  back(un2, ei2, d2, ms2, leps2, ld2, lu2);
}

// top boundary column (k_bc_top, boundary_pdmp_1_fftw3.C:402-474): ghosts
// g1 (lend+1) from u1 (lend) and g2 (lend+2) from u2 (lend-1)
MRM_HD void bc_top_vjp(const BcParams& B, const double u1[MRM_NV], const double u2[MRM_NV],
                       const double lg1[MRM_NV], const double lg2[MRM_NV], double lu1[MRM_NV], double lu2[MRM_NV]) {
  double s = -1.0;
  if (B.open_top == 1) s = (u1[1] < 0.0) ? -1.0 : 1.0;
  double a1[MRM_NV], a2[MRM_NV];
  for (int q = 0; q < MRM_NV; q++) { a1[q] = u1[q]; a2[q] = u2[q]; }
  a1[1] *= s;
  a2[1] *= s;
  double l1[MRM_NV], l2[MRM_NV];
  for (int q = 0; q < MRM_NV; q++) { l1[q] = lg1[q]; l2[q] = lg2[q]; }
  l1[6] = -lg1[6]; l1[7] = -lg1[7]; l2[6] = -lg2[6]; l2[7] = -lg2[7];
  if (B.eps_top > 0.0) {
    auto back = [&](const double* a, double* l) {
      const double le = l[4];
      const double msq = a[1] * a[1] + a[2] * a[2] + a[3] * a[3];
      l[0] += le * (B.eps_top - 0.5 * msq / (a[0] * a[0]));
      l[1] += le * a[1] / a[0];
      l[2] += le * a[2] / a[0];
      l[3] += le * a[3] / a[0];
      l[4] = 0.0;
    };
    back(a1, l1);
    back(a2, l2);
  }
  l1[1] *= s;
  l2[1] *= s;
  for (int q = 0; q < MRM_NV; q++) { lu1[q] += l1[q]; lu2[q] += l2[q]; }
}


// ================================================================== kernels
// All adjoint buffers are SoA over the whole grid buffer (field f of node n
// at buf[f*bs + n]).

__device__ __forceinline__ void ld8(const double* __restrict__ U, long bs, long n, double u[MRM_NV]) {
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) u[q] = U[q * bs + n];
}

// ---- stage, pass 1 (one thread per interior cell c): forward of the cell,
// then the cotangent of U_out(c) back through Tcheck, RK, sources, Boris and
// the sweeps' centre terms.  Writes G_D(c) (13 per sweep: cotangents of the
// flux differences), the centre-primitive cotangent lWc(c), and adds the U0
// cotangent to lX0(c).
__global__ void __launch_bounds__(MRM_BLOCK) k_adj_stage1(StageArgs A, const double* __restrict__ Xin,
                                                         const double* __restrict__ Pin,
                                                         const double* __restrict__ X0,
                                                         const double* __restrict__ lamY,
                                                         const double* __restrict__ Q, double* __restrict__ Gb,
                                                         double* __restrict__ lWc, double* __restrict__ lX0) {
  const Geom& G = A.G;
  const long bs = G.bs;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int i, j, k;
  if (!interior_ijk(G, t, i, j, k)) return;
  const long n = i + j * G.s1 + k * G.s2;
  const long str[3] = {1, G.s1, G.s2};
  double c[ST_NF];
  load_prim(Xin, Pin, bs, n, c);
  const double lf = lfac_of(c, A.RP.inv_va2max);
  double Dl[3][13];
  double acc[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) acc[q] = 0.0;
#pragma unroll
  for (int D = 0; D < 3; D++) {
    double wm2[ST_NF], wm1[ST_NF], wp1[ST_NF], wp2[ST_NF];
    load_prim(Xin, Pin, bs, n - 2 * str[D], wm2);
    load_prim(Xin, Pin, bs, n - str[D], wm1);
    load_prim(Xin, Pin, bs, n + str[D], wp1);
    load_prim(Xin, Pin, bs, n + 2 * str[D], wp2);
    double fm2[13], fm1[13], fp1[13], fp2[13], res[MRM_NV];
    auto get = [&](int s, double o[ST_NF]) {
      const double* src = s == 0 ? wm2 : s == 1 ? wm1 : s == 3 ? wp1 : wp2;
#pragma unroll
      for (int f = 0; f < ST_NF; f++) o[f] = src[f];
    };
    if (D == 0) { flux13<0>(wm2, fm2); flux13<0>(wm1, fm1); flux13<0>(wp1, fp1); flux13<0>(wp2, fp2);
                  sweep_residual_lr<0>(get, c, A.RP.w1[0], A.RP.w2[0], lf, res); }
    else if (D == 1) { flux13<1>(wm2, fm2); flux13<1>(wm1, fm1); flux13<1>(wp1, fp1); flux13<1>(wp2, fp2);
                       sweep_residual_lr<1>(get, c, A.RP.w1[1], A.RP.w2[1], lf, res); }
    else { flux13<2>(wm2, fm2); flux13<2>(wm1, fm1); flux13<2>(wp1, fp1); flux13<2>(wp2, fp2);
           sweep_residual_lr<2>(get, c, A.RP.w1[2], A.RP.w2[2], lf, res); }
#pragma unroll
    for (int q = 0; q < 13; q++) Dl[D][q] = A.RP.w1[D] * (fp1[q] - fm1[q]) + A.RP.w2[D] * (fp2[q] - fm2[q]);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
  }
  double u0[MRM_NV], lo[MRM_NV];
  ld8(X0, bs, n, u0);
  ld8(lamY, bs, n, lo);
  const double Qrad = Q ? Q[n] : 0.0;
  double lup[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0}, lR[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  double lu0[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  source_integrate_vjp(A.SP, c, acc, u0, Qrad, lo, lup, lR, lu0);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) lX0[q * bs + n] += lu0[q];
  double lc[ST_NF] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double llf = 0.0, Gd[13];
  sweep_centre_vjp<0>(c, Dl[0], lf, lR, Gd, lc, llf);
#pragma unroll
  for (int q = 0; q < 13; q++) Gb[(0 * 13 + q) * bs + n] = Gd[q];
  sweep_centre_vjp<1>(c, Dl[1], lf, lR, Gd, lc, llf);
#pragma unroll
  for (int q = 0; q < 13; q++) Gb[(1 * 13 + q) * bs + n] = Gd[q];
  sweep_centre_vjp<2>(c, Dl[2], lf, lR, Gd, lc, llf);
#pragma unroll
  for (int q = 0; q < 13; q++) Gb[(2 * 13 + q) * bs + n] = Gd[q];
  lfac_vjp(c, A.RP.inv_va2max, llf, lc);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) lWc[q * bs + n] = lup[q] + lc[q];
}

// ---- stage, pass 2 (one thread per buffer point p): gather the flux
// cotangents of p from the G of the interior cells whose stencils contain p
// (the transpose of the centred differences), add the centre term, and
// propagate through the flux and the cons->prim map into lX(p) and lP(p).
__global__ void __launch_bounds__(MRM_BLOCK) k_adj_stage2(Geom G, ResidParams RP, const double* __restrict__ Xin,
                                                         const double* __restrict__ Pin,
                                                         const double* __restrict__ Gb,
                                                         const double* __restrict__ lWc, double* __restrict__ lX,
                                                         double* __restrict__ lP) {
  const long bs = G.bs;
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= bs) return;
  const int i = (int)(n % G.n0);
  const long r = n / G.n0;
  const int j = (int)(r % G.n1), k = (int)(r / G.n1);
  const int ijk[3] = {i, j, k};
  const int lbv[3] = {G.lb0, G.lb1, G.lb2}, lev[3] = {G.le0, G.le1, G.le2};
  const long str[3] = {1, G.s1, G.s2};
  auto inside = [&](int d, int off) {   // is p + off e_d an interior cell?
    for (int e = 0; e < 3; e++) {
      const int x = ijk[e] + (e == d ? off : 0);
      if (x < lbv[e] || x > lev[e]) return false;
    }
    return true;
  };
  double lw[ST_NF] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  bool any = false;
  if (inside(0, 0)) {
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) lw[q] = lWc[q * bs + n];
    any = true;
  }
  double w[ST_NF];
  bool have_w = false;
#pragma unroll
  for (int D = 0; D < 3; D++) {
    const bool am1 = inside(D, -1), ap1 = inside(D, 1), am2 = inside(D, -2), ap2 = inside(D, 2);
    if (!(am1 || ap1 || am2 || ap2)) continue;
    double lf13[13];
#pragma unroll
    for (int q = 0; q < 13; q++) {
      const double* g = Gb + (long)(D * 13 + q) * bs + n;
      double v = 0.0;
      if (am1) v += RP.w1[D] * g[-str[D]];
      if (ap1) v -= RP.w1[D] * g[str[D]];
      if (am2) v += RP.w2[D] * g[-2 * str[D]];
      if (ap2) v -= RP.w2[D] * g[2 * str[D]];
      lf13[q] = v;
    }
    if (!have_w) { load_prim(Xin, Pin, bs, n, w); have_w = true; }
    if (D == 0) flux13_vjp<0>(w, lf13, lw);
    else if (D == 1) flux13_vjp<1>(w, lf13, lw);
    else flux13_vjp<2>(w, lf13, lw);
    any = true;
  }
  if (!any) return;
  double u[MRM_NV], lu[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  ld8(Xin, bs, n, u);
  prim_vjp(u, lw, lu);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) lX[q * bs + n] += lu[q];
  lP[n] += lw[8];
}

// ---- EOS pressure of every buffer point: lX += (dp/dX)^T lP
__global__ void __launch_bounds__(MRM_BLOCK) k_adj_eos(Geom G, EosDev E, const double* __restrict__ X,
                                                      const double* __restrict__ lP, double* __restrict__ lX) {
  const long bs = G.bs;
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= bs) return;
  const double lp = lP[n];
  if (lp == 0.0) return;
  double u[MRM_NV], lu[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  ld8(X, bs, n, u);
  eos_p_vjp(E, u, lp, lu);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) lX[q * bs + n] += lu[q];
}

// ---- TCheck of the interior, in place on the cotangent
__global__ void __launch_bounds__(MRM_BLOCK) k_adj_tcheck(Geom G, double rho_min, double eps_min, double eps_max,
                                                         double vmax, const double* __restrict__ U,
                                                         double* __restrict__ lam) {
  const long bs = G.bs;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int i, j, k;
  if (!interior_ijk(G, t, i, j, k)) return;
  const long n = i + j * G.s1 + k * G.s2;
  double u[MRM_NV], lo[MRM_NV], lu[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  ld8(U, bs, n, u);
  ld8(lam, bs, n, lo);
  tcheck_vjp(u, rho_min, eps_min, eps_max, vmax, lo, lu);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) lam[q * bs + n] = lu[q];
}

// ---- periodic ghost copies of dim DIM, reversed: lam(src) += lam(dst), lam(dst) = 0
template <int DIM>
__global__ void __launch_bounds__(MRM_BLOCK) k_adj_exchange(Geom G, double* __restrict__ lam) {
  const int nx = G.le0 - G.lb0 + 1;
  const int gh = DIM == 1 ? G.g1 : G.g2;
  const int lb = DIM == 1 ? G.lb1 : G.lb2, le = DIM == 1 ? G.le1 : G.le2;
  const int nb = DIM == 1 ? (G.le2 - G.lb2 + 1) : G.n1;
  const int ob = DIM == 1 ? G.lb2 : 0;
  const long sd = DIM == 1 ? G.s1 : G.s2, so = DIM == 1 ? G.s2 : G.s1;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nx * nb * gh) return;
  int i = G.lb0 + (int)(t % nx);
  long r = t / nx;
  int b = ob + (int)(r % nb);
  int igh = (int)(r / nb);
  long base = i + b * so;
  long src_r = base + (long)(lb + igh) * sd, dst_r = base + (long)(le + 1 + igh) * sd;
  long src_l = base + (long)(le - gh + 1 + igh) * sd, dst_l = base + (long)(lb - gh + igh) * sd;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    const long o = q * G.bs;
    lam[o + src_r] += lam[o + dst_r];
    lam[o + dst_r] = 0.0;
    lam[o + src_l] += lam[o + dst_l];
    lam[o + dst_l] = 0.0;
  }
}

// ---- bottom and top ghost layers, reversed, per (j,k) column of the whole
// transverse range.  Z is the state the boundary read; psum its layer sums.
// The cotangents of the two layer means are reduced into part[2][nblocks].
__global__ void __launch_bounds__(MRM_BLOCK) k_adj_bc(Geom G, EosDev E, BcParams B, const double* __restrict__ psum,
                                                     const double* __restrict__ Z, double* __restrict__ lam,
                                                     double* __restrict__ part) {
  __shared__ double lds[2 * MRM_BLOCK / 64];
  const long bs = G.bs;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double lpm[2] = {0.0, 0.0};
  if (t < (long)G.n1 * G.n2) {
    const int j = (int)(t % G.n1), k = (int)(t / G.n1);
    const long base = j * G.s1 + k * G.s2;
    {
      const long n1 = G.lb0 + base, n2 = n1 + 1, g1 = G.lb0 - 1 + base, g2 = G.lb0 - 2 + base;
      double u1[MRM_NV], u2[MRM_NV], l1[MRM_NV], l2[MRM_NV];
      double lu1[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0}, lu2[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
      ld8(Z, bs, n1, u1);
      ld8(Z, bs, n2, u2);
      ld8(lam, bs, g1, l1);
      ld8(lam, bs, g2, l2);
      bc_bottom_vjp(E, B, u1, u2, psum[0] * B.cc, psum[1] * B.cc, l1, l2, lu1, lu2, lpm[0], lpm[1]);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) {
        lam[q * bs + n1] += lu1[q];
        lam[q * bs + n2] += lu2[q];
        lam[q * bs + g1] = 0.0;
        lam[q * bs + g2] = 0.0;
      }
    }
    {
      const long n1 = G.le0 + base, n2 = n1 - 1, g1 = G.le0 + 1 + base, g2 = G.le0 + 2 + base;
      double u1[MRM_NV], u2[MRM_NV], l1[MRM_NV], l2[MRM_NV];
      double lu1[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0}, lu2[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
      ld8(Z, bs, n1, u1);
      ld8(Z, bs, n2, u2);
      ld8(lam, bs, g1, l1);
      ld8(lam, bs, g2, l2);
      bc_top_vjp(B, u1, u2, l1, l2, lu1, lu2);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) {
        lam[q * bs + n1] += lu1[q];
        lam[q * bs + n2] += lu2[q];
        lam[q * bs + g1] = 0.0;
        lam[q * bs + g2] = 0.0;
      }
    }
  }
  block_reduce<OpAdd, 2>(lpm, lds);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = lpm[0];
    part[gridDim.x + blockIdx.x] = lpm[1];
  }
}

// ---- the layer means' cotangents back into the two lowest interior layers
// (pmean = cc * sum over the interior columns of p_interp)
__global__ void __launch_bounds__(MRM_BLOCK) k_adj_bc_means(Geom G, EosDev E, double cc,
                                                           const double* __restrict__ lpm,
                                                           const double* __restrict__ Z, double* __restrict__ lam) {
  const long bs = G.bs;
  const int ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)ny * nz) return;
  const int j = G.lb1 + (int)(t % ny), k = G.lb2 + (int)(t / ny);
  const long n1 = G.lb0 + j * G.s1 + k * G.s2, n2 = n1 + 1;
  double u[MRM_NV], lu[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
  ld8(Z, bs, n1, u);
  bc_sum_p_vjp(E, u, cc * lpm[0], lu);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) { lam[q * bs + n1] += lu[q]; lu[q] = 0.0; }
  ld8(Z, bs, n2, u);
  bc_sum_p_vjp(E, u, cc * lpm[1], lu);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) lam[q * bs + n2] += lu[q];
}

// ---- TVD sweep D (tvdlimit_SR.C:33-695), transposed in gather form.  A
// point is on a line of sweep D as the forward defines the line sets; s is
// its coordinate along D, sz the line length.
struct TvdLine {
  bool on;
  int s, sz, gh, ix;
  long str;
};
template <int D>
__device__ __forceinline__ TvdLine tvd_line_of(const Geom& G, long n) {
  TvdLine L;
  const int i = (int)(n % G.n0);
  const long r = n / G.n0;
  const int j = (int)(r % G.n1), k = (int)(r / G.n1);
  L.ix = i;
  if (D == 0) { L.on = true; L.s = i; L.sz = G.n0; L.gh = G.g0; L.str = 1; }
  else if (D == 1) { L.on = i >= G.lb0 && i <= G.le0; L.s = j; L.sz = G.n1; L.gh = G.g1; L.str = G.s1; }
  else { L.on = i >= G.lb0 && i <= G.le0 && j >= G.lb1 && j <= G.le1; L.s = k; L.sz = G.n2; L.gh = G.g2; L.str = G.s2; }
  return L;
}

// scratch field offsets (units of bs)
enum {
  TA_W = 0,                      // TW_N point fields
  TA_S = TA_W + TW_N,            // 8 slopes
  TA_F = TA_S + MRM_NV,          // 8 interface fluxes (interface s: between s and s+1)
  TA_Q = TA_F + MRM_NV,          // resistive term of interface s
  TA_UFL = TA_Q + 1,             // update of point s: cotangent of F(s-1)
  TA_UFR = TA_UFL + MRM_NV,      //                    cotangent of F(s)
  TA_UQL = TA_UFR + MRM_NV,
  TA_UQR = TA_UQL + 1,
  TA_UW = TA_UQR + 1,            // TW_N point-field cotangents from the update
  TA_IA = TA_UW + TW_N,          // interface s: la (3)
  TA_IB = TA_IA + 3,             //              lb (TW_N)
  TA_IC = TA_IB + TW_N,          //              lc (TW_N)
  TA_ID = TA_IC + TW_N,          //              ld (3)
  TA_ISB = TA_ID + 3,            //              lSb (8)
  TA_ISC = TA_ISB + MRM_NV,      //              lSc (8)
  TA_SA = TA_ISC + MRM_NV,       // slope s: cotangents of points s-1, s, s+1 (8 each)
  TA_SB = TA_SA + MRM_NV,
  TA_SC = TA_SB + MRM_NV,
  TA_N = TA_SC + MRM_NV
};

template <int D>
__global__ void __launch_bounds__(MRM_BLOCK) k_adj_tvd_fwd(Geom G, TvdParams P, const double* __restrict__ U,
                                                          const double* __restrict__ Pr, double* __restrict__ T,
                                                          int phase) {
  const long bs = G.bs;
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= bs) return;
  const TvdLine L = tvd_line_of<D>(G, n);
  if (!L.on) return;
  const int ix = L.ix;
  if (phase == 0) {
    double u[MRM_NV];
    ld8(U, bs, n, u);
    const double tau = P.tau ? P.tau[n] : 0.0;
    TvdPt w;
    tvd_point(P, u, Pr[n], tau, P.hfb[ix], P.hft[ix], D == 0, w);
    tvdpt_store(w, T + TA_W * bs, bs, n);
  } else if (phase == 1) {
    if (L.s < 1 || L.s > L.sz - 2) return;
    TvdPt a, b, c;
    tvdpt_load(a, T + TA_W * bs, bs, n - L.str);
    tvdpt_load(b, T + TA_W * bs, bs, n);
    tvdpt_load(c, T + TA_W * bs, bs, n + L.str);
    double S[MRM_NV];
    tvd_slope(P, a, b, c, P.hfb[ix], P.hft[ix], S);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) T[(TA_S + q) * bs + n] = S[q];
  } else {
    if (L.s < 1 || L.s > L.sz - 3) return;
    TvdPt a, b, c, d;
    tvdpt_load(a, T + TA_W * bs, bs, n - L.str);
    tvdpt_load(b, T + TA_W * bs, bs, n);
    tvdpt_load(c, T + TA_W * bs, bs, n + L.str);
    tvdpt_load(d, T + TA_W * bs, bs, n + 2 * L.str);
    double Sb[MRM_NV], Sc[MRM_NV], F[MRM_NV];
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) { Sb[q] = T[(TA_S + q) * bs + n]; Sc[q] = T[(TA_S + q) * bs + n + L.str]; }
    const int istart = 1 + (P.is_gbeg0 ? L.gh : 0);
    const bool zb = (D == 0) && ((P.is_gend0 && L.s == L.sz - L.gh - 1) || (P.is_gbeg0 && L.s == L.gh - 1));
    const double q = tvd_iface<D>(P, a, b, c, d, Sb, Sc, P.hfb[ix], P.hft[ix], P.qft[ix], L.s >= istart, zb, F);
#pragma unroll
    for (int qq = 0; qq < MRM_NV; qq++) T[(TA_F + qq) * bs + n] = F[qq];
    T[TA_Q * bs + n] = q;
  }
}

// reverse phases: 0 update, 1 interfaces, 2 slopes, 3 points (writes lamIn, lP)
template <int D>
__global__ void __launch_bounds__(MRM_BLOCK) k_adj_tvd_rev(Geom G, TvdParams P, const double* __restrict__ U,
                                                          const double* __restrict__ Pr, double* __restrict__ T,
                                                          const double* __restrict__ lamOut,
                                                          double* __restrict__ lamIn, double* __restrict__ lP,
                                                          int phase) {
  const long bs = G.bs;
  long n = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (n >= bs) return;
  const TvdLine L = tvd_line_of<D>(G, n);
  if (phase == 3 && !L.on) {
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) lamIn[q * bs + n] = lamOut[q * bs + n];
    return;
  }
  if (!L.on) return;
  const int ix = L.ix;
  const bool updated = L.s >= L.gh && L.s <= L.sz - L.gh - 1;
  if (phase == 0) {
    if (!updated) return;
    TvdPt b;
    tvdpt_load(b, T + TA_W * bs, bs, n);
    double Fm[MRM_NV], F[MRM_NV], u[MRM_NV], lo[MRM_NV];
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) { Fm[q] = T[(TA_F + q) * bs + n - L.str]; F[q] = T[(TA_F + q) * bs + n]; }
    ld8(U, bs, n, u);
    ld8(lamOut, bs, n, lo);
    double lu[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0}, lF[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
    double lFm[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0}, lq = 0.0, lqm = 0.0, lb[TW_N];
#pragma unroll
    for (int q = 0; q < TW_N; q++) lb[q] = 0.0;
    tvd_update_vjp(P, b, Fm, F, T[TA_Q * bs + n - L.str], T[TA_Q * bs + n], P.qft[ix], u, lo, lu, lF, lFm, lq, lqm,
                   lb);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) {
      lamIn[q * bs + n] = lu[q];
      T[(TA_UFL + q) * bs + n] = lFm[q];
      T[(TA_UFR + q) * bs + n] = lF[q];
    }
    T[TA_UQL * bs + n] = lqm;
    T[TA_UQR * bs + n] = lq;
#pragma unroll
    for (int q = 0; q < TW_N; q++) T[(TA_UW + q) * bs + n] = lb[q];
  } else if (phase == 1) {
    if (L.s < 1 || L.s > L.sz - 3) return;
    TvdPt a, b, c, d;
    tvdpt_load(a, T + TA_W * bs, bs, n - L.str);
    tvdpt_load(b, T + TA_W * bs, bs, n);
    tvdpt_load(c, T + TA_W * bs, bs, n + L.str);
    tvdpt_load(d, T + TA_W * bs, bs, n + 2 * L.str);
    double Sb[MRM_NV], Sc[MRM_NV], lF[MRM_NV];
    const bool upd_here = L.s >= L.gh && L.s <= L.sz - L.gh - 1;
    const bool upd_next = L.s + 1 >= L.gh && L.s + 1 <= L.sz - L.gh - 1;
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) {
      Sb[q] = T[(TA_S + q) * bs + n];
      Sc[q] = T[(TA_S + q) * bs + n + L.str];
      lF[q] = (upd_here ? T[(TA_UFR + q) * bs + n] : 0.0) + (upd_next ? T[(TA_UFL + q) * bs + n + L.str] : 0.0);
    }
    const double lq = (upd_here ? T[TA_UQR * bs + n] : 0.0) + (upd_next ? T[TA_UQL * bs + n + L.str] : 0.0);
    const int istart = 1 + (P.is_gbeg0 ? L.gh : 0);
    const bool zb = (D == 0) && ((P.is_gend0 && L.s == L.sz - L.gh - 1) || (P.is_gbeg0 && L.s == L.gh - 1));
    TvdIfaceAdj A;
    tvd_iface_vjp<D>(P, a, b, c, d, Sb, Sc, P.hfb[ix], P.hft[ix], P.qft[ix], L.s >= istart, zb, lF, lq, A);
#pragma unroll
    for (int q = 0; q < 3; q++) { T[(TA_IA + q) * bs + n] = A.la[q]; T[(TA_ID + q) * bs + n] = A.ld[q]; }
#pragma unroll
    for (int q = 0; q < TW_N; q++) { T[(TA_IB + q) * bs + n] = A.lb[q]; T[(TA_IC + q) * bs + n] = A.lc[q]; }
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) { T[(TA_ISB + q) * bs + n] = A.lSb[q]; T[(TA_ISC + q) * bs + n] = A.lSc[q]; }
  } else if (phase == 2) {
    if (L.s < 1 || L.s > L.sz - 2) return;
    TvdPt a, b, c;
    tvdpt_load(a, T + TA_W * bs, bs, n - L.str);
    tvdpt_load(b, T + TA_W * bs, bs, n);
    tvdpt_load(c, T + TA_W * bs, bs, n + L.str);
    double lS[MRM_NV];
    const bool if_here = L.s <= L.sz - 3;     // interface s reads slope s as Sb
    const bool if_prev = L.s - 1 >= 1;        // interface s-1 reads slope s as Sc
#pragma unroll
    for (int q = 0; q < MRM_NV; q++)
      lS[q] = (if_here ? T[(TA_ISB + q) * bs + n] : 0.0) + (if_prev ? T[(TA_ISC + q) * bs + n - L.str] : 0.0);
    double la[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0}, lb[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0},
           lc[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0};
    tvd_slope_vjp(P, a, b, c, P.hfb[ix], P.hft[ix], lS, la, lb, lc);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) {
      T[(TA_SA + q) * bs + n] = la[q];
      T[(TA_SB + q) * bs + n] = lb[q];
      T[(TA_SC + q) * bs + n] = lc[q];
    }
  } else {
    // point: gather every cotangent of its TvdPt fields
    double lw[TW_N];
#pragma unroll
    for (int q = 0; q < TW_N; q++) lw[q] = updated ? T[(TA_UW + q) * bs + n] : 0.0;
    const int s = L.s, sz = L.sz;
    const long st = L.str;
    const int hk[3] = {0, 1, 4};
    if (s + 1 >= 1 && s + 1 <= sz - 3) {       // 'a' of interface s+1
#pragma unroll
      for (int t = 0; t < 3; t++) lw[TW_V + hk[t]] += T[(TA_IA + t) * bs + n + st];
    }
    if (s >= 1 && s <= sz - 3) {               // 'b' of interface s
#pragma unroll
      for (int q = 0; q < TW_N; q++) lw[q] += T[(TA_IB + q) * bs + n];
    }
    if (s - 1 >= 1 && s - 1 <= sz - 3) {       // 'c' of interface s-1
#pragma unroll
      for (int q = 0; q < TW_N; q++) lw[q] += T[(TA_IC + q) * bs + n - st];
    }
    if (s - 2 >= 1 && s - 2 <= sz - 3) {       // 'd' of interface s-2
#pragma unroll
      for (int t = 0; t < 3; t++) lw[TW_V + hk[t]] += T[(TA_ID + t) * bs + n - 2 * st];
    }
    if (s + 1 >= 1 && s + 1 <= sz - 2) {       // 'a' of slope s+1
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) lw[TW_V + q] += T[(TA_SA + q) * bs + n + st];
    }
    if (s >= 1 && s <= sz - 2) {               // 'b' of slope s
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) lw[TW_V + q] += T[(TA_SB + q) * bs + n];
    }
    if (s - 1 >= 1 && s - 1 <= sz - 2) {       // 'c' of slope s-1
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) lw[TW_V + q] += T[(TA_SC + q) * bs + n - st];
    }
    double u[MRM_NV], lu[MRM_NV] = {0, 0, 0, 0, 0, 0, 0, 0}, lp = 0.0;
    ld8(U, bs, n, u);
    const double tau = P.tau ? P.tau[n] : 0.0;
    tvd_point_vjp(P, u, Pr[n], tau, P.hfb[ix], P.hft[ix], D == 0, lw, lu, lp);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++)
      lamIn[q * bs + n] = (updated ? lamIn[q * bs + n] : lamOut[q * bs + n]) + lu[q];
    lP[n] += lp;
  }
}

}  // namespace MRM_NS
