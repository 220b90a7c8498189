// mrm_device.h -- device-side building blocks of the MHD step for gfx950.
//
// Grid state is structure-of-arrays in HBM: field v of cell `node` at
// U[v*bs + node], node = i + n0*(j + n1*k), dim 0 (vertical) fastest -- the
// reference's node order (grid.C:196-199) with cState's components
// (rho, M0, M1, M2, e, B0, B1, B2; mhd3d.H:24-27) split into separate arrays so
// that every wavefront load is a contiguous 512-byte segment.
//
// Arithmetic keeps the reference's operation order; the library is built with
// -ffp-contract=off so no a*b+c is fused.  smax/smin stand for the
// reference's std::max / std::min and compile to one v_max_f64 / v_min_f64
// (the compare-and-select form costs three instructions): for ordered
// operands the value is the same, only the sign of a zero result may differ
// (std::max(-0., +0.) is -0.), and a NaN operand gives the other operand.
// The states these kernels see hold no NaN (ConsToPrim's fallbacks replace
// invalid EOS values before any limiter reads them).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mrm_types.h"
#include "mrm_math.h"

#ifndef MRM_NS
#define MRM_NS exact
#endif

// cell functions compile for the host too, so that their derivatives can be
// checked against finite differences on the CPU
#ifndef MRM_HD
#define MRM_HD __host__ __device__ __forceinline__
#endif

namespace MRM_NS {




MRM_HD double smax(double a, double b) { return __builtin_fmax(a, b); }
MRM_HD double smin(double a, double b) { return __builtin_fmin(a, b); }
MRM_HD int imax_(int a, int b) { return (a < b) ? b : a; }
MRM_HD int imin_(int a, int b) { return (b < a) ? b : a; }

// (int)x as the reference's x86-64 build evaluates it: cvttsd2si returns
// INT_MIN for NaN and out-of-range values; the GPU conversion saturates.
MRM_HD int cvt_int(double x) {
  return (x >= -2147483648.0 && x < 2147483648.0) ? (int)x : (int)0x80000000;
}

// ------------------------------------------------------------------ EOS
// Bilinear lookups of eos_init.C:216-416 (scalar interpolations used by the
// boundary condition); term order follows each reference function.
MRM_HD double eos_fwd_raw(const EosDev& E, int t, double ee, double dd) {
  double ee1 = log(ee + E.eps_off);
  double logr = log(dd);
  int i = cvt_int((ee1 - E.xeps0) * E.inv_del_eps);
  if (i < 0) i = 0;
  if (i > E.N_eps - 2) i = E.N_eps - 2;
  int j = cvt_int((logr - E.xlr0) * E.inv_del_lr);
  if (j < 0) j = 0;
  if (j > E.N_lr - 2) j = E.N_lr - 2;
  const float* T = E.tab + (size_t)t * E.N_eps * E.N_lr;
  const long L = E.N_lr;
  double v = (logr - E.xlr[j]) * (ee1 - E.xeps[i]) * (double)T[(i + 1) * L + j + 1]
           + (logr - E.xlr[j]) * (E.xeps[i + 1] - ee1) * (double)T[i * L + j + 1]
           + (E.xlr[j + 1] - logr) * (ee1 - E.xeps[i]) * (double)T[(i + 1) * L + j]
           + (E.xlr[j + 1] - logr) * (E.xeps[i + 1] - ee1) * (double)T[i * L + j];
  return v;
}
MRM_HD double p_interp(const EosDev& E, double ee, double dd) {
  return exp(eos_fwd_raw(E, 0, ee, dd));
}
MRM_HD double s_interp(const EosDev& E, double ee, double dd) {
  return eos_fwd_raw(E, 2, ee, dd) - E.ss_off;
}
MRM_HD double inv_raw(const EosDev& E, int t, double pp, double ss) {
  double ss1 = ss + E.ss_off;
  double logp = log(pp);
  int i = cvt_int((logp - E.xlp30) * E.inv_del_lp3);
  if (i < 0) i = 0;
  if (i > E.N_lp3 - 2) i = E.N_lp3 - 2;
  int j = cvt_int((ss1 - E.xs30) * E.inv_del_s3);
  if (j < 0) j = 0;
  if (j > E.N_s3 - 2) j = E.N_s3 - 2;
  const float* T = E.itab + (size_t)t * E.N_lp3 * E.N_s3;
  const long L = E.N_s3;
  return (logp - E.xlp3[i]) * (ss1 - E.xs3[j]) * (double)T[(i + 1) * L + j + 1]
       + (logp - E.xlp3[i]) * (E.xs3[j + 1] - ss1) * (double)T[(i + 1) * L + j]
       + (E.xlp3[i + 1] - logp) * (ss1 - E.xs3[j]) * (double)T[i * L + j + 1]
       + (E.xlp3[i + 1] - logp) * (E.xs3[j + 1] - ss1) * (double)T[i * L + j];
}
MRM_HD double d3_interp(const EosDev& E, double pp, double ss) {
  return exp(inv_raw(E, 1, pp, ss));
}
MRM_HD double eps3_interp(const EosDev& E, double pp, double ss) {
  double ee = exp(inv_raw(E, 0, pp, ss));
  return ee - E.eps_off;
}

// ConsToPrim on one cell (cons_to_prim.C:108-204).  In: rho, M, e.
// Out: v (3), e_int, and p/T/ne/rhoi/amb in out[0..4].  With NOUT = 2 only
// p and T are produced (T decides the validity fallback of p); NOUT = 5
// produces all five.
// WANT_T with NOUT = 2: the temperature is always formed (out[1])
//
// The cell's lookup comes in two halves, so that a kernel can issue the
// memory reads of one cell (eos_issue: velocities, e_int, the node indices,
// the node coordinates and the table values) and do other work before the
// arithmetic that consumes them (eos_finish); cons_to_prim_cell is the two
// in a row.
template <int NOUT>
struct EosLook {
  double d, e, eps, lgr, ep;
  double xe0, xe1, xr0, xr1;      // xeps[i_e], xeps[i_e + 1], xlr[i_d], xlr[i_d + 1]
  float t11[NOUT], t10[NOUT], t01[NOUT], t00[NOUT];
};

// m1..m3 become the velocities and e the internal energy, as in ConsToPrim;
// FLOG: ln rho and ln eps through mrm_math::mlog (false: the library's log)
template <int NOUT, bool FLOG = true>
MRM_HD void eos_issue(const EosDev& E, double d, double& m1, double& m2, double& m3, double& e,
                      EosLook<NOUT>& g) {
  double dn = 1.0 / d;
  m1 = m1 * dn;
  m2 = m2 * dn;
  m3 = m3 * dn;
  double vv = (m1 * m1 + m2 * m2 + m3 * m3);
  e = e - 0.5 * vv * d;
  double eps = e * dn + E.eps_off;
  double lgr = FLOG ? mrm_math::mlog(d) : log(d);
  double ep = FLOG ? mrm_math::mlog(eps) : log(eps);
  double lr = lgr;
  int i_d = cvt_int((lr - E.xlr0) * E.inv_del_lr);
  i_d = imax_(0, i_d);
  i_d = imin_(E.N_lr - 2, i_d);
  int i_e = cvt_int((ep - E.xeps0) * E.inv_del_eps);
  i_e = imax_(0, i_e);
  i_e = imin_(E.N_eps - 2, i_e);
  g.d = d;
  g.e = e;
  g.eps = eps;
  g.lgr = lgr;
  g.ep = ep;
  g.xe0 = E.xeps[i_e];
  g.xe1 = E.xeps[i_e + 1];
  g.xr0 = E.xlr[i_d];
  g.xr1 = E.xlr[i_d + 1];
  const long L = E.N_lr;
  const long NT = (long)E.N_eps * E.N_lr;
  const long o11 = (i_e + 1) * L + i_d + 1, o10 = (i_e + 1) * L + i_d;
  const long o01 = i_e * L + i_d + 1, o00 = i_e * L + i_d;
  const int tabs[5] = {0, 1, 3, 4, 5};
#pragma unroll
  for (int q = 0; q < NOUT; q++) {
    const float* T = E.tab + tabs[q] * NT;
    g.t11[q] = T[o11];
    g.t10[q] = T[o10];
    g.t01[q] = T[o01];
    g.t00[q] = T[o00];
  }
}

template <int NOUT, bool WANT_T = false>
MRM_HD void eos_finish(const EosDev& E, const EosLook<NOUT>& g, double out[5]) {
  const double eos_gamma = 1.65, eos_mu = 0.62, eos_mu_n = 1.3;
  const double mh = 1.67262158e-24, kb = 1.380622e-16;
  const double Rgas = 8.314e7;
  const double c_temp = (eos_gamma - 1.) * eos_mu / Rgas;
  const double c_pres = (eos_gamma - 1.);
  const double d = g.d, e = g.e, eps = g.eps, lgr = g.lgr, ep = g.ep, lr = g.lgr;
  double coeff0 = (ep - g.xe0) * (lr - g.xr0);
  double coeff1 = (ep - g.xe0) * (g.xr1 - lr);
  double coeff2 = (g.xe1 - ep) * (lr - g.xr0);
  double coeff3 = (g.xe1 - ep) * (g.xr1 - lr);
  double r[5];
  double tt = 0.0;
#pragma unroll
  for (int q = 0; q < NOUT; q++) {
    double v = coeff0 * (double)g.t11[q];
    v += coeff1 * (double)g.t10[q];
    v += coeff2 * (double)g.t01[q];
    v += coeff3 * (double)g.t00[q];
    if (NOUT == 2 && q == 1) {
      tt = v;
      r[q] = 0.0;
    } else {
      r[q] = exp(v);
    }
  }
  double pres = r[0];
  double temp = r[1];
  if (NOUT == 2) {
    // The temperature only decides whether p is valid (cons_to_prim.C:195).
    // exp(tt) is finite and positive for |tt| < 700, so it is evaluated only
    // when a fallback branch uses it or tt lies outside that range.
    const bool plain = !(lgr < E.xlr0) && !(eps > 0.8 * E.eps_max);
    if (!WANT_T && plain && tt > -700.0 && tt < 700.0) {
      temp = 1.0;
    } else {
      temp = exp(tt);
    }
  }
  double nel = NOUT > 2 ? r[2] : 0.0;
  if (lgr < E.xlr0) {
    pres = c_pres * e;
    temp = c_temp * eps;
    nel = pres / (temp * kb) - d / (eos_mu_n * mh);
  }
  if (eps > 0.8 * E.eps_max) {
    double xx = smax(0.0, 5.0 * (E.eps_max - eps) / E.eps_max);
    pres = pres * xx + (1.0 - xx) * c_pres * e;
    temp = temp * xx + (1.0 - xx) * c_temp * eps;
    nel = pres / (temp * kb) - d / (eos_mu_n * mh);
  }
  int bad = 0;
  if ((pres != pres) || (pres <= 0.) || (pres == __builtin_inf())) bad = 1;
  if ((temp != temp) || (temp <= 0.) || (temp == __builtin_inf())) bad = 1;
  if (bad) {
    pres = c_pres * e;
    temp = c_temp * eps;
  }
  if (nel == __builtin_inf()) nel = pres / (temp * kb) - d / (eos_mu_n * mh);
  if ((nel < 1.0) || (nel != nel) || (nel == __builtin_inf())) nel = 1.0;
  out[0] = pres;
  out[1] = temp;
  if (NOUT > 2) {
    out[2] = nel;
    out[3] = r[3];
    out[4] = r[4];
  }
}

template <int NOUT, bool WANT_T = false, bool FLOG = true>
MRM_HD void cons_to_prim_cell(const EosDev& E, double d, double& m1, double& m2, double& m3, double& e,
                              double out[5]) {
  EosLook<NOUT> g;
  eos_issue<NOUT, FLOG>(E, d, m1, m2, m3, e, g);
  eos_finish<NOUT, WANT_T>(E, g, out);
}

// ------------------------------------------------------------------ residual
// Directional flux of one point (mhdres_SR.C:296-360), eta = Spitzer =
// ambipolar = 0.  D is the sweep direction d1; the del[d1][*] factors of the
// reference are kept as explicit 0/1 multipliers so NaN/inf propagate alike.
struct PtFlux {
  double f[MRM_NV];
  double ms[3];
};

template <int D>
MRM_HD void point_flux(const double var[MRM_NV], double pres, PtFlux& o) {
  const double p1 = D == 0 ? 1.0 : 0.0, p2 = D == 1 ? 1.0 : 0.0, p3 = D == 2 ? 1.0 : 0.0;
  double var0 = var[0], var1 = var[1], var2 = var[2], var3 = var[3];
  double var4 = var[4], var5 = var[5], var6 = var[6], var7 = var[7];
  double vel = var[1 + D];
  double mag = var[5 + D];
  double vv = var1 * var1 + var2 * var2 + var3 * var3;
  double bb = var5 * var5 + var6 * var6 + var7 * var7;
  double ekin = 0.5 * var0 * vv;
  double pmag = 0.5 * bb;
  double mflx = vel * var0;
  o.f[0] = mflx;
  o.f[1] = mflx * var1 + p1 * pres;
  o.f[2] = mflx * var2 + p2 * pres;
  o.f[3] = mflx * var3 + p3 * pres;
  o.f[4] = vel * (var4 + ekin + pres);
  o.f[5] = vel * var5 - var1 * mag;
  o.f[6] = vel * var6 - var2 * mag;
  o.f[7] = vel * var7 - var3 * mag;
  o.ms[0] = mag * var5 - p1 * pmag;
  o.ms[1] = mag * var6 - p2 * pmag;
  o.ms[2] = mag * var7 - p3 * pmag;
}

// lf = (1+x^2)/(1+x^2+x^4), x^2 = va^2 / va_max^2 (mhdres_SR.C:319-323)
MRM_HD double lfac_of(const double var[MRM_NV], double inv_va2max) {
  double dn = 1.0 / var[0];
  double bb = var[5] * var[5] + var[6] * var[6] + var[7] * var[7];
  double va2 = bb * dn;
  double x2 = va2 * inv_va2max;
  double x4 = x2 * x2;
  double s = 1.0 + x2;
  return s / (s + x4);
}

// fastest characteristic speed of one point (mhdres_SR.C:310-328)
MRM_HD double point_cmax(const double var[MRM_NV], double pres, double inv_va2max) {
  double var0 = var[0];
  double dn = 1.0 / var0;
  double vv = var[1] * var[1] + var[2] * var[2] + var[3] * var[3];
  double bb = var[5] * var[5] + var[6] * var[6] + var[7] * var[7];
  double va2 = bb * dn;
  double x2 = va2 * inv_va2max;
  double x4 = x2 * x2;
  double s = 1.0 + x2;
  double lf = s / (s + x4);
  double cs2 = 5.0 / 3.0 * pres * dn;
  vv = sqrt(vv);
  return vv + 0.0 + sqrt(smax(cs2, lf * (cs2 + va2)));
}

// cmax of a conservative point, primitives formed as ConsToPrim forms them
MRM_HD double point_cmax_cons(const double u[MRM_NV], double pres, double inv_va2max) {
  double w[MRM_NV];
  double dn = 1.0 / u[0];
  w[0] = u[0];
  w[1] = u[1] * dn;
  w[2] = u[2] * dn;
  w[3] = u[3] * dn;
  w[4] = 0.0;
  w[5] = u[5];
  w[6] = u[6];
  w[7] = u[7];
  return point_cmax(w, pres, inv_va2max);
}

// Residual contribution of sweep D at the centre point given the point data of
// the five points c-2..c+2 along D (mhdres_SR.C:361-412).  res must hold 0.
// DIAG: dg[0] += the energy flux divergence (tvar1, :368-383) and dg[1] +=
// v . Lorentz force (tvar3, :434-435)
template <int D, bool DIAG = false>
MRM_HD void sweep_residual(const double v[5][MRM_NV], const double p[5], double w1,
                                               double w2, double inv_va2max, double res[MRM_NV],
                                               double* dg = nullptr) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  PtFlux fm2, fm1, fp1, fp2;
  point_flux<D>(v[0], p[0], fm2);
  point_flux<D>(v[1], p[1], fm1);
  point_flux<D>(v[3], p[3], fp1);
  point_flux<D>(v[4], p[4], fp2);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    res[q] = 0.0;
    res[q] -= w1 * (fp1.f[q] - fm1.f[q]) + w2 * (fp2.f[q] - fm2.f[q]);
  }
  if (DIAG) dg[0] += res[4];
  double dB2 = w1 * (v[3][5 + D2] - v[1][5 + D2]) + w2 * (v[4][5 + D2] - v[0][5 + D2]);
  double dB3 = w1 * (v[3][5 + D3] - v[1][5 + D3]) + w2 * (v[4][5 + D3] - v[0][5 + D3]);
  double cb[3];
  cb[D] = -v[2][5 + D2] * dB2 - v[2][5 + D3] * dB3;
  cb[D2] = v[2][5 + D] * dB2;
  cb[D3] = v[2][5 + D] * dB3;
  double lrt0 = w1 * (fp1.ms[0] - fm1.ms[0]) + w2 * (fp2.ms[0] - fm2.ms[0]);
  double lrt1 = w1 * (fp1.ms[1] - fm1.ms[1]) + w2 * (fp2.ms[1] - fm2.ms[1]);
  double lrt2 = w1 * (fp1.ms[2] - fm1.ms[2]) + w2 * (fp2.ms[2] - fm2.ms[2]);
  double lf = lfac_of(v[2], inv_va2max);
  lrt0 = lf * lrt0 + (1.0 - lf) * cb[0];
  lrt1 = lf * lrt1 + (1.0 - lf) * cb[1];
  lrt2 = lf * lrt2 + (1.0 - lf) * cb[2];
  res[1] += lrt0;
  res[2] += lrt1;
  res[3] += lrt2;
  const double vl = v[2][1] * lrt0 + v[2][2] * lrt1 + v[2][3] * lrt2;
  res[4] += vl;
  if (DIAG) dg[1] += vl;
}

// Resistive part of one sweep along D (mhdres_SR.C:422-462, 742-766): the
// sweep's share of J = curl B, with the sweep signs -1, +1, -1 of
// mhdres_SR.C:94 (dB2, dB3: the 4th-order derivatives along D of B_D2, B_D3),
// and eta times the 4th-order second derivative of every B component along
// D; s1[q] = B_q(+1) + B_q(-1), s2[q] = B_q(+2) + B_q(-2), c = B at the point.
template <int D>
MRM_HD void eta_sweep(double dB2, double dB3, const double c[3], const double s1[3], const double s2[3],
                      double eta, double ww0, double ww1, double ww2, double J[3], double res[MRM_NV]) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  constexpr double sgn = D == 1 ? 1.0 : -1.0;
  J[D2] += sgn * dB3;
  J[D3] += -sgn * dB2;
#pragma unroll
  for (int q = 0; q < 3; q++) res[5 + q] += eta * (ww0 * c[q] + ww1 * s1[q] + ww2 * s2[q]);
}

template <int D, bool ETA>
MRM_HD void sweep_tail(const double e1[5], const double e2[5], const double bs1[3], const double bs2[3],
                       const double c[9], double w1, double w2, double lf, double res[MRM_NV],
                       const ResidParams* rp, double* J);

// Everything a residual sweep along D does after its flux differences
// (mhdres_SR.C:361-412): d1 = F(+1) - F(-1), d2 = F(+2) - F(-2) of the 8
// fluxes, 3 stresses, B_D2 and B_D3; bs1, bs2 (ETA) the sums of B at +-1,
// +-2; c the centre point's primitives
template <int D, bool ETA = false>
MRM_HD void sweep_finish(const double d1[13], const double d2[13], const double bs1[3], const double bs2[3],
                         const double c[9], double w1, double w2, double lf, double res[MRM_NV],
                         const ResidParams* rp, double* J) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    res[q] = 0.0;
    res[q] -= w1 * d1[q] + w2 * d2[q];
  }
  sweep_tail<D, ETA>(d1 + 8, d2 + 8, bs1, bs2, c, w1, w2, lf, res, rp, J);
}

// the part of sweep_finish after the flux terms: e1 = d1 + 8, e2 = d2 + 8
// (the stresses' and B_D2, B_D3's differences)
template <int D, bool ETA>
MRM_HD void sweep_tail(const double e1[5], const double e2[5], const double bs1[3], const double bs2[3],
                       const double c[9], double w1, double w2, double lf, double res[MRM_NV],
                       const ResidParams* rp, double* J) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  const double* d1 = e1 - 8;
  const double* d2 = e2 - 8;
  double dB2 = w1 * d1[11] + w2 * d2[11];
  double dB3 = w1 * d1[12] + w2 * d2[12];
  double cb[3];
  cb[D] = -c[5 + D2] * dB2 - c[5 + D3] * dB3;
  cb[D2] = c[5 + D] * dB2;
  cb[D3] = c[5 + D] * dB3;
  double lrt0 = w1 * d1[8] + w2 * d2[8];
  double lrt1 = w1 * d1[9] + w2 * d2[9];
  double lrt2 = w1 * d1[10] + w2 * d2[10];
  lrt0 = lf * lrt0 + (1.0 - lf) * cb[0];
  lrt1 = lf * lrt1 + (1.0 - lf) * cb[1];
  lrt2 = lf * lrt2 + (1.0 - lf) * cb[2];
  res[1] += lrt0;
  res[2] += lrt1;
  res[3] += lrt2;
  res[4] += c[1] * lrt0 + c[2] * lrt1 + c[3] * lrt2;
  if constexpr (ETA)
    eta_sweep<D>(dB2, dB3, c + 5, bs1, bs2, rp->eta, rp->ww0[D], rp->ww1[D], rp->ww2[D], J, res);
}

// The same residual as sweep_residual with a lower register peak: the four
// neighbour points are fetched through get(s, w[9]) (s = 0,1,3,4 for offsets
// -2,-1,+1,+2) one pair at a time, and only the differences
// F(+1)-F(-1) and F(+2)-F(-2) stay live.  Every arithmetic operation is the
// one sweep_residual (and mhdres_SR.C:361-412) performs.
// With ETA the resistive part of the sweep is added too (eta_sweep; J
// accumulates the sweep's share of curl B).
template <int D, bool ETA = false, class GET>
MRM_HD void sweep_residual_lr(GET get, const double c[9], double w1, double w2,
                                                  double lf, double res[MRM_NV], const ResidParams* rp = nullptr,
                                                  double* J = nullptr) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  double d1[13], d2[13];   // 8 fluxes, 3 stresses, B_D2, B_D3
  double bs1[3], bs2[3];   // B(+1) + B(-1), B(+2) + B(-2) (ETA only)
  {
    double pa[9], pb[9];
    PtFlux fa, fb;
    get(1, pa);
    point_flux<D>(pa, pa[8], fa);
    get(3, pb);
    point_flux<D>(pb, pb[8], fb);
#pragma unroll
    for (int q = 0; q < 8; q++) d1[q] = fb.f[q] - fa.f[q];
#pragma unroll
    for (int q = 0; q < 3; q++) d1[8 + q] = fb.ms[q] - fa.ms[q];
    d1[11] = pb[5 + D2] - pa[5 + D2];
    d1[12] = pb[5 + D3] - pa[5 + D3];
    if constexpr (ETA) {
#pragma unroll
      for (int q = 0; q < 3; q++) bs1[q] = pb[5 + q] + pa[5 + q];
    }
  }
  {
    double pa[9], pb[9];
    PtFlux fa, fb;
    get(0, pa);
    point_flux<D>(pa, pa[8], fa);
    get(4, pb);
    point_flux<D>(pb, pb[8], fb);
#pragma unroll
    for (int q = 0; q < 8; q++) d2[q] = fb.f[q] - fa.f[q];
#pragma unroll
    for (int q = 0; q < 3; q++) d2[8 + q] = fb.ms[q] - fa.ms[q];
    d2[11] = pb[5 + D2] - pa[5 + D2];
    d2[12] = pb[5 + D3] - pa[5 + D3];
    if constexpr (ETA) {
#pragma unroll
      for (int q = 0; q < 3; q++) bs2[q] = pb[5 + q] + pa[5 + q];
    }
  }
  sweep_finish<D, ETA>(d1, d2, bs1, bs2, c, w1, w2, lf, res, rp, J);
}

// A flux point as the residual sweeps along D use it: the 8 fluxes, the 3
// magnetic stresses, B_D2 and B_D3, and (ETA) B_D -- NFP_(ETA) values
template <bool ETA> constexpr int nfp_() { return ETA ? 14 : 13; }
template <int D, bool ETA>
MRM_HD void flux_pack(const double p[9], double o[14]) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  PtFlux f;
  point_flux<D>(p, p[8], f);
#pragma unroll
  for (int q = 0; q < 8; q++) o[q] = f.f[q];
#pragma unroll
  for (int q = 0; q < 3; q++) o[8 + q] = f.ms[q];
  o[11] = p[5 + D2];
  o[12] = p[5 + D3];
  if (ETA) o[13] = p[5 + D];
}

// sweep_residual_lr from packed flux points, one value at a time: getq(s, f)
// gives value f of the packed point at offset s - 2 (s = 0, 1, 3, 4), so
// that only the four values of one quantity are live at once; each flux
// term's residual is formed as sweep_finish forms it, the stresses' and B's
// differences go to sweep_tail
template <int D, bool ETA = false, class GETQ>
MRM_HD void sweep_residual_packed(GETQ getq, const double c[9], double w1, double w2, double lf,
                                  double res[MRM_NV], const ResidParams* rp = nullptr, double* J = nullptr) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    const double d1 = getq(3, q) - getq(1, q);
    const double d2 = getq(4, q) - getq(0, q);
    res[q] = 0.0;
    res[q] -= w1 * d1 + w2 * d2;
  }
  double e1[5], e2[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    e1[q] = getq(3, 8 + q) - getq(1, 8 + q);
    e2[q] = getq(4, 8 + q) - getq(0, 8 + q);
  }
  double bs1[3], bs2[3];
  if constexpr (ETA) {
    bs1[D2] = getq(3, 11) + getq(1, 11);
    bs1[D3] = getq(3, 12) + getq(1, 12);
    bs1[D] = getq(3, 13) + getq(1, 13);
    bs2[D2] = getq(4, 11) + getq(0, 11);
    bs2[D3] = getq(4, 12) + getq(0, 12);
    bs2[D] = getq(4, 13) + getq(0, 13);
  }
  sweep_tail<D, ETA>(e1, e2, bs1, bs2, c, w1, w2, lf, res, rp, J);
}

// ------------------------------------------------------------------ sources

// Tcheck clamp (Add_Sources_Integrate_SR.C:447-465)
MRM_HD void tcheck_vals(double u[MRM_NV], double rho_min, double eps_min, double eps_max,
                                            double vmax) {
  double dn = fmax(rho_min, u[0]);
  double vv = sqrt(u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / dn;
  double s = vmax / fmax(vmax, vv);
  double ekin = 0.5 * dn * vv * vv;
  double ee = u[4];
  double ee_low = eps_min * dn + ekin * s * s;
  double ee_up = eps_max * dn + ekin * s * s;
  u[4] = fmin(fmax(ee, ee_low), ee_up);
  u[0] = dn;
  u[1] = u[1] * s;
  u[2] = u[2] * s;
  u[3] = u[3] * s;
}

// Sources, Boris correction, RK update and Tcheck of one cell
// (Add_Sources_Integrate_SR.C:331-465).  up: primitive U at the cell; R: its
// residual (modified); u0: U0; dmpi: the damping profile at the cell's
// layer (0 with the damping off).  Returns the new conservative state in out.
// DIAG: dg[0..3] = efaci, e_dmpi, boris3 and the energy before the clamp
// (Add_Sources_Integrate_SR.C:168-170, 218, 292)
template <bool DIAG = false>
MRM_HD void source_integrate_cell(const SrcParams& P, const double up[MRM_NV],
                                                      double R[MRM_NV], const double u0[MRM_NV], double Qrad,
                                                      double dmpi, double out[MRM_NV], double* dg = nullptr) {
  double dn = P.invRdt * up[4];
  double efaci = dn / fmax(dn, fabs(Qrad));
  dn = P.g * up[0] - dmpi;
  double e_dmpi = dn * up[1];
  R[1] += dn;
  R[4] += e_dmpi + Qrad * efaci;
  double rh = up[0], rd = R[0];
  double vx = up[1], vy = up[2], vz = up[3];
  double bx = up[5], by = up[6], bz = up[7];
  double fsr_x = R[1], fsr_y = R[2], fsr_z = R[3];
  double bb = bx * bx + by * by + bz * bz;
  double va2 = bb / rh;
  double x2 = va2 * P.inv_va2max;
  double x4 = x2 * x2;
  double bf = x4 / (1 + x2 + x4);
  fsr_x -= vx * rd;
  fsr_y -= vy * rd;
  fsr_z -= vz * rd;
  dn = (fsr_x * bx + fsr_y * by + fsr_z * bz) / fmax(1e-100, bb);
  double boris0 = -bf * (fsr_x - dn * bx);
  double boris1 = -bf * (fsr_y - dn * by);
  double boris2 = -bf * (fsr_z - dn * bz);
  double boris3 = boris0 * vx + boris1 * vy + boris2 * vz;
  R[1] += boris0;
  R[2] += boris1;
  R[3] += boris2;
  R[4] += boris3;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) out[q] = u0[q] + P.wdt * R[q];
  if (DIAG) {
    dg[0] = efaci;
    dg[1] = e_dmpi;
    dg[2] = boris3;
    dg[3] = out[4];
  }
  tcheck_vals(out, P.rho_min, P.eps_min, P.eps_max, P.vmax);
}

// ------------------------------------------------------------------ TVD
// TVDlimit (tvdlimit_SR.C:33-695), diagnostics and ambipolar off.  One
// thread streams one grid line with a four-point window: at step t it
// evaluates point t, the slope at t-1, the interface flux at t-2 and updates
// point t-2 in place -- the same per-point, per-slope and per-interface
// arithmetic as the reference's phase loops over whole lines.

struct TvdPt {
  double v[MRM_NV];   // primitives; ln rho and ln eps when rho_log
  double lnr;         // ln rho (density-jump detector)
  double vsqr, bsqr, cm, hyp_e, hyp_v, hyp_B, BC2, CZ, hd;
};

// Limiter branch accounting (counting instantiations only; the sweep
// functions take a null TvdCnt* otherwise, and inlining removes the code).
// n: counts per choice (MRM_BR_* of muram_step.h); h: per branch point the
// wrapping sum of br_weight(position) x the choice's number; key: position of
// the point or interface being evaluated, (sweep * bufsize + node) * 8, to
// which the field index is added.
MRM_HD unsigned long long br_weight(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return (x ^ (x >> 31)) | 1ull;
}
struct TvdCnt {
  unsigned n[MRM_BR_NCOUNT];
  unsigned nnear[MRM_BK_N];
  unsigned nexact[MRM_BK_N];
  unsigned long long h[MRM_BK_N];
  unsigned long long key;
  double tie;         // near-tie margin relative to the scale
  bool on;            // the line being swept is counted
  MRM_HD void clear(double tie_) {
    for (int i = 0; i < MRM_BR_NCOUNT; i++) n[i] = 0;
    for (int i = 0; i < MRM_BK_N; i++) { h[i] = 0; nnear[i] = 0; nexact[i] = 0; }
    key = 0;
    tie = tie_;
    on = false;
  }
  // 0: decisive, 1: near tie (0 < margin < tie x scale), 2: exact tie
  // (margin 0); with tie = 0 every evaluation is decisive
  MRM_HD int cls(double margin, double scale) const {
    return !(tie > 0.0) ? 0 : margin == 0.0 ? 2 : (margin < tie * fabs(scale) ? 1 : 0);
  }
  // ties are counted as such and nothing else
  MRM_HD bool tied(int kind, int c) {
    nnear[kind] += (on && c == 1) ? 1u : 0u;
    nexact[kind] += (on && c == 2) ? 1u : 0u;
    return c != 0;
  }
  // a switch: counted and weighted when taken
  MRM_HD void hit(int c, int kind, int q, bool taken, int tc) {
    if (tied(kind, tc)) return;
    taken = taken && on;
    n[c] += taken ? 1u : 0u;
    h[kind] += taken ? br_weight(key + q) : 0ull;
  }
  // one of ncat choices (cat = 0 .. ncat-1, slots c0 ..)
  MRM_HD void cat(int c0, int ncat, int kind, int q, int cat, int tc) {
    if (tied(kind, tc)) return;
    for (int i = 0; i < ncat; i++) n[c0 + i] += (on && cat == i) ? 1u : 0u;
    h[kind] += on ? br_weight(key + q) * (unsigned long long)(cat + 1) : 0ull;
  }
};

MRM_HD double fmax3_(double a, double b, double c) { return fmax(fmax(a, b), c); }

// per-point quantities (tvdlimit_SR.C:269-385)
// FLOG: the logarithms through mrm_math::mlog (false: the library's log,
// for kernels whose registers are at their limit)
template <bool FLOG = true>
MRM_HD void tvd_point(const TvdParams& P, const double u[MRM_NV], double pres, double tau,
                                          double hfb1, double hft1, bool xsweep, TvdPt& o,
                                          TvdCnt* cn = nullptr) {
  double v0 = smax(P.rho_min, u[0]);
  double v1 = u[1], v2 = u[2], v3 = u[3], v4 = u[4];
  o.v[5] = u[5]; o.v[6] = u[6]; o.v[7] = u[7];
  o.CZ = (double)(tau > 1.0e-5);
  double dn = 1.0 / v0;
  v1 = v1 * dn;
  v2 = v2 * dn;
  v3 = v3 * dn;
  double vv = (v1 * v1 + v2 * v2 + v3 * v3);
  double bb = (o.v[5] * o.v[5] + o.v[6] * o.v[6] + o.v[7] * o.v[7]);
  v4 = v4 * dn - 0.5 * vv;
  o.vsqr = vv;
  o.bsqr = bb;
  vv = sqrt(vv);
  double va2 = bb * dn;
  double x2 = va2 * P.inv_va2max;
  double x4 = x2 * x2;
  double s = 1.0 + x2;
  double lf = s / (s + x4);
  double cs2 = 5.0 / 3.0 * pres * dn;
  double cfast = sqrt(va2 + cs2);
  double CME_mode = (1.0 - o.CZ) * ((double)(vv >= P.CME_thresh * cfast));
  if (cn) {
    cn->hit(MRM_BR_CZ, MRM_BK_CZ, 0, o.CZ != 0.0, cn->cls(fabs(tau - 1.0e-5), 1.0e-5));
    const double ct = P.CME_thresh * cfast;
    // with CZ = 1 the switch is off whatever vv is
    cn->hit(MRM_BR_CME, MRM_BK_CME, 0, CME_mode != 0.0, o.CZ != 0.0 ? 0 : cn->cls(fabs(vv - ct), fmax(vv, ct)));
  }
  dn = v0 * v0;
  double rf = smin(lf, dn / (dn + P.rho_lev));
  double hh = 1.0 - hfb1 - hft1;
  double cf = hfb1 * P.tvd_cs[0] + (rf * P.tvd_cs[1] + (1.0 - rf) * P.tvd_cs[2]) * hh + hft1 * P.tvd_cs[3];
  cf = smax(cf, CME_mode);
  cs2 *= cf * cf;
  double vv_amb = 0.0;
  o.cm = vv_amb + vv + sqrt(smax(cs2, lf * (cs2 + va2)));
  o.hyp_e = hfb1 * P.tvd_h[0] + (rf * P.tvd_h[1] + (1.0 - rf) * P.tvd_h[2]) * hh + hft1 * P.tvd_h[3];
  o.hyp_v = hfb1 * P.tvd_h[0] + (rf * P.tvd_h[1] + (1.0 - rf) * P.tvd_h[2] * P.tvd_pm_v) * hh +
            hft1 * P.tvd_h[3] * P.tvd_pm_v;
  o.hyp_B = hfb1 * P.tvd_h[0] + (rf * P.tvd_h[1] + (1.0 - rf) * P.tvd_h[2] * P.tvd_pm_B) * hh +
            hft1 * P.tvd_h[3] * P.tvd_pm_B;
  o.hyp_e *= (1.0 - CME_mode);
  o.hyp_v *= (1.0 - CME_mode);
  o.BC2 = 1.0 - lf;
  o.hd = 0.0;
  if (xsweep && (P.vhyp > 0.0)) {
    double h = P.vhyp * sqrt(o.vsqr);
    o.hd = 0.5 * P.idxd1 * smax(0.0, smin(h, P.cmax - o.cm));
  }
  if (P.rho_log) {
    v0 = FLOG ? mrm_math::mlog(v0) : log(v0);
    v4 = FLOG ? mrm_math::mlog(v4) : log(v4);
    o.lnr = v0;
  } else {
    o.lnr = FLOG ? mrm_math::mlog(v0) : log(v0);
  }
  o.v[0] = v0; o.v[1] = v1; o.v[2] = v2; o.v[3] = v3; o.v[4] = v4;
}

// limited slope at point i (tvdlimit_SR.C:388-434)
MRM_HD void tvd_slope(const TvdParams& P, const TvdPt& a, const TvdPt& b, const TvdPt& c,
                                          double hfb1, double hft1, double S[MRM_NV], TvdCnt* cn = nullptr) {
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    double sl = b.v[q] - a.v[q];
    double sr = c.v[q] - b.v[q];
    double slm = 0.25 * (sl + sr);
    double lower = smin(smin(sl, sr), slm);
    double upper = smax(smax(sl, sr), slm);
    S[q] = smax(0.0, lower) + smin(0.0, upper);
    if (cn) {
      // which of sl, sr, slm the min (positive) or max (negative) took;
      // margin: the sign's distance from zero and the gap to the runner-up
      int k;
      double mg;
      if (lower > 0.0) {
        const double m1 = (sr < sl) ? sr : sl;
        k = (slm < m1) ? 2 : (sr < sl) ? 1 : 0;
        const double x0 = k == 0 ? sl : k == 1 ? sr : slm;
        const double gap = fmin(k == 0 ? 1e300 : sl - x0, fmin(k == 1 ? 1e300 : sr - x0, k == 2 ? 1e300 : slm - x0));
        mg = fmin(lower, gap);
      } else if (upper < 0.0) {
        const double m1 = (sl < sr) ? sr : sl;
        k = (m1 < slm) ? 5 : (sl < sr) ? 4 : 3;
        const double x0 = k == 3 ? sl : k == 4 ? sr : slm;
        const double gap = fmin(k == 3 ? 1e300 : x0 - sl, fmin(k == 4 ? 1e300 : x0 - sr, k == 5 ? 1e300 : x0 - slm));
        mg = fmin(-upper, gap);
      } else {
        k = 6;
        mg = fmin(fabs(lower), fabs(upper));
      }
      cn->cat(MRM_BR_SLOPE_POS_L, 7, MRM_BK_SLOPE, q, k,
              cn->cls(mg, fmax3_(fabs(a.v[q]), fabs(b.v[q]), fabs(c.v[q]))));
    }
  }
  if ((P.visc_slope_bot > 0.0) || (P.visc_slope_top > 0.0)) {
    double cf = (1.0 - P.visc_slope_bot * hfb1) * (1.0 - P.visc_slope_top * hft1);
    S[1] *= cf; S[2] *= cf; S[3] *= cf;
  }
  if ((P.eta_slope_bot > 0.0) || (P.eta_slope_top > 0.0)) {
    double cf = (1.0 - P.eta_slope_bot * hfb1) * (1.0 - P.eta_slope_top * hft1);
    S[5] *= cf; S[6] *= cf; S[7] *= cf;
  }
  double qrho = fabs(b.lnr - c.lnr);
  if (qrho > P.q_rho_max) S[0] = 0.0;
  if (cn) cn->hit(MRM_BR_QRHO, MRM_BK_QRHO, 0, qrho > P.q_rho_max,
                  cn->cls(fabs(qrho - P.q_rho_max), fmax(fabs(b.lnr), fabs(c.lnr))));
  if (P.vmax_lim > 0.0) {
    if (b.vsqr > P.vsqr_diff) { S[1] = 0.0; S[2] = 0.0; S[3] = 0.0; }
    if (cn)
      cn->hit(MRM_BR_VFAST, MRM_BK_VFAST, 0, b.vsqr > P.vsqr_diff, cn->cls(fabs(b.vsqr - P.vsqr_diff), P.vsqr_diff));
  }
}

// diffusive flux at interface i (between points i = b and i+1 = c), with
// a = i-1 and d = i+2 (tvdlimit_SR.C:436-563); returns the resistive-heating
// interface term (:609-614); DIAG: *qv = the viscous-heating interface term
// (:627-634)
template <int D, bool DIAG = false>
MRM_HD double tvd_iface(const TvdParams& P, const TvdPt& a, const TvdPt& b, const TvdPt& c,
                                            const TvdPt& d, const double Sb[MRM_NV], const double Sc[MRM_NV],
                                            double hfb1, double hft1, double qft1, bool hyper_on,
                                            bool zero_bflux, double F[MRM_NV], double* qv = nullptr,
                                            TvdCnt* cn = nullptr) {
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    double sl = c.v[q] - b.v[q];
    double sl_lim = c.v[q] - b.v[q] - (Sb[q] + Sc[q]);
    double dn = (q >= 1 && q <= 3) ? b.hyp_v : (q >= 5 ? b.hyp_B : b.hyp_e);
    double rf = fabs(sl_lim) / smax(1e-100, fabs(sl));
    double cf = smax(0.0, smin(sl_lim, sl)) + smin(0.0, smax(sl_lim, sl));
    double cl = 1.0 + dn * (rf - 1.0);
    F[q] = smax(0.0, cl) * cf;
    if (cn) {
      // the minmod's choice: both positive (the smaller of sl_lim, sl), both
      // negative (the larger), or zero
      const double mn = (sl < sl_lim) ? sl : sl_lim, mx = (sl_lim < sl) ? sl : sl_lim;
      const int k = (mn > 0.0) ? ((sl < sl_lim) ? 1 : 0) : (mx < 0.0) ? ((sl_lim < sl) ? 3 : 2) : 4;
      // margin: the distance of the nearer of sl_lim, sl from zero (sl_lim
      // lies between 0 and sl, so which of the two the minmod takes is fixed
      // by the slopes' own choices)
      const double sc = fmax(fabs(b.v[q]), fabs(c.v[q]));
      cn->cat(MRM_BR_MM_POS_LIM, 5, MRM_BK_MM, q, k, cn->cls(fmin(fabs(mn), fabs(mx)), sc));
      // the clip's ratio rf = |sl_lim| / |sl| is noise where sl is (near)
      // zero: the clip is then tied with sl (the flux is ~0 either way)
      const int tsl = cn->cls(fabs(sl), sc);
      cn->hit(MRM_BR_CLIP, MRM_BK_CLIP, q, cl < 0.0, tsl ? tsl : cn->cls(fabs(cl), fmax(1.0, fabs(dn * (rf - 1.0)))));
    }
  }
  double tvd_faci = 0.5 * P.idxd1 * smin(P.cmax, smax(b.cm, c.cm));
  if (cn) {
    const double m = (b.cm < c.cm) ? c.cm : b.cm;
    const int k = (P.cmax < m) ? 0 : (b.cm < c.cm) ? 2 : 1;
    const double mg = k == 0 ? m - P.cmax : fmin(P.cmax - m, fabs(b.cm - c.cm));
    cn->cat(MRM_BR_CMAX_LIMIT, 3, MRM_BK_CMAX, 0, k, cn->cls(mg, fmax3_(P.cmax, b.cm, c.cm)));
  }
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) F[q] *= tvd_faci;
  F[5 + D] *= P.B_par_diff;
  if (P.need_tvd_coeff) {
    F[0] *= P.tc[0];
    double cf = smax(P.tc[1], P.visc_coeff_bot * hfb1 + P.visc_coeff_top * hft1);
    F[1] *= cf; F[2] *= cf; F[3] *= cf;
    F[4] *= P.tc[2];
    cf = smax(P.tc[3], P.eta_coeff_bot * hfb1 + P.eta_coeff_top * hft1);
    F[5] *= cf; F[6] *= cf; F[7] *= cf;
  }
  if (D == 0 && hyper_on) {
    double dn = smax(b.hd, c.hd);
    F[0] -= dn * (0.125 * (d.v[0] - a.v[0]) - 0.375 * (c.v[0] - b.v[0]));
    F[1] -= dn * (0.125 * (d.v[1] - a.v[1]) - 0.375 * (c.v[1] - b.v[1]));
    F[4] -= dn * (0.125 * (d.v[4] - a.v[4]) - 0.375 * (c.v[4] - b.v[4]));
  }
  double flx0 = F[0], flx1 = F[1], flx2 = F[2], flx3 = F[3], flx4 = F[4];
  double uif0 = 0.5 * (b.v[0] + Sb[0] + c.v[0] - Sc[0]);
  double uif1 = 0.5 * (b.v[1] + Sb[1] + c.v[1] - Sc[1]);
  double uif2 = 0.5 * (b.v[2] + Sb[2] + c.v[2] - Sc[2]);
  double uif3 = 0.5 * (b.v[3] + Sb[3] + c.v[3] - Sc[3]);
  double uif4 = 0.5 * (b.v[4] + Sb[4] + c.v[4] - Sc[4]);
  if (P.rho_log) {
    uif0 = exp(uif0);
    flx0 *= uif0;
    uif4 = exp(uif4);
    flx4 *= uif4;
  }
  double min_cz = flx0 * smin(b.CZ, c.CZ);
  flx1 *= uif0; flx1 += uif1 * min_cz;
  flx2 *= uif0; flx2 += uif2 * min_cz;
  flx3 *= uif0; flx3 += uif3 * min_cz;
  flx4 *= uif0;
  flx4 += (uif4 * min_cz) + qft1 * (uif1 * flx1 + uif2 * flx2 + uif3 * flx3);
  F[0] = flx0; F[1] = flx1; F[2] = flx2; F[3] = flx3; F[4] = flx4;
  if (zero_bflux) F[5] = 0.0;
  if (DIAG) *qv = (c.v[1] - b.v[1]) * F[1] + (c.v[2] - b.v[2]) * F[2] + (c.v[3] - b.v[3]) * F[3];
  return (c.v[5] - b.v[5]) * F[5] + (c.v[6] - b.v[6]) * F[6] + (c.v[7] - b.v[7]) * F[7];
}

// in-place update of point i from the fluxes at i-1 and i (tvdlimit_SR.C:565-657);
// DIAG: dg[0..3] = the energy diffusion (tvar8), boris3 (tvar6), qres and the
// energy before its floor (tvar7), :577-666
template <bool DIAG = false>
MRM_HD void tvd_update(const TvdParams& P, const TvdPt& b, const double Fm[MRM_NV],
                                           const double F[MRM_NV], double qm, double q, double qft1,
                                           double u[MRM_NV], double* dg = nullptr, TvdCnt* cn = nullptr) {
  double r[MRM_NV];
#pragma unroll
  for (int k = 0; k < MRM_NV; k++) r[k] = P.dt_tvd * (F[k] - Fm[k]);
  double v1 = b.v[1], v2 = b.v[2], v3 = b.v[3], b5 = b.v[5], b6 = b.v[6], b7 = b.v[7];
  r[4] += (1.0 - qft1) * (r[1] * v1 + r[2] * v2 + r[3] * v3);
  if (DIAG) dg[0] = r[4];
  double dn = (r[1] * b5 + r[2] * b6 + r[3] * b7) / smax(1e-100, b.bsqr);
  double boris0 = -b.BC2 * (r[1] - dn * b5);
  double boris1 = -b.BC2 * (r[2] - dn * b6);
  double boris2 = -b.BC2 * (r[3] - dn * b7);
  double boris3 = boris0 * v1 + boris1 * v2 + boris2 * v3;
  r[1] += boris0; r[2] += boris1; r[3] += boris2; r[4] += boris3;
  double qres = 0.5 * (q + qm) * qft1;
  r[4] += P.dt_tvd * qres;
#pragma unroll
  for (int k = 0; k < MRM_NV; k++) u[k] += r[k];
  if (cn) cn->hit(MRM_BR_RHO_FLOOR, MRM_BK_RHO_FLOOR, 0, u[0] < P.rho_min, cn->cls(fabs(u[0] - P.rho_min), P.rho_min));
  u[0] = smax(P.rho_min, u[0]);
  if (DIAG) {
    dg[1] = boris3;
    dg[2] = qres;
    dg[3] = u[4];
  }
  const double efl = P.eps_min * u[0] + 0.5 * (u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) / u[0];
  if (cn) cn->hit(MRM_BR_E_FLOOR, MRM_BK_E_FLOOR, 0, u[4] < efl, cn->cls(fabs(u[4] - efl), fmax(fabs(u[4]), fabs(efl))));
  u[4] = smax(u[4], efl);
}

}  // namespace MRM_NS
