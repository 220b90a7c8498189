// adjoint_cells.cpp -- CPU check of the adjoint's cell functions: for each
// vector-Jacobian product in muram_main_amd/csrc/mrm_adjoint.h, compare
// <lambda, (f(x + h d) - f(x - h d)) / 2h> with <vjp(x, lambda), d> over
// random primal points, cotangents and directions, where f is the forward
// cell function the GPU kernels call (compiled here for the host).
// Usage: adjoint_cells <eos table file>; prints one line per function and
// exits non-zero if any relative difference exceeds the tolerance.
#define MRM_NS adj
#include "mrm_adjoint.h"

#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <string>
#include <vector>

using namespace adj;
using Vec = std::vector<double>;
using Fwd = std::function<Vec(const Vec&)>;
using Vjp = std::function<Vec(const Vec&, const Vec&)>;

static std::mt19937_64 rng(12345);
static double urand(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
static double nrand() { return std::normal_distribution<double>(0.0, 1.0)(rng); }

// EOS table in the reference's format, transformed as mrm_eos_load does
struct HostEos {
  std::vector<double> ax;
  std::vector<float> tab, itab;
  EosDev E;
};
static bool load_eos(const char* path, HostEos& H) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return false;
  std::vector<float> buf;
  float x;
  while (std::fread(&x, 4, 1, f) == 1) buf.push_back(x);
  std::fclose(f);
  int Ne = (int)buf[0], Nr = (int)buf[1], Np = (int)buf[2], Ns = (int)buf[3];
  double e0 = buf[4], e1 = buf[5], r0 = buf[6], r1 = buf[7], p0 = buf[8], p1 = buf[9], s0 = buf[10], s1 = buf[11];
  double de = (e1 - e0) / (Ne - 1), dr = (r1 - r0) / (Nr - 1), dp = (p1 - p0) / (Np - 1), ds = (s1 - s0) / (Ns - 1);
  H.ax.resize(Ne + Nr + Np + Ns);
  for (int i = 0; i < Ne; i++) H.ax[i] = e0 + de * i;
  for (int i = 0; i < Nr; i++) H.ax[Ne + i] = r0 + dr * i;
  for (int i = 0; i < Np; i++) H.ax[Ne + Nr + i] = p0 + dp * i;
  for (int i = 0; i < Ns; i++) H.ax[Ne + Nr + Np + i] = s0 + ds * i;
  size_t nt = (size_t)Ne * Nr, ni = (size_t)Np * Ns;
  H.tab.resize(6 * nt);
  H.itab.resize(2 * ni);
  const float* fw = buf.data() + 14;
  for (int t = 0; t < 6; t++)
    for (int i = 0; i < Ne; i++)
      for (int j = 0; j < Nr; j++) H.tab[t * nt + (size_t)i * Nr + j] = (float)(((double)fw[t * nt + (size_t)j * Ne + i]) / (de * dr));
  const float* iv = fw + 6 * nt;
  for (int t = 0; t < 2; t++)
    for (int i = 0; i < Np; i++)
      for (int j = 0; j < Ns; j++) H.itab[t * ni + (size_t)i * Ns + j] = (float)(((double)iv[t * ni + (size_t)j * Np + i]) / (dp * ds));
  EosDev& E = H.E;
  E.N_eps = Ne; E.N_lr = Nr; E.N_lp3 = Np; E.N_s3 = Ns;
  E.inv_del_eps = 1.0 / de; E.inv_del_lr = 1.0 / dr; E.inv_del_lp3 = 1.0 / dp; E.inv_del_s3 = 1.0 / ds;
  E.eps_off = buf[12]; E.ss_off = buf[13];
  E.xeps = H.ax.data(); E.xlr = E.xeps + Ne; E.xlp3 = E.xlr + Nr; E.xs3 = E.xlp3 + Np;
  E.eps_max = std::exp(E.xeps[Ne - 1]) - E.eps_off;
  E.xeps0 = E.xeps[0]; E.xlr0 = E.xlr[0]; E.xlp30 = E.xlp3[0]; E.xs30 = E.xs3[0];
  E.rxeps = nullptr; E.rxlr = nullptr;
  E.tab = H.tab.data(); E.itab = H.itab.data();
  return true;
}

static int failures = 0;

// directional check of one VJP; scale[i] sets the size of x_i's perturbation
static void check(const char* name, const Vec& x, const Vec& scale, size_t m, Fwd f, Vjp vjp, int trials = 12,
                  double tol = 1e-6, double h = 1e-6) {
  double worst = 0.0;
  int used = 0;
  for (int t = 0; t < trials; t++) {
    Vec ly(m), d(x.size());
    for (auto& v : ly) v = nrand();
    for (size_t i = 0; i < x.size(); i++) d[i] = scale[i] * nrand();
    Vec xp = x, xm = x;
    for (size_t i = 0; i < x.size(); i++) { xp[i] += h * d[i]; xm[i] -= h * d[i]; }
    Vec yp = f(xp), ym = f(xm);
    double fd = 0.0, mag = 0.0;
    for (size_t i = 0; i < m; i++) {
      fd += ly[i] * (yp[i] - ym[i]) / (2 * h);
      mag += std::fabs(ly[i] * (yp[i] + ym[i]) / 2);
    }
    Vec lx = vjp(x, ly);
    double ad = 0.0, amag = 0.0;
    for (size_t i = 0; i < x.size(); i++) { ad += lx[i] * d[i]; amag += std::fabs(lx[i] * d[i]); }
    // floor: the dot products can cancel; FD resolves them to ~1e-8 of the summed terms
    const double den = std::max(std::max(std::fabs(fd), std::fabs(ad)), 1e-8 * std::max(amag, 1e-300));
    if (den == 0.0) continue;
    worst = std::max(worst, std::fabs(fd - ad) / den);
    used++;
  }
  const bool ok = worst <= tol && used > 0;
  if (!ok) failures++;
  std::printf("%-26s trials %2d  max rel diff %.3e  %s\n", name, used, worst, ok ? "ok" : "FAIL");
}

// a stratified-atmosphere-like conservative cell; with g_thin set, one from
// the tenuous top of a tall box (rho 1e-13 .. 1e-19 g/cm^3, fast flows, hot),
// where the Alfven limiter is strong
static bool g_thin = false;
static Vec cell(const EosDev& E) {
  const double rho = g_thin ? 1e-13 * std::exp(-urand(0.0, 14.0)) : 3e-7 * std::exp(-urand(0.0, 3.0));
  const double vs = g_thin ? 3e6 : 2e5;
  const double v[3] = {urand(-vs, vs), urand(-vs, vs), urand(-vs, vs)};
  const double eps = g_thin ? std::exp(urand(std::log(2e12), std::log(2e15))) : urand(1.5e12, 6e12);
  const double B[3] = {56.0 + urand(-15, 15), urand(-15, 15), urand(-15, 15)};
  Vec u(8);
  u[0] = rho;
  for (int k = 0; k < 3; k++) u[1 + k] = rho * v[k];
  u[4] = rho * eps + 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; k++) u[5 + k] = B[k];
  (void)E;
  return u;
}
static Vec cell_scale(const Vec& u) {
  Vec s(8);
  for (int q = 0; q < 8; q++) s[q] = std::max(std::fabs(u[q]), 1e-3 * std::fabs(u[0]) + 1e-300);
  s[1] = s[2] = s[3] = (g_thin ? 3e6 : 2e5) * u[0];
  s[5] = s[6] = s[7] = 10.0;
  return s;
}
// perturbation scales of a primitive point (rho, v, e_int, B, p)
static Vec prim_scale(const double* w) {
  Vec s(9);
  s[0] = std::fabs(w[0]);
  s[1] = s[2] = s[3] = g_thin ? 3e6 : 2e5;
  s[4] = std::fabs(w[4]);
  s[5] = s[6] = s[7] = 10.0;
  s[8] = std::fabs(w[8]);
  return s;
}
static double eos_p(const EosDev& E, const double* u) {
  double m1 = u[1], m2 = u[2], m3 = u[3], e = u[4], out[5];
  cons_to_prim_cell<2>(E, u[0], m1, m2, m3, e, out);
  return out[0];
}
static void prim9(const EosDev& E, const double* u, double* w) {
  const double d = u[0], dn = 1.0 / d;
  w[0] = d;
  w[1] = u[1] * dn; w[2] = u[2] * dn; w[3] = u[3] * dn;
  w[4] = u[4] - 0.5 * (w[1] * w[1] + w[2] * w[2] + w[3] * w[3]) * d;
  w[5] = u[5]; w[6] = u[6]; w[7] = u[7];
  w[8] = eos_p(E, u);
}

static TvdParams tvd_params() {
  TvdParams P{};
  P.rho_min = 1e-20; P.eps_min = 1e11; P.rho_lev = 1e-22; P.q_rho_max = std::log(10.0);
  P.visc_coeff_bot = 1.0; P.visc_slope_bot = 0.0; P.visc_coeff_top = 1.0; P.visc_slope_top = 0.0;
  P.eta_coeff_bot = 1.0; P.eta_slope_bot = 0.0; P.eta_coeff_top = 1.0; P.eta_slope_top = 0.0;
  P.B_par_diff = 0.2; P.vhyp = 1.0; P.tvd_pm_v = 1.0; P.tvd_pm_B = 4.0; P.vmax_lim = 0.75; P.CME_thresh = 0.5;
  const double h[4] = {2, 2, 1, 0}, cs[4] = {0.2, 0.2, 0.2, 1};
  for (int q = 0; q < 4; q++) { P.tvd_h[q] = h[q]; P.tvd_cs[q] = cs[q]; P.tc[q] = 1.0; }
  P.need_tvd_coeff = 0;
  P.inv_va2max = g_thin ? 1.0 / (3e8 * 3e8) : 1.0 / (3e8 * 3e8) * 1e4;   // limiter active
  P.vsqr_diff = std::pow(0.75 * 1e7, 2);
  P.dt_tvd = 0.2;
  P.cmax = 0.975 * 3.2e6 / 0.2;
  P.idxd1 = 1.0 / 3.2e6;
  P.rho_log = 1; P.is_gbeg0 = 1; P.is_gend0 = 1;
  return P;
}

static void run_all(const EosDev& E) {
  // ---- prim
  {
    Vec u = cell(E);
    check("prim", u, cell_scale(u), 8,
          [&](const Vec& x) { double w[9]; prim9(E, x.data(), w); return Vec(w, w + 8); },
          [&](const Vec& x, const Vec& l) { Vec lu(8, 0.0); prim_vjp(x.data(), l.data(), lu.data()); return lu; });
  }
  // ---- EOS pressure
  for (int r = 0; r < 3; r++) {
    Vec u = cell(E);
    check("eos_p", u, cell_scale(u), 1, [&](const Vec& x) { return Vec{eos_p(E, x.data())}; },
          [&](const Vec& x, const Vec& l) { Vec lu(8, 0.0); eos_p_vjp(E, x.data(), l[0], lu.data()); return lu; });
  }
  // ---- flux13 and lfac
  {
    Vec u = cell(E);
    Vec w(9);
    prim9(E, u.data(), w.data());
    Vec sc = prim_scale(w.data());
    auto fl = [&](auto D) {
      constexpr int d = decltype(D)::value;
      check(d == 0 ? "flux13<0>" : d == 1 ? "flux13<1>" : "flux13<2>", w, sc, 13,
            [&](const Vec& x) { Vec f(13); flux13<d>(x.data(), f.data()); return f; },
            [&](const Vec& x, const Vec& l) { Vec lx(9, 0.0); flux13_vjp<d>(x.data(), l.data(), lx.data()); return lx; });
    };
    fl(std::integral_constant<int, 0>());
    fl(std::integral_constant<int, 1>());
    fl(std::integral_constant<int, 2>());
    const double inv = 1.0 / (1.5e5 * 1.5e5);   // x^2 ~ 1: the limiter factor varies at first order
    check("lfac", w, sc, 1, [&](const Vec& x) { return Vec{lfac_of(x.data(), inv)}; },
          [&](const Vec& x, const Vec& l) { Vec lx(9, 0.0); lfac_vjp(x.data(), inv, l[0], lx.data()); return lx; });
  }
  // ---- one sweep's residual over a 5-point stencil (flux differences,
  // Boris-blended Lorentz force, work): x = 5 primitive points (45)
  {
    const double inv = g_thin ? 1.0 / (3e8 * 3e8) : 1.0 / (3e6 * 3e6);
    const double w1 = 8. / (12. * 3.2e6), w2 = -1. / (12. * 3.2e6);
    Vec x(45), sc(45);
    for (int p = 0; p < 5; p++) {
      Vec u = cell(E);
      prim9(E, u.data(), &x[9 * p]);
      Vec s9 = prim_scale(&x[9 * p]);
      for (int q = 0; q < 9; q++) sc[9 * p + q] = s9[q];
    }
    auto sw = [&](auto D) {
      constexpr int d = decltype(D)::value;
      auto fwd = [&](const Vec& xx) {
        auto get = [&](int s, double o[ST_NF]) { for (int f = 0; f < 9; f++) o[f] = xx[9 * s + f]; };
        double res[8];
        const double lf = lfac_of(&xx[18], inv);
        sweep_residual_lr<d>(get, &xx[18], w1, w2, lf, res);
        return Vec(res, res + 8);
      };
      auto vjp = [&](const Vec& xx, const Vec& l) {
        Vec lx(45, 0.0);
        double f[5][13];
        for (int p = 0; p < 5; p++) if (p != 2) flux13<d>(&xx[9 * p], f[p]);
        double Dl[13];
        for (int q = 0; q < 13; q++) Dl[q] = w1 * (f[3][q] - f[1][q]) + w2 * (f[4][q] - f[0][q]);
        const double lf = lfac_of(&xx[18], inv);
        double G[13], llf = 0.0;
        sweep_centre_vjp<d>(&xx[18], Dl, lf, l.data(), G, &lx[18], llf);
        lfac_vjp(&xx[18], inv, llf, &lx[18]);
        const int pt[4] = {0, 1, 3, 4};
        const double wt[4] = {-w2, -w1, w1, w2};
        for (int t = 0; t < 4; t++) {
          double lf13[13];
          for (int q = 0; q < 13; q++) lf13[q] = wt[t] * G[q];
          flux13_vjp<d>(&xx[9 * pt[t]], lf13, &lx[9 * pt[t]]);
        }
        return lx;
      };
      check(d == 0 ? "sweep<0>" : d == 1 ? "sweep<1>" : "sweep<2>", x, sc, 8, fwd, vjp);
    };
    sw(std::integral_constant<int, 0>());
    sw(std::integral_constant<int, 1>());
    sw(std::integral_constant<int, 2>());
  }
  // ---- Tcheck (three regimes: inside, energy clamped low, speed limited)
  for (int r = 0; r < 3; r++) {
    Vec u = cell(E);
    double vmax = 1e7, emin = 1e11, emax = 4e15;
    if (r == 1) emin = 1e13;                 // lower energy bound active
    if (r == 2) vmax = 1e5;                  // speed limit active
    check(r == 0 ? "tcheck" : r == 1 ? "tcheck (e clamp)" : "tcheck (v limit)", u, cell_scale(u), 8,
          [&](const Vec& x) { Vec y = x; tcheck_vals(y.data(), 1e-20, emin, emax, vmax); return y; },
          [&](const Vec& x, const Vec& l) { Vec lu(8, 0.0); tcheck_vjp(x.data(), 1e-20, emin, emax, vmax, l.data(), lu.data()); return lu; });
  }
  // ---- sources, Boris, RK, Tcheck: x = (up 8, R 8, u0 8)
  for (int r = 0; r < 2; r++) {
    SrcParams SP;
    SP.wdt = 0.05; SP.invRdt = 0.1; SP.g = -2.74e4; SP.inv_va2max = r == 0 ? 1.0 / (3e6 * 3e6) : 1.0 / (3e8 * 3e8);
    SP.rho_min = 1e-20; SP.eps_min = 1e11; SP.eps_max = 4e15; SP.vmax = 1e7;
    Vec u = cell(E), u0 = cell(E);
    double w[9];
    prim9(E, u.data(), w);
    Vec x(24), sc(24);
    Vec us = cell_scale(u);
    Vec s9 = prim_scale(w);
    for (int q = 0; q < 8; q++) {
      x[q] = w[q]; sc[q] = s9[q];
      x[8 + q] = 1e-3 * us[q] * nrand(); sc[8 + q] = 1e-3 * us[q];
      x[16 + q] = u0[q]; sc[16 + q] = cell_scale(u0)[q];
    }
    const double Qrad = r == 0 ? 0.0 : -1e-2 * w[4];
    check(r == 0 ? "source_integrate" : "source_integrate (Qrad)", x, sc, 8,
          [&](const Vec& xx) { double R[8], out[8]; for (int q = 0; q < 8; q++) R[q] = xx[8 + q];
                               source_integrate_cell(SP, &xx[0], R, &xx[16], Qrad, 0.0, out); return Vec(out, out + 8); },
          [&](const Vec& xx, const Vec& l) { Vec lx(24, 0.0);
                                             source_integrate_vjp(SP, &xx[0], &xx[8], &xx[16], Qrad, l.data(), &lx[0], &lx[8], &lx[16]);
                                             return lx; });
  }
  // ---- TVD point: x = (u 8, p)
  const TvdParams P = tvd_params();
  auto pt_fields = [](const TvdPt& o) {
    Vec y(TW_N);
    for (int q = 0; q < 8; q++) y[TW_V + q] = o.v[q];
    y[TW_LNR] = o.lnr; y[TW_VSQR] = o.vsqr; y[TW_BSQR] = o.bsqr; y[TW_CM] = o.cm; y[TW_HE] = o.hyp_e;
    y[TW_HV] = o.hyp_v; y[TW_HB] = o.hyp_B; y[TW_BC2] = o.BC2; y[TW_CZ] = 0.0; y[TW_HD] = o.hd;
    return y;
  };
  for (int r = 0; r < 3; r++) {
    Vec u = cell(E);
    Vec x(9), sc(9);
    Vec us = cell_scale(u);
    for (int q = 0; q < 8; q++) { x[q] = u[q]; sc[q] = us[q]; }
    x[8] = eos_p(E, u.data());
    sc[8] = x[8];
    const double hfb = r == 1 ? 0.3 : 0.0, hft = r == 2 ? 0.2 : 0.0;
    const bool xs = r != 1;
    check("tvd_point", x, sc, TW_N,
          [&](const Vec& xx) { TvdPt o; tvd_point(P, xx.data(), xx[8], 0.0, hfb, hft, xs, o); return pt_fields(o); },
          [&](const Vec& xx, const Vec& l) { Vec lx(9, 0.0); double lp = 0.0;
                                             tvd_point_vjp(P, xx.data(), xx[8], 0.0, hfb, hft, xs, l.data(), lx.data(), lp);
                                             lx[8] = lp; return lx; });
  }
  // ---- points of a short line for slopes / interfaces / update
  auto mkpt = [&](const Vec& u, TvdPt& o, bool xs) { tvd_point(P, u.data(), eos_p(E, u.data()), 0.0, 0.0, 0.0, xs, o); };
  {
    // slope: x = a.v, b.v, c.v (24); branch fields held at the primal values
    for (int r = 0; r < 3; r++) {
      TvdPt a, b, c;
      Vec ua = cell(E), ub = cell(E), uc = cell(E);
      if (r == 1) { ub = ua; for (int q = 0; q < 8; q++) { ub[q] *= 1.0 + 0.01 * nrand(); uc[q] = ub[q] * (1.0 + 0.02 * nrand()); } }
      mkpt(ua, a, true); mkpt(ub, b, true); mkpt(uc, c, true);
      Vec x(24), sc(24, 1.0);
      for (int q = 0; q < 8; q++) { x[q] = a.v[q]; x[8 + q] = b.v[q]; x[16 + q] = c.v[q]; }
      for (int q = 0; q < 24; q++) sc[q] = std::max(std::fabs(x[q]) * 1e-3, 1e-3);
      auto setv = [&](const Vec& xx, TvdPt& A, TvdPt& B, TvdPt& C) {
        A = a; B = b; C = c;
        for (int q = 0; q < 8; q++) { A.v[q] = xx[q]; B.v[q] = xx[8 + q]; C.v[q] = xx[16 + q]; }
      };
      check("tvd_slope", x, sc, 8,
            [&](const Vec& xx) { TvdPt A, B, C; setv(xx, A, B, C); double S[8]; tvd_slope(P, A, B, C, 0.0, 0.0, S); return Vec(S, S + 8); },
            [&](const Vec& xx, const Vec& l) { TvdPt A, B, C; setv(xx, A, B, C); Vec lx(24, 0.0);
                                               tvd_slope_vjp(P, A, B, C, 0.0, 0.0, l.data(), &lx[0], &lx[8], &lx[16]); return lx; },
            12, 1e-6, 1e-4);   // piecewise linear: a larger step has no truncation error
    }
  }
  // interface: x = a.v[0,1,4] (3), b fields (TW_N), c fields (TW_N), d.v[0,1,4] (3), Sb (8), Sc (8)
  auto iface_check = [&](auto Dc, bool hyper, const char* name) {
    constexpr int D = decltype(Dc)::value;
    for (int r = 0; r < 3; r++) {
      TvdPt a, b, c, d;
      Vec ua = cell(E), ub = cell(E), uc = cell(E), ud = cell(E);
      mkpt(ua, a, D == 0); mkpt(ub, b, D == 0); mkpt(uc, c, D == 0); mkpt(ud, d, D == 0);
      double Sb[8], Sc[8];
      for (int q = 0; q < 8; q++) { Sb[q] = 0.1 * (c.v[q] - b.v[q]) * urand(-1, 1); Sc[q] = 0.1 * (c.v[q] - b.v[q]) * urand(-1, 1); }
      const int hk[3] = {0, 1, 4};
      const int nb = 3 + 2 * TW_N + 3 + 16;
      Vec x(nb), sc(nb, 1.0);
      for (int t = 0; t < 3; t++) { x[t] = a.v[hk[t]]; x[3 + 2 * TW_N + t] = d.v[hk[t]]; }
      Vec bf = pt_fields(b), cf = pt_fields(c);
      for (int q = 0; q < TW_N; q++) { x[3 + q] = bf[q]; x[3 + TW_N + q] = cf[q]; }
      for (int q = 0; q < 8; q++) { x[6 + 2 * TW_N + q] = Sb[q]; x[14 + 2 * TW_N + q] = Sc[q]; }
      for (int q = 0; q < nb; q++) sc[q] = std::max(std::fabs(x[q]) * 1e-3, 1e-6);
      auto unpack = [&](const Vec& xx, TvdPt& A, TvdPt& B, TvdPt& C, TvdPt& Dd, double* sb, double* scc) {
        A = a; B = b; C = c; Dd = d;
        for (int t = 0; t < 3; t++) { A.v[hk[t]] = xx[t]; Dd.v[hk[t]] = xx[3 + 2 * TW_N + t]; }
        for (int q = 0; q < 8; q++) { B.v[q] = xx[3 + TW_V + q]; C.v[q] = xx[3 + TW_N + TW_V + q]; }
        B.cm = xx[3 + TW_CM]; B.hyp_e = xx[3 + TW_HE]; B.hyp_v = xx[3 + TW_HV]; B.hyp_B = xx[3 + TW_HB]; B.hd = xx[3 + TW_HD];
        C.cm = xx[3 + TW_N + TW_CM]; C.hd = xx[3 + TW_N + TW_HD];
        for (int q = 0; q < 8; q++) { sb[q] = xx[6 + 2 * TW_N + q]; scc[q] = xx[14 + 2 * TW_N + q]; }
      };
      const double qft = r == 2 ? 0.0 : 1.0;
      check(name, x, sc, 9,
            [&](const Vec& xx) { TvdPt A, B, C, Dd; double sb[8], scc[8], F[8]; unpack(xx, A, B, C, Dd, sb, scc);
                                 double q = tvd_iface<D>(P, A, B, C, Dd, sb, scc, 0.0, 0.0, qft, hyper, false, F);
                                 Vec y(F, F + 8); y.push_back(q); return y; },
            [&](const Vec& xx, const Vec& l) { TvdPt A, B, C, Dd; double sb[8], scc[8]; unpack(xx, A, B, C, Dd, sb, scc);
                                               TvdIfaceAdj Ad;
                                               tvd_iface_vjp<D>(P, A, B, C, Dd, sb, scc, 0.0, 0.0, qft, hyper, false, l.data(), l[8], Ad);
                                               Vec lx(nb, 0.0);
                                               for (int t = 0; t < 3; t++) { lx[t] = Ad.la[t]; lx[3 + 2 * TW_N + t] = Ad.ld[t]; }
                                               for (int q = 0; q < TW_N; q++) { lx[3 + q] = Ad.lb[q]; lx[3 + TW_N + q] = Ad.lc[q]; }
                                               for (int q = 0; q < 8; q++) { lx[6 + 2 * TW_N + q] = Ad.lSb[q]; lx[14 + 2 * TW_N + q] = Ad.lSc[q]; }
                                               return lx; },
            12, g_thin ? 1e-5 : 1e-6, 1e-5);   // tenuous cells: FD resolves to ~1e-5
    }
  };
  iface_check(std::integral_constant<int, 0>(), true, "tvd_iface<0> (hyper)");
  iface_check(std::integral_constant<int, 1>(), false, "tvd_iface<1>");
  iface_check(std::integral_constant<int, 2>(), false, "tvd_iface<2>");
  // update: x = u (8), F (8), Fm (8), q, qm, b.v1..3, b.v5..7, BC2, bsqr (8)
  for (int r = 0; r < 2; r++) {
    TvdPt b;
    Vec u = cell(E);
    mkpt(u, b, true);
    const int nb = 8 + 8 + 8 + 2 + 8;
    Vec x(nb), sc(nb);
    Vec us = cell_scale(u);
    for (int q = 0; q < 8; q++) {
      x[q] = u[q]; sc[q] = us[q];
      x[8 + q] = 1e-2 * us[q] * nrand(); sc[8 + q] = 1e-2 * us[q];
      x[16 + q] = 1e-2 * us[q] * nrand(); sc[16 + q] = 1e-2 * us[q];
    }
    x[24] = 1e-3 * us[4] * nrand(); sc[24] = 1e-3 * us[4];
    x[25] = 1e-3 * us[4] * nrand(); sc[25] = 1e-3 * us[4];
    const int bi[8] = {TW_V + 1, TW_V + 2, TW_V + 3, TW_V + 5, TW_V + 6, TW_V + 7, TW_BC2, TW_BSQR};
    Vec bfv = pt_fields(b);
    for (int t = 0; t < 8; t++) { x[26 + t] = bfv[bi[t]]; sc[26 + t] = std::max(std::fabs(bfv[bi[t]]) * 1e-2, 1e-6); }
    const double qft = r == 0 ? 1.0 : 0.0;
    auto setb = [&](const Vec& xx, TvdPt& B) {
      B = b;
      B.v[1] = xx[26]; B.v[2] = xx[27]; B.v[3] = xx[28]; B.v[5] = xx[29]; B.v[6] = xx[30]; B.v[7] = xx[31];
      B.BC2 = xx[32]; B.bsqr = xx[33];
    };
    check(r == 0 ? "tvd_update" : "tvd_update (qft 0)", x, sc, 8,
          [&](const Vec& xx) { TvdPt B; setb(xx, B); Vec y(xx.begin(), xx.begin() + 8);
                               tvd_update(P, B, &xx[16], &xx[8], xx[25], xx[24], qft, y.data()); return y; },
          [&](const Vec& xx, const Vec& l) { TvdPt B; setb(xx, B); Vec lx(nb, 0.0); double lq = 0.0, lqm = 0.0, lb[TW_N] = {0};
                                             tvd_update_vjp(P, B, &xx[16], &xx[8], xx[25], xx[24], qft, xx.data(), l.data(),
                                                            &lx[0], &lx[8], &lx[16], lq, lqm, lb);
                                             lx[24] = lq; lx[25] = lqm;
                                             for (int t = 0; t < 8; t++) lx[26 + t] = lb[bi[t]];
                                             return lx; });
  }
  // ---- boundary
  {
    Vec u = cell(E);
    check("bc_sum_p", u, cell_scale(u), 1,
          [&](const Vec& x) { const double d = x[0];
                              const double e = x[4] - 0.5 * ((x[1] * x[1] + x[2] * x[2] + x[3] * x[3]) / d);
                              return Vec{p_interp(E, e / d, d)}; },
          [&](const Vec& x, const Vec& l) { Vec lu(8, 0.0); bc_sum_p_vjp(E, x.data(), l[0], lu.data()); return lu; });
  }
  for (int r = 0; r < 2; r++) {
    BcParams B;
    Vec u1 = cell(E), u2 = cell(E);
    u1[1] = std::fabs(u1[1]) * (r == 0 ? 1.0 : -1.0);   // inflow / outflow branch
    const double pm1 = eos_p(E, u1.data()) * 1.01, pm2 = eos_p(E, u2.data()) * 0.99;
    B.p_bc = std::sqrt(pm1 * pm2) * 1.02;
    B.s_bc = s_interp(E, (u1[4] - 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3]) / u1[0]) / u1[0], u1[0]);
    B.B_crit = 300.0;
    B.eps_top = 0.0; B.cc = 1.0; B.open_top = 1;
    Vec x(18), sc(18);
    Vec s1 = cell_scale(u1), s2 = cell_scale(u2);
    for (int q = 0; q < 8; q++) { x[q] = u1[q]; x[8 + q] = u2[q]; sc[q] = s1[q]; sc[8 + q] = s2[q]; }
    x[16] = pm1; x[17] = pm2; sc[16] = pm1 * 0.1; sc[17] = pm2 * 0.1;
    check(r == 0 ? "bc_bottom (inflow)" : "bc_bottom (outflow)", x, sc, 16,
          [&](const Vec& xx) { double a1[8], a2[8]; for (int q = 0; q < 8; q++) { a1[q] = xx[q]; a2[q] = xx[8 + q]; }
                               bc_bottom_cells(E, B, xx[16], xx[17], a1, a2);
                               Vec y(a1, a1 + 8); y.insert(y.end(), a2, a2 + 8); return y; },
          [&](const Vec& xx, const Vec& l) { Vec lx(18, 0.0);
                                             bc_bottom_vjp(E, B, &xx[0], &xx[8], xx[16], xx[17], &l[0], &l[8], &lx[0], &lx[8], lx[16], lx[17]);
                                             return lx; });
  }
  for (int r = 0; r < 2; r++) {
    BcParams B{};
    B.open_top = 1;
    B.eps_top = r == 0 ? 0.0 : 3e12;
    Vec u1 = cell(E), u2 = cell(E);
    Vec x(16), sc(16);
    Vec s1 = cell_scale(u1), s2 = cell_scale(u2);
    for (int q = 0; q < 8; q++) { x[q] = u1[q]; x[8 + q] = u2[q]; sc[q] = s1[q]; sc[8 + q] = s2[q]; }
    check(r == 0 ? "bc_top" : "bc_top (eps_top)", x, sc, 16,
          [&](const Vec& xx) { double a1[8], a2[8]; for (int q = 0; q < 8; q++) { a1[q] = xx[q]; a2[q] = xx[8 + q]; }
                               bc_top_cells(B, a1, a2);
                               Vec y(a1, a1 + 8); y.insert(y.end(), a2, a2 + 8); return y; },
          [&](const Vec& xx, const Vec& l) { Vec lx(16, 0.0); bc_top_vjp(B, &xx[0], &xx[8], &l[0], &l[8], &lx[0], &lx[8]); return lx; });
  }
}

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: %s eos_table\n", argv[0]); return 2; }
  HostEos H;
  if (!load_eos(argv[1], H)) { std::fprintf(stderr, "cannot read %s\n", argv[1]); return 2; }
  const EosDev& E = H.E;

  run_all(E);
  g_thin = true;
  std::printf("-- tenuous cells (strong Alfven limiter)\n");
  run_all(E);
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
  return failures ? 1 : 0;
}
