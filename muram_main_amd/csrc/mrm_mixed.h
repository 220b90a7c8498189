// mrm_mixed.h -- FP32-mixed form of the fused MHD step.
//
// Storage: rho, M0..M2, B0..B2 and the pressure in FP32, the total energy e in
// FP64 (36 + 4 bytes per cell instead of 64 + 8).  Stencils, fluxes, the
// Lorentz force, sources and the TVD limiter run in FP32.  FP64 is kept where
// cancellation or accumulation needs it:
//  * internal-energy recovery e_int = e - M^2/(2 rho) (stage loads, the EOS,
//    the TVD's specific internal energy), which cancels in low-beta plasma;
//  * the RK accumulation of the energy, e = e0 + wdt*R_e, and the energy
//    clamps of Tcheck and TVDlimit;
//  * the characteristic-speed reduction that sets dt.
// The operation sequence of every formula is the FP64 path's (mhdres_SR.C,
// Add_Sources_Integrate_SR.C, tvdlimit_SR.C, cons_to_prim.C).
#pragma once
#include "mrm_fused.h"
#include "mrm_divb.h"
#include "mrm_spitzer.h"

namespace MRM_NS {

__device__ __forceinline__ int mx_slot(int q) { return q < 4 ? q : q - 1; }

__device__ __forceinline__ float fsmax(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ float fsmin(float a, float b) { return __builtin_fminf(a, b); }

// FP32 reciprocal, quotient and square root from the hardware estimates
// (v_rcp_f32, v_sqrt_f32: about 1 ulp), adequate for the FP32 stencils
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// FP64 reciprocal: v_rcp_f64 estimate refined by two Newton steps (error
// below 1 ulp of the FP64 quotient's rounding), for the FP64 islands
__device__ __forceinline__ double drcp(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// ------------------------------------------------------------------ point data
// primitive point (rho, v, e_int, B, p) in FP32 from the mixed state; e_int
// recovered in FP64
// the stored values of one point (rho, M, B, p in FP32; e in FP64), so that
// the load can be issued ahead of its use
struct MxRaw {
  float f[8];   // rho, M0, M1, M2, B0, B1, B2, p
  double e;
};
__device__ __forceinline__ void mx_load_raw(const MxState& S, const float* __restrict__ P, long bs, unsigned n,
                                            MxRaw& r) {
#pragma unroll
  for (int q = 0; q < 7; q++) r.f[q] = ld32(S.f + q * bs, n);
  r.f[7] = ld32(P, n);
  r.e = ld32(S.e, n);
}
// primitive point (rho, v, e_int, B, p) in FP32, e_int recovered in FP64
__device__ __forceinline__ void mx_prim_of_raw(const MxRaw& r, float w[ST_NF]) {
  const float d = r.f[0], m0 = r.f[1], m1 = r.f[2], m2 = r.f[3];
  double dd = (double)d;
  double mm = (double)m0 * (double)m0 + (double)m1 * (double)m1 + (double)m2 * (double)m2;
  double ei = r.e - 0.5 * mm * drcp(dd);
  float dn = frcp(d);
  w[0] = d;
  w[1] = m0 * dn;
  w[2] = m1 * dn;
  w[3] = m2 * dn;
  w[4] = (float)ei;
  w[5] = r.f[4];
  w[6] = r.f[5];
  w[7] = r.f[6];
  w[8] = r.f[7];
}
__device__ __forceinline__ void mx_load_prim(const MxState& S, const float* __restrict__ P, long bs, unsigned n,
                                             float w[ST_NF]) {
  MxRaw r;
  mx_load_raw(S, P, bs, n, r);
  mx_prim_of_raw(r, w);
}

template <int D>
__device__ __forceinline__ void mx_point_flux(const float var[ST_NF], float f[13], int d2, int d3) {
  const float p1 = D == 0 ? 1.0f : 0.0f, p2 = D == 1 ? 1.0f : 0.0f, p3 = D == 2 ? 1.0f : 0.0f;
  float var0 = var[0], var1 = var[1], var2 = var[2], var3 = var[3];
  float var4 = var[4], var5 = var[5], var6 = var[6], var7 = var[7], pres = var[8];
  float vel = var[1 + D];
  float mag = var[5 + D];
  float vv = var1 * var1 + var2 * var2 + var3 * var3;
  float bb = var5 * var5 + var6 * var6 + var7 * var7;
  float ekin = 0.5f * var0 * vv;
  float pmag = 0.5f * bb;
  float mflx = vel * var0;
  f[0] = mflx;
  f[1] = mflx * var1 + p1 * pres;
  f[2] = mflx * var2 + p2 * pres;
  f[3] = mflx * var3 + p3 * pres;
  f[4] = vel * (var4 + ekin + pres);
  f[5] = vel * var5 - var1 * mag;
  f[6] = vel * var6 - var2 * mag;
  f[7] = vel * var7 - var3 * mag;
  f[8] = mag * var5 - p1 * pmag;
  f[9] = mag * var6 - p2 * pmag;
  f[10] = mag * var7 - p3 * pmag;
  f[11] = var[5 + d2];
  f[12] = var[5 + d3];
}

__device__ __forceinline__ float mx_lfac(const float c[ST_NF], float inv_va2max) {
  float dn = frcp(c[0]);
  float bb = c[5] * c[5] + c[6] * c[6] + c[7] * c[7];
  float x2 = bb * dn * inv_va2max;
  float x4 = x2 * x2;
  float s = 1.0f + x2;
  return fdiv(s, s + x4);
}

// residual of sweep D at the centre c (mhdres_SR.C:361-412) in FP32.  The
// Lorentz force of the sweep is returned in lor[3] instead of being added to
// res[1..3] and its work to res[4]: the caller combines it with the Boris
// correction without cancellation.  ETA: the sweep's share of J = curl B and
// eta times the second derivative of B (mhdres_SR.C:422-462), as eta_sweep.
// the residual of sweep D from the flux differences d1 = F(+1) - F(-1),
// d2 = F(+2) - F(-2) and (ETA) the sums bs1 = B(+1) + B(-1), bs2 = B(+2) + B(-2)
template <int D, bool ETA>
__device__ __forceinline__ void mx_sweep_comb(const float d1[13], const float d2[13], const float bs1[3],
                                              const float bs2[3], const float c[ST_NF], float w1, float w2,
                                              float lf, float res[MRM_NV], float lor[3], const MxStageArgs* A,
                                              float* J) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) res[q] = 0.0f - (w1 * d1[q] + w2 * d2[q]);
  float dB2 = w1 * d1[11] + w2 * d2[11];
  float dB3 = w1 * d1[12] + w2 * d2[12];
  float cb[3];
  cb[D] = -c[5 + D2] * dB2 - c[5 + D3] * dB3;
  cb[D2] = c[5 + D] * dB2;
  cb[D3] = c[5 + D] * dB3;
  float lrt0 = w1 * d1[8] + w2 * d2[8];
  float lrt1 = w1 * d1[9] + w2 * d2[9];
  float lrt2 = w1 * d1[10] + w2 * d2[10];
  lor[0] = lf * lrt0 + (1.0f - lf) * cb[0];
  lor[1] = lf * lrt1 + (1.0f - lf) * cb[1];
  lor[2] = lf * lrt2 + (1.0f - lf) * cb[2];
  if constexpr (ETA) {
    constexpr float sgn = D == 1 ? 1.0f : -1.0f;
    J[D2] += sgn * dB3;
    J[D3] += -sgn * dB2;
#pragma unroll
    for (int q = 0; q < 3; q++)
      res[5 + q] += A->eta * (A->ww0[D] * c[5 + q] + A->ww1[D] * bs1[q] + A->ww2[D] * bs2[q]);
  }
}

template <int D, bool ETA = false, class GET>
__device__ __forceinline__ void mx_sweep(GET get, const float c[ST_NF], float w1, float w2, float lf,
                                         float res[MRM_NV], float lor[3], const MxStageArgs* A = nullptr,
                                         float* J = nullptr) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  float d1[13], d2[13];
  float bs1[3], bs2[3];   // B(+1) + B(-1), B(+2) + B(-2) (ETA only)
  {
    float pa[ST_NF], pb[ST_NF], fa[13], fb[13];
    get(1, pa);
    mx_point_flux<D>(pa, fa, D2, D3);
    get(3, pb);
    mx_point_flux<D>(pb, fb, D2, D3);
#pragma unroll
    for (int q = 0; q < 13; q++) d1[q] = fb[q] - fa[q];
    if constexpr (ETA) {
#pragma unroll
      for (int q = 0; q < 3; q++) bs1[q] = pb[5 + q] + pa[5 + q];
    }
  }
  {
    float pa[ST_NF], pb[ST_NF], fa[13], fb[13];
    get(0, pa);
    mx_point_flux<D>(pa, fa, D2, D3);
    get(4, pb);
    mx_point_flux<D>(pb, fb, D2, D3);
#pragma unroll
    for (int q = 0; q < 13; q++) d2[q] = fb[q] - fa[q];
    if constexpr (ETA) {
#pragma unroll
      for (int q = 0; q < 3; q++) bs2[q] = pb[5 + q] + pa[5 + q];
    }
  }
  mx_sweep_comb<D, ETA>(d1, d2, bs1, bs2, c, w1, w2, lf, res, lor, A, J);
}

// the point flux of sweep D packed for sharing through LDS: mx_point_flux's
// 13 values and, with ETA, B along D (the three B components are then
// f[13], f[11], f[12] for B[D], B[D2], B[D3])
template <bool ETA>
constexpr int mx_nfp() { return ETA ? 14 : 13; }
template <int D, bool ETA>
__device__ __forceinline__ void mx_flux_pack(const float var[ST_NF], float f[14]) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  mx_point_flux<D>(var, f, D2, D3);
  if (ETA) f[13] = var[5 + D];
}

// sweep D from packed neighbour fluxes: get(s, f) returns value f of the
// packed flux at offset s - 2 along D
template <int D, bool ETA, class GET>
__device__ __forceinline__ void mx_sweep_packed(GET get, const float c[ST_NF], float w1, float w2, float lf,
                                                float res[MRM_NV], float lor[3], const MxStageArgs* A,
                                                float* J) {
  constexpr int D2 = D == 0 ? 1 : 0;
  constexpr int D3 = D == 2 ? 1 : 2;
  float d1[13], d2[13], bs1[3], bs2[3];
#pragma unroll
  for (int q = 0; q < 13; q++) {
    d1[q] = get(3, q) - get(1, q);
    d2[q] = get(4, q) - get(0, q);
  }
  if constexpr (ETA) {
    bs1[D] = get(3, 13) + get(1, 13);
    bs1[D2] = get(3, 11) + get(1, 11);
    bs1[D3] = get(3, 12) + get(1, 12);
    bs2[D] = get(4, 13) + get(0, 13);
    bs2[D2] = get(4, 11) + get(0, 11);
    bs2[D3] = get(4, 12) + get(0, 12);
  }
  mx_sweep_comb<D, ETA>(d1, d2, bs1, bs2, c, w1, w2, lf, res, lor, A, J);
}

// pressure from the mixed conservative values (ConsToPrim's p and its
// validity fallback, cons_to_prim.C:110-198).  e_int is recovered in FP64;
// the bilinear lookup runs in FP32 on offsets from the cell's lower node,
// ln(x) - x_i = ln(x * exp(-x_i)) (a log of a ratio near one, accurate to
// FP32 rounding), with weights that sum to one; the interpolated ln p is
// split into the node value (FP64 product, hi + lo) and the FP32 increment.
// The lookup is split in two so that a caller can issue the memory reads of
// one cell (mx_eos_issue: the node indices and every table value) and do
// other work before the arithmetic that consumes them (mx_eos_finish).
struct MxEosG {
  double dd, ei, eps, rxe, rxr;
  int i_d;
  float p00, p10, p01, p11, t00, t10, t01, t11;   // ln p and ln T at the cell's four nodes
};

__device__ __forceinline__ void mx_eos_issue(const EosDev& E, float d, float m0, float m1, float m2, double e,
                                             MxEosG& g) {
  const double dd = (double)d;
  const double mm = (double)m0 * (double)m0 + (double)m1 * (double)m1 + (double)m2 * (double)m2;
  const double rd = drcp(dd);
  g.dd = dd;
  g.ei = e - 0.5 * mm * rd;
  g.eps = g.ei * rd + E.eps_off;
  // ln rho and ln eps only choose the table cell (the offsets from its node
  // are formed accurately below), so the hardware log2 serves; a cell chosen
  // one off at a node boundary interpolates to the same value
  const float lr = __builtin_amdgcn_logf(d) * 0.693147180559945f;
  const float ep = __builtin_amdgcn_logf((float)g.eps) * 0.693147180559945f;
  int i_d = cvt_int((double)(lr - (float)E.xlr0) * E.inv_del_lr);
  i_d = imax_(0, i_d);
  i_d = imin_(E.N_lr - 2, i_d);
  int i_e = cvt_int((double)(ep - (float)E.xeps0) * E.inv_del_eps);
  i_e = imax_(0, i_e);
  i_e = imin_(E.N_eps - 2, i_e);
  g.i_d = i_d;
  g.rxe = E.rxeps[i_e];
  g.rxr = E.rxlr[i_d];
  const long L = E.N_lr, NT = (long)E.N_eps * E.N_lr;
  const long o00 = i_e * L + i_d, o01 = o00 + 1, o10 = o00 + L, o11 = o10 + 1;
  g.p00 = E.tab[o00]; g.p10 = E.tab[o10]; g.p01 = E.tab[o01]; g.p11 = E.tab[o11];
  g.t00 = E.tab[NT + o00]; g.t10 = E.tab[NT + o10]; g.t01 = E.tab[NT + o01]; g.t11 = E.tab[NT + o11];
}

// WANT_T: the temperature is formed too and returned in *temp_out
template <bool WANT_T = false>
__device__ __forceinline__ float mx_eos_finish(const EosDev& E, const MxEosG& g, double* temp_out = nullptr) {
  const double eos_gamma = 1.65;
  const double c_pres = (eos_gamma - 1.);
  const double c_temp = (eos_gamma - 1.) * 0.62 / 8.314e7;
  const double ei = g.ei, eps = g.eps;
  const float te = logf((float)(eps * g.rxe));     // ln(eps) - xeps[i_e]
  const float tr = logf((float)(g.dd * g.rxr));    // ln(rho) - xlr[i_d]
  const float fe = te * (float)E.inv_del_eps;
  const float fr = tr * (float)E.inv_del_lr;
  const double area = 1.0 / (E.inv_del_eps * E.inv_del_lr);
  // T(fe, fr) = T00 + fe (T10 - T00) + fr (T01 - T00) + fe fr (T11 - T10 - T01 + T00)
  auto incr = [&](float t00, float t10, float t01, float t11) {
    float de = t10 - t00, dr = t01 - t00, dx = (t11 - t10) - (t01 - t00);
    return (fe * de + fr * dr + fe * fr * dx) * (float)area;
  };
  const double pb = (double)g.p00 * area;
  const float pb_hi = (float)pb, pb_lo = (float)(pb - (double)pb_hi);
  double pres = (double)expf(pb_hi) * (double)expf(pb_lo + incr(g.p00, g.p10, g.p01, g.p11));
  const float tt = (float)((double)g.t00 * area) + incr(g.t00, g.t10, g.t01, g.t11);
  double temp;
  const bool lowr = (g.i_d == 0) && (tr < 0.0f);            // ln(rho) < xlr[0]
  const bool hot = eps > 0.8 * E.eps_max;
  if (!WANT_T && !lowr && !hot && tt > -80.0f && tt < 80.0f) {
    temp = 1.0;
  } else {
    temp = (double)expf(tt);
  }
  if (lowr) {
    pres = c_pres * ei;
    temp = c_temp * eps;
  }
  if (hot) {
    double xx = smax(0.0, 5.0 * (E.eps_max - eps) / E.eps_max);
    pres = pres * xx + (1.0 - xx) * c_pres * ei;
    temp = temp * xx + (1.0 - xx) * c_temp * eps;
  }
  int bad = 0;
  if ((pres != pres) || (pres <= 0.) || (pres == __builtin_inf())) bad = 1;
  if ((temp != temp) || (temp <= 0.) || (temp == __builtin_inf())) bad = 1;
  if (bad) {
    pres = c_pres * ei;
    temp = c_temp * eps;
  }
  if (WANT_T) *temp_out = temp;
  return (float)pres;
}

template <bool WANT_T = false>
__device__ __forceinline__ float mx_pres(const EosDev& E, float d, float m0, float m1, float m2, double e,
                                         double* temp_out = nullptr) {
  MxEosG g;
  mx_eos_issue(E, d, m0, m1, m2, e, g);
  return mx_eos_finish<WANT_T>(E, g, temp_out);
}

// characteristic speed in FP64 from the mixed point (mhdres_SR.C:310-328)
__device__ __forceinline__ double mx_cmax(float d, float m0, float m1, float m2, float b0, float b1, float b2,
                                          float p, double inv_va2max) {
  double w[MRM_NV];
  double dn = 1.0 / (double)d;
  w[0] = d;
  w[1] = m0 * dn;
  w[2] = m1 * dn;
  w[3] = m2 * dn;
  w[4] = 0.0;
  w[5] = b0;
  w[6] = b1;
  w[7] = b2;
  return point_cmax(w, (double)p, inv_va2max);
}

// Tcheck (Add_Sources_Integrate_SR.C:447-465) with the energy bounds in FP64
__device__ __forceinline__ void mx_tcheck(float u[MRM_NV], double& e, float rho_min, double eps_min,
                                          double eps_max, float vmax) {
  float dnn = fmaxf(rho_min, u[0]);
  float vv = fsqrt(u[1] * u[1] + u[2] * u[2] + u[3] * u[3]) * frcp(dnn);
  float s = fdiv(vmax, fmaxf(vmax, vv));
  double ekin = 0.5 * (double)dnn * (double)vv * (double)vv;
  double ss = (double)s * (double)s;
  double lo = eps_min * (double)dnn + ekin * ss;
  double hi = eps_max * (double)dnn + ekin * ss;
  e = fmin(fmax(e, lo), hi);
  u[0] = dnn;
  u[1] *= s;
  u[2] *= s;
  u[3] *= s;
}

// the rest of a cell's stage after the three sweeps' sums acc (without the
// Lorentz force) and accL (the Lorentz force): eta |J|^2 and the conduction
// term, sources, Boris correction, the RK update (FP64 energy) and Tcheck;
// stores the new state of the cell and returns it in o (rho, M, B) and oe (e)
template <bool ETA, bool ERES>
__device__ __forceinline__ void mx_cell_finish(const MxStageArgs& A, const float w2[ST_NF], float acc[MRM_NV],
                                               const float accL[3], const float J[3], unsigned node, MxState S0,
                                               MxState Sout, const double* __restrict__ Qtot, float o[MRM_NV],
                                               double& oe) {
  const long bs = A.G.bs;
  if (ETA) acc[4] += A.eta * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
  if (ERES) acc[4] += (float)ld32(A.eres, node);

  // sources, Boris correction (Add_Sources_Integrate_SR.C:331-392), FP32
  const float* up = w2;
  const float Qrad = Qtot ? (float)ld32(Qtot, node) : 0.0f;
  float dn = A.invRdt * up[4];
  float efaci = fdiv(dn, fmaxf(dn, fabsf(Qrad)));
  dn = A.g * up[0];
  float e_dmpi = dn * up[1];
  acc[1] += dn;
  acc[4] += e_dmpi + Qrad * efaci;
  {
    // Lorentz force f_L (accL) and its work, Boris correction
    // (Add_Sources_Integrate_SR.C:355-392).  With f = f_h + f_L - v R_rho
    // the reference adds f_L, v.f_L, then -bf f_perp and -bf v.f_perp.
    // Where the Alfven limiter is strong bf -> 1 and those terms cancel
    // most digits; with lb = 1 - bf = (1+x2)/(1+x2+x4) the same sums are
    //   R_M = v R_rho + f_par + lb f_perp
    //   R_e += prL (v.b) + lb v.f_L,perp - bf v.(f_perp - f_L,perp)
    // (prL b = the part of f_L along b), which FP32 keeps accurate.
    float rh = up[0], rd = acc[0];
    float vx = up[1], vy = up[2], vz = up[3];
    float bx = up[5], by = up[6], bz = up[7];
    float fx = (acc[1] + accL[0]) - vx * rd, fy = (acc[2] + accL[1]) - vy * rd, fz = (acc[3] + accL[2]) - vz * rd;
    float bb = bx * bx + by * by + bz * bz;
    float x2 = bb * frcp(rh) * A.inv_va2max;
    float x4 = x2 * x2;
    float sx = 1 + x2 + x4;
    const float isx = frcp(sx);
    float bf = x4 * isx, lb = (1 + x2) * isx;
    float ib = frcp(fmaxf(1e-30f, bb));
    float pr = (fx * bx + fy * by + fz * bz) * ib;
    float px = fx - pr * bx, py = fy - pr * by, pz = fz - pr * bz;
    float prL = (accL[0] * bx + accL[1] * by + accL[2] * bz) * ib;
    float lx = accL[0] - prL * bx, ly = accL[1] - prL * by, lz = accL[2] - prL * bz;
    acc[1] = (vx * rd + pr * bx) + lb * px;
    acc[2] = (vy * rd + pr * by) + lb * py;
    acc[3] = (vz * rd + pr * bz) + lb * pz;
    acc[4] += prL * (vx * bx + vy * by + vz * bz) + lb * (vx * lx + vy * ly + vz * lz) -
              bf * (vx * (px - lx) + vy * (py - ly) + vz * (pz - lz));
  }
  // RK update: FP32 for rho, M, B; FP64 for e (Add_Sources_Integrate_SR.C:399)
  const float wdtf = (float)A.wdt;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) o[q] = q == 4 ? 0.0f : ld32(S0.f + mx_slot(q) * bs, node) + wdtf * acc[q];
  oe = ld32(S0.e, node) + A.wdt * (double)acc[4];
  mx_tcheck(o, oe, A.rho_min, A.eps_min, A.eps_max, A.vmax);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    if (q == 4) continue;
    st32(Sout.f + mx_slot(q) * bs, node, o[q]);
  }
  st32(Sout.e, node, oe);
}

// ------------------------------------------------------------------ stage kernel
// ETA: resistive terms in the sweeps and eta |J|^2 in the energy residual
// ERES: A.eres holds a further energy residual per cell (conduction), added
// after the sweeps
// WP: the pressure of the updated cells is written to Pout.  The pressure of
// plane k is formed during plane k+1's work: its table reads are issued
// before that plane's sweeps and consumed after them, so their latency hides
// behind the sweeps; the chunk's last plane is finished after the loop.
template <bool ETA = false, bool ERES = false, bool WP = true>
__global__ void __launch_bounds__(MX_SX* MX_SY, 2)
k_stage_mx(MxStageArgs A, MxState Sin, const float* __restrict__ Pin, MxState S0, MxState Sout,
           float* __restrict__ Pout, const double* __restrict__ Qtot) {
  __shared__ float lds[ST_NF][MX_LY][MX_LX];
  const Geom& G = A.G;
  const long bs = G.bs;
  const long nblocks = (long)A.ntx * A.nty * A.nkc;
  long lb = xcd_remap(blockIdx.x, nblocks);
  const int tx_t = (int)(lb % A.ntx);
  lb /= A.ntx;
  const int ty_t = (int)(lb % A.nty);
  const int kc = (int)(lb / A.nty);
  const int tx = threadIdx.x % MX_SX, ty = threadIdx.x / MX_SX;
  const int i0 = G.lb0 + tx_t * MX_SX, j0 = G.lb1 + ty_t * MX_SY;
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = (i <= G.le0) && (j <= G.le1);
  const int kbeg = G.lb2 + kc * MX_KC;
  const int kend = min(kbeg + MX_KC, G.le2 + 1);
  int hx_c = -1, hx_i = 0;
  if (tx < 2) { hx_c = tx; hx_i = i0 - 2 + tx; }
  else if (tx >= MX_SX - 2) { hx_c = tx + 4; hx_i = i0 + tx + 2; }
  int hy_r = -1, hy_j = 0;
  if (ty < 4) {
    hy_r = ty < 2 ? ty : MX_SY + ty;
    hy_j = ty < 2 ? j0 - 2 + ty : j0 + MX_SY + (ty - 2);
  }
  const bool hx_ok = hx_c >= 0 && hx_i < G.n0 && j < G.n1;
  const bool hy_ok = hy_r >= 0 && hy_j < G.n1 && i < G.n0;
  const bool own_ok = i < G.n0 && j < G.n1;
  const unsigned col = (unsigned)i + (unsigned)j * (unsigned)G.s1;
  const unsigned s2u = (unsigned)G.s2, s1u = (unsigned)G.s1;

  float w[5][ST_NF];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    int kk = kbeg - 2 + q;
    if (own_ok) mx_load_prim(Sin, Pin, bs, col + (unsigned)kk * s2u, w[q + 1]);
    else {
#pragma unroll
      for (int f = 0; f < ST_NF; f++) w[q + 1][f] = 1.0f;
    }
  }
  // the next plane's values (own column k+3, x and y halo of k+1) are loaded
  // during the current plane's work; kbeg + 2 <= lend2 + 2 = n2 - 1
  // the cell whose pressure is pending (plane k - 1 of this column); the
  // initial values are a harmless placeholder lookup that is never stored
  float pd_d = 1.0f, pd_m0 = 0.0f, pd_m1 = 0.0f, pd_m2 = 0.0f;
  double pd_e = 1.0;
  unsigned pd_node = 0;
  MxRaw rz, rhx, rhy;
  if (own_ok) mx_load_raw(Sin, Pin, bs, col + (unsigned)(kbeg + 2) * s2u, rz);
  if (hx_ok) mx_load_raw(Sin, Pin, bs, (unsigned)hx_i + (unsigned)j * s1u + (unsigned)kbeg * s2u, rhx);
  if (hy_ok) mx_load_raw(Sin, Pin, bs, (unsigned)i + (unsigned)hy_j * s1u + (unsigned)kbeg * s2u, rhy);
  for (int k = kbeg; k < kend; k++) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int f = 0; f < ST_NF; f++) w[q][f] = w[q + 1][f];
    if (own_ok) mx_prim_of_raw(rz, w[4]);
#pragma unroll
    for (int f = 0; f < ST_NF; f++) lds[f][ty + 2][tx + 2] = w[2][f];
    if (hx_ok) {
      float h[ST_NF];
      mx_prim_of_raw(rhx, h);
#pragma unroll
      for (int f = 0; f < ST_NF; f++) lds[f][ty + 2][hx_c] = h[f];
    }
    if (hy_ok) {
      float h[ST_NF];
      mx_prim_of_raw(rhy, h);
#pragma unroll
      for (int f = 0; f < ST_NF; f++) lds[f][hy_r][tx + 2] = h[f];
    }
    if (k + 1 < kend) {
      // k + 3 <= kend + 1 <= lend2 + 2 = n2 - 1
      if (own_ok) mx_load_raw(Sin, Pin, bs, col + (unsigned)(k + 3) * s2u, rz);
      if (hx_ok) mx_load_raw(Sin, Pin, bs, (unsigned)hx_i + (unsigned)j * s1u + (unsigned)(k + 1) * s2u, rhx);
      if (hy_ok) mx_load_raw(Sin, Pin, bs, (unsigned)i + (unsigned)hy_j * s1u + (unsigned)(k + 1) * s2u, rhy);
    }
    __syncthreads();
    if (valid) {
      MxEosG pg;
      if (WP) mx_eos_issue(A.E, pd_d, pd_m0, pd_m1, pd_m2, pd_e, pg);
      float acc[MRM_NV], res[MRM_NV], accL[3] = {0.0f, 0.0f, 0.0f}, lor[3];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] = 0.0f;
      const float lfc = mx_lfac(w[2], A.inv_va2max);
      auto getx = [&](int s, float o[ST_NF]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = lds[f][ty + 2][tx + s];
      };
      float J[3] = {0.0f, 0.0f, 0.0f};
      mx_sweep<0, ETA>(getx, w[2], A.w1[0], A.w2[0], lfc, res, lor, &A, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
#pragma unroll
      for (int q = 0; q < 3; q++) accL[q] += lor[q];
      auto gety = [&](int s, float o[ST_NF]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = lds[f][ty + s][tx + 2];
      };
      mx_sweep<1, ETA>(gety, w[2], A.w1[1], A.w2[1], lfc, res, lor, &A, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
#pragma unroll
      for (int q = 0; q < 3; q++) accL[q] += lor[q];
      auto getz = [&](int s, float o[ST_NF]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = w[s][f];
      };
      mx_sweep<2, ETA>(getz, w[2], A.w1[2], A.w2[2], lfc, res, lor, &A, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
#pragma unroll
      for (int q = 0; q < 3; q++) accL[q] += lor[q];
      const unsigned node = col + (unsigned)k * s2u;
      float o[MRM_NV];
      double oe;
      mx_cell_finish<ETA, ERES>(A, w[2], acc, accL, J, node, S0, Sout, Qtot, o, oe);
      if (WP) {
        const float pk = mx_eos_finish(A.E, pg);
        if (k > kbeg) st32(Pout, pd_node, pk);
        pd_d = o[0]; pd_m0 = o[1]; pd_m1 = o[2]; pd_m2 = o[3];
        pd_e = oe;
        pd_node = node;
      }
    }
    __syncthreads();
  }
  if (WP && valid && kend > kbeg) st32(Pout, pd_node, mx_pres(A.E, pd_d, pd_m0, pd_m1, pd_m2, pd_e));
}

// ------------------------------------------------------------------ stage kernel, fluxes shared
// k_stage_mx with each x and y point flux formed once per plane: every thread
// packs the fluxes of its own point along dims 0 and 1 into LDS, the ring
// points' fluxes are packed by one thread each (threads 0 .. 4*SX-1 the four
// dim-1 ring rows, the next 4*SY threads the dim-0 ring columns), and the x
// and y sweeps take the four neighbours' fluxes from LDS (4.5 point-flux
// evaluations per plane and wave instead of 12).  Dim-2 neighbours come from
// the five-plane register window.  Each term is formed with k_stage_mx's
// expressions.
template <bool ETA = false, bool ERES = false, bool WP = true>
__global__ void __launch_bounds__(MX_SX* MX_SY, 2)
k_stage_mxf(MxStageArgs A, MxState Sin, const float* __restrict__ Pin, MxState S0, MxState Sout,
            float* __restrict__ Pout, const double* __restrict__ Qtot) {
  constexpr int NFP = mx_nfp<ETA>();
  static_assert(4 * MX_SX + 4 * MX_SY <= MX_SX * MX_SY, "one ring point per thread at most");
  __shared__ float fx[NFP][MX_SY][MX_LX];   // dim-0 fluxes of row ty, columns i0-2 .. i0+SX+1
  __shared__ float fy[NFP][MX_LY][MX_SX];   // dim-1 fluxes of column tx, rows j0-2 .. j0+SY+1
  const Geom& G = A.G;
  const long bs = G.bs;
  const long nblocks = (long)A.ntx * A.nty * A.nkc;
  long lb = xcd_remap(blockIdx.x, nblocks);
  const int tx_t = (int)(lb % A.ntx);
  lb /= A.ntx;
  const int ty_t = (int)(lb % A.nty);
  const int kc = (int)(lb / A.nty);
  const int t = threadIdx.x;
  const int tx = t % MX_SX, ty = t / MX_SX;
  const int i0 = G.lb0 + tx_t * MX_SX, j0 = G.lb1 + ty_t * MX_SY;
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = (i <= G.le0) && (j <= G.le1);
  const int kbeg = G.lb2 + kc * MX_KC;
  const int kend = min(kbeg + MX_KC, G.le2 + 1);
  int ring = 0;                       // 1: dim-1 ring point, 2: dim-0 ring point
  int r_row = 0, r_col = 0, r_i = 0, r_j = 0;
  if (t < 4 * MX_SX) {
    const int rr_ = t / MX_SX;
    r_row = rr_ < 2 ? rr_ : MX_SY + rr_;
    r_col = t % MX_SX;
    r_i = i0 + r_col;
    r_j = j0 - 2 + r_row;
    ring = 1;
  } else if (t < 4 * MX_SX + 4 * MX_SY) {
    const int u = t - 4 * MX_SX;
    r_row = u / 4;
    const int c4 = u % 4;
    r_col = c4 < 2 ? c4 : MX_SX + c4;
    r_i = i0 - 2 + r_col;
    r_j = j0 + r_row;
    ring = 2;
  }
  const bool ring_ok = ring != 0 && r_i < G.n0 && r_j < G.n1;
  const bool own_ok = i < G.n0 && j < G.n1;
  const unsigned col = (unsigned)i + (unsigned)j * (unsigned)G.s1;
  const unsigned rcol = (unsigned)r_i + (unsigned)r_j * (unsigned)G.s1;
  const unsigned s2u = (unsigned)G.s2;

  float w[5][ST_NF];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    int kk = kbeg - 2 + q;
    if (own_ok) mx_load_prim(Sin, Pin, bs, col + (unsigned)kk * s2u, w[q + 1]);
    else {
#pragma unroll
      for (int f = 0; f < ST_NF; f++) w[q + 1][f] = 1.0f;
    }
  }
  float pd_d = 1.0f, pd_m0 = 0.0f, pd_m1 = 0.0f, pd_m2 = 0.0f;
  double pd_e = 1.0;
  unsigned pd_node = 0;
  MxRaw rz, rr;   // own column's plane k + 3, the ring point's plane k + 1
  if (own_ok) mx_load_raw(Sin, Pin, bs, col + (unsigned)(kbeg + 2) * s2u, rz);
  if (ring_ok) mx_load_raw(Sin, Pin, bs, rcol + (unsigned)kbeg * s2u, rr);
  for (int k = kbeg; k < kend; k++) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int f = 0; f < ST_NF; f++) w[q][f] = w[q + 1][f];
    if (own_ok) mx_prim_of_raw(rz, w[4]);
    {
      float o[14];
      mx_flux_pack<0, ETA>(w[2], o);
#pragma unroll
      for (int f = 0; f < NFP; f++) fx[f][ty][tx + 2] = o[f];
      mx_flux_pack<1, ETA>(w[2], o);
#pragma unroll
      for (int f = 0; f < NFP; f++) fy[f][ty + 2][tx] = o[f];
    }
    if (ring != 0) {
      float h[ST_NF], o[14];
      if (ring_ok) mx_prim_of_raw(rr, h);
      else {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) h[f] = 1.0f;
      }
      if (ring == 1) {
        mx_flux_pack<1, ETA>(h, o);
#pragma unroll
        for (int f = 0; f < NFP; f++) fy[f][r_row][r_col] = o[f];
      } else {
        mx_flux_pack<0, ETA>(h, o);
#pragma unroll
        for (int f = 0; f < NFP; f++) fx[f][r_row][r_col] = o[f];
      }
    }
    if (k + 1 < kend) {
      // k + 3 <= kend + 1 <= lend2 + 2 = n2 - 1
      if (own_ok) mx_load_raw(Sin, Pin, bs, col + (unsigned)(k + 3) * s2u, rz);
      if (ring_ok) mx_load_raw(Sin, Pin, bs, rcol + (unsigned)(k + 1) * s2u, rr);
    }
    __syncthreads();
    if (valid) {
      MxEosG pg;
      if (WP) mx_eos_issue(A.E, pd_d, pd_m0, pd_m1, pd_m2, pd_e, pg);
      float acc[MRM_NV], res[MRM_NV], accL[3] = {0.0f, 0.0f, 0.0f}, lor[3];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] = 0.0f;
      const float lfc = mx_lfac(w[2], A.inv_va2max);
      float J[3] = {0.0f, 0.0f, 0.0f};
      auto getx = [&](int s_, int f) { return fx[f][ty][tx + s_]; };
      mx_sweep_packed<0, ETA>(getx, w[2], A.w1[0], A.w2[0], lfc, res, lor, &A, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
#pragma unroll
      for (int q = 0; q < 3; q++) accL[q] += lor[q];
      auto gety = [&](int s_, int f) { return fy[f][ty + s_][tx]; };
      mx_sweep_packed<1, ETA>(gety, w[2], A.w1[1], A.w2[1], lfc, res, lor, &A, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
#pragma unroll
      for (int q = 0; q < 3; q++) accL[q] += lor[q];
      auto getz = [&](int s_, float o[ST_NF]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = w[s_][f];
      };
      mx_sweep<2, ETA>(getz, w[2], A.w1[2], A.w2[2], lfc, res, lor, &A, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
#pragma unroll
      for (int q = 0; q < 3; q++) accL[q] += lor[q];
      const unsigned node = col + (unsigned)k * s2u;
      float o[MRM_NV];
      double oe;
      mx_cell_finish<ETA, ERES>(A, w[2], acc, accL, J, node, S0, Sout, Qtot, o, oe);
      if (WP) {
        const float pk = mx_eos_finish(A.E, pg);
        if (k > kbeg) st32(Pout, pd_node, pk);
        pd_d = o[0]; pd_m0 = o[1]; pd_m1 = o[2]; pd_m2 = o[3];
        pd_e = oe;
        pd_node = node;
      }
    }
    __syncthreads();
  }
  if (WP && valid && kend > kbeg) st32(Pout, pd_node, mx_pres(A.E, pd_d, pd_m0, pd_m1, pd_m2, pd_e));
}

// ------------------------------------------------------------------ ghost fill
template <int DIM>
__global__ void __launch_bounds__(MRM_BLOCK) k_exchange_mx(Geom G, MxState S, float* __restrict__ P) {
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
  for (int q = 0; q < 7; q++) {
    S.f[q * G.bs + dst_r] = S.f[q * G.bs + src_r];
    S.f[q * G.bs + dst_l] = S.f[q * G.bs + src_l];
  }
  S.e[dst_r] = S.e[src_r];
  S.e[dst_l] = S.e[src_l];
  if (P) {
    P[dst_r] = P[src_r];
    P[dst_l] = P[src_l];
  }
}

// partial sums of p over the two lowest interior layers (FP64 from the mixed state)
__global__ void __launch_bounds__(MRM_BLOCK) k_bc_sums_mx(Geom G, EosDev E, MxState S, int mode,
                                                          double* __restrict__ part) {
  __shared__ double lds[5 * MRM_BLOCK / 64];
  const long bs = G.bs;
  const int ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < (long)ny * nz) {
    int j = G.lb1 + (int)(t % ny), k = G.lb2 + (int)(t / ny);
    long n1 = G.lb0 + j * G.s1 + k * G.s2, n2 = n1 + 1;
    double d1 = S.f[n1], d2 = S.f[n2];
    double m1x = S.f[bs + n1], m1y = S.f[2 * bs + n1], m1z = S.f[3 * bs + n1];
    double m2x = S.f[bs + n2], m2y = S.f[2 * bs + n2], m2z = S.f[3 * bs + n2];
    double e1 = S.e[n1] - 0.5 * ((m1x * m1x + m1y * m1y + m1z * m1z) / d1);
    double e2 = S.e[n2] - 0.5 * ((m2x * m2x + m2y * m2y + m2z * m2z) / d2);
    s[0] = p_interp(E, e1 / d1, d1);
    s[1] = p_interp(E, e2 / d2, d2);
    if (mode == 1) {
      double si = s_interp(E, e1 / d1, d1);
      if (m1x > 0.0) { s[2] = 1.0; s[3] = si; }
      s[4] = si;
    }
  }
  block_reduce<OpAdd, 5>(s, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 5; q++) part[q * gridDim.x + blockIdx.x] = s[q];
  }
}

// bottom/top ghost layers (boundary_pdmp_1_fftw3.C:240-318, 402-474) in FP64,
// stored in the mixed layout with their pressures
__global__ void __launch_bounds__(MRM_BLOCK) k_bc_mx(Geom G, EosDev E, BcParams B, const double* __restrict__ psum,
                                                     MxState S, float* __restrict__ P, double inv_va2max,
                                                     unsigned long long* cmax_out) {
  __shared__ double lds[MRM_BLOCK / 64];
  const long bs = G.bs;
  // one thread per (column, side): the bottom's inverse-EOS chain and the top
  // run in separate threads
  const long ncol = (long)G.n1 * G.n2;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double cm[1] = {0.0};
  const bool top = t >= ncol;
  if (top) t -= ncol;
  if (t < ncol) {
    int j = (int)(t % G.n1), k = (int)(t / G.n1);
    const bool inner = j >= G.lb1 && j <= G.le1 && k >= G.lb2 && k <= G.le2;
    long base = j * G.s1 + k * G.s2;
    auto ld = [&](long n, double u[MRM_NV]) {
      u[0] = S.f[n]; u[1] = S.f[bs + n]; u[2] = S.f[2 * bs + n]; u[3] = S.f[3 * bs + n];
      u[4] = S.e[n]; u[5] = S.f[4 * bs + n]; u[6] = S.f[5 * bs + n]; u[7] = S.f[6 * bs + n];
    };
    // both ghost cells' pressures are formed before any store, so that the
    // two EOS lookups overlap (the stores could alias the tables for the
    // compiler)
    auto st2 = [&](long na, const double ua[MRM_NV], long nb, const double ub[MRM_NV]) {
      float fa[MRM_NV], fb[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) { fa[q] = (float)ua[q]; fb[q] = (float)ub[q]; }
      float pa = 0.0f, pb = 0.0f;
      if (P) {
        pa = mx_pres(E, fa[0], fa[1], fa[2], fa[3], ua[4]);
        pb = mx_pres(E, fb[0], fb[1], fb[2], fb[3], ub[4]);
      }
      S.f[na] = fa[0]; S.f[bs + na] = fa[1]; S.f[2 * bs + na] = fa[2]; S.f[3 * bs + na] = fa[3];
      S.e[na] = ua[4];
      S.f[4 * bs + na] = fa[5]; S.f[5 * bs + na] = fa[6]; S.f[6 * bs + na] = fa[7];
      S.f[nb] = fb[0]; S.f[bs + nb] = fb[1]; S.f[2 * bs + nb] = fb[2]; S.f[3 * bs + nb] = fb[3];
      S.e[nb] = ub[4];
      S.f[4 * bs + nb] = fb[5]; S.f[5 * bs + nb] = fb[6]; S.f[6 * bs + nb] = fb[7];
      if (P) {
        P[na] = pa;
        P[nb] = pb;
        if (cmax_out && inner) {
          cm[0] = smax(cm[0], mx_cmax(fa[0], fa[1], fa[2], fa[3], fa[5], fa[6], fa[7], pa, inv_va2max));
          cm[0] = smax(cm[0], mx_cmax(fb[0], fb[1], fb[2], fb[3], fb[5], fb[6], fb[7], pb, inv_va2max));
        }
      }
    };
    if (!top) {
      long n1 = G.lb0 + base, n2 = n1 + 1;
      double u1[MRM_NV], u2[MRM_NV];
      ld(n1, u1);
      ld(n2, u2);
      bc_bottom_cells(E, B, psum[0] * B.cc, psum[1] * B.cc, u1, u2);
      st2(G.lb0 - 1 + base, u1, G.lb0 - 2 + base, u2);
    } else {
      long n1 = G.le0 + base, n2 = n1 - 1;
      double u1[MRM_NV], u2[MRM_NV];
      ld(n1, u1);
      ld(n2, u2);
      bc_top_cells(B, u1, u2);
      st2(G.le0 + 1 + base, u1, G.le0 + 2 + base, u2);
    }
  }
  if (cmax_out) {
    block_reduce<OpMax, 1>(cm, lds);
    if (threadIdx.x == 0) atomic_max_nonneg(cmax_out, cm[0]);
  }
}

// Adjust_Valf_Max's per-cell terms from a mixed cell: velocity and specific
// internal energy in FP64 from the stored FP32 density/momentum and FP64 energy
__device__ __forceinline__ void mx_valf_cell(ValfAcc& a, float d, float m0, float m1, float m2, double e, float p,
                                             double vlim95, double elim95) {
  const double dd = d, dn = 1.0 / dd;
  const double v0 = m0 * dn, v1 = m1 * dn, v2 = m2 * dn;
  const double eint = e - 0.5 * (v0 * v0 + v1 * v1 + v2 * v2) * dd;
  valf_cell(a, dd, v0, v1, v2, eint, p, vlim95, elim95);
}

// pressure of the whole buffer, cmax (interior + dim-d ghost faces) and,
// with valf, the Adjust_Valf_Max statistics of the interior
__global__ void __launch_bounds__(MRM_BLOCK) k_eos_p_mx(Geom G, EosDev E, MxState S, float* __restrict__ P,
                                                        double inv_va2max, unsigned long long* cmax_out,
                                                        unsigned long long* valf, double vlim95, double elim95) {
  __shared__ double lds[5 * MRM_BLOCK / 64];
  const long bs = G.bs;
  double cm[1] = {0.0};
  ValfAcc va;
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
    float d = S.f[n], m0 = S.f[bs + n], m1 = S.f[2 * bs + n], m2 = S.f[3 * bs + n];
    const double e = S.e[n];
    float p = mx_pres(E, d, m0, m1, m2, e);
    P[n] = p;
    int i = (int)(n % G.n0);
    long r = n / G.n0;
    int j = (int)(r % G.n1), k = (int)(r / G.n1);
    bool ii = i >= G.lb0 && i <= G.le0, ji = j >= G.lb1 && j <= G.le1, ki = k >= G.lb2 && k <= G.le2;
    bool vis = (ii && ji && ki) || (!ii && ji && ki) || (ii && !ji && ki) || (ii && ji && !ki);
    if (vis) cm[0] = smax(cm[0], mx_cmax(d, m0, m1, m2, S.f[4 * bs + n], S.f[5 * bs + n], S.f[6 * bs + n], p,
                                          inv_va2max));
    if (valf && ii && ji && ki) mx_valf_cell(va, d, m0, m1, m2, e, p, vlim95, elim95);
  }
  block_reduce<OpMax, 1>(cm, lds);
  if (threadIdx.x == 0) atomic_max_nonneg(cmax_out, cm[0]);
  if (valf) {
    __syncthreads();
    valf_flush(va, lds, valf);
  }
}

// cmax alone over the same points from the state and its pressure (the
// stage-1 bound after Adjust_Valf_Max has changed inv_va2max)
__global__ void __launch_bounds__(MRM_BLOCK) k_cmax_visited_mx(Geom G, MxState S, const float* __restrict__ P,
                                                               double inv_va2max, unsigned long long* cmax_out) {
  __shared__ double lds[MRM_BLOCK / 64];
  const long bs = G.bs;
  double cm[1] = {0.0};
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
    int i = (int)(n % G.n0);
    long r = n / G.n0;
    int j = (int)(r % G.n1), k = (int)(r / G.n1);
    bool ii = i >= G.lb0 && i <= G.le0, ji = j >= G.lb1 && j <= G.le1, ki = k >= G.lb2 && k <= G.le2;
    bool vis = (ii && ji && ki) || (!ii && ji && ki) || (ii && !ji && ki) || (ii && ji && !ki);
    if (vis)
      cm[0] = smax(cm[0], mx_cmax(S.f[n], S.f[bs + n], S.f[2 * bs + n], S.f[3 * bs + n], S.f[4 * bs + n],
                                  S.f[5 * bs + n], S.f[6 * bs + n], P[n], inv_va2max));
  }
  block_reduce<OpMax, 1>(cm, lds);
  if (threadIdx.x == 0) atomic_max_nonneg(cmax_out, cm[0]);
}

// FP64 SoA <-> mixed layout
__global__ void __launch_bounds__(MRM_BLOCK) k_to_mixed(long bs, const double* __restrict__ U, MxState S) {
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) {
      if (q == 4) S.e[n] = U[4 * bs + n];
      else S.f[mx_slot(q) * bs + n] = (float)U[q * bs + n];
    }
  }
}
__global__ void __launch_bounds__(MRM_BLOCK) k_from_mixed(long bs, MxState S, double* __restrict__ U) {
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) {
      if (q == 4) U[4 * bs + n] = S.e[n];
      else U[q * bs + n] = (double)S.f[mx_slot(q) * bs + n];
    }
  }
}


// ------------------------------------------------------------------ TVD
// TVDlimit (tvdlimit_SR.C:33-695) on the mixed state: the limiter in FP32,
// the specific internal energy recovered in FP64 and the energy update
// accumulated in FP64.
struct MxTvdPt {
  float v[MRM_NV];
  float lnr, vsqr, bsqr, cm, hyp_e, hyp_v, hyp_B, lf, CZ, hd;   // lf: Alfven limiter factor (BC2 = 1 - lf)
};

// per-point quantities (tvdlimit_SR.C:269-385)
__device__ __forceinline__ void mx_tvd_point(const MxTvdParams& P, const float u[MRM_NV], double e, float pres,
                                             float tau, float hfb1, float hft1, bool xsweep, MxTvdPt& o) {
  float v0 = fsmax(P.rho_min, u[0]);
  o.v[5] = u[5]; o.v[6] = u[6]; o.v[7] = u[7];
  o.CZ = (tau > 1.0e-5f) ? 1.0f : 0.0f;
  float dn = frcp(v0);
  float v1 = u[1] * dn, v2 = u[2] * dn, v3 = u[3] * dn;
  float vv = v1 * v1 + v2 * v2 + v3 * v3;
  float bb = o.v[5] * o.v[5] + o.v[6] * o.v[6] + o.v[7] * o.v[7];
  // specific internal energy e/rho - v^2/2 in FP64
  double dd = (double)v0;
  double mm = (double)u[1] * (double)u[1] + (double)u[2] * (double)u[2] + (double)u[3] * (double)u[3];
  const double rd = drcp(dd);
  double eps = (e - 0.5 * mm * rd) * rd;
  o.vsqr = vv;
  o.bsqr = bb;
  vv = fsqrt(vv);
  float va2 = bb * dn;
  float x2 = va2 * P.inv_va2max;
  float x4 = x2 * x2;
  float s = 1.0f + x2;
  float lf = fdiv(s, s + x4);
  float cs2 = 5.0f / 3.0f * pres * dn;
  float cfast = fsqrt(va2 + cs2);
  float CME_mode = (1.0f - o.CZ) * ((vv >= P.CME_thresh * cfast) ? 1.0f : 0.0f);
  dn = v0 * v0;
  float rf = fsmin(lf, fdiv(dn, dn + P.rho_lev));
  float hh = 1.0f - hfb1 - hft1;
  float cf = hfb1 * P.tvd_cs[0] + (rf * P.tvd_cs[1] + (1.0f - rf) * P.tvd_cs[2]) * hh + hft1 * P.tvd_cs[3];
  cf = fsmax(cf, CME_mode);
  cs2 *= cf * cf;
  o.cm = vv + fsqrt(fsmax(cs2, lf * (cs2 + va2)));
  o.hyp_e = hfb1 * P.tvd_h[0] + (rf * P.tvd_h[1] + (1.0f - rf) * P.tvd_h[2]) * hh + hft1 * P.tvd_h[3];
  o.hyp_v = hfb1 * P.tvd_h[0] + (rf * P.tvd_h[1] + (1.0f - rf) * P.tvd_h[2] * P.tvd_pm_v) * hh +
            hft1 * P.tvd_h[3] * P.tvd_pm_v;
  o.hyp_B = hfb1 * P.tvd_h[0] + (rf * P.tvd_h[1] + (1.0f - rf) * P.tvd_h[2] * P.tvd_pm_B) * hh +
            hft1 * P.tvd_h[3] * P.tvd_pm_B;
  o.hyp_e *= (1.0f - CME_mode);
  o.hyp_v *= (1.0f - CME_mode);
  o.lf = lf;
  o.hd = 0.0f;
  if (xsweep && (P.vhyp > 0.0f)) {
    float h = P.vhyp * fsqrt(o.vsqr);
    o.hd = 0.5f * P.idxd1 * fsmax(0.0f, fsmin(h, P.cmax - o.cm));
  }
  // the logarithms within half an ulp (mlogf): the slopes of ln rho and
  // ln eps are differences of values of 10 to 45, so a few ulps of error in
  // each log are a large error of the slope
  float v4;
  if (P.rho_log) {
    v0 = mrm_math::mlogf(v0);
    v4 = mrm_math::mlogf((float)eps);
    o.lnr = v0;
  } else {
    v4 = (float)eps;
    o.lnr = mrm_math::mlogf(v0);
  }
  o.v[0] = v0; o.v[1] = v1; o.v[2] = v2; o.v[3] = v3; o.v[4] = v4;
}

// limited slope at point i (tvdlimit_SR.C:388-434)
__device__ __forceinline__ void mx_tvd_slope(const MxTvdParams& P, const MxTvdPt& a, const MxTvdPt& b,
                                             const MxTvdPt& c, float hfb1, float hft1, float S[MRM_NV]) {
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    float sl = b.v[q] - a.v[q];
    float sr = c.v[q] - b.v[q];
    float slm = 0.25f * (sl + sr);
    float lower = fsmin(fsmin(sl, sr), slm);
    float upper = fsmax(fsmax(sl, sr), slm);
    S[q] = fsmax(0.0f, lower) + fsmin(0.0f, upper);
  }
  if ((P.visc_slope_bot > 0.0f) || (P.visc_slope_top > 0.0f)) {
    float cf = (1.0f - P.visc_slope_bot * hfb1) * (1.0f - P.visc_slope_top * hft1);
    S[1] *= cf; S[2] *= cf; S[3] *= cf;
  }
  if ((P.eta_slope_bot > 0.0f) || (P.eta_slope_top > 0.0f)) {
    float cf = (1.0f - P.eta_slope_bot * hfb1) * (1.0f - P.eta_slope_top * hft1);
    S[5] *= cf; S[6] *= cf; S[7] *= cf;
  }
  float qrho = fabsf(b.lnr - c.lnr);
  if (qrho > P.q_rho_max) S[0] = 0.0f;
  if (P.vmax_lim > 0.0f) {
    if (b.vsqr > P.vsqr_diff) { S[1] = 0.0f; S[2] = 0.0f; S[3] = 0.0f; }
  }
}

// diffusive flux at the interface between b (i) and c (i+1), a = i-1,
// d = i+2 (tvdlimit_SR.C:436-563); returns the resistive interface term
// (:609-614).  The ratio |sl_lim|/|sl| is capped at 1e30 so that a zero
// slope gives a zero flux in FP32 as it does in FP64.
template <int D>
__device__ __forceinline__ float mx_tvd_iface(const MxTvdParams& P, const MxTvdPt& a, const MxTvdPt& b,
                                              const MxTvdPt& c, const MxTvdPt& d, const float Sb[MRM_NV],
                                              const float Sc[MRM_NV], float hfb1, float hft1, float qft1,
                                              bool hyper_on, bool zero_bflux, float F[MRM_NV]) {
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    float sl = c.v[q] - b.v[q];
    float sl_lim = c.v[q] - b.v[q] - (Sb[q] + Sc[q]);
    float dn = (q >= 1 && q <= 3) ? b.hyp_v : (q >= 5 ? b.hyp_B : b.hyp_e);
    float rf = fminf(fabsf(sl_lim) * frcp(fsmax(1e-30f, fabsf(sl))), 1e30f);
    float cf = fsmax(0.0f, fsmin(sl_lim, sl)) + fsmin(0.0f, fsmax(sl_lim, sl));
    F[q] = fsmax(0.0f, 1.0f + dn * (rf - 1.0f)) * cf;
  }
  float tvd_faci = 0.5f * P.idxd1 * fsmin(P.cmax, fsmax(b.cm, c.cm));
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) F[q] *= tvd_faci;
  F[5 + D] *= P.B_par_diff;
  if (P.need_tvd_coeff) {
    F[0] *= P.tc[0];
    float cf = fsmax(P.tc[1], P.visc_coeff_bot * hfb1 + P.visc_coeff_top * hft1);
    F[1] *= cf; F[2] *= cf; F[3] *= cf;
    F[4] *= P.tc[2];
    cf = fsmax(P.tc[3], P.eta_coeff_bot * hfb1 + P.eta_coeff_top * hft1);
    F[5] *= cf; F[6] *= cf; F[7] *= cf;
  }
  if (D == 0 && hyper_on) {
    float dn = fsmax(b.hd, c.hd);
    F[0] -= dn * (0.125f * (d.v[0] - a.v[0]) - 0.375f * (c.v[0] - b.v[0]));
    F[1] -= dn * (0.125f * (d.v[1] - a.v[1]) - 0.375f * (c.v[1] - b.v[1]));
    F[4] -= dn * (0.125f * (d.v[4] - a.v[4]) - 0.375f * (c.v[4] - b.v[4]));
  }
  float flx0 = F[0], flx1 = F[1], flx2 = F[2], flx3 = F[3], flx4 = F[4];
  float uif0 = 0.5f * (b.v[0] + Sb[0] + c.v[0] - Sc[0]);
  float uif1 = 0.5f * (b.v[1] + Sb[1] + c.v[1] - Sc[1]);
  float uif2 = 0.5f * (b.v[2] + Sb[2] + c.v[2] - Sc[2]);
  float uif3 = 0.5f * (b.v[3] + Sb[3] + c.v[3] - Sc[3]);
  float uif4 = 0.5f * (b.v[4] + Sb[4] + c.v[4] - Sc[4]);
  if (P.rho_log) {
    uif0 = expf(uif0);
    flx0 *= uif0;
    uif4 = expf(uif4);
    flx4 *= uif4;
  }
  float min_cz = flx0 * fsmin(b.CZ, c.CZ);
  flx1 *= uif0; flx1 += uif1 * min_cz;
  flx2 *= uif0; flx2 += uif2 * min_cz;
  flx3 *= uif0; flx3 += uif3 * min_cz;
  flx4 *= uif0;
  flx4 += (uif4 * min_cz) + qft1 * (uif1 * flx1 + uif2 * flx2 + uif3 * flx3);
  F[0] = flx0; F[1] = flx1; F[2] = flx2; F[3] = flx3; F[4] = flx4;
  if (zero_bflux) F[5] = 0.0f;
  return (c.v[5] - b.v[5]) * F[5] + (c.v[6] - b.v[6]) * F[6] + (c.v[7] - b.v[7]) * F[7];
}

// in-place update of point i (tvdlimit_SR.C:565-657); u[4] unused, the energy
// is accumulated in e (FP64) and clamped there
__device__ __forceinline__ void mx_tvd_update(const MxTvdParams& P, const MxTvdPt& b, const float Fm[MRM_NV],
                                              const float F[MRM_NV], float qm, float q, float qft1,
                                              float u[MRM_NV], double& e) {
  float r[MRM_NV];
#pragma unroll
  for (int k = 0; k < MRM_NV; k++) r[k] = P.dt_tvd * (F[k] - Fm[k]);
  float v1 = b.v[1], v2 = b.v[2], v3 = b.v[3], b5 = b.v[5], b6 = b.v[6], b7 = b.v[7];
  // kinetic-energy term (1-qft) r.v and Boris correction r -= (1-lf) r_perp
  // (tvdlimit_SR.C:620-640), written as r_par + lf r_perp and
  // (1-qft) dn (b.v) + (lf - qft) r_perp.v: no cancellation where the
  // limiter is strong
  const float lf = b.lf;
  float dn = (r[1] * b5 + r[2] * b6 + r[3] * b7) * frcp(fsmax(1e-30f, b.bsqr));
  float q1 = r[1] - dn * b5, q2 = r[2] - dn * b6, q3 = r[3] - dn * b7;
  r[4] += (1.0f - qft1) * dn * (b5 * v1 + b6 * v2 + b7 * v3) + (lf - qft1) * (q1 * v1 + q2 * v2 + q3 * v3);
  r[1] = dn * b5 + lf * q1;
  r[2] = dn * b6 + lf * q2;
  r[3] = dn * b7 + lf * q3;
  float qres = 0.5f * (q + qm) * qft1;
  r[4] += P.dt_tvd * qres;
#pragma unroll
  for (int k = 0; k < MRM_NV; k++)
    if (k != 4) u[k] += r[k];
  e += (double)r[4];
  u[0] = fsmax(P.rho_min, u[0]);
  double d = u[0];
  e = smax(e, P.eps_min * d + 0.5 * ((double)u[1] * u[1] + (double)u[2] * u[2] + (double)u[3] * u[3]) * drcp(d));
}

__device__ __forceinline__ void mx_load_cons(const MxState& S, long bs, unsigned n, float u[MRM_NV], double& e) {
  u[0] = ld32(S.f, n); u[1] = ld32(S.f + bs, n); u[2] = ld32(S.f + 2 * bs, n); u[3] = ld32(S.f + 3 * bs, n);
  u[4] = 0.0f;
  u[5] = ld32(S.f + 4 * bs, n); u[6] = ld32(S.f + 5 * bs, n); u[7] = ld32(S.f + 6 * bs, n);
  e = ld32(S.e, n);
}
__device__ __forceinline__ void mx_store_cons(const MxState& S, long bs, unsigned n, const float u[MRM_NV], double e) {
  st32(S.f, n, u[0]); st32(S.f + bs, n, u[1]); st32(S.f + 2 * bs, n, u[2]); st32(S.f + 3 * bs, n, u[3]);
  st32(S.f + 4 * bs, n, u[5]); st32(S.f + 5 * bs, n, u[6]); st32(S.f + 6 * bs, n, u[7]);
  st32(S.e, n, e);
}

// one thread marches one grid line (dims 1 and 2, and dim 0 for long columns)
template <int D, bool HAS_TAU, bool EPI, bool VALF = false>
__global__ void __launch_bounds__(MRM_BLOCK) k_tvd_sweep_mx(Geom G, MxTvdParams P, MxState S,
                                                            const float* __restrict__ Pr, MxEpilogue X) {
  __shared__ double lds[(VALF ? 5 : 1) * MRM_BLOCK / 64];
  const long bs = G.bs;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long base = 0, str = 1;
  int sz = 0, gh = 0, lineX = 0;
  bool active;
  if (D == 0) {
    active = t < (long)G.n1 * G.n2;
    int j = (int)(t % G.n1), k = (int)(t / G.n1);
    base = j * G.s1 + k * G.s2; str = 1; sz = G.n0; gh = G.g0;
  } else {
    const int nx = G.le0 - G.lb0 + 1;
    const int nb = D == 1 ? P.y_nk : (G.le1 - G.lb1 + 1);
    active = t < (long)nx * nb;
    int i = G.lb0 + (int)(t % nx);
    int b = (int)(t / nx);
    lineX = i;
    if (D == 1) { base = i + (long)(P.y_k0 + b) * G.s2; str = G.s1; sz = G.n1; gh = G.g1; }
    else { base = i + (G.lb1 + b) * G.s1; str = G.s2; sz = G.n2; gh = G.g2; }
  }
  double cm[1] = {0.0};
  ValfAcc va;
  if (active) {
    const float lhfb = D == 0 ? 0.0f : (float)P.hfb[lineX];
    const float lhft = D == 0 ? 0.0f : (float)P.hft[lineX];
    const float lqft = D == 0 ? 0.0f : (float)P.qft[lineX];
    const int istart = 1 + (P.is_gbeg0 ? gh : 0);
    MxTvdPt W0, W1, W2, W3;
    float S0[MRM_NV], S1[MRM_NV], Fm[MRM_NV], F[MRM_NV];
    float qm = 0.0f;
    // the stored values of points s-2, s-1 stay in registers for the update
    // of point s-2; point s+1 is loaded while point s is processed
    float u1[MRM_NV], u2[MRM_NV], un[MRM_NV];
    double e1 = 0.0, e2 = 0.0, en;
    float pn_ = 0.0f, tn = 0.0f;
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) { u1[q] = 0.0f; u2[q] = 0.0f; }
    {
      const unsigned n0_ = (unsigned)base;
      mx_load_cons(S, bs, n0_, un, en);
      pn_ = ld32(Pr, n0_);
      tn = HAS_TAU ? (float)ld32(P.tau, n0_) : 0.0f;
    }
    for (int s = 0; s < sz; s++) {
      float u[MRM_NV];
      double e = en;
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) u[q] = un[q];
      const float pres = pn_, tau = tn;
      if (s + 1 < sz) {
        const unsigned nn = (unsigned)base + (unsigned)(s + 1) * (unsigned)str;
        mx_load_cons(S, bs, nn, un, en);
        pn_ = ld32(Pr, nn);
        if (HAS_TAU) tn = (float)ld32(P.tau, nn);
      }
      float hfb1 = D == 0 ? (float)P.hfb[s] : lhfb, hft1 = D == 0 ? (float)P.hft[s] : lhft;
      mx_tvd_point(P, u, e, pres, tau, hfb1, hft1, D == 0, W3);
      if (s >= 2) {
        const int ip = s - 1;
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) S0[q] = S1[q];
        mx_tvd_slope(P, W1, W2, W3, D == 0 ? (float)P.hfb[ip] : lhfb, D == 0 ? (float)P.hft[ip] : lhft, S1);
      }
      if (s >= 3) {
        const int ii = s - 2;
        const float ih = D == 0 ? (float)P.hfb[ii] : lhfb, it = D == 0 ? (float)P.hft[ii] : lhft;
        const float iq = D == 0 ? (float)P.qft[ii] : lqft;
        bool zb = (D == 0) && ((P.is_gend0 && ii == sz - gh - 1) || (P.is_gbeg0 && ii == gh - 1));
        float q = mx_tvd_iface<D>(P, W0, W1, W2, W3, S0, S1, ih, it, iq, ii >= istart, zb, F);
        if (ii >= gh && ii <= sz - gh - 1) {
          const unsigned nd = (unsigned)base + (unsigned)ii * (unsigned)str;
          float uu[MRM_NV];
          double ee = e2;
#pragma unroll
          for (int k = 0; k < MRM_NV; k++) uu[k] = u2[k];
          mx_tvd_update(P, W1, Fm, F, qm, q, iq, uu, ee);
          if (EPI) {
            mx_tcheck(uu, ee, X.rho_min, X.eps_min, X.eps_max, X.vmax);
            float pn = mx_pres(X.E, uu[0], uu[1], uu[2], uu[3], ee);
            st32(X.Pnew, nd, pn);
            cm[0] = smax(cm[0], mx_cmax(uu[0], uu[1], uu[2], uu[3], uu[5], uu[6], uu[7], pn, X.inv_va2max));
            if (VALF) mx_valf_cell(va, uu[0], uu[1], uu[2], uu[3], ee, pn, X.vlim95, X.elim95);
          }
          mx_store_cons(S, bs, nd, uu, ee);
        }
#pragma unroll
        for (int k = 0; k < MRM_NV; k++) Fm[k] = F[k];
        qm = q;
      }
      W0 = W1; W1 = W2; W2 = W3;
#pragma unroll
      for (int k = 0; k < MRM_NV; k++) { u2[k] = u1[k]; u1[k] = u[k]; }
      e2 = e1;
      e1 = e;
    }
  }
  if (EPI) {
    block_reduce<OpMax, 1>(cm, lds);
    if (threadIdx.x == 0) atomic_max_nonneg(X.cmax, cm[0]);
    if (VALF) {
      __syncthreads();
      valf_flush(va, lds, X.valf);
    }
  }
}

// one block per dim-0 grid line, phases separated by barriers, line data in LDS
__global__ void __launch_bounds__(TX_MAXN) k_tvd_x_lines_mx(Geom G, MxTvdParams P, MxState S,
                                                            const float* __restrict__ Pr) {
  extern __shared__ float smf[];
  const long bs = G.bs;
  const int sz = G.n0, gh = G.g0;
  const long nlines = (long)G.n1 * G.n2;
  const long line = xcd_remap(blockIdx.x, nlines);
  const int j = (int)(line % G.n1), k = (int)(line / G.n1);
  const unsigned base = (unsigned)j * (unsigned)G.s1 + (unsigned)k * (unsigned)G.s2;
  float* sv = smf;                     // [8][sz]
  float* sS = sv + MRM_NV * sz;        // [8][sz]
  float* sF = sS + MRM_NV * sz;        // [8][sz]
  float* scm = sF + MRM_NV * sz;
  float* shd = scm + sz;
  float* scz = shd + sz;
  float* slr = scz + sz;
  float* sq = slr + sz;
  const int t = threadIdx.x;
  MxTvdPt me;
  float u[MRM_NV];
  double e = 0.0;
  const float hfb = t < sz ? (float)P.hfb[t] : 0.0f, hft = t < sz ? (float)P.hft[t] : 0.0f;
  if (t < sz) {
    mx_load_cons(S, bs, base + t, u, e);
    float tau = P.tau ? (float)ld32(P.tau, base + t) : 0.0f;
    mx_tvd_point(P, u, e, ld32(Pr, base + t), tau, hfb, hft, true, me);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) sv[q * sz + t] = me.v[q];
    scm[t] = me.cm;
    shd[t] = me.hd;
    scz[t] = me.CZ;
    slr[t] = me.lnr;
  }
  __syncthreads();
  float Sl[MRM_NV];
  if (t >= 1 && t <= sz - 2) {
    MxTvdPt a, c;
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) { a.v[q] = sv[q * sz + t - 1]; c.v[q] = sv[q * sz + t + 1]; }
    c.lnr = slr[t + 1];
    mx_tvd_slope(P, a, me, c, hfb, hft, Sl);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) sS[q * sz + t] = Sl[q];
  }
  __syncthreads();
  float F[MRM_NV];
  const int istart = 1 + (P.is_gbeg0 ? gh : 0);
  if (t >= 1 && t <= sz - 3) {
    MxTvdPt a, c, d;
    float Sc[MRM_NV];
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) {
      a.v[q] = sv[q * sz + t - 1];
      c.v[q] = sv[q * sz + t + 1];
      d.v[q] = sv[q * sz + t + 2];
      Sc[q] = sS[q * sz + t + 1];
    }
    c.cm = scm[t + 1];
    c.hd = shd[t + 1];
    c.CZ = scz[t + 1];
    bool zb = (P.is_gend0 && t == sz - gh - 1) || (P.is_gbeg0 && t == gh - 1);
    float q = mx_tvd_iface<0>(P, a, me, c, d, Sl, Sc, hfb, hft, (float)P.qft[t], t >= istart, zb, F);
#pragma unroll
    for (int qq = 0; qq < MRM_NV; qq++) sF[qq * sz + t] = F[qq];
    sq[t] = q;
  }
  __syncthreads();
  if (t >= gh && t <= sz - gh - 1) {
    float Fm[MRM_NV];
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) Fm[q] = sF[q * sz + t - 1];
    mx_tvd_update(P, me, Fm, F, sq[t - 1], sq[t], (float)P.qft[t], u, e);
    mx_store_cons(S, bs, base + t, u, e);
  }
}

// ------------------------------------------------------------------ Clean_divB
// Clean_divB (clean_divB.C:35-369) on the mixed state: B is stored in FP32,
// the divergence, the potential phi (FP64 field) and the corrections are
// formed in FP64 from the widened values (k_divb / k_divb_update's
// arithmetic), B is rounded back to FP32 and e takes phi divB in FP64.  The
// last correction pass can finish the cells as the mixed TVD sweep would
// (TCheck, FP32 pressure, cmax, Adjust_Valf_Max statistics).
template <bool PHI>
__global__ void __launch_bounds__(DIVB_TX * DIVB_TY) k_divb_mx(DivbArgs A, MxState S, double* __restrict__ divB,
                                                               double* __restrict__ phi, unsigned long long* err) {
  __shared__ double lds[DIVB_TX * DIVB_TY / 64];
  const Geom& G = A.G;
  const float* Bx = S.f + 4 * G.bs;
  const float* By = S.f + 5 * G.bs;
  const float* Bz = S.f + 6 * G.bs;
  const long s1 = G.s1, s2 = G.s2;
  double m[1] = {0.0};
  const DivbTile t = divb_tile(G);
  if (t.ok) {
    double zw[5];
#pragma unroll
    for (int q = 0; q < 4; q++) zw[q] = (double)Bz[t.c0 + (long)(t.k0 - 2 + q) * s2];
    for (int k = t.k0; k < t.k1; k++) {
      const long n = t.c0 + (long)k * s2;
      zw[4] = (double)Bz[n + 2 * s2];
      double d = 0.0;
      d += (A.w1[0] * ((double)Bx[n + 1] - (double)Bx[n - 1]) + A.w2[0] * ((double)Bx[n + 2] - (double)Bx[n - 2]));
      d += (A.w1[1] * ((double)By[n + s1] - (double)By[n - s1]) +
            A.w2[1] * ((double)By[n + 2 * s1] - (double)By[n - 2 * s1]));
      d += (A.w1[2] * (zw[3] - zw[1]) + A.w2[2] * (zw[4] - zw[0]));
      divB[n] = d;
      m[0] = fmax(m[0], fabs(d));
      if (PHI) phi[n] += A.fac * d - A.damp * phi[n];
#pragma unroll
      for (int q = 0; q < 4; q++) zw[q] = zw[q + 1];
    }
  }
  block_reduce<OpMax, 1>(m, lds);
  if (threadIdx.x == 0) atomic_max_nonneg(err, m[0]);
}

template <bool EPI>
__global__ void __launch_bounds__(DIVB_TX * DIVB_TY) k_divb_update_mx(DivbArgs A, MxState S,
                                                                      const double* __restrict__ phi,
                                                                      const double* __restrict__ divB,
                                                                      MxEpilogue X) {
  __shared__ double lds[5 * DIVB_TX * DIVB_TY / 64];
  const Geom& G = A.G;
  const long bs = G.bs, s1 = G.s1, s2 = G.s2;
  float* Bx = S.f + 4 * bs;
  float* By = S.f + 5 * bs;
  float* Bz = S.f + 6 * bs;
  double cm[1] = {0.0};
  ValfAcc va;
  const DivbTile t = divb_tile(G);
  if (t.ok) {
    double pw[5];
#pragma unroll
    for (int q = 0; q < 4; q++) pw[q] = phi[t.c0 + (long)(t.k0 - 2 + q) * s2];
    for (int k = t.k0; k < t.k1; k++) {
      const long n = t.c0 + (long)k * s2;
      pw[4] = phi[n + 2 * s2];
      Bx[n] = (float)((double)Bx[n] + (A.w1[0] * (phi[n + 1] - phi[n - 1]) + A.w2[0] * (phi[n + 2] - phi[n - 2])));
      By[n] = (float)((double)By[n] +
                      (A.w1[1] * (phi[n + s1] - phi[n - s1]) + A.w2[1] * (phi[n + 2 * s1] - phi[n - 2 * s1])));
      Bz[n] = (float)((double)Bz[n] + (A.w1[2] * (pw[3] - pw[1]) + A.w2[2] * (pw[4] - pw[0])));
      double e = S.e[n] + pw[2] * divB[n];
      if (EPI) {
        float u[MRM_NV];
        double ee;
        mx_load_cons(S, bs, (unsigned)n, u, ee);
        mx_tcheck(u, e, X.rho_min, X.eps_min, X.eps_max, X.vmax);
        st32(S.f, (unsigned)n, u[0]);
        st32(S.f + bs, (unsigned)n, u[1]);
        st32(S.f + 2 * bs, (unsigned)n, u[2]);
        st32(S.f + 3 * bs, (unsigned)n, u[3]);
        const float pn = mx_pres(X.E, u[0], u[1], u[2], u[3], e);
        st32(X.Pnew, (unsigned)n, pn);
        cm[0] = smax(cm[0], mx_cmax(u[0], u[1], u[2], u[3], u[5], u[6], u[7], pn, X.inv_va2max));
        if (X.valf) mx_valf_cell(va, u[0], u[1], u[2], u[3], e, pn, X.vlim95, X.elim95);
      }
      S.e[n] = e;
#pragma unroll
      for (int q = 0; q < 4; q++) pw[q] = pw[q + 1];
    }
  }
  if (EPI) {
    block_reduce<OpMax, 1>(cm, lds);
    if (threadIdx.x == 0) atomic_max_nonneg(X.cmax, cm[0]);
    if (X.valf) {
      __syncthreads();
      valf_flush(va, lds, X.valf);
    }
  }
}

// ------------------------------------------------------------------ Spitzer conduction
// The conduction kernels of the fused step on the mixed state: the
// temperature from the FP32 lookup, the conduction terms in FP64 from the
// widened values (sflx, its update and the energy term stay FP64).
struct SptStateMx {
  const float* f;
  const double* e;
  long bs;
  __device__ __forceinline__ double rho(long n) const { return (double)f[n]; }
  __device__ __forceinline__ double mom(int q, long n) const { return (double)f[(1 + q) * bs + n]; }
  __device__ __forceinline__ double ene(long n) const { return e[n]; }
  __device__ __forceinline__ double mag(int q, long n) const { return (double)f[(4 + q) * bs + n]; }
};

__global__ void __launch_bounds__(MRM_BLOCK) k_eos_temp_mx(Geom G, EosDev E, MxState S, double* __restrict__ T) {
  const long bs = G.bs;
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
    double t;
    mx_pres<true>(E, S.f[n], S.f[bs + n], S.f[2 * bs + n], S.f[3 * bs + n], S.e[n], &t);
    T[n] = t;
  }
}

}  // namespace MRM_NS
