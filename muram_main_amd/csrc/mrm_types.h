// mrm_types.h -- parameter blocks shared by the host code and the kernels of
// the MHD step (plain structs passed by value to kernels).
#pragma once
#include "../../include/muram_step.h"

#define MRM_NV 8
#define RED_CMAX 0
#define RED_VMAX 1
#define RED_CS2MAX 2
#define RED_EMAX 3
#define RED_CNT_V 4
#define RED_CNT_E 5
#define RED_DIVB0 6    // max |div B| before and after Clean_divB's passes
#define RED_DIVB1 7
#define RED_SUM0 8
#define RED_SCRATCH 15 // reductions nobody reads
#define RED_NSLOT 16
#define MRM_BLOCK 256
#ifndef ST_SX
#define ST_SX 32           // tile of the FP64 stage kernels: dim 0 x dim 1, 256 threads
#define ST_SY 8
#endif
#ifndef MX_SX
#define MX_SX ST_SX        // tile of the FP32-mixed stage kernel
#define MX_SY ST_SY
#endif
#define MX_LX (MX_SX + 4)
#define MX_LY (MX_SY + 4)
#ifndef ST_KC
#define ST_KC 64           // dim-2 planes per block of the FP64 stage kernels
#endif
#ifndef MX_KC
#define MX_KC 32           // dim-2 planes per block of the FP32-mixed stage kernel
#endif
#define ST_LX (ST_SX + 4)
#define ST_LY (ST_SY + 4)
#define ST_NF 9            // rho, v0, v1, v2, e_int, B0, B1, B2, p
#define TX_MAXN 1024

struct Geom {
  int n0, n1, n2;          // buffer extents (gsize + 2*ghosts)
  int g0, g1, g2;          // ghosts
  int lb0, lb1, lb2;       // lbeg
  int le0, le1, le2;       // lend
  long s1, s2;             // strides of dims 1, 2 (stride of dim 0 is 1)
  long bs;                 // bufsize
};

struct EosDev {
  int N_eps, N_lr, N_lp3, N_s3;
  double inv_del_eps, inv_del_lr, inv_del_lp3, inv_del_s3;
  double eps_off, ss_off, eps_max;
  // the axes' first nodes xeps[0], xlr[0], xlp3[0], xs3[0]: kernel arguments,
  // so that a lookup's index needs no memory read before its table reads
  double xeps0, xlr0, xlp30, xs30;
  const double* xeps;
  const double* xlr;
  const double* xlp3;
  const double* xs3;
  const double* rxeps;  // exp(-xeps[i]) and exp(-xlr[j]): offsets of a point from its
  const double* rxlr;   // table node as logs of ratios near 1 (FP32-mixed lookup)
  const float* tab;    // [6][N_eps][N_lr]: ln p, ln T, s, ln ne, ln rhoi, ln amb (pre-divided)
  const float* itab;   // [2][N_lp3][N_s3]: ln(eps+eps_off), ln rho (pre-divided)
};

struct SrcParams {
  double wdt, invRdt, g, inv_va2max;
  double rho_min, eps_min, eps_max, vmax;
  // damping profile dmp[i] of the step (Add_Sources_Integrate_SR.C:92-108),
  // n0 doubles, or null with the damping off
  const double* dmp = nullptr;
};

struct TvdParams {
  double rho_min, eps_min, rho_lev, q_rho_max;
  double visc_coeff_bot, visc_slope_bot, visc_coeff_top, visc_slope_top;
  double eta_coeff_bot, eta_slope_bot, eta_coeff_top, eta_slope_top;
  double B_par_diff, vhyp, tvd_pm_v, tvd_pm_B, vmax_lim, CME_thresh;
  double tvd_h[4], tvd_cs[4], tc[4];
  double inv_va2max, vsqr_diff, cmax, idxd1, dt_tvd;
  int need_tvd_coeff, rho_log, is_gbeg0, is_gend0;
  int y_k0, y_nk;      // planes k0 .. k0 + nk - 1 of the dim-1 sweep's lines
  const double* hfb;
  const double* hft;
  const double* qft;
  const double* tau;   // NULL: Tau == 0 everywhere
  unsigned long long* bcount;   // limiter branch slots (counting instantiations), MRM_BR_NSLOT
  double btie;                  // near-tie margin relative to the operand scale
};

struct ResidParams {
  double w1[3], w2[3];
  double inv_va2max;
  // resistivity (param_eta > 0): eta and the 4th-order second-derivative
  // weights -30, 16, -1 over 12 dx^2 per dimension (mhdres_SR.C:205-207)
  double eta = 0.0;
  double ww0[3] = {0, 0, 0}, ww1[3] = {0, 0, 0}, ww2[3] = {0, 0, 0};
};

// diagnostics arrays tvar1..8, Qres, Qvis (grid.H:69-100): array t at
// dg + t * bs.  idt = 1 / Run.dt, dt_fac = dt_tvd / Run.dt (tvdlimit_SR.C:115-116)
#define DG_T1 0
#define DG_T2 1
#define DG_T3 2
#define DG_T4 3
#define DG_T5 4
#define DG_T6 5
#define DG_T7 6
#define DG_T8 7
#define DG_QRES 8
#define DG_QVIS 9
#define DG_N 10
struct DiagArgs {
  double* dg = nullptr;
  double idt = 0.0, dt_fac = 0.0;
};

// Spitzer conduction (mhdres_SR.C:37-49, 136-145)
struct SptParams {
  double w1[3], w2[3];
  double k0;          // param_spitzer
  double sat_flx;     // (1 + X_H) Rgas^1.5 sqrt(1836) / 8 (mhdres_SR.C:46)
  double hyp_diff, cmax_hyp, nu_hyp_max;   // from the call's dt and CFL (mhdres_SR.C:136-145)
};

// Clean_divB (clean_divB.C:35-369)
struct DivbArgs {
  Geom G;
  double w1[3], w2[3];     // 8/(12 dx), -1/(12 dx) per dimension (grid.C:339-340)
  double fac, damp;        // divB_fac, divB_damp (clean_divB.C:47-48, 110)
  double* qres = nullptr;  // diagnostics: Qres += phi divB / dt (clean_divB.C:264-270)
  double invdt = 0.0;
};

struct BcParams {
  double p_bc, s_bc, B_crit, eps_top, cc;
  int open_top;
  // potential-field top boundary: b_ext (8 planes of bplane = n1 n2 doubles,
  // boundary_pdmp_1_fftw3.C:422-436), or null for the vertical field
  const double* bext = nullptr;
  long bplane = 0;
};

struct TvdEpilogue {
  EosDev E;
  double rho_min, eps_min, eps_max, vmax, inv_va2max;
  double* Pnew;
  unsigned long long* cmax;
  // Adjust_Valf_Max statistics of the finished cells (va_adjust > 0): the
  // reduction slots red[RED_VMAX..RED_CNT_E], or null; 0.95 v_lim, 0.95 e_lim
  unsigned long long* valf;
  double vlim95, elim95;
  // the top interior layer (i = lend) stays out of cmax: the potential-field
  // boundary rewrites its B_y, B_z afterwards and reduces its speed then
  int skip_top_cmax = 0;
};

struct StageArgs {
  Geom G;
  EosDev E;
  ResidParams RP;
  SrcParams SP;
  int ntx, nty, nkc;       // tiles in dim 0, dim 1, chunks in dim 2
  int write_p;             // compute the EOS pressure of the new state
  const double* eres = nullptr;   // extra energy residual per cell (conduction), or null
};

struct TvdLineArgs {
  Geom G;
  TvdParams P;
};

// ---- FP32-mixed form: rho, M, B and p in FP32, the total energy in FP64
struct MxState {
  float* f;    // [7][bs]: rho, M0, M1, M2, B0, B1, B2
  double* e;   // [bs]
};

struct MxStageArgs {
  Geom G;
  EosDev E;
  float w1[3], w2[3];
  float inv_va2max;
  double wdt;
  float invRdt, g;
  float rho_min, vmax;
  double eps_min, eps_max;
  int ntx, nty, nkc;
  int write_p;
  const double* eres = nullptr;   // extra energy residual per cell (conduction), or null
  // resistivity (param_eta > 0): eta and the second-derivative weights
  float eta = 0.0f;
  float ww0[3] = {0, 0, 0}, ww1[3] = {0, 0, 0}, ww2[3] = {0, 0, 0};
};

struct MxTvdParams {
  float rho_min, rho_lev, q_rho_max;
  float visc_coeff_bot, visc_slope_bot, visc_coeff_top, visc_slope_top;
  float eta_coeff_bot, eta_slope_bot, eta_coeff_top, eta_slope_top;
  float B_par_diff, vhyp, tvd_pm_v, tvd_pm_B, vmax_lim, CME_thresh;
  float tvd_h[4], tvd_cs[4], tc[4];
  float inv_va2max, vsqr_diff, cmax, idxd1, dt_tvd;
  double eps_min;
  int need_tvd_coeff, rho_log, is_gbeg0, is_gend0;
  int y_k0, y_nk;
  const double* hfb;
  const double* hft;
  const double* qft;
  const double* tau;
};

struct MxEpilogue {
  EosDev E;
  float rho_min, vmax;
  double eps_min, eps_max, inv_va2max;
  float* Pnew;
  unsigned long long* cmax;
  unsigned long long* valf;   // Adjust_Valf_Max statistics (va_adjust > 0) or null
  double vlim95, elim95;
};
