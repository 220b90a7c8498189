/*
 * muram_oracle.c -- TEST INFRASTRUCTURE ONLY (parity checker and CPU baseline).
 *
 * Plain-C FP64 restatement of MURaM's per-timestep MHD update.  Each function
 * names the reference lines it follows; the arithmetic keeps the reference's
 * operation order (no FMA contraction: build with -ffp-contract=off) so that
 * results agree with the compiled reference bit for bit on the same libm.
 * std::max / std::min semantics (NaN handling included) are kept where the
 * reference uses them, fmax / fmin where it uses those.
 *
 * Covered besides the ideal-MHD step: Adjust_Valf_Max, Clean_divB,
 * resistivity, Spitzer conduction, restart of the state, and the
 * diagnostics arrays (tvar1..8, Qres, Qvis).  Out of scope: ambipolar
 * diffusion, damping, the potential-field top boundary.
 */
#include "muram_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* std::max(a,b) is (a<b)?b:a and std::min(a,b) is (b<a)?b:a */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }
static inline int imax(int a, int b) { return (a < b) ? b : a; }
static inline int imin(int a, int b) { return (b < a) ? b : a; }

#define NV 8
#define MAXSTAGE 4   /* src_int_tck.H:5 */

struct orc_ctx {
  int gsize[3], ghosts[3], periods[3];
  int lsize[3], lbeg[3], lend[3], n[3];
  long stride[3];
  long bufsize;
  int vsize;
  double dx[3], gxmin[3], gxmax[3], lxmin[3];
  mrm_phys_desc ph;

  /* EOS (eos_init.C:16-44) */
  int eos_ok;
  int N_eps, N_lr, N_lp3, N_s3;
  double del_eps, del_lr, del_lp3, del_s3;
  double inv_del_eps, inv_del_lr, inv_del_lp3, inv_del_s3;
  double eps_off, ss_off;
  double *xeps, *xlr, *xlp3, *xs3;
  float *tab[6];    /* p, T, s, ne, rhoi, amb: [i_eps*N_lr + i_rho] */
  float *itab[2];   /* eps3, d3: [i_lp*N_s3 + i_s] */

  double *U, *U0, *Res;  /* 8*bufsize each */
  double *fld[MRM_NFIELD];

  /* Adjust_Valf_Max.C:15,28,39 */
  double inv_va2max, v_lim, e_lim, va_max_old;
  int valf_ini;
  /* boundary_pdmp_1_fftw3.C:32,85 */
  double p_bc, s_bc;
  int bnd_ini;
  /* tvdlimit_SR.C:24-26,88 */
  double *hfb, *hft, *qft;
  int tvd_ini;
  /* stage-loop scalars */
  double last_cfl;
  /* Clean_divB's adaptive iteration count (clean_divB.C:50) */
  int divb_next_iter;
  /* GridData::curlB (grid.C:307-309), allocated with eta > 0 */
  double* curlB;
  /* Spitzer conduction (grid.C:311-318): sflx lives in fld[MRM_SFLX] */
  double *sflx0, *Rflx, *BgradT;
  /* diagnostics tvar1..8, Qres, Qvis (grid.C:319-329, grid.H:69-100), allocated at the
     first orc_diag_zero or first call with need_diagnostics */
  double* diag[10];
  /* potential-field top boundary (boundary_pdmp_1_fftw3.C:113-141, 335-442;
     potential_sd_fftw3.C:27-257): the 8 kernels (float, [gy][gx], file
     order z-1, z-2, x-0, x-1, x-2, y-0, y-1, y-2) and b_ext, 8 planes of
     n1 x n2 with periodic ghosts, computed once valid */
  float* pkern;
  double* bext;
  int bext_valid;
  /* damping (Add_Sources_Integrate_SR.C:92-108, 628-761): the vertical
     profile dmp (n0, zero outside the interior) formed at stage 1, and
     rho_max of the first call */
  double* dmp;
  double damp_rho_max;
  int damp_ini;
  /* limiter branch accounting (orc_branch_count): MRM_BR_NSLOT slots */
  int bcount_on;
  double btie;
  unsigned long long bcount[MRM_BR_NSLOT];
};

/* ---- limiter branch accounting: the same categories, position key and
   weights as the library's counters (muram_step.h, MRM_BR_*); a test of the
   checker, not of the reference, which keeps no such counts */
static inline unsigned long long br_weight(unsigned long long x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return (x ^ (x >> 31)) | 1ull;
}
typedef struct {
  unsigned long long n[MRM_BR_NCOUNT], h[MRM_BK_N], nnear[MRM_BK_N], nexact[MRM_BK_N];
  double tie;
} br_acc;
/* 0: decisive, 1: near tie (0 < margin < tie x scale), 2: exact tie (margin
   0); with tie = 0 every evaluation is decisive */
static inline int br_cls(const br_acc* a, double margin, double scale) {
  return !(a->tie > 0.0) ? 0 : margin == 0.0 ? 2 : (margin < a->tie * fabs(scale) ? 1 : 0);
}
/* ties are counted as such and nothing else */
static inline int br_tied(br_acc* a, int kind, int tc) {
  if (tc == 1) a->nnear[kind]++;
  if (tc == 2) a->nexact[kind]++;
  return tc != 0;
}
static inline void br_hit(br_acc* a, int c, int kind, unsigned long long key, int taken, int tc) {
  if (br_tied(a, kind, tc)) return;
  if (taken) { a->n[c]++; a->h[kind] += br_weight(key); }
}
static inline void br_cat(br_acc* a, int c0, int kind, unsigned long long key, int cat, int tc) {
  if (br_tied(a, kind, tc)) return;
  a->n[c0 + cat]++;
  a->h[kind] += br_weight(key) * (unsigned long long)(cat + 1);
}
static inline double fmax3(double a, double b, double c) { return fmax(fmax(a, b), c); }
static inline unsigned long long br_key(int d, long bs, long node) {
  return ((unsigned long long)d * (unsigned long long)bs + (unsigned long long)node) << 3;
}
void orc_branch_count(orc_ctx* c, int on, double tie) {
  if (on) memset(c->bcount, 0, sizeof c->bcount);
  c->bcount_on = on;
  c->btie = tie;
}
void orc_get_branch_counts(const orc_ctx* c, unsigned long long out[MRM_BR_NSLOT]) {
  memcpy(out, c->bcount, sizeof c->bcount);
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

long orc_bufsize(const orc_ctx* c) { return c->bufsize; }

/* GridData::Init on one rank (grid.C:147-348) */
orc_ctx* orc_create(const mrm_grid_desc* g, const mrm_phys_desc* p) {
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  if (!c) return NULL;
  c->bufsize = 1;
  c->vsize = 1;
  for (int d = 0; d < 3; d++) {
    c->gsize[d] = g->gsize[d];
    c->ghosts[d] = g->ghosts[d];
    c->periods[d] = g->periods[d];
    c->lsize[d] = g->gsize[d];
    c->lbeg[d] = g->ghosts[d];
    c->lend[d] = g->gsize[d] + g->ghosts[d] - 1;
    c->n[d] = g->gsize[d] + 2 * g->ghosts[d];
    c->bufsize *= c->n[d];
    c->vsize = imax(c->vsize, c->n[d]);
    c->gxmin[d] = g->gxmin[d];
    c->gxmax[d] = g->gxmax[d];
    c->dx[d] = (g->gxmax[d] - g->gxmin[d]) / g->gsize[d];
    c->lxmin[d] = g->gxmin[d];
  }
  c->stride[0] = 1;
  c->stride[1] = c->n[0];
  c->stride[2] = (long)c->n[0] * c->n[1];
  c->ph = *p;
  size_t bs = (size_t)c->bufsize;
  c->U = (double*)calloc(NV * bs, sizeof(double));
  c->U0 = (double*)calloc(NV * bs, sizeof(double));
  c->Res = (double*)calloc(NV * bs, sizeof(double));
  for (int f = 0; f < MRM_NFIELD; f++) c->fld[f] = (double*)calloc(bs, sizeof(double));
  c->inv_va2max = 0.0;
  c->v_lim = 0.0;
  c->e_lim = 0.0;
  c->va_max_old = 0.0;
  c->valf_ini = 1;
  c->p_bc = -1.0;
  c->s_bc = -1.0;
  c->bnd_ini = 1;
  c->tvd_ini = 1;
  c->divb_next_iter = 3;
  if (p->params[3] > 0.0) {
    c->sflx0 = (double*)calloc(bs, sizeof(double));
    c->Rflx = (double*)calloc(bs, sizeof(double));
    c->BgradT = (double*)calloc(bs, sizeof(double));
  }
  return c;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  free(c->U); free(c->U0); free(c->Res);
  for (int f = 0; f < MRM_NFIELD; f++) free(c->fld[f]);
  free(c->xeps); free(c->xlr); free(c->xlp3); free(c->xs3);
  for (int t = 0; t < 6; t++) free(c->tab[t]);
  for (int t = 0; t < 2; t++) free(c->itab[t]);
  free(c->hfb); free(c->hft); free(c->qft);
  free(c->curlB);
  free(c->sflx0); free(c->Rflx); free(c->BgradT);
  for (int t = 0; t < 10; t++) free(c->diag[t]);
  free(c->pkern);
  free(c->bext);
  free(c->dmp);
  free(c);
}

double* orc_state(orc_ctx* c, int which) {
  return which == MRM_U ? c->U : which == MRM_U0 ? c->U0 : c->Res;
}
double* orc_field(orc_ctx* c, int f) { return c->fld[f]; }
void orc_get_limits(const orc_ctx* c, double lim[4]) {
  lim[0] = c->inv_va2max; lim[1] = c->v_lim; lim[2] = c->e_lim; lim[3] = c->va_max_old;
}
void orc_set_limits(orc_ctx* c, const double lim[4]) {
  c->inv_va2max = lim[0]; c->v_lim = lim[1]; c->e_lim = lim[2]; c->va_max_old = lim[3];
}
void orc_get_bc(const orc_ctx* c, double* p_bc, double* s_bc) { *p_bc = c->p_bc; *s_bc = c->s_bc; }
void orc_set_bc(orc_ctx* c, double p_bc, double s_bc) { c->p_bc = p_bc; c->s_bc = s_bc; }

/* ------------------------------------------------------------------------ */
/* eos_init (eos_init.C:57-212): axes in double from the float header, the   */
/* forward tables transposed to [i_eps][i_rho] and divided by del_eps*del_lr */
/* in double before rounding back to float; same for the inverse tables.     */
int orc_eos_load(orc_ctx* c, const void* bytes, size_t nbytes) {
  const float* buf = (const float*)bytes;
  if (nbytes < 14 * sizeof(float)) return MRM_ERR_EOS;
  int N_eps = (int)buf[0], N_lr = (int)buf[1], N_lp3 = (int)buf[2], N_s3 = (int)buf[3];
  if (N_eps < 2 || N_lr < 2 || N_lp3 < 2 || N_s3 < 2) return MRM_ERR_EOS;
  size_t need = 14 + 6 * (size_t)N_eps * N_lr + 2 * (size_t)N_lp3 * N_s3;
  if (nbytes < need * sizeof(float)) return MRM_ERR_EOS;
  double eps0 = buf[4], eps1 = buf[5], lr0 = buf[6], lr1 = buf[7];
  double lp0 = buf[8], lp1 = buf[9], s0 = buf[10], s1 = buf[11];
  c->N_eps = N_eps; c->N_lr = N_lr; c->N_lp3 = N_lp3; c->N_s3 = N_s3;
  c->eps_off = buf[12];
  c->ss_off = buf[13];
  c->del_eps = (eps1 - eps0) / ((double)(N_eps - 1));
  c->del_lr = (lr1 - lr0) / ((double)(N_lr - 1));
  c->inv_del_eps = 1.0 / c->del_eps;
  c->inv_del_lr = 1.0 / c->del_lr;
  c->del_lp3 = (lp1 - lp0) / ((double)(N_lp3 - 1));
  c->del_s3 = (s1 - s0) / ((double)(N_s3 - 1));
  c->inv_del_lp3 = 1.0 / c->del_lp3;
  c->inv_del_s3 = 1.0 / c->del_s3;
  c->xeps = (double*)realloc(c->xeps, sizeof(double) * N_eps);
  c->xlr = (double*)realloc(c->xlr, sizeof(double) * N_lr);
  c->xlp3 = (double*)realloc(c->xlp3, sizeof(double) * N_lp3);
  c->xs3 = (double*)realloc(c->xs3, sizeof(double) * N_s3);
  for (int i = 0; i < N_eps; i++) c->xeps[i] = eps0 + c->del_eps * ((double)i);
  for (int i = 0; i < N_lr; i++) c->xlr[i] = lr0 + c->del_lr * ((double)i);
  for (int i = 0; i < N_lp3; i++) c->xlp3[i] = lp0 + c->del_lp3 * ((double)i);
  for (int i = 0; i < N_s3; i++) c->xs3[i] = s0 + c->del_s3 * ((double)i);
  const float* fw = buf + 14;
  size_t nt = (size_t)N_eps * N_lr;
  double sc = c->del_eps * c->del_lr;
  for (int t = 0; t < 6; t++) {
    c->tab[t] = (float*)realloc(c->tab[t], sizeof(float) * nt);
    for (int i = 0; i < N_eps; i++)
      for (int j = 0; j < N_lr; j++)
        c->tab[t][(size_t)i * N_lr + j] = (float)(((double)fw[t * nt + (size_t)j * N_eps + i]) / sc);
  }
  const float* iv = fw + 6 * nt;
  size_t ni = (size_t)N_lp3 * N_s3;
  double sci = c->del_lp3 * c->del_s3;
  for (int t = 0; t < 2; t++) {
    c->itab[t] = (float*)realloc(c->itab[t], sizeof(float) * ni);
    for (int i = 0; i < N_lp3; i++)
      for (int j = 0; j < N_s3; j++)
        c->itab[t][(size_t)i * N_s3 + j] = (float)(((double)iv[t * ni + (size_t)j * N_lp3 + i]) / sci);
  }
  c->eos_ok = 1;
  return MRM_OK;
}

#define TAB(c, t, ie, ir) ((double)(c)->tab[t][(size_t)(ie) * (c)->N_lr + (ir)])
#define ITAB(c, t, ip, is) ((double)(c)->itab[t][(size_t)(ip) * (c)->N_s3 + (is)])

/* d3_interp (eos_init.C:216-239) */
double orc_d3_interp(const orc_ctx* c, double pp, double ss) {
  double ss1 = ss + c->ss_off;
  double logp = log(pp);
  int i = (int)((logp - c->xlp3[0]) * c->inv_del_lp3);
  if (i < 0) i = 0;
  if (i > c->N_lp3 - 2) i = c->N_lp3 - 2;
  int j = (int)((ss1 - c->xs3[0]) * c->inv_del_s3);
  if (j < 0) j = 0;
  if (j > c->N_s3 - 2) j = c->N_s3 - 2;
  double dd = (logp - c->xlp3[i]) * (ss1 - c->xs3[j]) * ITAB(c, 1, i + 1, j + 1)
            + (logp - c->xlp3[i]) * (c->xs3[j + 1] - ss1) * ITAB(c, 1, i + 1, j)
            + (c->xlp3[i + 1] - logp) * (ss1 - c->xs3[j]) * ITAB(c, 1, i, j + 1)
            + (c->xlp3[i + 1] - logp) * (c->xs3[j + 1] - ss1) * ITAB(c, 1, i, j);
  return exp(dd);
}

/* eps3_interp (eos_init.C:242-265) */
double orc_eps3_interp(const orc_ctx* c, double pp, double ss) {
  double ss1 = ss + c->ss_off;
  double logp = log(pp);
  int i = (int)((logp - c->xlp3[0]) * c->inv_del_lp3);
  if (i < 0) i = 0;
  if (i > c->N_lp3 - 2) i = c->N_lp3 - 2;
  int j = (int)((ss1 - c->xs3[0]) * c->inv_del_s3);
  if (j < 0) j = 0;
  if (j > c->N_s3 - 2) j = c->N_s3 - 2;
  double ee = (logp - c->xlp3[i]) * (ss1 - c->xs3[j]) * ITAB(c, 0, i + 1, j + 1)
            + (logp - c->xlp3[i]) * (c->xs3[j + 1] - ss1) * ITAB(c, 0, i + 1, j)
            + (c->xlp3[i + 1] - logp) * (ss1 - c->xs3[j]) * ITAB(c, 0, i, j + 1)
            + (c->xlp3[i + 1] - logp) * (c->xs3[j + 1] - ss1) * ITAB(c, 0, i, j);
  ee = exp(ee);
  return ee - c->eps_off;
}

/* shared body of s_interp / T_interp / p_interp (eos_init.C:268-343) */
static inline double fwd_interp(const orc_ctx* c, int t, double ee, double dd, double* out_raw) {
  double ee1 = log(ee + c->eps_off);
  double logr = log(dd);
  int i = (int)((ee1 - c->xeps[0]) * c->inv_del_eps);
  if (i < 0) i = 0;
  if (i > c->N_eps - 2) i = c->N_eps - 2;
  int j = (int)((logr - c->xlr[0]) * c->inv_del_lr);
  if (j < 0) j = 0;
  if (j > c->N_lr - 2) j = c->N_lr - 2;
  double v = (logr - c->xlr[j]) * (ee1 - c->xeps[i]) * TAB(c, t, i + 1, j + 1)
           + (logr - c->xlr[j]) * (c->xeps[i + 1] - ee1) * TAB(c, t, i, j + 1)
           + (c->xlr[j + 1] - logr) * (ee1 - c->xeps[i]) * TAB(c, t, i + 1, j)
           + (c->xlr[j + 1] - logr) * (c->xeps[i + 1] - ee1) * TAB(c, t, i, j);
  *out_raw = v;
  return v;
}

double orc_s_interp(const orc_ctx* c, double ee, double dd) {
  double v; fwd_interp(c, 2, ee, dd, &v);
  return v - c->ss_off;
}
double orc_T_interp(const orc_ctx* c, double ee, double dd) {
  double v; fwd_interp(c, 1, ee, dd, &v);
  return exp(v);
}
double orc_p_interp(const orc_ctx* c, double ee, double dd) {
  double v; fwd_interp(c, 0, ee, dd, &v);
  return exp(v);
}

/* ------------------------------------------------------------------------ */
/* ConsToPrim (cons_to_prim.C:39-236) over the whole buffer                  */
static const double eos_gamma = 1.65, eos_mu = 0.62, eos_mu_n = 1.3;
static const double mh = 1.67262158e-24, kb = 1.380622e-16;

static inline int invalid_eos(double a, double b) {
  int status = 0;
  if ((a != a) || (a <= 0.) || (a == INFINITY)) status = 1;
  if ((b != b) || (b <= 0.) || (b == INFINITY)) status = 1;
  return status;
}

void orc_cons_to_prim(orc_ctx* c) {
  const long bs = c->bufsize;
  const double eps_max = exp(c->xeps[c->N_eps - 1]) - c->eps_off;
  const double Rgas = 8.314e7;
  const double c_temp = (eos_gamma - 1.) * eos_mu / Rgas;
  const double c_pres = (eos_gamma - 1.);
  double* U = c->U;
  double *P = c->fld[MRM_PRES], *T = c->fld[MRM_TEMP], *NE = c->fld[MRM_NE];
  double *RI = c->fld[MRM_RHOI], *AM = c->fld[MRM_AMB];
#pragma omp parallel for schedule(static)
  for (long node = 0; node < bs; node++) {
    double d = U[node], m1 = U[bs + node], m2 = U[2 * bs + node], m3 = U[3 * bs + node];
    double e = U[4 * bs + node];
    double dn = 1.0 / d;
    m1 = m1 * dn; m2 = m2 * dn; m3 = m3 * dn;
    double vv = (m1 * m1 + m2 * m2 + m3 * m3);
    e = e - 0.5 * vv * d;
    double eps = e * dn + c->eps_off;
    double lgr = log(d);
    double ep = log(eps);
    double lr = lgr;
    int i_d = (int)((lr - c->xlr[0]) * c->inv_del_lr);
    i_d = imax(0, i_d);
    i_d = imin(c->N_lr - 2, i_d);
    int i_e = (int)((ep - c->xeps[0]) * c->inv_del_eps);
    i_e = imax(0, i_e);
    i_e = imin(c->N_eps - 2, i_e);
    double coeff0 = (ep - c->xeps[i_e]) * (lr - c->xlr[i_d]);
    double coeff1 = (ep - c->xeps[i_e]) * (c->xlr[i_d + 1] - lr);
    double coeff2 = (c->xeps[i_e + 1] - ep) * (lr - c->xlr[i_d]);
    double coeff3 = (c->xeps[i_e + 1] - ep) * (c->xlr[i_d + 1] - lr);
    double r[6];
    for (int t = 0; t < 6; t++) {
      if (t == 2) continue;  /* entropy is not interpolated here */
      double v = coeff0 * TAB(c, t, i_e + 1, i_d + 1);
      v += coeff1 * TAB(c, t, i_e + 1, i_d);
      v += coeff2 * TAB(c, t, i_e, i_d + 1);
      v += coeff3 * TAB(c, t, i_e, i_d);
      r[t] = v;
    }
    double pres = exp(r[0]), temp = exp(r[1]), nel = exp(r[3]), rhoi = exp(r[4]), amb = exp(r[5]);
    if (lgr < c->xlr[0]) {
      pres = c_pres * e;
      temp = c_temp * eps;
      nel = pres / (temp * kb) - d / (eos_mu_n * mh);
    }
    if (eps > 0.8 * eps_max) {
      double xx = smax(0.0, 5.0 * (eps_max - eps) / eps_max);
      pres = pres * xx + (1.0 - xx) * c_pres * e;
      temp = temp * xx + (1.0 - xx) * c_temp * eps;
      nel = pres / (temp * kb) - d / (eos_mu_n * mh);
    }
    if (invalid_eos(pres, temp)) {
      pres = c_pres * e;
      temp = c_temp * eps;
    }
    if (nel == INFINITY) nel = pres / (temp * kb) - d / (eos_mu_n * mh);
    if ((nel < 1.0) || (nel != nel) || (nel == INFINITY)) nel = 1.0;
    U[bs + node] = m1; U[2 * bs + node] = m2; U[3 * bs + node] = m3; U[4 * bs + node] = e;
    P[node] = pres; T[node] = temp; NE[node] = nel; RI[node] = rhoi; AM[node] = amb;
  }
}

/* ------------------------------------------------------------------------ */
/* Adjust_Valf_Max (Adjust_Valf_Max.C:25-150) on one rank                    */
void orc_adjust_valf_max(orc_ctx* c, const mrm_run_desc* run, double stats[5]) {
  (void)run;
  const double max_fill = c->ph.params[5];
  const double cs_fac = 2.0, vv_fac = 3.0, eps_decay = 0.99;
  const long bs = c->bufsize;
  const double* U = c->U;
  const double* P = c->fld[MRM_PRES];
  double l_max0 = 0.0, l_max1 = 0.0, l_max2 = 0.0, l_sum0 = 0.0, l_sum1 = 0.0;
  const double vlim95 = c->v_lim * 0.95, elim95 = c->e_lim * 0.95;
#pragma omp parallel for collapse(2) reduction(max:l_max0, l_max1, l_max2) reduction(+:l_sum0, l_sum1)
  for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
    for (int j = c->lbeg[1]; j <= c->lend[1]; j++) {
      double m0 = 0.0, m1 = 0.0, m2 = 0.0, s0 = 0.0, s1 = 0.0;
      for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
        long node = i + c->stride[1] * j + c->stride[2] * k;
        double vx = U[bs + node], vy = U[2 * bs + node], vz = U[3 * bs + node];
        double vv = sqrt(vx * vx + vy * vy + vz * vz);
        double ee = U[4 * bs + node] / U[node];
        m0 = fmax(m0, vv);
        m1 = fmax(m1, 5 / 3. * P[node] / U[node]);
        m2 = fmax(m2, ee);
        if (vv >= vlim95) s0 += 1.0;
        if (ee >= elim95) s1 += 1.0;
      }
      l_max0 = fmax(l_max0, m0); l_max1 = fmax(l_max1, m1); l_max2 = fmax(l_max2, m2);
      l_sum0 += s0; l_sum1 += s1;
    }
  double g_max[3] = {l_max0, sqrt(l_max1), l_max2};
  double g_sum[2] = {l_sum0, l_sum1};
  double va_max;
  const double* prm = c->ph.params;
  if (prm[2] > 0.0) {
    if (c->valf_ini) {
      c->v_lim = fmin(g_max[0], c->ph.tchk[3]);
      c->e_lim = fmin(g_max[2], c->ph.tchk[2]);
      c->va_max_old = 0.0;
      c->valf_ini = 0;
    } else {
      if (g_sum[0] >= 4 * max_fill) c->v_lim *= 1.04;
      else if (g_sum[0] >= max_fill) c->v_lim *= 1.01;
      else if (g_sum[0] < 0.5 * max_fill) c->v_lim *= 0.99;
      if (g_sum[1] >= 4 * max_fill) c->e_lim *= 1.04;
      else if (g_sum[1] >= max_fill) c->e_lim *= 1.01;
      else if (g_sum[1] < 0.5 * max_fill) c->e_lim *= 0.99;
    }
    va_max = fmax(prm[1], prm[2] * vv_fac * g_max[0]);
    va_max = fmax(va_max, prm[2] * cs_fac * g_max[1]);
    va_max = fmax(va_max, c->va_max_old * eps_decay);
    c->v_lim = fmax(c->v_lim, prm[1] / vv_fac);
    c->v_lim = fmin(c->v_lim, c->ph.tchk[3]);
    c->e_lim = fmax(c->e_lim, pow(prm[1] / cs_fac, 2));
    c->e_lim = fmin(c->e_lim, c->ph.tchk[2]);
  } else {
    va_max = prm[1];
    c->v_lim = c->ph.tchk[3];
    c->e_lim = c->ph.tchk[2];
  }
  c->inv_va2max = 1.0 / pow(va_max, 2);
  c->va_max_old = va_max;
  if (stats) {
    stats[0] = g_max[0]; stats[1] = g_max[1]; stats[2] = g_max[2];
    stats[3] = g_sum[0]; stats[4] = g_sum[1];
  }
}

/* ------------------------------------------------------------------------ */
/* diagnostics arrays tvar1..8, Qres, Qvis: every operation called with     */
/* need_diagnostics adds its terms to them; the caller zeroes them at the    */
/* start of the stage that collects them, as solver.C:271-296 does.          */
enum { DG_T1, DG_T2, DG_T3, DG_T4, DG_T5, DG_T6, DG_T7, DG_T8, DG_QRES, DG_QVIS };
void orc_diag_zero(orc_ctx* c) {
  for (int t = 0; t < 10; t++) {
    if (!c->diag[t]) c->diag[t] = (double*)malloc(sizeof(double) * (size_t)c->bufsize);
    memset(c->diag[t], 0, sizeof(double) * (size_t)c->bufsize);
  }
}
static double** diag_of(orc_ctx* c, const mrm_run_desc* run) {
  if (!run || !run->need_diagnostics) return NULL;
  if (!c->diag[0]) orc_diag_zero(c);
  return c->diag;
}
double* orc_diag(orc_ctx* c, int t) { return (t >= 0 && t < 10) ? c->diag[t] : NULL; }

/* MHD_Residual (mhdres_SR.C:27-854) with Spitzer = ambipolar =             */
/* diagnostics = 0: three sweeps d = 0,1,2 (loop_order :86), each over the   */
/* interior transverse lines and the full line along d1 incl. ghosts.  With  */
/* eta > 0 each sweep adds its part of J = curl B to curlB (:742-749), the   */
/* last one adds eta |J|^2 to R.e (:440-451), every sweep adds eta times the */
/* 4th-order second derivative of B along d1 (:452-462, :757-766), and dt is */
/* limited by the diffusion time (:815-819).  With Spitzer conduction the    */
/* energy flux carries sflx B (:350-359), each sweep damps sflx with a 4th   */
/* difference and adds its part of B.grad T (:680-690), and the last sweep   */
/* relaxes sflx towards the saturation-limited Spitzer flux (:492-521).      */
double orc_mhd_residual(orc_ctx* c, const mrm_run_desc* run) {
  static const int loop_order[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
  static const double del[3][3] = {{1., 0., 0.}, {0., 1., 0.}, {0., 0., 1.}};
  static const double sign[3] = {-1., 1., -1.};
  const long bs = c->bufsize;
  const double* U = c->U;
  const double* Pg = c->fld[MRM_PRES];
  double* R = c->Res;
  const double inv_va2max = c->inv_va2max;
  const double eta = c->ph.params[4];
  const int needs_curlB = eta > 0.0;
  double dxmin = c->dx[0];
  for (int d = 1; d < 3; d++) dxmin = smin(dxmin, c->dx[d]);
  if (needs_curlB) {
    if (!c->curlB) c->curlB = (double*)malloc(sizeof(double) * 3 * (size_t)bs);
    memset(c->curlB, 0, sizeof(double) * 3 * (size_t)bs);
  }
  double* curlB = c->curlB;
  /* Spitzer constants (mhdres_SR.C:37-49, 136-145): from the dt of the call */
  const double k0_spt = c->ph.params[3];
  const int spitzer = k0_spt > 0.0;
  const double CFL_hyp = 1.8, avr_max = 0.25, eps_hyp = 0.02, X_H = 0.7, Rgas = 8.314e7;
  const double sat_flx = (1.0 + X_H) * pow(Rgas, 1.5) * sqrt(1836.) / 8.0;
  const double rdt = run ? run->dt : 0.0, rcfl = run ? run->CFL : 0.0;
  double cmax_hyp, nu_hyp_max, hyp_diff;
  if (rdt > 0.0) {
    cmax_hyp = smin(rcfl, CFL_hyp) * dxmin / rdt;
    nu_hyp_max = avr_max / rdt;
    hyp_diff = eps_hyp / rdt;
  } else {
    cmax_hyp = 0.0;
    nu_hyp_max = avr_max * cmax_hyp / (smin(rcfl, CFL_hyp) * dxmin);
    hyp_diff = eps_hyp * cmax_hyp / (smin(rcfl, CFL_hyp) * dxmin);
  }
  const double* sflxg = c->fld[MRM_SFLX];
  const double* temp = c->fld[MRM_TEMP];
  double** dg = diag_of(c, run);
  if (spitzer) memset(c->BgradT, 0, sizeof(double) * (size_t)bs);
  double cmax = 0.0;
  for (int d = 0; d < 3; d++) {
    const int d1 = loop_order[d][0], d2 = loop_order[d][1], d3 = loop_order[d][2];
    const double dx = c->dx[d1];
    const double w1 = 8. / (12. * dx);
    const double w2 = -1. / (12. * dx);
    const double ww0 = -30. / (12. * dx * dx);
    const double ww1 = 16. / (12. * dx * dx);
    const double ww2 = -1. / (12. * dx * dx);
    const double signd1 = sign[d1];
    const double p1 = del[d1][0], p2 = del[d1][1], p3 = del[d1][2];
    const int gh = c->ghosts[d1];
    const int sz = c->lend[d1] - c->lbeg[d1] + 1 + 2 * gh;
    const long str = c->stride[d1], strd2 = c->stride[d2], strd3 = c->stride[d3];
    const int kbeg = c->lbeg[d3], kend = c->lend[d3], jbeg = c->lbeg[d2], jend = c->lend[d2];
    double cmax_d = 0.0;
#pragma omp parallel reduction(max:cmax_d)
    {
      double* buf = (double*)malloc(sizeof(double) * (size_t)sz * (NV + NV + 7));
      double* var = buf;                 /* [NV][sz] */
      double* flx = var + NV * sz;       /* [NV][sz] */
      double* msx = flx + NV * sz;
      double* msy = msx + sz;
      double* msz = msy + sz;
      double* lf1 = msz + sz;
      double* sfl = lf1 + sz;            /* sflx along the line */
      double* vva = sfl + sz;            /* |v| */
      double* bba = vva + sz;            /* |B|^2 */
      double lcmax = 0.0;
#pragma omp for collapse(2) schedule(static)
      for (int k = kbeg; k <= kend; k++)
        for (int j = jbeg; j <= jend; j++) {
          long offset = j * strd2 + k * strd3;
          for (int i = 0; i < sz; i++) {
            long node = offset + i * str;
            for (int v = 0; v < NV; v++) var[v * sz + i] = U[v * bs + node];
            sfl[i] = spitzer ? sflxg[node] : 0.0;
          }
          for (int i = 0; i < sz; i++) {
            long node = offset + i * str;
            double var0 = var[i], var1 = var[sz + i], var2 = var[2 * sz + i], var3 = var[3 * sz + i];
            double var4 = var[4 * sz + i], var5 = var[5 * sz + i], var6 = var[6 * sz + i];
            double var7 = var[7 * sz + i];
            double vel = var[(1 + d1) * sz + i];
            double mag = var[(5 + d1) * sz + i];
            double dn = 1.0 / var0;
            double vv = var1 * var1 + var2 * var2 + var3 * var3;
            double bb = var5 * var5 + var6 * var6 + var7 * var7;
            double va2 = bb * dn;
            double ekin = 0.5 * var0 * vv;
            double pmag = 0.5 * bb;
            double x2 = va2 * inv_va2max;
            double x4 = x2 * x2;
            double s = 1.0 + x2;
            lf1[i] = s / (s + x4);
            double pres = Pg[node];
            double cs2 = 5.0 / 3.0 * pres * dn;
            vv = sqrt(vv);
            vva[i] = vv;
            bba[i] = bb;
            lcmax = smax(lcmax, vv + 0.0 + sqrt(smax(cs2, lf1[i] * (cs2 + va2))));
            double mflx = vel * var0;
            flx[0 * sz + i] = mflx;
            flx[1 * sz + i] = mflx * var1 + p1 * pres;
            flx[2 * sz + i] = mflx * var2 + p2 * pres;
            flx[3 * sz + i] = mflx * var3 + p3 * pres;
            flx[4 * sz + i] = vel * (var4 + ekin + pres);
            flx[5 * sz + i] = vel * var5 - var1 * mag;
            flx[6 * sz + i] = vel * var6 - var2 * mag;
            flx[7 * sz + i] = vel * var7 - var3 * mag;
            msx[i] = mag * var5 - p1 * pmag;
            msy[i] = mag * var6 - p2 * pmag;
            msz[i] = mag * var7 - p3 * pmag;
            if (spitzer) flx[4 * sz + i] += sfl[i] * var[(5 + d1) * sz + i];
          }
          for (int i = gh; i < sz - gh; i++) {
            long node = offset + i * str;
            double res[NV];
            for (int v = 0; v < NV; v++) {
              const double* F = flx + v * sz;
              res[v] = 0.0;
              res[v] -= w1 * (F[i + 1] - F[i - 1]) + w2 * (F[i + 2] - F[i - 2]);
            }
            if (dg) {   /* mhdres_SR.C:368-383, 688-697 */
              if (spitzer) {
                const double* Bd = var + (5 + d1) * sz;
                double dn = -(w1 * ((sfl[i + 1] * Bd[i + 1]) - (sfl[i - 1] * Bd[i - 1])) +
                              w2 * ((sfl[i + 2] * Bd[i + 2]) - (sfl[i - 2] * Bd[i - 2])));
                dg[DG_T1][node] += res[4] - dn;
                dg[DG_T2][node] += dn;
              } else {
                dg[DG_T1][node] += res[4];
              }
            }
            if (spitzer && d < 2) {
              const double* Bd = var + (5 + d1) * sz;
              c->Rflx[node] -= hyp_diff * ((sfl[i + 2] + sfl[i - 2]) - 4. * (sfl[i + 1] + sfl[i - 1]) + 6. * sfl[i]);
              c->BgradT[node] += (Bd[i] * (w1 * (temp[node + str] - temp[node - str]) +
                                           w2 * (temp[node + 2 * str] - temp[node - 2 * str])));
            }
            const double* B2 = var + (5 + d2) * sz;
            const double* B3 = var + (5 + d3) * sz;
            const double* B1 = var + (5 + d1) * sz;
            double dB2 = w1 * (B2[i + 1] - B2[i - 1]) + w2 * (B2[i + 2] - B2[i - 2]);
            double dB3 = w1 * (B3[i + 1] - B3[i - 1]) + w2 * (B3[i + 2] - B3[i - 2]);
            double curlBxB[3];
            curlBxB[d1] = -B2[i] * dB2 - B3[i] * dB3;
            curlBxB[d2] = B1[i] * dB2;
            curlBxB[d3] = B1[i] * dB3;
            double lrt0 = w1 * (msx[i + 1] - msx[i - 1]) + w2 * (msx[i + 2] - msx[i - 2]);
            double lrt1 = w1 * (msy[i + 1] - msy[i - 1]) + w2 * (msy[i + 2] - msy[i - 2]);
            double lrt2 = w1 * (msz[i + 1] - msz[i - 1]) + w2 * (msz[i + 2] - msz[i - 2]);
            double lf = lf1[i];
            lrt0 = lf * lrt0 + (1.0 - lf) * curlBxB[0];
            lrt1 = lf * lrt1 + (1.0 - lf) * curlBxB[1];
            lrt2 = lf * lrt2 + (1.0 - lf) * curlBxB[2];
            res[1] += lrt0;
            res[2] += lrt1;
            res[3] += lrt2;
            res[4] += var[sz + i] * lrt0 + var[2 * sz + i] * lrt1 + var[3 * sz + i] * lrt2;
            if (dg) dg[DG_T3][node] += var[sz + i] * lrt0 + var[2 * sz + i] * lrt1 + var[3 * sz + i] * lrt2;
            if (needs_curlB) {
              double cb[3];
              cb[d1] = 0;
              cb[d2] = signd1 * dB3;
              cb[d3] = -signd1 * dB2;
              for (int q = 0; q < 3; q++) curlB[q * bs + node] += cb[q];
              if (d == 2) {
                const double jx = curlB[node], jy = curlB[bs + node], jz = curlB[2 * bs + node];
                R[4 * bs + node] += eta * (jx * jx + jy * jy + jz * jz);
              }
              for (int v = 5; v <= 7; v++) {
                const double* X = var + v * sz;
                res[v] += eta * (ww0 * X[i] + ww1 * (X[i + 1] + X[i - 1]) + ww2 * (X[i + 2] + X[i - 2]));
              }
            }
            for (int v = 0; v < NV; v++) R[v * bs + node] += res[v];
            if (spitzer && d == 2) {
              const double* Bd = var + (5 + d1) * sz;
              const double tempi = temp[node];
              c->BgradT[node] += (Bd[i] * (w1 * (temp[node + str] - temp[node - str]) +
                                           w2 * (temp[node + 2 * str] - temp[node - 2 * str])));
              c->Rflx[node] -= hyp_diff * ((sfl[i + 2] + sfl[i - 2]) - 4. * (sfl[i + 1] + sfl[i - 1]) + 6. * sfl[i]);
              const double c_lim = cmax_hyp - vva[i];
              const double tt32 = tempi * sqrt(tempi);
              double kap_spt = k0_spt * tempi * tt32;
              const double F_sat = sat_flx * var[i] * tt32;
              const double invBB = 1.0 / smax(1e-100, bba[i]);
              double F_spt = -kap_spt * c->BgradT[node];
              const double f_lim = F_sat / (F_sat + fabs(F_spt * sqrt(invBB)));
              F_spt *= invBB * f_lim;
              kap_spt *= f_lim;
              const double nu_hyp = smin(c_lim * c_lim * var[4 * sz + i] / (tempi * kap_spt), nu_hyp_max);
              c->Rflx[node] += nu_hyp * (F_spt - sfl[i]);
            }
          }
        }
      cmax_d = smax(cmax_d, lcmax);
      free(buf);
    }
    cmax = smax(cmax, cmax_d);
  }
  double dt_cfl = dxmin / cmax;
  if (needs_curlB) {
    const double dt_diff_fac =
        0.5 / (1. / (c->dx[0] * c->dx[0]) + 1. / (c->dx[1] * c->dx[1]) + 1. / (c->dx[2] * c->dx[2]));
    double dt_res = dt_diff_fac / eta;
    if (run && run->CFL > 1) dt_res *= 0.5;
    dt_cfl = smin(dt_cfl, dt_res);
  }
  return dt_cfl;
}

/* Sync_timestep (solver.C:476-548) on one rank */
static int invalid_dt(double a) { return (a != a) || (a <= 0.) || (a == INFINITY); }

int orc_sync_timestep(orc_ctx* c, mrm_run_desc* run, double dt_rad, double dt_mhd, double* cfl) {
  double r0 = dt_rad, r1 = dt_mhd;
  if (invalid_dt(r0) || invalid_dt(r1)) return MRM_ERR_INVALID_DT;
  double dt_min_rt = c->ph.rt[3] * run->CFL * r1;
  if (r0 < dt_min_rt) r0 = smax(r0, dt_min_rt);
  double dt_min = smin(r0, r1);
  double dt = smin(r0, run->CFL * r1);
  if (run->outcad > 0) {
    double num = run->time / run->outcad;
    int ceil_ = (int)num;
    double rem = (num == (double)ceil_) ? 0.0 : (ceil_ + 1) * run->outcad - run->time;
    if (rem > 0.0) dt = smin(dt, rem);
  }
  if (run->autostep) run->dt = dt;
  *cfl = run->dt / dt_min;
  c->last_cfl = *cfl;
  if (*cfl >= 2.) return MRM_ERR_CFL;
  return MRM_OK;
}

/* ------------------------------------------------------------------------ */
/* Tcheck clamp (Add_Sources_Integrate_SR.C:447-465 and :603-617)            */
static inline void tcheck_cell(double* U, long bs, long node, double rho_min, double eps_min,
                               double eps_max, double vmax) {
  double dn = fmax(rho_min, U[node]);
  double mx = U[bs + node], my = U[2 * bs + node], mz = U[3 * bs + node];
  double vv = sqrt(mx * mx + my * my + mz * mz) / dn;
  double s = vmax / fmax(vmax, vv);
  double ekin = 0.5 * dn * vv * vv;
  double ee = U[4 * bs + node];
  double ee_low = eps_min * dn + ekin * s * s;
  double ee_up = eps_max * dn + ekin * s * s;
  U[4 * bs + node] = fmin(fmax(ee, ee_low), ee_up);
  U[node] = dn;
  U[bs + node] = mx * s; U[2 * bs + node] = my * s; U[3 * bs + node] = mz * s;
}

/* get_damping (Add_Sources_Integrate_SR.C:628-761) on one rank: rho_max over
   the interior at the first call; then per layer i the sum over the
   horizontal plane of M_x rho / hsize (k outer, j inner, as the loop nest
   accumulates), damped by hh = min(1/tau_min, (1/tau_ref) (mean / (vel_ref
   rho_max))^4) */
static void get_damping(orc_ctx* c) {
  const long bs = c->bufsize;
  const double* U = c->U;
  const int hsize = c->gsize[1] * c->gsize[2];
  const double tau_ref = 1.0 / c->ph.dmp[1], vel_ref = c->ph.dmp[2], tau_min = 1.0 / c->ph.dmp[3];
  if (!c->damp_ini) {
    double hh = 0.0;
    for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
      for (int j = c->lbeg[1]; j <= c->lend[1]; j++)
        for (int i = c->lbeg[0]; i <= c->lend[0]; i++)
          hh = fmax(hh, U[i + j * c->stride[1] + k * c->stride[2]]);
    c->damp_rho_max = hh;
    c->damp_ini = 1;
  }
  const int n0 = c->n[0];
  double* lbuf = (double*)calloc(n0, sizeof(double));
  for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
    for (int j = c->lbeg[1]; j <= c->lend[1]; j++)
      for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
        const long node = i + j * c->stride[1] + k * c->stride[2];
        lbuf[i] += U[bs + node] * U[node] / hsize;
      }
  for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
    double hh = tau_ref * pow(lbuf[i] / (vel_ref * c->damp_rho_max), 4);
    hh = fmin(tau_min, hh);
    c->dmp[i] = lbuf[i] * hh;
  }
  free(lbuf);
}

/* Source_Integrate_Tcheck, non-diagnostic branch                             */
/* (Add_Sources_Integrate_SR.C:38-110 and :299-467) with the damping profile  */
/* of stage 1 (:92-108), RT source from Qtot (+Qthin when ext_cor, which is   */
/* outside the path).                                                         */
void orc_source_integrate_tcheck(orc_ctx* c, const mrm_run_desc* run, int stage) {
  const long bs = c->bufsize;
  const double wdt = run->dt / (double)(MAXSTAGE + 1 - stage);
  const double eps_min = c->ph.tchk[0], rho_min = c->ph.tchk[1];
  const double vmax = (c->v_lim > 0.0) ? c->v_lim : c->ph.tchk[3];
  const double eps_max = (c->e_lim > 0.0) ? c->e_lim : c->ph.tchk[2];
  const double g = -c->ph.params[0];   /* Physics.g[0] (physics.H:201-203) */
  const double invRdt = c->ph.rt[4] / run->dt;
  const double inv_va2max = c->inv_va2max;
  double *U = c->U, *U0 = c->U0, *R = c->Res;
  const double* Qtot = c->fld[MRM_QTOT];
  double** dg = diag_of(c, run);   /* the diagnostic branch, :122-296 */
  const double invdt = 1.0 / run->dt;
  if (!c->dmp) c->dmp = (double*)calloc(c->n[0], sizeof(double));
  if (stage == 1 && (int)c->ph.dmp[0]) get_damping(c);
  const double* dmp = c->dmp;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
    for (int j = c->lbeg[1]; j <= c->lend[1]; j++) {
      long off0 = j * c->stride[1] + k * c->stride[2];
      for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
        long node = off0 + i;
        double Qrad = Qtot[node];
        double dn = invRdt * U[4 * bs + node];
        double efaci = dn / fmax(dn, fabs(Qrad));
        dn = g * U[node] - dmp[i];   /* gravity + damping (:164, :340) */
        double e_dmpi = dn * U[bs + node];
        R[bs + node] += dn;
        R[4 * bs + node] += e_dmpi + Qrad * efaci;
        if (dg) {
          dg[DG_T4][node] = efaci;
          dg[DG_T5][node] = e_dmpi;
        }
      }
      for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
        long node = off0 + i;
        double rh = U[node], rd = R[node];
        double vx = U[bs + node], vy = U[2 * bs + node], vz = U[3 * bs + node];
        double bx = U[5 * bs + node], by = U[6 * bs + node], bz = U[7 * bs + node];
        double fsr_x = R[bs + node], fsr_y = R[2 * bs + node], fsr_z = R[3 * bs + node];
        double bb = bx * bx + by * by + bz * bz;
        double va2 = bb / rh;
        double x2 = va2 * inv_va2max;
        double x4 = x2 * x2;
        double bf = x4 / (1 + x2 + x4);
        fsr_x -= vx * rd;
        fsr_y -= vy * rd;
        fsr_z -= vz * rd;
        double dn = (fsr_x * bx + fsr_y * by + fsr_z * bz) / fmax(1e-100, bb);
        double boris0 = -bf * (fsr_x - dn * bx);
        double boris1 = -bf * (fsr_y - dn * by);
        double boris2 = -bf * (fsr_z - dn * bz);
        double boris3 = boris0 * vx + boris1 * vy + boris2 * vz;
        R[bs + node] += boris0;
        R[2 * bs + node] += boris1;
        R[3 * bs + node] += boris2;
        R[4 * bs + node] += boris3;
        if (dg) dg[DG_T6][node] = boris3;
      }
      for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
        long node = off0 + i;
        for (int v = 0; v < NV; v++) {
          U[v * bs + node] = U0[v * bs + node] + wdt * R[v * bs + node];
          R[v * bs + node] = 0;
        }
      }
      for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
        const double ee = U[4 * bs + off0 + i];
        tcheck_cell(U, bs, off0 + i, rho_min, eps_min, eps_max, vmax);
        if (dg) dg[DG_T7][off0 + i] += (U[4 * bs + off0 + i] - ee) * invdt;
      }
      if (c->Rflx) {   /* hyperbolic conduction (Add_Sources_Integrate_SR.C:410-419) */
        double* sflx = c->fld[MRM_SFLX];
        for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
          const long node = off0 + i;
          sflx[node] = c->sflx0[node] + wdt * c->Rflx[node];
          c->Rflx[node] = 0;
        }
      }
    }
}

/* SaveCons (Add_Sources_Integrate_SR.C:475-531): U0 <- U over the interior */
void orc_save_cons(orc_ctx* c) {
  const long bs = c->bufsize;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
    for (int j = c->lbeg[1]; j <= c->lend[1]; j++) {
      long off0 = j * c->stride[1] + k * c->stride[2];
      for (int v = 0; v < NV; v++)
        for (int i = c->lbeg[0]; i <= c->lend[0]; i++)
          c->U0[v * bs + off0 + i] = c->U[v * bs + off0 + i];
      if (c->sflx0)   /* Add_Sources_Integrate_SR.C:504-517 */
        for (int i = c->lbeg[0]; i <= c->lend[0]; i++) c->sflx0[off0 + i] = c->fld[MRM_SFLX][off0 + i];
    }
}

/* TCheck (Add_Sources_Integrate_SR.C:534-625) */
void orc_tcheck(orc_ctx* c, const mrm_run_desc* run) {
  const long bs = c->bufsize;
  double** dg = diag_of(c, run);
  const double invdt = dg ? 1.0 / run->dt : 0.0;
  const double eps_min = c->ph.tchk[0], rho_min = c->ph.tchk[1];
  const double vmax = (c->v_lim > 0.0) ? c->v_lim : c->ph.tchk[3];
  const double eps_max = (c->e_lim > 0.0) ? c->e_lim : c->ph.tchk[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
    for (int j = c->lbeg[1]; j <= c->lend[1]; j++) {
      long off0 = j * c->stride[1] + k * c->stride[2];
      for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
        const double eei = c->U[4 * bs + off0 + i];
        tcheck_cell(c->U, bs, off0 + i, rho_min, eps_min, eps_max, vmax);
        if (dg) dg[DG_T7][off0 + i] += (c->U[4 * bs + off0 + i] - eei) * invdt;
      }
    }
}

/* ------------------------------------------------------------------------ */
/* TVDlimit (tvdlimit_SR.C:33-695), ambipolar off.  With need_diagnostics   */
/* it adds tvar8 (energy diffusion), tvar6 (its Boris projection), Qres,     */
/* tvar7 (the energy floor) and Qvis (viscous heating), :577-666.            */
void orc_tvdlimit(orc_ctx* c, const mrm_run_desc* run, double dt_tvd) {
  double** dg = diag_of(c, run);
  const double idt_full = dg ? 1.0 / run->dt : 0.0;
  const double dt_fac = dg ? dt_tvd / run->dt : 0.0;
  const mrm_phys_desc* P = &c->ph;
  const long bs = c->bufsize;
  const double rho_min = P->tchk[1], eps_min = P->tchk[0];
  const double* tvd_h = P->tvd_h;
  const double* tvd_cs = P->tvd_cs;
  const double rho_lev = pow(P->tvd[0], 2);
  const int rho_log = (int)P->tvd[1];
  const double q_rho_max = log(P->tvd[2]);
  const double visc_coeff_bot = smin(1.0, P->tvd_visc_bnd[0]);
  const double visc_slope_bot = smax(0.0, P->tvd_visc_bnd[0] - 1.0);
  const double visc_coeff_top = smin(1.0, P->tvd_visc_bnd[1]);
  const double visc_slope_top = smax(0.0, P->tvd_visc_bnd[1] - 1.0);
  const double eta_coeff_bot = smin(1.0, P->tvd_eta_bnd[0]);
  const double eta_slope_bot = smax(0.0, P->tvd_eta_bnd[0] - 1.0);
  const double eta_coeff_top = smin(1.0, P->tvd_eta_bnd[1]);
  const double eta_slope_top = smax(0.0, P->tvd_eta_bnd[1] - 1.0);
  const double B_par_diff = P->tvd[3];
  const double vhyp = P->tvd[4];
  const double Qdiff_bnd = P->tvd[5];
  const double tvd_pm_v = P->tvd[6];
  const double tvd_pm_B = P->tvd[7];
  const double vmax_lim = P->tvd[8];
  const double CME_thresh = P->tvd[9];
  const double h_bnd_bot = (P->tvd_h_bnd[0] < 1) ? P->tvd_h_bnd[0] : P->tvd_h_bnd[0] / (double)c->gsize[0];
  const double h_bnd_top = (P->tvd_h_bnd[1] < 1) ? P->tvd_h_bnd[1] : P->tvd_h_bnd[1] / (double)c->gsize[0];
  const double tc0 = P->tvd_coeff[0], tc1 = P->tvd_coeff[1], tc2 = P->tvd_coeff[2], tc3 = P->tvd_coeff[3];
  const int need_tvd_coeff = (tc0 * tc1 * tc2 * tc3 != 1.0);
  const double inv_va2max = c->inv_va2max;
  double idx[3] = {1. / c->dx[0], 1. / c->dx[1], 1.0 / c->dx[2]};

  if (c->tvd_ini) {  /* tvdlimit_SR.C:132-149, once */
    c->hfb = (double*)malloc(sizeof(double) * c->vsize);
    c->hft = (double*)malloc(sizeof(double) * c->vsize);
    c->qft = (double*)malloc(sizeof(double) * c->vsize);
    for (int i = c->lbeg[0] - c->ghosts[0]; i <= c->lend[0] + c->ghosts[0]; i++) {
      double coord = c->lxmin[0] + (0.5 + (i - c->lbeg[0])) * c->dx[0];
      double hh = coord / c->gxmax[0];
      double vv = smax(0.0, 1.0 - 0.5 * (1.0 - hh) / h_bnd_top);
      c->hft[i] = smin(1.0, vv * vv);
      vv = smax(0.0, 1.0 - 0.5 * hh / h_bnd_bot);
      c->hfb[i] = smin(1.0, vv * vv);
      c->qft[i] = (hh < 1 - 2 * h_bnd_top) ? 1.0 : Qdiff_bnd;
    }
    c->tvd_ini = 0;
  }
  double vsqr_diff;
  if (vmax_lim < 1.0) vsqr_diff = pow(vmax_lim * c->v_lim, 2.0);
  else vsqr_diff = pow(vmax_lim, 2.0);

  int bounds[3][2];
  for (int d = 0; d < 3; d++) {
    bounds[d][0] = c->lbeg[d] - c->ghosts[d];
    bounds[d][1] = c->lend[d] + c->ghosts[d];
  }
  static const int loop_order[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
  double* U = c->U;
  const double* Pg = c->fld[MRM_PRES];
  const double* Tau = c->fld[MRM_TAU];
  const double *hfb = c->hfb, *hft = c->hft, *qft = c->qft;
  const int is_gbeg0 = 1, is_gend0 = 1;

  for (int d = 0; d < 3; d++) {
    const int d1 = loop_order[d][0], d2 = loop_order[d][1], d3 = loop_order[d][2];
    const double cmax = 0.975 * c->dx[d1] / dt_tvd;
    const int sz = bounds[d1][1] - bounds[d1][0] + 1;
    const int gh = c->ghosts[d1];
    const long str = c->stride[d1], strided2 = c->stride[d2], strided3 = c->stride[d3];
    const double idxd1 = idx[d1];
    const int kbeg = bounds[d3][0], kend = bounds[d3][1], jbeg = bounds[d2][0], jend = bounds[d2][1];
#pragma omp parallel
    {
      br_acc bc;
      memset(&bc, 0, sizeof bc);
      bc.tie = c->btie;
      const int NA = 3 * NV + 16;
      double* mem = (double*)malloc(sizeof(double) * (size_t)sz * NA);
      double* var = mem;
      double* slp = var + NV * sz;
      double* flx = slp + NV * sz;
      double* a = flx + NV * sz;
      double *CZ = a, *hfb1 = a + sz, *hft1 = a + 2 * sz, *qft1 = a + 3 * sz, *vsqr = a + 4 * sz;
      double *bsqr = a + 5 * sz, *cm = a + 6 * sz, *hyp_e = a + 7 * sz, *hyp_v = a + 8 * sz;
      double *hyp_B = a + 9 * sz, *BC2 = a + 10 * sz, *hyperdiff = a + 11 * sz, *qrho = a + 12 * sz;
      double* res = a + 13 * sz;   /* only NV values per point used at a time */
      (void)res;
#pragma omp for collapse(2) schedule(static)
      for (int k = kbeg; k <= kend; k++)
        for (int j = jbeg; j <= jend; j++) {
          long offset = j * strided2 + k * strided3;
          /* branch accounting on lines with interior transverse indices */
          const int cnt = c->bcount_on && j >= c->lbeg[d2] && j <= c->lend[d2] && k >= c->lbeg[d3] &&
                          k <= c->lend[d3];
          for (int i = 0; i < sz; i++) {
            long node = offset + i * str;
            var[i] = smax(rho_min, U[node]);
            for (int v = 1; v < NV; v++) var[v * sz + i] = U[v * bs + node];
            CZ[i] = (double)(Tau[node] > 1.0e-5);
            int pi = (d1 == 0) ? i : (d2 == 0) ? j : k;
            hfb1[i] = hfb[pi]; hft1[i] = hft[pi]; qft1[i] = qft[pi];
          }
          for (int i = 0; i < sz; i++) {
            long node = offset + i * str;
            double dn = 1.0 / var[i];
            var[sz + i] = var[sz + i] * dn;
            var[2 * sz + i] = var[2 * sz + i] * dn;
            var[3 * sz + i] = var[3 * sz + i] * dn;
            double vv = (var[sz + i] * var[sz + i] + var[2 * sz + i] * var[2 * sz + i] +
                         var[3 * sz + i] * var[3 * sz + i]);
            double bb = (var[5 * sz + i] * var[5 * sz + i] + var[6 * sz + i] * var[6 * sz + i] +
                         var[7 * sz + i] * var[7 * sz + i]);
            var[4 * sz + i] = var[4 * sz + i] * dn - 0.5 * vv;
            vsqr[i] = vv;
            bsqr[i] = bb;
            vv = sqrt(vv);
            double va2 = bb * dn;
            double x2 = va2 * inv_va2max;
            double x4 = x2 * x2;
            double s = 1.0 + x2;
            double lf = s / (s + x4);
            double cs2 = 5.0 / 3.0 * Pg[node] * dn;
            double cfast = sqrt(va2 + cs2);
            double CME_mode = (1.0 - CZ[i]) * ((double)(vv >= CME_thresh * cfast));
            if (cnt) {
              br_hit(&bc, MRM_BR_CZ, MRM_BK_CZ, br_key(d1, bs, node), CZ[i] != 0.0,
                     br_cls(&bc, fabs(Tau[node] - 1.0e-5), 1.0e-5));
              const double ct = CME_thresh * cfast;
              br_hit(&bc, MRM_BR_CME, MRM_BK_CME, br_key(d1, bs, node), CME_mode != 0.0,
                     CZ[i] != 0.0 ? 0 : br_cls(&bc, fabs(vv - ct), fmax(vv, ct)));
            }
            dn = var[i] * var[i];
            double rf = smin(lf, dn / (dn + rho_lev));
            double hh = 1.0 - hfb1[i] - hft1[i];
            double cf = hfb1[i] * tvd_cs[0] + (rf * tvd_cs[1] + (1.0 - rf) * tvd_cs[2]) * hh + hft1[i] * tvd_cs[3];
            cf = smax(cf, CME_mode);
            cs2 *= cf * cf;
            double vv_amb = 0.0;
            cm[i] = vv_amb + vv + sqrt(smax(cs2, lf * (cs2 + va2)));
            hyp_e[i] = hfb1[i] * tvd_h[0] + (rf * tvd_h[1] + (1.0 - rf) * tvd_h[2]) * hh + hft1[i] * tvd_h[3];
            hyp_v[i] = hfb1[i] * tvd_h[0] + (rf * tvd_h[1] + (1.0 - rf) * tvd_h[2] * tvd_pm_v) * hh +
                       hft1[i] * tvd_h[3] * tvd_pm_v;
            hyp_B[i] = hfb1[i] * tvd_h[0] + (rf * tvd_h[1] + (1.0 - rf) * tvd_h[2] * tvd_pm_B) * hh +
                       hft1[i] * tvd_h[3] * tvd_pm_B;
            hyp_e[i] *= (1.0 - CME_mode);
            hyp_v[i] *= (1.0 - CME_mode);
            BC2[i] = 1.0 - lf;
          }
          if ((d1 == 0) && (vhyp > 0.0)) {
            for (int i = 0; i < sz; i++) {
              hyperdiff[i] = vhyp * sqrt(vsqr[i]);
              hyperdiff[i] = 0.5 * idxd1 * smax(0.0, smin(hyperdiff[i], cmax - cm[i]));
            }
          }
          if (rho_log) {
            for (int i = 0; i < sz; i++) {
              var[i] = log(var[i]);
              var[4 * sz + i] = log(var[4 * sz + i]);
            }
            for (int i = 0; i < sz - 1; i++) qrho[i] = fabs(var[i] - var[i + 1]);
          } else {
            for (int i = 0; i < sz - 1; i++) qrho[i] = fabs(log(var[i]) - log(var[i + 1]));
          }
          for (int v = 0; v < NV; v++) {
            const double* X = var + v * sz;
            double* S = slp + v * sz;
            for (int i = 1; i < sz - 1; i++) {
              double sl = X[i] - X[i - 1];
              double sr = X[i + 1] - X[i];
              double slm = 0.25 * (sl + sr);
              double lower = smin(smin(sl, sr), slm);
              double upper = smax(smax(sl, sr), slm);
              S[i] = smax(0.0, lower) + smin(0.0, upper);
              if (cnt) {
                /* which of sl, sr, slm the min (positive) or max (negative) took */
                /* margin: the sign's distance from zero and the gap to the
                   runner-up; scale: the variable at the three points */
                int kc;
                double mg;
                if (lower > 0.0) {
                  double m1 = (sr < sl) ? sr : sl;
                  kc = (slm < m1) ? 2 : (sr < sl) ? 1 : 0;
                  double x0 = kc == 0 ? sl : kc == 1 ? sr : slm;
                  double gap = fmin(kc == 0 ? 1e300 : sl - x0, fmin(kc == 1 ? 1e300 : sr - x0, kc == 2 ? 1e300 : slm - x0));
                  mg = fmin(lower, gap);
                } else if (upper < 0.0) {
                  double m1 = (sl < sr) ? sr : sl;
                  kc = (m1 < slm) ? 5 : (sl < sr) ? 4 : 3;
                  double x0 = kc == 3 ? sl : kc == 4 ? sr : slm;
                  double gap = fmin(kc == 3 ? 1e300 : x0 - sl, fmin(kc == 4 ? 1e300 : x0 - sr, kc == 5 ? 1e300 : x0 - slm));
                  mg = fmin(-upper, gap);
                } else {
                  kc = 6;
                  mg = fmin(fabs(lower), fabs(upper));
                }
                br_cat(&bc, MRM_BR_SLOPE_POS_L, MRM_BK_SLOPE, br_key(d1, bs, offset + i * str) + v, kc,
                       br_cls(&bc, mg, fmax3(fabs(X[i - 1]), fabs(X[i]), fabs(X[i + 1]))));
              }
            }
          }
          if ((visc_slope_bot > 0.0) || (visc_slope_top > 0.0)) {
            for (int i = 1; i < sz - 1; i++) {
              double cf = (1.0 - visc_slope_bot * hfb1[i]) * (1.0 - visc_slope_top * hft1[i]);
              slp[sz + i] *= cf; slp[2 * sz + i] *= cf; slp[3 * sz + i] *= cf;
            }
          }
          if ((eta_slope_bot > 0.0) || (eta_slope_top > 0.0)) {
            for (int i = 1; i < sz - 1; i++) {
              double cf = (1.0 - eta_slope_bot * hfb1[i]) * (1.0 - eta_slope_top * hft1[i]);
              slp[5 * sz + i] *= cf; slp[6 * sz + i] *= cf; slp[7 * sz + i] *= cf;
            }
          }
          for (int i = 1; i < sz - 1; i++) {
            if (qrho[i] > q_rho_max) slp[i] = 0.0;
            if (cnt) {
              double lr0 = rho_log ? var[i] : log(var[i]), lr1 = rho_log ? var[i + 1] : log(var[i + 1]);
              br_hit(&bc, MRM_BR_QRHO, MRM_BK_QRHO, br_key(d1, bs, offset + i * str), qrho[i] > q_rho_max,
                     br_cls(&bc, fabs(qrho[i] - q_rho_max), fmax(fabs(lr0), fabs(lr1))));
            }
          }
          if (vmax_lim > 0.0) {
            for (int i = 1; i < sz - 1; i++) {
              if (vsqr[i] > vsqr_diff) { slp[sz + i] = 0.0; slp[2 * sz + i] = 0.0; slp[3 * sz + i] = 0.0; }
              if (cnt)
                br_hit(&bc, MRM_BR_VFAST, MRM_BK_VFAST, br_key(d1, bs, offset + i * str), vsqr[i] > vsqr_diff,
                       br_cls(&bc, fabs(vsqr[i] - vsqr_diff), vsqr_diff));
            }
          }
          for (int v = 0; v < NV; v++) {
            const double* X = var + v * sz;
            const double* S = slp + v * sz;
            double* F = flx + v * sz;
            const double* H = (v >= 1 && v <= 3) ? hyp_v : (v >= 5 && v <= 7) ? hyp_B : hyp_e;
            for (int i = 1; i < sz - 2; i++) {
              double sl = X[i + 1] - X[i];
              double sl_lim = X[i + 1] - X[i] - (S[i] + S[i + 1]);
              double dn = H[i];
              double rf = fabs(sl_lim) / smax(1e-100, fabs(sl));
              double cf = smax(0.0, smin(sl_lim, sl)) + smin(0.0, smax(sl_lim, sl));
              double cl = 1.0 + dn * (rf - 1.0);
              F[i] = smax(0.0, cl) * cf;
              if (cnt) {
                /* the minmod's choice and the clip */
                double mn = (sl < sl_lim) ? sl : sl_lim, mx = (sl_lim < sl) ? sl : sl_lim;
                int kc = (mn > 0.0) ? ((sl < sl_lim) ? 1 : 0) : (mx < 0.0) ? ((sl_lim < sl) ? 3 : 2) : 4;
                unsigned long long key = br_key(d1, bs, offset + i * str) + v;
                /* margins: the nearer of sl_lim, sl to zero; the clip's ratio
                   is noise where sl is within a near tie of zero */
                double sc = fmax(fabs(X[i]), fabs(X[i + 1]));
                br_cat(&bc, MRM_BR_MM_POS_LIM, MRM_BK_MM, key, kc, br_cls(&bc, fmin(fabs(mn), fabs(mx)), sc));
                int tsl = br_cls(&bc, fabs(sl), sc);
                br_hit(&bc, MRM_BR_CLIP, MRM_BK_CLIP, key, cl < 0.0,
                       tsl ? tsl : br_cls(&bc, fabs(cl), fmax(1.0, fabs(dn * (rf - 1.0)))));
              }
            }
          }
          for (int i = 1; i < sz - 2; i++) {
            double tvd_faci = 0.5 * idxd1 * smin(cmax, smax(cm[i], cm[i + 1]));
            if (cnt) {
              double m = (cm[i] < cm[i + 1]) ? cm[i + 1] : cm[i];
              int kc = (cmax < m) ? 0 : (cm[i] < cm[i + 1]) ? 2 : 1;
              double mg = kc == 0 ? m - cmax : fmin(cmax - m, fabs(cm[i] - cm[i + 1]));
              br_cat(&bc, MRM_BR_CMAX_LIMIT, MRM_BK_CMAX, br_key(d1, bs, offset + i * str), kc,
                     br_cls(&bc, mg, fmax3(cmax, cm[i], cm[i + 1])));
            }
            for (int v = 0; v < NV; v++) flx[v * sz + i] *= tvd_faci;
            flx[(5 + d1) * sz + i] *= B_par_diff;
          }
          if (need_tvd_coeff) {
            for (int i = 1; i < sz - 2; i++) {
              flx[i] *= tc0;
              double cf = smax(tc1, visc_coeff_bot * hfb1[i] + visc_coeff_top * hft1[i]);
              flx[sz + i] *= cf; flx[2 * sz + i] *= cf; flx[3 * sz + i] *= cf;
              flx[4 * sz + i] *= tc2;
              cf = smax(tc3, eta_coeff_bot * hfb1[i] + eta_coeff_top * hft1[i]);
              flx[5 * sz + i] *= cf; flx[6 * sz + i] *= cf; flx[7 * sz + i] *= cf;
            }
          }
          if ((d1 == 0) && (vhyp > 0.0)) {
            int istart = 1;
            if (is_gbeg0) istart += gh;
            for (int i = istart; i < sz - 2; i++) {
              double dn = smax(hyperdiff[i], hyperdiff[i + 1]);
              static const int hv[3] = {0, 1, 4};
              for (int q = 0; q < 3; q++) {
                const double* X = var + hv[q] * sz;
                flx[hv[q] * sz + i] -= dn * (0.125 * (X[i + 2] - X[i - 1]) - 0.375 * (X[i + 1] - X[i]));
              }
            }
          }
          for (int i = 1; i < sz - 2; i++) {
            double flx0 = flx[i], flx1 = flx[sz + i], flx2 = flx[2 * sz + i], flx3 = flx[3 * sz + i];
            double flx4 = flx[4 * sz + i];
            double uif[5];
            for (int v = 0; v < 5; v++)
              uif[v] = 0.5 * (var[v * sz + i] + slp[v * sz + i] + var[v * sz + i + 1] - slp[v * sz + i + 1]);
            double uif0 = uif[0], uif1 = uif[1], uif2 = uif[2], uif3 = uif[3], uif4 = uif[4];
            if (rho_log) {
              uif0 = exp(uif0);
              flx0 *= uif0;
              uif4 = exp(uif4);
              flx4 *= uif4;
            }
            double min_cz = flx0 * smin(CZ[i], CZ[i + 1]);
            flx1 *= uif0; flx1 += uif1 * min_cz;
            flx2 *= uif0; flx2 += uif2 * min_cz;
            flx3 *= uif0; flx3 += uif3 * min_cz;
            flx4 *= uif0;
            flx4 += (uif4 * min_cz) + qft1[i] * (uif1 * flx1 + uif2 * flx2 + uif3 * flx3);
            flx[i] = flx0; flx[sz + i] = flx1; flx[2 * sz + i] = flx2; flx[3 * sz + i] = flx3;
            flx[4 * sz + i] = flx4;
          }
          if (d1 == 0) {
            if (is_gend0) flx[5 * sz + sz - gh - 1] = 0.0;
            if (is_gbeg0) flx[5 * sz + gh - 1] = 0.0;
          }
          /* resistive-heating interface term (reuses qrho, :609-614) */
          for (int i = 1; i < sz - 2; i++)
            qrho[i] = (var[5 * sz + i + 1] - var[5 * sz + i]) * flx[5 * sz + i] +
                      (var[6 * sz + i + 1] - var[6 * sz + i]) * flx[6 * sz + i] +
                      (var[7 * sz + i + 1] - var[7 * sz + i]) * flx[7 * sz + i];
          /* viscous-heating interface term for Qvis (:627-634), kept in hyp_e
             which is no longer needed */
          if (dg)
            for (int i = 1; i < sz - 2; i++)
              hyp_e[i] = (var[1 * sz + i + 1] - var[1 * sz + i]) * flx[1 * sz + i] +
                         (var[2 * sz + i + 1] - var[2 * sz + i]) * flx[2 * sz + i] +
                         (var[3 * sz + i + 1] - var[3 * sz + i]) * flx[3 * sz + i];
          for (int i = gh; i < sz - gh; i++) {
            long node = offset + i * str;
            double r[NV];
            for (int v = 0; v < NV; v++) r[v] = dt_tvd * (flx[v * sz + i] - flx[v * sz + i - 1]);
            double v1 = var[sz + i], v2 = var[2 * sz + i], v3 = var[3 * sz + i];
            double b5 = var[5 * sz + i], b6 = var[6 * sz + i], b7 = var[7 * sz + i];
            r[4] += (1.0 - qft1[i]) * (r[1] * v1 + r[2] * v2 + r[3] * v3);
            if (dg) dg[DG_T8][node] += r[4] * idt_full;
            double dn = (r[1] * b5 + r[2] * b6 + r[3] * b7) / smax(1e-100, bsqr[i]);
            double boris0 = -BC2[i] * (r[1] - dn * b5);
            double boris1 = -BC2[i] * (r[2] - dn * b6);
            double boris2 = -BC2[i] * (r[3] - dn * b7);
            double boris3 = boris0 * v1 + boris1 * v2 + boris2 * v3;
            r[1] += boris0; r[2] += boris1; r[3] += boris2; r[4] += boris3;
            double qres = 0.5 * (qrho[i] + qrho[i - 1]) * qft1[i];
            r[4] += dt_tvd * qres;
            for (int v = 0; v < NV; v++) U[v * bs + node] += r[v];
            if (cnt)
              br_hit(&bc, MRM_BR_RHO_FLOOR, MRM_BK_RHO_FLOOR, br_key(d1, bs, node), U[node] < rho_min,
                     br_cls(&bc, fabs(U[node] - rho_min), rho_min));
            U[node] = smax(rho_min, U[node]);
            double mx = U[bs + node], my = U[2 * bs + node], mz = U[3 * bs + node];
            const double e_pre = U[4 * bs + node];
            const double efl = eps_min * U[node] + 0.5 * (mx * mx + my * my + mz * mz) / U[node];
            if (cnt)
              br_hit(&bc, MRM_BR_E_FLOOR, MRM_BK_E_FLOOR, br_key(d1, bs, node), e_pre < efl,
                     br_cls(&bc, fabs(e_pre - efl), fmax(fabs(e_pre), fabs(efl))));
            U[4 * bs + node] = smax(U[4 * bs + node], efl);
            if (dg) {
              dg[DG_T6][node] += boris3 * idt_full;
              dg[DG_QRES][node] += dt_fac * qres;
              dg[DG_T7][node] += (U[4 * bs + node] - e_pre) * idt_full;
              dg[DG_QVIS][node] += dt_fac * (0.5 * (hyp_e[i] + hyp_e[i - 1]) * qft1[i]);
            }
          }
        }
      free(mem);
      if (c->bcount_on) {
#pragma omp critical(orc_bcount)
        {
          for (int q = 0; q < MRM_BR_NCOUNT; q++) c->bcount[q] += bc.n[q];
          for (int q = 0; q < MRM_BK_N; q++) {
            c->bcount[MRM_BR_NCOUNT + q] += bc.h[q];
            c->bcount[MRM_BR_NCOUNT + MRM_BK_N + q] += bc.nnear[q];
            c->bcount[MRM_BR_NCOUNT + 2 * MRM_BK_N + q] += bc.nexact[q];
          }
        }
      }
    }
    bounds[d1][0] = c->lbeg[d1];
    bounds[d1][1] = c->lend[d1];
  }
}

/* ------------------------------------------------------------------------ */
/* exchange_grid_acc (exchange.C:532-782) on one rank: dims in order 0,1,2;  */
/* a periodic dim copies the two opposite interior layers into its ghosts,   */
/* and the transverse range of later dims grows by the filled ghosts.        */
void orc_exchange_grid(orc_ctx* c) {
  static const int perm[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
  const long bs = c->bufsize;
  int gs[3][2];
  for (int d = 0; d < 3; d++) { gs[d][0] = c->lbeg[d]; gs[d][1] = c->lend[d]; }
  for (int dim = 0; dim < 3; dim++) {
    const int d1 = perm[dim][0], d2 = perm[dim][1], d3 = perm[dim][2];
    if (!c->periods[d1]) continue;   /* single rank: the non-periodic ends get no halo */
    const int gh = c->ghosts[d1], lbeg = c->lbeg[d1], lend = c->lend[d1];
    for (int igh = 0; igh < gh; igh++)
      for (int k = gs[d3][0]; k <= gs[d3][1]; k++)
        for (int j = gs[d2][0]; j <= gs[d2][1]; j++) {
          long base = j * c->stride[d2] + k * c->stride[d3];
          long src_r = base + (long)(lbeg + igh) * c->stride[d1];          /* -> right ghost */
          long dst_r = base + (long)(lend + 1 + igh) * c->stride[d1];
          long src_l = base + (long)(lend - gh + 1 + igh) * c->stride[d1]; /* -> left ghost */
          long dst_l = base + (long)(lbeg - gh + igh) * c->stride[d1];
          for (int v = 0; v < NV; v++) {
            c->U[v * bs + dst_r] = c->U[v * bs + src_r];
            c->U[v * bs + dst_l] = c->U[v * bs + src_l];
          }
          if (c->Rflx) {   /* sflx travels with U (exchange.C:111-120) */
            double* sf = c->fld[MRM_SFLX];
            sf[dst_r] = sf[src_r];
            sf[dst_l] = sf[src_l];
          }
        }
    gs[d1][1] += gh;
    gs[d1][0] -= gh;
  }
}

/* ------------------------------------------------------------------------ */
/* exchange_B_acc / exchange_single_acc (exchange.C:785-1120) on one rank:   */
/* the periodic fill of exchange_grid_acc for arrays a[0..na) only.          */
static void exchange_arrays(orc_ctx* c, double* const* a, int na) {
  static const int perm[3][3] = {{0, 1, 2}, {1, 0, 2}, {2, 0, 1}};
  int gs[3][2];
  for (int d = 0; d < 3; d++) { gs[d][0] = c->lbeg[d]; gs[d][1] = c->lend[d]; }
  for (int dim = 0; dim < 3; dim++) {
    const int d1 = perm[dim][0], d2 = perm[dim][1], d3 = perm[dim][2];
    if (!c->periods[d1]) continue;
    const int gh = c->ghosts[d1], lbeg = c->lbeg[d1], lend = c->lend[d1];
    for (int igh = 0; igh < gh; igh++)
      for (int k = gs[d3][0]; k <= gs[d3][1]; k++)
        for (int j = gs[d2][0]; j <= gs[d2][1]; j++) {
          long base = j * c->stride[d2] + k * c->stride[d3];
          long src_r = base + (long)(lbeg + igh) * c->stride[d1];
          long dst_r = base + (long)(lend + 1 + igh) * c->stride[d1];
          long src_l = base + (long)(lend - gh + 1 + igh) * c->stride[d1];
          long dst_l = base + (long)(lbeg - gh + igh) * c->stride[d1];
          for (int v = 0; v < na; v++) {
            a[v][dst_r] = a[v][src_r];
            a[v][dst_l] = a[v][src_l];
          }
        }
    gs[d1][1] += gh;
    gs[d1][0] -= gh;
  }
}

/* 4th-order div B of the interior into divB; returns max |div B|            */
/* (clean_divB.C:125-162)                                                    */
static double divb_interior(orc_ctx* c) {
  const long bs = c->bufsize;
  const double* U = c->U;
  double* divB = c->fld[MRM_DIVB];
  double err = 0.0;
  for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
    for (int j = c->lbeg[1]; j <= c->lend[1]; j++) {
      const long off0 = j * c->stride[1] + k * c->stride[2];
      for (int i = c->lbeg[0]; i <= c->lend[0]; i++) divB[off0 + i] = 0.0;
      for (int v = 0; v < 3; v++) {
        const double ww1 = 8. / (12. * c->dx[v]), ww2 = -1. / (12. * c->dx[v]);
        const long s = c->stride[v];
        const double* B = U + (5 + v) * bs;
        for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
          const long n = off0 + i;
          divB[n] += (ww1 * (B[n + s] - B[n - s]) + ww2 * (B[n + 2 * s] - B[n - 2 * s]));
        }
      }
      for (int i = c->lbeg[0]; i <= c->lend[0]; i++) err = smax(err, fabs(divB[off0 + i]));
    }
  return err;
}

/* Clean_divB (clean_divB.C:35-369) on one rank: a damped potential phi,     */
/* kept from call to call, is driven by the 4th-order div B and its gradient  */
/* is added to B (its work phi div B to e); the number of passes adapts to   */
/* how much each call reduces max |div B|.  err = {before, after} in the     */
/* reference's units.                                                        */
void orc_clean_divb(orc_ctx* c, const mrm_run_desc* run, double err_out[2]) {
  const long bs = c->bufsize;
  double** dg = diag_of(c, run);   /* Qres += phi divB / dt (clean_divB.C:264-270) */
  const double invdt = dg ? 1.0 / run->dt : 0.0;
  const int it_max = (int)c->ph.divB[1];
  const double divB_err = c->ph.divB[2];
  const double divB_coeff = 1.0, divB_damp = 0.65;
  double err_unit = 0.0;
  for (int v = 0; v < 3; v++) err_unit += 1.0 / c->dx[v];
  err_unit = 0.5 / err_unit * 3.5449;
  double dxmin = smin(c->dx[0], c->dx[1]);
  double dxmax = smax(c->dx[0], c->dx[1]);
  dxmin = smin(dxmin, c->dx[2]);
  dxmax = smax(dxmin, c->dx[2]);   /* clean_divB.C:95-98 compares dx[2] with dxmin */
  const double divB_fac = divB_coeff * smin(dxmin, 0.4 * dxmax) * dxmin;
  double* phi = c->fld[MRM_PHI];
  const double* divB = c->fld[MRM_DIVB];
  double* Bv[3] = {c->U + 5 * bs, c->U + 6 * bs, c->U + 7 * bs};

  exchange_arrays(c, Bv, 3);
  const double err0 = divb_interior(c);
  double err = err0;
  int iter;
  for (iter = 0; iter < c->divb_next_iter; iter++) {
    for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
      for (int j = c->lbeg[1]; j <= c->lend[1]; j++) {
        const long off0 = j * c->stride[1] + k * c->stride[2];
        for (int i = c->lbeg[0]; i <= c->lend[0]; i++)
          phi[off0 + i] += divB_fac * divB[off0 + i] - divB_damp * phi[off0 + i];
      }
    exchange_arrays(c, &phi, 1);
    /* vertical boundaries: phi antisymmetric (clean_divB.C:186-217) */
    for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
      for (int j = c->lbeg[1]; j <= c->lend[1]; j++) {
        const long off0 = j * c->stride[1] + k * c->stride[2];
        const int lb = c->lbeg[0], le = c->lend[0];
        phi[off0 + lb - 1] = -phi[off0 + lb];
        phi[off0 + lb - 2] = -phi[off0 + lb + 1];
        phi[off0 + le + 1] = -phi[off0 + le];
        phi[off0 + le + 2] = -phi[off0 + le - 1];
      }
    for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
      for (int j = c->lbeg[1]; j <= c->lend[1]; j++) {
        const long off0 = j * c->stride[1] + k * c->stride[2];
        for (int v = 0; v < 3; v++) {
          const double ww1 = 8. / (12. * c->dx[v]), ww2 = -1. / (12. * c->dx[v]);
          const long s = c->stride[v];
          for (int i = c->lbeg[0]; i <= c->lend[0]; i++) {
            const long n = off0 + i;
            Bv[v][n] += (ww1 * (phi[n + s] - phi[n - s]) + ww2 * (phi[n + 2 * s] - phi[n - 2 * s]));
          }
        }
        for (int i = c->lbeg[0]; i <= c->lend[0]; i++) c->U[4 * bs + off0 + i] += phi[off0 + i] * divB[off0 + i];
        if (dg)
          for (int i = c->lbeg[0]; i <= c->lend[0]; i++) dg[DG_QRES][off0 + i] += phi[off0 + i] * divB[off0 + i] * invdt;
      }
    exchange_arrays(c, Bv, 3);
    err = divb_interior(c);
  }
  const double e0 = err0 * err_unit, e1 = err * err_unit;
  const double ratio = e1 / e0;
  if (ratio > divB_err) c->divb_next_iter += 1;
  else c->divb_next_iter -= 1;
  c->divb_next_iter = imin(c->divb_next_iter, it_max);
  c->divb_next_iter = imax(1, c->divb_next_iter);
  if (err_out) { err_out[0] = e0; err_out[1] = e1; }
}

/* ------------------------------------------------------------------------ */
/* set_p_s_bc (boundary_pdmp_1_fftw3.C:496-565) on one rank                  */
static void set_p_s_bc(orc_ctx* c) {
  const long bs = c->bufsize;
  const double* U = c->U;
  double bnd0 = 0, bnd1 = 0, bnd2 = 0, bnd3 = 0, bnd4 = 0;
  for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
    for (int i = c->lbeg[1]; i <= c->lend[1]; i++) {
      long n1 = c->lbeg[0] + i * c->stride[1] + k * c->stride[2];
      long n2 = n1 + 1;
      double d1 = U[n1], d2 = U[n2];
      double m1x = U[bs + n1], m1y = U[2 * bs + n1], m1z = U[3 * bs + n1];
      double m2x = U[bs + n2], m2y = U[2 * bs + n2], m2z = U[3 * bs + n2];
      double e1 = U[4 * bs + n1] - 0.5 * (m1x * m1x + m1y * m1y + m1z * m1z) / d1;
      double e2 = U[4 * bs + n2] - 0.5 * (m2x * m2x + m2y * m2y + m2z * m2z) / d2;
      double eps1 = e1 / d1, eps2 = e2 / d2;
      bnd0 += orc_p_interp(c, eps1, d1);
      bnd1 += orc_p_interp(c, eps2, d2);
      if (m1x > 0.0) {
        bnd2 += 1.0;
        bnd3 += orc_s_interp(c, eps1, d1);
      }
      bnd4 += orc_s_interp(c, eps1, d1);
    }
  double nh = (double)(c->gsize[1] * c->gsize[2]);
  bnd0 /= nh;
  bnd1 /= nh;
  if (c->p_bc < 0.0) c->p_bc = bnd0 * sqrt(bnd0 / bnd1);
  if (c->s_bc < 0.0) {
    if (bnd2 > 0.0) c->s_bc = bnd3 / bnd2;
    else c->s_bc = bnd4 / nh;
  }
}

/* ---- potential-field extrapolation of the top boundary, one rank.
   b_ext[q] = the circular convolution of bz0 = B_x(lend) over the
   horizontal plane with kernel q (potential_sd_fftw3.C:178-216 forms it as
   IFFT(FFT(bz0) FFT(K_q) / (gx gy)); here the sum is taken directly),
   written into the interior of plane q, then periodic ghosts in dim 1
   (:218-228) and in dim 2 over the whole dim-1 range (y_exchange,
   :405-486). */
int orc_potential_kernels(orc_ctx* c, const float* k8) {
  const size_t n = (size_t)c->gsize[1] * c->gsize[2];
  if (!c->pkern) c->pkern = (float*)malloc(sizeof(float) * 8 * n);
  if (!c->bext) c->bext = (double*)calloc((size_t)8 * c->n[1] * c->n[2], sizeof(double));
  if (!c->pkern || !c->bext) return -1;
  memcpy(c->pkern, k8, sizeof(float) * 8 * n);
  c->bext_valid = 0;
  return 0;
}

const double* orc_potential_bext(const orc_ctx* c) { return c->bext; }

static void potential_update(orc_ctx* c) {
  const int gx = c->gsize[1], gy = c->gsize[2], g1 = c->ghosts[1], g2 = c->ghosts[2];
  const int n1 = c->n[1], n2 = c->n[2];
  const long bs = c->bufsize;
  double* bz = (double*)malloc(sizeof(double) * (size_t)gx * gy);
  /* bz0[j + k lsize1] = B_x at lend (boundary_pdmp_1_fftw3.C:339-349) */
  for (int k = 0; k < gy; k++)
    for (int j = 0; j < gx; j++)
      bz[j + (size_t)k * gx] = c->U[5 * bs + c->lend[0] + (j + c->lbeg[1]) * c->stride[1] + (k + c->lbeg[2]) * c->stride[2]];
#pragma omp parallel for collapse(3) schedule(static)
  for (int q = 0; q < 8; q++)
    for (int y = 0; y < gy; y++)
      for (int x = 0; x < gx; x++) {
        const float* K = c->pkern + (size_t)q * gx * gy;
        double sum = 0.0;
        for (int m2 = 0; m2 < gy; m2++) {
          const int yy = (y - m2 + gy) % gy;
          for (int m1 = 0; m1 < gx; m1++) {
            const int xx = (x - m1 + gx) % gx;
            sum += bz[m1 + (size_t)m2 * gx] * (double)K[xx + (size_t)yy * gx];
          }
        }
        c->bext[(size_t)q * n1 * n2 + (size_t)(y + g2) * n1 + (x + g1)] = sum;
      }
  for (int q = 0; q < 8; q++) {
    double* b = c->bext + (size_t)q * n1 * n2;
    for (int y = g2; y < gy + g2; y++)
      for (int i = 0; i < g1; i++) {
        b[(size_t)y * n1 + i] = b[(size_t)y * n1 + i + gx];
        b[(size_t)y * n1 + i + gx + g1] = b[(size_t)y * n1 + i + g1];
      }
    for (int j = 0; j < g2; j++)
      for (int i = 0; i < n1; i++) {
        b[(size_t)(j + gy + g2) * n1 + i] = b[(size_t)(j + g2) * n1 + i];
        b[(size_t)j * n1 + i] = b[(size_t)(j + gy) * n1 + i];
      }
  }
  free(bz);
  c->bext_valid = 1;
}

/* pt_update of the end-of-step call (solver.C:355-362) */
static int pt_update_rule(const orc_ctx* c, const mrm_run_desc* run, int stage) {
  const int pot = (int)c->ph.bnd[1];
  if (stage == 0) return 1;                       /* solver.C:81 */
  if (stage <= MAXSTAGE) return 0;                /* solver.C:320 */
  if (pot <= 2) return 1;
  return !(run->iteration % (pot - 1));
}

/* SetBoundaryConditions (boundary_pdmp_1_fftw3.C:81-493) with pt_update as
   solver.C passes it */
void orc_set_boundary_conditions(orc_ctx* c, const mrm_run_desc* run, int stage) {
  orc_set_boundary_conditions_pt(c, run, stage, pt_update_rule(c, run, stage));
}

void orc_set_boundary_conditions_pt(orc_ctx* c, const mrm_run_desc* run, int stage, int pt_update) {
  (void)run; (void)stage;
  const long bs = c->bufsize;
  double* U = c->U;
  const int OPEN_TOP = (int)c->ph.bnd[0];
  const double B_crit = c->ph.bnd[2] / sqrt(8.0 * asin(1.0));
  const double eps_top = c->ph.bnd[3];
  const double p1_dmp = 0.8;
  if (c->bnd_ini) {
    if ((c->p_bc < 0.0) || (c->s_bc < 0.0)) set_p_s_bc(c);
    c->bnd_ini = 0;
  }
  /* bottom: mean pressures of the two lowest interior layers (:172-226) */
  double lbuf5 = 0, lbuf6 = 0;
  for (int k = c->lbeg[2]; k <= c->lend[2]; k++)
    for (int j = c->lbeg[1]; j <= c->lend[1]; j++) {
      long n1 = c->lbeg[0] + j * c->stride[1] + k * c->stride[2];
      long n2 = n1 + 1;
      double d1 = U[n1], d2 = U[n2];
      double m1x = U[bs + n1], m1y = U[2 * bs + n1], m1z = U[3 * bs + n1];
      double m2x = U[bs + n2], m2y = U[2 * bs + n2], m2z = U[3 * bs + n2];
      double e1 = U[4 * bs + n1] - 0.5 * ((m1x * m1x + m1y * m1y + m1z * m1z) / d1);
      double e2 = U[4 * bs + n2] - 0.5 * ((m2x * m2x + m2y * m2y + m2z * m2z) / d2);
      lbuf5 += orc_p_interp(c, e1 / d1, d1);
      lbuf6 += orc_p_interp(c, e2 / d2, d2);
    }
  const double cc = 1.0 / ((double)(c->gsize[1] * c->gsize[2]));
  const double pmean1 = lbuf5 * cc, pmean2 = lbuf6 * cc;
  const double p_bc = c->p_bc, s_bc = c->s_bc;
  /* bottom ghost layers over the full transverse range (:240-318) */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = c->lbeg[2] - c->ghosts[2]; k <= c->lend[2] + c->ghosts[2]; k++)
    for (int j = c->lbeg[1] - c->ghosts[1]; j <= c->lend[1] + c->ghosts[1]; j++) {
      long base = j * c->stride[1] + k * c->stride[2];
      long n1 = c->lbeg[0] + base, n2 = n1 + 1;
      double u1[NV], u2[NV];
      for (int v = 0; v < NV; v++) { u1[v] = U[v * bs + n1]; u2[v] = U[v * bs + n2]; }
      u1[4] -= 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3]) / u1[0];
      u2[4] -= 0.5 * (u2[1] * u2[1] + u2[2] * u2[2] + u2[3] * u2[3]) / u2[0];
      double eps1 = u1[4] / u1[0];
      double eps2 = u2[4] / u2[0];
      double pmag = 0.5 * u1[5] * u1[5];
      double pmin1 = orc_p_interp(c, eps1, u1[0]) - pmean1;
      double pmin2 = orc_p_interp(c, eps2, u2[0]) - pmean2;
      (void)pmin2;
      double smin1 = orc_s_interp(c, eps1, u1[0]);
      double smin2 = orc_s_interp(c, eps2, u2[0]);
      double s1, s2;
      if (u1[1] >= 0.e0) { s1 = s_bc; s2 = s_bc; }
      else { s1 = smin1; s2 = smin2; }
      double cfac = p_bc / sqrt(pmean1 * pmean2);
      double p1 = pmean1 * cfac;
      double p2 = pmean1 * cfac * cfac;
      p1 += pmin1 * p1_dmp;
      p2 += pmin1 * p1_dmp * p1_dmp;
      p1 -= pmag * (1.0 - p1_dmp);
      p2 -= pmag * (1.0 - p1_dmp * p1_dmp);
      double cs = u1[5] / B_crit;
      cs = pow(cs, 4);
      cs = (1.0 - cs) / (1.0 + cs);
      for (int v = 1; v <= 3; v++) { u1[v] *= cs; u2[v] *= cs; }
      double rho1 = orc_d3_interp(c, p1, s1);
      double rho2 = orc_d3_interp(c, p2, s2);
      eps1 = orc_eps3_interp(c, p1, s1);
      eps2 = orc_eps3_interp(c, p2, s2);
      u1[0] = rho1;
      u2[0] = rho2;
      u1[4] = eps1 * rho1 + 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3]) / u1[0];
      u2[4] = eps2 * rho2 + 0.5 * (u2[1] * u2[1] + u2[2] * u2[2] + u2[3] * u2[3]) / u2[0];
      long g1 = c->lbeg[0] - 1 + base, g2 = c->lbeg[0] - 2 + base;
      for (int v = 0; v < NV; v++) { U[v * bs + g1] = u1[v]; U[v * bs + g2] = u2[v]; }
      if (c->Rflx) { c->fld[MRM_SFLX][g1] = 0.0; c->fld[MRM_SFLX][g2] = 0.0; }   /* :306-310 */
    }
  /* potential field (:335-381): b_ext recomputed on every call with
     bnd_pot = 1, on calls with pt_update with bnd_pot > 1 */
  const int POT = (int)c->ph.bnd[1];
  const int pot_on = POT > 0 && c->bext && c->pkern;
  if (pot_on && ((POT == 1) || (POT > 1 && pt_update == 1) || !c->bext_valid)) potential_update(c);
  const long lgs = (long)c->n[1] * c->n[2];
  /* top ghost layers (:402-474): vertical field (bnd_pot = 0, B_y and B_z
     mirrored) or the potential field of b_ext, which also sets B_y, B_z of
     the top interior layer */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = c->lbeg[2] - c->ghosts[2]; k <= c->lend[2] + c->ghosts[2]; k++)
    for (int j = c->lbeg[1] - c->ghosts[1]; j <= c->lend[1] + c->ghosts[1]; j++) {
      long base = j * c->stride[1] + k * c->stride[2];
      long n1 = c->lend[0] + base, n2 = n1 - 1;
      double u1[NV], u2[NV];
      for (int v = 0; v < NV; v++) { u1[v] = U[v * bs + n1]; u2[v] = U[v * bs + n2]; }
      if (OPEN_TOP == 1) {
        if (u1[1] < 0.0) { u1[1] *= -1; u2[1] *= -1; }
      } else {
        u1[1] *= -1; u2[1] *= -1;
      }
      if (eps_top > 0.0) {
        u1[4] = eps_top * u1[0] + 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3]) / u1[0];
        u2[4] = eps_top * u2[0] + 0.5 * (u2[1] * u2[1] + u2[2] * u2[2] + u2[3] * u2[3]) / u2[0];
      }
      if (pot_on) {
        const long il = j + (long)k * c->n[1];
        const double* b = c->bext;
        U[6 * bs + n1] = b[il + 2 * lgs];
        U[7 * bs + n1] = b[il + 5 * lgs];
        u1[5] = b[il + 0 * lgs]; u1[6] = b[il + 3 * lgs]; u1[7] = b[il + 6 * lgs];
        u2[5] = b[il + 1 * lgs]; u2[6] = b[il + 4 * lgs]; u2[7] = b[il + 7 * lgs];
      } else {
        u1[6] *= -1; u2[6] *= -1; u1[7] *= -1; u2[7] *= -1;
      }
      long g1 = c->lend[0] + 1 + base, g2 = c->lend[0] + 2 + base;
      for (int v = 0; v < NV; v++) { U[v * bs + g1] = u1[v]; U[v * bs + g2] = u2[v]; }
      if (c->Rflx) {   /* heat flux (:455-466): symmetric under a hot plate, else zero */
        double* sf = c->fld[MRM_SFLX];
        if (eps_top > 0.0) {
          sf[g1] = sf[n1];
          sf[g2] = sf[n2];
        } else {
          sf[n2] = 0.0;
          sf[n1] = 0.0;
          sf[g1] = 0.0;
          sf[g2] = 0.0;
        }
      }
    }
}

/* ------------------------------------------------------------------------ */
/* One ComputeSolution iteration (solver.C:115-365), RT and IO off          */
/* solver.C:102-380; with run->need_diagnostics the last step collects the  */
/* diagnostics from its stage 4 on, zeroed first (solver.C:264-296).         */
int orc_step(orc_ctx* c, mrm_run_desc* run, double dt_rad, int nsteps) {
  const int want_diag = run->need_diagnostics;
  for (int it = 0; it < nsteps; it++) {
    double cfl = 0.0;
    orc_save_cons(c);
    for (int stage = 1; stage <= MAXSTAGE; stage++) {
      orc_cons_to_prim(c);
      if (stage == 1) orc_adjust_valf_max(c, run, NULL);
      run->need_diagnostics = want_diag && (it == nsteps - 1) && (stage == MAXSTAGE);
      if (run->need_diagnostics) orc_diag_zero(c);
      double dt_mhd = orc_mhd_residual(c, run);
      if (stage == 1) {
        int rc = orc_sync_timestep(c, run, dt_rad, dt_mhd, &cfl);
        if (rc) {
          run->need_diagnostics = want_diag;
          return rc;
        }
      }
      orc_source_integrate_tcheck(c, run, stage);
      orc_exchange_grid(c);
      orc_set_boundary_conditions(c, run, stage);
    }
    if (cfl <= run->CFL_tvd) {
      orc_tvdlimit(c, run, run->dt);
    } else {
      orc_tvdlimit(c, run, 0.5 * run->dt);
      orc_exchange_grid(c);
      orc_tvdlimit(c, run, 0.5 * run->dt);
    }
    if (c->ph.divB[0] > 0) orc_clean_divb(c, run, NULL);
    orc_tcheck(c, run);
    orc_exchange_grid(c);
    orc_set_boundary_conditions(c, run, MAXSTAGE + 1);
    run->iteration++;
    run->globiter++;
    run->time += run->dt;
  }
  run->need_diagnostics = want_diag;
  return MRM_OK;
}
