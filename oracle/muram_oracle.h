/*
 * muram_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C11, FP64, OpenMP over grid lines) of MURaM's
 * per-timestep MHD update, used as the parity checker of the HIP path and as
 * the CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it; the product never does.
 *
 * State layout is structure-of-arrays: U[v][node], v in cState order
 * (rho, M0, M1, M2, e, B0, B1, B2), node = i + n0*(j + n1*k), dim 0 vertical
 * and fastest -- the reference's node order (grid.C:196-199) with the
 * components split out of the AoS cState.
 */
#ifndef MURAM_ORACLE_H
#define MURAM_ORACLE_H

#include "../include/muram_step.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

orc_ctx* orc_create(const mrm_grid_desc* g, const mrm_phys_desc* p);
void orc_destroy(orc_ctx* c);
long orc_bufsize(const orc_ctx* c);
int orc_eos_load(orc_ctx* c, const void* bytes, size_t nbytes);

/* direct access to the arrays (SoA, 8*bufsize doubles for U/U0/Res) */
double* orc_state(orc_ctx* c, int which);
double* orc_field(orc_ctx* c, int field);
void orc_get_limits(const orc_ctx* c, double lim[4]);
void orc_set_limits(orc_ctx* c, const double lim[4]);
void orc_get_bc(const orc_ctx* c, double* p_bc, double* s_bc);
void orc_set_bc(orc_ctx* c, double p_bc, double s_bc);
void orc_set_threads(int n);

void   orc_cons_to_prim(orc_ctx* c);
void   orc_adjust_valf_max(orc_ctx* c, const mrm_run_desc* run, double stats[5]);
double orc_mhd_residual(orc_ctx* c, const mrm_run_desc* run);
int    orc_sync_timestep(orc_ctx* c, mrm_run_desc* run, double dt_rad, double dt_mhd, double* cfl);
void   orc_source_integrate_tcheck(orc_ctx* c, const mrm_run_desc* run, int stage);
void   orc_save_cons(orc_ctx* c);
void   orc_tcheck(orc_ctx* c, const mrm_run_desc* run);
void   orc_tvdlimit(orc_ctx* c, const mrm_run_desc* run, double dt_tvd);
void   orc_exchange_grid(orc_ctx* c);
void   orc_set_boundary_conditions(orc_ctx* c, const mrm_run_desc* run, int stage);
/* with the reference's pt_update argument (the potential-field update);
   orc_set_boundary_conditions takes it from the stage as solver.C passes it */
void   orc_set_boundary_conditions_pt(orc_ctx* c, const mrm_run_desc* run, int stage, int pt_update);
/* potential-field kernels: 8 x gsize[2] x gsize[1] floats in the order
   z-1, z-2, x-0, x-1, x-2, y-0, y-1, y-2 (the PSF-kernel-*.dat files without
   their headers); b_ext: 8 planes of n1 x n2 doubles */
int    orc_potential_kernels(orc_ctx* c, const float* k8);
const double* orc_potential_bext(const orc_ctx* c);
void   orc_clean_divb(orc_ctx* c, const mrm_run_desc* run, double err[2]);
int    orc_step(orc_ctx* c, mrm_run_desc* run, double dt_rad, int nsteps);
/* diagnostics tvar1..8, Qres, Qvis (t = 0..9): zeroed by orc_diag_zero,
   added to by every operation called with run->need_diagnostics */
void   orc_diag_zero(orc_ctx* c);
double* orc_diag(orc_ctx* c, int t);
/* limiter branch accounting with the library's categories, checksums and
   near-tie rule (MRM_BR_* in muram_step.h): on = 1 zeroes and starts, 0 stops */
void   orc_branch_count(orc_ctx* c, int on, double tie);
void   orc_get_branch_counts(const orc_ctx* c, unsigned long long out[MRM_BR_NSLOT]);

/* scalar EOS interpolations of eos_init.C:216-416 */
double orc_p_interp(const orc_ctx* c, double ee, double dd);
double orc_s_interp(const orc_ctx* c, double ee, double dd);
double orc_T_interp(const orc_ctx* c, double ee, double dd);
double orc_d3_interp(const orc_ctx* c, double pp, double ss);
double orc_eps3_interp(const orc_ctx* c, double pp, double ss);

#ifdef __cplusplus
}
#endif
#endif
