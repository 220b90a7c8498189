/*
 * muram_step.h -- C ABI of the MI355X-native MURaM per-timestep MHD update.
 *
 * One opaque context per MURaM GridData (one GPU, one HIP stream).  The grid
 * state lives in HBM as structure-of-arrays (one array per cState component,
 * dim 0 fastest, the reference's node order); host AoS <-> device SoA
 * conversion happens only in mrm_set_state / mrm_get_state, which are the
 * counterparts of ACCH::UpdateGPU / ACCH::UpdateCPU (include/ACCH.h:49-50).
 *
 * Every function returns MRM_OK (0) or a negative MRM_ERR_* code; the message
 * of the last failure is available from mrm_last_error().  The reference aborts
 * the MPI job on the same conditions (solver.C:497-507, :529-534,
 * eos_init.C:74-76); the C++ shim maps the codes back onto MPI_Abort.
 * Calls on one context are not re-entrant; all device work of a context is
 * ordered on its stream.
 */
#ifndef MURAM_STEP_H
#define MURAM_STEP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRM_OK               0
#define MRM_ERR_ARG         -1   /* bad argument or unsupported parameter */
#define MRM_ERR_HIP         -2   /* HIP runtime failure (message has details) */
#define MRM_ERR_EOS         -3   /* EOS table missing or malformed (eos_init.C:74-76) */
#define MRM_ERR_INVALID_DT  -4   /* dt NaN, <= 0 or inf (solver.C:497-507) */
#define MRM_ERR_CFL         -5   /* cfl >= 2 (solver.C:529-534) */
#define MRM_ERR_STATE       -6   /* call order violated (e.g. no EOS loaded) */
#define MRM_ERR_UNSUPPORTED -7   /* physics option outside the implemented path */

/* storage precision of the grid state */
#define MRM_PREC_FP64        0   /* plain method: everything in FP64 */
#define MRM_PREC_FP32_MIXED  1   /* FP32 state/stencils; FP64 energy, e_int recovery, dt */

/* which state array (GridData::U, U0, Res; grid.H:52-55) */
#define MRM_U    0
#define MRM_U0   1
#define MRM_RES  2

/* scalar cell fields (grid.H:58-68) */
#define MRM_PRES 0
#define MRM_TEMP 1
#define MRM_NE   2
#define MRM_RHOI 3
#define MRM_AMB  4
#define MRM_QTOT 5
#define MRM_TAU  6
#define MRM_DIVB 7   /* Clean_divB's divergence and potential (grid.H:78-80) */
#define MRM_PHI  8
#define MRM_SFLX 9   /* Spitzer heat flux along B (grid.H:71, result_prim_8) */
#define MRM_NFIELD 10
/* diagnostics arrays (grid.H:69-70, 93-100), read and written through
 * mrm_get_field / mrm_set_field: tvar1..tvar8, Qres, Qvis */
#define MRM_TVAR1 10
#define MRM_TVAR8 17
#define MRM_QRES 18
#define MRM_QVIS 19

/* host array layouts for mrm_set_state / mrm_get_state */
#define MRM_LAYOUT_AOS 0   /* cState[bufsize]: 8 doubles per cell (mhd3d.H:24-27) */
#define MRM_LAYOUT_SOA 1   /* double[8][bufsize] */

/* Geometry: GridData fields set from parameters.dat (init.C:66-93). */
typedef struct mrm_grid_desc {
  int    gsize[3];    /* interior cells per dimension; dim 0 is vertical */
  int    ghosts[3];   /* ghost layers per side (2 in every config) */
  int    periods[3];  /* 1 = periodic; the bottom/top BC acts on dim 0 */
  double gxmin[3];
  double gxmax[3];
} mrm_grid_desc;

/* PhysicsData (physics.H:80-206), the members the hot path reads, with the
 * reference's index constants (i_param_*, i_bnd_*, i_tvd_*, i_tchk_*, ...). */
typedef struct mrm_phys_desc {
  double params[9];
  double bnd[5];
  double tvd_h[4];
  double tvd_cs[4];
  double tvd[10];
  double tvd_h_bnd[2];
  double tvd_visc_bnd[2];
  double tvd_eta_bnd[2];
  double tvd_coeff[4];
  double tchk[4];
  double dmp[4];
  double rt[10];
  double divB[3];     /* i_divB_switch, i_divB_itmax, i_divB_err (physics.H:43-45) */
  int    rt_ext_cor;
  int    pad_;
} mrm_phys_desc;

/* RunData (run.H:6-80), the members the hot path reads or updates. */
typedef struct mrm_run_desc {
  double dt;
  double time;
  double CFL;
  double CFL_tvd;
  double outcad;
  int    autostep;
  int    need_diagnostics;   /* the operations add to the diagnostics arrays */
  int    globiter;
  int    iteration;
} mrm_run_desc;

typedef struct mrm_ctx mrm_ctx;

/* Create a context on the current HIP device (GridData::Init, grid.C:147-348).
 * precision: MRM_PREC_FP64 or MRM_PREC_FP32_MIXED.  An FP32-mixed context
 * stores rho, M, B and p in FP32 and e in FP64; it runs whole steps
 * (mrm_step, with va_adjust, Clean_divB, resistivity and Spitzer conduction
 * as configured), mrm_set_state/mrm_get_state of U (converted on the
 * device), mrm_set_field/mrm_get_field, mrm_exchange_grid and
 * mrm_set_boundary_conditions; the other per-operation calls and
 * diagnostics return MRM_ERR_UNSUPPORTED there. */
int mrm_create(const mrm_grid_desc* grid, const mrm_phys_desc* phys, int precision,
               mrm_ctx** out);
int mrm_destroy(mrm_ctx* ctx);
const char* mrm_last_error(const mrm_ctx* ctx);
int mrm_bufsize(const mrm_ctx* ctx, long* bufsize, int dims[3]);
int mrm_synchronize(mrm_ctx* ctx);

/* eos_init (eos.H:38, eos_init.C:57-212): bytes of a table file in the
 * reference's binary format (14-float header, 6 forward, 2 inverse tables). */
int mrm_eos_load(mrm_ctx* ctx, const void* table_bytes, size_t nbytes);

/* ACCH::UpdateGPU / UpdateCPU counterparts for U, U0, Res (ACCH.h:49-50). */
int mrm_set_state(mrm_ctx* ctx, int which, const double* host, int layout);
int mrm_get_state(mrm_ctx* ctx, int which, double* host, int layout);
/* ... and for pres, temp, ne, rhoi, amb, Qtot, Tau, divB, phi, sflx and the
 * diagnostics arrays (one double per cell; diagnostics never collected read
 * as zero). */
int mrm_set_field(mrm_ctx* ctx, int field, const double* host);
int mrm_get_field(mrm_ctx* ctx, int field, double* host);

/* Globals the reference keeps outside GridData: inv_va2max, v_lim, e_lim,
 * va_max_old (Adjust_Valf_Max.C:15,39) and p_bc, s_bc
 * (boundary_pdmp_1_fftw3.C:32). lim[4] = {inv_va2max, v_lim, e_lim, va_max_old}. */
int mrm_get_limits(const mrm_ctx* ctx, double lim[4]);
int mrm_set_limits(mrm_ctx* ctx, const double lim[4]);
int mrm_get_bc(const mrm_ctx* ctx, double* p_bc, double* s_bc);
int mrm_set_bc(mrm_ctx* ctx, double p_bc, double s_bc);

/* Diagnostics (FP64 contexts): every operation called with
 * run->need_diagnostics adds its terms to tvar1..8, Qres, Qvis
 * (mhdres_SR.C:368-383, 434-435, 688-702; Add_Sources_Integrate_SR.C:122-296,
 * 618-620; tvdlimit_SR.C:577-666; clean_divB.C:264-270).  mrm_diag_zero is the
 * zeroing loop solver.C:271-296 runs where the flag turns on (stage 4 of an
 * output step); the same zeroing happens by itself on the first call that
 * takes a run descriptor with the flag set after one with it clear, so a
 * caller that sets the flag as the solver does needs no explicit call.  The
 * arrays are allocated by their first use. */
int mrm_diag_zero(mrm_ctx* ctx);

/* ConsToPrim (eos.H:39, cons_to_prim.C:39-236): U -> primitive in place over
 * the whole buffer, pres/temp/ne/rhoi/amb filled. */
int mrm_cons_to_prim(mrm_ctx* ctx);

/* Adjust_Valf_Max (limit_va.H:7, Adjust_Valf_Max.C:25-150).  If stats is not
 * NULL it receives {max|v|, max c_s, max eps, n(|v|>=.95 v_lim), n(eps>=.95 e_lim)}. */
int mrm_adjust_valf_max(mrm_ctx* ctx, const mrm_run_desc* run, double stats[5]);

/* MHD_Residual (mhd_tvd.H:5, mhdres_SR.C:27-854): Res += -div F + Lorentz
 * force over the interior; returns dt_cfl = dxmin / cmax (dt_cfl may be
 * NULL: no host round trip then).
 *
 * Deferred stage: SaveCons, ConsToPrim and MHD_Residual are recorded rather
 * than run when the next calls allow it (FP64, no diagnostics, no Spitzer
 * conduction; ConsToPrim stays recorded through Adjust_Valf_Max only with
 * va_adjust = 0 and stats NULL).  MHD_Residual then forms ConsToPrim's EOS
 * fields and cmax (what it returns), and the next Source_Integrate_Tcheck
 * runs the whole stage -- residual, sources, RK update, Tcheck -- in one
 * kernel from the conservative state.  The result is that of the calls run
 * one by one, to rounding; any other call (a state or field read included)
 * first carries the recorded operations out, so every call sees what it
 * would see without the deferral.  MRM_DEFER=0 (read at mrm_create) runs
 * every call as it comes. */
int mrm_mhd_residual(mrm_ctx* ctx, const mrm_run_desc* run, double* dt_cfl);

/* MHD_Residual on caller-owned device arrays, allocating nothing.  It has no
 * counterpart in the reference; it is for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc.  It needs no context: U (8 x bufsize doubles, SoA,
 * primitive as ConsToPrim leaves it, ghosts filled), pres (bufsize) and res
 * (8 x bufsize, added to over the interior as mrm_mhd_residual adds to Res)
 * are device pointers of the grid's buffer, dt_cfl a device pointer that
 * receives dxmin / cmax (limited as mrm_mhd_residual limits it with
 * resistivity; run may be NULL without it).  Every kernel is enqueued on
 * `stream` (a hipStream_t, NULL for the default stream) and nothing is
 * synchronized.  The characteristic-speed reduction lives in a 256-byte
 * device array of the library, one per GPU, held until the process ends and
 * cleared by each call: calls that overlap on one GPU race.  The arithmetic
 * and the kernels are those of mrm_mhd_residual (bit-identical results).
 * Limits: grids of at most 2^29 buffer cells (MRM_ERR_ARG beyond), ghosts 2;
 * Spitzer conduction and ambipolar diffusion, whose residual terms have
 * kernels of their own, return MRM_ERR_UNSUPPORTED. */
int mrm_mhd_residual_noalloc(const mrm_grid_desc* grid, const mrm_phys_desc* phys, const mrm_run_desc* run,
                             double inv_va2max, const double* U, const double* pres, double* res,
                             double* dt_cfl, void* stream);

/* Sync_timestep (solver.C:487-548), single rank: updates run->dt if autostep,
 * writes cfl; MRM_ERR_INVALID_DT / MRM_ERR_CFL on the reference's aborts. */
int mrm_sync_timestep(mrm_ctx* ctx, mrm_run_desc* run, double dt_rad, double dt_mhd,
                      double* cfl);

/* Source_Integrate_Tcheck (src_int_tck.H:8, Add_Sources_Integrate_SR.C:38-472). */
int mrm_source_integrate_tcheck(mrm_ctx* ctx, const mrm_run_desc* run, int stage);
/* SaveCons (src_int_tck.H:10, Add_Sources_Integrate_SR.C:475-531). */
int mrm_save_cons(mrm_ctx* ctx);
/* TCheck (src_int_tck.H:7, Add_Sources_Integrate_SR.C:534-625). */
int mrm_tcheck(mrm_ctx* ctx, const mrm_run_desc* run);
/* TVDlimit (mhd_tvd.H:6, tvdlimit_SR.C:33-695). */
int mrm_tvdlimit(mrm_ctx* ctx, const mrm_run_desc* run, double dt_tvd);
/* Clean_divB (divB.H, clean_divB.C:35-369) on one rank: the damped potential
 * phi (MRM_PHI, zero at mrm_create, kept between calls) corrects B and e over
 * the adaptive number of passes; err (may be NULL) receives max |div B|
 * before and after in the reference's units.  FP64 contexts; mrm_step calls
 * it between TVDlimit and TCheck when phys.divB[0] > 0. */
int mrm_clean_divB(mrm_ctx* ctx, const mrm_run_desc* run, double err[2]);
/* exchange_grid_acc (exchange.H:8, exchange.C:532-782) on one rank:
 * periodic ghost layers of U in dims 1 and 2, filled on the device. */
int mrm_exchange_grid(mrm_ctx* ctx);
/* SetBoundaryConditions (icbc.H:10-12, boundary_pdmp_1_fftw3.C:81-493),
 * plus set_p_s_bc (:496-565) on the first call.  The top boundary holds the
 * vertical field (phys.bnd[1] = bnd_pot = 0) or the potential field
 * (bnd_pot > 0, FP64 contexts, after mrm_potential_kernels): B of the two
 * top ghost layers and B_y, B_z of the top interior layer from b_ext
 * (:402-474).  b_ext is recomputed from B_x of the top interior layer on
 * every call with bnd_pot = 1, on calls with pt_update = 1 with bnd_pot > 1
 * (:335-381), and on the first call after the kernels were loaded.
 * mrm_set_boundary_conditions takes pt_update from stage as solver.C passes
 * it: 1 at stage 0 (:81), 0 at stages 1-4 (:320), at stage 5 (the end of
 * the step) 1 with bnd_pot <= 2 or run->iteration % (bnd_pot - 1) == 0
 * (:355-363); mrm_set_boundary_conditions_pt takes the reference's argument. */
int mrm_set_boundary_conditions(mrm_ctx* ctx, const mrm_run_desc* run, int stage);
int mrm_set_boundary_conditions_pt(mrm_ctx* ctx, const mrm_run_desc* run, int stage, int pt_update);

/* Kernels of the potential-field extrapolation (potential_sd_fftw3.C:83-183,
 * read there from PSF-kernel-<c>.dat): 8 planes of gsize[2] x gsize[1]
 * floats, dim 1 fastest, in the order z-1, z-2, x-0, x-1, x-2, y-0, y-1,
 * y-2.  b_ext[q] is the circular convolution of B_x(lend) over the
 * horizontal plane with kernel q (potential_sd_fftw3.C:178-216), formed
 * through 2D DFTs on the device.  Needs dx[1] = dx[2] (:71-75), gsize[1]
 * and gsize[2] in ghosts..2048 and products of 2, 3, 5 and 7.  FP64
 * contexts.  Replaces the kernel reading and spectra of potential_ext_fftw3;
 * mrm_potential_kernels_load reads the eight files of dir (each a 3-float
 * header gsize[1], gsize[2], dx[0]/dx[1], checked against the grid as at
 * :140-152, then the plane).  mrm_get_potential_bext copies b_ext (8 planes
 * of n1 x n2 doubles with their periodic ghosts, boundary_pdmp_1_fftw3.C:
 * 124-125) to the host. */
int mrm_potential_kernels(mrm_ctx* ctx, const float* k8);
int mrm_potential_kernels_load(mrm_ctx* ctx, const char* dir);
int mrm_get_potential_bext(mrm_ctx* ctx, double* out);

/* nsteps iterations of the ComputeSolution loop body (solver.C:109-365) with
 * RT and IO off: SaveCons, 4 x (ConsToPrim, Adjust_Valf_Max at stage 1,
 * MHD_Residual, Sync_timestep at stage 1, Source_Integrate_Tcheck, exchange,
 * BC), TVDlimit (split in two halves when cfl > CFL_tvd), Clean_divB when
 * phys.divB[0] > 0, TCheck, exchange, BC, Run.Advance.  dt_rad is the
 * radiative time step the caller supplies (RT is outside the path).  With
 * resistivity the stage kernels add the resistive terms; with Spitzer
 * conduction two kernels per stage form the temperature and the conduction
 * terms, which the stage kernel adds.  With
 * run->need_diagnostics the last step collects the diagnostics from its
 * stage 4 on, zeroed first (solver.C:264-296), through the per-operation
 * kernels; run->need_diagnostics is left as it was. */
int mrm_step(mrm_ctx* ctx, mrm_run_desc* run, double dt_rad, int nsteps);

/* Restart files of one rank holding the whole grid, host side (no context):
 * result_prim_<v>.<iter> (OutputSolution / RestoreSolution, io_xysl.C:99-209,
 * 272-412), Header.<iter> (WriteHeaderFile, io_xysl.C:39-48) and backup.dat
 * (Write/ReadBackupFile, boundary_pdmp_1_fftw3.C:44-62).  dir includes the
 * trailing '/', resfile is Run.resfile (e.g. "result").  U is SoA FP64
 * conservative, 8 x bufsize with ghosts; reading zeroes the ghosts. */
int mrm_io_read_prim(const char* dir, const char* resfile, int iter, const mrm_grid_desc* grid, double* U);
int mrm_io_write_prim(const char* dir, const char* resfile, int iter, const mrm_grid_desc* grid, const double* U);
/* One variable file result_prim_<v>.<iter> to or from a field of bufsize
 * doubles (v = 8: Spitzer's heat flux sflx, io_xysl.C:322-336, 389-392). */
int mrm_io_read_var(const char* dir, const char* resfile, int v, int iter, const mrm_grid_desc* grid, double* f);
int mrm_io_write_var(const char* dir, const char* resfile, int v, int iter, const mrm_grid_desc* grid,
                     const double* f);
int mrm_io_write_header(const char* dir, int iter, const mrm_grid_desc* grid, double time, double dt,
                        double va_max);
int mrm_io_read_header(const char* dir, int iter, int gsize[3], double dx[3], double* time, double* dt,
                       double* va_max);
int mrm_io_write_backup(const char* file, int iter, double time, double p_bc, double s_bc);
int mrm_io_read_backup(const char* file, int* iter, double* time, double* p_bc, double* s_bc);
/* The eight PSF-kernel-<c>.dat files of dir (potential_sd_fftw3.C:77-86,
 * 129-163) into k8 in the order of mrm_potential_kernels; MRM_ERR_ARG when
 * a file is missing or short, or a header differs from gx, gy, dx_ratio. */
int mrm_io_read_psf_kernels(const char* dir, int gx, int gy, float dx_ratio, float* k8);

/* Reverse mode of mrm_step.  The reference has no adjoint; these entry points
 * differentiate the step it defines (solver.C:109-365, as mrm_step runs it)
 * for callers that need sensitivities of a scalar objective to the state.
 * mrm_adjoint_forward: nsteps steps as mrm_step, keeping the state every
 * `every` steps and each step's dt and TVD branch (va_adjust = 0, interior
 * extents >= 4).  mrm_adjoint_backward: seed = dJ/dU after the steps
 * (SoA, 8 x bufsize doubles, host); grad receives dJ/dU before them.  Each
 * segment of `every` steps is recomputed from its checkpoint.  In an
 * FP32-mixed context the forward is the mixed step, the checkpoints hold the
 * mixed state (36 bytes per cell), the backward replays each segment with
 * the mixed step and applies the FP64 transposes at the widened mixed states
 * (the FP64 step's derivative along the mixed trajectory), and leaves the
 * context's state as the forward left it.  dt, the TVD
 * branch, p_bc/s_bc, Qtot and Tau are held fixed; min/max, clamps and table
 * cells are differentiated along the forward's branch.  With
 * fold_initial_ghosts != 0 the initial state's ghost cells are taken as filled
 * by mrm_exchange_grid + mrm_set_boundary_conditions, and their cotangents
 * are folded into the interior. */
int mrm_adjoint_forward(mrm_ctx* ctx, mrm_run_desc* run, double dt_rad, int nsteps, int every);
int mrm_adjoint_backward(mrm_ctx* ctx, const double* seed, double* grad, int fold_initial_ghosts);

/* Per-kernel device timers: the counterpart of the reference's per-routine
 * timers printed at verbose > 1 (solver.C:41-43, :428-471).  When enabled,
 * every launch is bracketed by HIP events on the context's stream.
 * mrm_timers(ctx, 1) clears and starts, mrm_timers(ctx, 0) stops;
 * mrm_timers(ctx, 2) times only the volume kernels (not the exchange, the
 * boundary kernels and the cmax faces): each event record idles the GPU for
 * some microseconds, which is noticeable around the short surface kernels;
 * mrm_get_timers synchronizes and returns the accumulated milliseconds and
 * launch counts per timer id (MRM_NTIMER entries; names from mrm_timer_name). */
#define MRM_NTIMER 24
int mrm_timers(mrm_ctx* ctx, int enable);
int mrm_get_timers(mrm_ctx* ctx, double ms[MRM_NTIMER], long launches[MRM_NTIMER]);
const char* mrm_timer_name(int id);

/* Limiter branch accounting (FP64 contexts).  No counterpart in the
 * reference: it counts the branch choices TVDlimit makes (tvdlimit_SR.C)
 * so that a caller can check that two implementations chose alike.
 * Counted are the evaluations on lines whose transverse indices are
 * interior (every point, slope, interface and update of such a line, in
 * each of the three sweeps), summed over every TVDlimit call while the
 * accounting is on.  Slots 0..MRM_BR_NCOUNT-1 are counts; slots
 * MRM_BR_NCOUNT.. are position-weighted checksums, one per branch point:
 * the wrapping 64-bit sum over the evaluations of w(sweep, node, field) x
 * code, w an odd hash of the position and code the choice's number (1 for a
 * switch taken), so that two runs with equal counts but a choice moved from
 * one cell to another still differ.  Categories:
 *   CZ        Tau > 1e-5 at a point                        (:281)
 *   CME       CME_mode = 1 at a point                      (:336)
 *   SLOPE_*   the reconstruction slope per point and field (:392-396):
 *             positive from sl, sr or slm, negative from sl, sr or slm, zero
 *   QRHO      density slope zeroed by the jump detector    (:419-421)
 *   VFAST     velocity slopes zeroed for fast flow         (:427-433)
 *   MM_*      the minmod choice per interface and field    (:449): positive
 *             from sl_lim or sl, negative from sl_lim or sl, zero
 *   CLIP      max(0, 1 + dn (rf - 1)) clipped to 0         (:450)
 *   CMAX_*    min(cmax, max(cm_i, cm_i+1)): cmax, cm_i or cm_i+1 (:457)
 *   RHO_FLOOR max(rho_min, rho) raised rho after the update (:653)
 *   E_FLOOR   the energy floor raised e after the update   (:656)
 * Near ties: every evaluation has a margin (the distance of the deciding
 * quantity from the decision boundary: |lower| and the gap between the two
 * smallest of sl, sr, slm for a positive slope, |x - x0| for a switch x > x0,
 * ...) and a scale (the magnitude of its operands: the variable's values at
 * the points involved, the threshold, ...).  An evaluation whose margin is
 * below tie x scale is a near tie: either choice gives the same value to
 * within tie of the scale, and two implementations whose inputs differ in
 * the last bits may choose differently.  A margin of exactly zero is an
 * exact tie (identical operands: copies, mirrors, a limited slope equal to
 * the difference it came from -- or values that happen to round alike).
 * Near ties and exact ties are counted per branch point in slots
 * MRM_BR_NCOUNT + MRM_BK_N + kind and MRM_BR_NCOUNT + 2 MRM_BK_N + kind and
 * left out of the counts and checksums, which then hold the decisive choices
 * only (tie = 0: no tie accounting, every evaluation is decisive).  The
 * clip of an interface whose difference sl is itself tied with zero is tied
 * with it (its ratio |sl_lim|/|sl| is then rounding noise).
 * mrm_branch_count(ctx, 1, tie) zeroes the slots and turns the accounting on
 * (the TVD sweeps then run counting instantiations of the same kernels);
 * mrm_branch_count(ctx, 0, 0) turns it off; mrm_get_branch_counts
 * synchronizes and copies the MRM_BR_NSLOT slots.  The diagnostics path (an
 * output step's collecting sweeps) is not counted. */
enum {
  MRM_BR_CZ = 0, MRM_BR_CME,
  MRM_BR_SLOPE_POS_L, MRM_BR_SLOPE_POS_R, MRM_BR_SLOPE_POS_M,
  MRM_BR_SLOPE_NEG_L, MRM_BR_SLOPE_NEG_R, MRM_BR_SLOPE_NEG_M, MRM_BR_SLOPE_ZERO,
  MRM_BR_QRHO, MRM_BR_VFAST,
  MRM_BR_MM_POS_LIM, MRM_BR_MM_POS_SL, MRM_BR_MM_NEG_LIM, MRM_BR_MM_NEG_SL, MRM_BR_MM_ZERO,
  MRM_BR_CLIP,
  MRM_BR_CMAX_LIMIT, MRM_BR_CMAX_LEFT, MRM_BR_CMAX_RIGHT,
  MRM_BR_RHO_FLOOR, MRM_BR_E_FLOOR,
  MRM_BR_NCOUNT
};
/* checksum slots: MRM_BR_NCOUNT + kind */
enum {
  MRM_BK_CZ = 0, MRM_BK_CME, MRM_BK_SLOPE, MRM_BK_QRHO, MRM_BK_VFAST, MRM_BK_MM, MRM_BK_CLIP,
  MRM_BK_CMAX, MRM_BK_RHO_FLOOR, MRM_BK_E_FLOOR, MRM_BK_N
};
#define MRM_BR_NSLOT (MRM_BR_NCOUNT + 3 * MRM_BK_N)
int mrm_branch_count(mrm_ctx* ctx, int on, double tie);
int mrm_get_branch_counts(mrm_ctx* ctx, unsigned long long out[MRM_BR_NSLOT]);

#ifdef __cplusplus
}
#endif

#endif /* MURAM_STEP_H */
