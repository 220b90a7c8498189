// mrm_step.hip -- C ABI (include/muram_step.h) of the MI355X-native MURaM
// per-timestep MHD update: context, device memory, launches, and the host-side
// scalar logic of Adjust_Valf_Max, Sync_timestep and set_p_s_bc.
//
// One context per GridData; all device work on the context's stream; the
// state stays in HBM between calls.  Host round trips happen only where the
// reference returns a scalar (dt_cfl, the limiter statistics, the first
// p_bc/s_bc) and in mrm_set_state/mrm_get_state.
#include <hip/hip_runtime.h>
#include <math.h>
#include <algorithm>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <vector>
#include <mutex>

#include "../../include/muram_step.h"
#include "mrm_launch.h"
#include "mrm_divb.h"
#include "mrm_spitzer.h"
#include "mrm_mixed_api.h"
#include "mrm_adjoint_api.h"
#include "mrm_potential_api.h"

using namespace exact;

// the contracted build of the fused-step kernels (mrm_fast.hip)
namespace fast {
extern hipError_t (*p_launch_stage)(hipStream_t, unsigned, int, const StageArgs&, const double*, const double*,
                                    const double*, double*, double*, const double*);
extern hipError_t (*p_launch_tvd_x_lines)(hipStream_t, unsigned, unsigned, size_t, const TvdLineArgs&, double*,
                                          const double*);
extern hipError_t (*p_launch_tvd_sweep)(hipStream_t, int, bool, const TvdEpilogue*, unsigned, const Geom&,
                                        const TvdParams&, double*, const double*);
}  // namespace fast

struct mrm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  mrm_grid_desc grid;
  mrm_phys_desc phys;
  int precision = MRM_PREC_FP64;
  Geom G;
  int lsize[3];
  int vsize = 1;
  double dx[3];
  double dxmin = 0.0;

  double* S[3] = {nullptr, nullptr, nullptr};   // state buffers (fused step rotates them)
  int cur = 0;                                   // S[cur] is the current state U
  double* dU = nullptr;                          // = S[cur]
  double* dU0 = nullptr;                         // = S[c0]
  int c0 = 1;
  // the drop-in path's deferred stage (MRM_DEFER=0 turns it off): a
  // ConsToPrim recorded but not run (U still holds the conservative state),
  // an MHD_Residual recorded after forming only what it returns (the EOS
  // outputs and cmax; the residual is left to the stage kernel that
  // Source_Integrate_Tcheck then runs), a SaveCons recorded (U0 = U, no copy)
  bool defer = true;
  bool c2p_def = false, res_def = false, u0_alias = false;
  // the state buffer (index into S) whose ConsToPrim outputs T, ne, rhoi and
  // amb are still to be formed (a recorded MHD_Residual forms the pressure
  // alone); -1: the fields hold the last ConsToPrim's outputs
  int eosf_pend = -1;
  // a drop-in TVDlimit's dim-2 sweep, recorded: carried out with TCheck as
  // its epilogue when TCheck is the next call, plain before any other
  bool tvd2_def = false;
  double tvd2_dt = 0.0;
  bool tvd2_periodic = false;
  // the ghost cells of a deferred Source_Integrate_Tcheck, recorded: those of
  // gp_src in ConsToPrim's primitive form, into gp_dst.  1: every ghost
  // cell; 2: the dim-0 ghost layers alone (an exchange has since written
  // every ghost cell at interior i).  SetBoundaryConditions after an
  // exchange writes the rest itself
  int gp_pend = 0;
  const double* gp_src = nullptr;
  double* gp_dst = nullptr;
  double def_inv_va2max = 0.0;                   // inv_va2max of the recorded MHD_Residual
  double* dR = nullptr;
  double* dP2 = nullptr;                         // second pressure buffer of the fused step
  bool fast_valid = false;                       // pres and cmax describe S[cur]
  // the interior of dR is zero by Source_Integrate_Tcheck's Res := 0, which
  // was not stored: the next residual starts from zero, every other reader
  // stores the zeros first (res_settle)
  bool res_zero = false;
  bool force_sequence = false;
  bool tvd_x_march = false;                      // dim-0 TVD always by the marching kernel
  size_t lds_max = 64 * 1024;                    // LDS bytes one block may use (device attribute)
  int stage_variant = 4;                         // stage kernel: 4 three wave groups, 3 two, 2 shared
                                                 // fluxes, 1 register window, 0 re-read
  bool fma = false;                              // fused step uses the contracted build (MRM_FMA=1)
  double last_dt = 0.0;                          // dt and TVD split of the last fused step
  bool last_split = false;
  double* dF = nullptr;   // MRM_NFIELD arrays of bs doubles
  // FP32-mixed contexts: three mixed state buffers and two FP32 pressures in
  // place of S[], dP2 and the FP64 pressure; dR is the FP64 staging buffer
  MxState M[3] = {};
  // adjoint: checkpoints of the recorded forward, per-step records, work buffers
  struct AdjRec { double dt; bool split; };
  std::vector<AdjRec> adj_rec;
  std::vector<double*> adj_ckpt;
  int adj_every = 0;
  bool adj_fill0 = false;
  double* adj_mem = nullptr;   // one allocation for the work buffers below
  size_t adj_bytes = 0;
  char* adj_mx_mem = nullptr;  // FP32-mixed backward: a segment's mixed states + the current state's copy
  size_t adj_mx_bytes = 0;
  void* mxmem[3] = {nullptr, nullptr, nullptr};
  float* Pm[2] = {nullptr, nullptr};
  bool qtot_set = false, tau_set = false;

  bool eos_ok = false;
  EosDev E;
  void* d_eos = nullptr;  // one allocation for axes and tables

  double inv_va2max = 0.0, v_lim = 0.0, e_lim = 0.0, va_max_old = 0.0;
  int valf_ini = 1;
  double p_bc = -1.0, s_bc = -1.0;
  int bnd_ini = 1;

  double* d_prof = nullptr;  // hfb, hft, qft (3 * n0)
  // potential-field top boundary (bnd_pot > 0): the FFT plan with the
  // kernels' spectra, and b_ext (8 planes of n1 x n2), valid once updated
  pot::Plan* pot = nullptr;
  double* d_bext = nullptr;
  bool bext_valid = false;
  // damping (dmp_switch): the profile dmp[i] (n0), rho_max's bits (one
  // slot), DMP_SETS x n0 partial layer sums; rho_max is taken at the first call
  double* d_damp = nullptr;
  bool damp_ini = true;
  int tvd_ini = 1;

  // Clean_divB (clean_divB.C:50): passes of the next call; the last call's
  // errors wait in d_red[RED_DIVB0/1] until the next host round trip
  int divb_next_iter = 3;
  bool divb_pending = false;

  // Spitzer conduction (param_spitzer > 0): sflx0, Rflx and, for the fused
  // step, the second sflx buffer and the conductive energy term (4 x bs);
  // sflx is field MRM_SFLX
  double* d_spt = nullptr;

  // diagnostics tvar1..8, Qres, Qvis (DG_N arrays of bs doubles, fields
  // MRM_TVAR1..MRM_QVIS), allocated by their first use
  double* d_diag = nullptr;
  int diag_last = 0;   // need_diagnostics of the last call that took a run descriptor

  // limiter branch accounting (mrm_branch_count): MRM_BR_NSLOT slots
  unsigned long long* d_bcount = nullptr;
  bool bcount_on = false;
  double btie = 0.0;

  unsigned long long* d_red = nullptr;   // RED_NSLOT slots
  double* d_part = nullptr;
  long npart = 0;
  unsigned long long* h_red = nullptr;   // pinned mirror of d_red

  // per-kernel timers (solver.C:428-471 counterpart)
  int timing = 0;                                // 0 off, 1 every timer, 2 volume kernels only
  struct Rec { int id; hipEvent_t a, b; };
  std::vector<hipEvent_t> evpool;
  std::vector<Rec> pending;
  double tms[MRM_NTIMER] = {0};
  long tcnt[MRM_NTIMER] = {0};
  hipEvent_t ev_get() {
    if (!evpool.empty()) { hipEvent_t e = evpool.back(); evpool.pop_back(); return e; }
    hipEvent_t e = nullptr;
    hipEventCreate(&e);
    return e;
  }
  void harvest() {
    hipStreamSynchronize(stream);
    for (auto& r : pending) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) { tms[r.id] += ms; tcnt[r.id]++; }
      evpool.push_back(r.a);
      evpool.push_back(r.b);
    }
    pending.clear();
  }

  std::string err;
};

enum {
  T_C2P = 0, T_VALF, T_RESID, T_CMAXF, T_SRCINT, T_SAVECONS, T_TCHECK, T_EXCH, T_BCSUM, T_BCBOT, T_BCTOP,
  T_TVD0, T_TVD1, T_TVD2, T_NAMED
};
static const char* const timer_names[MRM_NTIMER] = {
  "cons_to_prim", "valf_reduce", "residual", "cmax_faces", "source_integrate", "save_cons", "tcheck",
  "exchange", "bc_sums", "bc_bottom", "bc_top", "tvd_sweep_x", "tvd_sweep_y", "tvd_sweep_z",
  "stage_fused", "eos_p", "clean_divb", "spitzer", "eos_temp", "potential", "damping", "t21", "t22", "t23"};
static const int T_DIVB = T_NAMED + 2;
static const int T_SPT = T_NAMED + 3;
static const int T_EOST = T_NAMED + 4;
static const int T_POT = T_NAMED + 5;
static const int T_DAMP = T_NAMED + 6;

// brackets the launches of one timer id with events on the context's stream
struct TScope {
  mrm_ctx* c;
  int id;
  hipEvent_t a = nullptr;
  TScope(mrm_ctx* c_, int id_) : c(c_), id(id_) {
    const bool surface = id == T_EXCH || id == T_BCSUM || id == T_BCBOT || id == T_BCTOP || id == T_CMAXF;
    if (c->timing == 1 || (c->timing == 2 && !surface)) { a = c->ev_get(); hipEventRecord(a, c->stream); }
  }
  ~TScope() {
    if (a) {
      hipEvent_t b = c->ev_get();
      hipEventRecord(b, c->stream);
      c->pending.push_back({id, a, b});
      if (c->pending.size() > 8192) c->harvest();
    }
  }
};

static int fail(mrm_ctx* c, int code, const char* what, hipError_t e = hipSuccess) {
  if (c) {
    char buf[512];
    if (e != hipSuccess)
      snprintf(buf, sizeof buf, "%s: %s (%s)", what, hipGetErrorString(e), hipGetErrorName(e));
    else
      snprintf(buf, sizeof buf, "%s", what);
    c->err = buf;
  }
  return code;
}

#define HIPCHK(c, x)                                  \
  do {                                                \
    hipError_t e_ = (x);                              \
    if (e_ != hipSuccess) return fail(c, MRM_ERR_HIP, #x, e_); \
  } while (0)

// drop-in per-operation calls work on FP64 state; FP32-mixed contexts run whole steps
#define NEED_FP64(c, fn)                              \
  do {                                                \
    if ((c)->precision != MRM_PREC_FP64)              \
      return fail(c, MRM_ERR_UNSUPPORTED, fn ": FP32-mixed contexts run whole steps (mrm_step) only"); \
  } while (0)

#define LAUNCHCHK(c, name)                            \
  do {                                                \
    hipError_t e_ = hipGetLastError();                \
    if (e_ != hipSuccess) return fail(c, MRM_ERR_HIP, name, e_); \
  } while (0)

static inline unsigned nblk(long n) { return (unsigned)((n + MRM_BLOCK - 1) / MRM_BLOCK); }
static inline long ninterior(const mrm_ctx* c) { return (long)c->lsize[0] * c->lsize[1] * c->lsize[2]; }
static inline double smaxh(double a, double b) { return (a < b) ? b : a; }
static inline double sminh(double a, double b) { return (b < a) ? b : a; }
static inline double bits_to_double(unsigned long long b) {
  double d;
  memcpy(&d, &b, 8);
  return d;
}

// ------------------------------------------------------------------ lifecycle
static int scratch_addr(unsigned long long** out);

// buffer geometry and cell sizes of a grid descriptor (GridData::Init,
// grid.C:147-264, on one rank)
static void geom_of(const mrm_grid_desc& grid, Geom& G, double dx[3]) {
  int n[3];
  for (int d = 0; d < 3; d++) {
    n[d] = grid.gsize[d] + 2 * grid.ghosts[d];
    dx[d] = (grid.gxmax[d] - grid.gxmin[d]) / grid.gsize[d];
  }
  G.n0 = n[0]; G.n1 = n[1]; G.n2 = n[2];
  G.g0 = grid.ghosts[0]; G.g1 = grid.ghosts[1]; G.g2 = grid.ghosts[2];
  G.lb0 = G.g0; G.lb1 = G.g1; G.lb2 = G.g2;
  G.le0 = grid.gsize[0] + G.g0 - 1; G.le1 = grid.gsize[1] + G.g1 - 1; G.le2 = grid.gsize[2] + G.g2 - 1;
  G.s1 = n[0];
  G.s2 = (long)n[0] * n[1];
  G.bs = (long)n[0] * n[1] * n[2];
}

extern "C" int mrm_create(const mrm_grid_desc* grid, const mrm_phys_desc* phys, int precision, mrm_ctx** out) {
  if (!grid || !phys || !out) return MRM_ERR_ARG;
  *out = nullptr;
  if (precision != MRM_PREC_FP64 && precision != MRM_PREC_FP32_MIXED) return MRM_ERR_UNSUPPORTED;
  // physics outside the implemented path
  if (phys->params[6] > 0.0) return MRM_ERR_UNSUPPORTED;
  if (phys->rt_ext_cor != 0) return MRM_ERR_UNSUPPORTED;
  // the damping: FP64 contexts
  if ((int)phys->dmp[0] != 0 && precision != MRM_PREC_FP64) return MRM_ERR_UNSUPPORTED;
  // the potential-field top boundary: FP64 contexts
  if ((int)phys->bnd[1] < 0) return MRM_ERR_ARG;
  if ((int)phys->bnd[1] > 0 && precision != MRM_PREC_FP64) return MRM_ERR_UNSUPPORTED;
  if (phys->divB[0] > 0.0 && phys->divB[1] < 1.0) return MRM_ERR_UNSUPPORTED;
  for (int d = 0; d < 3; d++) {
    if (grid->gsize[d] < 1 || grid->ghosts[d] != 2) return MRM_ERR_ARG;
    if (!(grid->gxmax[d] > grid->gxmin[d])) return MRM_ERR_ARG;
  }
  if (grid->periods[0] || !grid->periods[1] || !grid->periods[2]) return MRM_ERR_UNSUPPORTED;
  if (grid->gsize[0] < 3) return MRM_ERR_ARG;
  mrm_ctx* c = new mrm_ctx();
  c->grid = *grid;
  c->phys = *phys;
  c->precision = precision;
  if (hipGetDevice(&c->device) != hipSuccess) { delete c; return MRM_ERR_HIP; }
  {
    // resolve the no-allocation entry point's scratch on this device now
    unsigned long long* sp = nullptr;
    if (scratch_addr(&sp) != MRM_OK) { delete c; return MRM_ERR_HIP; }
  }
  Geom& G = c->G;
  geom_of(*grid, G, c->dx);
  for (int d = 0; d < 3; d++) {
    c->lsize[d] = grid->gsize[d];
    const int n = grid->gsize[d] + 2 * grid->ghosts[d];
    c->vsize = n > c->vsize ? n : c->vsize;
  }
  c->dxmin = c->dx[0];
  for (int d = 1; d < 3; d++) c->dxmin = sminh(c->dxmin, c->dx[d]);
  const size_t bs = (size_t)G.bs;
  int rc = MRM_OK;
  hipError_t e;
  if ((e = hipStreamCreate(&c->stream)) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "hipStreamCreate", e);
  const bool mixed_prec = precision == MRM_PREC_FP32_MIXED;
  // the step kernels address cells with 32-bit byte offsets of 8-byte values
  if (8 * bs >= (size_t(1) << 32)) {
    fail(c, MRM_ERR_UNSUPPORTED, "at most 2^29 cells per buffer");
    fprintf(stderr, "mrm_create: %s\n", c->err.c_str());
    mrm_destroy(c);
    return MRM_ERR_UNSUPPORTED;
  }
  const size_t fbytes = (7 * bs * sizeof(float) + 255) / 256 * 256;
  for (int q = 0; q < 3 && !rc; q++) {
    if (mixed_prec) {
      if ((e = hipMalloc(&c->mxmem[q], fbytes + bs * sizeof(double))) != hipSuccess)
        rc = fail(c, MRM_ERR_HIP, "hipMalloc mixed state", e);
      else {
        c->M[q].f = (float*)c->mxmem[q];
        c->M[q].e = (double*)((char*)c->mxmem[q] + fbytes);
      }
    } else if ((e = hipMalloc(&c->S[q], 8 * bs * sizeof(double))) != hipSuccess) {
      rc = fail(c, MRM_ERR_HIP, "hipMalloc state", e);
    }
  }
  c->dU = c->S[0];
  c->dU0 = c->S[1];
  for (int q = 0; q < 2 && mixed_prec && !rc; q++)
    if ((e = hipMalloc(&c->Pm[q], bs * sizeof(float))) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "hipMalloc P", e);
  if (!rc && !mixed_prec && (e = hipMalloc(&c->dP2, bs * sizeof(double))) != hipSuccess)
    rc = fail(c, MRM_ERR_HIP, "hipMalloc P2", e);
  if (!rc && (e = hipMalloc(&c->dR, 8 * bs * sizeof(double))) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "hipMalloc Res", e);
  if (!rc && (e = hipMalloc(&c->dF, MRM_NFIELD * bs * sizeof(double))) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "hipMalloc fields", e);
  c->npart = 8 * (long)nblk((long)G.n1 * G.n2) + 64;
  if (!rc && (e = hipMalloc(&c->d_part, c->npart * sizeof(double))) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "hipMalloc part", e);
  if (!rc && (e = hipMalloc(&c->d_red, RED_NSLOT * sizeof(unsigned long long))) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "hipMalloc red", e);
  if (!rc && (e = hipHostMalloc(&c->h_red, RED_NSLOT * sizeof(unsigned long long), hipHostMallocDefault)) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "hipHostMalloc", e);
  if (!rc && (e = hipMalloc(&c->d_prof, 3 * G.n0 * sizeof(double))) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "hipMalloc prof", e);
  // Res starts at zero (the reference relies on it, grid.C:279); all arrays defined
  for (int q = 0; q < 3 && !rc; q++) {
    if (mixed_prec) {
      if ((e = hipMemsetAsync(c->mxmem[q], 0, fbytes + bs * sizeof(double), c->stream)) != hipSuccess)
        rc = fail(c, MRM_ERR_HIP, "memset", e);
    } else if ((e = hipMemsetAsync(c->S[q], 0, 8 * bs * sizeof(double), c->stream)) != hipSuccess) {
      rc = fail(c, MRM_ERR_HIP, "memset", e);
    }
  }
  for (int q = 0; q < 2 && mixed_prec && !rc; q++)
    if ((e = hipMemsetAsync(c->Pm[q], 0, bs * sizeof(float), c->stream)) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "memset", e);
  if (!rc && !mixed_prec && (e = hipMemsetAsync(c->dP2, 0, bs * sizeof(double), c->stream)) != hipSuccess)
    rc = fail(c, MRM_ERR_HIP, "memset", e);
  const char* path = getenv("MRM_STEP_PATH");
  c->force_sequence = path && strcmp(path, "sequence") == 0;
  const char* zw = getenv("MRM_STAGE_ZWIN");
  c->stage_variant = (zw && zw[0] >= '0' && zw[0] <= '4') ? zw[0] - '0' : 4;
  const char* df = getenv("MRM_DEFER");
  c->defer = !(df && df[0] == '0');
  const char* fm = getenv("MRM_FMA");
  c->fma = fm && fm[0] == '1';      // contracted build on request only (the exact one is the parity build)
  const char* tx = getenv("MRM_TVD_X");
  c->tvd_x_march = tx && strcmp(tx, "march") == 0;
  // LDS one block may use: the larger of the default and the opt-in limit
  // (the kernels above 64 KB ask for it with hipFuncSetAttribute)
  for (hipDeviceAttribute_t a : {hipDeviceAttributeMaxSharedMemoryPerBlock, hipDeviceAttributeSharedMemPerBlockOptin}) {
    int v = 0;
    if (!rc && hipDeviceGetAttribute(&v, a, c->device) == hipSuccess && (size_t)v > c->lds_max) c->lds_max = (size_t)v;
  }
  (void)hipGetLastError();
  if (!rc && (e = hipMemsetAsync(c->dR, 0, 8 * bs * sizeof(double), c->stream)) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "memset", e);
  if (!rc && (e = hipMemsetAsync(c->dF, 0, MRM_NFIELD * bs * sizeof(double), c->stream)) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "memset", e);
  if (!rc && phys->params[3] > 0.0) {   // Spitzer: sflx0, Rflx
    if ((e = hipMalloc(&c->d_spt, 4 * bs * sizeof(double))) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "hipMalloc sflx", e);
    else if ((e = hipMemsetAsync(c->d_spt, 0, 4 * bs * sizeof(double), c->stream)) != hipSuccess)
      rc = fail(c, MRM_ERR_HIP, "memset sflx", e);
  }
  if (!rc && (e = hipStreamSynchronize(c->stream)) != hipSuccess) rc = fail(c, MRM_ERR_HIP, "sync", e);
  if (rc) {
    fprintf(stderr, "mrm_create: %s\n", c->err.c_str());
    mrm_destroy(c);
    return rc;
  }
  *out = c;
  return MRM_OK;
}

extern "C" int mrm_destroy(mrm_ctx* c) {
  if (!c) return MRM_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  c->harvest();
  for (hipEvent_t e : c->evpool) hipEventDestroy(e);
  for (int q = 0; q < 3; q++) { hipFree(c->S[q]); hipFree(c->mxmem[q]); }
  for (int q = 0; q < 2; q++) hipFree(c->Pm[q]);
  hipFree(c->dR); hipFree(c->dF); hipFree(c->dP2); hipFree(c->d_spt); hipFree(c->d_diag);
  hipFree(c->d_eos); hipFree(c->d_prof); hipFree(c->d_red); hipFree(c->d_part); hipFree(c->d_bcount);
  pot::plan_destroy(c->pot);
  hipFree(c->d_bext);
  hipFree(c->d_damp);
  for (double* p : c->adj_ckpt) hipFree(p);
  hipFree(c->adj_mem);
  hipFree(c->adj_mx_mem);
  if (c->h_red) hipHostFree(c->h_red);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return MRM_OK;
}

extern "C" const char* mrm_last_error(const mrm_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int mrm_bufsize(const mrm_ctx* c, long* bufsize, int dims[3]) {
  if (!c) return MRM_ERR_ARG;
  if (bufsize) *bufsize = c->G.bs;
  if (dims) { dims[0] = c->G.n0; dims[1] = c->G.n1; dims[2] = c->G.n2; }
  return MRM_OK;
}

extern "C" int mrm_synchronize(mrm_ctx* c) {
  if (!c) return MRM_ERR_ARG;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRM_OK;
}

// ------------------------------------------------------------------ EOS
// eos_init (eos_init.C:57-212): axes in double from the float header; the
// forward tables transposed to [i_eps][i_rho] and divided by del_eps*del_lr in
// double before rounding to float (:170-175); inverse tables likewise (:195-196).
static int drop_settle(mrm_ctx* c);
static int drop_settle_k(mrm_ctx* c, const double* w, bool gp_keep = false);
static int gp_settle(mrm_ctx* c);

extern "C" int mrm_eos_load(mrm_ctx* c, const void* bytes, size_t nbytes) {
  if (!c || !bytes) return MRM_ERR_ARG;
  if (int rc = drop_settle(c)) return rc;
  const float* buf = (const float*)bytes;
  c->fast_valid = false;
  if (nbytes < 14 * sizeof(float)) return fail(c, MRM_ERR_EOS, "EOS table: short header");
  int N_eps = (int)buf[0], N_lr = (int)buf[1], N_lp3 = (int)buf[2], N_s3 = (int)buf[3];
  if (N_eps < 2 || N_lr < 2 || N_lp3 < 2 || N_s3 < 2) return fail(c, MRM_ERR_EOS, "EOS table: bad sizes");
  size_t nt = (size_t)N_eps * N_lr, ni = (size_t)N_lp3 * N_s3;
  if (nbytes < (14 + 6 * nt + 2 * ni) * sizeof(float)) return fail(c, MRM_ERR_EOS, "EOS table: truncated");
  double eps0 = buf[4], eps1 = buf[5], lr0 = buf[6], lr1 = buf[7];
  double lp0 = buf[8], lp1 = buf[9], s0 = buf[10], s1 = buf[11];
  double del_eps = (eps1 - eps0) / ((double)(N_eps - 1));
  double del_lr = (lr1 - lr0) / ((double)(N_lr - 1));
  double del_lp3 = (lp1 - lp0) / ((double)(N_lp3 - 1));
  double del_s3 = (s1 - s0) / ((double)(N_s3 - 1));
  std::vector<double> ax(2 * N_eps + 2 * N_lr + N_lp3 + N_s3);
  double* xeps = ax.data();
  double* xlr = xeps + N_eps;
  double* xlp3 = xlr + N_lr;
  double* xs3 = xlp3 + N_lp3;
  for (int i = 0; i < N_eps; i++) xeps[i] = eps0 + del_eps * ((double)i);
  for (int i = 0; i < N_lr; i++) xlr[i] = lr0 + del_lr * ((double)i);
  for (int i = 0; i < N_lp3; i++) xlp3[i] = lp0 + del_lp3 * ((double)i);
  for (int i = 0; i < N_s3; i++) xs3[i] = s0 + del_s3 * ((double)i);
  double* rxeps = xs3 + N_s3;
  double* rxlr = rxeps + N_eps;
  for (int i = 0; i < N_eps; i++) rxeps[i] = exp(-xeps[i]);
  for (int i = 0; i < N_lr; i++) rxlr[i] = exp(-xlr[i]);
  std::vector<float> tab(6 * nt), itab(2 * ni);
  const float* fw = buf + 14;
  double sc = del_eps * del_lr;
  for (int t = 0; t < 6; t++)
    for (int i = 0; i < N_eps; i++)
      for (int j = 0; j < N_lr; j++)
        tab[t * nt + (size_t)i * N_lr + j] = (float)(((double)fw[t * nt + (size_t)j * N_eps + i]) / sc);
  const float* iv = fw + 6 * nt;
  double sci = del_lp3 * del_s3;
  for (int t = 0; t < 2; t++)
    for (int i = 0; i < N_lp3; i++)
      for (int j = 0; j < N_s3; j++)
        itab[t * ni + (size_t)i * N_s3 + j] = (float)(((double)iv[t * ni + (size_t)j * N_lp3 + i]) / sci);
  size_t axb = ax.size() * sizeof(double);
  size_t total = axb + (tab.size() + itab.size()) * sizeof(float);
  if (c->d_eos) { hipFree(c->d_eos); c->d_eos = nullptr; }
  HIPCHK(c, hipMalloc(&c->d_eos, total));
  char* base = (char*)c->d_eos;
  HIPCHK(c, hipMemcpy(base, ax.data(), axb, hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(base + axb, tab.data(), tab.size() * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(base + axb + tab.size() * sizeof(float), itab.data(), itab.size() * sizeof(float),
                      hipMemcpyHostToDevice));
  EosDev& E = c->E;
  E.N_eps = N_eps; E.N_lr = N_lr; E.N_lp3 = N_lp3; E.N_s3 = N_s3;
  E.inv_del_eps = 1.0 / del_eps; E.inv_del_lr = 1.0 / del_lr;
  E.inv_del_lp3 = 1.0 / del_lp3; E.inv_del_s3 = 1.0 / del_s3;
  E.eps_off = buf[12];
  E.ss_off = buf[13];
  E.eps_max = exp(xeps[N_eps - 1]) - E.eps_off;   // cons_to_prim.C:52, host libm
  E.xeps0 = xeps[0]; E.xlr0 = xlr[0]; E.xlp30 = xlp3[0]; E.xs30 = xs3[0];
  double* dax = (double*)base;
  E.xeps = dax; E.xlr = dax + N_eps; E.xlp3 = dax + N_eps + N_lr; E.xs3 = dax + N_eps + N_lr + N_lp3;
  E.rxeps = E.xs3 + N_s3;
  E.rxlr = E.rxeps + N_eps;
  E.tab = (const float*)(base + axb);
  E.itab = (const float*)(base + axb + tab.size() * sizeof(float));
  c->eos_ok = true;
  return MRM_OK;
}

// ------------------------------------------------------------------ host <-> device
static double* state_ptr(mrm_ctx* c, int which) {
  return which == MRM_U ? c->dU : which == MRM_U0 ? c->dU0 : which == MRM_RES ? c->dR : nullptr;
}

static int res_settle(mrm_ctx* c);
static int drop_settle(mrm_ctx* c);

// In FP32-mixed contexts U is staged through dR (FP64) and converted on the
// device: rho, M, B rounded to FP32, e kept in FP64.
extern "C" int mrm_set_state(mrm_ctx* c, int which, const double* host, int layout) {
  if (!c || !host) return MRM_ERR_ARG;
  if (int rc = drop_settle_k(c, state_ptr(c, which))) return rc;
  const bool mx = c->precision == MRM_PREC_FP32_MIXED;
  if (mx && which != MRM_U) return fail(c, MRM_ERR_UNSUPPORTED, "mrm_set_state: FP32-mixed contexts hold U only");
  double* d = mx ? c->dR : state_ptr(c, which);
  if (!d) return fail(c, MRM_ERR_ARG, "mrm_set_state: bad array id");
  c->fast_valid = false;
  if (which == MRM_RES) c->res_zero = false;
  const size_t bs = c->G.bs;
  if (layout == MRM_LAYOUT_SOA) {
    HIPCHK(c, hipMemcpyAsync(d, host, 8 * bs * sizeof(double), hipMemcpyHostToDevice, c->stream));
  } else if (layout == MRM_LAYOUT_AOS) {
    std::vector<double> soa(8 * bs);
    for (size_t n = 0; n < bs; n++)
      for (int v = 0; v < 8; v++) soa[v * bs + n] = host[8 * n + v];
    HIPCHK(c, hipMemcpyAsync(d, soa.data(), 8 * bs * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else {
    return fail(c, MRM_ERR_ARG, "mrm_set_state: bad layout");
  }
  if (mx) HIPCHK(c, mixed::launch_convert(c->stream, true, c->G.bs, c->dR, c->M[c->cur]));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRM_OK;
}

extern "C" int mrm_get_state(mrm_ctx* c, int which, double* host, int layout) {
  if (!c || !host) return MRM_ERR_ARG;
  if (int rc = drop_settle_k(c, nullptr)) return rc;
  const bool mx = c->precision == MRM_PREC_FP32_MIXED;
  if (mx && which != MRM_U) return fail(c, MRM_ERR_UNSUPPORTED, "mrm_get_state: FP32-mixed contexts hold U only");
  double* d = mx ? c->dR : state_ptr(c, which);
  if (!d) return fail(c, MRM_ERR_ARG, "mrm_get_state: bad array id");
  if (which == MRM_RES) {
    int rc = res_settle(c);
    if (rc) return rc;
  }
  const size_t bs = c->G.bs;
  if (mx) HIPCHK(c, mixed::launch_convert(c->stream, false, c->G.bs, c->dR, c->M[c->cur]));
  if (layout == MRM_LAYOUT_SOA) {
    HIPCHK(c, hipMemcpyAsync(host, d, 8 * bs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  } else if (layout == MRM_LAYOUT_AOS) {
    std::vector<double> soa(8 * bs);
    HIPCHK(c, hipMemcpyAsync(soa.data(), d, 8 * bs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t n = 0; n < bs; n++)
      for (int v = 0; v < 8; v++) host[8 * n + v] = soa[v * bs + n];
  } else {
    return fail(c, MRM_ERR_ARG, "mrm_get_state: bad layout");
  }
  return MRM_OK;
}

static int diag_ensure(mrm_ctx* c);

extern "C" int mrm_set_field(mrm_ctx* c, int f, const double* host) {
  if (!c || !host || f < 0 || f >= MRM_TVAR1 + DG_N) return MRM_ERR_ARG;
  if (int rc = (f >= MRM_TEMP && f <= MRM_AMB) ? drop_settle(c) : drop_settle_k(c, nullptr)) return rc;
  const size_t bs = c->G.bs;
  if (f >= MRM_TVAR1) {
    int rc = diag_ensure(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_diag + (f - MRM_TVAR1) * bs, host, bs * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRM_OK;
  }
  c->fast_valid = false;
  HIPCHK(c, hipMemcpyAsync(c->dF + f * bs, host, bs * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (f == MRM_QTOT) c->qtot_set = true;
  if (f == MRM_TAU) c->tau_set = true;
  return MRM_OK;
}

extern "C" int mrm_get_field(mrm_ctx* c, int f, double* host) {
  if (!c || !host || f < 0 || f >= MRM_TVAR1 + DG_N) return MRM_ERR_ARG;
  if (int rc = (f >= MRM_TEMP && f <= MRM_AMB) ? drop_settle(c) : drop_settle_k(c, nullptr)) return rc;
  const size_t bs = c->G.bs;
  if (f >= MRM_TVAR1) {
    // diagnostics never collected read as zero
    if (!c->d_diag) {
      std::fill(host, host + bs, 0.0);
      return MRM_OK;
    }
    HIPCHK(c, hipMemcpyAsync(host, c->d_diag + (f - MRM_TVAR1) * bs, bs * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MRM_OK;
  }
  if (c->precision == MRM_PREC_FP32_MIXED && f == MRM_PRES) {
    // the FP32 pressure carried by the mixed step, widened
    std::vector<float> pf(bs);
    HIPCHK(c, hipMemcpyAsync(pf.data(), c->Pm[0], bs * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t n = 0; n < bs; n++) host[n] = (double)pf[n];
    return MRM_OK;
  }
  HIPCHK(c, hipMemcpyAsync(host, c->dF + f * bs, bs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRM_OK;
}

extern "C" int mrm_get_limits(const mrm_ctx* c, double lim[4]) {
  if (!c || !lim) return MRM_ERR_ARG;
  lim[0] = c->inv_va2max; lim[1] = c->v_lim; lim[2] = c->e_lim; lim[3] = c->va_max_old;
  return MRM_OK;
}
static int tvd2_settle(mrm_ctx* c);

extern "C" int mrm_set_limits(mrm_ctx* c, const double lim[4]) {
  if (!c || !lim) return MRM_ERR_ARG;
  // a recorded TVD sweep runs with the limits it was recorded under
  if (int rc = tvd2_settle(c)) return rc;
  c->inv_va2max = lim[0]; c->v_lim = lim[1]; c->e_lim = lim[2]; c->va_max_old = lim[3];
  c->fast_valid = false;
  return MRM_OK;
}
extern "C" int mrm_get_bc(const mrm_ctx* c, double* p_bc, double* s_bc) {
  if (!c) return MRM_ERR_ARG;
  if (p_bc) *p_bc = c->p_bc;
  if (s_bc) *s_bc = c->s_bc;
  return MRM_OK;
}
extern "C" int mrm_set_bc(mrm_ctx* c, double p_bc, double s_bc) {
  if (!c) return MRM_ERR_ARG;
  c->p_bc = p_bc; c->s_bc = s_bc;
  return MRM_OK;
}

// ------------------------------------------------------------------ ConsToPrim
static int c2p_launch(mrm_ctx* c);

extern "C" int mrm_cons_to_prim(mrm_ctx* c) {
  if (!c) return MRM_ERR_ARG;
  NEED_FP64(c, "mrm_cons_to_prim");
  if (int rc = gp_settle(c)) return rc;
  c->fast_valid = false;
  if (!c->eos_ok) return fail(c, MRM_ERR_STATE, "mrm_cons_to_prim: no EOS table loaded");
  // recorded: the next MHD_Residual and Source_Integrate_Tcheck read the
  // conservative state themselves, anything else runs it first
  if (c->c2p_def || c->res_def) {
    int rc = drop_settle_k(c, nullptr);
    if (rc) return rc;
  }
  if (c->defer) {
    c->c2p_def = true;
    return MRM_OK;
  }
  return c2p_launch(c);
}

// ------------------------------------------------------------------ Adjust_Valf_Max
static void valf_apply(mrm_ctx* c, double stats[5]);

// device reductions + the host-side update of Adjust_Valf_Max.C:97-144
static int diag_track(mrm_ctx* c, const mrm_run_desc* run);

static void valf_constants(mrm_ctx* c);

extern "C" int mrm_adjust_valf_max(mrm_ctx* c, const mrm_run_desc* run, double stats[5]) {
  if (!c) return MRM_ERR_ARG;
  NEED_FP64(c, "mrm_adjust_valf_max");
  if (int rc = gp_settle(c)) return rc;
  c->fast_valid = false;
  int rc = diag_track(c, run);
  if (rc) return rc;
  // va_adjust = 0 with no statistics asked for: the limits are constants
  // (Adjust_Valf_Max.C:136-144), and a recorded ConsToPrim can stay recorded
  if (c->c2p_def && !c->res_def && !(c->phys.params[2] > 0.0) && !stats) {
    valf_constants(c);
    return MRM_OK;
  }
  if ((rc = drop_settle_k(c, nullptr))) return rc;
  // slots RED_CMAX..RED_CNT_E only: a fused step's Clean_divB maxima may
  // still wait in RED_DIVB0/1 for their host update
  HIPCHK(c, hipMemsetAsync(c->d_red, 0, (RED_CNT_E + 1) * sizeof(unsigned long long), c->stream));
  long n = ninterior(c);
  unsigned nb = nblk(n);
  if (nb > 4096) nb = 4096;
  { TScope ts(c, T_VALF);
  k_valf_reduce<<<nb, MRM_BLOCK, 0, c->stream>>>(c->G, c->dU, c->dF + MRM_PRES * c->G.bs, c->v_lim * 0.95,
                                                 c->e_lim * 0.95, c->d_red);
  }
  LAUNCHCHK(c, "k_valf_reduce");
  HIPCHK(c, hipMemcpyAsync(c->h_red, c->d_red, RED_NSLOT * sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  valf_apply(c, stats);
  return MRM_OK;
}

// the host-side update of Adjust_Valf_Max.C:86-144 from the reduction slots
// in c->h_red (max |v|, max (5/3) p/rho, max e_int/rho, the two counts)
static void valf_apply(mrm_ctx* c, double stats[5]) {
  double g_max[3] = {bits_to_double(c->h_red[RED_VMAX]), sqrt(bits_to_double(c->h_red[RED_CS2MAX])),
                     bits_to_double(c->h_red[RED_EMAX])};
  double g_sum[2] = {(double)c->h_red[RED_CNT_V], (double)c->h_red[RED_CNT_E]};
  const double* prm = c->phys.params;
  const double max_fill = prm[5], cs_fac = 2.0, vv_fac = 3.0, eps_decay = 0.99;
  double va_max;
  if (prm[2] > 0.0) {
    if (c->valf_ini) {
      c->v_lim = fmin(g_max[0], c->phys.tchk[3]);
      c->e_lim = fmin(g_max[2], c->phys.tchk[2]);
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
    c->v_lim = fmin(c->v_lim, c->phys.tchk[3]);
    c->e_lim = fmax(c->e_lim, pow(prm[1] / cs_fac, 2));
    c->e_lim = fmin(c->e_lim, c->phys.tchk[2]);
  } else {
    va_max = prm[1];
    c->v_lim = c->phys.tchk[3];
    c->e_lim = c->phys.tchk[2];
  }
  c->inv_va2max = 1.0 / pow(va_max, 2);
  c->va_max_old = va_max;
  if (stats) {
    stats[0] = g_max[0]; stats[1] = g_max[1]; stats[2] = g_max[2];
    stats[3] = g_sum[0]; stats[4] = g_sum[1];
  }
}

// ------------------------------------------------------------------ MHD_Residual
// the resistive members of ResidParams (mhdres_SR.C:57, 205-207)
static void eta_params(const mrm_ctx* c, ResidParams& RP) {
  RP.eta = c->phys.params[4];
  for (int d = 0; d < 3; d++) {
    const double dx = c->dx[d];
    RP.ww0[d] = -30. / (12. * dx * dx);
    RP.ww1[d] = 16. / (12. * dx * dx);
    RP.ww2[d] = -1. / (12. * dx * dx);
  }
}

// dt_cfl with the resistive limit dt_diff_fac / eta, halved for CFL > 1
// (mhdres_SR.C:147-160, 815-819)
static double eta_dt_limit(const mrm_ctx* c, const mrm_run_desc* run, double dt_cfl) {
  const double eta = c->phys.params[4];
  if (!(eta > 0.0)) return dt_cfl;
  const double dt_diff_fac =
      0.5 / (1. / (c->dx[0] * c->dx[0]) + 1. / (c->dx[1] * c->dx[1]) + 1. / (c->dx[2] * c->dx[2]));
  double dt_res = dt_diff_fac / eta;
  if (run && run->CFL > 1) dt_res *= 0.5;
  return std::min(dt_cfl, dt_res);
}

// ------------------------------------------------------------------ diagnostics
static int diag_ensure(mrm_ctx* c) {
  if (c->d_diag) return MRM_OK;
  const size_t n = (size_t)DG_N * c->G.bs;
  hipError_t e = hipMalloc(&c->d_diag, n * sizeof(double));
  if (e != hipSuccess) {
    c->d_diag = nullptr;
    return fail(c, MRM_ERR_HIP, "hipMalloc diagnostics", e);
  }
  HIPCHK(c, hipMemsetAsync(c->d_diag, 0, n * sizeof(double), c->stream));
  return MRM_OK;
}

// solver.C:264-296 zeroes the arrays where need_diagnostics turns on (stage 4
// of an output step) and turns it off at the next step's stage 1.  A caller
// that sets the flag as the solver does gets the same zeroing: the first
// call that sees it set after calls that did not zeroes the arrays.
static int diag_track(mrm_ctx* c, const mrm_run_desc* run) {
  const int on = run && run->need_diagnostics;
  int rc = MRM_OK;
  if (on && !c->diag_last && c->precision == MRM_PREC_FP64) rc = mrm_diag_zero(c);
  c->diag_last = on;
  return rc;
}

// the arrays a call adds its diagnostics to: none without need_diagnostics;
// idt = 1 / Run.dt, dt_fac = dt_tvd / Run.dt
static int diag_args(mrm_ctx* c, const mrm_run_desc* run, DiagArgs& DA, double dt_tvd = 0.0) {
  DA = DiagArgs{};
  int rc = diag_track(c, run);
  if (rc) return rc;
  if (!run || !run->need_diagnostics) return MRM_OK;
  rc = diag_ensure(c);
  if (rc) return rc;
  DA.dg = c->d_diag;
  DA.idt = 1.0 / run->dt;
  DA.dt_fac = dt_tvd / run->dt;
  return MRM_OK;
}

// the zeroing of solver.C:271-296 at the start of the stage that collects them
extern "C" int mrm_diag_zero(mrm_ctx* c) {
  if (!c) return MRM_ERR_ARG;
  if (int rc = drop_settle(c)) return rc;
  NEED_FP64(c, "mrm_diag_zero");
  c->diag_last = 1;
  if (!c->d_diag) return diag_ensure(c);
  HIPCHK(c, hipMemsetAsync(c->d_diag, 0, (size_t)DG_N * c->G.bs * sizeof(double), c->stream));
  return MRM_OK;
}

static int launch_spitzer_resid(mrm_ctx* c, const mrm_run_desc* run);

// store the Res := 0 that k_source_integrate<..., false> left out
static int res_settle(mrm_ctx* c) {
  if (!c->res_zero) return MRM_OK;
  k_res_zero<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, c->dR);
  LAUNCHCHK(c, "k_res_zero");
  c->res_zero = false;
  return MRM_OK;
}

// va: the inv_va2max to use (the context's when negative)
static int launch_residual(mrm_ctx* c, const mrm_run_desc* run, double va = -1.0) {
  ResidParams RP;
  for (int d = 0; d < 3; d++) {
    RP.w1[d] = 8. / (12. * c->dx[d]);
    RP.w2[d] = -1. / (12. * c->dx[d]);
  }
  RP.inv_va2max = va < 0.0 ? c->inv_va2max : va;
  eta_params(c, RP);
  DiagArgs DA;
  int rc = diag_args(c, run, DA);
  if (rc) return rc;
  HIPCHK(c, hipMemsetAsync(c->d_red + RED_CMAX, 0, sizeof(unsigned long long), c->stream));
  const double* P = c->dF + MRM_PRES * c->G.bs;
  { TScope ts(c, T_RESID);
  if (DA.dg) {
    if ((rc = res_settle(c))) return rc;
    k_residual<true><<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, RP, c->dU, P, c->dR, c->d_red, DA);
  } else {
    // the tiled kernel (plane in LDS, dim-2 register window)
    StageArgs A;
    A.G = c->G;
    A.RP = RP;
    A.ntx = (c->lsize[0] + ST_SX - 1) / ST_SX;
    A.nty = (c->lsize[1] + ST_SY - 1) / ST_SY;
    A.nkc = (c->lsize[2] + ST_KC - 1) / ST_KC;
    const unsigned nb = (unsigned)((long)A.ntx * A.nty * A.nkc);
    const unsigned nt = ST_SX * ST_SY;
    if (c->res_zero) {
      if (RP.eta > 0.0) k_resid_tile<true, true><<<nb, nt, 0, c->stream>>>(A, c->dU, P, c->dR, c->d_red);
      else k_resid_tile<false, true><<<nb, nt, 0, c->stream>>>(A, c->dU, P, c->dR, c->d_red);
    } else {
      if (RP.eta > 0.0) k_resid_tile<true><<<nb, nt, 0, c->stream>>>(A, c->dU, P, c->dR, c->d_red);
      else k_resid_tile<false><<<nb, nt, 0, c->stream>>>(A, c->dU, P, c->dR, c->d_red);
    }
    c->res_zero = false;
  }
  }
  LAUNCHCHK(c, "k_residual");
  long nf = 2L * c->G.g0 * c->lsize[1] * c->lsize[2] + 2L * c->G.g1 * c->lsize[0] * c->lsize[2] +
            2L * c->G.g2 * c->lsize[0] * c->lsize[1];
  unsigned nb = nblk(nf);
  if (nb > 2048) nb = 2048;
  { TScope ts(c, T_CMAXF);
  k_cmax_faces<<<nb, MRM_BLOCK, 0, c->stream>>>(c->G, c->inv_va2max, c->dU, P, c->d_red);
  }
  LAUNCHCHK(c, "k_cmax_faces");
  if (c->d_spt) return launch_spitzer_resid(c, run);
  return MRM_OK;
}

// Spitzer terms of MHD_Residual on the primitive state: constants from the
// dt and CFL of the call (mhdres_SR.C:37-49, 136-145; Run.dt is still the
// previous step's at stage 1)
static SptParams spitzer_params(const mrm_ctx* c, double rdt, double rcfl);

static int launch_spitzer_resid(mrm_ctx* c, const mrm_run_desc* run) {
  const SptParams S = spitzer_params(c, run ? run->dt : 0.0, run ? run->CFL : 0.0);
  const long bs = c->G.bs;
  DiagArgs DA;
  int rc = diag_args(c, run, DA);
  if (rc) return rc;
  k_spitzer_resid<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, S, c->dU, c->dF + MRM_TEMP * bs,
                                                                   c->dF + MRM_SFLX * bs, c->dR, c->d_spt + bs,
                                                                   DA.dg ? DA.dg + DG_T2 * bs : nullptr);
  LAUNCHCHK(c, "k_spitzer_resid");
  return MRM_OK;
}

static SptParams spitzer_params(const mrm_ctx* c, double rdt, double rcfl) {
  SptParams S;
  for (int d = 0; d < 3; d++) {
    S.w1[d] = 8. / (12. * c->dx[d]);
    S.w2[d] = -1. / (12. * c->dx[d]);
  }
  const double CFL_hyp = 1.8, avr_max = 0.25, eps_hyp = 0.02, X_H = 0.7, Rgas = 8.314e7;
  S.k0 = c->phys.params[3];
  S.sat_flx = (1.0 + X_H) * pow(Rgas, 1.5) * sqrt(1836.) / 8.0;
  if (rdt > 0.0) {
    S.cmax_hyp = std::min(rcfl, CFL_hyp) * c->dxmin / rdt;
    S.nu_hyp_max = avr_max / rdt;
    S.hyp_diff = eps_hyp / rdt;
  } else {
    S.cmax_hyp = 0.0;
    S.nu_hyp_max = avr_max * S.cmax_hyp / (std::min(rcfl, CFL_hyp) * c->dxmin);
    S.hyp_diff = eps_hyp * S.cmax_hyp / (std::min(rcfl, CFL_hyp) * c->dxmin);
  }
  return S;
}

static int read_cmax(mrm_ctx* c, double* cmax) {
  HIPCHK(c, hipMemcpyAsync(c->h_red, c->d_red, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *cmax = bits_to_double(c->h_red[RED_CMAX]);
  return MRM_OK;
}

static int c2p_launch(mrm_ctx* c) {
  unsigned nb = nblk(c->G.bs);
  if (nb > 65536u * 4) nb = 65536u * 4;
  TScope ts(c, T_C2P);
  k_cons_to_prim<5><<<nb, MRM_BLOCK, 0, c->stream>>>(c->G, c->E, c->dU, c->dF);
  LAUNCHCHK(c, "k_cons_to_prim");
  // all five outputs of the newest ConsToPrim: nothing stays pending
  c->eosf_pend = -1;
  return MRM_OK;
}

// the pending ConsToPrim outputs T, ne, rhoi, amb formed from their state
static int eosf_settle(mrm_ctx* c) {
  if (c->eosf_pend < 0) return MRM_OK;
  const double* U = c->S[c->eosf_pend];
  c->eosf_pend = -1;
  unsigned nb = nblk(c->G.bs);
  if (nb > 65536u) nb = 65536u;
  TScope ts(c, T_C2P);
  k_eos_fields<false><<<nb, MRM_BLOCK, 0, c->stream>>>(c->G, c->E, U, c->dF, 0.0, c->d_red + RED_SCRATCH);
  LAUNCHCHK(c, "k_eos_fields");
  return MRM_OK;
}

// drop_settle for an operation that reads no pending EOS field and writes
// only the state buffer w (or none): the pending fields stay pending unless
// their state is w.  gp_keep: the caller takes recorded ghost cells over
// (exchange, SetBoundaryConditions), unless another record is pending
static int drop_settle_k(mrm_ctx* c, const double* w, bool gp_keep);

// the recorded ghost cells of a deferred Source_Integrate_Tcheck written
static int gp_settle(mrm_ctx* c) {
  if (!c->gp_pend) return MRM_OK;
  const Geom& G = c->G;
  const long na = 4L * G.n1 * G.n2;
  if (c->gp_pend == 2) {
    k_ghosts_prim<true><<<nblk(na), MRM_BLOCK, 0, c->stream>>>(G, c->gp_src, c->gp_dst);
  } else {
    const long ng = na + 4L * ((long)G.n0 * G.n2 + (long)G.n0 * G.n1);
    k_ghosts_prim<false><<<nblk(ng), MRM_BLOCK, 0, c->stream>>>(G, c->gp_src, c->gp_dst);
  }
  c->gp_pend = 0;
  LAUNCHCHK(c, "k_ghosts_prim");
  return MRM_OK;
}

// every recorded drop-in operation carried out: the pending EOS outputs,
// the stage output's ghost cells, TVDlimit's last sweep, then in the order
// they were called SaveCons's copy, ConsToPrim and MHD_Residual's residual.
// Every entry point that reads or changes the state outside the deferred
// stage calls this, or drop_settle_k for what it touches, first.
static int drop_settle(mrm_ctx* c) {
  if (int rc = eosf_settle(c)) return rc;
  return drop_settle_k(c, nullptr);
}

static int tvd_buf(mrm_ctx* c, double* U, const double* Pr, double dt_tvd, const TvdEpilogue* epi,
                   bool contracted, int only_d, bool y_interior);

// the recorded dim-2 sweep of a drop-in TVDlimit carried out plain
static int tvd2_settle(mrm_ctx* c) {
  if (!c->tvd2_def) return MRM_OK;
  c->tvd2_def = false;
  return tvd_buf(c, c->dU, c->dF + MRM_PRES * c->G.bs, c->tvd2_dt, nullptr, false, 2, c->tvd2_periodic);
}

static int drop_settle_k(mrm_ctx* c, const double* w, bool gp_keep) {
  int rc;
  if (c->gp_pend && !(gp_keep && !c->c2p_def && !c->res_def && !c->u0_alias && !c->tvd2_def) &&
      (rc = gp_settle(c)))
    return rc;
  if ((rc = tvd2_settle(c))) return rc;
  if (c->eosf_pend >= 0) {
    // the pending state is about to change: the operation writes it, or
    // SaveCons's recorded copy goes into it
    const double* P = c->S[c->eosf_pend];
    if (((w && w == P) || (c->u0_alias && c->dU0 == P)) && (rc = eosf_settle(c))) return rc;
  }
  if (c->u0_alias) {
    TScope ts(c, T_SAVECONS);
    k_save_cons<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, c->dU, c->dU0);
    LAUNCHCHK(c, "k_save_cons");
    c->u0_alias = false;
  }
  if (c->c2p_def) {
    c->c2p_def = false;
    if ((rc = c2p_launch(c))) return rc;
  }
  if (c->res_def) {
    c->res_def = false;
    mrm_run_desc r{};
    r.need_diagnostics = 0;
    if ((rc = launch_residual(c, &r, c->def_inv_va2max))) return rc;
  }
  return MRM_OK;
}

// whether MHD_Residual may be recorded: FP64, a recorded ConsToPrim, no
// diagnostics and no Spitzer conduction (whose terms need ConsToPrim's
// temperature and the stored residual)
static bool res_deferrable(const mrm_ctx* c, const mrm_run_desc* run) {
  return c->defer && c->c2p_def && !c->res_def && !(run && run->need_diagnostics) && !c->d_spt &&
         c->precision == MRM_PREC_FP64;
}

extern "C" int mrm_mhd_residual(mrm_ctx* c, const mrm_run_desc* run, double* dt_cfl) {
  if (!c) return MRM_ERR_ARG;
  NEED_FP64(c, "mrm_mhd_residual");
  int rc;
  if ((rc = gp_settle(c))) return rc;
  if (res_deferrable(c, run)) {
    // the EOS outputs of the recorded ConsToPrim and cmax; the residual is
    // formed by the stage kernel of the next Source_Integrate_Tcheck
    if ((rc = diag_track(c, run))) return rc;
    HIPCHK(c, hipMemsetAsync(c->d_red + RED_CMAX, 0, sizeof(unsigned long long), c->stream));
    unsigned nb = nblk(c->G.bs);
    if (nb > 65536u) nb = 65536u;
    {
      TScope ts(c, T_C2P);
      k_eos_fields<true><<<nb, MRM_BLOCK, 0, c->stream>>>(c->G, c->E, c->dU, c->dF, c->inv_va2max,
                                                          c->d_red + RED_CMAX);
    }
    LAUNCHCHK(c, "k_eos_fields");
    // T, ne, rhoi, amb of this state are formed when a caller reads them
    c->eosf_pend = c->cur;
    c->res_def = true;
    c->def_inv_va2max = c->inv_va2max;
    if (!dt_cfl) return MRM_OK;
    double cmax;
    if ((rc = read_cmax(c, &cmax))) return rc;
    *dt_cfl = eta_dt_limit(c, run, c->dxmin / cmax);   // mhdres_SR.C:795-819
    return MRM_OK;
  }
  if ((rc = drop_settle_k(c, nullptr))) return rc;
  rc = launch_residual(c, run);
  if (rc || !dt_cfl) return rc;
  double cmax;
  rc = read_cmax(c, &cmax);
  if (rc) return rc;
  *dt_cfl = eta_dt_limit(c, run, c->dxmin / cmax);   // mhdres_SR.C:795-819
  return MRM_OK;
}

// ------------------------------------------------------------------ MHD_Residual without allocation
// mrm_mhd_residual_noalloc runs MHD_Residual's kernels (k_resid_tile,
// k_cmax_faces) on caller-owned device arrays.  The only device memory it
// needs beyond its inputs and outputs, the characteristic-speed reduction
// slot, lives in `scratch`, a file-scope device array (one copy per GPU that
// loads the library, held until the process ends); its address on the
// current device is resolved once per device with hipGetSymbolAddress, at
// mrm_create or at the entry point's first call on that device, and only
// that address is ever passed to kernels.  Every call on a GPU shares that
// GPU's array: calls that overlap on one GPU race.
#define MRM_NOALLOC_SLOTS 32     // 256 bytes; RED_CMAX is the slot these kernels use
#define MRM_MAX_DEVICES 64
static __device__ __align__(256) unsigned long long scratch[MRM_NOALLOC_SLOTS];
static unsigned long long* scratch_dev[MRM_MAX_DEVICES];
static std::mutex scratch_mu;

// the device address of `scratch` on the current device (resolved once)
static int scratch_addr(unsigned long long** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MRM_MAX_DEVICES) return MRM_ERR_HIP;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (!scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return MRM_ERR_HIP;
    scratch_dev[dev] = static_cast<unsigned long long*>(p);
  }
  *out = scratch_dev[dev];
  return MRM_OK;
}

// zeroes n slots (a kernel, not hipMemsetAsync, so that the call can be
// captured in a HIP graph)
__global__ void k_slots_zero(unsigned long long* p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0ull;
}

// dt_cfl = dxmin / cmax, limited by the resistive dt (mhdres_SR.C:795-819),
// written to device memory
__global__ void k_dt_from_cmax(const unsigned long long* red, double dxmin, double dt_res, double* dt_out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const double dt = dxmin / __longlong_as_double((long long)red[RED_CMAX]);
    *dt_out = (dt_res < dt) ? dt_res : dt;
  }
}

extern "C" int mrm_mhd_residual_noalloc(const mrm_grid_desc* grid, const mrm_phys_desc* phys,
                                        const mrm_run_desc* run, double inv_va2max, const double* U,
                                        const double* pres, double* res, double* dt_cfl, void* stream) {
  if (!grid || !phys || !U || !pres || !res || !dt_cfl) return MRM_ERR_ARG;
  for (int d = 0; d < 3; d++) {
    if (grid->gsize[d] < 1 || grid->ghosts[d] != 2 || !(grid->gxmax[d] > grid->gxmin[d])) return MRM_ERR_ARG;
  }
  // conduction and ambipolar diffusion add residual terms of their own
  if (phys->params[3] > 0.0 || phys->params[6] > 0.0) return MRM_ERR_UNSUPPORTED;
  Geom G;
  double dx[3];
  geom_of(*grid, G, dx);
  // the kernels address cells with 32-bit byte offsets of 8-byte values
  if (8 * (size_t)G.bs >= (size_t(1) << 32)) return MRM_ERR_ARG;
  unsigned long long* red = nullptr;
  int rc = scratch_addr(&red);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  StageArgs A;
  A.G = G;
  for (int d = 0; d < 3; d++) {
    A.RP.w1[d] = 8. / (12. * dx[d]);
    A.RP.w2[d] = -1. / (12. * dx[d]);
  }
  A.RP.inv_va2max = inv_va2max;
  A.RP.eta = phys->params[4];
  for (int d = 0; d < 3; d++) {
    A.RP.ww0[d] = -30. / (12. * dx[d] * dx[d]);
    A.RP.ww1[d] = 16. / (12. * dx[d] * dx[d]);
    A.RP.ww2[d] = -1. / (12. * dx[d] * dx[d]);
  }
  const int n[3] = {grid->gsize[0], grid->gsize[1], grid->gsize[2]};
  A.ntx = (n[0] + ST_SX - 1) / ST_SX;
  A.nty = (n[1] + ST_SY - 1) / ST_SY;
  A.nkc = (n[2] + ST_KC - 1) / ST_KC;
  k_slots_zero<<<1, 64, 0, st>>>(red, MRM_NOALLOC_SLOTS);
  const unsigned nb = (unsigned)((long)A.ntx * A.nty * A.nkc);
  if (A.RP.eta > 0.0) k_resid_tile<true><<<nb, ST_SX * ST_SY, 0, st>>>(A, U, pres, res, red);
  else k_resid_tile<false><<<nb, ST_SX * ST_SY, 0, st>>>(A, U, pres, res, red);
  long nf = 2L * G.g0 * n[1] * n[2] + 2L * G.g1 * n[0] * n[2] + 2L * G.g2 * n[0] * n[1];
  unsigned nbf = nblk(nf);
  if (nbf > 2048) nbf = 2048;
  k_cmax_faces<<<nbf, MRM_BLOCK, 0, st>>>(G, inv_va2max, U, pres, red);
  double dxmin = dx[0];
  for (int d = 1; d < 3; d++) dxmin = sminh(dxmin, dx[d]);
  double dt_res = INFINITY;
  if (A.RP.eta > 0.0) {
    const double dt_diff_fac = 0.5 / (1. / (dx[0] * dx[0]) + 1. / (dx[1] * dx[1]) + 1. / (dx[2] * dx[2]));
    dt_res = dt_diff_fac / A.RP.eta;
    if (run && run->CFL > 1) dt_res *= 0.5;
  }
  k_dt_from_cmax<<<1, 64, 0, st>>>(red, dxmin, dt_res, dt_cfl);
  return hipGetLastError() == hipSuccess ? MRM_OK : MRM_ERR_HIP;
}

// ------------------------------------------------------------------ Sync_timestep
static int invalid_dt(double a) { return (a != a) || (a <= 0.) || (a == INFINITY); }

extern "C" int mrm_sync_timestep(mrm_ctx* c, mrm_run_desc* run, double dt_rad, double dt_mhd, double* cfl) {
  if (!c || !run) return MRM_ERR_ARG;
  int rc = diag_track(c, run);
  if (rc) return rc;
  double r0 = dt_rad, r1 = dt_mhd;
  if (invalid_dt(r0)) return fail(c, MRM_ERR_INVALID_DT, "Invalid time step: dt_rad");
  if (invalid_dt(r1)) return fail(c, MRM_ERR_INVALID_DT, "Invalid time step: dt_mhd");
  double dt_min_rt = c->phys.rt[3] * run->CFL * r1;
  if (r0 < dt_min_rt) r0 = smaxh(r0, dt_min_rt);
  double dt_min = sminh(r0, r1);
  double dt = sminh(r0, run->CFL * r1);
  if (run->outcad > 0) {
    double num = run->time / run->outcad;
    int ceil_ = (int)num;
    double rem = (num == (double)ceil_) ? 0.0 : (ceil_ + 1) * run->outcad - run->time;
    if (rem > 0.0) dt = sminh(dt, rem);
  }
  if (run->autostep) run->dt = dt;
  double cf = run->dt / dt_min;
  if (cfl) *cfl = cf;
  if (cf >= 2.) return fail(c, MRM_ERR_CFL, "Invalid time step: CFL >= 2");
  return MRM_OK;
}

// ------------------------------------------------------------------ sources / integrate
static SrcParams src_params(const mrm_ctx* c, const mrm_run_desc* run, int stage) {
  SrcParams SP;
  SP.wdt = run->dt / double(4 + 1 - stage);
  SP.invRdt = c->phys.rt[4] / run->dt;
  SP.g = -c->phys.params[0];
  SP.inv_va2max = c->inv_va2max;
  SP.rho_min = c->phys.tchk[1];
  SP.eps_min = c->phys.tchk[0];
  SP.vmax = (c->v_lim > 0.0) ? c->v_lim : c->phys.tchk[3];
  SP.eps_max = (c->e_lim > 0.0) ? c->e_lim : c->phys.tchk[2];
  SP.dmp = (int)c->phys.dmp[0] ? c->d_damp : nullptr;
  return SP;
}

// get_damping at stage 1 (Add_Sources_Integrate_SR.C:105-108): the profile
// from U (velocities with prim, else momenta); the first call also takes
// rho_max.  A no-op with the damping off.
static int damping_update(mrm_ctx* c, const double* U, bool prim) {
  if (!(int)c->phys.dmp[0]) return MRM_OK;
  const Geom& G = c->G;
  if (!c->d_damp) {
    const size_t bytes = ((size_t)G.n0 * (1 + DMP_SETS) + 1) * sizeof(double);
    HIPCHK(c, hipMalloc(&c->d_damp, bytes));
    HIPCHK(c, hipMemsetAsync(c->d_damp, 0, bytes, c->stream));
  }
  double* prof = c->d_damp;
  unsigned long long* rmax = (unsigned long long*)(c->d_damp + G.n0);
  double* part = c->d_damp + G.n0 + 1;
  const int nx = G.le0 - G.lb0 + 1;
  const dim3 grid((nx + 63) / 64, DMP_SETS / 4);
  TScope ts(c, T_DAMP);
  if (c->damp_ini) {
    if (prim) k_damp_partials<true, true><<<grid, 256, 0, c->stream>>>(G, U, part, rmax);
    else k_damp_partials<false, true><<<grid, 256, 0, c->stream>>>(G, U, part, rmax);
    c->damp_ini = false;
  } else {
    if (prim) k_damp_partials<true, false><<<grid, 256, 0, c->stream>>>(G, U, part, rmax);
    else k_damp_partials<false, false><<<grid, 256, 0, c->stream>>>(G, U, part, rmax);
  }
  LAUNCHCHK(c, "k_damp_partials");
  k_damp_finish<<<nblk(nx), MRM_BLOCK, 0, c->stream>>>(G, part, rmax, 1.0 / c->phys.dmp[1], c->phys.dmp[2],
                                                       1.0 / c->phys.dmp[3], prof);
  LAUNCHCHK(c, "k_damp_finish");
  return MRM_OK;
}

static int stage_launch(mrm_ctx* c, const mrm_run_desc* run, int stage, const double* Uin, const double* Pin,
                        const double* U0, double* Uout, double* Pout, const double* eres = nullptr,
                        double rp_va = -1.0);

// Source_Integrate_Tcheck after a recorded ConsToPrim and MHD_Residual: the
// stage kernel forms the residual from the conservative state and the
// recorded EOS pressure and applies sources, Boris correction, RK update and
// Tcheck, into the spare state buffer, whose ghost cells are to take the
// input's in ConsToPrim's primitive form (what ConsToPrim and the in-place
// update leave there: recorded, since the exchange and SetBoundaryConditions
// that follow in the stage loop overwrite every one of them); the buffers
// change roles, and Res := 0 is recorded
static int srcint_deferred(mrm_ctx* c, const mrm_run_desc* run, int stage) {
  int rc;
  if (stage == 1 && (rc = damping_update(c, c->dU, false))) return rc;
  const int spare = 3 - c->cur - c->c0;
  // the stage kernel writes the spare buffer (the pending EOS fields' state
  // is the input, or an older state superseded by it)
  if (c->eosf_pend == spare && (rc = eosf_settle(c))) return rc;
  double* Uin = c->dU;
  double* Uout = c->S[spare];
  const double* U0 = c->u0_alias ? c->dU : c->dU0;
  if ((rc = stage_launch(c, run, stage, Uin, c->dF + MRM_PRES * c->G.bs, U0, Uout, nullptr, nullptr,
                         c->def_inv_va2max)))
    return rc;
  c->gp_pend = 1;
  c->gp_src = Uin;
  c->gp_dst = Uout;
  if (c->u0_alias) {
    c->c0 = c->cur;     // the input state is U0
    c->u0_alias = false;
  }
  c->cur = spare;
  c->dU = c->S[c->cur];
  c->dU0 = c->S[c->c0];
  c->c2p_def = false;
  c->res_def = false;
  c->res_zero = true;
  return MRM_OK;
}

extern "C" int mrm_source_integrate_tcheck(mrm_ctx* c, const mrm_run_desc* run, int stage) {
  if (!c || !run || stage < 1 || stage > 4) return MRM_ERR_ARG;
  NEED_FP64(c, "mrm_source_integrate_tcheck");
  if (int rc = gp_settle(c)) return rc;
  c->fast_valid = false;
  if (c->res_def && !run->need_diagnostics) return srcint_deferred(c, run, stage);
  if (int rc = drop_settle_k(c, c->dU)) return rc;
  // the damping profile of the step, from the primitive state of stage 1
  if (stage == 1) {
    int rc = damping_update(c, c->dU, true);
    if (rc) return rc;
  }
  SrcParams SP = src_params(c, run, stage);
  const double* Q = c->qtot_set ? c->dF + MRM_QTOT * c->G.bs : nullptr;
  DiagArgs DA;
  int rc = diag_args(c, run, DA);
  if (rc) return rc;
  // the kernel reads Res: zeros left unstored by a previous call go in first
  if ((rc = res_settle(c))) return rc;
  TScope ts(c, T_SRCINT);
  if (DA.dg) {
    k_source_integrate<true><<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, SP, c->dU, c->dU0, c->dR, Q, DA);
  } else {
    k_source_integrate<false, false><<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, SP, c->dU, c->dU0, c->dR,
                                                                                     Q, DA);
    c->res_zero = true;
  }
  LAUNCHCHK(c, "k_source_integrate");
  if (c->d_spt) {
    const long bs = c->G.bs;
    k_spitzer_integrate<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, SP.wdt, c->dF + MRM_SFLX * bs,
                                                                         c->d_spt, c->d_spt + bs);
    LAUNCHCHK(c, "k_spitzer_integrate");
  }
  return MRM_OK;
}

extern "C" int mrm_save_cons(mrm_ctx* c) {
  if (!c) return MRM_ERR_ARG;
  NEED_FP64(c, "mrm_save_cons");
  c->fast_valid = false;
  if (int rc = drop_settle_k(c, c->defer ? nullptr : c->dU0)) return rc;
  if (c->defer) {
    c->u0_alias = true;   // copied only if U changes in place before a deferred stage
  } else {
    TScope ts(c, T_SAVECONS);
    k_save_cons<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, c->dU, c->dU0);
    LAUNCHCHK(c, "k_save_cons");
  }
  if (c->d_spt) {
    k_spitzer_save<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, c->dF + MRM_SFLX * c->G.bs, c->d_spt);
    LAUNCHCHK(c, "k_spitzer_save");
  }
  return MRM_OK;
}

extern "C" int mrm_tcheck(mrm_ctx* c, const mrm_run_desc* run) {
  if (!c) return MRM_ERR_ARG;
  NEED_FP64(c, "mrm_tcheck");
  if (int rc = gp_settle(c)) return rc;
  double vmax = (c->v_lim > 0.0) ? c->v_lim : c->phys.tchk[3];
  double eps_max = (c->e_lim > 0.0) ? c->e_lim : c->phys.tchk[2];
  if (c->tvd2_def && !(run && run->need_diagnostics)) {
    // the recorded dim-2 sweep with TCheck as its epilogue: the clamp of
    // Add_Sources_Integrate_SR.C:447-465 applied to each cell right after
    // its last TVD update (every interior cell is updated by that sweep)
    c->tvd2_def = false;
    if (int rc = drop_settle_k(c, c->dU)) return rc;
    c->fast_valid = false;
    TvdEpilogue X{};
    X.E = c->E;
    X.rho_min = c->phys.tchk[1];
    X.eps_min = c->phys.tchk[0];
    X.eps_max = eps_max;
    X.vmax = vmax;
    X.Pnew = nullptr;
    X.cmax = nullptr;
    X.valf = nullptr;
    return tvd_buf(c, c->dU, c->dF + MRM_PRES * c->G.bs, c->tvd2_dt, &X, false, 2, c->tvd2_periodic);
  }
  if (int rc = drop_settle_k(c, c->dU)) return rc;
  c->fast_valid = false;
  DiagArgs DA;
  int rc = diag_args(c, run, DA);
  if (rc) return rc;
  TScope ts(c, T_TCHECK);
  k_tcheck<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(c->G, c->phys.tchk[1], c->phys.tchk[0], eps_max, vmax,
                                                          c->dU, DA);
  LAUNCHCHK(c, "k_tcheck");
  return MRM_OK;
}

// ------------------------------------------------------------------ exchange
// periodic ghosts of U (and of the pressure P when given) in dims 1 and 2
static int exchange_buf(mrm_ctx* c, double* U, double* P) {
  const Geom& G = c->G;
  TScope ts(c, T_EXCH);
  long n1 = (long)c->lsize[0] * c->lsize[2] * G.g1;
  k_exchange_p<1><<<nblk(n1), MRM_BLOCK, 0, c->stream>>>(G, U, P);
  LAUNCHCHK(c, "k_exchange_p<1>");
  long n2 = (long)c->lsize[0] * G.n1 * G.g2;
  k_exchange_p<2><<<nblk(n2), MRM_BLOCK, 0, c->stream>>>(G, U, P);
  LAUNCHCHK(c, "k_exchange_p<2>");
  return MRM_OK;
}

static int exchange_mx(mrm_ctx* c, MxState M, float* P) {
  TScope ts(c, T_EXCH);
  HIPCHK(c, mixed::launch_exchange(c->stream, c->G, M, P));
  return MRM_OK;
}

static int exchange_arrays(mrm_ctx* c, double* base, int na);

extern "C" int mrm_exchange_grid(mrm_ctx* c) {
  if (!c) return MRM_ERR_ARG;
  if (int rc = drop_settle_k(c, c->dU, true)) return rc;
  c->fast_valid = false;
  if (c->precision == MRM_PREC_FP32_MIXED) return exchange_mx(c, c->M[c->cur], nullptr);
  int rc = exchange_buf(c, c->dU, nullptr);
  if (rc) return rc;
  // every ghost cell at interior i written: the dim-0 ghost layers remain
  if (c->gp_pend == 1) c->gp_pend = 2;
  if (!c->d_spt) return rc;
  return exchange_arrays(c, c->dF + MRM_SFLX * c->G.bs, 1);
}

// ------------------------------------------------------------------ Clean_divB
static DivbArgs divb_args(const mrm_ctx* c) {
  DivbArgs A;
  A.G = c->G;
  for (int d = 0; d < 3; d++) {
    A.w1[d] = 8. / (12. * c->dx[d]);
    A.w2[d] = -1. / (12. * c->dx[d]);
  }
  // clean_divB.C:92-110 (the 3-D branch compares dx[2] with the updated dxmin)
  double dxmin = std::min(c->dx[0], c->dx[1]);
  double dxmax = std::max(c->dx[0], c->dx[1]);
  dxmin = std::min(dxmin, c->dx[2]);
  dxmax = std::max(dxmin, c->dx[2]);
  A.fac = 1.0 * std::min(dxmin, 0.4 * dxmax) * dxmin;
  A.damp = 0.65;
  return A;
}

// the pass-count update of clean_divB.C:330-345 from the maxima in h_red
static void divb_settle_host(mrm_ctx* c, double err_out[2]) {
  double err_unit = 0.0;
  for (int v = 0; v < 3; v++) err_unit += 1.0 / c->dx[v];
  err_unit = 0.5 / err_unit * 3.5449;
  const double e0 = bits_to_double(c->h_red[RED_DIVB0]) * err_unit;
  const double e1 = bits_to_double(c->h_red[RED_DIVB1]) * err_unit;
  const double err = e1 / e0;
  if (err > c->phys.divB[2]) c->divb_next_iter += 1;
  else c->divb_next_iter -= 1;
  c->divb_next_iter = std::min(c->divb_next_iter, (int)c->phys.divB[1]);
  c->divb_next_iter = std::max(1, c->divb_next_iter);
  c->divb_pending = false;
  if (err_out) { err_out[0] = e0; err_out[1] = e1; }
}

static int divb_settle(mrm_ctx* c, double err_out[2]) {
  if (!c->divb_pending) return MRM_OK;
  HIPCHK(c, hipMemcpyAsync(c->h_red, c->d_red, (RED_DIVB1 + 1) * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  divb_settle_host(c, err_out);
  return MRM_OK;
}

static int exchange_arrays(mrm_ctx* c, double* base, int na) {
  const Geom& G = c->G;
  long n1 = (long)c->lsize[0] * c->lsize[2] * G.g1;
  k_exchange_arr<1><<<nblk(n1), MRM_BLOCK, 0, c->stream>>>(G, base, na, G.bs);
  LAUNCHCHK(c, "k_exchange_arr<1>");
  long n2 = (long)c->lsize[0] * G.n1 * G.g2;
  k_exchange_arr<2><<<nblk(n2), MRM_BLOCK, 0, c->stream>>>(G, base, na, G.bs);
  LAUNCHCHK(c, "k_exchange_arr<2>");
  return MRM_OK;
}

// Clean_divB's passes on U (clean_divB.C:118-328).  The first divergence
// kernel also makes the first potential update, each later one the next
// pass's; with epi the last correction pass finishes every cell.  The two
// maxima land in d_red[RED_DIVB0/1] and c->divb_pending is set.
static int clean_divb_dev(mrm_ctx* c, double* U, const TvdEpilogue* epi, const DiagArgs* DA = nullptr) {
  int rc;
  if ((rc = divb_settle(c, nullptr))) return rc;
  const Geom& G = c->G;
  DivbArgs A = divb_args(c);
  if (DA && DA->dg) {
    A.qres = DA->dg + DG_QRES * G.bs;
    A.invdt = DA->idt;
  }
  double* divB = c->dF + MRM_DIVB * G.bs;
  double* phi = c->dF + MRM_PHI * G.bs;
  unsigned long long* red = c->d_red;
  const unsigned nb = divb_blocks(G);
  const int npass = c->divb_next_iter;
  TScope ts(c, T_DIVB);
  HIPCHK(c, hipMemsetAsync(red + RED_DIVB0, 0, 2 * sizeof(unsigned long long), c->stream));
  if ((rc = exchange_arrays(c, U + 5 * G.bs, 3))) return rc;
  k_divb<true><<<nb, DIVB_TX * DIVB_TY, 0, c->stream>>>(A, U, divB, phi, red + RED_DIVB0);
  LAUNCHCHK(c, "k_divb");
  for (int it = 0; it < npass; it++) {
    const bool last = it == npass - 1;
    if ((rc = exchange_arrays(c, phi, 1))) return rc;
    const long ncol = (long)c->lsize[1] * c->lsize[2];
    k_phi_vertical<<<nblk(ncol), MRM_BLOCK, 0, c->stream>>>(G, phi);
    LAUNCHCHK(c, "k_phi_vertical");
    if (last && epi) k_divb_update<true><<<nb, DIVB_TX * DIVB_TY, 0, c->stream>>>(A, U, phi, divB, *epi);
    else k_divb_update<false><<<nb, DIVB_TX * DIVB_TY, 0, c->stream>>>(A, U, phi, divB, TvdEpilogue{});
    LAUNCHCHK(c, "k_divb_update");
    if ((rc = exchange_arrays(c, U + 5 * G.bs, 3))) return rc;
    if (last) k_divb<false><<<nb, DIVB_TX * DIVB_TY, 0, c->stream>>>(A, U, divB, phi, red + RED_DIVB1);
    else k_divb<true><<<nb, DIVB_TX * DIVB_TY, 0, c->stream>>>(A, U, divB, phi, red + RED_SCRATCH);
    LAUNCHCHK(c, "k_divb");
  }
  c->divb_pending = true;
  return MRM_OK;
}

// Clean_divB on a mixed state (FP32 B): the passes of clean_divb_dev with
// the mixed kernels; phi and divB stay FP64 fields
static int clean_divb_mx(mrm_ctx* c, MxState M, const MxEpilogue* epi) {
  int rc;
  if ((rc = divb_settle(c, nullptr))) return rc;
  const Geom& G = c->G;
  const DivbArgs A = divb_args(c);
  double* divB = c->dF + MRM_DIVB * G.bs;
  double* phi = c->dF + MRM_PHI * G.bs;
  unsigned long long* red = c->d_red;
  const int npass = c->divb_next_iter;
  TScope ts(c, T_DIVB);
  HIPCHK(c, hipMemsetAsync(red + RED_DIVB0, 0, 2 * sizeof(unsigned long long), c->stream));
  HIPCHK(c, mixed::launch_exchange_b(c->stream, G, M));
  HIPCHK(c, mixed::launch_divb(c->stream, A, M, divB, phi, red + RED_DIVB0, true));
  for (int it = 0; it < npass; it++) {
    const bool last = it == npass - 1;
    if ((rc = exchange_arrays(c, phi, 1))) return rc;
    const long ncol = (long)c->lsize[1] * c->lsize[2];
    k_phi_vertical<<<nblk(ncol), MRM_BLOCK, 0, c->stream>>>(G, phi);
    LAUNCHCHK(c, "k_phi_vertical");
    HIPCHK(c, mixed::launch_divb_update(c->stream, A, M, phi, divB, last ? epi : nullptr));
    HIPCHK(c, mixed::launch_exchange_b(c->stream, G, M));
    HIPCHK(c, mixed::launch_divb(c->stream, A, M, divB, phi, last ? red + RED_DIVB1 : red + RED_SCRATCH, !last));
  }
  c->divb_pending = true;
  return MRM_OK;
}

extern "C" int mrm_clean_divB(mrm_ctx* c, const mrm_run_desc* run, double err[2]) {
  if (!c) return MRM_ERR_ARG;
  if (int rc = drop_settle_k(c, c->dU)) return rc;
  NEED_FP64(c, "mrm_clean_divB");
  if (c->phys.divB[1] < 1.0) return fail(c, MRM_ERR_ARG, "mrm_clean_divB: divB_itmax < 1");
  c->fast_valid = false;
  DiagArgs DA;
  int rc = diag_args(c, run, DA);
  if (rc) return rc;
  rc = clean_divb_dev(c, c->dU, nullptr, &DA);
  if (rc) return rc;
  return divb_settle(c, err);
}

// ------------------------------------------------------------------ boundary conditions
// partial sums over the two lowest layers of U (FP64) or of the mixed state M
static int bottom_sums(mrm_ctx* c, int mode, const double* U, double* dst_dev, const MxState* M = nullptr) {
  const long nyz = (long)c->lsize[1] * c->lsize[2];
  unsigned nb = nblk(nyz);
  TScope ts(c, T_BCSUM);
  if (M) {
    HIPCHK(c, mixed::launch_bc_sums(c->stream, nb, c->G, c->E, *M, mode, c->d_part));
  } else if (mode == 0) {
    k_bc_sums<0><<<nb, MRM_BLOCK, 0, c->stream>>>(c->G, c->E, U, c->d_part);
    LAUNCHCHK(c, "k_bc_sums<0>");
  } else {
    k_bc_sums<1><<<nb, MRM_BLOCK, 0, c->stream>>>(c->G, c->E, U, c->d_part);
    LAUNCHCHK(c, "k_bc_sums<1>");
  }
  k_sum_partials<<<1, MRM_BLOCK, 0, c->stream>>>(c->d_part, (int)nb, mode == 0 ? 2 : 5, dst_dev);
  LAUNCHCHK(c, "k_sum_partials");
  return MRM_OK;
}

// set_p_s_bc (boundary_pdmp_1_fftw3.C:496-565), one rank
static int set_p_s_bc(mrm_ctx* c, const double* U, const MxState* M = nullptr) {
  double* dsum = (double*)(c->d_red + RED_SUM0);
  int rc = bottom_sums(c, 1, U, dsum, M);
  if (rc) return rc;
  double h[5];
  HIPCHK(c, hipMemcpyAsync(h, dsum, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double nh = (double)(c->grid.gsize[1] * c->grid.gsize[2]);
  h[0] /= nh;
  h[1] /= nh;
  if (c->p_bc < 0.0) c->p_bc = h[0] * sqrt(h[0] / h[1]);
  if (c->s_bc < 0.0) {
    if (h[2] > 0.0) c->s_bc = h[3] / h[2];
    else c->s_bc = h[4] / nh;
  }
  return MRM_OK;
}

static BcParams bc_params(const mrm_ctx* c) {
  BcParams B;
  B.p_bc = c->p_bc;
  B.s_bc = c->s_bc;
  B.B_crit = c->phys.bnd[2] / sqrt(8.0 * asin(1.0));
  B.eps_top = c->phys.bnd[3];
  B.cc = 1.0 / ((double)(c->grid.gsize[1] * c->grid.gsize[2]));
  B.open_top = (int)c->phys.bnd[0];
  if ((int)c->phys.bnd[1] > 0) {
    B.bext = c->d_bext;
    B.bplane = (long)c->G.n1 * c->G.n2;
  }
  return B;
}

// pt_update as solver.C passes it: 1 on the first call (stage 0), 0 inside
// the step (stages 1-4), at the end of a step 1 with bnd_pot <= 2 and on
// every (bnd_pot - 1)-th iteration otherwise (solver.C:81, 320, 355-363)
static int pt_update_rule(const mrm_ctx* c, const mrm_run_desc* run, int stage) {
  const int pot = (int)c->phys.bnd[1];
  if (stage == 0) return 1;
  if (stage <= 4) return 0;
  if (pot <= 2 || !run) return 1;
  return !(run->iteration % (pot - 1));
}

// the potential-field extrapolation of the top boundary when it is due
// (boundary_pdmp_1_fftw3.C:335-381): on every call with bnd_pot = 1, on
// calls with pt_update with bnd_pot > 1, and whenever b_ext has not been
// formed yet
static int potential_due(mrm_ctx* c, const double* U, int pt_update) {
  const int pot = (int)c->phys.bnd[1];
  if (pot <= 0) return MRM_OK;
  if (!c->pot || !c->pot->kernels)
    return fail(c, MRM_ERR_STATE, "SetBoundaryConditions: bnd_pot > 0 needs the potential-field kernels "
                                  "(mrm_potential_kernels)");
  if (!(pot == 1 || pt_update == 1 || !c->bext_valid)) return MRM_OK;
  TScope ts(c, T_POT);
  hipError_t e = pot::plan_update(c->pot, c->stream, c->G, U, c->d_bext);
  if (e != hipSuccess) return fail(c, MRM_ERR_HIP, "potential field update", e);
  c->bext_valid = true;
  return MRM_OK;
}

// SetBoundaryConditions on buffer U.  With P the ghost pressures are written
// too (fused step); with cmax the dim-0 ghost faces enter the cmax reduction.
static int bc_buf(mrm_ctx* c, double* U, double* P, unsigned long long* cmax, int pt_update) {
  int rc;
  if (c->bnd_ini) {
    if ((c->p_bc < 0.0) || (c->s_bc < 0.0)) {
      rc = set_p_s_bc(c, U);
      if (rc) return rc;
    }
    c->bnd_ini = 0;
  }
  double* dsum = (double*)(c->d_red + RED_SUM0);
  rc = bottom_sums(c, 0, U, dsum);
  if (rc) return rc;
  // bz0 is B_x(lend) before this call's top boundary (boundary_pdmp_1_fftw3.C:339-349)
  if ((rc = potential_due(c, U, pt_update))) return rc;
  BcParams B = bc_params(c);
  const long nyz = (long)c->G.n1 * c->G.n2;
  if (P) {
    TScope ts(c, T_BCBOT);
    k_bc_fused<<<nblk(2 * nyz), MRM_BLOCK, 0, c->stream>>>(c->G, c->E, B, dsum, U, P, c->inv_va2max, cmax);
    LAUNCHCHK(c, "k_bc_fused");
  } else {
    {
      TScope ts(c, T_BCBOT);
      k_bc_bottom<<<nblk(nyz), MRM_BLOCK, 0, c->stream>>>(c->G, c->E, B, dsum, U);
    }
    LAUNCHCHK(c, "k_bc_bottom");
    {
      TScope ts(c, T_BCTOP);
      k_bc_top<<<nblk(nyz), MRM_BLOCK, 0, c->stream>>>(c->G, B, U);
    }
    LAUNCHCHK(c, "k_bc_top");
  }
  return MRM_OK;
}

// SetBoundaryConditions on the mixed state M (ghost pressures with P, the
// dim-0 ghost faces in the cmax reduction with cmax)
static int bc_mx(mrm_ctx* c, MxState M, float* P, unsigned long long* cmax) {
  int rc;
  if (c->bnd_ini) {
    if ((c->p_bc < 0.0) || (c->s_bc < 0.0)) {
      rc = set_p_s_bc(c, nullptr, &M);
      if (rc) return rc;
    }
    c->bnd_ini = 0;
  }
  double* dsum = (double*)(c->d_red + RED_SUM0);
  rc = bottom_sums(c, 0, nullptr, dsum, &M);
  if (rc) return rc;
  BcParams B = bc_params(c);
  const long nyz = (long)c->G.n1 * c->G.n2;
  TScope ts(c, T_BCBOT);
  HIPCHK(c, mixed::launch_bc(c->stream, nblk(2 * nyz), c->G, c->E, B, dsum, M, P, c->inv_va2max, cmax));
  return MRM_OK;
}

extern "C" int mrm_set_boundary_conditions(mrm_ctx* c, const mrm_run_desc* run, int stage) {
  if (!c) return MRM_ERR_ARG;
  return mrm_set_boundary_conditions_pt(c, run, stage, pt_update_rule(c, run, stage));
}

extern "C" int mrm_set_boundary_conditions_pt(mrm_ctx* c, const mrm_run_desc* run, int stage, int pt_update) {
  (void)stage;
  if (!c) return MRM_ERR_ARG;
  // after an exchange the recorded ghost cells are the dim-0 layers, which
  // the boundaries write in every column from interior cells alone
  if (int rc = drop_settle_k(c, c->dU, c->gp_pend == 2)) return rc;
  if (!c->eos_ok) return fail(c, MRM_ERR_STATE, "mrm_set_boundary_conditions: no EOS table loaded");
  c->fast_valid = false;
  if (int rc = diag_track(c, run)) return rc;
  if (c->precision == MRM_PREC_FP32_MIXED) return bc_mx(c, c->M[c->cur], nullptr, nullptr);
  int rc = bc_buf(c, c->dU, nullptr, nullptr, pt_update);
  if (rc) return rc;
  c->gp_pend = 0;
  if (!c->d_spt) return rc;
  const long nyz = (long)c->G.n1 * c->G.n2;
  k_spitzer_bc<<<nblk(nyz), MRM_BLOCK, 0, c->stream>>>(c->G, c->phys.bnd[3] > 0.0, c->dF + MRM_SFLX * c->G.bs);
  LAUNCHCHK(c, "k_spitzer_bc");
  return MRM_OK;
}

// ------------------------------------------------------------------ potential field
// the kernels of the potential-field extrapolation (potential_sd_fftw3.C:83-183):
// plan, device b_ext, spectra
extern "C" int mrm_potential_kernels(mrm_ctx* c, const float* k8) {
  if (!c || !k8) return MRM_ERR_ARG;
  NEED_FP64(c, "mrm_potential_kernels");
  const int gx = c->grid.gsize[1], gy = c->grid.gsize[2];
  if (gy > 1 && c->dx[1] != c->dx[2])
    return fail(c, MRM_ERR_UNSUPPORTED, "mrm_potential_kernels: dx[1] = dx[2] required for the potential field");
  if (gx < c->G.g1 || gy < c->G.g2)
    return fail(c, MRM_ERR_UNSUPPORTED, "mrm_potential_kernels: horizontal extents below the ghost width");
  if (!c->pot) {
    std::string err;
    int rc = pot::plan_create(&c->pot, gx, gy, err);
    if (rc) return fail(c, rc, err.c_str());
  }
  if (!c->d_bext) {
    const size_t bytes = (size_t)8 * c->G.n1 * c->G.n2 * sizeof(double);
    HIPCHK(c, hipMalloc(&c->d_bext, bytes));
    HIPCHK(c, hipMemsetAsync(c->d_bext, 0, bytes, c->stream));
  }
  hipError_t e = pot::plan_kernels(c->pot, k8, c->stream);
  if (e != hipSuccess) return fail(c, MRM_ERR_HIP, "mrm_potential_kernels", e);
  c->bext_valid = false;
  return MRM_OK;
}

// the eight PSF-kernel-<c>.dat files of dir, headers checked against the grid
extern "C" int mrm_potential_kernels_load(mrm_ctx* c, const char* dir) {
  if (!c || !dir) return MRM_ERR_ARG;
  const int gx = c->grid.gsize[1], gy = c->grid.gsize[2];
  std::vector<float> k8((size_t)8 * gx * gy);
  const int rc = mrm_io_read_psf_kernels(dir, gx, gy, (float)(c->dx[0] / c->dx[1]), k8.data());
  if (rc) return fail(c, rc, "mrm_potential_kernels_load: missing PSF-kernel file or header inconsistent with the grid");
  return mrm_potential_kernels(c, k8.data());
}

extern "C" int mrm_get_potential_bext(mrm_ctx* c, double* out) {
  if (!c || !out) return MRM_ERR_ARG;
  if (!c->d_bext) return fail(c, MRM_ERR_STATE, "mrm_get_potential_bext: no potential-field kernels loaded");
  HIPCHK(c, hipMemcpyAsync(out, c->d_bext, (size_t)8 * c->G.n1 * c->G.n2 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRM_OK;
}

// ------------------------------------------------------------------ TVD
static int tvd_profiles(mrm_ctx* c, double h_bnd_bot, double h_bnd_top, double Qdiff_bnd) {
  // tvdlimit_SR.C:137-149, evaluated once
  const int n0 = c->G.n0;
  std::vector<double> prof(3 * n0);
  const double gxmax0 = c->grid.gxmax[0], lxmin0 = c->grid.gxmin[0], dx0 = c->dx[0];
  for (int i = 0; i < n0; i++) {
    double coord = lxmin0 + (0.5 + (i - c->G.lb0)) * dx0;
    double hh = coord / gxmax0;
    double vv = smaxh(0.0, 1.0 - 0.5 * (1.0 - hh) / h_bnd_top);
    prof[n0 + i] = sminh(1.0, vv * vv);                 // hft
    vv = smaxh(0.0, 1.0 - 0.5 * hh / h_bnd_bot);
    prof[i] = sminh(1.0, vv * vv);                      // hfb
    prof[2 * n0 + i] = (hh < 1 - 2 * h_bnd_top) ? 1.0 : Qdiff_bnd;   // qft
  }
  HIPCHK(c, hipMemcpy(c->d_prof, prof.data(), prof.size() * sizeof(double), hipMemcpyHostToDevice));
  c->tvd_ini = 0;
  return MRM_OK;
}

static int tvd_params(mrm_ctx* c, double dt_tvd, TvdParams& P) {
  const mrm_phys_desc& Ph = c->phys;
  P.rho_min = Ph.tchk[1];
  P.eps_min = Ph.tchk[0];
  P.rho_lev = pow(Ph.tvd[0], 2);
  P.rho_log = (int)Ph.tvd[1];
  P.q_rho_max = log(Ph.tvd[2]);
  P.visc_coeff_bot = sminh(1.0, Ph.tvd_visc_bnd[0]);
  P.visc_slope_bot = smaxh(0.0, Ph.tvd_visc_bnd[0] - 1.0);
  P.visc_coeff_top = sminh(1.0, Ph.tvd_visc_bnd[1]);
  P.visc_slope_top = smaxh(0.0, Ph.tvd_visc_bnd[1] - 1.0);
  P.eta_coeff_bot = sminh(1.0, Ph.tvd_eta_bnd[0]);
  P.eta_slope_bot = smaxh(0.0, Ph.tvd_eta_bnd[0] - 1.0);
  P.eta_coeff_top = sminh(1.0, Ph.tvd_eta_bnd[1]);
  P.eta_slope_top = smaxh(0.0, Ph.tvd_eta_bnd[1] - 1.0);
  P.B_par_diff = Ph.tvd[3];
  P.vhyp = Ph.tvd[4];
  P.tvd_pm_v = Ph.tvd[6];
  P.tvd_pm_B = Ph.tvd[7];
  P.vmax_lim = Ph.tvd[8];
  P.CME_thresh = Ph.tvd[9];
  for (int q = 0; q < 4; q++) { P.tvd_h[q] = Ph.tvd_h[q]; P.tvd_cs[q] = Ph.tvd_cs[q]; P.tc[q] = Ph.tvd_coeff[q]; }
  P.need_tvd_coeff = (P.tc[0] * P.tc[1] * P.tc[2] * P.tc[3] != 1.0);
  P.inv_va2max = c->inv_va2max;
  P.vsqr_diff = (P.vmax_lim < 1.0) ? pow(P.vmax_lim * c->v_lim, 2.0) : pow(P.vmax_lim, 2.0);
  P.dt_tvd = dt_tvd;
  P.is_gbeg0 = 1;
  P.is_gend0 = 1;
  if (c->tvd_ini) {
    double hb = (Ph.tvd_h_bnd[0] < 1) ? Ph.tvd_h_bnd[0] : Ph.tvd_h_bnd[0] / double(c->grid.gsize[0]);
    double ht = (Ph.tvd_h_bnd[1] < 1) ? Ph.tvd_h_bnd[1] : Ph.tvd_h_bnd[1] / double(c->grid.gsize[0]);
    int rc = tvd_profiles(c, hb, ht, Ph.tvd[5]);
    if (rc) return rc;
  }
  const int n0 = c->G.n0;
  P.hfb = c->d_prof;
  P.hft = c->d_prof + n0;
  P.qft = c->d_prof + 2 * n0;
  P.tau = c->tau_set ? c->dF + MRM_TAU * c->G.bs : nullptr;
  P.bcount = c->bcount_on ? c->d_bcount : nullptr;
  P.btie = c->btie;
  P.y_k0 = 0;        // every plane, ghosts included (tvdlimit_SR.C:199-205)
  P.y_nk = c->G.n2;
  return MRM_OK;
}

static size_t tvd_x_lds(const mrm_ctx* c) { return (size_t)(2 * MRM_NV + 5) * c->G.n0 * sizeof(double); }

// whether the dim-0 sweep runs one block per line (k_tvd_x_lines: the line's
// points, slopes and fluxes in LDS, lds bytes of it): lines of at most TX_MAXN
// points whose arrays fit one block's LDS (160 KB on gfx950); otherwise the
// marching kernel, one thread per line
static bool tvd_x_lines_fit(const mrm_ctx* c, size_t lds) {
  return !c->tvd_x_march && c->G.n0 <= TX_MAXN && lds <= c->lds_max;
}

// the three sweeps of TVDlimit on buffer U with pressure Pr; with epi the
// dim-2 sweep finishes each cell (TCheck, new pressure, cmax)
// the parameters of sweep d (tvdlimit_SR.C:41-86, 199-212)
static int tvd_sweep_params(mrm_ctx* c, double dt_tvd, int d, TvdParams& P) {
  int rc = tvd_params(c, dt_tvd, P);
  if (rc) return rc;
  P.cmax = 0.975 * c->dx[d] / dt_tvd;
  P.idxd1 = 1. / c->dx[d];
  return MRM_OK;
}

// With y_interior (the whole-step paths, which fill the ghosts before
// TVDlimit) and no Tau field, the dim-1 sweep runs on the interior planes only
// and the periodic copy fills the dim-2 ghost planes afterwards: those planes
// hold copies of interior planes (exchange and BC made them so, and the
// dim-0 sweep treats each line alone), so sweeping them would compute the
// same values, and the interior lines alone fill one round of the GPU's
// threads.  A caller-supplied Tau need not be periodic in its ghost planes,
// so with Tau every plane is swept as the reference does; a drop-in
// TVDlimit call gets whatever ghosts the caller left, and takes the interior
// sweep only after k_ghost2_periodic has found the ghost planes to be exact
// copies (ghosts_periodic).
static int tvd_buf(mrm_ctx* c, double* U, const double* Pr, double dt_tvd, const TvdEpilogue* epi,
                   bool contracted = false, int only_d = -1, bool y_interior = false) {
  TvdParams P;
  int rc = tvd_params(c, dt_tvd, P);
  if (rc) return rc;
  const Geom& G = c->G;
  const bool y_copy = y_interior && P.tau == nullptr;
  if (y_copy) {
    P.y_k0 = G.lb2;
    P.y_nk = c->lsize[2];
  }
  const double idx[3] = {1. / c->dx[0], 1. / c->dx[1], 1.0 / c->dx[2]};
  const bool tau = P.tau != nullptr;
  for (int d = 0; d < 3; d++) {
    if (only_d >= 0 && d != only_d) continue;
    P.cmax = 0.975 * c->dx[d] / dt_tvd;
    P.idxd1 = idx[d];
    long nl = d == 0 ? (long)G.n1 * G.n2 : d == 1 ? (long)c->lsize[0] * P.y_nk : (long)c->lsize[0] * c->lsize[1];
    unsigned nb = nblk(nl);
    {
      TScope ts(c, T_TVD0 + d);
      hipError_t e;
      size_t lds = tvd_x_lds(c);
      if (d == 0 && tvd_x_lines_fit(c, lds)) {
        TvdLineArgs A{G, P};
        unsigned bsz = (unsigned)((G.n0 + 63) / 64 * 64);
        e = contracted ? fast::p_launch_tvd_x_lines(c->stream, (unsigned)nl, bsz, lds, A, U, Pr)
                       : exact::launch_tvd_x_lines(c->stream, (unsigned)nl, bsz, lds, A, U, Pr);
      } else {
        const TvdEpilogue* x = d == 2 ? epi : nullptr;
        e = contracted ? fast::p_launch_tvd_sweep(c->stream, d, tau, x, nb, G, P, U, Pr)
                       : exact::launch_tvd_sweep(c->stream, d, tau, x, nb, G, P, U, Pr);
      }
      if (e != hipSuccess) return fail(c, MRM_ERR_HIP, "tvd sweep", e);
    }
    if (d == 1 && y_copy) {
      // the copy of the dim-2 ghost planes counts as an exchange launch
      TScope ts(c, T_EXCH);
      const long n2 = (long)c->lsize[0] * G.n1 * G.g2;
      k_exchange_p<2><<<nblk(n2), MRM_BLOCK, 0, c->stream>>>(G, U, nullptr);
      LAUNCHCHK(c, "k_exchange_p<2>");
    }
  }
  return MRM_OK;
}

// whether U's and Pr's dim-2 ghost planes are exact copies of their periodic
// images (one device pass over the ghost planes and a host read of the flag)
static int ghosts_periodic(mrm_ctx* c, const double* U, const double* Pr, bool* yes) {
  unsigned long long* flag = c->d_red + RED_SCRATCH;
  HIPCHK(c, hipMemsetAsync(flag, 0, sizeof(unsigned long long), c->stream));
  const long n = (long)c->G.n0 * c->G.n1 * c->G.g2;
  k_ghost2_periodic<<<nblk(n), MRM_BLOCK, 0, c->stream>>>(c->G, U, Pr, flag);
  LAUNCHCHK(c, "k_ghost2_periodic");
  unsigned long long h = 1;
  HIPCHK(c, hipMemcpyAsync(&h, flag, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *yes = h == 0;
  return MRM_OK;
}

// the three sweeps with the diagnostics of every updated point: the marching
// kernels in every dimension, exact build
static int tvd_buf_diag(mrm_ctx* c, double* U, const double* Pr, double dt_tvd, const DiagArgs& DA) {
  TvdParams P;
  int rc = tvd_params(c, dt_tvd, P);
  if (rc) return rc;
  const Geom& G = c->G;
  const double idx[3] = {1. / c->dx[0], 1. / c->dx[1], 1.0 / c->dx[2]};
  for (int d = 0; d < 3; d++) {
    P.cmax = 0.975 * c->dx[d] / dt_tvd;
    P.idxd1 = idx[d];
    long nl = d == 0 ? (long)G.n1 * G.n2 : d == 1 ? (long)c->lsize[0] * G.n2 : (long)c->lsize[0] * c->lsize[1];
    TScope ts(c, T_TVD0 + d);
    hipError_t e = exact::launch_tvd_sweep_diag(c->stream, d, P.tau != nullptr, nblk(nl), G, P, U, Pr, DA);
    if (e != hipSuccess) return fail(c, MRM_ERR_HIP, "tvd sweep", e);
  }
  return MRM_OK;
}

// limiter branch accounting: the sweeps of tvd_buf run their counting
// instantiations while it is on (the FP32-mixed sweeps and the diagnostics
// sweeps are not counted)
extern "C" int mrm_branch_count(mrm_ctx* c, int on, double tie) {
  if (!c || !(tie >= 0.0)) return MRM_ERR_ARG;
  if (int rc = drop_settle_k(c, nullptr)) return rc;   // a recorded sweep counts under the old setting
  c->btie = tie;
  if (on) {
    NEED_FP64(c, "mrm_branch_count");
    if (!c->d_bcount)
      HIPCHK(c, hipMalloc((void**)&c->d_bcount, MRM_BR_NSLOT * sizeof(unsigned long long)));
    HIPCHK(c, hipMemsetAsync(c->d_bcount, 0, MRM_BR_NSLOT * sizeof(unsigned long long), c->stream));
  }
  c->bcount_on = on != 0;
  return MRM_OK;
}

extern "C" int mrm_get_branch_counts(mrm_ctx* c, unsigned long long out[MRM_BR_NSLOT]) {
  if (!c || !out) return MRM_ERR_ARG;
  if (int rc = drop_settle_k(c, nullptr)) return rc;
  if (!c->d_bcount) {
    memset(out, 0, MRM_BR_NSLOT * sizeof(unsigned long long));
    return MRM_OK;
  }
  HIPCHK(c, hipMemcpyAsync(out, c->d_bcount, MRM_BR_NSLOT * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRM_OK;
}

extern "C" int mrm_tvdlimit(mrm_ctx* c, const mrm_run_desc* run, double dt_tvd) {
  if (!c || !run) return MRM_ERR_ARG;
  if (int rc = drop_settle_k(c, c->dU)) return rc;
  NEED_FP64(c, "mrm_tvdlimit");
  c->fast_valid = false;
  DiagArgs DA;
  int rc = diag_args(c, run, DA, dt_tvd);
  if (rc) return rc;
  if (DA.dg) return tvd_buf_diag(c, c->dU, c->dF + MRM_PRES * c->G.bs, dt_tvd, DA);
  const double* Pr = c->dF + MRM_PRES * c->G.bs;
  bool periodic = false;
  if (!c->tau_set && (rc = ghosts_periodic(c, c->dU, Pr, &periodic))) return rc;
  if (!c->defer) return tvd_buf(c, c->dU, Pr, dt_tvd, nullptr, false, -1, periodic);
  // the dim-0 and dim-1 sweeps now, the dim-2 sweep recorded
  if ((rc = tvd_buf(c, c->dU, Pr, dt_tvd, nullptr, false, 0, periodic))) return rc;
  if ((rc = tvd_buf(c, c->dU, Pr, dt_tvd, nullptr, false, 1, periodic))) return rc;
  c->tvd2_def = true;
  c->tvd2_dt = dt_tvd;
  c->tvd2_periodic = periodic;
  return MRM_OK;
}

// ------------------------------------------------------------------ full step
// the reference's call sequence, one drop-in operation after the other
// with diag, stage 4 and everything after it collect the diagnostics, zeroed
// first (solver.C:264-296); run->need_diagnostics is left set
static int step_sequence(mrm_ctx* c, mrm_run_desc* run, double dt_rad, bool diag = false) {
  int rc;
#define STEPCHK(x) do { rc = (x); if (rc) return rc; } while (0)
  double cfl = 0.0;
  STEPCHK(mrm_save_cons(c));
  for (int stage = 1; stage <= 4; stage++) {
    STEPCHK(mrm_cons_to_prim(c));
    if (stage == 1) STEPCHK(mrm_adjust_valf_max(c, run, nullptr));
    if (diag && stage == 4) {
      run->need_diagnostics = 1;
      STEPCHK(mrm_diag_zero(c));
    }
    if (stage == 1) {
      double dt_mhd;
      STEPCHK(mrm_mhd_residual(c, run, &dt_mhd));
      STEPCHK(mrm_sync_timestep(c, run, dt_rad, dt_mhd, &cfl));
    } else {
      STEPCHK(mrm_mhd_residual(c, run, nullptr));   // dt_cfl of stages 2-4 is not used (solver.C:302-309)
    }
    STEPCHK(mrm_source_integrate_tcheck(c, run, stage));
    STEPCHK(mrm_exchange_grid(c));
    STEPCHK(mrm_set_boundary_conditions(c, run, stage));
  }
  if (cfl <= run->CFL_tvd) {
    STEPCHK(mrm_tvdlimit(c, run, run->dt));
  } else {
    STEPCHK(mrm_tvdlimit(c, run, 0.5 * run->dt));
    STEPCHK(mrm_exchange_grid(c));
    STEPCHK(mrm_tvdlimit(c, run, 0.5 * run->dt));
  }
  if (c->phys.divB[0] > 0.0) STEPCHK(mrm_clean_divB(c, run, nullptr));
  STEPCHK(mrm_tcheck(c, run));
  STEPCHK(mrm_exchange_grid(c));
  STEPCHK(mrm_set_boundary_conditions(c, run, 5));
#undef STEPCHK
  return MRM_OK;
}

// Adjust_Valf_Max with va_adjust = 0 (Adjust_Valf_Max.C:136-144): its
// reductions only feed the va_adjust branch, the limits are constants
static void valf_constants(mrm_ctx* c) {
  double va_max = c->phys.params[1];
  c->v_lim = c->phys.tchk[3];
  c->e_lim = c->phys.tchk[2];
  c->inv_va2max = 1.0 / pow(va_max, 2);
  c->va_max_old = va_max;
}

// Spitzer conduction in the fused step, one stage: ConsToPrim's temperature
// of the stage input, then the conduction terms from sflx_in (constants from
// dt and CFL of the stage's residual) into d_spt + 3 bs for the stage kernel
// and sflx_out = sflx0 + wdt Rflx
static int spitzer_stage(mrm_ctx* c, double dt_res, double cfl, double wdt, const double* U, const double* sflx_in,
                         double* sflx_out) {
  const Geom& G = c->G;
  const SptParams S = spitzer_params(c, dt_res, cfl);
  double* T = c->dF + MRM_TEMP * G.bs;
  unsigned nb = nblk(G.bs);
  if (nb > 65536u) nb = 65536u;
  {
    TScope ts(c, T_EOST);
    k_eos_temp<<<nb, MRM_BLOCK, 0, c->stream>>>(G, c->E, U, T);
  }
  LAUNCHCHK(c, "k_eos_temp");
  TScope ts(c, T_SPT);
  k_spitzer_stage<<<spitzer_blocks(G), SPT_TX * SPT_TY, 0, c->stream>>>(G, S, wdt, SptStateF64{U, G.bs}, T, sflx_in,
                                                                        c->d_spt, sflx_out, c->d_spt + 3 * G.bs);
  LAUNCHCHK(c, "k_spitzer_stage");
  return MRM_OK;
}

// ghosts of a heat-flux buffer: periodic copies, then the vertical BC
// (exchange.C:111-120; boundary_pdmp_1_fftw3.C:306-310, 455-466)
static int spitzer_ghosts(mrm_ctx* c, double* f) {
  int rc = exchange_arrays(c, f, 1);
  if (rc) return rc;
  const long nyz = (long)c->G.n1 * c->G.n2;
  k_spitzer_bc<<<nblk(nyz), MRM_BLOCK, 0, c->stream>>>(c->G, c->phys.bnd[3] > 0.0, f);
  LAUNCHCHK(c, "k_spitzer_bc");
  return MRM_OK;
}

static int stage_launch(mrm_ctx* c, const mrm_run_desc* run, int stage, const double* Uin, const double* Pin,
                        const double* U0, double* Uout, double* Pout, const double* eres, double rp_va) {
  StageArgs A;
  A.eres = eres;
  A.G = c->G;
  A.E = c->E;
  for (int d = 0; d < 3; d++) {
    A.RP.w1[d] = 8. / (12. * c->dx[d]);
    A.RP.w2[d] = -1. / (12. * c->dx[d]);
  }
  A.RP.inv_va2max = rp_va < 0.0 ? c->inv_va2max : rp_va;
  eta_params(c, A.RP);
  A.SP = src_params(c, run, stage);
  A.ntx = (c->lsize[0] + ST_SX - 1) / ST_SX;
  A.nty = (c->lsize[1] + ST_SY - 1) / ST_SY;
  A.nkc = (c->lsize[2] + ST_KC - 1) / ST_KC;
  A.write_p = Pout != nullptr;
  const double* Q = c->qtot_set ? c->dF + MRM_QTOT * c->G.bs : nullptr;
  unsigned nb = (unsigned)((long)A.ntx * A.nty * A.nkc);
  TScope ts(c, T_NAMED);
  hipError_t e = c->fma ? fast::p_launch_stage(c->stream, nb, c->stage_variant, A, Uin, Pin, U0, Uout, Pout, Q)
                        : exact::launch_stage(c->stream, nb, c->stage_variant, A, Uin, Pin, U0, Uout, Pout, Q);
  if (e != hipSuccess) return fail(c, MRM_ERR_HIP, "k_stage", e);
  return MRM_OK;
}

// One time step as device-resident fused kernels: four stage kernels with
// ghost fills between them, three TVD sweeps (the last one finishing each
// cell), and the final ghost fill.  Pressures travel with the state, so
// ConsToPrim never runs as a separate pass; the only host round trip is the
// dt reduction (solver.C:487-548 needs it on the host).
static int step_fused(mrm_ctx* c, mrm_run_desc* run, double dt_rad) {
  const Geom& G = c->G;
  double* Pc = c->dF + MRM_PRES * G.bs;
  double* Pa = c->dP2;
  unsigned long long* cmax = c->d_red + RED_CMAX;
  int rc;
  // va_adjust > 0: Adjust_Valf_Max's statistics of the step's input come
  // from the previous step's last TVD sweep (or from k_eos_p), its update
  // runs on the host, and cmax is reduced again with the new inv_va2max
  const bool va_adj = c->phys.params[2] > 0.0;
  if (!va_adj) valf_constants(c);
  if (!c->fast_valid) {
    HIPCHK(c, hipMemsetAsync(c->d_red, 0, (RED_CNT_E + 1) * sizeof(unsigned long long), c->stream));
    unsigned nb = nblk(G.bs);
    if (nb > 65536u) nb = 65536u;
    TScope ts(c, T_NAMED + 1);
    k_eos_p<<<nb, MRM_BLOCK, 0, c->stream>>>(G, c->E, c->S[c->cur], Pc, c->inv_va2max, cmax,
                                             va_adj ? c->d_red : nullptr, 0.95 * c->v_lim, 0.95 * c->e_lim);
    LAUNCHCHK(c, "k_eos_p");
  }
  // one round trip brings cmax, the limiter statistics and the previous
  // Clean_divB's maxima
  double cm;
  const int nread = c->divb_pending ? RED_DIVB1 + 1 : va_adj ? RED_CNT_E + 1 : 1;
  HIPCHK(c, hipMemcpyAsync(c->h_red, c->d_red, nread * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->divb_pending) divb_settle_host(c, nullptr);
  cm = bits_to_double(c->h_red[RED_CMAX]);
  if (va_adj) {
    valf_apply(c, nullptr);
    HIPCHK(c, hipMemsetAsync(cmax, 0, sizeof(unsigned long long), c->stream));
    unsigned nb = nblk(G.bs);
    if (nb > 65536u) nb = 65536u;
    { TScope ts(c, T_VALF);
      k_cmax_visited<<<nb, MRM_BLOCK, 0, c->stream>>>(G, c->S[c->cur], Pc, c->inv_va2max, cmax); }
    LAUNCHCHK(c, "k_cmax_visited");
    if ((rc = read_cmax(c, &cm))) return rc;
  }
  double cfl = 0.0;
  // stage 1's conduction constants use Run.dt before Sync_timestep updates it
  // (MHD_Residual runs first, mhdres_SR.C:136-145)
  const double dt_prev = run->dt;
  if ((rc = mrm_sync_timestep(c, run, dt_rad, eta_dt_limit(c, run, c->dxmin / cm), &cfl))) return rc;
  const int i0 = c->cur, i1 = (c->cur + 1) % 3, i2 = (c->cur + 2) % 3;
  double* U0 = c->S[i0];
  // the damping profile of the step from U^n (stage 1's state)
  if ((rc = damping_update(c, U0, false))) return rc;
  struct { double *in, *pin, *out, *pout; } st[4] = {
      {c->S[i0], Pc, c->S[i1], Pa}, {c->S[i1], Pa, c->S[i2], Pc},
      {c->S[i2], Pc, c->S[i1], Pa}, {c->S[i1], Pa, c->S[i2], nullptr}};
  // Spitzer conduction: sflx alternates between MRM_SFLX and a second buffer
  // (an even number of stages ends in MRM_SFLX); sflx0 is saved first
  const bool spt = c->d_spt != nullptr;
  double* sflx[2] = {c->dF + MRM_SFLX * G.bs, c->d_spt + 2 * G.bs};
  if (spt) {
    TScope ts(c, T_SPT);
    k_spitzer_save<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(G, sflx[0], c->d_spt);
    LAUNCHCHK(c, "k_spitzer_save");
  }
  for (int s = 0; s < 4; s++) {
    if (spt && (rc = spitzer_stage(c, s == 0 ? dt_prev : run->dt, run->CFL, run->dt / double(4 - s), st[s].in,
                                   sflx[s % 2], sflx[(s + 1) % 2])))
      return rc;
    if ((rc = stage_launch(c, run, s + 1, st[s].in, st[s].pin, U0, st[s].out, st[s].pout,
                           spt ? c->d_spt + 3 * G.bs : nullptr)))
      return rc;
    if ((rc = exchange_buf(c, st[s].out, st[s].pout))) return rc;
    if ((rc = bc_buf(c, st[s].out, st[s].pout, nullptr, 0))) return rc;
    if (spt && (rc = spitzer_ghosts(c, sflx[(s + 1) % 2]))) return rc;
  }
  // TVD on U4 with the pressure of the stage-4 input (tvdlimit_SR.C:332 reads
  // Grid.pres of the last ConsToPrim)
  double* Uf = c->S[i2];
  HIPCHK(c, hipMemsetAsync(c->d_red, 0, (va_adj ? RED_CNT_E + 1 : 1) * sizeof(unsigned long long), c->stream));
  TvdEpilogue X;
  X.E = c->E;
  X.rho_min = c->phys.tchk[1];
  X.eps_min = c->phys.tchk[0];
  X.vmax = (c->v_lim > 0.0) ? c->v_lim : c->phys.tchk[3];
  X.eps_max = (c->e_lim > 0.0) ? c->e_lim : c->phys.tchk[2];
  X.inv_va2max = c->inv_va2max;
  X.Pnew = Pc;
  X.cmax = cmax;
  X.valf = va_adj ? c->d_red : nullptr;
  X.vlim95 = 0.95 * c->v_lim;
  X.elim95 = 0.95 * c->e_lim;
  X.skip_top_cmax = (int)c->phys.bnd[1] > 0;
  c->last_dt = run->dt;
  c->last_split = !(cfl <= run->CFL_tvd);
  // with Clean_divB (between TVDlimit and TCheck, solver.C:341-349) its last
  // correction pass finishes the cells instead of the last TVD sweep
  const bool divb = c->phys.divB[0] > 0.0;
  const TvdEpilogue* tvd_epi = divb ? nullptr : &X;
  if (cfl <= run->CFL_tvd) {
    if ((rc = tvd_buf(c, Uf, Pa, run->dt, tvd_epi, c->fma, -1, true))) return rc;
  } else {
    if ((rc = tvd_buf(c, Uf, Pa, 0.5 * run->dt, nullptr, c->fma, -1, true))) return rc;
    if ((rc = exchange_buf(c, Uf, nullptr))) return rc;
    if ((rc = tvd_buf(c, Uf, Pa, 0.5 * run->dt, tvd_epi, c->fma, -1, true))) return rc;
  }
  if (divb && (rc = clean_divb_dev(c, Uf, &X))) return rc;
  if ((rc = exchange_buf(c, Uf, Pc))) return rc;
  if ((rc = bc_buf(c, Uf, Pc, cmax, pt_update_rule(c, run, 5)))) return rc;
  c->cur = i2;
  c->c0 = (i2 + 1) % 3;
  c->dU = c->S[i2];
  c->dU0 = c->S[c->c0];
  c->fast_valid = true;
  return MRM_OK;
}

// ------------------------------------------------------------------ FP32-mixed step
static MxTvdParams mx_tvd_params(const TvdParams& P) {
  MxTvdParams Q;
  Q.rho_min = (float)P.rho_min;
  Q.rho_lev = (float)P.rho_lev;
  Q.q_rho_max = (float)P.q_rho_max;
  Q.visc_coeff_bot = (float)P.visc_coeff_bot; Q.visc_slope_bot = (float)P.visc_slope_bot;
  Q.visc_coeff_top = (float)P.visc_coeff_top; Q.visc_slope_top = (float)P.visc_slope_top;
  Q.eta_coeff_bot = (float)P.eta_coeff_bot; Q.eta_slope_bot = (float)P.eta_slope_bot;
  Q.eta_coeff_top = (float)P.eta_coeff_top; Q.eta_slope_top = (float)P.eta_slope_top;
  Q.B_par_diff = (float)P.B_par_diff; Q.vhyp = (float)P.vhyp;
  Q.tvd_pm_v = (float)P.tvd_pm_v; Q.tvd_pm_B = (float)P.tvd_pm_B;
  Q.vmax_lim = (float)P.vmax_lim; Q.CME_thresh = (float)P.CME_thresh;
  for (int q = 0; q < 4; q++) {
    Q.tvd_h[q] = (float)P.tvd_h[q]; Q.tvd_cs[q] = (float)P.tvd_cs[q]; Q.tc[q] = (float)P.tc[q];
  }
  Q.inv_va2max = (float)P.inv_va2max;
  Q.vsqr_diff = (float)P.vsqr_diff;
  Q.cmax = (float)P.cmax;
  Q.idxd1 = (float)P.idxd1;
  Q.dt_tvd = (float)P.dt_tvd;
  Q.eps_min = P.eps_min;
  Q.need_tvd_coeff = P.need_tvd_coeff; Q.rho_log = P.rho_log;
  Q.is_gbeg0 = P.is_gbeg0; Q.is_gend0 = P.is_gend0;
  Q.y_k0 = P.y_k0; Q.y_nk = P.y_nk;
  Q.hfb = P.hfb; Q.hft = P.hft; Q.qft = P.qft; Q.tau = P.tau;
  return Q;
}

// the three TVD sweeps on the mixed state; with epi the dim-2 sweep finishes
// each cell (TCheck, new pressure, cmax)
static int tvd_mx(mrm_ctx* c, MxState M, const float* Pr, double dt_tvd, const MxEpilogue* epi) {
  TvdParams P;
  int rc = tvd_params(c, dt_tvd, P);
  if (rc) return rc;
  const Geom& G = c->G;
  const double idx[3] = {1. / c->dx[0], 1. / c->dx[1], 1.0 / c->dx[2]};
  const bool tau = P.tau != nullptr;
  for (int d = 0; d < 3; d++) {
    P.cmax = 0.975 * c->dx[d] / dt_tvd;
    P.idxd1 = idx[d];
    const bool y_copy = P.tau == nullptr;   // as in tvd_buf
    if (y_copy) {
      P.y_k0 = G.lb2;
      P.y_nk = c->lsize[2];
    }
    MxTvdParams Q = mx_tvd_params(P);
    long nl = d == 0 ? (long)G.n1 * G.n2 : d == 1 ? (long)c->lsize[0] * P.y_nk : (long)c->lsize[0] * c->lsize[1];
    TScope ts(c, T_TVD0 + d);
    hipError_t e;
    size_t lds = (size_t)(3 * MRM_NV + 5) * G.n0 * sizeof(float);
    if (d == 0 && tvd_x_lines_fit(c, lds)) {
      unsigned bsz = (unsigned)((G.n0 + 63) / 64 * 64);
      e = mixed::launch_tvd_x_lines(c->stream, (unsigned)nl, bsz, lds, G, Q, M, Pr);
    } else {
      e = mixed::launch_tvd_sweep(c->stream, d, tau, d == 2 ? epi : nullptr, nblk(nl), G, Q, M, Pr);
    }
    if (e != hipSuccess) return fail(c, MRM_ERR_HIP, "tvd sweep (mixed)", e);
    if (d == 1 && y_copy) HIPCHK(c, mixed::launch_exchange_dim2(c->stream, G, M));
  }
  return MRM_OK;
}

static int stage_mx(mrm_ctx* c, const mrm_run_desc* run, int stage, MxState in, const float* Pin, MxState s0,
                    MxState out, float* Pout, const double* eres = nullptr) {
  MxStageArgs A;
  A.eres = eres;
  A.G = c->G;
  A.E = c->E;
  for (int d = 0; d < 3; d++) {
    A.w1[d] = (float)(8. / (12. * c->dx[d]));
    A.w2[d] = (float)(-1. / (12. * c->dx[d]));
  }
  A.inv_va2max = (float)c->inv_va2max;
  SrcParams SP = src_params(c, run, stage);
  A.wdt = SP.wdt;
  A.invRdt = (float)SP.invRdt;
  A.g = (float)SP.g;
  A.rho_min = (float)SP.rho_min;
  A.vmax = (float)SP.vmax;
  A.eps_min = SP.eps_min;
  A.eps_max = SP.eps_max;
  A.ntx = (c->lsize[0] + MX_SX - 1) / MX_SX;
  A.nty = (c->lsize[1] + MX_SY - 1) / MX_SY;
  A.nkc = (c->lsize[2] + MX_KC - 1) / MX_KC;
  A.write_p = Pout != nullptr;
  {
    ResidParams RP;
    eta_params(c, RP);
    A.eta = (float)RP.eta;
    for (int d = 0; d < 3; d++) {
      A.ww0[d] = (float)RP.ww0[d];
      A.ww1[d] = (float)RP.ww1[d];
      A.ww2[d] = (float)RP.ww2[d];
    }
  }
  const double* Q = c->qtot_set ? c->dF + MRM_QTOT * c->G.bs : nullptr;
  unsigned nb = (unsigned)((long)A.ntx * A.nty * A.nkc);
  TScope ts(c, T_NAMED);
  HIPCHK(c, mixed::launch_stage(c->stream, nb, A, in, Pin, s0, out, Pout, Q));
  return MRM_OK;
}

static int step_mixed_body(mrm_ctx* c, const mrm_run_desc* run, double dt_prev, bool split);

// One time step in the FP32-mixed form: the fused step's schedule on mixed
// state buffers (FP32 rho, M, B, p; FP64 e), with the cmax reduction and dt
// on the host in FP64.
static int step_mixed(mrm_ctx* c, mrm_run_desc* run, double dt_rad) {
  const Geom& G = c->G;
  float* Pc = c->Pm[0];
  float* Pa = c->Pm[1];
  unsigned long long* cmax = c->d_red + RED_CMAX;
  int rc;
  // va_adjust > 0 as in step_fused: statistics from the previous step's last
  // TVD sweep (or k_eos_p_mx), host update, cmax again with the new limit
  const bool va_adj = c->phys.params[2] > 0.0;
  if (!va_adj) valf_constants(c);
  if (!c->fast_valid) {
    HIPCHK(c, hipMemsetAsync(c->d_red, 0, (RED_CNT_E + 1) * sizeof(unsigned long long), c->stream));
    unsigned nb = nblk(G.bs);
    if (nb > 65536u) nb = 65536u;
    TScope ts(c, T_NAMED + 1);
    HIPCHK(c, mixed::launch_eos_p(c->stream, nb, G, c->E, c->M[c->cur], Pc, c->inv_va2max, cmax,
                                  va_adj ? c->d_red : nullptr, 0.95 * c->v_lim, 0.95 * c->e_lim));
  }
  // one round trip brings cmax, the limiter statistics and the previous
  // Clean_divB's maxima
  double cm;
  const int nread = c->divb_pending ? RED_DIVB1 + 1 : va_adj ? RED_CNT_E + 1 : 1;
  HIPCHK(c, hipMemcpyAsync(c->h_red, c->d_red, nread * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->divb_pending) divb_settle_host(c, nullptr);
  cm = bits_to_double(c->h_red[RED_CMAX]);
  if (va_adj) {
    valf_apply(c, nullptr);
    HIPCHK(c, hipMemsetAsync(cmax, 0, sizeof(unsigned long long), c->stream));
    unsigned nb = nblk(G.bs);
    if (nb > 65536u) nb = 65536u;
    {
      TScope ts(c, T_VALF);
      HIPCHK(c, mixed::launch_cmax(c->stream, nb, G, c->M[c->cur], Pc, c->inv_va2max, cmax));
    }
    if ((rc = read_cmax(c, &cm))) return rc;
  }
  double cfl = 0.0;
  const double dt_prev = run->dt;   // stage 1's conduction constants (mhdres_SR.C:136-145)
  if ((rc = mrm_sync_timestep(c, run, dt_rad, eta_dt_limit(c, run, c->dxmin / cm), &cfl))) return rc;
  c->last_dt = run->dt;
  c->last_split = !(cfl <= run->CFL_tvd);
  return step_mixed_body(c, run, dt_prev, c->last_split);
}

// the stages, TVDlimit and ghost fills of a mixed step whose dt (run->dt) and
// TVD split are decided; the current state's pressure must be valid
static int step_mixed_body(mrm_ctx* c, const mrm_run_desc* run, double dt_prev, bool split) {
  const Geom& G = c->G;
  float* Pc = c->Pm[0];
  float* Pa = c->Pm[1];
  unsigned long long* cmax = c->d_red + RED_CMAX;
  int rc;
  const bool va_adj = c->phys.params[2] > 0.0;
  const int i0 = c->cur, i1 = (c->cur + 1) % 3, i2 = (c->cur + 2) % 3;
  MxState* M = c->M;
  struct { int in; float* pin; int out; float* pout; } st[4] = {
      {i0, Pc, i1, Pa}, {i1, Pa, i2, Pc}, {i2, Pc, i1, Pa}, {i1, Pa, i2, nullptr}};
  // Spitzer conduction as in step_fused: sflx alternates between two FP64
  // buffers, the conduction kernels read the mixed state
  const bool spt = c->d_spt != nullptr;
  double* sflx[2] = {c->dF + MRM_SFLX * G.bs, c->d_spt + 2 * G.bs};
  double* eres = spt ? c->d_spt + 3 * G.bs : nullptr;
  if (spt) {
    TScope ts(c, T_SPT);
    k_spitzer_save<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(G, sflx[0], c->d_spt);
    LAUNCHCHK(c, "k_spitzer_save");
  }
  for (int s = 0; s < 4; s++) {
    if (spt) {
      const SptParams S = spitzer_params(c, s == 0 ? dt_prev : run->dt, run->CFL);
      double* T = c->dF + MRM_TEMP * G.bs;
      {
        TScope ts(c, T_EOST);
        HIPCHK(c, mixed::launch_eos_temp(c->stream, G, c->E, M[st[s].in], T));
      }
      TScope ts(c, T_SPT);
      HIPCHK(c, mixed::launch_spitzer_stage(c->stream, G, S, run->dt / double(4 - s), M[st[s].in], T, sflx[s % 2],
                                            c->d_spt, sflx[(s + 1) % 2], eres));
    }
    if ((rc = stage_mx(c, run, s + 1, M[st[s].in], st[s].pin, M[i0], M[st[s].out], st[s].pout, eres))) return rc;
    if ((rc = exchange_mx(c, M[st[s].out], st[s].pout))) return rc;
    if ((rc = bc_mx(c, M[st[s].out], st[s].pout, nullptr))) return rc;
    if (spt && (rc = spitzer_ghosts(c, sflx[(s + 1) % 2]))) return rc;
  }
  HIPCHK(c, hipMemsetAsync(c->d_red, 0, (va_adj ? RED_CNT_E + 1 : 1) * sizeof(unsigned long long), c->stream));
  MxEpilogue X;
  X.E = c->E;
  X.rho_min = (float)c->phys.tchk[1];
  X.eps_min = c->phys.tchk[0];
  X.vmax = (float)((c->v_lim > 0.0) ? c->v_lim : c->phys.tchk[3]);
  X.eps_max = (c->e_lim > 0.0) ? c->e_lim : c->phys.tchk[2];
  X.inv_va2max = c->inv_va2max;
  X.Pnew = Pc;
  X.cmax = cmax;
  X.valf = va_adj ? c->d_red : nullptr;
  X.vlim95 = 0.95 * c->v_lim;
  X.elim95 = 0.95 * c->e_lim;
  // with Clean_divB its last correction pass finishes the cells
  const bool divb = c->phys.divB[0] > 0.0;
  const MxEpilogue* tvd_epi = divb ? nullptr : &X;
  if (!split) {
    if ((rc = tvd_mx(c, M[i2], Pa, run->dt, tvd_epi))) return rc;
  } else {
    if ((rc = tvd_mx(c, M[i2], Pa, 0.5 * run->dt, nullptr))) return rc;
    if ((rc = exchange_mx(c, M[i2], nullptr))) return rc;
    if ((rc = tvd_mx(c, M[i2], Pa, 0.5 * run->dt, tvd_epi))) return rc;
  }
  if (divb && (rc = clean_divb_mx(c, M[i2], &X))) return rc;
  if ((rc = exchange_mx(c, M[i2], Pc))) return rc;
  if ((rc = bc_mx(c, M[i2], Pc, cmax))) return rc;
  c->cur = i2;
  c->fast_valid = true;
  return MRM_OK;
}

extern "C" int mrm_step(mrm_ctx* c, mrm_run_desc* run, double dt_rad, int nsteps) {
  if (!c || !run || nsteps < 0) return MRM_ERR_ARG;
  // the fused step rotates every state buffer; the per-operation sequence's
  // calls settle what they touch themselves (ConsToPrim's pending outputs
  // stay pending, as between the reference's own calls)
  if (int rc = c->force_sequence ? drop_settle_k(c, nullptr) : drop_settle(c)) return rc;
  if (!c->eos_ok) return fail(c, MRM_ERR_STATE, "mrm_step: no EOS table loaded");
  const bool mx = c->precision == MRM_PREC_FP32_MIXED;
  const int want_diag = run->need_diagnostics;
  if (want_diag && mx) return fail(c, MRM_ERR_UNSUPPORTED, "mrm_step: diagnostics need an FP64 context");
  // Spitzer conduction, and the step that collects diagnostics, run through
  // the per-operation kernels
  const bool fused = !c->force_sequence;
  run->need_diagnostics = 0;
  int rc = MRM_OK;
  for (int it = 0; it < nsteps; it++) {
    const bool diag = want_diag && it == nsteps - 1;
    if (!diag) c->diag_last = 0;
    rc = mx ? step_mixed(c, run, dt_rad) : (fused && !diag) ? step_fused(c, run, dt_rad)
                                                            : step_sequence(c, run, dt_rad, diag);
    if (rc) break;
    run->iteration++;
    run->globiter++;
    run->time += run->dt;
  }
  run->need_diagnostics = want_diag;
  return rc;
}

// ------------------------------------------------------------------ adjoint
// Reverse mode of mrm_step (FP64, fused schedule): mrm_adjoint_forward runs
// the steps, keeping the state every `every` steps and each step's dt and
// TVD branch; mrm_adjoint_backward walks the segments from the last one,
// recomputes each segment's step states from its checkpoint, and applies the
// transpose of each step from the end.  dt, the TVD branch, p_bc/s_bc, Qtot
// and Tau are held at their forward values.

// work buffers, FP64 SoA over the grid buffer
struct AdjBufs {
  double *X[5], *P[4], *V[6], *W, *Z, *seg, *lamA, *lamB, *lamX0, *lamP, *lamP3, *Gb, *lWc, *T, *small;
};

static int adj_buffers(mrm_ctx* c, AdjBufs& B) {
  const size_t bs = c->G.bs, f8 = 8 * bs;
  const size_t nT = (size_t)adj::tvd_scratch_fields();
  const size_t n = 5 * f8 + 4 * bs + 6 * f8 + 2 * f8 + (size_t)c->adj_every * f8 + 3 * f8 + 2 * bs + 39 * bs +
                   8 * bs + nT * bs + 64;
  const size_t bytes = n * sizeof(double);
  if (c->adj_bytes < bytes) {
    hipFree(c->adj_mem);
    c->adj_mem = nullptr;
    c->adj_bytes = 0;
    HIPCHK(c, hipMalloc(&c->adj_mem, bytes));
    c->adj_bytes = bytes;
  }
  double* p = c->adj_mem;
  for (int q = 0; q < 5; q++) { B.X[q] = p; p += f8; }
  for (int q = 0; q < 4; q++) { B.P[q] = p; p += bs; }
  for (int q = 0; q < 6; q++) { B.V[q] = p; p += f8; }
  B.W = p; p += f8;
  B.Z = p; p += f8;
  B.seg = p; p += (size_t)c->adj_every * f8;
  B.lamA = p; p += f8;
  B.lamB = p; p += f8;
  B.lamX0 = p; p += f8;
  B.lamP = p; p += bs;
  B.lamP3 = p; p += bs;
  B.Gb = p; p += 39 * bs;
  B.lWc = p; p += 8 * bs;
  B.T = p; p += nT * bs;
  B.small = p;   // [0..7] stage layer sums, [8..9] final, [10..11] initial fill, [16..17] mean cotangents
  return MRM_OK;
}

static StageArgs adj_stage_args(mrm_ctx* c, const mrm_run_desc* r, int stage) {
  StageArgs A;
  A.G = c->G;
  A.E = c->E;
  for (int d = 0; d < 3; d++) {
    A.RP.w1[d] = 8. / (12. * c->dx[d]);
    A.RP.w2[d] = -1. / (12. * c->dx[d]);
  }
  A.RP.inv_va2max = c->inv_va2max;
  A.SP = src_params(c, r, stage);
  A.ntx = A.nty = A.nkc = 0;
  A.write_p = 0;
  return A;
}

// forward of step n from Un, keeping the stage states, the TVD sweep inputs,
// the pre-boundary state and the layer sums every boundary fill read; with
// Unext, also the new state
static int adj_recompute(mrm_ctx* c, int n, const double* Un, AdjBufs& B, double* Unext) {
  const Geom& G = c->G;
  const size_t f8b = 8 * (size_t)G.bs * sizeof(double);
  mrm_run_desc r = {};
  r.dt = c->adj_rec[n].dt;
  const bool split = c->adj_rec[n].split;
  int rc;
  HIPCHK(c, hipMemcpyAsync(B.X[0], Un, f8b, hipMemcpyDeviceToDevice, c->stream));
  {
    unsigned nb = nblk(G.bs);
    if (nb > 65536u) nb = 65536u;
    k_eos_p<<<nb, MRM_BLOCK, 0, c->stream>>>(G, c->E, B.X[0], B.P[0], c->inv_va2max, nullptr);
    LAUNCHCHK(c, "k_eos_p");
  }
  for (int s = 1; s <= 4; s++) {
    double* Po = s < 4 ? B.P[s] : nullptr;
    if ((rc = stage_launch(c, &r, s, B.X[s - 1], B.P[s - 1], B.X[0], B.X[s], Po))) return rc;
    if ((rc = exchange_buf(c, B.X[s], Po))) return rc;
    if ((rc = bc_buf(c, B.X[s], Po, nullptr, 0))) return rc;
    if ((rc = bottom_sums(c, 0, B.X[s], B.small + 2 * (s - 1)))) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(B.W, B.X[4], f8b, hipMemcpyDeviceToDevice, c->stream));
  const int passes = split ? 2 : 1;
  const double dtv = split ? 0.5 * r.dt : r.dt;
  for (int pass = 0; pass < passes; pass++) {
    for (int d = 0; d < 3; d++) {
      HIPCHK(c, hipMemcpyAsync(B.V[3 * pass + d], B.W, f8b, hipMemcpyDeviceToDevice, c->stream));
      if ((rc = tvd_buf(c, B.W, B.P[3], dtv, nullptr, c->fma, d, true))) return rc;
    }
    if (split && pass == 0 && (rc = exchange_buf(c, B.W, nullptr))) return rc;
  }
  HIPCHK(c, hipMemcpyAsync(B.Z, B.W, f8b, hipMemcpyDeviceToDevice, c->stream));
  const double vmax = (c->v_lim > 0.0) ? c->v_lim : c->phys.tchk[3];
  const double eps_max = (c->e_lim > 0.0) ? c->e_lim : c->phys.tchk[2];
  k_tcheck<<<nblk(ninterior(c)), MRM_BLOCK, 0, c->stream>>>(G, c->phys.tchk[1], c->phys.tchk[0], eps_max, vmax, B.Z,
                                                          DiagArgs{});
  LAUNCHCHK(c, "k_tcheck");
  if ((rc = exchange_buf(c, B.Z, nullptr))) return rc;
  if ((rc = bottom_sums(c, 0, B.Z, B.small + 8))) return rc;
  if (Unext) {
    HIPCHK(c, hipMemcpyAsync(Unext, B.Z, f8b, hipMemcpyDeviceToDevice, c->stream));
    if ((rc = bc_buf(c, Unext, nullptr, nullptr, 0))) return rc;
  }
  return MRM_OK;
}

// transpose of step n: B.lamA holds the cotangent of U(n+1) on entry and of
// U(n) on return (adj_recompute(n) must have run)
static int adj_step_back(mrm_ctx* c, int n, AdjBufs& B) {
  const Geom& G = c->G;
  const size_t bs = G.bs, f8b = 8 * bs * sizeof(double);
  hipStream_t st = c->stream;
  mrm_run_desc r = {};
  r.dt = c->adj_rec[n].dt;
  const bool split = c->adj_rec[n].split;
  const BcParams BP = bc_params(c);
  double* lpm = B.small + 16;
  // final ghost fill, TCheck
  HIPCHK(c, adj::launch_bc(st, G, c->E, BP, B.small + 8, B.Z, B.lamA, c->d_part, lpm));
  HIPCHK(c, adj::launch_exchange(st, G, B.lamA));
  const double vmax = (c->v_lim > 0.0) ? c->v_lim : c->phys.tchk[3];
  const double eps_max = (c->e_lim > 0.0) ? c->e_lim : c->phys.tchk[2];
  HIPCHK(c, adj::launch_tcheck(st, G, c->phys.tchk[1], c->phys.tchk[0], eps_max, vmax, B.W, B.lamA));
  // TVD sweeps in reverse
  HIPCHK(c, hipMemsetAsync(B.lamP3, 0, bs * sizeof(double), st));
  const int passes = split ? 2 : 1;
  const double dtv = split ? 0.5 * r.dt : r.dt;
  for (int pass = passes - 1; pass >= 0; pass--) {
    if (split && pass == 0) HIPCHK(c, adj::launch_exchange(st, G, B.lamA));
    for (int d = 2; d >= 0; d--) {
      TvdParams P;
      int rc = tvd_sweep_params(c, dtv, d, P);
      if (rc) return rc;
      HIPCHK(c, adj::launch_tvd(st, d, G, P, B.V[3 * pass + d], B.P[3], B.T, B.lamA, B.lamB, B.lamP3));
      std::swap(B.lamA, B.lamB);
    }
  }
  // stages 4..1
  HIPCHK(c, hipMemsetAsync(B.lamX0, 0, f8b, st));
  const double* Q = c->qtot_set ? c->dF + MRM_QTOT * G.bs : nullptr;
  for (int s = 4; s >= 1; s--) {
    HIPCHK(c, adj::launch_bc(st, G, c->E, BP, B.small + 2 * (s - 1), B.X[s], B.lamA, c->d_part, lpm));
    HIPCHK(c, adj::launch_exchange(st, G, B.lamA));
    double* lprev = s > 1 ? B.lamB : B.lamX0;
    if (s > 1) HIPCHK(c, hipMemsetAsync(B.lamB, 0, f8b, st));
    double* lP = (s - 1 == 3) ? B.lamP3 : B.lamP;
    if (s - 1 != 3) HIPCHK(c, hipMemsetAsync(B.lamP, 0, bs * sizeof(double), st));
    const StageArgs A = adj_stage_args(c, &r, s);
    HIPCHK(c, adj::launch_stage(st, A, B.X[s - 1], B.P[s - 1], B.X[0], B.lamA, Q, B.Gb, B.lWc, B.lamX0, lprev, lP));
    HIPCHK(c, adj::launch_eos(st, G, c->E, B.X[s - 1], lP, lprev));
    if (s > 1) std::swap(B.lamA, B.lamB);
  }
  HIPCHK(c, hipMemcpyAsync(B.lamA, B.lamX0, f8b, hipMemcpyDeviceToDevice, st));
  return MRM_OK;
}

// bytes of one mixed state (7 FP32 fields, padded, and the FP64 energy), as
// allocated at context creation
static size_t mx_state_bytes(const mrm_ctx* c) {
  const size_t bs = (size_t)c->G.bs;
  return (7 * bs * sizeof(float) + 255) / 256 * 256 + bs * sizeof(double);
}

// In an FP32-mixed context the forward runs the mixed step and keeps its
// checkpoints in the mixed format (36 instead of 64 bytes per cell); the
// backward replays each segment with the mixed step (recorded dt and TVD
// branch) and applies the FP64 transposes at the widened mixed states: the
// adjoint of the FP64 step along the mixed trajectory.
extern "C" int mrm_adjoint_forward(mrm_ctx* c, mrm_run_desc* run, double dt_rad, int nsteps, int every) {
  if (!c || !run || nsteps < 1 || every < 1) return MRM_ERR_ARG;
  if (int rc = drop_settle(c)) return rc;
  const bool mx = c->precision == MRM_PREC_FP32_MIXED;
  if (!c->eos_ok) return fail(c, MRM_ERR_STATE, "mrm_adjoint_forward: no EOS table loaded");
  if (run->need_diagnostics) return fail(c, MRM_ERR_UNSUPPORTED, "diagnostics are outside the path");
  if (c->force_sequence || c->phys.params[2] > 0.0 || c->phys.divB[0] > 0.0 || c->phys.params[4] > 0.0 ||
      c->d_spt || (int)c->phys.bnd[1] > 0 || (int)c->phys.dmp[0] != 0)
    return fail(c, MRM_ERR_UNSUPPORTED,
                "mrm_adjoint_forward: the adjoint follows the fused step with va_adjust = 0, eta = 0, the "
                "vertical-field top boundary and no damping, Clean_divB or Spitzer conduction");
  for (int d = 0; d < 3; d++)
    if (c->lsize[d] < 4) return fail(c, MRM_ERR_UNSUPPORTED, "mrm_adjoint_forward: interior extents below 4");
  for (double* p : c->adj_ckpt) hipFree(p);
  c->adj_ckpt.clear();
  c->adj_rec.clear();
  c->adj_every = every;
  const size_t f8b = 8 * (size_t)c->G.bs * sizeof(double);
  const size_t cb = mx ? mx_state_bytes(c) : f8b;
  for (int n = 0; n < nsteps; n++) {
    if (n % every == 0) {
      double* p = nullptr;
      HIPCHK(c, hipMalloc(&p, cb));
      c->adj_ckpt.push_back(p);
      HIPCHK(c, hipMemcpyAsync(p, mx ? c->mxmem[c->cur] : (void*)c->S[c->cur], cb, hipMemcpyDeviceToDevice,
                               c->stream));
    }
    int rc = mx ? step_mixed(c, run, dt_rad) : step_fused(c, run, dt_rad);
    if (rc) return rc;
    c->adj_rec.push_back({c->last_dt, c->last_split});
    run->iteration++;
    run->globiter++;
    run->time += run->dt;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRM_OK;
}

// the mixed state in buffer src (mx_state_bytes) widened to FP64 U
static int mx_widen(mrm_ctx* c, const void* src, double* U) {
  MxState S;
  S.f = (float*)src;
  S.e = (double*)((char*)src + (mx_state_bytes(c) - (size_t)c->G.bs * sizeof(double)));
  HIPCHK(c, mixed::launch_convert(c->stream, false, c->G.bs, U, S));
  return MRM_OK;
}

// the mixed step n of the recorded forward from the mixed state src into dst,
// with the step's dt and TVD branch (the context's current state buffer is
// the working space)
static int mx_replay(mrm_ctx* c, int n, const void* src, void* dst) {
  const size_t mb = mx_state_bytes(c);
  HIPCHK(c, hipMemcpyAsync(c->mxmem[c->cur], src, mb, hipMemcpyDeviceToDevice, c->stream));
  unsigned nb = nblk(c->G.bs);
  if (nb > 65536u) nb = 65536u;
  // (the cmax slot takes a maximum that nothing reads)
  HIPCHK(c, mixed::launch_eos_p(c->stream, nb, c->G, c->E, c->M[c->cur], c->Pm[0], c->inv_va2max,
                                c->d_red + RED_CMAX, nullptr, 0.0, 0.0));
  mrm_run_desc r = {};
  r.dt = c->adj_rec[n].dt;
  int rc = step_mixed_body(c, &r, r.dt, c->adj_rec[n].split);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(dst, c->mxmem[c->cur], mb, hipMemcpyDeviceToDevice, c->stream));
  return MRM_OK;
}

static int adjoint_backward_mixed(mrm_ctx* c, const double* seed, double* grad, int fold_initial_ghosts);

extern "C" int mrm_adjoint_backward(mrm_ctx* c, const double* seed, double* grad, int fold_initial_ghosts) {
  if (!c || !seed || !grad) return MRM_ERR_ARG;
  if (c->adj_rec.empty()) return fail(c, MRM_ERR_STATE, "mrm_adjoint_backward: no recorded forward");
  if (int rc = gp_settle(c)) return rc;
  if (c->precision == MRM_PREC_FP32_MIXED) return adjoint_backward_mixed(c, seed, grad, fold_initial_ghosts);
  AdjBufs B;
  int rc = adj_buffers(c, B);
  if (rc) return rc;
  const size_t f8b = 8 * (size_t)c->G.bs * sizeof(double);
  HIPCHK(c, hipMemcpyAsync(B.lamA, seed, f8b, hipMemcpyHostToDevice, c->stream));
  const int N = (int)c->adj_rec.size(), every = c->adj_every;
  const int nseg = (N + every - 1) / every;
  for (int sg = nseg - 1; sg >= 0; sg--) {
    const int n0 = sg * every, n1 = std::min(N, n0 + every);
    double* segs = B.seg;
    HIPCHK(c, hipMemcpyAsync(segs, c->adj_ckpt[sg], f8b, hipMemcpyDeviceToDevice, c->stream));
    for (int n = n0; n < n1 - 1; n++)
      if ((rc = adj_recompute(c, n, segs + (size_t)(n - n0) * 8 * c->G.bs, B,
                              segs + (size_t)(n - n0 + 1) * 8 * c->G.bs))) return rc;
    for (int n = n1 - 1; n >= n0; n--) {
      if ((rc = adj_recompute(c, n, segs + (size_t)(n - n0) * 8 * c->G.bs, B, nullptr))) return rc;
      if ((rc = adj_step_back(c, n, B))) return rc;
    }
  }
  if (fold_initial_ghosts) {
    // the initial state's ghosts came from mrm_exchange_grid + mrm_set_boundary_conditions
    if ((rc = bottom_sums(c, 0, c->adj_ckpt[0], B.small + 10))) return rc;
    HIPCHK(c, adj::launch_bc(c->stream, c->G, c->E, bc_params(c), B.small + 10, c->adj_ckpt[0], B.lamA, c->d_part,
                             B.small + 16));
    HIPCHK(c, adj::launch_exchange(c->stream, c->G, B.lamA));
  }
  HIPCHK(c, hipMemcpyAsync(grad, B.lamA, f8b, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MRM_OK;
}

// The backward of an FP32-mixed context: the segment's mixed states come from
// replaying its checkpoint with the mixed step, each is widened to FP64 for
// the step's recompute and transpose; the context's state is left as the
// forward left it (its pressure is formed again by the next step).
static int adjoint_backward_mixed(mrm_ctx* c, const double* seed, double* grad, int fold_initial_ghosts) {
  AdjBufs B;
  const int every = c->adj_every;
  c->adj_every = 1;   // one FP64 segment slot: the widened state of the step being transposed
  int rc = adj_buffers(c, B);
  c->adj_every = every;
  if (rc) return rc;
  const size_t mb = mx_state_bytes(c), f8b = 8 * (size_t)c->G.bs * sizeof(double);
  // the mixed segment states and a copy of the context's current state (kept
  // by the context, like the FP64 work buffers, for the next backward)
  const size_t need = (size_t)(every + 1) * mb;
  if (c->adj_mx_bytes < need) {
    hipFree(c->adj_mx_mem);
    c->adj_mx_mem = nullptr;
    c->adj_mx_bytes = 0;
    HIPCHK(c, hipMalloc(&c->adj_mx_mem, need));
    c->adj_mx_bytes = need;
  }
  char* ms = c->adj_mx_mem;
  auto slot = [&](int q) { return (void*)(ms + (size_t)q * mb); };
  const int cur0 = c->cur;
  HIPCHK(c, hipMemcpyAsync(slot(every), c->mxmem[cur0], mb, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(B.lamA, seed, f8b, hipMemcpyHostToDevice, c->stream));
  const int N = (int)c->adj_rec.size();
  const int nseg = (N + every - 1) / every;
  for (int sg = nseg - 1; sg >= 0 && !rc; sg--) {
    const int n0 = sg * every, n1 = std::min(N, n0 + every);
    if (hipMemcpyAsync(slot(0), c->adj_ckpt[sg], mb, hipMemcpyDeviceToDevice, c->stream) != hipSuccess) {
      rc = fail(c, MRM_ERR_HIP, "hipMemcpyAsync checkpoint");
      break;
    }
    for (int n = n0; n < n1 - 1 && !rc; n++) rc = mx_replay(c, n, slot(n - n0), slot(n - n0 + 1));
    for (int n = n1 - 1; n >= n0 && !rc; n--) {
      if ((rc = mx_widen(c, slot(n - n0), B.seg))) break;
      if ((rc = adj_recompute(c, n, B.seg, B, nullptr))) break;
      rc = adj_step_back(c, n, B);
    }
  }
  if (!rc && fold_initial_ghosts) {
    // the initial state's ghosts came from mrm_exchange_grid + mrm_set_boundary_conditions
    if (!(rc = mx_widen(c, c->adj_ckpt[0], B.seg)) && !(rc = bottom_sums(c, 0, B.seg, B.small + 10))) {
      if (adj::launch_bc(c->stream, c->G, c->E, bc_params(c), B.small + 10, B.seg, B.lamA, c->d_part,
                         B.small + 16) != hipSuccess ||
          adj::launch_exchange(c->stream, c->G, B.lamA) != hipSuccess)
        rc = fail(c, MRM_ERR_HIP, "adjoint initial ghosts");
    }
  }
  if (!rc && hipMemcpyAsync(grad, B.lamA, f8b, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
    rc = fail(c, MRM_ERR_HIP, "hipMemcpyAsync gradient");
  // the forward's final state back in the context
  c->cur = cur0;
  (void)hipMemcpyAsync(c->mxmem[cur0], slot(every), mb, hipMemcpyDeviceToDevice, c->stream);
  c->fast_valid = false;
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = fail(c, MRM_ERR_HIP, "hipStreamSynchronize");
  return rc;
}

// ------------------------------------------------------------------ timers
extern "C" int mrm_timers(mrm_ctx* c, int enable) {
  if (!c) return MRM_ERR_ARG;
  c->harvest();
  if (enable) {
    for (int i = 0; i < MRM_NTIMER; i++) { c->tms[i] = 0.0; c->tcnt[i] = 0; }
  }
  c->timing = enable == 2 ? 2 : enable != 0 ? 1 : 0;
  return MRM_OK;
}

extern "C" int mrm_get_timers(mrm_ctx* c, double ms[MRM_NTIMER], long launches[MRM_NTIMER]) {
  if (!c) return MRM_ERR_ARG;
  c->harvest();
  for (int i = 0; i < MRM_NTIMER; i++) {
    if (ms) ms[i] = c->tms[i];
    if (launches) launches[i] = c->tcnt[i];
  }
  return MRM_OK;
}

extern "C" const char* mrm_timer_name(int id) {
  return (id >= 0 && id < MRM_NTIMER) ? timer_names[id] : "";
}
