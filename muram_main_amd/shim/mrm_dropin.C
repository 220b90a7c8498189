// mrm_dropin.C -- MURaM's MHD-step functions forwarded to the MI355X library.
#include <mpi.h>
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iterator>
#include <map>
#include <vector>
#include "ACCH.h"
#include "divB.H"
#include "grid.H"
#include "run.H"
#include "physics.H"
#include "eos.H"
#include "mhd_tvd.H"
#include "src_int_tck.H"
#include "limit_va.H"
#include "exchange.H"
#include "icbc.H"
extern "C" {
#include "muram_step.h"   // <repo>/include
}

double inv_va2max, v_lim, e_lim;          // limit_va.H:5 (kept in sync below)
static std::map<const GridData*, mrm_ctx*> g_ctx;

static void chk(mrm_ctx* c, int rc, const char* what) {
  if (rc == MRM_OK) return;
  std::fprintf(stderr, "%s: %d %s\n", what, rc, mrm_last_error(c));
  MPI_Abort(MPI_COMM_WORLD, 1);
}

static mrm_ctx* ctx_of(const GridData& G, const PhysicsData& P) {
  auto it = g_ctx.find(&G);
  if (it != g_ctx.end()) return it->second;
  mrm_grid_desc gd;
  mrm_phys_desc pd = {};
  for (int d = 0; d < 3; d++) {
    gd.gsize[d] = G.gsize[d]; gd.ghosts[d] = G.ghosts[d]; gd.periods[d] = G.periods[d];
    gd.gxmin[d] = G.gxmin[d]; gd.gxmax[d] = G.gxmax[d];
  }
  // PhysicsData's parameter arrays map one to one (physics.H:84-96)
  std::copy(P.params, P.params + 9, pd.params);   std::copy(P.bnd, P.bnd + 5, pd.bnd);
  std::copy(P.tvd_h, P.tvd_h + 4, pd.tvd_h);      std::copy(P.tvd_cs, P.tvd_cs + 4, pd.tvd_cs);
  std::copy(P.tvd, P.tvd + 10, pd.tvd);           std::copy(P.tvd_h_bnd, P.tvd_h_bnd + 2, pd.tvd_h_bnd);
  std::copy(P.tvd_visc_bnd, P.tvd_visc_bnd + 2, pd.tvd_visc_bnd);
  std::copy(P.tvd_eta_bnd, P.tvd_eta_bnd + 2, pd.tvd_eta_bnd);
  std::copy(P.tvd_coeff, P.tvd_coeff + 4, pd.tvd_coeff);
  std::copy(P.tchk, P.tchk + 4, pd.tchk);         std::copy(P.dmp, P.dmp + 4, pd.dmp);
  std::copy(P.rt, P.rt + 10, pd.rt);              pd.rt_ext_cor = P.rt_ext[0];
  std::copy(P.divB, P.divB + 3, pd.divB);
  mrm_ctx* c = nullptr;
  chk(nullptr, mrm_create(&gd, &pd, MRM_PREC_FP64, &c), "mrm_create");
  // potential-field top boundary: the PSF kernels from the run directory,
  // where potential_sd_fftw3.C:77-86 reads them
  if (P.bnd[1] > 0) chk(c, mrm_potential_kernels_load(c, "."), "mrm_potential_kernels_load");
  g_ctx[&G] = c;
  return c;
}

static mrm_run_desc run_of(const RunData& R) {
  mrm_run_desc r;
  r.dt = R.dt; r.time = R.time; r.CFL = R.CFL; r.CFL_tvd = R.CFL_tvd; r.outcad = R.outcad;
  r.autostep = R.autostep; r.need_diagnostics = R.need_diagnostics;
  r.globiter = R.globiter; r.iteration = R.iteration;
  return r;
}

static void pull_limits(mrm_ctx* c) {        // context -> reference globals
  double lim[4];
  mrm_get_limits(c, lim);
  inv_va2max = lim[0]; v_lim = lim[1]; e_lim = lim[2];
}

static std::vector<char> read_file(const char* path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) { std::fprintf(stderr, "cannot open %s\n", path); MPI_Abort(MPI_COMM_WORLD, 1); }
  return std::vector<char>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

// eos.H:38 -- the reference loader keeps running for its host-side interps
void eos_init_reference(GridData&, RunData&, const PhysicsData&);   // eos_init.C, renamed with -D
void eos_init(GridData& G, RunData& R, const PhysicsData& P) {
  eos_init_reference(G, R, P);
  mrm_ctx* c = ctx_of(G, P);
  std::vector<char> bytes = read_file(R.eos_name);   // the same table file
  chk(c, mrm_eos_load(c, bytes.data(), bytes.size()), "mrm_eos_load");
}

void ConsToPrim(GridData& G, const PhysicsData& P, const RunData&) {               // eos.H:39
  mrm_ctx* c = ctx_of(G, P);
  chk(c, mrm_cons_to_prim(c), "ConsToPrim");
}

void Adjust_Valf_Max(const RunData& R, const GridData& G, const PhysicsData& P) {  // limit_va.H:7
  mrm_ctx* c = ctx_of(G, P);
  mrm_run_desc r = run_of(R);
  chk(c, mrm_adjust_valf_max(c, &r, nullptr), "Adjust_Valf_Max");
  pull_limits(c);
}

double MHD_Residual(const RunData& R, GridData& G, const PhysicsData& P) {         // mhd_tvd.H:5
  mrm_ctx* c = ctx_of(G, P);
  mrm_run_desc r = run_of(R);
  double dt_cfl = 0.0;
  chk(c, mrm_mhd_residual(c, &r, &dt_cfl), "MHD_Residual");
  return dt_cfl;
}

void Source_Integrate_Tcheck(const RunData& R, GridData& G, const PhysicsData& P, const int stage) {
  mrm_ctx* c = ctx_of(G, P);                                                         // src_int_tck.H:8
  mrm_run_desc r = run_of(R);
  chk(c, mrm_source_integrate_tcheck(c, &r, stage), "Source_Integrate_Tcheck");
}

void SaveCons(GridData& G, const PhysicsData& P) {                                  // src_int_tck.H:10
  mrm_ctx* c = ctx_of(G, P);
  chk(c, mrm_save_cons(c), "SaveCons");
}

void TCheck(const RunData& R, GridData& G, const PhysicsData& P) {                  // src_int_tck.H:7
  mrm_ctx* c = ctx_of(G, P);
  mrm_run_desc r = run_of(R);
  chk(c, mrm_tcheck(c, &r), "TCheck");
}

void TVDlimit(const RunData& R, GridData& G, const PhysicsData& P, const double dt_tvd) {  // mhd_tvd.H:6
  mrm_ctx* c = ctx_of(G, P);
  mrm_run_desc r = run_of(R);
  chk(c, mrm_tvdlimit(c, &r, dt_tvd), "TVDlimit");
}

void Clean_divB(const RunData& R, GridData& G, const PhysicsData& P) {               // divB.H
  mrm_ctx* c = ctx_of(G, P);     // phi lives on the device; Grid.phi/divB mirror via UpdateCPU
  mrm_run_desc r = run_of(R);
  chk(c, mrm_clean_divB(c, &r, nullptr), "Clean_divB");
}

void exchange_grid_acc(GridData& G, const PhysicsData& P, const int) {             // exchange.H:8
  mrm_ctx* c = ctx_of(G, P);     // one rank: periodic dims 1, 2 filled on the device
  chk(c, mrm_exchange_grid(c), "exchange_grid_acc");
}

void SetBoundaryConditions(const RunData& R, GridData& G, const PhysicsData& P,    // icbc.H:10-12
                           const int stage, const int pt_update, RTS*) {
  mrm_ctx* c = ctx_of(G, P);
  mrm_run_desc r = run_of(R);
  chk(c, mrm_set_boundary_conditions_pt(c, &r, stage, pt_update), "SetBoundaryConditions");
}

// ACCH.h:49-50 -- the solver's mirror hooks (ACCH.C's, renamed with -D):
// find the array among the contexts' grids and copy it to or from the device
static bool mirror(void* p, bool to_device) {
  for (auto& kv : g_ctx) {
    const GridData& G = *kv.first;
    mrm_ctx* c = kv.second;
    const struct { void* a; int which; } st[] = {{G.U, MRM_U}, {G.U0, MRM_U0}, {G.Res, MRM_RES}};
    for (const auto& x : st)
      if (x.a && p == x.a) {
        chk(c, to_device ? mrm_set_state(c, x.which, (const double*)p, MRM_LAYOUT_AOS)
                         : mrm_get_state(c, x.which, (double*)p, MRM_LAYOUT_AOS), "ACCH::Update");
        return true;
      }
    const struct { double* a; int field; } fl[] = {
        {G.pres, MRM_PRES}, {G.temp, MRM_TEMP}, {G.ne, MRM_NE}, {G.rhoi, MRM_RHOI}, {G.amb, MRM_AMB},
        {G.Qtot, MRM_QTOT}, {G.Tau, MRM_TAU}, {G.divB, MRM_DIVB}, {G.phi, MRM_PHI}, {G.sflx, MRM_SFLX},
        {G.tvar1, MRM_TVAR1}, {G.tvar2, MRM_TVAR1 + 1}, {G.tvar3, MRM_TVAR1 + 2}, {G.tvar4, MRM_TVAR1 + 3},
        {G.tvar5, MRM_TVAR1 + 4}, {G.tvar6, MRM_TVAR1 + 5}, {G.tvar7, MRM_TVAR1 + 6}, {G.tvar8, MRM_TVAR8},
        {G.Qres, MRM_QRES}, {G.Qvis, MRM_QVIS}};
    for (const auto& x : fl)
      if (x.a && p == x.a) {
        chk(c, to_device ? mrm_set_field(c, x.field, x.a) : mrm_get_field(c, x.field, x.a), "ACCH::Update");
        return true;
      }
  }
  return false;   // an array the library does not hold: nothing to mirror
}
void ACCH::UpdateGPU(void* p, std::size_t) { mirror(p, true); }
void ACCH::UpdateCPU(void* p, std::size_t) { mirror(p, false); }
