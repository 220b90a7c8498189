// Test infrastructure only: a driver for the reference's own hot-path
// translation units, compiled from the sources where they lie under the
// reference tree (see oracle/Makefile).  It is the "compiled reference" that
// pins the C restatement in oracle/muram_oracle.c.  Nothing in the product
// links or runs it.
//
// Reads a case file (geometry, PhysicsData/RunData subset, EOS table path,
// initial state in SoA layout, and a list of operations), runs the operations
// through the reference functions, and writes snapshots after the operations
// flagged for dumping.
//
// The operation list mirrors the stage loop of ComputeSolution
// (solver.C:109-365); Sync_timestep is file-static in solver.C and is
// restated here (solver.C:487-548).  RT is outside the path: Qtot = Tau = 0
// and dt_rad is supplied by the case.  SetBoundaryConditions is called with
// stage 2 where the loop passes stage 1: the stage argument only selects the
// BC.log line that needs an RTS object (boundary_pdmp_1_fftw3.C:155-156).
#include <mpi.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cmath>
#include <limits>
#include <vector>
#include <algorithm>

#include "physics.H"
#include "grid.H"
#include "run.H"
#include "eos.H"
#include "limit_va.H"
#include "mhd_tvd.H"
#include "src_int_tck.H"
#include "exchange.H"
#include "icbc.H"
#include "comm_split.H"
#include "divB.H"

#include "../include/muram_step.h"

extern double p_bc, s_bc;

enum {
  OP_STARTUP = 0, OP_SAVECONS = 1, OP_C2P = 2, OP_VALF = 3, OP_RESID = 4,
  OP_SYNC = 5, OP_SRCINT = 6, OP_EXCHANGE = 7, OP_BC = 8, OP_TVD = 9,
  OP_TCHECK = 10, OP_STEP = 11, OP_ADVANCE = 12, OP_DIVB = 13, OP_DIAG = 14,
  OP_STEPDIAG = 15
};
#define OP_DUMP 0x100

static double g_dt_rad = 1e30, g_dt_mhd = 0.0, g_cfl = 0.0;

static int invalid_dt(double a) {
  return (a != a) || (a <= 0.) || (a == std::numeric_limits<double>::infinity());
}

// Sync_timestep (solver.C:487-548) on one rank.
static void sync_timestep(RunData& Run, const PhysicsData& Physics, double dt_rad,
                          double dt_mhd, double* cfl) {
  double r0 = dt_rad, r1 = dt_mhd;
  if (invalid_dt(r0) || invalid_dt(r1)) {
    fprintf(stderr, "harness: invalid dt %g %g\n", r0, r1);
    MPI_Abort(MPI_COMM_WORLD, 1);
  }
  double dt_min_rt = Physics.rt[i_rt_tstep] * Run.CFL * r1;
  if (r0 < dt_min_rt) r0 = std::max(r0, dt_min_rt);
  double dt_min = std::min(r0, r1);
  double dt = std::min(r0, Run.CFL * r1);
  if (Run.outcad > 0) {
    double rem = Run.dt_rem();
    if (rem > 0.0) dt = std::min(dt, rem);
  }
  if (Run.autostep) Run.dt = dt;
  *cfl = Run.dt / dt_min;
  if (*cfl >= 2.) {
    fprintf(stderr, "harness: CFL %g\n", *cfl);
    MPI_Abort(MPI_COMM_WORLD, 1);
  }
}

static void bc(const RunData& Run, GridData& Grid, const PhysicsData& Physics, int stage) {
  SetBoundaryConditions(Run, Grid, Physics, stage == 1 ? 2 : stage, 1, nullptr);
}

// the zeroing of the diagnostics arrays at the start of the stage that
// collects them (solver.C:271-296)
static void diag_zero(GridData& Grid) {
  for (double* a : {Grid.tvar1, Grid.tvar2, Grid.tvar3, Grid.tvar4, Grid.tvar5, Grid.tvar6, Grid.tvar7,
                    Grid.tvar8, Grid.Qres, Grid.Qvis})
    memset(a, 0, sizeof(double) * Grid.bufsize);
}

// one iteration of the ComputeSolution loop body (solver.C:115-365); with
// diag, stage 4 and what follows it collect the diagnostics (solver.C:264-267)
static void full_step(RunData& Run, GridData& Grid, const PhysicsData& Physics, bool diag = false) {
  SaveCons(Grid, Physics);
  for (int stage = 1; stage <= maxstage; stage++) {
    ConsToPrim(Grid, Physics, Run);
    if (stage == 1) Adjust_Valf_Max(Run, Grid, Physics);
    Run.need_diagnostics = diag && stage == maxstage;
    if (Run.need_diagnostics) diag_zero(Grid);
    g_dt_mhd = MHD_Residual(Run, Grid, Physics);
    if (stage == 1) sync_timestep(Run, Physics, g_dt_rad, g_dt_mhd, &g_cfl);
    Source_Integrate_Tcheck(Run, Grid, Physics, stage);
    exchange_grid_acc(Grid, Physics, 1);
    bc(Run, Grid, Physics, stage);
  }
  if (g_cfl <= Run.CFL_tvd) {
    TVDlimit(Run, Grid, Physics, Run.dt);
  } else {
    TVDlimit(Run, Grid, Physics, 0.5 * Run.dt);
    exchange_grid_acc(Grid, Physics, 0);
    TVDlimit(Run, Grid, Physics, 0.5 * Run.dt);
  }
  if (Physics.divB[i_divB_switch] > 0) Clean_divB(Run, Grid, Physics);
  TCheck(Run, Grid, Physics);
  exchange_grid_acc(Grid, Physics, 0);
  bc(Run, Grid, Physics, maxstage + 1);
  Run.Advance();
  Run.need_diagnostics = 0;
}

static void rd(void* p, size_t sz, size_t n, FILE* f) {
  if (fread(p, sz, n, f) != n) { fprintf(stderr, "harness: short read\n"); MPI_Abort(MPI_COMM_WORLD, 2); }
}

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  if (argc < 3) { fprintf(stderr, "usage: %s case.bin out.bin\n", argv[0]); MPI_Abort(MPI_COMM_WORLD, 2); }
  FILE* f = fopen(argv[1], "rb");
  if (!f) { fprintf(stderr, "harness: cannot open %s\n", argv[1]); MPI_Abort(MPI_COMM_WORLD, 2); }
  int magic = 0;
  rd(&magic, 4, 1, f);
  if (magic != 0x4D524D31) { fprintf(stderr, "harness: bad magic\n"); MPI_Abort(MPI_COMM_WORLD, 2); }
  mrm_grid_desc gd; mrm_phys_desc pd; mrm_run_desc rdsc;
  rd(&gd, sizeof gd, 1, f); rd(&pd, sizeof pd, 1, f); rd(&rdsc, sizeof rdsc, 1, f);
  double bcv[3]; rd(bcv, 8, 3, f);
  char eos_path[256]; rd(eos_path, 1, 256, f);
  int nops = 0; rd(&nops, 4, 1, f);
  std::vector<int> ops(2 * nops); rd(ops.data(), 4, 2 * nops, f);

  RunData Run;
  Run.rank = 0; Run.zrank = 0;
  Run.dt = rdsc.dt; Run.time = rdsc.time; Run.CFL = rdsc.CFL; Run.CFL_tvd = rdsc.CFL_tvd;
  Run.outcad = rdsc.outcad; Run.autostep = rdsc.autostep;
  // diagnostics = 1 only allocates tvar1..8, Qres, Qvis (grid.C:319-335);
  // need_diagnostics is set by OP_DIAG and OP_STEPDIAG
  Run.need_diagnostics = 0; Run.diagnostics = 1; Run.verbose = 0;
  Run.globiter = rdsc.globiter; Run.iteration = rdsc.iteration;
  strncpy(Run.eos_name, eos_path, 127); Run.eos_name[127] = 0;

  PhysicsData Physics;
  memcpy(Physics.params, pd.params, sizeof pd.params);
  memcpy(Physics.bnd, pd.bnd, sizeof pd.bnd);
  memcpy(Physics.tvd_h, pd.tvd_h, sizeof pd.tvd_h);
  memcpy(Physics.tvd_cs, pd.tvd_cs, sizeof pd.tvd_cs);
  memcpy(Physics.tvd, pd.tvd, sizeof pd.tvd);
  memcpy(Physics.tvd_h_bnd, pd.tvd_h_bnd, sizeof pd.tvd_h_bnd);
  memcpy(Physics.tvd_visc_bnd, pd.tvd_visc_bnd, sizeof pd.tvd_visc_bnd);
  memcpy(Physics.tvd_eta_bnd, pd.tvd_eta_bnd, sizeof pd.tvd_eta_bnd);
  memcpy(Physics.tvd_coeff, pd.tvd_coeff, sizeof pd.tvd_coeff);
  memcpy(Physics.tchk, pd.tchk, sizeof pd.tchk);
  memcpy(Physics.dmp, pd.dmp, sizeof pd.dmp);
  memcpy(Physics.rt, pd.rt, sizeof pd.rt);
  Physics.rt_ext[i_ext_cor] = pd.rt_ext_cor;
  memcpy(Physics.divB, pd.divB, sizeof pd.divB);
  Physics.Init();

  GridData Grid;
  for (int d = 0; d < 3; d++) {
    Grid.gsize[d] = gd.gsize[d]; Grid.ghosts[d] = gd.ghosts[d];
    Grid.periods[d] = gd.periods[d]; Grid.procs[d] = 1;
    Grid.gxmin[d] = gd.gxmin[d]; Grid.gxmax[d] = gd.gxmax[d];
  }
  Grid.Init(Run, Physics);
  comm_split_init(Run, Grid);
  const int bs = Grid.bufsize;
  // arrays the hot path reads and the harness must define (RT off: zero)
  memset(Grid.Res, 0, sizeof(cState) * bs);
  memset(Grid.U0, 0, sizeof(cState) * bs);
  memset(Grid.Qtot, 0, sizeof(double) * bs);
  memset(Grid.Tau, 0, sizeof(double) * bs);
  for (double* a : {Grid.pres, Grid.temp, Grid.ne, Grid.rhoi, Grid.amb, Grid.divB, Grid.phi})
    memset(a, 0, sizeof(double) * bs);
  // Spitzer conduction's arrays exist when param_spitzer > 0 (grid.C:311-318)
  for (double* a : {Grid.sflx, Grid.sflx0, Grid.Rflx, Grid.BgradT})
    if (a) memset(a, 0, sizeof(double) * bs);
  std::vector<double> zeros(bs, 0.0);
  diag_zero(Grid);

  std::vector<double> soa(8L * bs);
  rd(soa.data(), 8, 8L * bs, f);
  fclose(f);
  double* Ud = (double*)Grid.U;
  for (long n = 0; n < bs; n++)
    for (int v = 0; v < 8; v++) Ud[8 * n + v] = soa[(long)v * bs + n];

  p_bc = bcv[0]; s_bc = bcv[1]; g_dt_rad = bcv[2];
  eos_init(Grid, Run, Physics);

  FILE* out = fopen(argv[2], "wb");
  int ndump = 0;
  for (int o = 0; o < nops; o++) if (ops[2 * o] & OP_DUMP) ndump++;
  fwrite(&ndump, 4, 1, out);
  fwrite(&bs, 4, 1, out);

  for (int o = 0; o < nops; o++) {
    int op = ops[2 * o] & 0xff, arg = ops[2 * o + 1];
    switch (op) {
      case OP_STARTUP: exchange_grid_acc(Grid, Physics, 1); bc(Run, Grid, Physics, 0); break;
      case OP_SAVECONS: SaveCons(Grid, Physics); break;
      case OP_C2P: ConsToPrim(Grid, Physics, Run); break;
      case OP_VALF: Adjust_Valf_Max(Run, Grid, Physics); break;
      case OP_RESID: g_dt_mhd = MHD_Residual(Run, Grid, Physics); break;
      case OP_SYNC: sync_timestep(Run, Physics, g_dt_rad, g_dt_mhd, &g_cfl); break;
      case OP_SRCINT: Source_Integrate_Tcheck(Run, Grid, Physics, arg); break;
      case OP_EXCHANGE: exchange_grid_acc(Grid, Physics, arg); break;
      case OP_BC: bc(Run, Grid, Physics, arg); break;
      case OP_TVD: TVDlimit(Run, Grid, Physics, arg ? 0.5 * Run.dt : Run.dt); break;
      case OP_TCHECK: TCheck(Run, Grid, Physics); break;
      case OP_STEP: for (int s = 0; s < arg; s++) full_step(Run, Grid, Physics); break;
      case OP_ADVANCE: Run.Advance(); break;
      case OP_DIVB: Clean_divB(Run, Grid, Physics); break;
      case OP_DIAG:
        Run.need_diagnostics = arg;
        if (arg) diag_zero(Grid);
        break;
      case OP_STEPDIAG:
        for (int s = 0; s < arg; s++) full_step(Run, Grid, Physics, s == arg - 1);
        break;
      default: fprintf(stderr, "harness: bad op %d\n", op); MPI_Abort(MPI_COMM_WORLD, 2);
    }
    if (ops[2 * o] & OP_DUMP) {
      double sc[16] = {0};
      sc[0] = Run.dt; sc[1] = Run.time; sc[2] = g_dt_mhd; sc[3] = g_cfl;
      sc[4] = p_bc; sc[5] = s_bc; sc[6] = inv_va2max; sc[7] = v_lim; sc[8] = e_lim;
      sc[9] = Run.globiter;
      fwrite(&o, 4, 1, out);
      fwrite(sc, 8, 16, out);
      double* Rd = (double*)Grid.Res;
      for (int v = 0; v < 8; v++) for (long n = 0; n < bs; n++) soa[(long)v * bs + n] = Ud[8 * n + v];
      fwrite(soa.data(), 8, 8L * bs, out);
      for (int v = 0; v < 8; v++) for (long n = 0; n < bs; n++) soa[(long)v * bs + n] = Rd[8 * n + v];
      fwrite(soa.data(), 8, 8L * bs, out);
      for (double* a : {Grid.pres, Grid.temp, Grid.ne, Grid.rhoi, Grid.amb, Grid.divB, Grid.phi})
        fwrite(a, 8, bs, out);
      fwrite(Grid.sflx ? Grid.sflx : zeros.data(), 8, bs, out);
      for (double* a : {Grid.tvar1, Grid.tvar2, Grid.tvar3, Grid.tvar4, Grid.tvar5, Grid.tvar6, Grid.tvar7,
                        Grid.tvar8, Grid.Qres, Grid.Qvis})
        fwrite(a, 8, bs, out);
    }
  }
  fclose(out);
  MPI_Finalize();
  return 0;
}
