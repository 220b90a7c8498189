// mrm_adjoint.hip -- the adjoint kernels of the FP64 step (no FMA
// contraction, like the exact forward build) and their launchers.
#define MRM_NS adj
#include "mrm_adjoint.h"
#include "mrm_adjoint_api.h"

namespace adj {

static inline unsigned nb_(long n) { return (unsigned)((n + MRM_BLOCK - 1) / MRM_BLOCK); }

hipError_t launch_stage(hipStream_t st, const StageArgs& A, const double* Xin, const double* Pin, const double* X0,
                        const double* lamY, const double* Q, double* Gb, double* lWc, double* lX0, double* lXprev,
                        double* lP) {
  const Geom& G = A.G;
  const long ni = (long)(G.le0 - G.lb0 + 1) * (G.le1 - G.lb1 + 1) * (G.le2 - G.lb2 + 1);
  k_adj_stage1<<<nb_(ni), MRM_BLOCK, 0, st>>>(A, Xin, Pin, X0, lamY, Q, Gb, lWc, lX0);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_adj_stage2<<<nb_(G.bs), MRM_BLOCK, 0, st>>>(G, A.RP, Xin, Pin, Gb, lWc, lXprev, lP);
  return hipGetLastError();
}

hipError_t launch_eos(hipStream_t st, const Geom& G, const EosDev& E, const double* X, const double* lP, double* lX) {
  k_adj_eos<<<nb_(G.bs), MRM_BLOCK, 0, st>>>(G, E, X, lP, lX);
  return hipGetLastError();
}

hipError_t launch_tcheck(hipStream_t st, const Geom& G, double rho_min, double eps_min, double eps_max, double vmax,
                         const double* U, double* lam) {
  const long ni = (long)(G.le0 - G.lb0 + 1) * (G.le1 - G.lb1 + 1) * (G.le2 - G.lb2 + 1);
  k_adj_tcheck<<<nb_(ni), MRM_BLOCK, 0, st>>>(G, rho_min, eps_min, eps_max, vmax, U, lam);
  return hipGetLastError();
}

hipError_t launch_exchange(hipStream_t st, const Geom& G, double* lam) {
  const int nx = G.le0 - G.lb0 + 1;
  // the forward fills dim 1, then dim 2 (whose copies include the dim-1 ghost rows)
  long n2 = (long)nx * G.n1 * G.g2;
  k_adj_exchange<2><<<nb_(n2), MRM_BLOCK, 0, st>>>(G, lam);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  long n1 = (long)nx * (G.le2 - G.lb2 + 1) * G.g1;
  k_adj_exchange<1><<<nb_(n1), MRM_BLOCK, 0, st>>>(G, lam);
  return hipGetLastError();
}

hipError_t launch_bc(hipStream_t st, const Geom& G, const EosDev& E, const BcParams& B, const double* psum,
                     const double* Z, double* lam, double* part, double* lpm) {
  const unsigned nb = nb_((long)G.n1 * G.n2);
  k_adj_bc<<<nb, MRM_BLOCK, 0, st>>>(G, E, B, psum, Z, lam, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_sum_partials<<<1, MRM_BLOCK, 0, st>>>(part, (int)nb, 2, lpm);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const long nyz = (long)(G.le1 - G.lb1 + 1) * (G.le2 - G.lb2 + 1);
  k_adj_bc_means<<<nb_(nyz), MRM_BLOCK, 0, st>>>(G, E, B.cc, lpm, Z, lam);
  return hipGetLastError();
}

int tvd_scratch_fields() { return TA_N; }

template <int D>
static hipError_t tvd_d(hipStream_t st, const Geom& G, const TvdParams& P, const double* U, const double* Pr,
                        double* T, const double* lamOut, double* lamIn, double* lP) {
  const unsigned nb = nb_(G.bs);
  for (int ph = 0; ph < 3; ph++) {
    k_adj_tvd_fwd<D><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, T, ph);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  for (int ph = 0; ph < 4; ph++) {
    k_adj_tvd_rev<D><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, T, lamOut, lamIn, lP, ph);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_tvd(hipStream_t st, int d, const Geom& G, const TvdParams& P, const double* U, const double* Pr,
                      double* T, const double* lamOut, double* lamIn, double* lP) {
  if (d == 0) return tvd_d<0>(st, G, P, U, Pr, T, lamOut, lamIn, lP);
  if (d == 1) return tvd_d<1>(st, G, P, U, Pr, T, lamOut, lamIn, lP);
  return tvd_d<2>(st, G, P, U, Pr, T, lamOut, lamIn, lP);
}

}  // namespace adj
