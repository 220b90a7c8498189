// mrm_adjoint_api.h -- host-side launchers of the adjoint kernels (compiled in
// mrm_adjoint.hip, namespace adj).
#pragma once
#include <hip/hip_runtime.h>
#include "mrm_types.h"

namespace adj {
// stage transpose: lX0 += U0 cotangents, lXprev += input-state cotangents,
// lP += input-pressure cotangents; Gb (39 x bs) and lWc (8 x bs) are scratch
hipError_t launch_stage(hipStream_t st, const StageArgs& A, const double* Xin, const double* Pin, const double* X0,
                        const double* lamY, const double* Q, double* Gb, double* lWc, double* lX0, double* lXprev,
                        double* lP);
hipError_t launch_eos(hipStream_t st, const Geom& G, const EosDev& E, const double* X, const double* lP, double* lX);
hipError_t launch_tcheck(hipStream_t st, const Geom& G, double rho_min, double eps_min, double eps_max, double vmax,
                         const double* U, double* lam);
hipError_t launch_exchange(hipStream_t st, const Geom& G, double* lam);
// boundary transpose on state Z with layer sums psum[2]; part (>= 2 x nblocks)
// and lpm[2] are scratch
hipError_t launch_bc(hipStream_t st, const Geom& G, const EosDev& E, const BcParams& B, const double* psum,
                     const double* Z, double* lam, double* part, double* lpm);
// one TVD sweep transposed: lamIn = (sweep)^T lamOut, lP += pressure cotangents;
// T is scratch of tvd_scratch_fields() x bs doubles
int tvd_scratch_fields();
hipError_t launch_tvd(hipStream_t st, int d, const Geom& G, const TvdParams& P, const double* U, const double* Pr,
                      double* T, const double* lamOut, double* lamIn, double* lP);
}  // namespace adj
