// mrm_mixed_api.h -- host-side launchers of the FP32-mixed step kernels
// (compiled in mrm_mixed.hip, namespace mixed).
#pragma once
#include <hip/hip_runtime.h>
#include "mrm_types.h"

namespace mixed {
hipError_t launch_stage(hipStream_t st, unsigned nb, const MxStageArgs& A, MxState in, const float* Pin, MxState s0,
                        MxState out, float* Pout, const double* Q);
hipError_t launch_exchange(hipStream_t st, const Geom& G, MxState S, float* P);
// the periodic copy of the dim-2 ghost planes only (after the dim-1 TVD sweep)
hipError_t launch_exchange_dim2(hipStream_t st, const Geom& G, MxState S);
hipError_t launch_bc_sums(hipStream_t st, unsigned nb, const Geom& G, const EosDev& E, MxState S, int mode,
                          double* part);
hipError_t launch_bc(hipStream_t st, unsigned nb, const Geom& G, const EosDev& E, const BcParams& B,
                     const double* psum, MxState S, float* P, double inv_va2max, unsigned long long* cmax);
hipError_t launch_eos_p(hipStream_t st, unsigned nb, const Geom& G, const EosDev& E, MxState S, float* P,
                        double inv_va2max, unsigned long long* cmax, unsigned long long* valf, double vlim95,
                        double elim95);
hipError_t launch_cmax(hipStream_t st, unsigned nb, const Geom& G, MxState S, const float* P, double inv_va2max,
                       unsigned long long* cmax);
hipError_t launch_tvd_x_lines(hipStream_t st, unsigned nl, unsigned bsz, size_t lds, const Geom& G,
                              const MxTvdParams& P, MxState S, const float* Pr);
hipError_t launch_tvd_sweep(hipStream_t st, int d, bool tau, const MxEpilogue* epi, unsigned nb, const Geom& G,
                            const MxTvdParams& P, MxState S, const float* Pr);
hipError_t launch_convert(hipStream_t st, bool to_mixed, long bs, double* U, MxState S);
// Clean_divB on the mixed state: div B (with the next potential update when
// phi_update), the B/e correction (finishing the cells when epi is given),
// and the periodic ghosts of the three FP32 B components
hipError_t launch_divb(hipStream_t st, const DivbArgs& A, MxState S, double* divB, double* phi,
                       unsigned long long* err, bool phi_update);
hipError_t launch_divb_update(hipStream_t st, const DivbArgs& A, MxState S, const double* phi, const double* divB,
                              const MxEpilogue* epi);
hipError_t launch_exchange_b(hipStream_t st, const Geom& G, MxState S);
// Spitzer conduction of one stage on the mixed state (the fused step's
// k_eos_temp / k_spitzer_stage)
hipError_t launch_eos_temp(hipStream_t st, const Geom& G, const EosDev& E, MxState S, double* T);
hipError_t launch_spitzer_stage(hipStream_t st, const Geom& G, const SptParams& P, double wdt, MxState S,
                                const double* T, const double* sflx_in, const double* sflx0, double* sflx_out,
                                double* eres);
}  // namespace mixed
