// mrm_launch.h -- host-side launchers of the fused-step kernels, compiled in
// each translation unit's kernel namespace (exact: no FMA contraction;
// fast: contraction allowed) so the host code can pick either build.
#pragma once
#include "mrm_fused.h"

namespace MRM_NS {

template <bool ZWIN>
inline void launch_stage_z(hipStream_t st, unsigned nb, const StageArgs& A, const double* Uin, const double* Pin,
                           const double* U0, double* Uout, double* Pout, const double* Q) {
  const bool eta = A.RP.eta > 0.0, eres = A.eres != nullptr;
  if (eta && eres) k_stage<ZWIN, true, true><<<nb, ST_SX * ST_SY, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
  else if (eta) k_stage<ZWIN, true><<<nb, ST_SX * ST_SY, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
  else if (eres) k_stage<ZWIN, false, true><<<nb, ST_SX * ST_SY, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
  else k_stage<ZWIN><<<nb, ST_SX * ST_SY, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
}

// variant 4: k_stage_w with the EOS pressure in a third wave group (three
// waves per SIMD; the two-group kernel with resistivity); 3: k_stage_f's
// work over two wave groups (k_stage_w); 2: fluxes shared through LDS
// (k_stage_f); 1: register window of primitives (k_stage<true>); 0: dim-2
// neighbours re-read (k_stage<false>)
inline hipError_t launch_stage(hipStream_t st, unsigned nb, int variant, const StageArgs& A, const double* Uin,
                               const double* Pin, const double* U0, double* Uout, double* Pout,
                               const double* Q) {
  // with resistivity the three-group kernel would not fit three waves per
  // SIMD without spilling: the two-group kernel runs instead
  if (variant == 4 && A.RP.eta > 0.0) variant = 3;
  if (variant == 4) {
    const unsigned nt = 3 * ST_SX * ST_SY;
    if (A.eres != nullptr) k_stage_w<false, true, true><<<nb, nt, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
    else k_stage_w<false, false, true><<<nb, nt, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
  } else if (variant == 3) {
    const bool eta = A.RP.eta > 0.0, eres = A.eres != nullptr;
    const unsigned nt = 2 * ST_SX * ST_SY;
    if (eta && eres) k_stage_w<true, true><<<nb, nt, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
    else if (eta) k_stage_w<true><<<nb, nt, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
    else if (eres) k_stage_w<false, true><<<nb, nt, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
    else k_stage_w<><<<nb, nt, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
  } else if (variant == 2) {
    const bool eta = A.RP.eta > 0.0, eres = A.eres != nullptr;
    if (eta && eres) k_stage_f<true, true><<<nb, ST_SX * ST_SY, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
    else if (eta) k_stage_f<true><<<nb, ST_SX * ST_SY, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
    else if (eres) k_stage_f<false, true><<<nb, ST_SX * ST_SY, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
    else k_stage_f<><<<nb, ST_SX * ST_SY, 0, st>>>(A, Uin, Pin, U0, Uout, Pout, Q);
  } else if (variant == 1) {
    launch_stage_z<true>(st, nb, A, Uin, Pin, U0, Uout, Pout, Q);
  } else {
    launch_stage_z<false>(st, nb, A, Uin, Pin, U0, Uout, Pout, Q);
  }
  return hipGetLastError();
}

inline hipError_t launch_tvd_x_lines(hipStream_t st, unsigned nl, unsigned bsz, size_t lds, const TvdLineArgs& A,
                                     double* U, const double* Pr) {
  if (A.P.bcount) {
    allow_lds(k_tvd_x_lines<TX_MAXN, true>, lds);
    k_tvd_x_lines<TX_MAXN, true><<<nl, bsz, lds, st>>>(A, U, Pr);
  } else {
    allow_lds(k_tvd_x_lines<TX_MAXN>, lds);
    k_tvd_x_lines<TX_MAXN><<<nl, bsz, lds, st>>>(A, U, Pr);
  }
  return hipGetLastError();
}

// the counting instantiations of the marching sweeps (limiter branch accounting)
inline hipError_t launch_tvd_sweep_cnt(hipStream_t st, int d, bool tau, const TvdEpilogue* epi, unsigned nb,
                                       const Geom& G, const TvdParams& P, double* U, const double* Pr) {
  TvdEpilogue X{};
  const DiagArgs N{};
  if (d == 0) {
    if (tau) k_tvd_sweep<0, true, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, N);
    else k_tvd_sweep<0, false, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, N);
  } else if (d == 1) {
    if (tau) k_tvd_sweep<1, true, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, N);
    else k_tvd_sweep<1, false, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, N);
  } else if (epi && epi->valf) {
    if (tau) k_tvd_sweep<2, true, true, true, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, N);
    else k_tvd_sweep<2, false, true, true, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, N);
  } else if (epi) {
    if (tau) k_tvd_sweep<2, true, true, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, N);
    else k_tvd_sweep<2, false, true, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, N);
  } else {
    if (tau) k_tvd_sweep<2, true, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, N);
    else k_tvd_sweep<2, false, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, N);
  }
  return hipGetLastError();
}

// marching sweep along dim d (0, 1, 2); epi applies the finishing epilogue (d = 2)
inline hipError_t launch_tvd_sweep(hipStream_t st, int d, bool tau, const TvdEpilogue* epi, unsigned nb,
                                   const Geom& G, const TvdParams& P, double* U, const double* Pr) {
  if (P.bcount) return launch_tvd_sweep_cnt(st, d, tau, epi, nb, G, P, U, Pr);
  TvdEpilogue X{};
  if (d == 2 && epi && !epi->Pnew) {
    // TCheck alone as the epilogue
    if (tau) k_tvd_sweep<2, true, false, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, DiagArgs{});
    else k_tvd_sweep<2, false, false, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, DiagArgs{});
    return hipGetLastError();
  }
  if (d == 0) {
    if (tau) k_tvd_sweep<0, true, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DiagArgs{});
    else k_tvd_sweep<0, false, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DiagArgs{});
  } else if (d == 1) {
    if (tau) k_tvd_sweep<1, true, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DiagArgs{});
    else k_tvd_sweep<1, false, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DiagArgs{});
  } else if (epi && epi->valf) {
    if (tau) k_tvd_sweep<2, true, true, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, DiagArgs{});
    else k_tvd_sweep<2, false, true, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, DiagArgs{});
  } else if (epi) {
    if (tau) k_tvd_sweep<2, true, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, DiagArgs{});
    else k_tvd_sweep<2, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, *epi, DiagArgs{});
  } else {
    if (tau) k_tvd_sweep<2, true, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DiagArgs{});
    else k_tvd_sweep<2, false, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DiagArgs{});
  }
  return hipGetLastError();
}

// the same sweeps with the diagnostics of every updated point (no epilogue)
inline hipError_t launch_tvd_sweep_diag(hipStream_t st, int d, bool tau, unsigned nb, const Geom& G,
                                        const TvdParams& P, double* U, const double* Pr, const DiagArgs& DA) {
  TvdEpilogue X{};
  if (d == 0) {
    if (tau) k_tvd_sweep<0, true, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DA);
    else k_tvd_sweep<0, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DA);
  } else if (d == 1) {
    if (tau) k_tvd_sweep<1, true, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DA);
    else k_tvd_sweep<1, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DA);
  } else {
    if (tau) k_tvd_sweep<2, true, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DA);
    else k_tvd_sweep<2, false, false, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, U, Pr, X, DA);
  }
  return hipGetLastError();
}

}  // namespace MRM_NS
