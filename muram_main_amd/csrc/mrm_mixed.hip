// mrm_mixed.hip -- the FP32-mixed step kernels (built with FMA contraction)
// and their launchers.
#define MRM_NS mixed
#include "mrm_mixed.h"
#include "mrm_mixed_api.h"

namespace mixed {

static inline unsigned nblk_(long n) { return (unsigned)((n + MRM_BLOCK - 1) / MRM_BLOCK); }

// stage kernel of the mixed step: 1 fluxes shared through LDS (k_stage_mxf),
// 0 every sweep's fluxes formed in place (k_stage_mx); MRM_MX_STAGE
static int stage_variant() {
  static int v = -1;
  if (v < 0) {
    const char* s = getenv("MRM_MX_STAGE");
    v = (s && s[0] == '0') ? 0 : 1;
  }
  return v;
}

hipError_t launch_stage(hipStream_t st, unsigned nb, const MxStageArgs& A, MxState in, const float* Pin, MxState s0,
                        MxState out, float* Pout, const double* Q) {
  const bool eta = A.eta > 0.0f, eres = A.eres != nullptr, wp = Pout != nullptr;
  const unsigned nt = MX_SX * MX_SY;
  if (stage_variant() == 1) {
    if (wp) {
      if (eta && eres) k_stage_mxf<true, true, true><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
      else if (eta) k_stage_mxf<true, false, true><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
      else if (eres) k_stage_mxf<false, true, true><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
      else k_stage_mxf<false, false, true><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
    } else {
      if (eta && eres) k_stage_mxf<true, true, false><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
      else if (eta) k_stage_mxf<true, false, false><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
      else if (eres) k_stage_mxf<false, true, false><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
      else k_stage_mxf<false, false, false><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
    }
    return hipGetLastError();
  }
  if (wp) {
    if (eta && eres) k_stage_mx<true, true, true><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
    else if (eta) k_stage_mx<true, false, true><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
    else if (eres) k_stage_mx<false, true, true><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
    else k_stage_mx<false, false, true><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
  } else {
    if (eta && eres) k_stage_mx<true, true, false><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
    else if (eta) k_stage_mx<true, false, false><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
    else if (eres) k_stage_mx<false, true, false><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
    else k_stage_mx<false, false, false><<<nb, nt, 0, st>>>(A, in, Pin, s0, out, Pout, Q);
  }
  return hipGetLastError();
}

hipError_t launch_exchange(hipStream_t st, const Geom& G, MxState S, float* P) {
  const int nx = G.le0 - G.lb0 + 1;
  long n1 = (long)nx * (G.le2 - G.lb2 + 1) * G.g1;
  k_exchange_mx<1><<<nblk_(n1), MRM_BLOCK, 0, st>>>(G, S, P);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  long n2 = (long)nx * G.n1 * G.g2;
  k_exchange_mx<2><<<nblk_(n2), MRM_BLOCK, 0, st>>>(G, S, P);
  return hipGetLastError();
}

hipError_t launch_exchange_dim2(hipStream_t st, const Geom& G, MxState S) {
  const int nx = G.le0 - G.lb0 + 1;
  long n2 = (long)nx * G.n1 * G.g2;
  k_exchange_mx<2><<<nblk_(n2), MRM_BLOCK, 0, st>>>(G, S, nullptr);
  return hipGetLastError();
}

hipError_t launch_bc_sums(hipStream_t st, unsigned nb, const Geom& G, const EosDev& E, MxState S, int mode,
                          double* part) {
  k_bc_sums_mx<<<nb, MRM_BLOCK, 0, st>>>(G, E, S, mode, part);
  return hipGetLastError();
}

hipError_t launch_bc(hipStream_t st, unsigned nb, const Geom& G, const EosDev& E, const BcParams& B,
                     const double* psum, MxState S, float* P, double inv_va2max, unsigned long long* cmax) {
  k_bc_mx<<<nb, MRM_BLOCK, 0, st>>>(G, E, B, psum, S, P, inv_va2max, cmax);
  return hipGetLastError();
}

hipError_t launch_eos_p(hipStream_t st, unsigned nb, const Geom& G, const EosDev& E, MxState S, float* P,
                        double inv_va2max, unsigned long long* cmax, unsigned long long* valf, double vlim95,
                        double elim95) {
  k_eos_p_mx<<<nb, MRM_BLOCK, 0, st>>>(G, E, S, P, inv_va2max, cmax, valf, vlim95, elim95);
  return hipGetLastError();
}

hipError_t launch_cmax(hipStream_t st, unsigned nb, const Geom& G, MxState S, const float* P, double inv_va2max,
                       unsigned long long* cmax) {
  k_cmax_visited_mx<<<nb, MRM_BLOCK, 0, st>>>(G, S, P, inv_va2max, cmax);
  return hipGetLastError();
}

hipError_t launch_tvd_x_lines(hipStream_t st, unsigned nl, unsigned bsz, size_t lds, const Geom& G,
                              const MxTvdParams& P, MxState S, const float* Pr) {
  allow_lds(k_tvd_x_lines_mx, lds);
  k_tvd_x_lines_mx<<<nl, bsz, lds, st>>>(G, P, S, Pr);
  return hipGetLastError();
}

hipError_t launch_tvd_sweep(hipStream_t st, int d, bool tau, const MxEpilogue* epi, unsigned nb, const Geom& G,
                            const MxTvdParams& P, MxState S, const float* Pr) {
  MxEpilogue X{};
  if (d == 0) {
    if (tau) k_tvd_sweep_mx<0, true, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, X);
    else k_tvd_sweep_mx<0, false, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, X);
  } else if (d == 1) {
    if (tau) k_tvd_sweep_mx<1, true, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, X);
    else k_tvd_sweep_mx<1, false, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, X);
  } else if (epi && epi->valf) {
    if (tau) k_tvd_sweep_mx<2, true, true, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, *epi);
    else k_tvd_sweep_mx<2, false, true, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, *epi);
  } else if (epi) {
    if (tau) k_tvd_sweep_mx<2, true, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, *epi);
    else k_tvd_sweep_mx<2, false, true><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, *epi);
  } else {
    if (tau) k_tvd_sweep_mx<2, true, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, X);
    else k_tvd_sweep_mx<2, false, false><<<nb, MRM_BLOCK, 0, st>>>(G, P, S, Pr, X);
  }
  return hipGetLastError();
}

hipError_t launch_convert(hipStream_t st, bool to_mixed, long bs, double* U, MxState S) {
  unsigned nb = nblk_(bs);
  if (nb > 65536u) nb = 65536u;
  if (to_mixed) k_to_mixed<<<nb, MRM_BLOCK, 0, st>>>(bs, U, S);
  else k_from_mixed<<<nb, MRM_BLOCK, 0, st>>>(bs, S, U);
  return hipGetLastError();
}

hipError_t launch_divb(hipStream_t st, const DivbArgs& A, MxState S, double* divB, double* phi,
                       unsigned long long* err, bool phi_update) {
  const unsigned nb = divb_blocks(A.G);
  if (phi_update) k_divb_mx<true><<<nb, DIVB_TX * DIVB_TY, 0, st>>>(A, S, divB, phi, err);
  else k_divb_mx<false><<<nb, DIVB_TX * DIVB_TY, 0, st>>>(A, S, divB, phi, err);
  return hipGetLastError();
}

hipError_t launch_divb_update(hipStream_t st, const DivbArgs& A, MxState S, const double* phi, const double* divB,
                              const MxEpilogue* epi) {
  const unsigned nb = divb_blocks(A.G);
  if (epi) k_divb_update_mx<true><<<nb, DIVB_TX * DIVB_TY, 0, st>>>(A, S, phi, divB, *epi);
  else k_divb_update_mx<false><<<nb, DIVB_TX * DIVB_TY, 0, st>>>(A, S, phi, divB, MxEpilogue{});
  return hipGetLastError();
}

hipError_t launch_eos_temp(hipStream_t st, const Geom& G, const EosDev& E, MxState S, double* T) {
  unsigned nb = nblk_(G.bs);
  if (nb > 65536u) nb = 65536u;
  k_eos_temp_mx<<<nb, MRM_BLOCK, 0, st>>>(G, E, S, T);
  return hipGetLastError();
}

hipError_t launch_spitzer_stage(hipStream_t st, const Geom& G, const SptParams& P, double wdt, MxState S,
                                const double* T, const double* sflx_in, const double* sflx0, double* sflx_out,
                                double* eres) {
  k_spitzer_stage<<<spitzer_blocks(G), SPT_TX * SPT_TY, 0, st>>>(G, P, wdt, SptStateMx{S.f, S.e, G.bs}, T, sflx_in,
                                                                 sflx0, sflx_out, eres);
  return hipGetLastError();
}

hipError_t launch_exchange_b(hipStream_t st, const Geom& G, MxState S) {
  const int nx = G.le0 - G.lb0 + 1;
  long n1 = (long)nx * (G.le2 - G.lb2 + 1) * G.g1;
  k_exchange_arr<1><<<nblk_(n1), MRM_BLOCK, 0, st>>>(G, S.f + 4 * G.bs, 3, G.bs);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  long n2 = (long)nx * G.n1 * G.g2;
  k_exchange_arr<2><<<nblk_(n2), MRM_BLOCK, 0, st>>>(G, S.f + 4 * G.bs, 3, G.bs);
  return hipGetLastError();
}

}  // namespace mixed
