// mrm_spitzer.h -- Spitzer heat conduction in hyperbolic form (param_spitzer
// > 0) on the device, FP64, for the per-operation entry points: the heat flux
// along B, sflx, is a cell field with its own residual Rflx and RK update.
//   residual (mhdres_SR.C:350-359, 486-521, 680-690): the energy flux carries
//     sflx B_d along every sweep; each sweep damps sflx with a 4th difference
//     and adds B_d dT/dx_d to B.grad T; then sflx relaxes towards the
//     saturation-limited Spitzer flux -kappa B.grad T / |B|^2 at the rate
//     nu_hyp, bounded by the characteristic-speed margin and by avr_max/dt;
//   integrate / SaveCons (Add_Sources_Integrate_SR.C:410-419, 504-517);
//   ghosts (boundary_pdmp_1_fftw3.C:306-310, 455-466; exchange.C:111-120).
#pragma once
#include "mrm_kernels.h"

namespace MRM_NS {

// one interior cell per thread; U primitive (after ConsToPrim), T its
// temperature over the whole buffer.  With t2 (diagnostics), tvar2 += the
// conductive energy flux divergence (mhdres_SR.C:368-376, 689-694)
__global__ void __launch_bounds__(MRM_BLOCK) k_spitzer_resid(Geom G, SptParams S, const double* __restrict__ U,
                                                             const double* __restrict__ T,
                                                             const double* __restrict__ sflx, double* __restrict__ R,
                                                             double* __restrict__ Rflx, double* __restrict__ t2) {
  const long bs = G.bs;
  int i, j, k;
  if (!interior_ijk(G, blockIdx.x * (long)blockDim.x + threadIdx.x, i, j, k)) return;
  const long n = i + j * G.s1 + (long)k * G.s2;
  const long st[3] = {1, G.s1, G.s2};
  double re = 0.0, rf = Rflx[n], bgt = 0.0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const long s = st[d];
    const double* Bd = U + (5 + d) * bs;
    const double w1 = S.w1[d], w2 = S.w2[d];
    re -= w1 * (sflx[n + s] * Bd[n + s] - sflx[n - s] * Bd[n - s]) +
          w2 * (sflx[n + 2 * s] * Bd[n + 2 * s] - sflx[n - 2 * s] * Bd[n - 2 * s]);
    rf -= S.hyp_diff * ((sflx[n + 2 * s] + sflx[n - 2 * s]) - 4. * (sflx[n + s] + sflx[n - s]) + 6. * sflx[n]);
    bgt += Bd[n] * (w1 * (T[n + s] - T[n - s]) + w2 * (T[n + 2 * s] - T[n - 2 * s]));
  }
  R[4 * bs + n] += re;
  if (t2) t2[n] += re;
  const double v0 = U[bs + n], v1 = U[2 * bs + n], v2 = U[3 * bs + n];
  const double b0 = U[5 * bs + n], b1 = U[6 * bs + n], b2 = U[7 * bs + n];
  const double vv = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
  const double bb = b0 * b0 + b1 * b1 + b2 * b2;
  const double tempi = T[n];
  const double c_lim = S.cmax_hyp - vv;
  const double tt32 = tempi * sqrt(tempi);
  double kap = S.k0 * tempi * tt32;
  const double F_sat = S.sat_flx * U[n] * tt32;
  const double invBB = 1.0 / ((1e-100 < bb) ? bb : 1e-100);
  double F = -kap * bgt;
  const double f_lim = F_sat / (F_sat + fabs(F * sqrt(invBB)));
  F *= invBB * f_lim;
  kap *= f_lim;
  const double nu_c = c_lim * c_lim * U[4 * bs + n] / (tempi * kap);
  const double nu = (S.nu_hyp_max < nu_c) ? S.nu_hyp_max : nu_c;
  Rflx[n] = rf + nu * (F - sflx[n]);
}

// sflx = sflx0 + wdt Rflx, Rflx = 0 (interior)
__global__ void __launch_bounds__(MRM_BLOCK) k_spitzer_integrate(Geom G, double wdt, double* __restrict__ sflx,
                                                                 const double* __restrict__ sflx0,
                                                                 double* __restrict__ Rflx) {
  int i, j, k;
  if (!interior_ijk(G, blockIdx.x * (long)blockDim.x + threadIdx.x, i, j, k)) return;
  const long n = i + j * G.s1 + (long)k * G.s2;
  sflx[n] = sflx0[n] + wdt * Rflx[n];
  Rflx[n] = 0.0;
}

// sflx0 = sflx (interior)
__global__ void __launch_bounds__(MRM_BLOCK) k_spitzer_save(Geom G, const double* __restrict__ sflx,
                                                            double* __restrict__ sflx0) {
  int i, j, k;
  if (!interior_ijk(G, blockIdx.x * (long)blockDim.x + threadIdx.x, i, j, k)) return;
  const long n = i + j * G.s1 + (long)k * G.s2;
  sflx0[n] = sflx[n];
}

// vertical ghosts of sflx over every (j,k): zero at the bottom; at the top
// symmetric under a hot plate, else the top two interior layers and the
// ghosts zero (zero flux through the boundary)
__global__ void __launch_bounds__(MRM_BLOCK) k_spitzer_bc(Geom G, int hot_plate, double* __restrict__ sflx) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)G.n1 * G.n2) return;
  const long base = (t % G.n1) * G.s1 + (t / G.n1) * G.s2;
  sflx[base + G.lb0 - 1] = 0.0;
  sflx[base + G.lb0 - 2] = 0.0;
  const long n1 = base + G.le0, n2 = n1 - 1;
  if (hot_plate) {
    sflx[n1 + 1] = sflx[n1];
    sflx[n1 + 2] = sflx[n2];
  } else {
    sflx[n2] = 0.0;
    sflx[n1] = 0.0;
    sflx[n1 + 1] = 0.0;
    sflx[n1 + 2] = 0.0;
  }
}

// ---------------------------------------------------------------- fused step
// The fused step keeps conservative states and carries only the pressure, so
// conduction gets two kernels per stage around the stage kernel:
// k_eos_temp forms ConsToPrim's temperature of the stage's input state over
// the whole buffer (cons_to_prim.C:108-198), and k_spitzer_stage does what
// k_spitzer_resid and k_spitzer_integrate do for one stage -- the conductive
// energy term goes to eres, which the stage kernel adds to the energy
// residual after its three sweeps, and sflx_out = sflx0 + wdt Rflx with Rflx
// formed from sflx_in (the integrator leaves Rflx zero between stages, so it
// never needs storing).

__global__ void __launch_bounds__(MRM_BLOCK) k_eos_temp(Geom G, EosDev E, const double* __restrict__ U,
                                                        double* __restrict__ T) {
  const long bs = G.bs;
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
    double m1 = U[bs + n], m2 = U[2 * bs + n], m3 = U[3 * bs + n], e = U[4 * bs + n], pr[5];
    cons_to_prim_cell<2, true>(E, U[n], m1, m2, m3, e, pr);
    T[n] = pr[1];
  }
}

// A 64 x 4 (i, j) tile of columns per 256-thread block, one column per
// thread, marching along k through a chunk of SPT_KC planes: the
// k-neighbours of sflx, B_z and T come from five-plane register windows, the
// i- and j-neighbours through the caches.  U conservative (the stage's
// input), T its temperature, sflx_in with its ghosts filled.
constexpr int SPT_TX = 64, SPT_TY = 4, SPT_KC = 32;

inline unsigned spitzer_blocks(const Geom& G) {
  const int nx = G.le0 - G.lb0 + 1, ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  return (unsigned)(((nx + SPT_TX - 1) / SPT_TX) * ((ny + SPT_TY - 1) / SPT_TY) *
                    ((nz + SPT_KC - 1) / SPT_KC));
}

// the conservative state as the kernel reads it: FP64 SoA U
struct SptStateF64 {
  const double* U;
  long bs;
  __device__ __forceinline__ double rho(long n) const { return U[n]; }
  __device__ __forceinline__ double mom(int q, long n) const { return U[(1 + q) * bs + n]; }
  __device__ __forceinline__ double ene(long n) const { return U[4 * bs + n]; }
  __device__ __forceinline__ double mag(int q, long n) const { return U[(5 + q) * bs + n]; }
};

template <class ST>
__global__ void __launch_bounds__(SPT_TX * SPT_TY) k_spitzer_stage(Geom G, SptParams S, double wdt, ST U,
                                                                   const double* __restrict__ T,
                                                                   const double* __restrict__ sflx,
                                                                   const double* __restrict__ sflx0,
                                                                   double* __restrict__ sflx_out,
                                                                   double* __restrict__ eres) {
  const long s1 = G.s1, s2 = G.s2;
  const int nx = G.le0 - G.lb0 + 1, ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  const int ntx = (nx + SPT_TX - 1) / SPT_TX, nty = (ny + SPT_TY - 1) / SPT_TY;
  const int b = blockIdx.x;
  const int bx = b % ntx, by = (b / ntx) % nty, bz = b / (ntx * nty);
  const int i = G.lb0 + bx * SPT_TX + (int)(threadIdx.x % SPT_TX);
  const int j = G.lb1 + by * SPT_TY + (int)(threadIdx.x / SPT_TX);
  if (i > G.le0 || j > G.le1) return;
  const int k0 = G.lb2 + bz * SPT_KC, k1 = min(k0 + SPT_KC, G.lb2 + nz);
  const long c0 = i + (long)j * s1;
  auto B0 = [&](long m) { return U.mag(0, m); };
  auto B1 = [&](long m) { return U.mag(1, m); };
  auto B2 = [&](long m) { return U.mag(2, m); };
  double fw[5], bw[5], tw[5];   // sflx, B_z, T at planes k-2 .. k+2
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const long m = c0 + (long)(k0 - 2 + q) * s2;
    fw[q + 1] = sflx[m];
    bw[q + 1] = B2(m);
    tw[q + 1] = T[m];
  }
  for (int k = k0; k < k1; k++) {
    const long n = c0 + (long)k * s2;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      fw[q] = fw[q + 1];
      bw[q] = bw[q + 1];
      tw[q] = tw[q + 1];
    }
    fw[4] = sflx[n + 2 * s2];
    bw[4] = B2(n + 2 * s2);
    tw[4] = T[n + 2 * s2];
    const double fc = fw[2];
    double re = 0.0, rf = 0.0, bgt = 0.0;
    {   // dim 0
      const double w1 = S.w1[0], w2 = S.w2[0];
      re -= w1 * (sflx[n + 1] * B0(n + 1) - sflx[n - 1] * B0(n - 1)) +
            w2 * (sflx[n + 2] * B0(n + 2) - sflx[n - 2] * B0(n - 2));
      rf -= S.hyp_diff * ((sflx[n + 2] + sflx[n - 2]) - 4. * (sflx[n + 1] + sflx[n - 1]) + 6. * fc);
      bgt += B0(n) * (w1 * (T[n + 1] - T[n - 1]) + w2 * (T[n + 2] - T[n - 2]));
    }
    {   // dim 1
      const double w1 = S.w1[1], w2 = S.w2[1];
      re -= w1 * (sflx[n + s1] * B1(n + s1) - sflx[n - s1] * B1(n - s1)) +
            w2 * (sflx[n + 2 * s1] * B1(n + 2 * s1) - sflx[n - 2 * s1] * B1(n - 2 * s1));
      rf -= S.hyp_diff * ((sflx[n + 2 * s1] + sflx[n - 2 * s1]) - 4. * (sflx[n + s1] + sflx[n - s1]) + 6. * fc);
      bgt += B1(n) * (w1 * (T[n + s1] - T[n - s1]) + w2 * (T[n + 2 * s1] - T[n - 2 * s1]));
    }
    {   // dim 2, from the windows
      const double w1 = S.w1[2], w2 = S.w2[2];
      re -= w1 * (fw[3] * bw[3] - fw[1] * bw[1]) + w2 * (fw[4] * bw[4] - fw[0] * bw[0]);
      rf -= S.hyp_diff * ((fw[4] + fw[0]) - 4. * (fw[3] + fw[1]) + 6. * fc);
      bgt += bw[2] * (w1 * (tw[3] - tw[1]) + w2 * (tw[4] - tw[0]));
    }
    eres[n] = re;
    // primitives of the cell as ConsToPrim forms them (cons_to_prim.C:110-118)
    const double d = U.rho(n);
    const double dn = 1.0 / d;
    const double v0 = U.mom(0, n) * dn, v1 = U.mom(1, n) * dn, v2 = U.mom(2, n) * dn;
    const double eint = U.ene(n) - 0.5 * (v0 * v0 + v1 * v1 + v2 * v2) * d;
    const double b0 = B0(n), b1 = B1(n), b2 = bw[2];
    const double vv = sqrt(v0 * v0 + v1 * v1 + v2 * v2);
    const double bb = b0 * b0 + b1 * b1 + b2 * b2;
    const double tempi = tw[2];
    const double c_lim = S.cmax_hyp - vv;
    const double tt32 = tempi * sqrt(tempi);
    double kap = S.k0 * tempi * tt32;
    const double F_sat = S.sat_flx * d * tt32;
    const double invBB = 1.0 / ((1e-100 < bb) ? bb : 1e-100);
    double F = -kap * bgt;
    const double f_lim = F_sat / (F_sat + fabs(F * sqrt(invBB)));
    F *= invBB * f_lim;
    kap *= f_lim;
    const double nu_c = c_lim * c_lim * eint / (tempi * kap);
    const double nu = (S.nu_hyp_max < nu_c) ? S.nu_hyp_max : nu_c;
    rf = rf + nu * (F - fc);
    sflx_out[n] = sflx0[n] + wdt * rf;
  }
}

}  // namespace MRM_NS
