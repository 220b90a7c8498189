// mrm_divb.h -- Clean_divB (clean_divB.C:35-369) on the device, FP64: the
// 4th-order div B of the interior, the damped potential phi (kept from call to
// call), its periodic and antisymmetric vertical ghosts, and the correction
// B += grad phi, e += phi div B.  The two cell kernels march along k; the
// update kernel can finish each cell as the step's last TVD sweep would (TCheck, EOS
// pressure, characteristic speed, Adjust_Valf_Max statistics) when the
// cleaning runs between TVDlimit and TCheck (solver.C:325-345).
#pragma once
#include "mrm_kernels.h"

namespace MRM_NS {

// Tiling of the interior for the two cell kernels: a 64 x 4 (i, j) tile per
// 256-thread block, one thread per column, marching along k through a chunk
// of DIVB_KC planes; the k-neighbours of the marched array sit in a five-plane
// register window, the i- and j-neighbours come through the caches.
constexpr int DIVB_TX = 64, DIVB_TY = 4, DIVB_KC = 32;

struct DivbTile {
  long c0;        // column offset i + j s1
  int k0, k1;     // planes [k0, k1) of this block
  bool ok;
};
__device__ __forceinline__ DivbTile divb_tile(const Geom& G) {
  const int nx = G.le0 - G.lb0 + 1, ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  const int ntx = (nx + DIVB_TX - 1) / DIVB_TX, nty = (ny + DIVB_TY - 1) / DIVB_TY;
  const int b = blockIdx.x;
  const int bx = b % ntx, by = (b / ntx) % nty, bz = b / (ntx * nty);
  const int i = G.lb0 + bx * DIVB_TX + (int)(threadIdx.x % DIVB_TX);
  const int j = G.lb1 + by * DIVB_TY + (int)(threadIdx.x / DIVB_TX);
  DivbTile t;
  t.k0 = G.lb2 + bz * DIVB_KC;
  t.k1 = min(t.k0 + DIVB_KC, G.lb2 + nz);
  t.ok = i <= G.le0 && j <= G.le1;
  t.c0 = i + (long)j * G.s1;
  return t;
}
inline unsigned divb_blocks(const Geom& G) {
  const int nx = G.le0 - G.lb0 + 1, ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  return (unsigned)(((nx + DIVB_TX - 1) / DIVB_TX) * ((ny + DIVB_TY - 1) / DIVB_TY) *
                    ((nz + DIVB_KC - 1) / DIVB_KC));
}

// divB of the interior (clean_divB.C:125-162) and max |divB| into err; with
// PHI, the next pass's potential update phi += fac divB - damp phi
// (clean_divB.C:168-178) of the same cell
template <bool PHI>
__global__ void __launch_bounds__(DIVB_TX * DIVB_TY) k_divb(DivbArgs A, const double* __restrict__ U,
                                                            double* __restrict__ divB, double* __restrict__ phi,
                                                            unsigned long long* err) {
  __shared__ double lds[DIVB_TX * DIVB_TY / 64];
  const Geom& G = A.G;
  const double* Bx = U + 5 * G.bs;
  const double* By = U + 6 * G.bs;
  const double* Bz = U + 7 * G.bs;
  const long s1 = G.s1, s2 = G.s2;
  double m[1] = {0.0};
  const DivbTile t = divb_tile(G);
  if (t.ok) {
    double zw[5];
#pragma unroll
    for (int q = 0; q < 4; q++) zw[q] = Bz[t.c0 + (long)(t.k0 - 2 + q) * s2];
    for (int k = t.k0; k < t.k1; k++) {
      const long n = t.c0 + (long)k * s2;
      zw[4] = Bz[n + 2 * s2];
      double d = 0.0;
      d += (A.w1[0] * (Bx[n + 1] - Bx[n - 1]) + A.w2[0] * (Bx[n + 2] - Bx[n - 2]));
      d += (A.w1[1] * (By[n + s1] - By[n - s1]) + A.w2[1] * (By[n + 2 * s1] - By[n - 2 * s1]));
      d += (A.w1[2] * (zw[3] - zw[1]) + A.w2[2] * (zw[4] - zw[0]));
      divB[n] = d;
      m[0] = fmax(m[0], fabs(d));
      if (PHI) phi[n] += A.fac * d - A.damp * phi[n];
#pragma unroll
      for (int q = 0; q < 4; q++) zw[q] = zw[q + 1];
    }
  }
  block_reduce<OpMax, 1>(m, lds);
  if (threadIdx.x == 0) atomic_max_nonneg(err, m[0]);
}

// periodic ghosts of na arrays base + q * vstride along dim DIM (the ranges of
// exchange_B_acc / exchange_single_acc on one rank, exchange.C:785-1120)
template <int DIM, class T>
__global__ void __launch_bounds__(MRM_BLOCK) k_exchange_arr(Geom G, T* __restrict__ base, int na,
                                                            long vstride) {
  const int nx = G.le0 - G.lb0 + 1;
  const int gh = DIM == 1 ? G.g1 : G.g2;
  const int lb = DIM == 1 ? G.lb1 : G.lb2, le = DIM == 1 ? G.le1 : G.le2;
  const int nb = DIM == 1 ? (G.le2 - G.lb2 + 1) : G.n1;
  const int ob = DIM == 1 ? G.lb2 : 0;
  const long sd = DIM == 1 ? G.s1 : G.s2, so = DIM == 1 ? G.s2 : G.s1;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nx * nb * gh) return;
  int i = G.lb0 + (int)(t % nx);
  long r = t / nx;
  int b = ob + (int)(r % nb);
  int igh = (int)(r / nb);
  long c0 = i + b * so;
  long src_r = c0 + (long)(lb + igh) * sd, dst_r = c0 + (long)(le + 1 + igh) * sd;
  long src_l = c0 + (long)(le - gh + 1 + igh) * sd, dst_l = c0 + (long)(lb - gh + igh) * sd;
  for (int q = 0; q < na; q++) {
    T* a = base + q * vstride;
    a[dst_r] = a[src_r];
    a[dst_l] = a[src_l];
  }
}

// phi antisymmetric across the bottom and top of every interior column
// (clean_divB.C:186-217)
__global__ void __launch_bounds__(MRM_BLOCK) k_phi_vertical(Geom G, double* __restrict__ phi) {
  const int ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)ny * nz) return;
  const long c0 = (G.lb1 + (int)(t % ny)) * G.s1 + (long)(G.lb2 + (int)(t / ny)) * G.s2;
  const int lb = G.lb0, le = G.le0;
  phi[c0 + lb - 1] = -phi[c0 + lb];
  phi[c0 + lb - 2] = -phi[c0 + lb + 1];
  phi[c0 + le + 1] = -phi[c0 + le];
  phi[c0 + le + 2] = -phi[c0 + le - 1];
}

// B += grad phi, e += phi divB (clean_divB.C:219-262).  With EPI the cell is
// then finished: TCheck (Add_Sources_Integrate_SR.C:591-624), its EOS
// pressure, its characteristic speed and, with X.valf, the Adjust_Valf_Max
// statistics for the next step.
template <bool EPI>
__global__ void __launch_bounds__(DIVB_TX * DIVB_TY) k_divb_update(DivbArgs A, double* __restrict__ U,
                                                                   const double* __restrict__ phi,
                                                                   const double* __restrict__ divB, TvdEpilogue X) {
  __shared__ double lds[5 * DIVB_TX * DIVB_TY / 64];
  const Geom& G = A.G;
  const long bs = G.bs, s1 = G.s1, s2 = G.s2;
  double cm[1] = {0.0};
  ValfAcc va;
  const DivbTile t = divb_tile(G);
  if (t.ok) {
    double pw[5];
#pragma unroll
    for (int q = 0; q < 4; q++) pw[q] = phi[t.c0 + (long)(t.k0 - 2 + q) * s2];
    for (int k = t.k0; k < t.k1; k++) {
      const long n = t.c0 + (long)k * s2;
      pw[4] = phi[n + 2 * s2];
      U[5 * bs + n] += (A.w1[0] * (phi[n + 1] - phi[n - 1]) + A.w2[0] * (phi[n + 2] - phi[n - 2]));
      U[6 * bs + n] += (A.w1[1] * (phi[n + s1] - phi[n - s1]) + A.w2[1] * (phi[n + 2 * s1] - phi[n - 2 * s1]));
      U[7 * bs + n] += (A.w1[2] * (pw[3] - pw[1]) + A.w2[2] * (pw[4] - pw[0]));
      U[4 * bs + n] += pw[2] * divB[n];
      if (A.qres) A.qres[n] += pw[2] * divB[n] * A.invdt;
      if (EPI) {
        double uu[MRM_NV];
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) uu[q] = U[q * bs + n];
        tcheck_vals(uu, X.rho_min, X.eps_min, X.eps_max, X.vmax);
#pragma unroll
        for (int q = 0; q < 5; q++) U[q * bs + n] = uu[q];
        double m1 = uu[1], m2 = uu[2], m3 = uu[3], ei = uu[4], pr[5];
        cons_to_prim_cell<2>(X.E, uu[0], m1, m2, m3, ei, pr);
        X.Pnew[n] = pr[0];
        if (!(X.skip_top_cmax && t.c0 % s1 == G.le0)) cm[0] = smax(cm[0], point_cmax_cons(uu, pr[0], X.inv_va2max));
        if (X.valf) valf_cell(va, uu[0], m1, m2, m3, ei, pr[0], X.vlim95, X.elim95);
      }
#pragma unroll
      for (int q = 0; q < 4; q++) pw[q] = pw[q + 1];
    }
  }
  if (EPI) {
    block_reduce<OpMax, 1>(cm, lds);
    if (threadIdx.x == 0) atomic_max_nonneg(X.cmax, cm[0]);
    if (X.valf) {
      __syncthreads();
      valf_flush(va, lds, X.valf);
    }
  }
}

}  // namespace MRM_NS
