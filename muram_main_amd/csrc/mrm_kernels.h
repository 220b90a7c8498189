// mrm_kernels.h -- HIP kernels of the MHD step (FP64 path), one per
// reference operation.  Launch geometry is chosen by the host code in
// mrm_step.hip; every kernel checks its own index range.
#pragma once
#include "mrm_device.h"

namespace MRM_NS {



// ---------------------------------------------------------------- helpers
// non-negative doubles order like their bit patterns: max via integer atomics
// Loads and stores at a 32-bit byte offset from a wave-uniform base pointer:
// the address is base + zext(offset), which the compiler issues as an
// SGPR-base + VGPR-offset access instead of 64-bit address arithmetic per
// lane.  The host keeps 8 * bufsize below 2^32.
template <class T>
__device__ __forceinline__ T ld32(const T* base, unsigned n) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + (unsigned)(n * (unsigned)sizeof(T)));
}
template <class T>
__device__ __forceinline__ void st32(T* base, unsigned n, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + (unsigned)(n * (unsigned)sizeof(T))) = v;
}

__device__ __forceinline__ void atomic_max_nonneg(unsigned long long* slot, double v) {
  if (v > 0.0) atomicMax(slot, (unsigned long long)__double_as_longlong(v));
}

struct OpAdd { static __device__ __forceinline__ double f(double a, double b) { return a + b; } };
struct OpMax { static __device__ __forceinline__ double f(double a, double b) { return fmax(a, b); } };

// block-wide reduction of N values; the result is valid in thread 0.
// lds must hold N * blockDim.x / 64 doubles.
template <class OP, int N>
__device__ __forceinline__ void block_reduce(double (&v)[N], double* lds) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < N; q++) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) x = OP::f(x, __shfl_xor(x, o, 64));
    if (lane == 0) lds[q * nw + wid] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < N; q++) {
      double x = lds[q * nw];
      for (int w = 1; w < nw; w++) x = OP::f(x, lds[q * nw + w]);
      v[q] = x;
    }
  }
}

// interior index -> (i, j, k)
__device__ __forceinline__ bool interior_ijk(const Geom& G, long t, int& i, int& j, int& k) {
  const int nx = G.le0 - G.lb0 + 1, ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  if (t >= (long)nx * ny * nz) return false;
  i = (int)(t % nx);
  long r = t / nx;
  j = (int)(r % ny);
  k = (int)(r / ny);
  i += G.lb0; j += G.lb1; k += G.lb2;
  return true;
}

// ---------------------------------------------------------------- ConsToPrim
template <int NOUT>
__global__ void __launch_bounds__(MRM_BLOCK) k_cons_to_prim(Geom G, EosDev E, double* __restrict__ U,
                                                            double* __restrict__ F) {
  const long bs = G.bs;
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
    double d = U[n], m1 = U[bs + n], m2 = U[2 * bs + n], m3 = U[3 * bs + n], e = U[4 * bs + n];
    double out[5];
    cons_to_prim_cell<NOUT>(E, d, m1, m2, m3, e, out);
    U[bs + n] = m1; U[2 * bs + n] = m2; U[3 * bs + n] = m3; U[4 * bs + n] = e;
    F[n] = out[0];
    if (NOUT > 2) {
#pragma unroll
      for (int q = 1; q < 5; q++) F[q * bs + n] = out[q];
    }
  }
}

// ---------------------------------------------------------------- Adjust_Valf_Max
__global__ void __launch_bounds__(MRM_BLOCK) k_valf_reduce(Geom G, const double* __restrict__ U,
                                                           const double* __restrict__ pres, double vlim95,
                                                           double elim95, unsigned long long* red) {
  __shared__ double lds[3 * MRM_BLOCK / 64];
  __shared__ unsigned long long cnt[2];
  if (threadIdx.x < 2) cnt[threadIdx.x] = 0;
  __syncthreads();
  const long bs = G.bs;
  double m[3] = {0.0, 0.0, 0.0};
  unsigned long long c0 = 0, c1 = 0;
  const long nint = (long)(G.le0 - G.lb0 + 1) * (G.le1 - G.lb1 + 1) * (G.le2 - G.lb2 + 1);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nint; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    interior_ijk(G, t, i, j, k);
    long node = i + j * G.s1 + k * G.s2;
    double vx = U[bs + node], vy = U[2 * bs + node], vz = U[3 * bs + node];
    double vv = sqrt(vx * vx + vy * vy + vz * vz);
    double ee = U[4 * bs + node] / U[node];
    m[0] = fmax(m[0], vv);
    m[1] = fmax(m[1], 5 / 3. * pres[node] / U[node]);
    m[2] = fmax(m[2], ee);
    if (vv >= vlim95) c0++;
    if (ee >= elim95) c1++;
  }
  atomicAdd(&cnt[0], c0);
  atomicAdd(&cnt[1], c1);
  block_reduce<OpMax, 3>(m, lds);
  if (threadIdx.x == 0) {
    atomic_max_nonneg(&red[RED_VMAX], m[0]);
    atomic_max_nonneg(&red[RED_CS2MAX], m[1]);
    atomic_max_nonneg(&red[RED_EMAX], m[2]);
    atomicAdd(&red[RED_CNT_V], cnt[0]);
    atomicAdd(&red[RED_CNT_E], cnt[1]);
  }
}

// Adjust_Valf_Max's per-cell terms (Adjust_Valf_Max.C:58-81) of one interior
// cell from ConsToPrim's primitive values, accumulated per thread: max |v|,
// max (5/3) p/rho, max e_int/rho, and the counts of |v| >= 0.95 v_lim and of
// e_int/rho >= 0.95 e_lim (kept as doubles: exact integers)
struct ValfAcc {
  double m[3] = {0.0, 0.0, 0.0};
  double n[2] = {0.0, 0.0};
};
__device__ __forceinline__ void valf_cell(ValfAcc& a, double d, double v1, double v2, double v3, double eint,
                                          double p, double vlim95, double elim95) {
  const double vv = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
  const double ee = eint / d;
  a.m[0] = fmax(a.m[0], vv);
  a.m[1] = fmax(a.m[1], 5 / 3. * p / d);
  a.m[2] = fmax(a.m[2], ee);
  if (vv >= vlim95) a.n[0] += 1.0;
  if (ee >= elim95) a.n[1] += 1.0;
}
// the same accumulators held in LDS, one column per thread (acc[q * stride]),
// for kernels whose registers are all taken by the main loop
__device__ __forceinline__ void valf_lds_init(double* acc, int stride) {
#pragma unroll
  for (int q = 0; q < 5; q++) acc[q * stride] = 0.0;
}
__device__ __forceinline__ void valf_lds_cell(double* acc, int stride, double d, double v1, double v2, double v3,
                                              double eint, double p, double vlim95, double elim95) {
  const double vv = sqrt(v1 * v1 + v2 * v2 + v3 * v3);
  const double ee = eint / d;
  acc[0] = fmax(acc[0], vv);
  acc[stride] = fmax(acc[stride], 5 / 3. * p / d);
  acc[2 * stride] = fmax(acc[2 * stride], ee);
  if (vv >= vlim95) acc[3 * stride] += 1.0;
  if (ee >= elim95) acc[4 * stride] += 1.0;
}
__device__ __forceinline__ ValfAcc valf_lds_load(const double* acc, int stride) {
  ValfAcc a;
#pragma unroll
  for (int q = 0; q < 3; q++) a.m[q] = acc[q * stride];
  a.n[0] = acc[3 * stride];
  a.n[1] = acc[4 * stride];
  return a;
}

// block reduction of the accumulators into red[RED_VMAX..RED_CNT_E]; every
// thread of the block calls it; lds holds 5 * blockDim.x / 64 doubles
__device__ __forceinline__ void valf_flush(ValfAcc& a, double* lds, unsigned long long* red) {
  const int nw = blockDim.x >> 6;
  block_reduce<OpMax, 3>(a.m, lds);
  block_reduce<OpAdd, 2>(a.n, lds + 3 * nw);
  if (threadIdx.x == 0) {
    atomic_max_nonneg(&red[RED_VMAX], a.m[0]);
    atomic_max_nonneg(&red[RED_CS2MAX], a.m[1]);
    atomic_max_nonneg(&red[RED_EMAX], a.m[2]);
    atomicAdd(&red[RED_CNT_V], (unsigned long long)a.n[0]);
    atomicAdd(&red[RED_CNT_E], (unsigned long long)a.n[1]);
  }
}

// ---------------------------------------------------------------- MHD_Residual

__device__ __forceinline__ void load_pt(const double* __restrict__ U, const double* __restrict__ P, long bs,
                                        long node, double v[MRM_NV], double& p) {
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) v[q] = U[q * bs + node];
  p = P[node];
}

template <int D, bool DIAG = false>
__device__ __forceinline__ void resid_dir(const double* __restrict__ U, const double* __restrict__ Pr, long bs,
                                          long node, long str, const double vc[MRM_NV], double pc,
                                          const ResidParams& RP, double acc[MRM_NV], double J[3],
                                          double* dg = nullptr) {
  double v[5][MRM_NV], p[5];
  load_pt(U, Pr, bs, node - 2 * str, v[0], p[0]);
  load_pt(U, Pr, bs, node - str, v[1], p[1]);
  load_pt(U, Pr, bs, node + str, v[3], p[3]);
  load_pt(U, Pr, bs, node + 2 * str, v[4], p[4]);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) v[2][q] = vc[q];
  p[2] = pc;
  double res[MRM_NV];
  sweep_residual<D, DIAG>(v, p, RP.w1[D], RP.w2[D], RP.inv_va2max, res, dg);
  if (RP.eta > 0.0) {
    constexpr int D2 = D == 0 ? 1 : 0;
    constexpr int D3 = D == 2 ? 1 : 2;
    const double dB2 = RP.w1[D] * (v[3][5 + D2] - v[1][5 + D2]) + RP.w2[D] * (v[4][5 + D2] - v[0][5 + D2]);
    const double dB3 = RP.w1[D] * (v[3][5 + D3] - v[1][5 + D3]) + RP.w2[D] * (v[4][5 + D3] - v[0][5 + D3]);
    double s1[3], s2[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      s1[q] = v[3][5 + q] + v[1][5 + q];
      s2[q] = v[4][5 + q] + v[0][5 + q];
    }
    eta_sweep<D>(dB2, dB3, v[2] + 5, s1, s2, RP.eta, RP.ww0[D], RP.ww1[D], RP.ww2[D], J, res);
  }
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
}

// R += residual over the interior; cmax of the interior points into red.
// DIAG: tvar1 and tvar3 of the cell (mhdres_SR.C:368-383, 434-435)
template <bool DIAG = false>
__global__ void __launch_bounds__(MRM_BLOCK) k_residual(Geom G, ResidParams RP, const double* __restrict__ U,
                                                        const double* __restrict__ Pr, double* __restrict__ R,
                                                        unsigned long long* red, DiagArgs DA) {
  __shared__ double lds[MRM_BLOCK / 64];
  const long bs = G.bs;
  double cm[1] = {0.0};
  int i, j, k;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (interior_ijk(G, t, i, j, k)) {
    long node = i + j * G.s1 + k * G.s2;
    double vc[MRM_NV], pc;
    load_pt(U, Pr, bs, node, vc, pc);
    double acc[MRM_NV];
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) acc[q] = R[q * bs + node];
    double J[3] = {0.0, 0.0, 0.0};
    double dg[2] = {0.0, 0.0};
    resid_dir<0, DIAG>(U, Pr, bs, node, 1, vc, pc, RP, acc, J, dg);
    resid_dir<1, DIAG>(U, Pr, bs, node, G.s1, vc, pc, RP, acc, J, dg);
    resid_dir<2, DIAG>(U, Pr, bs, node, G.s2, vc, pc, RP, acc, J, dg);
    if (RP.eta > 0.0) acc[4] += RP.eta * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) R[q * bs + node] = acc[q];
    if (DIAG) {
      DA.dg[DG_T1 * bs + node] += dg[0];
      DA.dg[DG_T3 * bs + node] += dg[1];
    }
    cm[0] = smax(0.0, point_cmax(vc, pc, RP.inv_va2max));
  }
  block_reduce<OpMax, 1>(cm, lds);
  if (threadIdx.x == 0) atomic_max_nonneg(&red[RED_CMAX], cm[0]);
}

// cmax over the ghost points that the sweeps of MHD_Residual visit: the two
// ghost layers at each end of every interior line (mhdres_SR.C:237-328).
__global__ void __launch_bounds__(MRM_BLOCK) k_cmax_faces(Geom G, double inv_va2max, const double* __restrict__ U,
                                                          const double* __restrict__ Pr, unsigned long long* red) {
  __shared__ double lds[MRM_BLOCK / 64];
  const int nx = G.le0 - G.lb0 + 1, ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  const long f0 = (long)2 * G.g0 * ny * nz, f1 = (long)2 * G.g1 * nx * nz, f2 = (long)2 * G.g2 * nx * ny;
  double cm[1] = {0.0};
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < f0 + f1 + f2; t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    if (t < f0) {                          // dim-0 ghosts at interior (j,k)
      long r = t;
      int a = (int)(r % (2 * G.g0)); r /= 2 * G.g0;
      j = G.lb1 + (int)(r % ny); k = G.lb2 + (int)(r / ny);
      i = a < G.g0 ? a : G.le0 + 1 + (a - G.g0);
    } else if (t < f0 + f1) {              // dim-1 ghosts at interior (i,k)
      long r = t - f0;
      i = G.lb0 + (int)(r % nx); r /= nx;
      int a = (int)(r % (2 * G.g1)); k = G.lb2 + (int)(r / (2 * G.g1));
      j = a < G.g1 ? a : G.le1 + 1 + (a - G.g1);
    } else {                               // dim-2 ghosts at interior (i,j)
      long r = t - f0 - f1;
      i = G.lb0 + (int)(r % nx); r /= nx;
      j = G.lb1 + (int)(r % ny); int a = (int)(r / ny);
      k = a < G.g2 ? a : G.le2 + 1 + (a - G.g2);
    }
    long node = i + j * G.s1 + k * G.s2;
    double v[MRM_NV], p;
    load_pt(U, Pr, G.bs, node, v, p);
    cm[0] = smax(cm[0], point_cmax(v, p, inv_va2max));
  }
  block_reduce<OpMax, 1>(cm, lds);
  if (threadIdx.x == 0) atomic_max_nonneg(&red[RED_CMAX], cm[0]);
}

// ---------------------------------------------------------------- sources / integrate
// DIAG: tvar4..6 set and tvar7 added to (Add_Sources_Integrate_SR.C:122-296)
// ZERO_R: Res := 0 is stored; without it the caller records that the
// interior of Res is zero and the next residual starts from zero instead of
// reading it
template <bool DIAG = false, bool ZERO_R = true>
__global__ void __launch_bounds__(MRM_BLOCK) k_source_integrate(Geom G, SrcParams SP, double* __restrict__ U,
                                                                const double* __restrict__ U0, double* __restrict__ R,
                                                                const double* __restrict__ Qtot, DiagArgs DA) {
  const long bs = G.bs;
  int i, j, k;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (!interior_ijk(G, t, i, j, k)) return;
  long node = i + j * G.s1 + k * G.s2;
  double up[MRM_NV], r[MRM_NV], u0[MRM_NV], out[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    up[q] = U[q * bs + node];
    r[q] = R[q * bs + node];
    u0[q] = U0[q * bs + node];
  }
  double Qrad = Qtot ? Qtot[node] : 0.0;
  double dg[4];
  source_integrate_cell<DIAG>(SP, up, r, u0, Qrad, SP.dmp ? SP.dmp[i] : 0.0, out, dg);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    U[q * bs + node] = out[q];
    if (ZERO_R) R[q * bs + node] = 0.0;
  }
  if (DIAG) {
    DA.dg[DG_T4 * bs + node] = dg[0];
    DA.dg[DG_T5 * bs + node] = dg[1];
    DA.dg[DG_T6 * bs + node] = dg[2];
    DA.dg[DG_T7 * bs + node] += (out[4] - dg[3]) * DA.idt;
  }
}

// Res := 0 over the interior (the store k_source_integrate<..., false> left out)
__global__ void __launch_bounds__(MRM_BLOCK) k_res_zero(Geom G, double* __restrict__ R) {
  int i, j, k;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (!interior_ijk(G, t, i, j, k)) return;
  long node = i + j * G.s1 + k * G.s2;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) R[q * G.bs + node] = 0.0;
}

__global__ void __launch_bounds__(MRM_BLOCK) k_save_cons(Geom G, const double* __restrict__ U, double* __restrict__ U0) {
  int i, j, k;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (!interior_ijk(G, t, i, j, k)) return;
  long node = i + j * G.s1 + k * G.s2;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) U0[q * G.bs + node] = U[q * G.bs + node];
}

// with DA.dg, tvar7 += the clamp's change of e / dt (Add_Sources_Integrate_SR.C:618-620)
__global__ void __launch_bounds__(MRM_BLOCK) k_tcheck(Geom G, double rho_min, double eps_min, double eps_max,
                                                      double vmax, double* __restrict__ U, DiagArgs DA) {
  int i, j, k;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (!interior_ijk(G, t, i, j, k)) return;
  long node = i + j * G.s1 + k * G.s2;
  double u[MRM_NV];
#pragma unroll
  for (int q = 0; q < 5; q++) u[q] = U[q * G.bs + node];
  const double eei = u[4];
  tcheck_vals(u, rho_min, eps_min, eps_max, vmax);
#pragma unroll
  for (int q = 0; q < 5; q++) U[q * G.bs + node] = u[q];
  if (DA.dg) DA.dg[DG_T7 * G.bs + node] += (u[4] - eei) * DA.idt;
}

// ---------------------------------------------------------------- exchange (periodic, one rank)
// dim 1 ghosts at interior (i,k); dim 2 ghosts at interior i and all j
template <int DIM>
__global__ void __launch_bounds__(MRM_BLOCK) k_exchange(Geom G, double* __restrict__ U) {
  const int nx = G.le0 - G.lb0 + 1;
  const int gh = DIM == 1 ? G.g1 : G.g2;
  const int lb = DIM == 1 ? G.lb1 : G.lb2, le = DIM == 1 ? G.le1 : G.le2;
  const int nb = DIM == 1 ? (G.le2 - G.lb2 + 1) : G.n1;     // extent of the other transverse dim
  const int ob = DIM == 1 ? G.lb2 : 0;
  const long sd = DIM == 1 ? G.s1 : G.s2, so = DIM == 1 ? G.s2 : G.s1;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nx * nb * gh) return;
  int i = G.lb0 + (int)(t % nx);
  long r = t / nx;
  int b = ob + (int)(r % nb);
  int igh = (int)(r / nb);
  long base = i + b * so;
  long src_r = base + (long)(lb + igh) * sd, dst_r = base + (long)(le + 1 + igh) * sd;
  long src_l = base + (long)(le - gh + 1 + igh) * sd, dst_l = base + (long)(lb - gh + igh) * sd;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    U[q * G.bs + dst_r] = U[q * G.bs + src_r];
    U[q * G.bs + dst_l] = U[q * G.bs + src_l];
  }
}

// ---------------------------------------------------------------- boundary conditions
// per-block partial sums over the interior (j,k) of the two bottom layers:
// mode 0: p_interp of layers lbeg, lbeg+1 (boundary_pdmp_1_fftw3.C:172-202)
// mode 1: set_p_s_bc's five sums (boundary_pdmp_1_fftw3.C:516-536)
template <int MODE>
__global__ void __launch_bounds__(MRM_BLOCK) k_bc_sums(Geom G, EosDev E, const double* __restrict__ U,
                                                       double* __restrict__ part) {
  constexpr int NS = MODE == 0 ? 2 : 5;
  __shared__ double lds[NS * MRM_BLOCK / 64];
  const long bs = G.bs;
  const int ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  double s[NS];
#pragma unroll
  for (int q = 0; q < NS; q++) s[q] = 0.0;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < (long)ny * nz) {
    int j = G.lb1 + (int)(t % ny), k = G.lb2 + (int)(t / ny);
    long n1 = G.lb0 + j * G.s1 + k * G.s2, n2 = n1 + 1;
    double d1 = U[n1], d2 = U[n2];
    double m1x = U[bs + n1], m1y = U[2 * bs + n1], m1z = U[3 * bs + n1];
    double m2x = U[bs + n2], m2y = U[2 * bs + n2], m2z = U[3 * bs + n2];
    if (MODE == 0) {
      double e1 = U[4 * bs + n1] - 0.5 * ((m1x * m1x + m1y * m1y + m1z * m1z) / d1);
      double e2 = U[4 * bs + n2] - 0.5 * ((m2x * m2x + m2y * m2y + m2z * m2z) / d2);
      s[0] = p_interp(E, e1 / d1, d1);
      s[1] = p_interp(E, e2 / d2, d2);
    } else {
      double e1 = U[4 * bs + n1] - 0.5 * (m1x * m1x + m1y * m1y + m1z * m1z) / d1;
      double e2 = U[4 * bs + n2] - 0.5 * (m2x * m2x + m2y * m2y + m2z * m2z) / d2;
      double eps1 = e1 / d1, eps2 = e2 / d2;
      s[0] = p_interp(E, eps1, d1);
      s[1] = p_interp(E, eps2, d2);
      double si = s_interp(E, eps1, d1);
      if (m1x > 0.0) { s[2] = 1.0; s[3] = si; }
      s[NS - 1] = si;
    }
  }
  block_reduce<OpAdd, NS>(s, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NS; q++) part[q * gridDim.x + blockIdx.x] = s[q];
  }
}

// fixed-order sum of per-block partials: out[q] = sum_b part[q*nb + b]
__global__ void __launch_bounds__(MRM_BLOCK) k_sum_partials(const double* __restrict__ part, int nb, int nq,
                                                            double* __restrict__ out) {
  __shared__ double lds[MRM_BLOCK];
  for (int q = 0; q < nq; q++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[q * nb + b];
    lds[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
      double x = 0.0;
      for (int w = 0; w < (int)blockDim.x; w++) x += lds[w];
      out[q] = x;
    }
    __syncthreads();
  }
}


// bottom ghost cells of one (j,k) column (boundary_pdmp_1_fftw3.C:245-318):
// in u1 = U(lbeg), u2 = U(lbeg+1); out u1 = U(lbeg-1), u2 = U(lbeg-2)
MRM_HD void bc_bottom_cells(const EosDev& E, const BcParams& B, double pmean1, double pmean2,
                            double u1[MRM_NV], double u2[MRM_NV]) {
  const double p1_dmp = 0.8;
  u1[4] -= 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3]) / u1[0];
  u2[4] -= 0.5 * (u2[1] * u2[1] + u2[2] * u2[2] + u2[3] * u2[3]) / u2[0];
  double eps1 = u1[4] / u1[0];
  double eps2 = u2[4] / u2[0];
  double pmag = 0.5 * u1[5] * u1[5];
  double pmin1 = p_interp(E, eps1, u1[0]) - pmean1;
  double smin1 = s_interp(E, eps1, u1[0]);
  double smin2 = s_interp(E, eps2, u2[0]);
  double s1, s2;
  if (u1[1] >= 0.e0) { s1 = B.s_bc; s2 = B.s_bc; }
  else { s1 = smin1; s2 = smin2; }
  double c = B.p_bc / sqrt(pmean1 * pmean2);
  double p1 = pmean1 * c;
  double p2 = pmean1 * c * c;
  p1 += pmin1 * p1_dmp;
  p2 += pmin1 * p1_dmp * p1_dmp;
  p1 -= pmag * (1.0 - p1_dmp);
  p2 -= pmag * (1.0 - p1_dmp * p1_dmp);
  c = u1[5] / B.B_crit;
  c = pow(c, 4.0);
  c = (1.0 - c) / (1.0 + c);
  u1[1] *= c; u1[2] *= c; u1[3] *= c;
  u2[1] *= c; u2[2] *= c; u2[3] *= c;
  double rho1 = d3_interp(E, p1, s1);
  double rho2 = d3_interp(E, p2, s2);
  eps1 = eps3_interp(E, p1, s1);
  eps2 = eps3_interp(E, p2, s2);
  u1[0] = rho1;
  u2[0] = rho2;
  u1[4] = eps1 * rho1 + 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3]) / u1[0];
  u2[4] = eps2 * rho2 + 0.5 * (u2[1] * u2[1] + u2[2] * u2[2] + u2[3] * u2[3]) / u2[0];
}

// top ghost cells of one column, vertical field (boundary_pdmp_1_fftw3.C:402-474,
// bnd_pot = 0): in u1 = U(lend), u2 = U(lend-1); out u1 = U(lend+1), u2 = U(lend+2)
MRM_HD void bc_top_cells(const BcParams& B, double u1[MRM_NV], double u2[MRM_NV]) {
  if (B.open_top == 1) {
    if (u1[1] < 0.0) { u1[1] *= -1; u2[1] *= -1; }
  } else {
    u1[1] *= -1; u2[1] *= -1;
  }
  if (B.eps_top > 0.0) {
    u1[4] = B.eps_top * u1[0] + 0.5 * (u1[1] * u1[1] + u1[2] * u1[2] + u1[3] * u1[3]) / u1[0];
    u2[4] = B.eps_top * u2[0] + 0.5 * (u2[1] * u2[1] + u2[2] * u2[2] + u2[3] * u2[3]) / u2[0];
  }
  u1[6] *= -1; u2[6] *= -1; u1[7] *= -1; u2[7] *= -1;
}

// potential-field top boundary (boundary_pdmp_1_fftw3.C:422-436), after
// bc_top_cells: B_y, B_z of the top interior cell (*by, *bz) and the whole B
// of both ghost cells from b_ext; il = j + k n1 is the column's offset in
// each b_ext plane
MRM_HD void bc_top_potential(const BcParams& B, long il, double* by, double* bz, double u1[MRM_NV],
                             double u2[MRM_NV]) {
  const double* b = B.bext + il;
  const long p = B.bplane;
  *by = b[2 * p];
  *bz = b[5 * p];
  u1[5] = b[0];
  u1[6] = b[3 * p];
  u1[7] = b[6 * p];
  u2[5] = b[p];
  u2[6] = b[4 * p];
  u2[7] = b[7 * p];
}

// bottom ghost layers, all (j,k) incl. ghosts; psum[0..1] are the layer sums
// of p written by k_sum_partials
__global__ void __launch_bounds__(MRM_BLOCK) k_bc_bottom(Geom G, EosDev E, BcParams B, const double* __restrict__ psum,
                                                         double* __restrict__ U) {
  const long bs = G.bs;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)G.n1 * G.n2) return;
  const double pmean1 = psum[0] * B.cc, pmean2 = psum[1] * B.cc;
  int j = (int)(t % G.n1), k = (int)(t / G.n1);
  long base = j * G.s1 + k * G.s2;
  long n1 = G.lb0 + base, n2 = n1 + 1;
  double u1[MRM_NV], u2[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) { u1[q] = U[q * bs + n1]; u2[q] = U[q * bs + n2]; }
  bc_bottom_cells(E, B, pmean1, pmean2, u1, u2);
  long g1 = G.lb0 - 1 + base, g2 = G.lb0 - 2 + base;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) { U[q * bs + g1] = u1[q]; U[q * bs + g2] = u2[q]; }
}

// top ghost layers
__global__ void __launch_bounds__(MRM_BLOCK) k_bc_top(Geom G, BcParams B, double* __restrict__ U) {
  const long bs = G.bs;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)G.n1 * G.n2) return;
  int j = (int)(t % G.n1), k = (int)(t / G.n1);
  long base = j * G.s1 + k * G.s2;
  long n1 = G.le0 + base, n2 = n1 - 1;
  double u1[MRM_NV], u2[MRM_NV];
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) { u1[q] = U[q * bs + n1]; u2[q] = U[q * bs + n2]; }
  bc_top_cells(B, u1, u2);
  if (B.bext) bc_top_potential(B, t, &U[6 * bs + n1], &U[7 * bs + n1], u1, u2);
  long g1 = G.le0 + 1 + base, g2 = G.le0 + 2 + base;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) { U[q * bs + g1] = u1[q]; U[q * bs + g2] = u2[q]; }
}

// ---------------------------------------------------------------- damping
// get_damping (Add_Sources_Integrate_SR.C:628-761) on one rank.  Layer sums
// of rho v_x / hsize in two passes: k_damp_partials gives each wave 64
// consecutive layers (coalesced in i) and one of DMP_SETS column sets
// (columns set, set + DMP_SETS, ...), k_damp_finish adds the sets of a layer
// in a fixed order and forms dmp[i] = sum hh, hh = min(1/tau_min,
// (1/tau_ref) (sum / (vel_ref rho_max))^4).  PRIM: U holds velocities (after
// ConsToPrim), else momenta, whose velocity is M (1/rho) as ConsToPrim forms
// it.  FIRST also reduces max rho into *rmax (the first call's rho_max).
constexpr int DMP_SETS = 1024;

template <bool PRIM, bool FIRST>
__global__ void __launch_bounds__(256) k_damp_partials(Geom G, const double* __restrict__ U,
                                                       double* __restrict__ part, unsigned long long* rmax) {
  const int nx = G.le0 - G.lb0 + 1, ny = G.le1 - G.lb1 + 1, nz = G.le2 - G.lb2 + 1;
  const int lane = threadIdx.x & 63;
  const int set = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int ii = blockIdx.x * 64 + lane;
  const long ncol = (long)ny * nz;
  const double hsize = (double)ncol;
  double acc = 0.0, rm = 0.0;
  if (ii < nx) {
    const long i = G.lb0 + ii;
    for (long c = set; c < ncol; c += DMP_SETS) {
      const long node = i + (G.lb1 + c % ny) * G.s1 + (G.lb2 + c / ny) * G.s2;
      const double d = U[node];
      double v = U[G.bs + node];
      if (!PRIM) v = v * (1.0 / d);
      acc += v * d / hsize;
      if (FIRST) rm = fmax(rm, d);
    }
    part[(long)set * G.n0 + i] = acc;
  }
  if (FIRST) {
    for (int o = 32; o > 0; o >>= 1) rm = fmax(rm, __shfl_xor(rm, o, 64));
    if (lane == 0) atomic_max_nonneg(rmax, rm);
  }
}

__global__ void __launch_bounds__(256) k_damp_finish(Geom G, const double* __restrict__ part,
                                                     const unsigned long long* __restrict__ rmax, double tau_ref,
                                                     double vel_ref, double tau_min, double* __restrict__ dmp) {
  const int ii = blockIdx.x * blockDim.x + threadIdx.x;
  if (ii > G.le0 - G.lb0) return;
  const long i = G.lb0 + ii;
  double sum = 0.0;
  for (int set = 0; set < DMP_SETS; set++) sum += part[(long)set * G.n0 + i];
  const double rho_max = __longlong_as_double((long long)*rmax);
  double hh = tau_ref * pow(sum / (vel_ref * rho_max), 4.0);
  hh = fmin(tau_min, hh);
  dmp[i] = sum * hh;
}

// ---------------------------------------------------------------- TVD sweeps
// Line sets (tvdlimit_SR.C:199-205, 671-672): sweep 0 covers every (j,k)
// incl. ghosts; sweep 1 interior i and all k; sweep 2 interior i and j.
// With EPI, the last sweep also applies TCheck to each finished cell
// (Add_Sources_Integrate_SR.C:591-624; every interior cell is finished by the
// dim-2 sweep), writes its EOS pressure and reduces its characteristic speed,
// which is what the next step's ConsToPrim and MHD_Residual compute from it.

__device__ double ghost_pres_(const EosDev& E, const double u[MRM_NV]) {
  double m1 = u[1], m2 = u[2], m3 = u[3], e = u[4], pr[5];
  cons_to_prim_cell<2>(E, u[0], m1, m2, m3, e, pr);
  return pr[0];
}

// VALF: the dim-2 epilogue also reduces Adjust_Valf_Max's statistics
// (va_adjust > 0); that variant is held to two waves per SIMD like the others.
// DIAG: tvar8, tvar6, Qres, tvar7 and Qvis of every updated point
// (tvdlimit_SR.C:577-666)
// The conserved values of the last three points of each line wait in an LDS
// ring ([slot][field][thread], 48 KB per block) for their update two points
// later, instead of a second read from memory.
// limiter branch slots of one thread added to the context's totals: one
// atomic per slot and wave (every lane of the wave must call it)
__device__ inline void tvd_cnt_flush(const TvdCnt& c, unsigned long long* out) {
  const bool lead = (threadIdx.x & 63) == 0;
#pragma unroll
  for (int i = 0; i < MRM_BR_NCOUNT; i++) {
    unsigned long long v = c.n[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lead && v) atomicAdd(out + i, v);
  }
#pragma unroll
  for (int i = 0; i < MRM_BK_N; i++) {
    unsigned long long v = c.h[i];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if (lead && v) atomicAdd(out + MRM_BR_NCOUNT + i, v);
    unsigned long long w = c.nnear[i];
    for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, 64);
    if (lead && w) atomicAdd(out + MRM_BR_NCOUNT + MRM_BK_N + i, w);
    unsigned long long x = c.nexact[i];
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    if (lead && x) atomicAdd(out + MRM_BR_NCOUNT + 2 * MRM_BK_N + i, x);
  }
}

// position key of node in sweep d (see TvdCnt)
__device__ __forceinline__ unsigned long long br_key(int d, long bs, long node) {
  return ((unsigned long long)d * (unsigned long long)bs + (unsigned long long)node) << 3;
}

// TCK: TCheck alone after each update (X's limits; no pressure, no cmax)
template <int D, bool HAS_TAU, bool EPI, bool VALF = false, bool DIAG = false, bool CNT = false, bool TCK = false>
__global__ void __launch_bounds__(MRM_BLOCK) __attribute__((amdgpu_waves_per_eu(VALF ? 2 : 1)))
k_tvd_sweep(Geom G, TvdParams P, double* __restrict__ U, const double* __restrict__ Pr, TvdEpilogue X,
            DiagArgs DA) {
  __shared__ double lds[(VALF ? 5 : 1) * MRM_BLOCK / 64];
  __shared__ double vacc[VALF ? 5 * MRM_BLOCK : 1];
  __shared__ double ring[3][MRM_NV][MRM_BLOCK];
  if (VALF) valf_lds_init(vacc + threadIdx.x, MRM_BLOCK);
  TvdCnt cnt;
  if (CNT) cnt.clear(P.btie);
  TvdCnt* const cn = CNT ? &cnt : nullptr;
  const long bs = G.bs;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int tl = threadIdx.x;
  long base = 0, str = 1;
  int sz = 0, gh = 0, lineX = 0;
  bool active;
  if (D == 0) {
    active = t < (long)G.n1 * G.n2;
    int j = (int)(t % G.n1), k = (int)(t / G.n1);
    base = j * G.s1 + k * G.s2; str = 1; sz = G.n0; gh = G.g0;
    if (CNT) cnt.on = active && j >= G.lb1 && j <= G.le1 && k >= G.lb2 && k <= G.le2;
  } else {
    const int nx = G.le0 - G.lb0 + 1;
    const int nb = D == 1 ? P.y_nk : (G.le1 - G.lb1 + 1);
    active = t < (long)nx * nb;
    int i = G.lb0 + (int)(t % nx);
    int b = (int)(t / nx);
    lineX = i;
    if (D == 1) { base = i + (long)(P.y_k0 + b) * G.s2; str = G.s1; sz = G.n1; gh = G.g1; }
    else { base = i + (G.lb1 + b) * G.s1; str = G.s2; sz = G.n2; gh = G.g2; }
    if (CNT) cnt.on = active && (D == 2 || (P.y_k0 + b >= G.lb2 && P.y_k0 + b <= G.le2));
  }
  double cm[1] = {0.0};
  if (active) {
    const double lhfb = D == 0 ? 0.0 : P.hfb[lineX];
    const double lhft = D == 0 ? 0.0 : P.hft[lineX];
    const double lqft = D == 0 ? 0.0 : P.qft[lineX];
    const int istart = 1 + (P.is_gbeg0 ? gh : 0);
    TvdPt W0, W1, W2, W3;      // points t-3, t-2, t-1, t
    double S0[MRM_NV], S1[MRM_NV], Fm[MRM_NV], F[MRM_NV];
    double qm = 0.0, qv = 0.0, qvm = 0.0;
    int slot = 0;    // ring slot of point s; point s-2 is in slot (s+1) % 3
    // PF: point s's conserved values and pressure are loaded during point
    // s-1's work (the last iteration loads point sz-1 again and does not use
    // it); the sweeps with the epilogue or the diagnostics have no registers
    // to spare for them
    constexpr bool PF = !EPI && !DIAG;
    double un[MRM_NV], pn = 0.0;
    if (PF) {
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) un[q] = U[q * bs + base];
      pn = Pr[base];
    }
    for (int s = 0; s < sz; s++) {
      long node = base + s * str;
      double u[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) {
        u[q] = PF ? un[q] : U[q * bs + node];
        ring[slot][q][tl] = u[q];
      }
      const double pr = PF ? pn : Pr[node];
      if (PF) {
        const long nn = base + (long)min(s + 1, sz - 1) * str;
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) un[q] = U[q * bs + nn];
        pn = Pr[nn];
      }
      const int slot2 = slot == 2 ? 0 : slot + 1;
      slot = slot2;
      double tau = HAS_TAU ? P.tau[node] : 0.0;
      double hfb1 = D == 0 ? P.hfb[s] : lhfb, hft1 = D == 0 ? P.hft[s] : lhft;
      if (CNT) cnt.key = br_key(D, bs, node);
      tvd_point<!EPI>(P, u, pr, tau, hfb1, hft1, D == 0, W3, cn);
      if (s >= 2) {
        const int ip = s - 1;   // slope at point s-1
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) S0[q] = S1[q];
        if (CNT) cnt.key = br_key(D, bs, base + ip * str);
        tvd_slope(P, W1, W2, W3, D == 0 ? P.hfb[ip] : lhfb, D == 0 ? P.hft[ip] : lhft, S1, cn);
      }
      if (s >= 3) {
        const int ii = s - 2;   // interface between ii and ii+1; update of point ii
        const double ih = D == 0 ? P.hfb[ii] : lhfb, it = D == 0 ? P.hft[ii] : lhft;
        const double iq = D == 0 ? P.qft[ii] : lqft;
        bool zb = (D == 0) && ((P.is_gend0 && ii == sz - gh - 1) || (P.is_gbeg0 && ii == gh - 1));
        if (CNT) cnt.key = br_key(D, bs, base + ii * str);
        double q = tvd_iface<D, DIAG>(P, W0, W1, W2, W3, S0, S1, ih, it, iq, ii >= istart, zb, F, &qv, cn);
        if (ii >= gh && ii <= sz - gh - 1) {
          long nd = base + ii * str;
          double uu[MRM_NV];
#pragma unroll
          for (int k = 0; k < MRM_NV; k++) uu[k] = ring[slot2][k][tl];
          double dg[4];
          tvd_update<DIAG>(P, W1, Fm, F, qm, q, iq, uu, dg, cn);
          if (DIAG) {
            DA.dg[DG_T8 * bs + nd] += dg[0] * DA.idt;
            DA.dg[DG_T6 * bs + nd] += dg[1] * DA.idt;
            DA.dg[DG_QRES * bs + nd] += DA.dt_fac * dg[2];
            DA.dg[DG_T7 * bs + nd] += (uu[4] - dg[3]) * DA.idt;
            DA.dg[DG_QVIS * bs + nd] += DA.dt_fac * (0.5 * (qv + qvm) * iq);
          }
          if (TCK) tcheck_vals(uu, X.rho_min, X.eps_min, X.eps_max, X.vmax);
          if (EPI) {
            tcheck_vals(uu, X.rho_min, X.eps_min, X.eps_max, X.vmax);
            // without X.Pnew the epilogue is TCheck alone (the drop-in
            // TVDlimit's last sweep carried out by TCheck)
            if (X.Pnew) {
              const double pn = ghost_pres_(X.E, uu);
              if (VALF) {
                // v and e_int as ConsToPrim forms them (cons_to_prim.C:110-118)
                const double dn = 1.0 / uu[0];
                const double v1 = uu[1] * dn, v2 = uu[2] * dn, v3 = uu[3] * dn;
                const double ei = uu[4] - 0.5 * (v1 * v1 + v2 * v2 + v3 * v3) * uu[0];
                valf_lds_cell(vacc + threadIdx.x, MRM_BLOCK, uu[0], v1, v2, v3, ei, pn, X.vlim95, X.elim95);
              }
              X.Pnew[nd] = pn;
              if (!(X.skip_top_cmax && lineX == G.le0)) cm[0] = smax(cm[0], point_cmax_cons(uu, pn, X.inv_va2max));
            }
          }
#pragma unroll
          for (int k = 0; k < MRM_NV; k++) U[k * bs + nd] = uu[k];
        }
#pragma unroll
        for (int k = 0; k < MRM_NV; k++) Fm[k] = F[k];
        qm = q;
        qvm = qv;
      }
      W0 = W1; W1 = W2; W2 = W3;
    }
  }
  if (CNT) tvd_cnt_flush(cnt, P.bcount);
  if (EPI) {
    block_reduce<OpMax, 1>(cm, lds);
    if (threadIdx.x == 0 && X.Pnew) atomic_max_nonneg(X.cmax, cm[0]);
    if (VALF) {
      __syncthreads();
      ValfAcc va = valf_lds_load(vacc + threadIdx.x, MRM_BLOCK);
      valf_flush(va, lds, X.valf);
    }
  }
}

}  // namespace MRM_NS
