// mrm_fused.h -- kernels of the device-resident fused step.
//
// One RK stage is one kernel: the MHD residual of all three sweeps, the
// sources, the Boris correction, U = U0 + wdt*R, the Tcheck clamp and the EOS
// pressure of the new state, written to a second state buffer (the stage's
// input stays intact, so U0 needs no copy and Res never touches HBM).  The
// arithmetic per cell is the drop-in path's, term for term: the residual is
// accumulated in the reference's order ((0 + R_x) + R_y) + R_z
// (mhdres_SR.C:176, :456-466), primitives are formed with ConsToPrim's
// expressions (cons_to_prim.C:110-118).
//
// Tiling: a 32x8 (dim0 x dim1) column tile per block, marched along dim 2.
// The plane being updated sits in LDS with a two-cell ring in dims 0 and 1;
// dim 2 neighbours come from a five-plane register window per thread.
#pragma once
#include "mrm_kernels.h"

namespace MRM_NS {



// conservative point -> primitive (ConsToPrim's arithmetic, cons_to_prim.C:110-118)
__device__ __forceinline__ void load_prim(const double* __restrict__ U, const double* __restrict__ P, long bs,
                                          long node, double w[ST_NF]) {
  double d = U[node];
  double dn = 1.0 / d;
  double m1 = U[bs + node] * dn, m2 = U[2 * bs + node] * dn, m3 = U[3 * bs + node] * dn;
  double vv = (m1 * m1 + m2 * m2 + m3 * m3);
  w[0] = d;
  w[1] = m1;
  w[2] = m2;
  w[3] = m3;
  w[4] = U[4 * bs + node] - 0.5 * vv * d;
  w[5] = U[5 * bs + node];
  w[6] = U[6 * bs + node];
  w[7] = U[7 * bs + node];
  w[8] = P[node];
}

// the nine values load_prim reads, as stored (conservative U, p), so that a
// load can be issued ahead of its use; prim_of_raw then forms the primitives
// with load_prim's arithmetic
__device__ __forceinline__ void load_raw(const double* __restrict__ U, const double* __restrict__ P, long bs,
                                         unsigned node, double r[ST_NF]) {
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) r[q] = ld32(U + q * bs, node);
  r[8] = ld32(P, node);
}
__device__ __forceinline__ void prim_of_raw(const double r[ST_NF], double w[ST_NF]) {
  double d = r[0];
  double dn = 1.0 / d;
  double m1 = r[1] * dn, m2 = r[2] * dn, m3 = r[3] * dn;
  double vv = (m1 * m1 + m2 * m2 + m3 * m3);
  w[0] = d;
  w[1] = m1;
  w[2] = m2;
  w[3] = m3;
  w[4] = r[4] - 0.5 * vv * d;
  w[5] = r[5];
  w[6] = r[6];
  w[7] = r[7];
  w[8] = r[8];
}

// map block index to a tile so that consecutive tiles share an XCD (blocks are
// dealt to the 8 XCDs round robin)
__device__ __forceinline__ long xcd_remap(long b, long nb) {
  long q = nb / 8;
  if (b < 8 * q) return (b % 8) * q + b / 8;
  return b;
}

// ZWIN = true: dim-2 neighbours from a register window (one wave per SIMD);
// ZWIN = false: dim-2 neighbours re-read through the caches each plane.
// ETA: resistive terms (param_eta > 0) in the three sweeps and eta |J|^2 in
// the energy residual (mhdres_SR.C:422-462).  ERES: A.eres holds a further
// energy residual per cell (Spitzer conduction's), added after the sweeps.
// With A.write_p the pressure of plane k is formed during plane k+1's work:
// its EOS reads are issued before that plane's sweeps and consumed after
// them, so their latency hides behind the sweeps; the chunk's last plane is
// finished after the loop.
template <bool ZWIN, bool ETA = false, bool ERES = false>
__global__ void __launch_bounds__(ST_SX* ST_SY, ZWIN ? 1 : 2)
k_stage(StageArgs A, const double* __restrict__ Uin, const double* __restrict__ Pin, const double* __restrict__ U0,
        double* __restrict__ Uout, double* __restrict__ Pout, const double* __restrict__ Qtot) {
  __shared__ double lds[ST_NF][ST_LY][ST_LX];
  const Geom& G = A.G;
  const long bs = G.bs;
  const long nblocks = (long)A.ntx * A.nty * A.nkc;
  long lb = xcd_remap(blockIdx.x, nblocks);
  const int tx_t = (int)(lb % A.ntx);
  lb /= A.ntx;
  const int ty_t = (int)(lb % A.nty);
  const int kc = (int)(lb / A.nty);
  const int tx = threadIdx.x % ST_SX, ty = threadIdx.x / ST_SX;
  const int i0 = G.lb0 + tx_t * ST_SX, j0 = G.lb1 + ty_t * ST_SY;
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = (i <= G.le0) && (j <= G.le1);
  const int kbeg = G.lb2 + kc * ST_KC;
  const int kend = min(kbeg + ST_KC, G.le2 + 1);     // exclusive
  // halo assignment: tx<2 / tx>=SX-2 load the dim-0 ring of their row,
  // ty<4 load one point of the dim-1 ring rows
  int hx_c = -1, hx_i = 0;
  if (tx < 2) { hx_c = tx; hx_i = i0 - 2 + tx; }
  else if (tx >= ST_SX - 2) { hx_c = tx + 4; hx_i = i0 + tx + 2; }
  int hy_r = -1, hy_j = 0;
  if (ty < 4) {
    hy_r = ty < 2 ? ty : ST_SY + ty;          // rows 0,1 and SY+2, SY+3
    hy_j = ty < 2 ? j0 - 2 + ty : j0 + ST_SY + (ty - 2);
  }
  const bool hx_ok = hx_c >= 0 && hx_i < G.n0 && j < G.n1;
  const bool hy_ok = hy_r >= 0 && hy_j < G.n1 && i < G.n0;
  const bool own_ok = i < G.n0 && j < G.n1;
  const unsigned col = (unsigned)i + (unsigned)j * (unsigned)G.s1;
  const unsigned s1u = (unsigned)G.s1, s2u = (unsigned)G.s2;

  // the cell whose pressure is pending (plane k - 1 of this column); the
  // initial values are a harmless placeholder lookup that is never stored
  double pd_u[5] = {1.0, 0.0, 0.0, 0.0, 1.0};
  unsigned pd_node = 0;
  double w[5][ST_NF];   // planes k-2 .. k+2 of this thread's column (w[2] is the centre)
  // with the register window, the next plane's raw values (own column k+3,
  // the x and y halo of k+1) are loaded during the current plane's work
  double rz[ST_NF], rhx[ST_NF], rhy[ST_NF];
  if (ZWIN) {
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int kk = kbeg - 2 + q;
      if (own_ok) { double r[ST_NF]; load_raw(Uin, Pin, bs, col + (unsigned)kk * s2u, r); prim_of_raw(r, w[q + 1]); }
      else {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) w[q + 1][f] = 1.0;
      }
    }
    // kbeg + 2 <= lend2 + 2 = n2 - 1
    if (own_ok) load_raw(Uin, Pin, bs, col + (unsigned)(kbeg + 2) * s2u, rz);
    if (hx_ok) load_raw(Uin, Pin, bs, (unsigned)hx_i + (unsigned)j * s1u + (unsigned)kbeg * s2u, rhx);
    if (hy_ok) load_raw(Uin, Pin, bs, (unsigned)i + (unsigned)hy_j * s1u + (unsigned)kbeg * s2u, rhy);
  }
  for (int k = kbeg; k < kend; k++) {
    if (ZWIN) {
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int f = 0; f < ST_NF; f++) w[q][f] = w[q + 1][f];
      if (own_ok) prim_of_raw(rz, w[4]);     // plane k + 2, loaded during plane k - 1
    } else {
      if (own_ok) { double r[ST_NF]; load_raw(Uin, Pin, bs, col + (unsigned)k * s2u, r); prim_of_raw(r, w[2]); }
      else {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) w[2][f] = 1.0;
      }
    }
    // plane k into LDS
#pragma unroll
    for (int f = 0; f < ST_NF; f++) lds[f][ty + 2][tx + 2] = w[2][f];
    if (hx_ok) {
      double h[ST_NF];
      if (ZWIN) prim_of_raw(rhx, h);
      else { double r[ST_NF]; load_raw(Uin, Pin, bs, (unsigned)hx_i + (unsigned)j * s1u + (unsigned)k * s2u, r); prim_of_raw(r, h); }
#pragma unroll
      for (int f = 0; f < ST_NF; f++) lds[f][ty + 2][hx_c] = h[f];
    }
    if (hy_ok) {
      double h[ST_NF];
      if (ZWIN) prim_of_raw(rhy, h);
      else { double r[ST_NF]; load_raw(Uin, Pin, bs, (unsigned)i + (unsigned)hy_j * s1u + (unsigned)k * s2u, r); prim_of_raw(r, h); }
#pragma unroll
      for (int f = 0; f < ST_NF; f++) lds[f][hy_r][tx + 2] = h[f];
    }
    if (ZWIN && k + 1 < kend) {
      // k + 3 <= kend + 1 <= lend2 + 2 = n2 - 1: inside the buffer
      if (own_ok) load_raw(Uin, Pin, bs, col + (unsigned)(k + 3) * s2u, rz);
      if (hx_ok) load_raw(Uin, Pin, bs, (unsigned)hx_i + (unsigned)j * s1u + (unsigned)(k + 1) * s2u, rhx);
      if (hy_ok) load_raw(Uin, Pin, bs, (unsigned)i + (unsigned)hy_j * s1u + (unsigned)(k + 1) * s2u, rhy);
    }
    __syncthreads();
    if (valid) {
      EosLook<2> pg;
      if (A.write_p) eos_issue<2>(A.E, pd_u[0], pd_u[1], pd_u[2], pd_u[3], pd_u[4], pg);
      double acc[MRM_NV], res[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] = 0.0;
      const double lfc = lfac_of(w[2], A.RP.inv_va2max);
      auto getx = [&](int s, double o[9]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = lds[f][ty + 2][tx + s];
      };
      double J[3] = {0.0, 0.0, 0.0};
      sweep_residual_lr<0, ETA>(getx, w[2], A.RP.w1[0], A.RP.w2[0], lfc, res, &A.RP, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
      auto gety = [&](int s, double o[9]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = lds[f][ty + s][tx + 2];
      };
      sweep_residual_lr<1, ETA>(gety, w[2], A.RP.w1[1], A.RP.w2[1], lfc, res, &A.RP, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
      auto getz = [&](int s, double o[9]) {
        if (ZWIN) {
#pragma unroll
          for (int f = 0; f < ST_NF; f++) o[f] = w[s][f];
        } else {
          double r[ST_NF];
          load_raw(Uin, Pin, bs, col + (unsigned)(k + s - 2) * s2u, r);
          prim_of_raw(r, o);
        }
      };
      sweep_residual_lr<2, ETA>(getz, w[2], A.RP.w1[2], A.RP.w2[2], lfc, res, &A.RP, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
      if (ETA) acc[4] += A.RP.eta * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
      // sources, Boris, RK update, Tcheck
      const unsigned node = col + (unsigned)k * s2u;
      if (ERES) acc[4] += ld32(A.eres, node);
      double u0[MRM_NV], out[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) u0[q] = ld32(U0 + q * bs, node);
      double Qrad = Qtot ? ld32(Qtot, node) : 0.0;
      const double dmpi = A.SP.dmp ? A.SP.dmp[i] : 0.0;
      source_integrate_cell(A.SP, w[2], acc, u0, Qrad, dmpi, out);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) st32(Uout + q * bs, node, out[q]);
      if (A.write_p) {
        double pr[5];
        eos_finish<2>(A.E, pg, pr);
        if (k > kbeg) st32(Pout, pd_node, pr[0]);
#pragma unroll
        for (int q = 0; q < 5; q++) pd_u[q] = out[q];
        pd_node = node;
      }
    }
    __syncthreads();
  }
  if (A.write_p && valid && kend > kbeg) {
    double pr[5];
    cons_to_prim_cell<2>(A.E, pd_u[0], pd_u[1], pd_u[2], pd_u[3], pd_u[4], pr);
    st32(Pout, pd_node, pr[0]);
  }
}

// ------------------------------------------------------------------ stage kernel, fluxes shared
// k_stage with each x and y point flux formed once per plane: every thread
// packs the fluxes of its own point along dims 0 and 1 into LDS, the ring
// points' fluxes are packed by one thread each (threads 0..127 the four dim-1
// ring rows, 128..159 the dim-0 ring columns), and the x and y sweeps take
// their four neighbours' fluxes from LDS instead of forming them again (each
// residual term is formed as in k_stage, so the result is bit for bit
// k_stage's).  Dim-2 neighbours come from the five-plane register window of
// primitives, as in k_stage<true>.  Per plane and wave this is 4.5 point-flux
// evaluations instead of 12: box512 stage 5.78 -> 5.33 ms (exact build).
// One block per CU: the register window and the packed fluxes need more than
// the 256 registers of two waves per SIMD (forcing two spills to scratch and
// runs four times slower on box512).
template <bool ETA = false, bool ERES = false>
__global__ void __launch_bounds__(ST_SX* ST_SY, 1)
k_stage_f(StageArgs A, const double* __restrict__ Uin, const double* __restrict__ Pin, const double* __restrict__ U0,
          double* __restrict__ Uout, double* __restrict__ Pout, const double* __restrict__ Qtot) {
  constexpr int NFP = nfp_<ETA>();
  static_assert(4 * ST_SX + 4 * ST_SY <= ST_SX * ST_SY, "one ring point per thread at most");
  __shared__ double fx[NFP][ST_SY][ST_LX];   // dim-0 fluxes of row ty, columns i0-2 .. i0+SX+1
  __shared__ double fy[NFP][ST_LY][ST_SX];   // dim-1 fluxes of column tx, rows j0-2 .. j0+SY+1
  const Geom& G = A.G;
  const long bs = G.bs;
  const long nblocks = (long)A.ntx * A.nty * A.nkc;
  long lb = xcd_remap(blockIdx.x, nblocks);
  const int tx_t = (int)(lb % A.ntx);
  lb /= A.ntx;
  const int ty_t = (int)(lb % A.nty);
  const int kc = (int)(lb / A.nty);
  const int t = threadIdx.x;
  const int tx = t % ST_SX, ty = t / ST_SX;
  const int i0 = G.lb0 + tx_t * ST_SX, j0 = G.lb1 + ty_t * ST_SY;
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = (i <= G.le0) && (j <= G.le1);
  const int kbeg = G.lb2 + kc * ST_KC;
  const int kend = min(kbeg + ST_KC, G.le2 + 1);     // exclusive
  // the one ring point of this thread: threads 0 .. 4*SX-1 the dim-1 ring
  // rows (0, 1, SY+2, SY+3), threads 4*SX .. 4*SX+4*SY-1 the dim-0 ring
  // columns (0, 1, SX+2, SX+3) of row (t - 4*SX) / 4
  int ring = 0;                       // 1: dim-1 ring point, 2: dim-0 ring point
  int r_row = 0, r_col = 0, r_i = 0, r_j = 0;
  if (t < 4 * ST_SX) {
    const int rr = t / ST_SX;
    r_row = rr < 2 ? rr : ST_SY + rr;
    r_col = t % ST_SX;
    r_i = i0 + r_col;
    r_j = j0 - 2 + r_row;
    ring = 1;
  } else if (t < 4 * ST_SX + 4 * ST_SY) {
    const int u = t - 4 * ST_SX;
    r_row = u / 4;
    const int c4 = u % 4;
    r_col = c4 < 2 ? c4 : ST_SX + c4;
    r_i = i0 - 2 + r_col;
    r_j = j0 + r_row;
    ring = 2;
  }
  const bool ring_ok = ring != 0 && r_i < G.n0 && r_j < G.n1;
  const bool own_ok = i < G.n0 && j < G.n1;
  const unsigned col = (unsigned)i + (unsigned)j * (unsigned)G.s1;
  const unsigned rcol = (unsigned)r_i + (unsigned)r_j * (unsigned)G.s1;
  const unsigned s2u = (unsigned)G.s2;

  double pd_u[5] = {1.0, 0.0, 0.0, 0.0, 1.0};
  unsigned pd_node = 0;
  double w[5][ST_NF];   // planes k-2 .. k+2 of this thread's column
  double rz[ST_NF], rr[ST_NF];   // raw values of the own column's plane k+3 and of the ring point's plane k+1
#pragma unroll
  for (int q = 0; q < 4; q++) {
    int kk = kbeg - 2 + q;
    if (own_ok) { double r[ST_NF]; load_raw(Uin, Pin, bs, col + (unsigned)kk * s2u, r); prim_of_raw(r, w[q + 1]); }
    else {
#pragma unroll
      for (int f = 0; f < ST_NF; f++) w[q + 1][f] = 1.0;
    }
  }
  if (own_ok) load_raw(Uin, Pin, bs, col + (unsigned)(kbeg + 2) * s2u, rz);
  if (ring_ok) load_raw(Uin, Pin, bs, rcol + (unsigned)kbeg * s2u, rr);
  for (int k = kbeg; k < kend; k++) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int f = 0; f < ST_NF; f++) w[q][f] = w[q + 1][f];
    if (own_ok) prim_of_raw(rz, w[4]);     // plane k + 2, loaded during plane k - 1
    {
      double o[14];
      flux_pack<0, ETA>(w[2], o);
#pragma unroll
      for (int f = 0; f < NFP; f++) fx[f][ty][tx + 2] = o[f];
      flux_pack<1, ETA>(w[2], o);
#pragma unroll
      for (int f = 0; f < NFP; f++) fy[f][ty + 2][tx] = o[f];
    }
    if (ring == 1) {
      double h[ST_NF], o[14];
      if (ring_ok) prim_of_raw(rr, h);
      else {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) h[f] = 1.0;
      }
      flux_pack<1, ETA>(h, o);
#pragma unroll
      for (int f = 0; f < NFP; f++) fy[f][r_row][r_col] = o[f];
    } else if (ring == 2) {
      double h[ST_NF], o[14];
      if (ring_ok) prim_of_raw(rr, h);
      else {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) h[f] = 1.0;
      }
      flux_pack<0, ETA>(h, o);
#pragma unroll
      for (int f = 0; f < NFP; f++) fx[f][r_row][r_col] = o[f];
    }
    if (k + 1 < kend) {
      // k + 3 <= kend + 1 <= lend2 + 2 = n2 - 1: inside the buffer
      if (own_ok) load_raw(Uin, Pin, bs, col + (unsigned)(k + 3) * s2u, rz);
      if (ring_ok) load_raw(Uin, Pin, bs, rcol + (unsigned)(k + 1) * s2u, rr);
    }
    __syncthreads();
    if (valid) {
      EosLook<2> pg;
      if (A.write_p) eos_issue<2>(A.E, pd_u[0], pd_u[1], pd_u[2], pd_u[3], pd_u[4], pg);
      double acc[MRM_NV], res[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] = 0.0;
      const double lfc = lfac_of(w[2], A.RP.inv_va2max);
      double J[3] = {0.0, 0.0, 0.0};
      auto getx = [&](int s, int f) { return fx[f][ty][tx + s]; };
      sweep_residual_packed<0, ETA>(getx, w[2], A.RP.w1[0], A.RP.w2[0], lfc, res, &A.RP, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
      auto gety = [&](int s, int f) { return fy[f][ty + s][tx]; };
      sweep_residual_packed<1, ETA>(gety, w[2], A.RP.w1[1], A.RP.w2[1], lfc, res, &A.RP, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
      auto getz = [&](int s, double o[9]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = w[s][f];
      };
      sweep_residual_lr<2, ETA>(getz, w[2], A.RP.w1[2], A.RP.w2[2], lfc, res, &A.RP, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
      if (ETA) acc[4] += A.RP.eta * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
      const unsigned node = col + (unsigned)k * s2u;
      if (ERES) acc[4] += ld32(A.eres, node);
      double u0[MRM_NV], out[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) u0[q] = ld32(U0 + q * bs, node);
      double Qrad = Qtot ? ld32(Qtot, node) : 0.0;
      const double dmpi = A.SP.dmp ? A.SP.dmp[i] : 0.0;
      source_integrate_cell(A.SP, w[2], acc, u0, Qrad, dmpi, out);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) st32(Uout + q * bs, node, out[q]);
      if (A.write_p) {
        double pr[5];
        eos_finish<2>(A.E, pg, pr);
        if (k > kbeg) st32(Pout, pd_node, pr[0]);
#pragma unroll
        for (int q = 0; q < 5; q++) pd_u[q] = out[q];
        pd_node = node;
      }
    }
    __syncthreads();
  }
  if (A.write_p && valid && kend > kbeg) {
    double pr[5];
    cons_to_prim_cell<2>(A.E, pd_u[0], pd_u[1], pd_u[2], pd_u[3], pd_u[4], pr);
    st32(Pout, pd_node, pr[0]);
  }
}

// ------------------------------------------------------------------ stage kernel, two wave groups
// k_stage_f's work split over two groups of four waves on the same tile, so
// that each group needs about half of the registers and the block runs two
// waves per SIMD (k_stage_f: one):
//   xy group (threads 0 .. 255), plane kbeg + m in iteration m: packs the
//     dim-0 and dim-1 point fluxes of its point and ring point into LDS and
//     forms the x and y residual terms (and their part of J), which it hands
//     to the z group in LDS; it also forms the EOS pressure of the new state
//     of plane kbeg + m - 2, which the z group handed over;
//   z group (threads 256 .. 511), plane kbeg + n - 1 in iteration n: keeps
//     the five-plane dim-2 window, adds the dim-2 terms to the xy group's sum
//     and applies sources, Boris correction, RK update and Tcheck.
// Each iteration has two barriers; the handover buffers are double
// buffered by iteration parity.  The sums are formed in k_stage_f's order
// ((0 + R_x) + R_y) + R_z, so the result is bit for bit k_stage_f's.
// EG: a third group of four waves (threads 512 .. 767) forms the EOS
// pressure instead of the xy group, so the block runs three waves per SIMD.
template <bool ETA = false, bool ERES = false, bool EG = false>
__global__ void __launch_bounds__((EG ? 3 : 2) * ST_SX * ST_SY, 1)
k_stage_w(StageArgs A, const double* __restrict__ Uin, const double* __restrict__ Pin, const double* __restrict__ U0,
          double* __restrict__ Uout, double* __restrict__ Pout, const double* __restrict__ Qtot) {
  constexpr int NFP = nfp_<ETA>();
  constexpr int NT = ST_SX * ST_SY;
  // xy group -> z group: the x + y sums, the point's Alfven limiter factor
  // (the same value the z group would form from the same primitives), and J
  // with ETA
  constexpr int NA = ETA ? MRM_NV + 4 : MRM_NV + 1;
  static_assert(4 * ST_SX + 4 * ST_SY <= NT, "one ring point per thread at most");
  __shared__ double fx[NFP][ST_SY][ST_LX];   // dim-0 fluxes of row ty, columns i0-2 .. i0+SX+1
  __shared__ double fy[NFP][ST_LY][ST_SX];   // dim-1 fluxes of column tx, rows j0-2 .. j0+SY+1
  __shared__ double axy[2][NA][NT];          // xy group -> z group
  __shared__ double aout[2][5][NT];          // z group -> xy group: the new rho, M, e
  const Geom& G = A.G;
  const long bs = G.bs;
  const long nblocks = (long)A.ntx * A.nty * A.nkc;
  long lb = xcd_remap(blockIdx.x, nblocks);
  const int tx_t = (int)(lb % A.ntx);
  lb /= A.ntx;
  const int ty_t = (int)(lb % A.nty);
  const int kc = (int)(lb / A.nty);
  const int grp = threadIdx.x / NT;   // 0: xy group, 1: z group, 2: EOS group (EG)
  const bool zg = grp == 1;
  const int t = threadIdx.x - grp * NT;
  const int tx = t % ST_SX, ty = t / ST_SX;
  const int i0 = G.lb0 + tx_t * ST_SX, j0 = G.lb1 + ty_t * ST_SY;
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = (i <= G.le0) && (j <= G.le1);
  const int kbeg = G.lb2 + kc * ST_KC;
  const int kend = min(kbeg + ST_KC, G.le2 + 1);     // exclusive
  const int NK = kend - kbeg;
  const bool own_ok = i < G.n0 && j < G.n1;
  const unsigned col = (unsigned)i + (unsigned)j * (unsigned)G.s1;
  const unsigned s2u = (unsigned)G.s2;
  if (EG && grp == 2) {
    // ---- EOS group: the pressure of plane kbeg + m - 2's new state
    for (int m = 0; m <= NK + 1; m++) {
      const bool eos = A.write_p && valid && m >= 2;
      EosLook<2> pg;
      if (eos) {
        const int sl = (m - 1) & 1;
        double d = aout[sl][0][t], m1 = aout[sl][1][t], m2 = aout[sl][2][t], m3 = aout[sl][3][t],
               e = aout[sl][4][t];
        eos_issue<2>(A.E, d, m1, m2, m3, e, pg);
      }
      __syncthreads();
      if (eos) {
        double pr[5];
        eos_finish<2>(A.E, pg, pr);
        st32(Pout, col + (unsigned)(kbeg + m - 2) * s2u, pr[0]);
      }
      __syncthreads();
    }
  } else if (!zg) {
    // ---- xy group: the ring point of this thread as in k_stage_f
    int ring = 0;
    int r_row = 0, r_col = 0, r_i = 0, r_j = 0;
    if (t < 4 * ST_SX) {
      const int rr_ = t / ST_SX;
      r_row = rr_ < 2 ? rr_ : ST_SY + rr_;
      r_col = t % ST_SX;
      r_i = i0 + r_col;
      r_j = j0 - 2 + r_row;
      ring = 1;
    } else if (t < 4 * ST_SX + 4 * ST_SY) {
      const int u = t - 4 * ST_SX;
      r_row = u / 4;
      const int c4 = u % 4;
      r_col = c4 < 2 ? c4 : ST_SX + c4;
      r_i = i0 - 2 + r_col;
      r_j = j0 + r_row;
      ring = 2;
    }
    const bool ring_ok = ring != 0 && r_i < G.n0 && r_j < G.n1;
    const unsigned rcol = (unsigned)r_i + (unsigned)r_j * (unsigned)G.s1;
    double rc[ST_NF], rr[ST_NF];   // raw values of the own point and the ring point, plane kbeg + m
    if (own_ok) load_raw(Uin, Pin, bs, col + (unsigned)kbeg * s2u, rc);
    if (ring_ok) load_raw(Uin, Pin, bs, rcol + (unsigned)kbeg * s2u, rr);
    for (int m = 0; m <= NK + 1; m++) {
      double c[ST_NF];
      if (m < NK) {
        if (own_ok) prim_of_raw(rc, c);
        else {
#pragma unroll
          for (int f = 0; f < ST_NF; f++) c[f] = 1.0;
        }
        double o[14];
        flux_pack<0, ETA>(c, o);
#pragma unroll
        for (int f = 0; f < NFP; f++) fx[f][ty][tx + 2] = o[f];
        flux_pack<1, ETA>(c, o);
#pragma unroll
        for (int f = 0; f < NFP; f++) fy[f][ty + 2][tx] = o[f];
        if (ring != 0) {
          double h[ST_NF];
          if (ring_ok) prim_of_raw(rr, h);
          else {
#pragma unroll
            for (int f = 0; f < ST_NF; f++) h[f] = 1.0;
          }
          if (ring == 1) {
            flux_pack<1, ETA>(h, o);
#pragma unroll
            for (int f = 0; f < NFP; f++) fy[f][r_row][r_col] = o[f];
          } else {
            flux_pack<0, ETA>(h, o);
#pragma unroll
            for (int f = 0; f < NFP; f++) fx[f][r_row][r_col] = o[f];
          }
        }
        if (m + 1 < NK) {
          if (own_ok) load_raw(Uin, Pin, bs, col + (unsigned)(kbeg + m + 1) * s2u, rc);
          if (ring_ok) load_raw(Uin, Pin, bs, rcol + (unsigned)(kbeg + m + 1) * s2u, rr);
        }
      }
      // the pressure of plane kbeg + m - 2's new state (handed over in iteration m - 1)
      const bool eos = !EG && A.write_p && valid && m >= 2;
      EosLook<2> pg;
      if (eos) {
        const int sl = (m - 1) & 1;
        double d = aout[sl][0][t], m1 = aout[sl][1][t], m2 = aout[sl][2][t], m3 = aout[sl][3][t],
               e = aout[sl][4][t];
        eos_issue<2>(A.E, d, m1, m2, m3, e, pg);
      }
      __syncthreads();
      if (m < NK && valid) {
        double acc[MRM_NV], res[MRM_NV];
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) acc[q] = 0.0;
        const double lfc = lfac_of(c, A.RP.inv_va2max);
        double J[3] = {0.0, 0.0, 0.0};
        auto getx = [&](int s, int f) { return fx[f][ty][tx + s]; };
        sweep_residual_packed<0, ETA>(getx, c, A.RP.w1[0], A.RP.w2[0], lfc, res, &A.RP, J);
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
        auto gety = [&](int s, int f) { return fy[f][ty + s][tx]; };
        sweep_residual_packed<1, ETA>(gety, c, A.RP.w1[1], A.RP.w2[1], lfc, res, &A.RP, J);
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
        const int sl = m & 1;
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) axy[sl][q][t] = acc[q];
        axy[sl][MRM_NV][t] = lfc;
        if (ETA) {
#pragma unroll
          for (int q = 0; q < 3; q++) axy[sl][(MRM_NV + 1 + q) % NA][t] = J[q];
        }
      }
      if (eos) {
        double pr[5];
        eos_finish<2>(A.E, pg, pr);
        st32(Pout, col + (unsigned)(kbeg + m - 2) * s2u, pr[0]);
      }
      __syncthreads();
    }
  } else {
    // ---- z group
    double w[5][ST_NF];   // planes k-2 .. k+2 of this thread's column
    double rz[ST_NF];     // raw values of plane k + 3
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int kk = kbeg - 2 + q;
      if (own_ok) { double r[ST_NF]; load_raw(Uin, Pin, bs, col + (unsigned)kk * s2u, r); prim_of_raw(r, w[q + 1]); }
      else {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) w[q + 1][f] = 1.0;
      }
    }
    if (!EG && own_ok) load_raw(Uin, Pin, bs, col + (unsigned)(kbeg + 2) * s2u, rz);
    auto getz = [&](int s, double o[9]) {
#pragma unroll
      for (int f = 0; f < ST_NF; f++) o[f] = w[s][f];
    };
    for (int n = 0; n <= NK + 1; n++) {
      const int k = kbeg + n - 1;
      const bool act = n >= 1 && n <= NK;
      double res[MRM_NV];
      if (act) {
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
          for (int f = 0; f < ST_NF; f++) w[q][f] = w[q + 1][f];
        if (own_ok) {
          if (EG) {
            // three groups: no load ahead (the registers of rz), plane k + 2
            // loaded here; k + 2 <= lend2 + 2 = n2 - 1
            load_raw(Uin, Pin, bs, col + (unsigned)(k + 2) * s2u, rz);
            prim_of_raw(rz, w[4]);
          } else {
            prim_of_raw(rz, w[4]);     // plane k + 2
            // k + 3 <= kend + 1 <= lend2 + 2 = n2 - 1: inside the buffer
            if (k + 1 < kend) load_raw(Uin, Pin, bs, col + (unsigned)(k + 3) * s2u, rz);
          }
        }
        // without resistivity the dim-2 terms do not depend on the xy
        // group's sums: form them while that group packs its fluxes (with
        // it, J passes through the three sweeps in order, so they wait)
        if (!ETA && valid) {
          // the xy group handed this plane's sums over in the previous iteration
          const double lfc = axy[(n - 1) & 1][MRM_NV][t];
          sweep_residual_lr<2, false>(getz, w[2], A.RP.w1[2], A.RP.w2[2], lfc, res, &A.RP, nullptr);
        }
      }
      __syncthreads();
      if (act && valid) {
        const int sl = (n - 1) & 1;
        double acc[MRM_NV];
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) acc[q] = axy[sl][q][t];
        double J[3] = {0.0, 0.0, 0.0};
        if (ETA) {
#pragma unroll
          for (int q = 0; q < 3; q++) J[q] = axy[sl][(MRM_NV + 1 + q) % NA][t];
          const double lfc = axy[sl][MRM_NV][t];
          sweep_residual_lr<2, ETA>(getz, w[2], A.RP.w1[2], A.RP.w2[2], lfc, res, &A.RP, J);
        }
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
        if (ETA) acc[4] += A.RP.eta * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
        const unsigned node = col + (unsigned)k * s2u;
        if (ERES) acc[4] += ld32(A.eres, node);
        double u0[MRM_NV], out[MRM_NV];
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) u0[q] = ld32(U0 + q * bs, node);
        double Qrad = Qtot ? ld32(Qtot, node) : 0.0;
        const double dmpi = A.SP.dmp ? A.SP.dmp[i] : 0.0;
        source_integrate_cell(A.SP, w[2], acc, u0, Qrad, dmpi, out);
#pragma unroll
        for (int q = 0; q < MRM_NV; q++) st32(Uout + q * bs, node, out[q]);
        if (A.write_p) {
#pragma unroll
          for (int q = 0; q < 5; q++) aout[n & 1][q][t] = out[q];
        }
      }
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------ drop-in residual, tiled
// MHD_Residual of the per-operation path (mhdres_SR.C:27-854) with the stage
// kernel's tiling: U holds ConsToPrim's primitives, P the pressure; the
// plane sits in LDS with its two-cell ring in dims 0 and 1, dim-2
// neighbours come from a five-plane register window whose next plane is
// loaded during the current plane's work.  R += ((R_x) + R_y) + R_z per
// interior cell in the reference's order, and the interior points' cmax goes
// to red[RED_CMAX] (the ghost faces are k_cmax_faces').  RZERO: the interior
// of R is known to be zero; the sums start from 0.0 without reading it.
template <bool ETA, bool RZERO = false>
__global__ void __launch_bounds__(ST_SX* ST_SY, 1)
k_resid_tile(StageArgs A, const double* __restrict__ U, const double* __restrict__ P, double* __restrict__ R,
             unsigned long long* red) {
  __shared__ double lds[ST_NF][ST_LY][ST_LX];
  __shared__ double lred[ST_SX * ST_SY / 64];
  const Geom& G = A.G;
  const long bs = G.bs;
  const long nblocks = (long)A.ntx * A.nty * A.nkc;
  long lb = xcd_remap(blockIdx.x, nblocks);
  const int tx_t = (int)(lb % A.ntx);
  lb /= A.ntx;
  const int ty_t = (int)(lb % A.nty);
  const int kc = (int)(lb / A.nty);
  const int tx = threadIdx.x % ST_SX, ty = threadIdx.x / ST_SX;
  const int i0 = G.lb0 + tx_t * ST_SX, j0 = G.lb1 + ty_t * ST_SY;
  const int i = i0 + tx, j = j0 + ty;
  const bool valid = (i <= G.le0) && (j <= G.le1);
  const int kbeg = G.lb2 + kc * ST_KC;
  const int kend = min(kbeg + ST_KC, G.le2 + 1);     // exclusive
  int hx_c = -1, hx_i = 0;
  if (tx < 2) { hx_c = tx; hx_i = i0 - 2 + tx; }
  else if (tx >= ST_SX - 2) { hx_c = tx + 4; hx_i = i0 + tx + 2; }
  int hy_r = -1, hy_j = 0;
  if (ty < 4) {
    hy_r = ty < 2 ? ty : ST_SY + ty;
    hy_j = ty < 2 ? j0 - 2 + ty : j0 + ST_SY + (ty - 2);
  }
  const bool hx_ok = hx_c >= 0 && hx_i < G.n0 && j < G.n1;
  const bool hy_ok = hy_r >= 0 && hy_j < G.n1 && i < G.n0;
  const bool own_ok = i < G.n0 && j < G.n1;
  const unsigned col = (unsigned)i + (unsigned)j * (unsigned)G.s1;
  const unsigned s1u = (unsigned)G.s1, s2u = (unsigned)G.s2;
  double cm[1] = {0.0};
  double w[5][ST_NF];
  double rz[ST_NF], rhx[ST_NF], rhy[ST_NF];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if (own_ok) load_raw(U, P, bs, col + (unsigned)(kbeg - 2 + q) * s2u, w[q + 1]);
    else {
#pragma unroll
      for (int f = 0; f < ST_NF; f++) w[q + 1][f] = 1.0;
    }
  }
  if (own_ok) load_raw(U, P, bs, col + (unsigned)(kbeg + 2) * s2u, rz);
  if (hx_ok) load_raw(U, P, bs, (unsigned)hx_i + (unsigned)j * s1u + (unsigned)kbeg * s2u, rhx);
  if (hy_ok) load_raw(U, P, bs, (unsigned)i + (unsigned)hy_j * s1u + (unsigned)kbeg * s2u, rhy);
  for (int k = kbeg; k < kend; k++) {
#pragma unroll
    for (int q = 0; q < 4; q++)
#pragma unroll
      for (int f = 0; f < ST_NF; f++) w[q][f] = w[q + 1][f];
    if (own_ok) {
#pragma unroll
      for (int f = 0; f < ST_NF; f++) w[4][f] = rz[f];
    }
#pragma unroll
    for (int f = 0; f < ST_NF; f++) lds[f][ty + 2][tx + 2] = w[2][f];
    if (hx_ok) {
#pragma unroll
      for (int f = 0; f < ST_NF; f++) lds[f][ty + 2][hx_c] = rhx[f];
    }
    if (hy_ok) {
#pragma unroll
      for (int f = 0; f < ST_NF; f++) lds[f][hy_r][tx + 2] = rhy[f];
    }
    if (k + 1 < kend) {
      if (own_ok) load_raw(U, P, bs, col + (unsigned)(k + 3) * s2u, rz);
      if (hx_ok) load_raw(U, P, bs, (unsigned)hx_i + (unsigned)j * s1u + (unsigned)(k + 1) * s2u, rhx);
      if (hy_ok) load_raw(U, P, bs, (unsigned)i + (unsigned)hy_j * s1u + (unsigned)(k + 1) * s2u, rhy);
    }
    __syncthreads();
    if (valid) {
      const unsigned node = col + (unsigned)k * s2u;
      double acc[MRM_NV], res[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] = RZERO ? 0.0 : ld32(R + q * bs, node);
      const double lfc = lfac_of(w[2], A.RP.inv_va2max);
      auto getx = [&](int s, double o[9]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = lds[f][ty + 2][tx + s];
      };
      double J[3] = {0.0, 0.0, 0.0};
      sweep_residual_lr<0, ETA>(getx, w[2], A.RP.w1[0], A.RP.w2[0], lfc, res, &A.RP, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
      auto gety = [&](int s, double o[9]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = lds[f][ty + s][tx + 2];
      };
      sweep_residual_lr<1, ETA>(gety, w[2], A.RP.w1[1], A.RP.w2[1], lfc, res, &A.RP, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
      auto getz = [&](int s, double o[9]) {
#pragma unroll
        for (int f = 0; f < ST_NF; f++) o[f] = w[s][f];
      };
      sweep_residual_lr<2, ETA>(getz, w[2], A.RP.w1[2], A.RP.w2[2], lfc, res, &A.RP, J);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) acc[q] += res[q];
      if (ETA) acc[4] += A.RP.eta * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) st32(R + q * bs, node, acc[q]);
      cm[0] = smax(cm[0], smax(0.0, point_cmax(w[2], w[2][8], A.RP.inv_va2max)));
    }
    __syncthreads();
  }
  block_reduce<OpMax, 1>(cm, lred);
  if (threadIdx.x == 0) atomic_max_nonneg(red + RED_CMAX, cm[0]);
}

// ------------------------------------------------------------------ ghost fill of the fused path
// periodic copies of U and p (exchange.C:532-782 on one rank): dim 1 at
// interior (i,k), dim 2 at interior i and all j
template <int DIM>
__global__ void __launch_bounds__(MRM_BLOCK) k_exchange_p(Geom G, double* __restrict__ U, double* __restrict__ P) {
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
  long base = i + b * so;
  long src_r = base + (long)(lb + igh) * sd, dst_r = base + (long)(le + 1 + igh) * sd;
  long src_l = base + (long)(le - gh + 1 + igh) * sd, dst_l = base + (long)(lb - gh + igh) * sd;
#pragma unroll
  for (int q = 0; q < MRM_NV; q++) {
    U[q * G.bs + dst_r] = U[q * G.bs + src_r];
    U[q * G.bs + dst_l] = U[q * G.bs + src_l];
  }
  if (P) {
    P[dst_r] = P[src_r];
    P[dst_l] = P[src_l];
  }
}

// whether the dim-2 ghost planes of U and P (every i: the dim-0 sweep reads
// the dim-0 ghosts too; every j) hold exact copies of their periodic
// images: any difference (a NaN included) sets *diff to 1
__global__ void __launch_bounds__(MRM_BLOCK) k_ghost2_periodic(Geom G, const double* __restrict__ U,
                                                               const double* __restrict__ P,
                                                               unsigned long long* __restrict__ diff) {
  const int nx = G.n0;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)nx * G.n1 * G.g2) return;
  int i = (int)(t % nx);
  long r = t / nx;
  int b = (int)(r % G.n1);
  int igh = (int)(r / G.n1);
  long base = i + b * G.s1;
  long src_r = base + (long)(G.lb2 + igh) * G.s2, dst_r = base + (long)(G.le2 + 1 + igh) * G.s2;
  long src_l = base + (long)(G.le2 - G.g2 + 1 + igh) * G.s2, dst_l = base + (long)(G.lb2 - G.g2 + igh) * G.s2;
  // bit patterns, so that -0.0 against +0.0 (equal as numbers, but the
  // reference sweeps each plane with its own bits) counts as a difference
  auto eq = [](double a, double b) { return __double_as_longlong(a) == __double_as_longlong(b); };
  bool same = eq(P[dst_r], P[src_r]) && eq(P[dst_l], P[src_l]);
#pragma unroll
  for (int q = 0; q < MRM_NV; q++)
    same = same && eq(U[q * G.bs + dst_r], U[q * G.bs + src_r]) && eq(U[q * G.bs + dst_l], U[q * G.bs + src_l]);
  if (!same) *diff = 1ull;
}

// pressure of a freshly written ghost cell (ConsToPrim's p, cons_to_prim.C:108-198)
__device__ __forceinline__ double ghost_pres(const EosDev& E, const double u[MRM_NV]) {
  double m1 = u[1], m2 = u[2], m3 = u[3], e = u[4], pr[5];
  cons_to_prim_cell<2>(E, u[0], m1, m2, m3, e, pr);
  return pr[0];
}

// bottom and top ghost layers with their pressures; when cmax_out is set, the
// characteristic speed of the dim-0 ghost cells on interior (j,k) lines is
// reduced into it (mhdres_SR.C:237-328 visits them)
__global__ void __launch_bounds__(MRM_BLOCK) k_bc_fused(Geom G, EosDev E, BcParams B, const double* __restrict__ psum,
                                                        double* __restrict__ U, double* __restrict__ P,
                                                        double inv_va2max, unsigned long long* cmax_out) {
  __shared__ double lds[MRM_BLOCK / 64];
  const long bs = G.bs;
  // one thread per (column, side): bottom and top run in separate threads
  const long ncol = (long)G.n1 * G.n2;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  double cm[1] = {0.0};
  const bool top = t >= ncol;
  if (top) t -= ncol;
  if (t < ncol) {
    const double pmean1 = psum[0] * B.cc, pmean2 = psum[1] * B.cc;
    int j = (int)(t % G.n1), k = (int)(t / G.n1);
    const bool inner = j >= G.lb1 && j <= G.le1 && k >= G.lb2 && k <= G.le2;
    long base = j * G.s1 + k * G.s2;
    // ---- bottom (boundary_pdmp_1_fftw3.C:240-318)
    if (!top) {
      long n1 = G.lb0 + base, n2 = n1 + 1;
      double u1[MRM_NV], u2[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) { u1[q] = U[q * bs + n1]; u2[q] = U[q * bs + n2]; }
      bc_bottom_cells(E, B, pmean1, pmean2, u1, u2);
      long g1 = G.lb0 - 1 + base, g2 = G.lb0 - 2 + base;
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) { U[q * bs + g1] = u1[q]; U[q * bs + g2] = u2[q]; }
      double pg1 = ghost_pres(E, u1), pg2 = ghost_pres(E, u2);
      P[g1] = pg1;
      P[g2] = pg2;
      if (cmax_out && inner) {
        cm[0] = smax(cm[0], point_cmax_cons(u1, pg1, inv_va2max));
        cm[0] = smax(cm[0], point_cmax_cons(u2, pg2, inv_va2max));
      }
    }
    // ---- top (boundary_pdmp_1_fftw3.C:402-474), vertical or potential field
    else {
      long n1 = G.le0 + base, n2 = n1 - 1;
      double u1[MRM_NV], u2[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) { u1[q] = U[q * bs + n1]; u2[q] = U[q * bs + n2]; }
      double u0[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) u0[q] = u1[q];
      bc_top_cells(B, u1, u2);
      if (B.bext) {
        // the top interior cell's new B_y, B_z; its speed enters cmax here
        // (the last TVD sweep leaves it out, skip_top_cmax)
        bc_top_potential(B, t, &u0[6], &u0[7], u1, u2);
        U[6 * bs + n1] = u0[6];
        U[7 * bs + n1] = u0[7];
        if (cmax_out && inner) cm[0] = smax(cm[0], point_cmax_cons(u0, P[n1], inv_va2max));
      }
      long g1 = G.le0 + 1 + base, g2 = G.le0 + 2 + base;
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) { U[q * bs + g1] = u1[q]; U[q * bs + g2] = u2[q]; }
      double pg1 = ghost_pres(E, u1), pg2 = ghost_pres(E, u2);
      P[g1] = pg1;
      P[g2] = pg2;
      if (cmax_out && inner) {
        cm[0] = smax(cm[0], point_cmax_cons(u1, pg1, inv_va2max));
        cm[0] = smax(cm[0], point_cmax_cons(u2, pg2, inv_va2max));
      }
    }
  }
  if (cmax_out) {
    block_reduce<OpMax, 1>(cm, lds);
    if (threadIdx.x == 0) atomic_max_nonneg(cmax_out, cm[0]);
  }
}

// cell n of the buffer: interior, or a ghost cell on an interior line of its
// own direction (the points whose speeds MHD_Residual's sweeps reduce)
__device__ __forceinline__ void cmax_cell_class(const Geom& G, long n, bool& interior, bool& visited) {
  int i = (int)(n % G.n0);
  long r = n / G.n0;
  int j = (int)(r % G.n1), k = (int)(r / G.n1);
  bool ii = i >= G.lb0 && i <= G.le0, ji = j >= G.lb1 && j <= G.le1, ki = k >= G.lb2 && k <= G.le2;
  interior = ii && ji && ki;
  visited = interior || (!ii && ji && ki) || (ii && !ji && ki) || (ii && ji && !ki);
}

// pressure (and optionally cmax over the interior and the ghost faces, and
// the Adjust_Valf_Max statistics of the interior) of a conservative state
// over the whole buffer, U left untouched
__global__ void __launch_bounds__(MRM_BLOCK) k_eos_p(Geom G, EosDev E, const double* __restrict__ U,
                                                     double* __restrict__ P, double inv_va2max,
                                                     unsigned long long* cmax_out, unsigned long long* valf = nullptr,
                                                     double vlim95 = 0.0, double elim95 = 0.0) {
  __shared__ double lds[5 * MRM_BLOCK / 64];
  const long bs = G.bs;
  double cm[1] = {0.0};
  ValfAcc va;
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
    double u[MRM_NV];
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) u[q] = U[q * bs + n];
    double m1 = u[1], m2 = u[2], m3 = u[3], ei = u[4], pr[5];
    cons_to_prim_cell<2>(E, u[0], m1, m2, m3, ei, pr);
    const double p = pr[0];
    P[n] = p;
    if (cmax_out || valf) {
      bool inner, vis;
      cmax_cell_class(G, n, inner, vis);
      if (cmax_out && vis) cm[0] = smax(cm[0], point_cmax_cons(u, p, inv_va2max));
      if (valf && inner) valf_cell(va, u[0], m1, m2, m3, ei, p, vlim95, elim95);
    }
  }
  if (cmax_out) {
    block_reduce<OpMax, 1>(cm, lds);
    if (threadIdx.x == 0) atomic_max_nonneg(cmax_out, cm[0]);
  }
  if (valf) {
    __syncthreads();
    valf_flush(va, lds, valf);
  }
}

// ConsToPrim's EOS outputs of the conservative state U written to F without
// converting U (the drop-in path's deferred stage).  PRES: the pressure alone
// and the characteristic speed of the points MHD_Residual's sweeps visit
// (k_eos_p's set) into cmax_out -- what the stage kernel and the host's dt
// need of a ConsToPrim + MHD_Residual pair; the pressure is ConsToPrim's, bit
// for bit (the temperature only decides its validity).  !PRES: the other
// four outputs (T, ne, rhoi, amb), formed when a caller reads them.
template <bool PRES>
__global__ void __launch_bounds__(MRM_BLOCK) k_eos_fields(Geom G, EosDev E, const double* __restrict__ U,
                                                          double* __restrict__ F, double inv_va2max,
                                                          unsigned long long* cmax_out) {
  __shared__ double lds[MRM_BLOCK / 64];
  const long bs = G.bs;
  double cm[1] = {0.0};
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
    double u[MRM_NV];
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) u[q] = U[q * bs + n];
    double m1 = u[1], m2 = u[2], m3 = u[3], ei = u[4], pr[5];
    if (PRES) {
      cons_to_prim_cell<2>(E, u[0], m1, m2, m3, ei, pr);
      F[MRM_PRES * bs + n] = pr[0];
      bool inner, vis;
      cmax_cell_class(G, n, inner, vis);
      if (vis) cm[0] = smax(cm[0], point_cmax_cons(u, pr[0], inv_va2max));
    } else {
      cons_to_prim_cell<5>(E, u[0], m1, m2, m3, ei, pr);
#pragma unroll
      for (int q = 1; q < 5; q++) F[q * bs + n] = pr[q];
    }
  }
  if (PRES) {
    block_reduce<OpMax, 1>(cm, lds);
    if (threadIdx.x == 0) atomic_max_nonneg(cmax_out, cm[0]);
  }
}

// every ghost cell (outside the interior) of the conservative state src
// written to dst in ConsToPrim's primitive form (v = M (1/rho),
// e_int = e - rho v^2 / 2, cons_to_prim.C:110-118; rho and B as they are): the
// ghost cells a ConsToPrim followed by an in-place Source_Integrate_Tcheck
// leaves.  The ghost width is 2 in every dimension.  Index space: the dim-0
// ghost layers over all (j, k), then the dim-1 ghost rows over all (i, k),
// then the dim-2 ghost planes over all (i, j); edges and corners are written
// more than once, with the same values.  DIM0: the dim-0 ghost layers alone.
template <bool DIM0>
__global__ void __launch_bounds__(MRM_BLOCK) k_ghosts_prim(Geom G, const double* __restrict__ src,
                                                           double* __restrict__ dst) {
  const long na = 4L * G.n1 * G.n2, nb = 4L * G.n0 * G.n2, nc = 4L * G.n0 * G.n1;
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  int i, j, k;
  if (DIM0 && t >= na) return;
  if (t < na) {
    const int l = (int)(t % 4);
    const long r = t / 4;
    j = (int)(r % G.n1); k = (int)(r / G.n1); i = l < 2 ? l : G.n0 - 4 + l;
  } else if ((t -= na) < nb) {
    const int l = (int)(t % 4);
    const long r = t / 4;
    i = (int)(r % G.n0); k = (int)(r / G.n0); j = l < 2 ? l : G.n1 - 4 + l;
  } else if ((t -= nb) < nc) {
    const int l = (int)(t % 4);
    const long r = t / 4;
    i = (int)(r % G.n0); j = (int)(r / G.n0); k = l < 2 ? l : G.n2 - 4 + l;
  } else {
    return;
  }
  const long n = i + j * G.s1 + k * G.s2, bs = G.bs;
  const double d = src[n];
  const double dn = 1.0 / d;
  const double m1 = src[bs + n] * dn, m2 = src[2 * bs + n] * dn, m3 = src[3 * bs + n] * dn;
  const double vv = (m1 * m1 + m2 * m2 + m3 * m3);
  dst[n] = d;
  dst[bs + n] = m1;
  dst[2 * bs + n] = m2;
  dst[3 * bs + n] = m3;
  dst[4 * bs + n] = src[4 * bs + n] - 0.5 * vv * d;
#pragma unroll
  for (int q = 5; q < MRM_NV; q++) dst[q * bs + n] = src[q * bs + n];
}

// cmax alone over the same points, from a state and its pressure: the
// stage-1 speed bound after Adjust_Valf_Max has changed inv_va2max
__global__ void __launch_bounds__(MRM_BLOCK) k_cmax_visited(Geom G, const double* __restrict__ U,
                                                            const double* __restrict__ P, double inv_va2max,
                                                            unsigned long long* cmax_out) {
  __shared__ double lds[MRM_BLOCK / 64];
  const long bs = G.bs;
  double cm[1] = {0.0};
  for (long n = blockIdx.x * (long)blockDim.x + threadIdx.x; n < bs; n += (long)gridDim.x * blockDim.x) {
    bool inner, vis;
    cmax_cell_class(G, n, inner, vis);
    if (vis) {
      double u[MRM_NV];
#pragma unroll
      for (int q = 0; q < MRM_NV; q++) u[q] = U[q * bs + n];
      cm[0] = smax(cm[0], point_cmax_cons(u, P[n], inv_va2max));
    }
  }
  block_reduce<OpMax, 1>(cm, lds);
  if (threadIdx.x == 0) atomic_max_nonneg(cmax_out, cm[0]);
}

// ------------------------------------------------------------------ TVD, dim-0 sweep, one block per line
// thread t owns point t of the line; the per-point, per-slope, per-interface
// and per-update arithmetic is tvd_point / tvd_slope / tvd_iface / tvd_update

// lines whose arrays exceed 64 KB of LDS ask for the larger dynamic LDS
// first (gfx950 gives one block up to 160 KB)
template <class K>
inline void allow_lds(K kernel, size_t lds) {
  if (lds > 64 * 1024) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    (void)hipGetLastError();
  }
}

// MAXT: the block size bound (lines of up to MAXT points).  At 1024 the
// kernel fits 128 VGPRs (4 waves per SIMD) without spilling; a 320- or
// 512-thread bound lets it use more registers at fewer waves, which
// measured slower on box512 lines
template <int MAXT, bool CNT = false>
__global__ void __launch_bounds__(MAXT) k_tvd_x_lines(TvdLineArgs A, double* __restrict__ U,
                                                        const double* __restrict__ Pr) {
  extern __shared__ double sm[];
  const Geom& G = A.G;
  const TvdParams& P = A.P;
  const long bs = G.bs;
  const int sz = G.n0, gh = G.g0;
  const long nlines = (long)G.n1 * G.n2;
  const long line = xcd_remap(blockIdx.x, nlines);
  const int j = (int)(line % G.n1), k = (int)(line / G.n1);
  const long base = j * G.s1 + k * G.s2;
  double* sv = sm;                     // [8][sz] primitives (log rho/eps)
  double* sS = sv + MRM_NV * sz;       // [8][sz] slopes, then the interface fluxes
  double* sF = sS;
  double* scm = sS + MRM_NV * sz;      // cm
  double* shd = scm + sz;              // hyperdiffusion coefficient
  double* scz = shd + sz;              // CZ
  double* slr = scz + sz;              // ln rho
  double* sq = slr + sz;               // resistive interface term
  const int t = threadIdx.x;
  TvdCnt cnt;
  if (CNT) {
    cnt.clear(P.btie);
    cnt.on = j >= G.lb1 && j <= G.le1 && k >= G.lb2 && k <= G.le2;
    cnt.key = br_key(0, bs, base + t);
  }
  TvdCnt* const cn = CNT ? &cnt : nullptr;
  TvdPt me;
  if (t < sz) {
    double u[MRM_NV];
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) u[q] = U[q * bs + base + t];
    double tau = P.tau ? P.tau[base + t] : 0.0;
    tvd_point(P, u, Pr[base + t], tau, P.hfb[t], P.hft[t], true, me, cn);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) sv[q * sz + t] = me.v[q];
    scm[t] = me.cm;
    shd[t] = me.hd;
    scz[t] = me.CZ;
    slr[t] = me.lnr;
  }
  __syncthreads();
  // slopes need v at t-1, t+1 and ln rho at t+1: rebuild the neighbour
  // points' fields that tvd_slope reads
  double S[MRM_NV];
  if (t >= 1 && t <= sz - 2) {
    TvdPt a, c;
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) { a.v[q] = sv[q * sz + t - 1]; c.v[q] = sv[q * sz + t + 1]; }
    c.lnr = slr[t + 1];
    tvd_slope(P, a, me, c, P.hfb[t], P.hft[t], S, cn);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) sS[q * sz + t] = S[q];
  }
  __syncthreads();
  double F[MRM_NV];
  const int istart = 1 + (P.is_gbeg0 ? gh : 0);
  const bool iface = t >= 1 && t <= sz - 3;
  // the fluxes overwrite the slopes: every slope is read before any flux is stored
  double Sc[MRM_NV];
  if (iface) {
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) Sc[q] = sS[q * sz + t + 1];
  }
  __syncthreads();
  if (iface) {
    TvdPt a, c, d;
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) {
      a.v[q] = sv[q * sz + t - 1];
      c.v[q] = sv[q * sz + t + 1];
      d.v[q] = sv[q * sz + t + 2];
    }
    c.cm = scm[t + 1];
    c.hd = shd[t + 1];
    c.CZ = scz[t + 1];
    bool zb = (P.is_gend0 && t == sz - gh - 1) || (P.is_gbeg0 && t == gh - 1);
    double q = tvd_iface<0>(P, a, me, c, d, S, Sc, P.hfb[t], P.hft[t], P.qft[t], t >= istart, zb, F, nullptr, cn);
#pragma unroll
    for (int qq = 0; qq < MRM_NV; qq++) sF[qq * sz + t] = F[qq];
    sq[t] = q;
  }
  __syncthreads();
  if (t >= gh && t <= sz - gh - 1) {
    // the conserved values are read again (an L2 hit) rather than kept in
    // registers through the three barriers
    double Fm[MRM_NV], u[MRM_NV];
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) {
      Fm[q] = sF[q * sz + t - 1];
      u[q] = U[q * bs + base + t];
    }
    tvd_update(P, me, Fm, F, sq[t - 1], sq[t], P.qft[t], u, nullptr, cn);
#pragma unroll
    for (int q = 0; q < MRM_NV; q++) U[q * bs + base + t] = u[q];
  }
  if (CNT) tvd_cnt_flush(cnt, P.bcount);
}

}  // namespace MRM_NS
