// mrm_potential.hip -- the potential-field extrapolation of the top boundary
// on the GPU (potential_sd_fftw3.C:27-257): mixed-radix Stockham FFTs, one
// workgroup per line with the line, its ping-pong copy and the twiddle table
// in LDS.
//
//   k_pot_rows_fwd   row DFTs (dim 1) of bz0 = B_x(lend), or of the kernels
//   k_pot_cols       column DFTs (dim 2); for an update also the 8 spectrum
//                    products and their inverse column DFTs
//   k_pot_rows_inv   inverse row DFTs, real part into b_ext with its ghosts
//
// A Stockham pass of radix R over a line of N points with Ns = the product
// of the radices already applied reads x[j + r N/R] (r < R), multiplies
// point r by w^(r k) (k = j mod Ns, w = e^{-2 pi i / (Ns R)}), takes the
// R-point DFT and writes it to y[(j / Ns) Ns R + k + s Ns] (s < R); the
// last pass leaves the transform in natural order.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "mrm_potential_api.h"

namespace pot {
namespace {

constexpr int BLK = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// twiddle e^{-+2 pi i t / N} from the table of the forward direction
template <bool INV>
__device__ __forceinline__ double2 twv(const double2* tw, int t) {
  double2 w = tw[t];
  if (INV) w.y = -w.y;
  return w;
}

// R-point DFT in place; NR = N / R indexes w_R = tw[NR]
template <int R, bool INV>
__device__ __forceinline__ void small_dft(double2 v[R], const double2* tw, int NR) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const double2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
    const double2 c = cadd(v[1], v[3]), d = csub(v[1], v[3]);
    // -i d (forward) or +i d (inverse)
    const double2 jd = INV ? make_double2(-d.y, d.x) : make_double2(d.y, -d.x);
    v[0] = cadd(a, c);
    v[2] = csub(a, c);
    v[1] = cadd(b, jd);
    v[3] = csub(b, jd);
  } else {
    double2 o[R];
#pragma unroll
    for (int s = 0; s < R; s++) {
      o[s] = v[0];
#pragma unroll
      for (int r = 1; r < R; r++) o[s] = cadd(o[s], cmul(v[r], twv<INV>(tw, ((r * s) % R) * NR)));
    }
#pragma unroll
    for (int s = 0; s < R; s++) v[s] = o[s];
  }
}

template <int R, bool INV>
__device__ __forceinline__ void pass(const double2* __restrict__ x, double2* __restrict__ y, const double2* tw, int N,
                                     int Ns, int j) {
  const int NR = N / R;
  const int k = j % Ns;
  const int step = NR / Ns;   // N / (Ns R)
  double2 v[R];
#pragma unroll
  for (int r = 0; r < R; r++) v[r] = x[j + r * NR];
#pragma unroll
  for (int r = 1; r < R; r++) v[r] = cmul(v[r], twv<INV>(tw, r * k * step));
  small_dft<R, INV>(v, tw, NR);
  const int d = (j / Ns) * Ns * R + k;
#pragma unroll
  for (int s = 0; s < R; s++) y[d + s * Ns] = v[s];
}

// transform of the line in a (LDS), b is the ping-pong buffer; returns the
// buffer holding the result.  Every thread of the block calls it.
template <bool INV>
__device__ double2* fft_lds(double2* a, double2* b, const double2* tw, const Fac& F) {
  const int N = F.n;
  int Ns = 1;
  for (int s = 0; s < F.ns; s++) {
    const int R = F.r[s];
    for (int j = threadIdx.x; j < N / R; j += blockDim.x) {
      switch (R) {
        case 4: pass<4, INV>(a, b, tw, N, Ns, j); break;
        case 2: pass<2, INV>(a, b, tw, N, Ns, j); break;
        case 3: pass<3, INV>(a, b, tw, N, Ns, j); break;
        case 5: pass<5, INV>(a, b, tw, N, Ns, j); break;
        default: pass<7, INV>(a, b, tw, N, Ns, j); break;
      }
    }
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// row transforms: KER reads kernel rows K[row][x] (row = q gy + y), else
// B_x(lend) of row y
template <bool KER>
__global__ void __launch_bounds__(BLK) k_pot_rows_fwd(Geom G, Fac F, const double2* __restrict__ twg,
                                                      const double* __restrict__ U, const float* __restrict__ K,
                                                      double2* __restrict__ W) {
  extern __shared__ double2 sm[];
  const int gx = F.n;
  double2 *a = sm, *b = sm + gx, *tw = sm + 2 * gx;
  const long row = blockIdx.x;
  for (int x = threadIdx.x; x < gx; x += blockDim.x) {
    tw[x] = twg[x];
    double v;
    if (KER) v = (double)K[row * gx + x];
    else v = U[5 * G.bs + G.le0 + (long)(x + G.lb1) * G.s1 + (row + G.lb2) * G.s2];
    a[x] = make_double2(v, 0.0);
  }
  __syncthreads();
  const double2* r = fft_lds<false>(a, b, tw, F);
  for (int x = threadIdx.x; x < gx; x += blockDim.x) W[row * gx + x] = r[x];
}

// column transforms.  KER: block (q, kx) of 8 gx; the forward column DFT of
// plane q of W divided by gx gy into S.  Else block kx: the forward column
// DFT of W, then for each kernel q its product with S[q] and the inverse
// column DFT into Z[q].
template <bool KER>
__global__ void __launch_bounds__(BLK) k_pot_cols(Fac F, int gx, const double2* __restrict__ twg,
                                                  const double2* __restrict__ W, double2* __restrict__ S,
                                                  double2* __restrict__ Z) {
  extern __shared__ double2 sm[];
  const int gy = F.n;
  double2 *a = sm, *b = sm + gy, *tw = sm + 2 * gy, *c = sm + 3 * gy;
  const long plane = (long)gx * gy;
  const int kx = KER ? (int)(blockIdx.x % gx) : (int)blockIdx.x;
  const long q0 = KER ? (long)(blockIdx.x / gx) : 0;
  for (int y = threadIdx.x; y < gy; y += blockDim.x) {
    tw[y] = twg[y];
    a[y] = W[q0 * plane + (long)y * gx + kx];
  }
  __syncthreads();
  double2* f = fft_lds<false>(a, b, tw, F);
  if (KER) {
    const double nrm = (double)gy * (double)gx;
    for (int y = threadIdx.x; y < gy; y += blockDim.x)
      S[q0 * plane + (long)y * gx + kx] = make_double2(f[y].x / nrm, f[y].y / nrm);
    return;
  }
  // the spectrum stays in f; the products go through the other two buffers
  double2* p0 = (f == a) ? b : a;
  double2* p1 = c;
  for (int q = 0; q < 8; q++) {
    for (int y = threadIdx.x; y < gy; y += blockDim.x) p0[y] = cmul(f[y], S[q * plane + (long)y * gx + kx]);
    __syncthreads();
    const double2* r = fft_lds<true>(p0, p1, tw, F);
    for (int y = threadIdx.x; y < gy; y += blockDim.x) Z[q * plane + (long)y * gx + kx] = r[y];
    __syncthreads();
  }
}

// inverse row transforms of Z, block (q, y) of 8 gy: the real part into
// plane q of b_ext at (x + g1, y + g2), with the periodic ghost columns and,
// for rows within g2 of either edge, their ghost-row copies
__global__ void __launch_bounds__(BLK) k_pot_rows_inv(Geom G, Fac F, int gy, const double2* __restrict__ twg,
                                                      const double2* __restrict__ Z, double* __restrict__ bext) {
  extern __shared__ double2 sm[];
  const int gx = F.n;
  double2 *a = sm, *b = sm + gx, *tw = sm + 2 * gx;
  const long row = blockIdx.x;
  const int q = (int)(row / gy), y = (int)(row % gy);
  for (int x = threadIdx.x; x < gx; x += blockDim.x) {
    tw[x] = twg[x];
    a[x] = Z[row * gx + x];
  }
  __syncthreads();
  const double2* r = fft_lds<true>(a, b, tw, F);
  const int n1 = G.n1, g1 = G.lb1, g2 = G.lb2;
  double* P = bext + (long)q * n1 * G.n2;
  // destination rows: y + g2, and y + g2 + gy (y < g2) or y + g2 - gy (y >= gy - g2)
  int rows[3] = {y + g2, -1, -1};
  if (y < g2) rows[1] = y + g2 + gy;
  if (y >= gy - g2) rows[2] = y + g2 - gy;
  for (int x = threadIdx.x; x < gx; x += blockDim.x) {
    const double v = r[x].x;
#pragma unroll
    for (int m = 0; m < 3; m++) {
      if (rows[m] < 0) continue;
      double* d = P + (long)rows[m] * n1;
      d[x + g1] = v;
      if (x >= gx - g1) d[x + g1 - gx] = v;
      if (x < g1) d[x + g1 + gx] = v;
    }
  }
}

bool factor(int N, Fac& F) {
  F = Fac();
  F.n = N;
  int m = N;
  while (m % 4 == 0 && F.ns < MAXST) { F.r[F.ns++] = 4; m /= 4; }
  for (int r : {2, 3, 5, 7})
    while (m % r == 0 && F.ns < MAXST) { F.r[F.ns++] = r; m /= r; }
  return m == 1;
}

// e^{-2 pi i t / N} for t < N: the angle reduced to a quadrant and to its
// first half, in long double, so that quarter turns are exact
std::vector<double2> twiddles(int N) {
  std::vector<double2> w(N);
  const long double hp = 1.5707963267948966192313216916397514L;   // pi / 2
  for (int t = 0; t < N; t++) {
    const long t4 = 4L * t;
    const int quad = (int)(t4 / N);
    const long rem = t4 - (long)quad * N;   // angle within the quadrant: (pi/2) rem / N
    long double c, s;
    if (2 * rem <= N) {
      const long double th = hp * (long double)rem / (long double)N;
      c = cosl(th);
      s = sinl(th);
    } else {
      const long double th = hp * (long double)(N - rem) / (long double)N;
      c = sinl(th);
      s = cosl(th);
    }
    long double cc, ss;
    switch (quad) {
      case 0: cc = c; ss = s; break;
      case 1: cc = -s; ss = c; break;
      case 2: cc = -c; ss = -s; break;
      default: cc = s; ss = -c; break;
    }
    w[t] = make_double2((double)cc, -(double)ss);
  }
  return w;
}

size_t lds_rows(int gx) { return (size_t)3 * gx * sizeof(double2); }
size_t lds_cols(int gy, bool ker) { return (size_t)(ker ? 3 : 4) * gy * sizeof(double2); }

hipError_t allow(const void* fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

}  // namespace

int plan_create(Plan** out, int gx, int gy, std::string& err) {
  *out = nullptr;
  err.clear();
  if (gx < 1 || gy < 1 || gx > MAXN || gy > MAXN) {
    err = "potential field: horizontal extents must lie in 1.." + std::to_string(MAXN);
    return MRM_ERR_UNSUPPORTED;
  }
  Plan* p = new Plan();
  p->gx = gx;
  p->gy = gy;
  if (!factor(gx, p->fx) || !factor(gy, p->fy)) {
    delete p;
    err = "potential field: horizontal extents must be products of 2, 3, 5 and 7";
    return MRM_ERR_UNSUPPORTED;
  }
  const size_t plane = (size_t)gx * gy;
  hipError_t e = hipMalloc(&p->twx, gx * sizeof(double2));
  if (e == hipSuccess) e = hipMalloc(&p->twy, gy * sizeof(double2));
  if (e == hipSuccess) e = hipMalloc(&p->kspec, 8 * plane * sizeof(double2));
  if (e == hipSuccess) e = hipMalloc(&p->W, plane * sizeof(double2));
  if (e == hipSuccess) e = hipMalloc(&p->Z, 8 * plane * sizeof(double2));
  std::vector<double2> wx = twiddles(gx), wy = twiddles(gy);
  if (e == hipSuccess) e = hipMemcpy(p->twx, wx.data(), gx * sizeof(double2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(p->twy, wy.data(), gy * sizeof(double2), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = allow((const void*)&k_pot_rows_fwd<true>, lds_rows(gx));
  if (e == hipSuccess) e = allow((const void*)&k_pot_rows_fwd<false>, lds_rows(gx));
  if (e == hipSuccess) e = allow((const void*)&k_pot_rows_inv, lds_rows(gx));
  if (e == hipSuccess) e = allow((const void*)&k_pot_cols<true>, lds_cols(gy, true));
  if (e == hipSuccess) e = allow((const void*)&k_pot_cols<false>, lds_cols(gy, false));
  if (e != hipSuccess) {
    err = std::string("potential field: ") + hipGetErrorString(e);
    plan_destroy(p);
    return MRM_ERR_HIP;
  }
  *out = p;
  return MRM_OK;
}

void plan_destroy(Plan* p) {
  if (!p) return;
  hipFree(p->twx);
  hipFree(p->twy);
  hipFree(p->kspec);
  hipFree(p->W);
  hipFree(p->Z);
  delete p;
}

hipError_t plan_kernels(Plan* p, const float* k8, hipStream_t s) {
  const size_t plane = (size_t)p->gx * p->gy;
  float* dk = nullptr;
  hipError_t e = hipMalloc(&dk, 8 * plane * sizeof(float));
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(dk, k8, 8 * plane * sizeof(float), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    Geom G{};
    k_pot_rows_fwd<true><<<8 * p->gy, BLK, lds_rows(p->gx), s>>>(G, p->fx, p->twx, nullptr, dk, p->Z);
    k_pot_cols<true><<<8 * p->gx, BLK, lds_cols(p->gy, true), s>>>(p->fy, p->gx, p->twy, p->Z, p->kspec, nullptr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  hipFree(dk);
  if (e == hipSuccess) p->kernels = true;
  return e;
}

hipError_t plan_update(const Plan* p, hipStream_t s, const Geom& G, const double* U, double* bext) {
  k_pot_rows_fwd<false><<<p->gy, BLK, lds_rows(p->gx), s>>>(G, p->fx, p->twx, U, nullptr, p->W);
  k_pot_cols<false><<<p->gx, BLK, lds_cols(p->gy, false), s>>>(p->fy, p->gx, p->twy, p->W, p->kspec, p->Z);
  k_pot_rows_inv<<<8 * p->gy, BLK, lds_rows(p->gx), s>>>(G, p->fx, p->gy, p->twx, p->Z, bext);
  return hipGetLastError();
}

}  // namespace pot
