// mrm_potential_api.h -- potential-field extrapolation of the top boundary
// (potential_sd_fftw3.C:27-257, boundary_pdmp_1_fftw3.C:335-381), one rank
// holding the whole horizontal plane.
//
// b_ext[q] = the circular convolution of bz0 = B_x(lend) over the
// horizontal plane with kernel q, for the 8 kernels in the order z-1, z-2,
// x-0, x-1, x-2, y-0, y-1, y-2.  As in the reference the convolution runs
// through 2D DFTs: the kernels' spectra are formed once and divided by
// gx*gy, then every update is one forward transform of bz0, 8 spectrum
// products and 8 inverse transforms, of which the real part is kept.  The
// transforms are mixed-radix (2, 3, 4, 5, 7) Stockham FFTs with the whole
// line in LDS, one workgroup per line: rows (dim 1) first, then columns
// (dim 2).  b_ext holds 8 planes of n1 x n2 (dim 1 fastest) with the
// periodic ghosts filled, dim 1 first and then dim 2 over the whole dim-1
// range (potential_sd_fftw3.C:218-228, 405-486).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "mrm_types.h"

namespace pot {

constexpr int MAXN = 2048;    // longest horizontal extent (LDS: 64 * N bytes per column)
constexpr int MAXST = 16;

struct Fac {
  int n = 1, ns = 0;
  int r[MAXST] = {};
};

struct Plan {
  Fac fx, fy;
  int gx = 0, gy = 0;
  double2* twx = nullptr;     // e^{-2 pi i t / gx}, t < gx
  double2* twy = nullptr;
  double2* kspec = nullptr;   // [8][gy][gx]: kernel spectra / (gx gy)
  double2* W = nullptr;       // [gy][gx] row transforms of bz0
  double2* Z = nullptr;       // [8][gy][gx] column-inverse products
  bool kernels = false;
};

// plan for a gx x gy plane; "" or an error message in err
int plan_create(Plan** out, int gx, int gy, std::string& err);
void plan_destroy(Plan* p);
// k8: 8 x gy x gx floats on the host (dim 1 fastest); forms the spectra
hipError_t plan_kernels(Plan* p, const float* k8, hipStream_t s);
// bz0 = U[5](lend, interior j, k) -> b_ext (8 x n1 x n2 with ghosts)
hipError_t plan_update(const Plan* p, hipStream_t s, const Geom& G, const double* U, double* bext);

}  // namespace pot
