// mrm_io.cpp -- MURaM's restart files on the host: result_prim_<v>.<iter>,
// Header.<iter> and backup.dat (io_xysl.C:39-48, 99-209, 224-412;
// boundary_pdmp_1_fftw3.C:44-62), for one rank holding the whole grid.
//
// result_prim_<v>: float32, the interior, i (dim 0) fastest, one file per
// variable v = rho, v0, v1, v2, e_int, B0, B1, B2 (primitive).  Reading
// converts to conservative (e += rho v^2 / 2, M = rho v) in FP64 as
// RestoreSolution does; writing converts back in FP64 and rounds to float as
// OutputSolution does.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/muram_step.h"

namespace {

std::string prim_name(const char* dir, const char* resfile, int v, int iter) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "_prim_%d.%06d", v, iter);
  return std::string(dir) + resfile + buf;
}

struct Dims {
  int n[3], g[3], gs[3];
  long bs;
};
bool dims_of(const mrm_grid_desc* grid, Dims& D) {
  if (!grid) return false;
  for (int d = 0; d < 3; d++) {
    D.gs[d] = grid->gsize[d];
    D.g[d] = grid->ghosts[d];
    if (D.gs[d] < 1 || D.g[d] < 0) return false;
    D.n[d] = D.gs[d] + 2 * D.g[d];
  }
  D.bs = (long)D.n[0] * D.n[1] * D.n[2];
  return true;
}
inline long node(const Dims& D, int i, int j, int k) {
  return (long)(i + D.g[0]) + (long)D.n[0] * ((long)(j + D.g[1]) + (long)D.n[1] * (k + D.g[2]));
}

}  // namespace

extern "C" int mrm_io_read_prim(const char* dir, const char* resfile, int iter, const mrm_grid_desc* grid,
                                double* U) {
  Dims D;
  if (!dir || !resfile || !U || !dims_of(grid, D)) return MRM_ERR_ARG;
  const long n_int = (long)D.gs[0] * D.gs[1] * D.gs[2];
  std::vector<float> buf(n_int);
  std::memset(U, 0, sizeof(double) * 8 * D.bs);
  for (int v = 0; v < 8; v++) {
    std::string f = prim_name(dir, resfile, v, iter);
    FILE* fh = std::fopen(f.c_str(), "rb");
    if (!fh) return MRM_ERR_ARG;
    const size_t got = std::fread(buf.data(), sizeof(float), n_int, fh);
    std::fclose(fh);
    if ((long)got != n_int) return MRM_ERR_ARG;
    for (int k = 0; k < D.gs[2]; k++)
      for (int j = 0; j < D.gs[1]; j++)
        for (int i = 0; i < D.gs[0]; i++)
          U[v * D.bs + node(D, i, j, k)] = (double)buf[i + (long)D.gs[0] * (j + (long)D.gs[1] * k)];
  }
  // io_xysl.C:401-406: e += M^2 d / 2 with M holding v, then M *= d
  for (int k = 0; k < D.gs[2]; k++)
    for (int j = 0; j < D.gs[1]; j++)
      for (int i = 0; i < D.gs[0]; i++) {
        const long n = node(D, i, j, k);
        const double d = U[n];
        const double v0 = U[D.bs + n], v1 = U[2 * D.bs + n], v2 = U[3 * D.bs + n];
        U[4 * D.bs + n] += 0.5 * (v0 * v0 + v1 * v1 + v2 * v2) * d;
        U[D.bs + n] = v0 * d;
        U[2 * D.bs + n] = v1 * d;
        U[3 * D.bs + n] = v2 * d;
      }
  return MRM_OK;
}

extern "C" int mrm_io_write_prim(const char* dir, const char* resfile, int iter, const mrm_grid_desc* grid,
                                 const double* U) {
  Dims D;
  if (!dir || !resfile || !U || !dims_of(grid, D)) return MRM_ERR_ARG;
  const long n_int = (long)D.gs[0] * D.gs[1] * D.gs[2];
  std::vector<float> buf(n_int);
  for (int v = 0; v < 8; v++) {
    for (int k = 0; k < D.gs[2]; k++)
      for (int j = 0; j < D.gs[1]; j++)
        for (int i = 0; i < D.gs[0]; i++) {
          const long n = node(D, i, j, k);
          // io_xysl.C:154-159: M /= d, e -= d M^2 / 2
          const double d = U[n];
          const double v0 = U[D.bs + n] / d, v1 = U[2 * D.bs + n] / d, v2 = U[3 * D.bs + n] / d;
          double x;
          if (v == 0) x = d;
          else if (v <= 3) x = v == 1 ? v0 : v == 2 ? v1 : v2;
          else if (v == 4) x = U[4 * D.bs + n] - 0.5 * d * (v0 * v0 + v1 * v1 + v2 * v2);
          else x = U[v * D.bs + n];
          buf[i + (long)D.gs[0] * (j + (long)D.gs[1] * k)] = (float)x;
        }
    std::string f = prim_name(dir, resfile, v, iter);
    FILE* fh = std::fopen(f.c_str(), "wb");
    if (!fh) return MRM_ERR_ARG;
    const size_t put = std::fwrite(buf.data(), sizeof(float), n_int, fh);
    std::fclose(fh);
    if ((long)put != n_int) return MRM_ERR_ARG;
  }
  return MRM_OK;
}

// one result_prim_<v> file (v = 8: Spitzer's heat flux, io_xysl.C:322-336,
// 389-392) to or from a field of bufsize doubles, ghosts zero on reading
extern "C" int mrm_io_read_var(const char* dir, const char* resfile, int v, int iter, const mrm_grid_desc* grid,
                               double* f) {
  Dims D;
  if (!dir || !resfile || !f || v < 0 || !dims_of(grid, D)) return MRM_ERR_ARG;
  const long n_int = (long)D.gs[0] * D.gs[1] * D.gs[2];
  std::vector<float> buf(n_int);
  FILE* fh = std::fopen(prim_name(dir, resfile, v, iter).c_str(), "rb");
  if (!fh) return MRM_ERR_ARG;
  const size_t got = std::fread(buf.data(), sizeof(float), n_int, fh);
  std::fclose(fh);
  if ((long)got != n_int) return MRM_ERR_ARG;
  std::memset(f, 0, sizeof(double) * D.bs);
  for (int k = 0; k < D.gs[2]; k++)
    for (int j = 0; j < D.gs[1]; j++)
      for (int i = 0; i < D.gs[0]; i++)
        f[node(D, i, j, k)] = (double)buf[i + (long)D.gs[0] * (j + (long)D.gs[1] * k)];
  return MRM_OK;
}

extern "C" int mrm_io_write_var(const char* dir, const char* resfile, int v, int iter, const mrm_grid_desc* grid,
                                const double* f) {
  Dims D;
  if (!dir || !resfile || !f || v < 0 || !dims_of(grid, D)) return MRM_ERR_ARG;
  const long n_int = (long)D.gs[0] * D.gs[1] * D.gs[2];
  std::vector<float> buf(n_int);
  for (int k = 0; k < D.gs[2]; k++)
    for (int j = 0; j < D.gs[1]; j++)
      for (int i = 0; i < D.gs[0]; i++)
        buf[i + (long)D.gs[0] * (j + (long)D.gs[1] * k)] = (float)f[node(D, i, j, k)];
  FILE* fh = std::fopen(prim_name(dir, resfile, v, iter).c_str(), "wb");
  if (!fh) return MRM_ERR_ARG;
  const size_t put = std::fwrite(buf.data(), sizeof(float), n_int, fh);
  std::fclose(fh);
  return (long)put == n_int ? MRM_OK : MRM_ERR_ARG;
}

extern "C" int mrm_io_write_header(const char* dir, int iter, const mrm_grid_desc* grid, double time, double dt,
                                   double va_max) {
  if (!dir || !grid) return MRM_ERR_ARG;
  char name[64];
  std::snprintf(name, sizeof name, "Header.%06d", iter);
  FILE* fh = std::fopen((std::string(dir) + name).c_str(), "w");
  if (!fh) return MRM_ERR_ARG;
  double dx[3];
  for (int d = 0; d < 3; d++) dx[d] = (grid->gxmax[d] - grid->gxmin[d]) / grid->gsize[d];
  std::fprintf(fh, "%d  %d  %d  %f  %f  %f  %f  %f  %f", grid->gsize[0], grid->gsize[1], grid->gsize[2], dx[0],
               dx[1], dx[2], time, dt, va_max);
  std::fclose(fh);
  return MRM_OK;
}

extern "C" int mrm_io_read_header(const char* dir, int iter, int gsize[3], double dx[3], double* time,
                                  double* dt, double* va_max) {
  if (!dir || !gsize || !dx) return MRM_ERR_ARG;
  char name[64];
  std::snprintf(name, sizeof name, "Header.%06d", iter);
  std::ifstream in(std::string(dir) + name);
  if (!in) return MRM_ERR_ARG;
  double t = 0.0, d = 0.0, va = 0.0;
  in >> gsize[0] >> gsize[1] >> gsize[2] >> dx[0] >> dx[1] >> dx[2];
  if (!in) return MRM_ERR_ARG;
  in >> t >> d >> va;   // the initial Header of a run may stop after dx
  if (time) *time = in ? t : 0.0;
  if (dt) *dt = in ? d : 0.0;
  if (va_max) *va_max = in ? va : 0.0;
  return MRM_OK;
}

// backup.dat: "iter  time  p_bc  s_bc" (boundary_pdmp_1_fftw3.C:44-62)
extern "C" int mrm_io_write_backup(const char* file, int iter, double time, double p_bc, double s_bc) {
  if (!file) return MRM_ERR_ARG;
  FILE* fh = std::fopen(file, "w");
  if (!fh) return MRM_ERR_ARG;
  std::fprintf(fh, "%d  %.16lg  %lg  %lg", iter, time, p_bc, s_bc);
  std::fclose(fh);
  return MRM_OK;
}

extern "C" int mrm_io_read_backup(const char* file, int* iter, double* time, double* p_bc, double* s_bc) {
  if (!file || !iter || !time) return MRM_ERR_ARG;
  std::ifstream in(file);
  if (!in) return MRM_ERR_ARG;
  double p = -1.0, s = -1.0;
  in >> *iter >> *time;
  if (!in) return MRM_ERR_ARG;
  in >> p >> s;
  if (p_bc) *p_bc = in ? p : -1.0;
  if (s_bc) *s_bc = in ? s : -1.0;
  return MRM_OK;
}

// PSF-kernel-<c>.dat (potential_sd_fftw3.C:77-86, 129-163): a header of
// three floats (gsize[1], gsize[2], dx[0]/dx[1]), then the plane, dim 1
// fastest; the files in the order z-1, z-2, x-0, x-1, x-2, y-0, y-1, y-2
extern "C" int mrm_io_read_psf_kernels(const char* dir, int gx, int gy, float dx_ratio, float* k8) {
  if (!dir || !k8 || gx < 1 || gy < 1) return MRM_ERR_ARG;
  static const char* const names[8] = {"z-1", "z-2", "x-0", "x-1", "x-2", "y-0", "y-1", "y-2"};
  std::string d(dir);
  if (!d.empty() && d.back() != '/') d += '/';
  const size_t plane = (size_t)gx * gy;
  for (int q = 0; q < 8; q++) {
    std::ifstream in(d + "PSF-kernel-" + names[q] + ".dat", std::ios::binary);
    if (!in) return MRM_ERR_ARG;
    float h[3];
    in.read(reinterpret_cast<char*>(h), sizeof h);
    if (!in || (int)h[0] != gx || (int)h[1] != gy || h[2] != dx_ratio) return MRM_ERR_ARG;
    in.read(reinterpret_cast<char*>(k8 + q * plane), (std::streamsize)(plane * sizeof(float)));
    if (!in) return MRM_ERR_ARG;
  }
  return MRM_OK;
}
