// Microstructure analysis of one state on the GPU: the per-voxel quantities and
// the boundary / precipitate-coverage counts of the reference's post-processor
// (src/paraview.f90:188-313).
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "../../include/pf.h"
#include "pf_internal.hpp"

namespace pf {
namespace {

constexpr int kAnaBlock = 256;

// One thread per voxel (x fastest).  phi: (var, nx, ny, nz), var fastest.
__global__ void __launch_bounds__(kAnaBlock)
k_analyze(const double* __restrict__ phi, long P, int var, double threshold, double* __restrict__ phi_sq,
          double* __restrict__ phi_3, double* __restrict__ boundary, int* __restrict__ var_num,
          long* __restrict__ part) {
  __shared__ long red[2][kAnaBlock / 64];
  long nb = 0, npb = 0;
  for (long p = blockIdx.x * (long)kAnaBlock + threadIdx.x; p < P; p += (long)gridDim.x * kAnaBlock) {
    const double* v = phi + p * var;
    double sq = 0.0;
    int vmax = 0, above = 0;
    for (int i = 0; i < var; ++i) {
      const double f = v[i];
      sq += f * f;
      if (f > 0.8) vmax = i + 1;  // the last grain above 0.8
      if (f > threshold) ++above;
    }
    phi_sq[p] = sq;
    var_num[p] = vmax;
    phi_3[p] = (var >= 3) ? v[2] : 0.0;
    const bool b = above > 1;  // more than one grain present: a boundary voxel
    boundary[p] = b ? 1.0 : 0.0;
    if (b) {
      ++nb;
      if (v[var - 1] > threshold) ++npb;  // the precipitate phase on the boundary
    }
  }
  // block partials (no atomics): wave sums, then one value per block
  for (int o = 32; o > 0; o >>= 1) {
    nb += __shfl_down(nb, o, 64);
    npb += __shfl_down(npb, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = nb;
    red[1][threadIdx.x >> 6] = npb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    long a = 0, c = 0;
    for (int w = 0; w < kAnaBlock / 64; ++w) { a += red[0][w]; c += red[1][w]; }
    part[2 * blockIdx.x] = c;
    part[2 * blockIdx.x + 1] = a;
  }
}

// |grad phi_3| by central differences on interior voxels; faces are 0.
__global__ void __launch_bounds__(kAnaBlock)
k_grad_mag(const double* __restrict__ f, int nx, int ny, int nz, double* __restrict__ g) {
  const long P = (long)nx * ny * nz;
  for (long p = blockIdx.x * (long)kAnaBlock + threadIdx.x; p < P; p += (long)gridDim.x * kAnaBlock) {
    const int i = (int)(p % nx);
    const long r = p / nx;
    const int j = (int)(r % ny), k = (int)(r / ny);
    double out = 0.0;
    if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1 && k > 0 && k < nz - 1) {
      const long sx = 1, sy = nx, sz = (long)nx * ny;
      const double gx = (f[p + sx] - f[p - sx]) / 2.0;
      const double gy = (f[p + sy] - f[p - sy]) / 2.0;
      const double gz = (f[p + sz] - f[p - sz]) / 2.0;
      out = sqrt(gx * gx + gy * gy + gz * gz);
    }
    g[p] = out;
  }
}

}  // namespace
}  // namespace pf

using namespace pf;

extern "C" int pf_analyze_state(const double* phi, int nx, int ny, int nz, int var, double threshold,
                                double* phi_sq, double* phi_3, double* grad_phi_3, double* boundary,
                                int* var_num, long* counts) {
  if (!phi || !phi_sq || !phi_3 || !grad_phi_3 || !boundary || !var_num || !counts)
    return fail(PF_EINVAL, "null argument");
  if (nx < 1 || ny < 1 || nz < 1 || var < 1) return fail(PF_EINVAL, "bad grid or var");
  const long P = (long)nx * ny * nz;
  const long grid = std::min<long>((P + kAnaBlock - 1) / kAnaBlock, 16384);
  double *dphi = nullptr, *dout = nullptr;
  int* dvar = nullptr;
  long* dpart = nullptr;
  auto release = [&]() {
    if (dphi) (void)hipFree(dphi);
    if (dout) (void)hipFree(dout);
    if (dvar) (void)hipFree(dvar);
    if (dpart) (void)hipFree(dpart);
  };
#define PF_ANA(call)                                                             \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      release();                                                                 \
      return fail(PF_EHIP, std::string(#call " failed: ") + hipGetErrorString(e_)); \
    }                                                                            \
  } while (0)
  PF_ANA(hipMalloc(&dphi, sizeof(double) * (size_t)P * var));
  PF_ANA(hipMalloc(&dout, sizeof(double) * (size_t)P * 4));
  PF_ANA(hipMalloc(&dvar, sizeof(int) * (size_t)P));
  PF_ANA(hipMalloc(&dpart, sizeof(long) * 2 * (size_t)grid));
  PF_ANA(hipMemcpy(dphi, phi, sizeof(double) * (size_t)P * var, hipMemcpyHostToDevice));
  double* dsq = dout;
  double* d3 = dout + P;
  double* dg = dout + 2 * P;
  double* db = dout + 3 * P;
  hipLaunchKernelGGL(k_analyze, dim3((unsigned)grid), dim3(kAnaBlock), 0, 0, dphi, P, var, threshold, dsq, d3, db,
                     dvar, dpart);
  PF_ANA(hipGetLastError());
  hipLaunchKernelGGL(k_grad_mag, dim3((unsigned)grid), dim3(kAnaBlock), 0, 0, d3, nx, ny, nz, dg);
  PF_ANA(hipGetLastError());
  std::vector<long> part(2 * (size_t)grid);
  PF_ANA(hipMemcpy(part.data(), dpart, sizeof(long) * part.size(), hipMemcpyDeviceToHost));
  PF_ANA(hipMemcpy(phi_sq, dsq, sizeof(double) * (size_t)P, hipMemcpyDeviceToHost));
  PF_ANA(hipMemcpy(phi_3, d3, sizeof(double) * (size_t)P, hipMemcpyDeviceToHost));
  PF_ANA(hipMemcpy(grad_phi_3, dg, sizeof(double) * (size_t)P, hipMemcpyDeviceToHost));
  PF_ANA(hipMemcpy(boundary, db, sizeof(double) * (size_t)P, hipMemcpyDeviceToHost));
  PF_ANA(hipMemcpy(var_num, dvar, sizeof(int) * (size_t)P, hipMemcpyDeviceToHost));
#undef PF_ANA
  release();
  long c0 = 0, c1 = 0;
  for (long b = 0; b < grid; ++b) { c0 += part[2 * b]; c1 += part[2 * b + 1]; }
  counts[0] = c0;  // precipitate voxels on boundaries
  counts[1] = c1;  // boundary voxels
  return 0;
}
