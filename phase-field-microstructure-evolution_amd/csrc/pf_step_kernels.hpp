// Pointwise (real-space) and k-space kernels of the timestep.
#pragma once
#include "pf_fft.hpp"

namespace pf {

// Per-axis wavenumber tables (device, FP64): a_d and the Nyquist-zeroed a'_d
// for kx in [0, nx/2], ky in [0, ny), kz in [0, nz)   (src/k_space.f90:21-59).
struct KTab {
  const double *ax, *axp, *ay, *ayp, *az, *azp;
  double uz;  // 2 pi / nz: az[m] = (double)(float)m_s * uz, m_s = m or m - nz
};
// The same tables rounded to FP32, for the FP32-mixed update arithmetic.
struct KTabF {
  const float *ax, *axp, *ay, *ayp, *az, *azp;
};
template <typename R> struct KTabOf;
template <> struct KTabOf<double> { using T = KTab; };
template <> struct KTabOf<float> { using T = KTabF; };

// Per-column (kx, ky fixed) partial sums of the wavenumber combinations and the
// kz-dependent completion, in precision R: s1 = kf_sum/i, k2 = k_sq, k3 =
// k_cube/i, k4 = k_4, kp2 = |kf|^2.  Each completion adds the z term last, the
// order of the per-point expressions of src/var_diff.f90:932-935, 978-981, 1019,
// so hoisting the (x, y) parts changes no bit.
template <typename R> struct KCol {
  R s1, k2, k3, k4, kp2;  // x + y parts
};
template <typename R, typename KT>
__device__ __forceinline__ KCol<R> kcol(const KT& t, int kx, int ky) {
  const R ax = t.ax[kx], axp = t.axp[kx], ay = t.ay[ky], ayp = t.ayp[ky];
  const R x2 = ax * ax, y2 = ay * ay;
  KCol<R> c;
  c.s1 = axp + ayp;
  c.k2 = -x2 + -y2;
  c.k3 = -(axp * x2) + -(ayp * y2);
  c.k4 = (x2 * ax) * ax + (y2 * ay) * ay;
  c.kp2 = axp * axp + ayp * ayp;
  return c;
}
template <typename R> struct KCoefR {
  R s1, k2, k3, k4, kp2;
};
template <typename R, typename KT>
__device__ __forceinline__ KCoefR<R> kcoef_col(const KT& t, const KCol<R>& c, int kz) {
  const R az = t.az[kz], azp = t.azp[kz];
  const R z2 = az * az;
  KCoefR<R> k;
  k.s1 = c.s1 + azp;
  k.k2 = c.k2 + -z2;
  k.k3 = c.k3 + -(azp * z2);
  k.k4 = c.k4 + (z2 * az) * az;
  k.kp2 = c.kp2 + azp * azp;
  return k;
}

// kcoef_col with a_z and a'_z evaluated instead of read from the table.  The
// table entry is the one product (double)(float)m_s * (2 pi / nz) (kaxis); the
// product here is kept out of FMA contraction, so it is rounded on its own and
// equals the table value bit for bit.  The z-line kernel so avoids a load per
// slot whose wait would also wait for the stores issued before it.  (What the
// compiler then contracts downstream, e.g. 1 + D dt kappa k^4 into one FMA or
// not, depends on the surrounding code: the FP64 state moved by ~1e-16.)
__device__ __forceinline__ double az_of(int kz, int nz, double uz) {
#pragma clang fp contract(off)
  return (double)(float)((2 * kz <= nz) ? kz : kz - nz) * uz;
}
__device__ __forceinline__ KCoefR<double> kcoef_z(const KCol<double>& c, int kz, int nz, double uz) {
  const double az = az_of(kz, nz, uz), azp = (2 * kz == nz) ? 0.0 : az;
  const double z2 = az * az;
  KCoefR<double> k;
  k.s1 = c.s1 + azp;
  k.k2 = c.k2 + -z2;
  k.k3 = c.k3 + -(azp * z2);
  k.k4 = c.k4 + (z2 * az) * az;
  k.kp2 = c.kp2 + azp * azp;
  return k;
}

// Per-point combinations in the reference's evaluation order
// (src/var_diff.f90:932-935, 978-981, 1019):
//   kf_sum = i*s1, k_sq = k2, k_cube = i*k3, k_4 = k4;  kp2 = |kf|^2 (Nyquist-zeroed).
struct KCoef {
  double s1, k2, k3, k4, kp2;
};
__device__ __forceinline__ KCoef kcoef(const KTab& t, int kx, int ky, int kz) {
  const double ax = t.ax[kx], ay = t.ay[ky], az = t.az[kz];
  const double axp = t.axp[kx], ayp = t.ayp[ky], azp = t.azp[kz];
  KCoef k;
  k.s1 = __dadd_rn(__dadd_rn(axp, ayp), azp);
  const double x2 = __dmul_rn(ax, ax), y2 = __dmul_rn(ay, ay), z2 = __dmul_rn(az, az);
  k.k2 = __dadd_rn(__dadd_rn(-x2, -y2), -z2);
  k.k3 = __dadd_rn(__dadd_rn(-__dmul_rn(axp, x2), -__dmul_rn(ayp, y2)), -__dmul_rn(azp, z2));
  const double x4 = __dmul_rn(__dmul_rn(x2, ax), ax), y4 = __dmul_rn(__dmul_rn(y2, ay), ay),
               z4 = __dmul_rn(__dmul_rn(z2, az), az);
  k.k4 = __dadd_rn(__dadd_rn(x4, y4), z4);
  k.kp2 = __dadd_rn(__dadd_rn(__dmul_rn(axp, axp), __dmul_rn(ayp, ayp)), __dmul_rn(azp, azp));
  return k;
}

// A bank of nf fields in the split-Nyquist layout.
template <typename R> struct Bank {
  c2<R>* main;  // nf * M*ny*nz complex (== nf * nx*ny*nz reals)
  c2<R>* nyq;   // nf * ny*nz complex
};

// Geometry shared by the kernels.
struct Geo {
  int nx, ny, nz, M, var, F;
  long P, plane, hmain;  // nx*ny*nz, ny*nz, M*ny*nz
  // column-tile layout of the reduced forms' carried spectrum S: tiles of stx
  // columns (the z-line kernel's tile width), stm of them over the main block's
  // M*ny columns, stpf per field (stm + those of the ny Nyquist-plane columns)
  int stx;
  long stm, stpf;
};

// Row decode for k-space kernels: one block per (ky, kz, field) row, threads over
// kx in [0, M] (kx == M is the Nyquist entry).
struct RowIdx {
  int ky, kz, f;
};
__device__ __forceinline__ RowIdx row_decode(const Geo& g, long row) {
  RowIdx r;
  r.ky = (int)(row % g.ny);
  const long t = row / g.ny;
  r.kz = (int)(t % g.nz);
  r.f = (int)(t / g.nz);
  return r;
}
template <typename R>
__device__ __forceinline__ c2<R>* elem(const Bank<R>& b, const Geo& g, int f, int kx, int ky, int kz) {
  if (kx < g.M) return b.main + (long)f * g.hmain + ((long)ky + (long)g.ny * kz) * g.M + kx;
  return b.nyq + (long)f * g.plane + ky + (long)g.ny * kz;
}
// The same element of a spectrum in the column-tile layout (all of it in b.main).
template <typename R>
__device__ __forceinline__ c2<R>* elem_tiled(const Bank<R>& b, const Geo& g, int f, int kx, int ky, int kz) {
  const long col = (kx < g.M) ? (long)ky * g.M + kx : (long)ky;
  const long tt = (kx < g.M ? 0 : g.stm) + col / g.stx;
  return b.main + (((long)f * g.stpf + tt) * g.nz + kz) * g.stx + col % g.stx;
}

// Deterministic block reduction helpers (fixed tree order).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}
template <int NT, bool MAX>
__device__ __forceinline__ double block_ext(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int s = NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = sh[threadIdx.x + s];
      if (MAX ? (o > sh[threadIdx.x]) : (o < sh[threadIdx.x])) sh[threadIdx.x] = o;
    }
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

}  // namespace pf
