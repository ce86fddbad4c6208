// The timestep: handle, state I/O, the FP64 faithful sequence and the reduced
// (spectral-state) FP64 / FP32-mixed sequences, diagnostics and stage timers.
//
// Field order inside the library: field 0 = c, fields 1..var = eta_1..eta_var.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pf.h"
#include "pf_internal.hpp"
#include "pf_fused.hpp"
#include "pf_step_kernels.hpp"

namespace pf {
double kaxis(int m, int n, bool zero_nyquist);

enum Stage {
  ST_NONLINEAR = 0, ST_X_R2C, ST_Y_FWD, ST_Z_FWD, ST_UPDATE, ST_Z_INV, ST_Y_INV, ST_X_C2R,
  ST_DIAG, ST_MOBILITY, ST_GRAD, NSTAGE
};
static const char* kStageNames[NSTAGE] = {
    "nonlinear", "x_r2c", "y_fwd", "z_fwd", "kspace_update", "z_inv", "y_inv", "x_c2r",
    "diag_reduce", "mobility", "grad_energy"};

constexpr int NB = 256;        // block size of pointwise kernels
constexpr int NPART_NL = 6;    // f_ch, f_eta, c_max, c_min, phi_min, phi_3
constexpr long kMaxGrid = 8192;
// The z-line kernel streams four banks with short (TX-column) row segments;
// more resident tiles per CU keep more rows in flight (measured 2.5% faster
// at 32768 than at 8192 blocks, 256^3 x 17 fields).
constexpr long kZlineMaxGrid = 32768;
// Largest grids a tuning may ask for (the block-partial buffers are sized for them).
constexpr long kXlineGridCap = 16384;
constexpr long kZlineGridCap = 65536;

using UCoef = UPar;

// ===========================================================================
// Kernels
// ===========================================================================

// AoS host layout (phi var-fastest, con) <-> bank of real fields (c, eta_1..).
template <typename R>
__global__ void k_scatter_state(const double* __restrict__ phi, const double* __restrict__ con,
                                R* __restrict__ dst, long P, int var) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < P; t += (long)gridDim.x * blockDim.x) {
    dst[t] = (R)con[t];
    for (int i = 0; i < var; ++i) dst[(long)(i + 1) * P + t] = (R)phi[t * var + i];
  }
}
template <typename R>
__global__ void k_gather_state(const R* __restrict__ src, double* __restrict__ phi,
                               double* __restrict__ con, long P, int var) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < P; t += (long)gridDim.x * blockDim.x) {
    con[t] = (double)src[t];
    for (int i = 0; i < var; ++i) phi[t * var + i] = (double)src[(long)(i + 1) * P + t];
  }
}

// Voronoi polycrystal built on the device (the synthetic initial condition of
// pf_ic_voronoi, same draw order): point t = (i, j, k), x fastest, belongs to
// the nearest seed by periodic minimum-image distance, the lowest index on ties;
// eta_{(s mod var)+1} = 1 there and the other eta 0; c is the host's draw.  The
// distances use the host generator's FP64 operations in its order, with no
// contraction into FMAs, so every owner equals the host's.  counts[v] gains the
// number of points owned by field v + 1 (exact: the field means).
__device__ __forceinline__ double vor_wrap(double d, int n) {
  d = fabs(d);
  const double alt = (double)n - d;
  return alt < d ? alt : d;
}
template <typename R>
__global__ void __launch_bounds__(NB) k_voronoi_state(Geo g, const double* __restrict__ seeds, int nseeds,
                                                      const double* __restrict__ con, R* __restrict__ dst,
                                                      unsigned long long* __restrict__ counts) {
  extern __shared__ unsigned int hist[];  // g.var counters
  for (int v = threadIdx.x; v < g.var; v += NB) hist[v] = 0u;
  __syncthreads();
  for (long t = (long)blockIdx.x * NB + threadIdx.x; t < g.P; t += (long)gridDim.x * NB) {
    const int i = (int)(t % g.nx), j = (int)((t / g.nx) % g.ny), k = (int)(t / ((long)g.nx * g.ny));
    int best = 0;
    double bd = 0.0;
    for (int s = 0; s < nseeds; ++s) {
#pragma clang fp contract(off)
      const double dx = vor_wrap((double)i - seeds[3 * s], g.nx);
      const double dy = vor_wrap((double)j - seeds[3 * s + 1], g.ny);
      const double dz = vor_wrap((double)k - seeds[3 * s + 2], g.nz);
      const double d2 = (dx * dx + dy * dy) + dz * dz;
      if (s == 0 || d2 < bd) {
        bd = d2;
        best = s;
      }
    }
    const int own = best % g.var;
    dst[t] = (R)con[t];
    for (int v = 0; v < g.var; ++v) dst[(long)(v + 1) * g.P + t] = (v == own) ? R(1) : R(0);
    atomicAdd(&hist[own], 1u);
  }
  __syncthreads();
  for (int v = threadIdx.x; v < g.var; v += NB)
    if (hist[v]) atomicAdd(&counts[v], (unsigned long long)hist[v]);
}

// ---------------------------------------------------------------------------
// Loops 5-10, faithful form (src/var_diff.f90:685-905), FP64, O(var^2).
// Reads the real state bank (c, eta_i); writes df_dc, df_dphi_i, mob_c.
// ---------------------------------------------------------------------------
// gamma_ij of src/var_diff.f90:206-218 (the matrix pf_gamma builds): sigma_1
// between two distinct matrix grains, sigma_2 between a grain and the
// precipitate field eta_var (index mat), 0 on the diagonal.  Selected from (i, j)
// instead of loaded: the same double values, without 2 V^2 loads per point.
__device__ __forceinline__ double gam_ij(const UCoef& u, int mat, int i, int j) {
  return (i == j) ? 0.0 : ((i == mat || j == mat) ? u.s2g : u.s1g);
}

template <bool DIAG>
__global__ void __launch_bounds__(NB)
k_nl_faithful(Geo g, UCoef u, const double* __restrict__ st,
              double* __restrict__ dfdc, double* __restrict__ dfdphi, double* __restrict__ mob,
              double* __restrict__ part) {
  extern __shared__ double sh[];  // var * NB values, then NB reduction scratch
  double* red = sh + (long)g.var * NB;
  const int V = g.var, mat = V - 1;
  double a_ch = 0.0, a_eta = 0.0, cmax = 0.0, cmin = 1.0, pmin = 1.0, p3 = 0.0;
  for (long base = (long)blockIdx.x * NB; base < g.P; base += (long)gridDim.x * NB) {
    const long t = base + threadIdx.x;
    const bool ok = t < g.P;
    __syncthreads();
    for (int i = 0; i < V; ++i) sh[i * NB + threadIdx.x] = ok ? st[(long)(i + 1) * g.P + t] : 0.0;
    __syncthreads();
    if (!ok) continue;
    const double* ph = sh + threadIdx.x;
    const double con = st[t];
    // loop 5
    double fe = 0.0;
    for (int i = 0; i < V; ++i) {
      const double fi = ph[i * NB], fi2 = fi * fi;
      fe = (fe - 0.5 * fi2) + 0.25 * (fi2 * fi2);
      for (int j = i + 1; j < V; ++j) {
        const double fj = ph[j * NB];
        fe = fe + ((gam_ij(u, mat, i, j) * fi2) * (fj * fj));
      }
    }
    // loop 6
    double num = 0.0, den = 0.0;
    for (int i = 0; i < V; ++i) {
      const double f2 = ph[i * NB] * ph[i * NB], f4 = f2 * f2;
      if (i < mat) num = num + f4;
      den = den + f4;
    }
    const double hm = num / den, hp = 1.0 - hm;
    const double dm = con - u.cm, dp = con - u.cp;
    const double fch = (hm * u.am) * (dm * dm) + (hp * u.ap) * (dp * dp);
    dfdc[t] = ((2.0 * hm) * u.am) * dm + ((2.0 * hp) * u.ap) * dp;
    // loops 7-9
    const double fm = u.am * (dm * dm), fp = u.ap * (dp * dp);
    const double fv = ph[mat * NB], fv2 = fv * fv, fv4 = fv2 * fv2;
    const double t2v = (4.0 * (fv2 * fv)) * (den - fv4);
    const double den2 = den * den;
    double ps = 0.0;
    for (int i = 0; i < V; ++i) {
      const double fi = ph[i * NB];
      const double term1 = (fi * fi) * fi - fi;
      double term2 = 0.0;
      for (int j = 0; j < V; ++j) {
        if (j == i) continue;
        const double fj = ph[j * NB];
        term2 = term2 + ((2.0 * gam_ij(u, mat, i, j)) * fi) * (fj * fj);
      }
      const double term4 = (i < mat) ? (((4.0 * ((fi * fi) * fi)) * fv4) * (fm - fp)) / den2
                                     : (t2v * (fp - fm)) / den2;
      dfdphi[(long)i * g.P + t] = (term1 + term2) + term4;
      ps = ps + fi * fi;
    }
    // loop 10: mob_c = D_bulk(var) + 2(1 - sum phi^2) D_gb, D_bulk(var) = D_L
    mob[t] = u.D + ((2.0 * (1.0 - ps)) * u.Dgb);
    if (DIAG) {
      a_ch += fch;
      a_eta += fe;
      if (fv > p3) p3 = fv;
      if (con <= cmin) cmin = con;
      if (con > cmax) cmax = con;
      if (ps < pmin) pmin = ps;
    }
  }
  if (DIAG) {
    __syncthreads();
    const double s0 = block_sum<NB>(a_ch, red), s1 = block_sum<NB>(a_eta, red);
    const double e0 = block_ext<NB, true>(cmax, red), e1 = block_ext<NB, false>(cmin, red);
    const double e2 = block_ext<NB, false>(pmin, red), e3 = block_ext<NB, true>(p3, red);
    if (threadIdx.x == 0) {
      double* o = part + (long)blockIdx.x * NPART_NL;
      o[0] = s0; o[1] = s1; o[2] = e0; o[3] = e1; o[4] = e2; o[5] = e3;
    }
  }
}

// ---------------------------------------------------------------------------
// Loops 5-10, reduced form: O(var) per point using the block structure of gamma
// (sigma_1 between matrix grains, sigma_2 between a grain and eta_var,
// src/var_diff.f90:206-218).  In place on the real work bank:
// (c, eta_1..eta_V) -> (df_dc, df_dphi_1..df_dphi_V); optional (m - D_mean).
// ---------------------------------------------------------------------------
template <typename R, bool DIAG, bool MOB, bool WRITE = true>
__global__ void __launch_bounds__(NB)
k_nl_reduced(Geo g, UCoef u, R* __restrict__ w, R* __restrict__ mobd, double* __restrict__ part) {
  extern __shared__ unsigned char shraw[];
  R* sh = reinterpret_cast<R*>(shraw);
  double* red = reinterpret_cast<double*>(shraw + sizeof(R) * (size_t)g.var * NB + 16);
  const int V = g.var, mat = V - 1;
  const R am = (R)u.am, ap = (R)u.ap, cm = (R)u.cm, cp = (R)u.cp, s1g = (R)u.s1g, s2g = (R)u.s2g;
  double a_ch = 0.0, a_eta = 0.0, cmax = 0.0, cmin = 1.0, pmin = 1.0, p3 = 0.0;
  for (long base = (long)blockIdx.x * NB; base < g.P; base += (long)gridDim.x * NB) {
    const long t = base + threadIdx.x;
    const bool ok = t < g.P;
    __syncthreads();
    for (int i = 0; i < V; ++i) sh[i * NB + threadIdx.x] = ok ? w[(long)(i + 1) * g.P + t] : R(0);
    __syncthreads();
    if (!ok) continue;
    const R* ph = sh + threadIdx.x;
    const R con = w[t];
    R sm = 0, num = 0, pair = 0;
    R fe = 0;
    for (int i = 0; i < mat; ++i) {
      const R a = ph[i * NB] * ph[i * NB];
      pair += a * sm;
      sm += a;
      num += a * a;
      if (DIAG) fe += a * (R(0.25) * a - R(0.5));
    }
    const R fv = ph[mat * NB], av = fv * fv, av2 = av * av;
    const R d4 = num + av2;
    const R hm = num / d4, hp = R(1) - hm;
    const R dm = con - cm, dp = con - cp;
    const R fm = am * (dm * dm), fpp = ap * (dp * dp);
    if (WRITE) w[t] = R(2) * hm * am * dm + R(2) * hp * ap * dp;
    const R inv_d4 = R(1) / d4;
    const R cT4m = R(4) * (fm - fpp) * (av2 * inv_d4) * inv_d4;
    if (WRITE) {
      for (int i = 0; i < mat; ++i) {
        const R fi = ph[i * NB], a = fi * fi;
        const R ex = s1g * (sm - a) + s2g * av;
        w[(long)(i + 1) * g.P + t] = fi * (a - R(1)) + R(2) * fi * ex + cT4m * fi * a;
      }
      const R T4v = R(4) * fv * av * hm * (fpp - fm) * inv_d4;
      w[(long)V * g.P + t] = fv * (av - R(1)) + R(2) * fv * s2g * sm + T4v;
    }
    const R ps = sm + av;
    if (MOB && WRITE) mobd[t] = R(2) * (R(1) - ps) * (R)u.Dgb;
    if (DIAG) {
      fe += av * (R(0.25) * av - R(0.5)) + s1g * pair + s2g * av * sm;
      a_ch += (double)(hm * fm + hp * fpp);
      a_eta += (double)fe;
      const double dfv = fv, dc = con, dps = ps;
      if (dfv > p3) p3 = dfv;
      if (dc <= cmin) cmin = dc;
      if (dc > cmax) cmax = dc;
      if (dps < pmin) pmin = dps;
    }
  }
  if (DIAG) {
    __syncthreads();
    const double s0 = block_sum<NB>(a_ch, red), s1 = block_sum<NB>(a_eta, red);
    const double e0 = block_ext<NB, true>(cmax, red), e1 = block_ext<NB, false>(cmin, red);
    const double e2 = block_ext<NB, false>(pmin, red), e3 = block_ext<NB, true>(p3, red);
    if (threadIdx.x == 0) {
      double* o = part + (long)blockIdx.x * NPART_NL;
      o[0] = s0; o[1] = s1; o[2] = e0; o[3] = e1; o[4] = e2; o[5] = e3;
    }
  }
}

// ---------------------------------------------------------------------------
// Faithful k-space kernels (FP64).  One block per (ky, kz, field) row, threads
// over kx in [0, M] (kx = M: Nyquist plane).  Grid-stride over rows.
// ---------------------------------------------------------------------------

#define PF_FAITHFUL_SIZES(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(160) X(192) X(256) X(512)

// Loops 1 and 3 (dft_grad_d = dft * kf_d, src/var_diff.f90:577-579, 645-647)
// fused with the first inverse pass of the gradients: for
// every z column of state field f, the three gradient spectra i a'_d u^
// (a' = Nyquist-zeroed wavenumbers) are formed in registers, transformed back
// along z and stored as gradient fields 3f+d; the spectra themselves are never
// stored.  NYQ: the Nyquist planes (columns = ky, a'_x = 0).
template <int NZ, bool NYQ>
__global__ void __launch_bounds__(col_tile(NZ, sizeof(double2)) * plan_T(NZ))
k_grad_zinv(Geo g, KTab kt, const double2* __restrict__ src, double2* __restrict__ dst, long W,
            long tiles_per_field, const double2* __restrict__ tw) {
  constexpr int E = plan_E(NZ);
  constexpr int TX = col_tile(NZ, sizeof(double2));
  __shared__ double2 sh[(plan_np(NZ) > 1) ? NZ * TX : 1];
  const long f = blockIdx.x / tiles_per_field;
  const long col0 = (blockIdx.x - f * tiles_per_field) * TX;
  const int ncol = (int)((W - col0) < TX ? (W - col0) : TX);
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const bool act = c < ncol;
  const long col = col0 + (act ? c : 0);
  const double ax = NYQ ? kt.axp[g.M] : kt.axp[col % g.M];
  const double ay = NYQ ? kt.ayp[col] : kt.ayp[col / g.M];
  const double2* in = src + f * NZ * W + col;
  double2 u[E], v[E];
#pragma unroll
  for (int s = 0; s < kInSlots<NZ>; ++s)
    if (act && rfft_in_valid<NZ>(t, s)) u[s] = in[(long)rfft_in_idx<NZ>(t, s) * W];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
#pragma unroll
    for (int s = 0; s < kInSlots<NZ>; ++s) {
      if (!(act && rfft_in_valid<NZ>(t, s))) continue;
      const double a = (d == 0) ? ax : (d == 1) ? ay : kt.azp[rfft_in_idx<NZ>(t, s)];
      v[s] = make_double2(-(u[s].y * a), u[s].x * a);
    }
    if (d > 0) __syncthreads();  // the exchange tile is reused
    rfft<double, NZ, true, TX>(v, t, act, sh + c, tw);
    double2* out = dst + (3 * f + d) * NZ * W + col;
#pragma unroll
    for (int s = 0; s < kOutSlots<NZ>; ++s)
      if (act && rfft_out_valid<NZ>(t, s)) out[(long)rfft_out_idx<NZ>(t, s) * W] = v[s];
  }
}


// loop 11: dft_dummy = kf_sum*dft_df_dc - grad_coeff_c*dft_con*k_cube  (src/var_diff.f90:932-936)
template <typename R>
__global__ void k_ghat(Geo g, UCoef u, KTab kt, Bank<R> fhat, int ff, Bank<R> chat, int cf,
                       Bank<R> out, long nrows) {
  for (long row = blockIdx.x; row < nrows; row += gridDim.x) {
    const RowIdx r = row_decode(g, row);
    for (int kx = threadIdx.x; kx <= g.M; kx += blockDim.x) {
      const KCoef k = kcoef(kt, kx, r.ky, r.kz);
      const c2<R> f = *elem(fhat, g, ff, kx, r.ky, r.kz);
      const c2<R> c = *elem(chat, g, cf, kx, r.ky, r.kz);
      const double kcr = u.kc * (double)c.x, kci = u.kc * (double)c.y;
      const double re = -(k.s1 * (double)f.y) - (-(kci * k.k3));
      const double im = (k.s1 * (double)f.x) - (kcr * k.k3);
      *elem(out, g, 0, kx, r.ky, r.kz) = mk<R>((R)re, (R)im);
    }
  }
}

// loop 12: dummy = dummy*(mob_c - D_mean)  (src/var_diff.f90:956)
__global__ void k_mult_faithful(long P, double D, double* __restrict__ x, const double* __restrict__ mob) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < P; t += (long)gridDim.x * blockDim.x)
    x[t] = x[t] * (mob[t] - D);
}
template <typename R>
__global__ void k_mult(long P, R* __restrict__ x, const R* __restrict__ m) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < P; t += (long)gridDim.x * blockDim.x)
    x[t] = x[t] * m[t];
}

// loop 13: dft_con = (dft_con + term11 + term33) / term22  (src/var_diff.f90:978-987)
__global__ void k_ch_faithful(Geo g, UCoef u, KTab kt, Bank<double> st, Bank<double> hhat,
                              Bank<double> fhat, long nrows) {
  for (long row = blockIdx.x; row < nrows; row += gridDim.x) {
    const RowIdx r = row_decode(g, row);
    for (int kx = threadIdx.x; kx <= g.M; kx += blockDim.x) {
      const KCoef k = kcoef(kt, kx, r.ky, r.kz);
      double2* cp = elem(st, g, 0, kx, r.ky, r.kz);
      const double2 c = *cp, h = *elem(hhat, g, 0, kx, r.ky, r.kz), f = *elem(fhat, g, 0, kx, r.ky, r.kz);
      const double t11r = -((h.y * u.dt) * k.s1), t11i = (h.x * u.dt) * k.s1;
      const double t22 = 1.0 + ((u.D * u.dt) * u.kc) * k.k4;
      const double r33 = (u.D * k.k2) * u.dt;
      const double nr = (c.x + t11r) + r33 * f.x, ni = (c.y + t11i) + r33 * f.y;
      *cp = make_double2(nr / t22, ni / t22);
    }
  }
}

// loop 14: dft_dummy = (dft_dummy - mob_phi*t_step*dft_dummy2) / (1 - mob_phi*kappa*t_step*k_sq)
// (src/var_diff.f90:1019-1027); fields 1..var of the state bank, df_dphi bank 0..var-1.
__global__ void k_ac_faithful(Geo g, UCoef u, KTab kt, Bank<double> st, Bank<double> dpsi, long nrows) {
  const double ldt = u.L * u.dt, lkd = (u.L * u.kphi) * u.dt;
  for (long row = blockIdx.x; row < nrows; row += gridDim.x) {
    const RowIdx r = row_decode(g, row);  // r.f in [0, var)
    for (int kx = threadIdx.x; kx <= g.M; kx += blockDim.x) {
      const KCoef k = kcoef(kt, kx, r.ky, r.kz);
      double2* sp = elem(st, g, r.f + 1, kx, r.ky, r.kz);
      const double2 s = *sp, p = *elem(dpsi, g, r.f, kx, r.ky, r.kz);
      const double t22 = 1.0 - lkd * k.k2;
      *sp = make_double2((s.x - ldt * p.x) / t22, (s.y - ldt * p.y) / t22);
    }
  }
}

// ---------------------------------------------------------------------------
// Reduced k-space update for all fields, in place on S, copy into W for the
// inverse transform; optional Parseval partial sums of the new state for the
// next step's f_int.  FP32 form keeps each field's k = 0 mode in FP64 (dc64).
// STILED: S is in the column-tile layout (elem_tiled).
// ---------------------------------------------------------------------------
template <typename R, bool DC64, bool MOB, bool PARSEVAL, bool UPD, bool STILED = false>
__global__ void __launch_bounds__(NB)
k_update(Geo g, UCoef u, KTab kt, Bank<R> S, Bank<R> W, Bank<R> H, double* __restrict__ dc64,
         double* __restrict__ part, long nrows) {
  __shared__ double red[NB];
  double acc = 0.0;
  for (long row = blockIdx.x; row < nrows; row += gridDim.x) {
    const RowIdx r = row_decode(g, row);
    const double kap = (r.f == 0) ? u.kc : u.kphi;
    for (int kx = threadIdx.x; kx <= g.M; kx += blockDim.x) {
      const KCoef k = kcoef(kt, kx, r.ky, r.kz);
      c2<R>* sp = STILED ? elem_tiled(S, g, r.f, kx, r.ky, r.kz) : elem(S, g, r.f, kx, r.ky, r.kz);
      c2<R> sn = *sp;
      if (UPD) {
        const c2<R> w = *elem(W, g, r.f, kx, r.ky, r.kz);
        if (r.f == 0) {
          const double den = 1.0 + ((u.D * u.dt) * u.kc) * k.k4;
          const double A = 1.0 / den, B = ((u.D * k.k2) * u.dt) * A;
          double re = A * (double)sn.x + B * (double)w.x;
          double im = A * (double)sn.y + B * (double)w.y;
          if (MOB) {
            const c2<R> h = *elem(H, g, 0, kx, r.ky, r.kz);
            const double C = (u.dt * k.s1) * A;
            re += -C * (double)h.y;
            im += C * (double)h.x;
          }
          sn = mk<R>((R)re, (R)im);
        } else {
          const double A = 1.0 / (1.0 - ((u.L * u.kphi) * u.dt) * k.k2);
          const double B = -(u.L * u.dt) * A;
          sn = mk<R>((R)(A * (double)sn.x + B * (double)w.x), (R)(A * (double)sn.y + B * (double)w.y));
        }
        if (DC64 && kx == 0 && r.ky == 0 && r.kz == 0) {
          if (r.f != 0) dc64[r.f] = dc64[r.f] - (u.L * u.dt) * (double)w.x;
          sn = mk<R>((R)dc64[r.f], R(0));
        }
        *sp = sn;
        *elem(W, g, r.f, kx, r.ky, r.kz) = sn;
      }
      if (PARSEVAL) {
        const double wgt = (kx == 0 || kx == g.M) ? 1.0 : 2.0;
        const double m2 = (double)sn.x * (double)sn.x + (double)sn.y * (double)sn.y;
        acc += (0.5 * kap) * wgt * k.kp2 * m2;
      }
    }
  }
  if (PARSEVAL) {
    const double s = block_sum<NB>(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s * (double)g.P;
  }
}

// ---------------------------------------------------------------------------
// A spectrum in the natural layout (main block, Nyquist planes) -> the
// column-tile layout of S; the tiles' padding columns are zero.
// ---------------------------------------------------------------------------
template <typename R>
__global__ void __launch_bounds__(NB) k_tile_spectrum(Geo g, const c2<R>* __restrict__ main,
                                                      const c2<R>* __restrict__ nyq, c2<R>* __restrict__ st, long n) {
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < n; o += (long)gridDim.x * blockDim.x) {
    const long c = o % g.stx, rest = o / g.stx;
    const long kz = rest % g.nz, tf = rest / g.nz;
    const long f = tf / g.stpf, tt = tf % g.stpf;
    c2<R> v = mk<R>(R(0), R(0));
    if (tt < g.stm) {
      const long col = tt * g.stx + c;
      if (col < (long)g.M * g.ny) v = main[f * g.hmain + kz * ((long)g.M * g.ny) + col];
    } else {
      const long col = (tt - g.stm) * g.stx + c;
      if (col < g.ny) v = nyq[f * g.plane + kz * g.ny + col];
    }
    st[o] = v;
  }
}

// ---------------------------------------------------------------------------
// Final diagnostics reduction (one block): Neumaier-compensated sums over the
// block partials in fixed order, extrema with the reference's initial values.
// ---------------------------------------------------------------------------
struct DiagSrc {
  const double* nl; long n_nl;  // NPART_NL per block
  const double* fi; long n_fi;  // one per block
  const double* cmean;          // device pointer to mean(c), or null
  int cmean_is_float;
  const float* cmean_f;
};
__device__ __forceinline__ void neum(double& s, double& c, double x) {
  const double t = s + x;
  if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
  s = t;
}
// out: the row, or (ctr != null) the base of the rows: row *ctr is written and
// *ctr advanced by ctr_step, so that a captured step writes each replay's row.
// One block of kDiagThreads: the partials of a 256^3 step (8192 x 6 from the
// x line, 32768 from the z line) are read by 1024 threads instead of 256.
constexpr int kDiagThreads = 1024;
__global__ void __launch_bounds__(kDiagThreads) k_diag_final(DiagSrc d, double* __restrict__ out, int* __restrict__ ctr,
                                                   int ctr_step) {
  __shared__ double red[kDiagThreads];
  double sch = 0, cch = 0, set = 0, cet = 0, sfi = 0, cfi = 0;
  double cmax = 0.0, cmin = 1.0, pmin = 1.0, p3 = 0.0;
  for (long b = threadIdx.x; b < d.n_nl; b += kDiagThreads) {
    const double* o = d.nl + b * NPART_NL;
    neum(sch, cch, o[0]);
    neum(set, cet, o[1]);
    if (o[2] > cmax) cmax = o[2];
    if (o[3] <= cmin) cmin = o[3];
    if (o[4] < pmin) pmin = o[4];
    if (o[5] > p3) p3 = o[5];
  }
  for (long b = threadIdx.x; b < d.n_fi; b += kDiagThreads) neum(sfi, cfi, d.fi[b]);
  const double fch = block_sum<kDiagThreads>(sch + cch, red);
  const double feta = block_sum<kDiagThreads>(set + cet, red);
  const double fint = block_sum<kDiagThreads>(sfi + cfi, red);
  const double e0 = block_ext<kDiagThreads, true>(cmax, red), e1 = block_ext<kDiagThreads, false>(cmin, red);
  const double e2 = block_ext<kDiagThreads, false>(pmin, red), e3 = block_ext<kDiagThreads, true>(p3, red);
  if (threadIdx.x == 0) {
    if (ctr) {
      const int row = *ctr;
      out += 9 * (long)row;
      *ctr = row + ctr_step;
    }
    out[0] = fint; out[1] = fch; out[2] = feta; out[3] = (fint + fch) + feta;
    out[4] = e0; out[5] = e1; out[6] = e2; out[7] = e3;
    out[8] = d.cmean_is_float ? (double)*d.cmean_f : (d.cmean ? *d.cmean : 0.0);
  }
}

__global__ void k_set_int(int* __restrict__ p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// p[0 .. n) = 0 (a kernel rather than a memset, so it may sit in a captured graph)
__global__ void k_fill_zero(double* __restrict__ p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

}  // namespace pf

// ===========================================================================
// Handle
// ===========================================================================
using namespace pf;

struct pf_handle {
  Geo g{};
  int mode = 0;
  pf_params p{};
  UCoef u{};
  hipStream_t s = nullptr;
  int dev = 0;
  long step = 0;  // completed steps

  double* dk = nullptr;
  float* dkf = nullptr;
  KTab kt{};
  KTabF ktf{};

  // faithful (FP64) banks
  Bank<double> fs{}, fgrad{}, fdfdc{}, fdum{}, fpsi{};
  double* fmob = nullptr;
  // reduced banks (precision R, stored untyped)
  void *S_main = nullptr, *W_main = nullptr, *W_nyq = nullptr;  // S: column-tile layout
  void *G_main = nullptr, *G_nyq = nullptr, *mob = nullptr;
  void *Fc_main = nullptr, *Fc_nyq = nullptr, *X_main = nullptr, *X_nyq = nullptr;
  double* dc64 = nullptr;

  double* part_nl = nullptr;
  double* part_fi = nullptr;
  long n_part_nl = 0, n_part_fi = 0;
  bool fint_ready = false;  // part_fi holds f_int of the current state
  double* ddiag = nullptr;  // 9 doubles per record
  long ddiag_cap = 0;

  // Captured step period (hipGraph): gx replays gx_period steps whose pattern
  // of diagnostics is that of diag_every = gx_key (0: none).  While capturing,
  // diagnostic rows go through the device counter dctr.
  bool graphs = true;
  hipGraphExec_t gx = nullptr;
  int gx_key = -1, gx_period = 0;
  int* dctr = nullptr;
  bool capturing = false;
  int cap_step = 0;

  double* staging = nullptr;  // (var+1)*P doubles for state I/O

  // launch shapes of the fused kernels (pf_set_tuning; from the tuned table at
  // pf_create): x-line lines per block (0 = automatic) and grid, z-line grid
  // and XCD run length in main-block rows
  pf_tuning tune{0, 0, 0, -1, 0};

  std::vector<void*> allocs;

  // timing
  bool timing = false;
  unsigned timing_mask = ~0u;  // stages timed while timing is on
  bool ev_open = false;        // tbeg recorded a start event that tend must close
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_stage;
  size_t ev_used = 0;
  double st_ms[NSTAGE] = {};
  long st_cnt[NSTAGE] = {};
};

// Makes the handle's device current for the duration of an entry point and
// restores the caller's device afterwards: kernels, streams and allocations of
// a handle all belong to the device it was created on.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

namespace {

template <typename R> typename KTabOf<R>::T ktab(const pf_handle* h);
template <> KTab ktab<double>(const pf_handle* h) { return h->kt; }
template <> KTabF ktab<float>(const pf_handle* h) { return h->ktf; }

int dalloc(pf_handle* h, void** p, size_t bytes) {
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(PF_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
  }
  h->allocs.push_back(*p);
  return 0;
}
template <typename R> int alloc_bank(pf_handle* h, Bank<R>& b, int nf) {
  int rc = dalloc(h, (void**)&b.main, sizeof(c2<R>) * (size_t)h->g.hmain * nf);
  if (!rc) rc = dalloc(h, (void**)&b.nyq, sizeof(c2<R>) * (size_t)h->g.plane * nf);
  return rc;
}
template <typename R> Bank<R> sub(const Bank<R>& b, const Geo& g, int f) {
  return Bank<R>{b.main + (long)f * g.hmain, b.nyq + (long)f * g.plane};
}

// ---- timers ----
void tbeg(pf_handle* h, int stage) {
  if (!h->timing) return;
  h->ev_open = (h->timing_mask >> stage) & 1;
  if (!h->ev_open) return;
  if (h->ev_used + 2 > h->ev.size()) {
    for (int i = 0; i < 64; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) { h->timing = false; return; }
      h->ev.push_back(e);
    }
    h->ev_stage.resize(h->ev.size());
  }
  h->ev_stage[h->ev_used] = stage;
  (void)hipEventRecord(h->ev[h->ev_used], h->s);
}
void tend(pf_handle* h) {
  if (!h->timing || !h->ev_open) return;
  h->ev_open = false;
  (void)hipEventRecord(h->ev[h->ev_used + 1], h->s);
  h->ev_used += 2;
}
void tflush(pf_handle* h) {
  if (h->ev_used == 0) return;
  (void)hipEventSynchronize(h->ev[h->ev_used - 1]);
  for (size_t i = 0; i < h->ev_used; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1]) == hipSuccess) {
      h->st_ms[h->ev_stage[i]] += ms;
      h->st_cnt[h->ev_stage[i]] += 1;
    }
  }
  h->ev_used = 0;
}

unsigned grid_for(long n, long per) { return (unsigned)std::min<long>(kMaxGrid, std::max<long>(1, (n + per - 1) / per)); }
unsigned row_block(const Geo& g) {
  int b = ((g.M + 1 + 63) / 64) * 64;
  return (unsigned)std::min(b, NB);
}

// Final diag reduction into ddiag record `idx`.
int diag_final(pf_handle* h, long idx, bool with_nl, bool with_fi, const double* cmean, const float* cmean_f) {
  DiagSrc d{};
  d.nl = h->part_nl; d.n_nl = with_nl ? h->n_part_nl : 0;
  d.fi = h->part_fi; d.n_fi = with_fi ? h->n_part_fi : 0;
  d.cmean = cmean;
  d.cmean_f = cmean_f;
  d.cmean_is_float = cmean_f != nullptr;
  if (h->capturing)
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kDiagThreads), 0, h->s, d, h->ddiag, h->dctr, h->cap_step);
  else
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(kDiagThreads), 0, h->s, d, h->ddiag + 9 * idx, (int*)nullptr, 0);
  return 0;
}

void drop_graph(pf_handle* h) {
  if (h->gx) (void)hipGraphExecDestroy(h->gx);
  h->gx = nullptr;
  h->gx_key = -1;
  h->gx_period = 0;
}

int ensure_diag(pf_handle* h, long n) {
  if (n <= h->ddiag_cap) return 0;
  drop_graph(h);  // a captured period holds the old row buffer
  if (h->ddiag) {
    (void)hipFree(h->ddiag);
    h->allocs.erase(std::remove(h->allocs.begin(), h->allocs.end(), (void*)h->ddiag), h->allocs.end());
  }
  h->ddiag = nullptr;
  h->ddiag_cap = 0;
  int rc = dalloc(h, (void**)&h->ddiag, sizeof(double) * 9 * (size_t)n);
  if (!rc) h->ddiag_cap = n;
  return rc;
}

// ---------------------------------------------------------------------------
// Faithful step (FP64): the reference's loop sequence.
// ---------------------------------------------------------------------------
int launch_grad_zinv(pf_handle* h) {
  const Geo& g = h->g;
  const double2* tw = twiddle_table<double>(g.nz);
  if (!tw) return fail(PF_EHIP, "twiddle table allocation failed");
  switch (g.nz) {
#define PF_CASE(N)                                                                                      \
  case N: {                                                                                             \
    constexpr int TX = col_tile(N, sizeof(double2));                                                    \
    const long Wm = (long)g.M * g.ny, tm = (Wm + TX - 1) / TX, tn = ((long)g.ny + TX - 1) / TX;         \
    hipLaunchKernelGGL((k_grad_zinv<N, false>), dim3((unsigned)(tm * g.F)), dim3(TX * plan_T(N)), 0,    \
                       h->s, g, h->kt, h->fs.main, h->fgrad.main, Wm, tm, tw);                           \
    hipLaunchKernelGGL((k_grad_zinv<N, true>), dim3((unsigned)(tn * g.F)), dim3(TX * plan_T(N)), 0,     \
                       h->s, g, h->kt, h->fs.nyq, h->fgrad.nyq, (long)g.ny, tn, tw);                     \
    return 0;                                                                                           \
  }
    PF_FAITHFUL_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported nz");
  }
}

int step_faithful(pf_handle* h, bool diag, long rec) {
  const Geo& g = h->g;
  hipStream_t s = h->s;
  const unsigned rb = row_block(g);
  int rc;
  double* st = (double*)h->fs.main;
  // loops 5-10 (read the real state before it is transformed in place)
  h->n_part_nl = grid_for(g.P, NB);
  tbeg(h, ST_NONLINEAR);
  const size_t shb = sizeof(double) * ((size_t)g.var * NB + NB);
  if (diag)
    hipLaunchKernelGGL((k_nl_faithful<true>), dim3(grid_for(g.P, NB)), dim3(NB), shb, s, g, h->u, st,
                       (double*)h->fdfdc.main, (double*)h->fpsi.main, h->fmob, h->part_nl);
  else
    hipLaunchKernelGGL((k_nl_faithful<false>), dim3(grid_for(g.P, NB)), dim3(NB), shb, s, g, h->u, st,
                       (double*)h->fdfdc.main, (double*)h->fpsi.main, h->fmob, h->part_nl);
  tend(h);
  // loops 1-4: F(phi_i), F(c); gradients; f_int
  tbeg(h, ST_X_R2C);
  if ((rc = fwd3d<double>(h->fs.main, h->fs.nyq, g.nx, g.ny, g.nz, g.F, 1.0 / (double)(float)g.P, s))) return rc;
  tend(h);
  // The gradients (3F inverse transforms) feed nothing but f_int, a summary-row
  // quantity: they run on the steps whose row is reported.  The state sequence
  // is the reference's either way.
  if (diag) {
    tbeg(h, ST_GRAD);
    if ((rc = launch_grad_zinv(h))) return rc;
    // loops 2 and 4 (src/var_diff.f90:620-622, 670-671): inverse y of the 3F
    // gradient fields, then the inverse x pass summed straight into
    // f_int = sum 0.5*kappa*|grad u|^2 (the real gradients are never stored)
    const long M = g.M, nb3 = 3L * g.F;
    if ((rc = col_pass<double>(g.ny, true, h->fgrad.main, M, (long)g.nz * nb3, s))) return rc;
    if ((rc = row_pass<double>(g.ny, true, h->fgrad.nyq, (long)g.nz * nb3, s))) return rc;
    long nblk = 0;
    if ((rc = x_c2r_sumsq<double>(g.nx, h->fgrad.main, h->fgrad.nyq, (long)g.plane * nb3, g.plane,
                                  0.5 * h->u.kc, 0.5 * h->u.kphi, h->part_fi,
                                  2 * std::max(kMaxGrid, kZlineMaxGrid), &nblk, s)))
      return rc;
    h->n_part_fi = nblk;
    tend(h);
  }
  if (diag) {
    tbeg(h, ST_DIAG);
    diag_final(h, rec, true, true, (const double*)h->fs.main, nullptr);  // Re F(c)(0) = mean(c)
    tend(h);
  }
  // loop 10: F(df_dc)  (F(c) of loop 10 is the same transform as loop 3's; reused)
  tbeg(h, ST_MOBILITY);
  if ((rc = fwd3d<double>(h->fdfdc.main, h->fdfdc.nyq, g.nx, g.ny, g.nz, 1, 1.0 / (double)(float)g.P, s))) return rc;
  // loop 11
  hipLaunchKernelGGL((k_ghat<double>), dim3(grid_for(g.plane, 1)), dim3(rb), 0, s, g, h->u, h->kt, h->fdfdc, 0,
                     h->fs, 0, h->fdum, (long)g.plane);
  if ((rc = inv3d<double>(h->fdum.main, h->fdum.nyq, g.nx, g.ny, g.nz, 1, s))) return rc;
  // loop 12
  hipLaunchKernelGGL(k_mult_faithful, dim3(grid_for(g.P, NB)), dim3(NB), 0, s, g.P, h->u.D,
                     (double*)h->fdum.main, h->fmob);
  if ((rc = fwd3d<double>(h->fdum.main, h->fdum.nyq, g.nx, g.ny, g.nz, 1, 1.0 / (double)(float)g.P, s))) return rc;
  tend(h);
  // loop 13
  tbeg(h, ST_UPDATE);
  hipLaunchKernelGGL(k_ch_faithful, dim3(grid_for(g.plane, 1)), dim3(rb), 0, s, g, h->u, h->kt, h->fs, h->fdum,
                     h->fdfdc, (long)g.plane);
  tend(h);
  // loop 14: F(df_dphi_i) (F(phi_i) is loop 1's transform; reused)
  tbeg(h, ST_Y_FWD);
  if ((rc = fwd3d<double>(h->fpsi.main, h->fpsi.nyq, g.nx, g.ny, g.nz, g.var, 1.0 / (double)(float)g.P, s))) return rc;
  tend(h);
  tbeg(h, ST_UPDATE);
  hipLaunchKernelGGL(k_ac_faithful, dim3(grid_for((long)g.plane * g.var, 1)), dim3(rb), 0, s, g, h->u, h->kt,
                     h->fs, h->fpsi, (long)g.plane * g.var);
  tend(h);
  // inverse transforms of c (loop 13) and phi_i (loop 14), batched
  tbeg(h, ST_Z_INV);
  if ((rc = inv3d<double>(h->fs.main, h->fs.nyq, g.nx, g.ny, g.nz, g.F, s))) return rc;
  tend(h);
  return 0;
}

// ---------------------------------------------------------------------------
// Reduced step (fused kernels).  Between steps W holds the x-half-line state
// (the spectrum S transformed back along z and y), S the full spectrum.
// ---------------------------------------------------------------------------
#define PF_GRID_SIZES(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(160) X(192) X(256) X(512)

// Lines per x-line block: the smallest of 1, 2, 4, 8 that gives the block's
// transforms at least two waves of work (F*LINES*T >= 128 threads).  Small
// blocks keep several of them resident per CU, which is what hides the
// latency of each block's staging and barriers: at 256^3 x 17 fields one
// line per block (3 waves, 19 KB of LDS) runs 1.6x faster than two.
constexpr int kXlineMinThreads = 128;
// LDS one x-line block may take (gfx950: 160 KiB per workgroup); above 64 KiB
// the kernel is opted in with hipFuncAttributeMaxDynamicSharedMemorySize.
constexpr size_t kXlineMaxLds = 160 * 1024;

template <typename R, int NX, int LINES, bool XONLY>
bool xline_fits(const Geo& g, long* nt_out, size_t* shb_out) {
  constexpr int M = NX / 2, T = plan_T(M);
  const int F = XONLY ? 1 : g.F;
  if (g.plane % LINES) return false;
  const long jobs = (long)F * LINES;
  const long MP = xline_row_pitch(M, sizeof(c2<R>));
  long need = std::max<long>(64, (jobs * T + 63) / 64 * 64);
  const size_t body = sizeof(c2<R>) * ((size_t)xline_rows(M, (int)jobs) * MP + NX) +
                      (xline_pair_rows(M) ? 0 : sizeof(R) * (size_t)(jobs + 1));
  const size_t shb = ((body + 15) & ~(size_t)15) + 16 * sizeof(double);
  if (need > kXlineMaxThreadsN<R, NX> || shb > kXlineMaxLds) return false;
  *nt_out = need;
  *shb_out = shb;
  return true;
}

template <typename R, int NX, int LINES, bool XONLY>
int launch_xline_n(pf_handle* h, c2<R>* wm, c2<R>* wn, R* mobd, double* part, long* grid_out, bool diag,
                   bool mob, long nt, size_t shb) {
  const Geo& g = h->g;
  const long ngroups = g.plane / LINES;
  const long cap = (!XONLY && h->tune.xline_grid > 0) ? h->tune.xline_grid : kMaxGrid;
  long grid = std::min<long>(ngroups, cap);
  if (grid >= 8) grid -= grid % 8;
  const c2<R>* tw = twiddle_table<R>(NX);
  if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");
  if (shb > 65536 &&
      hipFuncSetAttribute(reinterpret_cast<const void*>(&k_xline<R, NX, LINES, XONLY>),
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)shb) != hipSuccess)
    return fail(PF_EHIP, "x-line kernel: " + std::to_string(shb) + " bytes of LDS refused");
  hipLaunchKernelGGL((k_xline<R, NX, LINES, XONLY>), dim3((unsigned)grid), dim3((unsigned)nt), shb, h->s, g, h->u,
                     wm, wn, mobd, tw, part, ngroups, (int)diag, (int)mob, (const Member<R>*)nullptr);
  if (grid_out) *grid_out = grid;
  return 0;
}

template <typename R, bool XONLY>
int launch_xline(pf_handle* h, c2<R>* wm, c2<R>* wn, R* mobd, double* part, long* grid_out, bool diag,
                 bool mob) {
  long nt = 0;
  size_t shb = 0;
  const long F = XONLY ? 1 : h->g.F;
  switch (h->g.nx) {
#define PF_TRY(N, L)                                                                                     \
  if (xline_fits<R, N, L, XONLY>(h->g, &nt, &shb))                                                       \
    return launch_xline_n<R, N, L, XONLY>(h, wm, wn, mobd, part, grid_out, diag, mob, nt, shb);
#define PF_CASE(N)                                                                                       \
  case N: {                                                                                              \
    const long T = plan_T(N / 2);                                                                        \
    int want = 1;                                                                                        \
    while (want < 8 && F * want * T < kXlineMinThreads) want *= 2;                                       \
    const int tl = XONLY ? 0 : h->tune.xline_lines;                                                      \
    if (tl > 0) want = tl;                                                                               \
    if (want >= 8) { PF_TRY(N, 8) }                                                                      \
    if (want >= 4 && want < 8) { PF_TRY(N, 4) }                                                          \
    if (want >= 2 && want < 4) { PF_TRY(N, 2) }                                                          \
    if (tl == 0 || want == 1) { PF_TRY(N, 1) }                                                           \
    return fail(PF_ENOTIMPL, "x-line kernel: " + std::to_string(F) + " fields of length " #N " in " +    \
                                 std::to_string(want) + "-line blocks exceed one block's threads or LDS"); \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
#undef PF_TRY
    default: return fail(PF_EINVAL, "unsupported nx");
  }
}

// Whether some x-line block shape fits this grid (checked at pf_create, so
// that an unsupported combination fails there rather than at the first step).
template <typename R, bool XONLY>
bool xline_supported(const Geo& g) {
  long nt = 0;
  size_t shb = 0;
  switch (g.nx) {
#define PF_CASE(N)                                                                                 \
  case N:                                                                                          \
    return xline_fits<R, N, 8, XONLY>(g, &nt, &shb) || xline_fits<R, N, 4, XONLY>(g, &nt, &shb) ||   \
           xline_fits<R, N, 2, XONLY>(g, &nt, &shb) || xline_fits<R, N, 1, XONLY>(g, &nt, &shb);
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return false;
  }
}

// Fused y pass (k_ypass): main block and Nyquist planes in one launch.
template <typename R, bool INV>
int launch_ypass_fused(pf_handle* h, c2<R>* main, c2<R>* nyq, int nf) {
  const Geo& g = h->g;
  switch (g.ny) {
#define PF_CASE(N)                                                                              \
  case N: {                                                                                     \
    constexpr int TX = col_tile(N, sizeof(c2<R>));                                              \
    const long npanels = (long)g.nz * nf, nlines = (long)g.nz * nf;                             \
    const long tpp = ((long)g.M + TX - 1) / TX;                                                 \
    /* a multiple of 8 blocks for the XCD order (xcd_block); the extra blocks fall on   */       \
    /* Nyquist lines past the last one and touch no memory                              */       \
    const long grid = (tpp * npanels + (nlines + TX - 1) / TX + 7) / 8 * 8;                     \
    const c2<R>* tw = twiddle_table<R>(N);                                                      \
    if (!tw) return fail(PF_EHIP, "twiddle table allocation failed");                           \
    hipLaunchKernelGGL((k_ypass<R, N, INV>), dim3((unsigned)grid), dim3(TX * plan_T(N)), 0, \
                       h->s, main, nyq, (long)g.M, npanels, nlines, tw, (const Member<R>*)nullptr, 0L); \
    return 0;                                                                                   \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported ny");
  }
}

// y transform of nf fields of a bank.  FP32: the fused pass.  FP64: the column
// pass over the main block and the row pass over the Nyquist planes; in FP64 the
// fused kernel measured 1.10 ms (16 columns per tile) / 0.97 ms (8) per pass at
// 256^3 x 17 against 0.85 ms for the plain column pass, while in FP32 the two
// are equal (0.45 ms).
template <typename R, bool INV>
int launch_ypass(pf_handle* h, c2<R>* main, c2<R>* nyq, int nf) {
  if constexpr (sizeof(R) == 8) {
    const Geo& g = h->g;
    const long nlines = (long)g.nz * nf;
    int rc = col_pass<R>(g.ny, INV, main, (long)g.M, nlines, h->s);
    if (!rc) rc = row_pass<R>(g.ny, INV, nyq, nlines, h->s);
    return rc;
  } else {
    return launch_ypass_fused<R, INV>(h, main, nyq, nf);
  }
}

// The z-line's plan variant: the tuned one for the FP32 update of every field
// (ROLE_REG), the default for FP64 and for the mobility chain's roles.
template <typename R, int ROLE> int zline_plan_of(const pf_handle* h, int n) {
  if (sizeof(R) != 4 || ROLE != ROLE_REG || h->tune.zline_plan <= 0) return 0;
  return plan_alt(n, h->tune.zline_plan, 0) != 0 ? h->tune.zline_plan : 0;
}

template <typename R, int ROLE, bool DC64, bool PARSEVAL>
int launch_zline(pf_handle* h, Bank<R> S, Bank<R> W, Bank<R> X, Bank<R> Y, double* part, int f0, int nf,
                 long* grid_out) {
  const Geo& g = h->g;
  switch (g.nz) {
#define PF_CASE(N)                                                                                   \
  case N: {                                                                                          \
    constexpr int TX = col_tile(N, sizeof(c2<R>));                                                 \
    const long tm = ((long)g.M * g.ny + TX - 1) / TX, tn = ((long)g.ny + TX - 1) / TX;               \
    const long cap = h->tune.zline_grid > 0 ? h->tune.zline_grid : kZlineMaxGrid;                    \
    long grid = std::min<long>((tm + tn) * nf, cap);                                                 \
    /* whole XCD runs of main-block rows (xcd_runs): extra blocks find no tile */                    \
    const long rows = h->tune.zline_run_rows < 0 ? 1 : h->tune.zline_run_rows;                      \
    const long runt = rows * (((long)g.M + TX - 1) / TX), run = 8 * runt;                           \
    if (run > 0 && (grid + run - 1) / run * run <= cap) grid = (grid + run - 1) / run * run;        \
    const c2<R>* tw = twiddle_table<R>(N);                                                           \
    if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");                                    \
    const int plan = zline_plan_of<R, ROLE>(h, N);                                                   \
    if constexpr (sizeof(R) == 4 && ROLE == ROLE_REG && plan_alt(N, 1, 0) != 0) {                    \
      if (plan == 1) {                                                                               \
        hipLaunchKernelGGL((k_zline<R, N, ROLE, DC64, PARSEVAL, false, 1>), dim3((unsigned)grid),    \
                           dim3(TX * plan_T(N, 1)), 0, h->s, g, h->u, h->kt, S, W, X, Y, h->dc64, part, \
                           tw, f0, nf, tm, tn, runt, (const Member<R>*)nullptr);                     \
        if (grid_out) *grid_out = grid;                                                              \
        return 0;                                                                                    \
      }                                                                                              \
      if (plan == 2) {                                                                               \
        hipLaunchKernelGGL((k_zline<R, N, ROLE, DC64, PARSEVAL, false, 2>), dim3((unsigned)grid),    \
                           dim3(TX * plan_T(N, 2)), 0, h->s, g, h->u, h->kt, S, W, X, Y, h->dc64, part, \
                           tw, f0, nf, tm, tn, runt, (const Member<R>*)nullptr);                     \
        if (grid_out) *grid_out = grid;                                                              \
        return 0;                                                                                    \
      }                                                                                              \
    }                                                                                                \
    hipLaunchKernelGGL((k_zline<R, N, ROLE, DC64, PARSEVAL>), dim3((unsigned)grid),                  \
                       dim3(TX * plan_T(N)), 0, h->s, g, h->u, h->kt, S, W, X, Y, h->dc64, part, tw,      \
                       f0, nf, tm, tn, runt, (const Member<R>*)nullptr);                             \
    if (grid_out) *grid_out = grid;                                                                  \
    return 0;                                                                                        \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported nz");
  }
}

// ---------------------------------------------------------------------------
// Batched ensemble: one launch per stage over every member of a group (same
// grid and var, FP32-mixed form, constant mobility), each member's banks and
// parameters read from the device table mt.  The kernels and their per-point
// arithmetic are those of a single member's step, so each member's result is
// bit-identical to stepping it alone.
// ---------------------------------------------------------------------------
int xline_batch(const Geo& g, const UPar& u, const pf_tuning& tn, const Member<float>* mt, long nmem,
                hipStream_t s) {
  long nt = 0;
  size_t shb = 0;
  const long cap = tn.xline_grid > 0 ? tn.xline_grid : kMaxGrid;
  switch (g.nx) {
#define PF_TRY(N, L)                                                                                   \
  if (xline_fits<float, N, L, false>(g, &nt, &shb)) {                                                  \
    const long ngroups = nmem * (g.plane / L);                                                         \
    long grid = std::min<long>(ngroups, cap);                                                          \
    if (grid >= 8) grid -= grid % 8;                                                                   \
    const c2<float>* tw = twiddle_table<float>(N);                                                     \
    if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");                                      \
    if (shb > 65536 &&                                                                                 \
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_xline<float, N, L, false, true>),         \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shb) != hipSuccess)       \
      return fail(PF_EHIP, "x-line kernel: " + std::to_string(shb) + " bytes of LDS refused");         \
    hipLaunchKernelGGL((k_xline<float, N, L, false, true>), dim3((unsigned)grid), dim3((unsigned)nt), shb, \
                       s, g, u, (c2<float>*)nullptr, (c2<float>*)nullptr, (float*)nullptr, tw,         \
                       (double*)nullptr, ngroups, 0, 0, mt);                                            \
    return 0;                                                                                          \
  }
#define PF_CASE(N)                                                                                     \
  case N: {                                                                                            \
    const long T = plan_T(N / 2);                                                                      \
    int want = 1;                                                                                      \
    while (want < 8 && g.F * want * T < kXlineMinThreads) want *= 2;                                   \
    if (tn.xline_lines > 0) want = tn.xline_lines;                                                     \
    if (want >= 8) { PF_TRY(N, 8) }                                                                    \
    if (want >= 4 && want < 8) { PF_TRY(N, 4) }                                                        \
    if (want >= 2 && want < 4) { PF_TRY(N, 2) }                                                        \
    if (tn.xline_lines == 0 || want == 1) { PF_TRY(N, 1) }                                             \
    return fail(PF_ENOTIMPL, "x-line kernel: " + std::to_string(g.F) +                                \
                                 " fields of length " #N " exceed one block's threads or LDS");        \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
#undef PF_TRY
    default: return fail(PF_EINVAL, "unsupported nx");
  }
}

template <bool INV> int ypass_batch(const Geo& g, const Member<float>* mt, long nmem, hipStream_t s) {
  switch (g.ny) {
#define PF_CASE(N)                                                                                    \
  case N: {                                                                                           \
    constexpr int TX = col_tile(N, sizeof(c2<float>));                                                \
    const long ppm = (long)g.nz * g.F, npanels = nmem * ppm, nlines = npanels;                        \
    const long tpp = ((long)g.M + TX - 1) / TX;                                                       \
    const long grid = (tpp * npanels + (nlines + TX - 1) / TX + 7) / 8 * 8;                           \
    const c2<float>* tw = twiddle_table<float>(N);                                                    \
    if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");                                     \
    hipLaunchKernelGGL((k_ypass<float, N, INV, true>), dim3((unsigned)grid), dim3(TX * plan_T(N)), 0, s, \
                       (c2<float>*)nullptr, (c2<float>*)nullptr, (long)g.M, npanels, nlines, tw, mt, ppm); \
    return 0;                                                                                         \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported ny");
  }
}

int zline_batch(const Geo& g, const UPar& u, const pf_tuning& tun, const KTab& kt, const Member<float>* mt,
                long nmem, hipStream_t s) {
  const Bank<float> none{nullptr, nullptr};
  const long cap = tun.zline_grid > 0 ? tun.zline_grid : kZlineMaxGrid;
  const long rows = tun.zline_run_rows < 0 ? 1 : tun.zline_run_rows;
  switch (g.nz) {
#define PF_CASE(N)                                                                                   \
  case N: {                                                                                          \
    constexpr int TX = col_tile(N, sizeof(c2<float>));                                             \
    const long tm = ((long)g.M * g.ny + TX - 1) / TX, tn = ((long)g.ny + TX - 1) / TX;               \
    const long nf = nmem * g.F;                                                                      \
    long grid = std::min<long>((tm + tn) * nf, cap);                                                 \
    const long runt = rows * (((long)g.M + TX - 1) / TX), run = 8 * runt;                            \
    if (run > 0 && (grid + run - 1) / run * run <= cap) grid = (grid + run - 1) / run * run;        \
    const c2<float>* tw = twiddle_table<float>(N);                                                   \
    if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");                                    \
    const int plan = (tun.zline_plan > 0 && plan_alt(N, tun.zline_plan, 0) != 0) ? tun.zline_plan : 0; \
    if constexpr (plan_alt(N, 1, 0) != 0) {                                                          \
      if (plan == 1) {                                                                               \
        hipLaunchKernelGGL((k_zline<float, N, ROLE_REG, true, false, true, 1>), dim3((unsigned)grid), \
                           dim3(TX * plan_T(N, 1)), 0, s, g, u, kt, none, none, none, none,          \
                           (double*)nullptr, (double*)nullptr, tw, 0, (int)nf, tm, tn, runt, mt);     \
        return 0;                                                                                    \
      }                                                                                              \
      if (plan == 2) {                                                                               \
        hipLaunchKernelGGL((k_zline<float, N, ROLE_REG, true, false, true, 2>), dim3((unsigned)grid), \
                           dim3(TX * plan_T(N, 2)), 0, s, g, u, kt, none, none, none, none,          \
                           (double*)nullptr, (double*)nullptr, tw, 0, (int)nf, tm, tn, runt, mt);     \
        return 0;                                                                                    \
      }                                                                                              \
    }                                                                                                \
    hipLaunchKernelGGL((k_zline<float, N, ROLE_REG, true, false, true>), dim3((unsigned)grid),       \
                       dim3(TX * plan_T(N)), 0, s, g, u, kt, none, none, none, none, (double*)nullptr, \
                       (double*)nullptr, tw, 0, (int)nf, tm, tn, runt, mt);                          \
    return 0;                                                                                        \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported nz");
  }
}

template <typename R>
int step_reduced(pf_handle* h, bool diag, long rec, bool fint_next) {
  const Geo& g = h->g;
  hipStream_t s = h->s;
  const bool mob = h->p.D_gb != 0.0;
  constexpr bool DC = sizeof(R) == 4;
  Bank<R> S{(c2<R>*)h->S_main, nullptr}, W{(c2<R>*)h->W_main, (c2<R>*)h->W_nyq};
  Bank<R> G{(c2<R>*)h->G_main, (c2<R>*)h->G_nyq}, Fc{(c2<R>*)h->Fc_main, (c2<R>*)h->Fc_nyq};
  int rc;

  if (diag && !h->fint_ready) {
    tbeg(h, ST_DIAG);
    const long rows = (long)g.plane * g.F;
    hipLaunchKernelGGL((k_update<R, false, false, true, false, true>), dim3(grid_for(rows, 1)), dim3(NB), 0, s, g,
                       h->u, h->kt, S, W, G, h->dc64, h->part_fi, rows);
    h->n_part_fi = grid_for(rows, 1);
    tend(h);
  }
  // x: C2R -> nonlinear -> R2C, all fields
  tbeg(h, ST_NONLINEAR);
  {
    long gr = 0;
    R* md = (R*)h->mob;
    rc = launch_xline<R, false>(h, W.main, W.nyq, md, h->part_nl, &gr, diag, mob);
    if (rc) return rc;
    if (diag) h->n_part_nl = gr;
  }
  tend(h);
  if (diag) {
    tbeg(h, ST_DIAG);
    diag_final(h, rec, true, true, DC ? h->dc64 : (const double*)h->S_main, nullptr);
    tend(h);
  }
  tbeg(h, ST_Y_FWD);
  if ((rc = launch_ypass<R, false>(h, W.main, W.nyq, g.F))) return rc;
  tend(h);
  tbeg(h, ST_UPDATE);
  long g1 = 0, g2 = 0;
  if (!mob) {
    rc = fint_next ? launch_zline<R, ROLE_REG, DC, true>(h, S, W, W, W, h->part_fi, 0, g.F, &g1)
                   : launch_zline<R, ROLE_REG, DC, false>(h, S, W, W, W, h->part_fi, 0, g.F, &g1);
    if (rc) return rc;
  } else {
    if ((rc = launch_zline<R, ROLE_GHAT, false, false>(h, S, W, Fc, G, h->part_fi, 0, 1, nullptr))) return rc;
    rc = fint_next ? launch_zline<R, ROLE_REG, DC, true>(h, S, W, W, W, h->part_fi, 1, g.var, &g1)
                   : launch_zline<R, ROLE_REG, DC, false>(h, S, W, W, W, h->part_fi, 1, g.var, &g1);
    if (rc) return rc;
    tend(h);
    tbeg(h, ST_MOBILITY);
    if ((rc = launch_ypass<R, true>(h, G.main, G.nyq, 1))) return rc;
    if ((rc = launch_xline<R, true>(h, G.main, G.nyq, (R*)h->mob, nullptr, nullptr, false, false))) return rc;
    if ((rc = launch_ypass<R, false>(h, G.main, G.nyq, 1))) return rc;
    tend(h);
    tbeg(h, ST_UPDATE);
    rc = fint_next ? launch_zline<R, ROLE_CMOB, DC, true>(h, S, W, Fc, G, h->part_fi + g1, 0, 1, &g2)
                   : launch_zline<R, ROLE_CMOB, DC, false>(h, S, W, Fc, G, h->part_fi + g1, 0, 1, &g2);
    if (rc) return rc;
  }
  if (fint_next) h->n_part_fi = g1 + g2;
  h->fint_ready = fint_next;
  tend(h);
  tbeg(h, ST_Y_INV);
  if ((rc = launch_ypass<R, true>(h, W.main, W.nyq, g.F))) return rc;
  tend(h);
  return 0;
}

template <typename R> int alloc_reduced(pf_handle* h) {
  Geo& g = h->g;
  g.stx = col_tile(g.nz, sizeof(c2<R>));
  g.stm = ((long)g.M * g.ny + g.stx - 1) / g.stx;
  g.stpf = g.stm + ((long)g.ny + g.stx - 1) / g.stx;
  Bank<R> S{}, W{}, G{};
  int rc = dalloc(h, (void**)&S.main, sizeof(c2<R>) * (size_t)g.stpf * g.nz * g.stx * g.F);
  if (!rc) rc = alloc_bank<R>(h, W, h->g.F);
  Bank<R> Fc{}, X{};
  if (!rc && h->p.D_gb != 0.0) {
    rc = alloc_bank<R>(h, G, 1);
    if (!rc) rc = alloc_bank<R>(h, Fc, 1);
    if (!rc) rc = dalloc(h, &h->mob, sizeof(R) * (size_t)h->g.P);
  }
  if (!rc) rc = alloc_bank<R>(h, X, h->g.F);
  h->S_main = S.main; h->W_main = W.main; h->W_nyq = W.nyq;
  h->G_main = G.main; h->G_nyq = G.nyq; h->Fc_main = Fc.main; h->Fc_nyq = Fc.nyq;
  h->X_main = X.main; h->X_nyq = X.nyq;
  if (!rc) rc = dalloc(h, (void**)&h->dc64, sizeof(double) * h->g.F);
  return rc;
}

constexpr int kMeanChunks = 64;

struct Neumaier {
  double s = 0.0, c = 0.0;
  void add(double x) {
    const double t = s + x;
    if (std::fabs(s) >= std::fabs(x)) c += (s - t) + x; else c += (x - t) + s;
    s = t;
  }
  double value() const { return s + c; }
};

std::vector<double> host_means(const double* phi, const double* con, long P, int var) {
  const int F = var + 1;
  std::vector<Neumaier> part((size_t)kMeanChunks * F);
  auto run = [&](int ch) {
    const long t0 = P * ch / kMeanChunks, t1 = P * (ch + 1) / kMeanChunks;
    Neumaier* acc = &part[(size_t)ch * F];
    for (long t = t0; t < t1; ++t) {
      acc[0].add(con[t]);
      const double* p = phi + t * var;
      for (int f = 1; f < F; ++f) acc[f].add(p[f - 1]);
    }
  };
  const int nth = (int)std::max(1u, std::min<unsigned>(16, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (int w = 0; w < nth; ++w)
    th.emplace_back([&, w] { for (int ch = w; ch < kMeanChunks; ch += nth) run(ch); });
  for (auto& t : th) t.join();
  std::vector<double> dc(F);
  for (int f = 0; f < F; ++f) {
    Neumaier tot;
    for (int ch = 0; ch < kMeanChunks; ++ch) {
      tot.add(part[(size_t)ch * F + f].s);
      tot.add(part[(size_t)ch * F + f].c);
    }
    dc[f] = tot.value() / (double)(float)P;
  }
  return dc;
}

// Reduced forms, once W holds the real state (SoA): S <- F(W) with f_trans
// normalisation (computed in X, then put in the column-tile layout), W <- x-half
// lines, dc64 <- the FP64 field means dc.
template <typename R> int finish_put_reduced(pf_handle* h, const std::vector<double>& dc) {
  const Geo& g = h->g;
  PF_HIP(hipMemcpyAsync(h->X_main, h->W_main, sizeof(R) * (size_t)g.P * g.F, hipMemcpyDeviceToDevice, h->s));
  int rc = fwd3d<R>((c2<R>*)h->X_main, (c2<R>*)h->X_nyq, g.nx, g.ny, g.nz, g.F, (R)(1.0 / (double)(float)g.P), h->s);
  if (rc) return rc;
  const long nst = g.stpf * g.nz * g.stx * g.F;
  hipLaunchKernelGGL((k_tile_spectrum<R>), dim3(grid_for(nst, NB)), dim3(NB), 0, h->s, g,
                     (const c2<R>*)h->X_main, (const c2<R>*)h->X_nyq, (c2<R>*)h->S_main, nst);
  // W <- x-half lines: (1/nx) R2C along x of the real state == z,y-inverse of S
  rc = x_r2c<R>(g.nx, (c2<R>*)h->W_main, (c2<R>*)h->W_nyq, (long)g.ny * g.nz * g.F, (R)(1.0 / (double)g.nx), h->s);
  if (rc) return rc;
  PF_HIP(hipMemcpyAsync(h->dc64, dc.data(), sizeof(double) * g.F, hipMemcpyHostToDevice, h->s));
  PF_HIP(hipStreamSynchronize(h->s));
  return 0;
}

template <typename R> int put_reduced(pf_handle* h, const double* phi, const double* con) {
  const Geo& g = h->g;
  // W <- real state
  hipLaunchKernelGGL((k_scatter_state<R>), dim3(grid_for(g.P, NB)), dim3(NB), 0, h->s, h->staging,
                     h->staging + (size_t)g.var * g.P, (R*)h->W_main, g.P, g.var);
  // FP64 means (k = 0 modes), Neumaier-compensated: one pass over the host
  // arrays in kMeanChunks fixed chunks (threads take whole chunks), chunk
  // partials combined in chunk order, so the result does not depend on the
  // thread count
  return finish_put_reduced<R>(h, host_means(phi, con, g.P, g.var));
}

}  // namespace

// ---------------------------------------------------------------------------
// Tuned launch shapes: entries written by the tuner (autotune.py) from timings
// on an MI355X, one per (grid, var, precision); pf_create takes the matching
// entry, else the defaults.
// ---------------------------------------------------------------------------
struct TunedShape {
  int nx, ny, nz, var, prec;  // prec: PF_PREC_FP64 / PF_PREC_FP32
  pf_tuning t;
};
const TunedShape kTuned[] = {
#include "pf_tuned.inc"
    {0, 0, 0, 0, -1, {0, 0, 0, -1, 0}},
};

// Whether the x-line can run blocks of `lines` lines for this handle.
bool xline_lines_fit(const pf_handle* h, int lines) {
  const Geo& g = h->g;
  long nt = 0;
  size_t shb = 0;
  const bool f32 = h->mode == PF_MODE_FP32_MIXED;
  switch (g.nx) {
#define PF_CASE(N)                                                                                      \
  case N:                                                                                               \
    switch (lines) {                                                                                    \
      case 1: return f32 ? xline_fits<float, N, 1, false>(g, &nt, &shb) : xline_fits<double, N, 1, false>(g, &nt, &shb); \
      case 2: return f32 ? xline_fits<float, N, 2, false>(g, &nt, &shb) : xline_fits<double, N, 2, false>(g, &nt, &shb); \
      case 4: return f32 ? xline_fits<float, N, 4, false>(g, &nt, &shb) : xline_fits<double, N, 4, false>(g, &nt, &shb); \
      case 8: return f32 ? xline_fits<float, N, 8, false>(g, &nt, &shb) : xline_fits<double, N, 8, false>(g, &nt, &shb); \
      default: return false;                                                                            \
    }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return false;
  }
}

int check_tuning(const pf_handle* h, const pf_tuning& t) {
  if (!(t.xline_lines == 0 || t.xline_lines == 1 || t.xline_lines == 2 || t.xline_lines == 4 || t.xline_lines == 8))
    return fail(PF_EINVAL, "xline_lines must be 0, 1, 2, 4 or 8");
  if (t.xline_lines > 0 && !xline_lines_fit(h, t.xline_lines))
    return fail(PF_EINVAL, "x-line blocks of " + std::to_string(t.xline_lines) + " lines do not fit this grid");
  if (t.xline_grid != 0 && (t.xline_grid < 256 || t.xline_grid > kXlineGridCap))
    return fail(PF_EINVAL, "xline_grid must be 0 or in [256, " + std::to_string(kXlineGridCap) + "]");
  if (t.zline_grid != 0 && (t.zline_grid < 1024 || t.zline_grid > kZlineGridCap))
    return fail(PF_EINVAL, "zline_grid must be 0 or in [1024, " + std::to_string(kZlineGridCap) + "]");
  if (t.zline_run_rows < -1 || t.zline_run_rows > 8) return fail(PF_EINVAL, "zline_run_rows must be in [-1, 8]");
  if (t.zline_plan < 0 || t.zline_plan > 2) return fail(PF_EINVAL, "zline_plan must be 0, 1 or 2");
  if (t.zline_plan > 0 && (h->mode != PF_MODE_FP32_MIXED || plan_alt(h->g.nz, t.zline_plan, 0) == 0))
    return fail(PF_EINVAL, "no z-line plan variant " + std::to_string(t.zline_plan) + " for nz = " +
                               std::to_string(h->g.nz) + " in this form (variants exist for FP32-mixed at "
                               "nz = 64, 128, 160, 192, 256, 512)");
  return 0;
}

void apply_tuned_table(pf_handle* h) {
  const char* env = std::getenv("PF_TUNING");
  if (env && std::string(env) == "default") return;
  const int prec = h->mode == PF_MODE_FP32_MIXED ? PF_PREC_FP32 : PF_PREC_FP64;
  for (const TunedShape& e : kTuned)
    if (e.nx == h->g.nx && e.ny == h->g.ny && e.nz == h->g.nz && e.var == h->g.var && e.prec == prec) {
      if (check_tuning(h, e.t) == 0) h->tune = e.t;
      return;
    }
}

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" int pf_create(const pf_params* p, int nx, int ny, int nz, int var, int mode, pf_handle** out) {
  if (!p || !out) return fail(PF_EINVAL, "null argument");
  *out = nullptr;
  if (!grid_size_supported(nx) || !grid_size_supported(ny) || !grid_size_supported(nz))
    return fail(PF_EINVAL, std::string("unsupported grid; each axis must be one of ") + supported_sizes_text());
  if (var < 1 || var > 256) return fail(PF_EINVAL, "var must be in [1, 256]");
  if (mode < PF_MODE_FP64_FAITHFUL || mode > PF_MODE_FP32_MIXED) return fail(PF_EINVAL, "bad mode");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(PF_EHIP, "no HIP device available");
  pf_handle* h = new pf_handle();
  h->mode = mode;
  h->p = *p;
  Geo& g = h->g;
  g.nx = nx; g.ny = ny; g.nz = nz; g.M = nx / 2; g.var = var; g.F = var + 1;
  g.P = (long)nx * ny * nz; g.plane = (long)ny * nz; g.hmain = (long)g.M * ny * nz;
  UCoef& u = h->u;
  u.dt = p->t_step; u.kphi = p->grad_coeff_phi; u.kc = p->grad_coeff_c; u.D = p->D_L; u.L = p->mob_phi;
  u.Dgb = p->D_gb; u.am = p->am; u.ap = p->ap; u.cm = p->con_0_mat; u.cp = p->con_0_ppt;
  u.s1g = p->sigma_1; u.s2g = p->sigma_2; u.invP = 1.0 / (double)(float)g.P;
  int rc = 0;
  auto bail = [&](int code) { pf_destroy(h); return code; };
  if (hipGetDevice(&h->dev) != hipSuccess) return bail(fail(PF_EHIP, "hipGetDevice failed"));
  if (hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(PF_EHIP, "stream creation failed"));
  // wavenumber tables
  {
    const int M = g.M;
    std::vector<double> tab(2 * (M + 1) + 2 * ny + 2 * nz);
    double* q = tab.data();
    for (int i = 0; i <= M; ++i) { q[i] = kaxis(i, nx, false); q[M + 1 + i] = kaxis(i, nx, true); }
    q += 2 * (M + 1);
    for (int i = 0; i < ny; ++i) { q[i] = kaxis(i, ny, false); q[ny + i] = kaxis(i, ny, true); }
    q += 2 * ny;
    for (int i = 0; i < nz; ++i) { q[i] = kaxis(i, nz, false); q[nz + i] = kaxis(i, nz, true); }
    if ((rc = dalloc(h, (void**)&h->dk, sizeof(double) * tab.size()))) return bail(rc);
    if (hipMemcpy(h->dk, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(PF_EHIP, "copy of wavenumber tables failed"));
    const double* d = h->dk;
    h->kt.ax = d; h->kt.axp = d + M + 1; d += 2 * (M + 1);
    h->kt.ay = d; h->kt.ayp = d + ny; d += 2 * ny;
    h->kt.az = d; h->kt.azp = d + nz;
    h->kt.uz = (2.0 * (std::atan(1.0) * 4.0)) / (double)(float)nz;  // the unit of kaxis
    std::vector<float> tabf(tab.begin(), tab.end());
    if ((rc = dalloc(h, (void**)&h->dkf, sizeof(float) * tabf.size()))) return bail(rc);
    if (hipMemcpy(h->dkf, tabf.data(), sizeof(float) * tabf.size(), hipMemcpyHostToDevice) != hipSuccess)
      return bail(fail(PF_EHIP, "copy of wavenumber tables failed"));
    const float* df = h->dkf;
    h->ktf.ax = df; h->ktf.axp = df + M + 1; df += 2 * (M + 1);
    h->ktf.ay = df; h->ktf.ayp = df + ny; df += 2 * ny;
    h->ktf.az = df; h->ktf.azp = df + nz;
  }
  if (mode == PF_MODE_FP64_FAITHFUL) {
    if ((rc = alloc_bank<double>(h, h->fs, g.F)) || (rc = alloc_bank<double>(h, h->fgrad, 3 * g.F)) ||
        (rc = alloc_bank<double>(h, h->fdfdc, 1)) || (rc = alloc_bank<double>(h, h->fdum, 1)) ||
        (rc = alloc_bank<double>(h, h->fpsi, var)) || (rc = dalloc(h, (void**)&h->fmob, sizeof(double) * g.P)))
      return bail(rc);
  } else if (mode == PF_MODE_FP64_REDUCED) {
    if (!xline_supported<double, false>(g))
      return bail(fail(PF_ENOTIMPL, "FP64 reduced form: " + std::to_string(g.F) + " fields of " +
                                        std::to_string(nx) + "-point x lines exceed one block's threads or LDS"));
    if ((rc = alloc_reduced<double>(h))) return bail(rc);
  } else {
    if (!xline_supported<float, false>(g))
      return bail(fail(PF_ENOTIMPL, "FP32-mixed form: " + std::to_string(g.F) + " fields of " +
                                        std::to_string(nx) + "-point x lines exceed one block's threads or LDS"));
    if ((rc = alloc_reduced<float>(h))) return bail(rc);
  }
  if ((rc = dalloc(h, (void**)&h->part_nl, sizeof(double) * NPART_NL * std::max(kMaxGrid, kXlineGridCap))))
    return bail(rc);
  if ((rc = dalloc(h, (void**)&h->part_fi, sizeof(double) * 2 * std::max(kMaxGrid, kZlineGridCap)))) return bail(rc);
  if ((rc = dalloc(h, (void**)&h->staging, sizeof(double) * (size_t)g.P * g.F))) return bail(rc);
  h->n_part_nl = grid_for(g.P, NB);
  if ((rc = ensure_diag(h, 64))) return bail(rc);
  if ((rc = dalloc(h, (void**)&h->dctr, sizeof(int)))) return bail(rc);
  // the twiddle tables are set up now (their first use on a device copies
  // them), not inside a step that a caller may capture
  for (int n : {nx, ny, nz})
    if (!twiddle_table<float>(n) || !twiddle_table<double>(n))
      return bail(fail(PF_EHIP, "twiddle table set-up failed"));
  if (mode != PF_MODE_FP64_FAITHFUL) apply_tuned_table(h);
  *out = h;
  return 0;
}

extern "C" int pf_set_tuning(pf_handle* h, const pf_tuning* t) {
  if (!h || !t) return fail(PF_EINVAL, "null argument");
  DeviceGuard dg(h->dev);
  const int rc = check_tuning(h, *t);
  if (rc) return rc;
  drop_graph(h);  // a captured period holds the old launch shapes
  h->tune = *t;
  return 0;
}

extern "C" int pf_get_tuning(const pf_handle* h, pf_tuning* t) {
  if (!h || !t) return fail(PF_EINVAL, "null argument");
  *t = h->tune;
  return 0;
}

extern "C" int pf_destroy(pf_handle* h) {
  if (!h) return 0;
  DeviceGuard dg(h->dev);
  if (h->s) (void)hipStreamSynchronize(h->s);
  drop_graph(h);
  for (void* p : h->allocs) (void)hipFree(p);
  for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
  return 0;
}

extern "C" int pf_put_state(pf_handle* h, const double* phi, const double* con) {
  if (!h || !phi || !con) return fail(PF_EINVAL, "null argument");
  DeviceGuard dg(h->dev);
  const Geo& g = h->g;
  int rc = to_device(h->staging, phi, sizeof(double) * (size_t)g.P * g.var, h->s);
  if (!rc) rc = to_device(h->staging + (size_t)g.P * g.var, con, sizeof(double) * (size_t)g.P, h->s);
  if (rc) return rc;
  h->fint_ready = false;
  if (h->mode == PF_MODE_FP64_FAITHFUL) {
    hipLaunchKernelGGL((k_scatter_state<double>), dim3(grid_for(g.P, NB)), dim3(NB), 0, h->s, h->staging,
                       h->staging + (size_t)g.var * g.P, (double*)h->fs.main, g.P, g.var);
    PF_HIP(hipStreamSynchronize(h->s));
    return 0;
  }
  if (h->mode == PF_MODE_FP64_REDUCED) return put_reduced<double>(h, phi, con);
  return put_reduced<float>(h, phi, con);
}

// Largest seed count of the device-built Voronoi state (seeds are read from
// global memory by every thread; this bounds the per-point search).
constexpr int kMaxVoronoiSeeds = 1 << 16;

extern "C" int pf_put_voronoi(pf_handle* h, int nseeds, int iseed, double c0, double c_amp) {
  if (!h) return fail(PF_EINVAL, "null handle");
  if (nseeds < 1 || nseeds > kMaxVoronoiSeeds)
    return fail(PF_EINVAL, "nseeds must be in [1, " + std::to_string(kMaxVoronoiSeeds) + "]");
  DeviceGuard dg(h->dev);
  const Geo& g = h->g;
  // the draws of pf_ic_voronoi: 3 per seed (x, y, z), then one c per point
  pf_ran2_state st;
  pf_ran2_init(&st);
  int idum = iseed;
  std::vector<double> seeds(3 * (size_t)nseeds);
  for (int s = 0; s < nseeds; ++s) {
    seeds[3 * s] = pf_ran2(&st, &idum) * g.nx;
    seeds[3 * s + 1] = pf_ran2(&st, &idum) * g.ny;
    seeds[3 * s + 2] = pf_ran2(&st, &idum) * g.nz;
  }
  std::vector<double> con((size_t)g.P);
  for (long t = 0; t < g.P; ++t) con[t] = c0 + c_amp * (2.0 * pf_ran2(&st, &idum) - 1.0);
  // staging: c at its usual place (after the var*P phi values), seeds and the
  // owner counts in the (unused here) phi area
  double* dcon = h->staging + (size_t)g.var * g.P;
  double* dseeds = h->staging;
  unsigned long long* dcounts = reinterpret_cast<unsigned long long*>(h->staging + 3 * (size_t)nseeds);
  if ((size_t)3 * nseeds + g.var > (size_t)g.var * g.P)
    return fail(PF_EINVAL, "too many seeds for this grid");
  int rc = to_device(dcon, con.data(), sizeof(double) * (size_t)g.P, h->s);
  if (rc) return rc;
  PF_HIP(hipMemcpyAsync(dseeds, seeds.data(), sizeof(double) * seeds.size(), hipMemcpyHostToDevice, h->s));
  hipLaunchKernelGGL(k_fill_zero, dim3(1), dim3(NB), 0, h->s, (double*)dcounts, (long)g.var);
  h->fint_ready = false;
  const size_t shb = sizeof(unsigned int) * (size_t)g.var;
  const unsigned grid = grid_for(g.P, NB);
  if (h->mode == PF_MODE_FP64_FAITHFUL) {
    hipLaunchKernelGGL((k_voronoi_state<double>), dim3(grid), dim3(NB), shb, h->s, g, dseeds, nseeds, dcon,
                       (double*)h->fs.main, dcounts);
    PF_HIP(hipGetLastError());
    PF_HIP(hipStreamSynchronize(h->s));
    return 0;
  }
  if (h->mode == PF_MODE_FP64_REDUCED)
    hipLaunchKernelGGL((k_voronoi_state<double>), dim3(grid), dim3(NB), shb, h->s, g, dseeds, nseeds, dcon,
                       (double*)h->W_main, dcounts);
  else
    hipLaunchKernelGGL((k_voronoi_state<float>), dim3(grid), dim3(NB), shb, h->s, g, dseeds, nseeds, dcon,
                       (float*)h->W_main, dcounts);
  PF_HIP(hipGetLastError());
  std::vector<unsigned long long> cnt(g.var);
  PF_HIP(hipMemcpyAsync(cnt.data(), dcounts, sizeof(unsigned long long) * g.var, hipMemcpyDeviceToHost, h->s));
  PF_HIP(hipStreamSynchronize(h->s));
  // the means pf_put_state would take from the host arrays: c's by the same
  // chunked compensated sum, each eta's an exact count of ones over P
  std::vector<double> dc = host_means(nullptr, con.data(), g.P, 0);
  dc.resize(g.F);
  for (int v = 0; v < g.var; ++v) dc[v + 1] = (double)cnt[v] / (double)(float)g.P;
  if (h->mode == PF_MODE_FP64_REDUCED) return finish_put_reduced<double>(h, dc);
  return finish_put_reduced<float>(h, dc);
}

namespace {
// Reduced modes: real state of the current step = C2R along x of a copy of W.
template <typename R> int real_state_reduced(pf_handle* h) {
  const Geo& g = h->g;
  PF_HIP(hipMemcpyAsync(h->X_main, h->W_main, sizeof(R) * (size_t)g.P * g.F, hipMemcpyDeviceToDevice, h->s));
  PF_HIP(hipMemcpyAsync(h->X_nyq, h->W_nyq, sizeof(c2<R>) * (size_t)g.plane * g.F, hipMemcpyDeviceToDevice, h->s));
  return x_c2r<R>(g.nx, (c2<R>*)h->X_main, (const c2<R>*)h->X_nyq, (long)g.ny * g.nz * g.F, h->s);
}
}  // namespace

extern "C" int pf_get_state(pf_handle* h, double* phi, double* con) {
  if (!h || !phi || !con) return fail(PF_EINVAL, "null argument");
  DeviceGuard dg(h->dev);
  const Geo& g = h->g;
  double* sphi = h->staging;
  double* scon = h->staging + (size_t)g.var * g.P;
  int rc = 0;
  if (h->mode == PF_MODE_FP64_FAITHFUL) {
    hipLaunchKernelGGL((k_gather_state<double>), dim3(grid_for(g.P, NB)), dim3(NB), 0, h->s,
                       (const double*)h->fs.main, sphi, scon, g.P, g.var);
  } else if (h->mode == PF_MODE_FP64_REDUCED) {
    if ((rc = real_state_reduced<double>(h))) return rc;
    hipLaunchKernelGGL((k_gather_state<double>), dim3(grid_for(g.P, NB)), dim3(NB), 0, h->s,
                       (const double*)h->X_main, sphi, scon, g.P, g.var);
  } else {
    if ((rc = real_state_reduced<float>(h))) return rc;
    hipLaunchKernelGGL((k_gather_state<float>), dim3(grid_for(g.P, NB)), dim3(NB), 0, h->s,
                       (const float*)h->X_main, sphi, scon, g.P, g.var);
  }
  PF_HIP(hipGetLastError());
  rc = to_host(phi, sphi, sizeof(double) * (size_t)g.P * g.var, h->s);
  if (!rc) rc = to_host(con, scon, sizeof(double) * (size_t)g.P, h->s);
  return rc;
}

namespace {

// One step of the handle's form.  rec: the diagnostics row (when diag).
int step_once(pf_handle* h, bool diag, long rec, bool next_diag) {
  if (h->mode == PF_MODE_FP64_FAITHFUL) return step_faithful(h, diag, rec);
  if (h->mode == PF_MODE_FP64_REDUCED) return step_reduced<double>(h, diag, rec, next_diag);
  return step_reduced<float>(h, diag, rec, next_diag);
}

// Captures one period of steps with the diagnostics pattern of diag_every = de
// (0: none) into h->gx:
//   de = 0 : one step, no diagnostics
//   de = 1 : one step with diagnostics that also carries f_int into the next
//   de >= 2: de steps, the step before last carrying f_int, the last reporting
// The capture only records; nothing runs.  On a capture failure graphs are
// switched off for the handle and the caller steps eagerly.
int capture_period(pf_handle* h, int de) {
  drop_graph(h);
  const int P = de > 0 ? de : 1;
  const bool ready0 = h->fint_ready;
  h->fint_ready = (de == 1);
  h->capturing = true;
  h->cap_step = P;
  int rc = 0;
  if (hipStreamBeginCapture(h->s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    h->capturing = false;
    h->fint_ready = ready0;
    h->graphs = false;
    (void)hipGetLastError();
    return 0;
  }
  for (int j = 0; j < P && !rc; ++j) {
    const bool diag = de > 0 && j == P - 1;
    const bool next = de == 1 || (de >= 2 && j == P - 2);
    rc = step_once(h, diag, 0, next);
  }
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(h->s, &graph);
  h->capturing = false;
  h->fint_ready = ready0;
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess || !graph || hipGraphInstantiate(&h->gx, graph, nullptr, nullptr, 0) != hipSuccess) {
    h->gx = nullptr;
    h->graphs = false;
    (void)hipGetLastError();
  } else {
    h->gx_key = de;
    h->gx_period = P;
  }
  if (graph) (void)hipGraphDestroy(graph);
  return 0;
}

// Fewest replays of a period for which a call captures (or reuses) a graph.
constexpr long kMinGraphReplays = 2;

}  // namespace

extern "C" int pf_set_graphs(pf_handle* h, int on) {
  if (!h) return fail(PF_EINVAL, "null handle");
  DeviceGuard dg(h->dev);
  if (!on) drop_graph(h);
  h->graphs = on != 0;
  return 0;
}

namespace {
// Members that a batched launch can step: FP32-mixed, constant mobility, no
// stage timing.
bool batchable(const pf_handle* h) {
  return h->mode == PF_MODE_FP32_MIXED && h->p.D_gb == 0.0 && !h->timing;
}
bool same_shape(const pf_handle* a, const pf_handle* b) {
  const pf_tuning &ta = a->tune, &tb = b->tune;
  return a->g.nx == b->g.nx && a->g.ny == b->g.ny && a->g.nz == b->g.nz && a->g.var == b->g.var &&
         ta.xline_lines == tb.xline_lines && ta.xline_grid == tb.xline_grid && ta.zline_grid == tb.zline_grid &&
         ta.zline_run_rows == tb.zline_run_rows && ta.zline_plan == tb.zline_plan;
}

// A group of batchable members of one shape: its device member table, stepped
// on the first member's stream (the members' own streams are idle by then).
struct Group {
  std::vector<pf_handle*> mem;
  Member<float>* table = nullptr;
};

int upload_table(Group& gr) {
  std::vector<Member<float>> tab(gr.mem.size());
  for (size_t i = 0; i < gr.mem.size(); ++i) {
    pf_handle* h = gr.mem[i];
    tab[i] = Member<float>{(c2<float>*)h->W_main, (c2<float>*)h->W_nyq, (c2<float>*)h->S_main,
                           h->dc64, h->u};
  }
  if (hipMalloc(&gr.table, sizeof(Member<float>) * tab.size()) != hipSuccess) {
    gr.table = nullptr;
    return fail(PF_ENOMEM, "member table allocation failed");
  }
  PF_HIP(hipMemcpyAsync(gr.table, tab.data(), sizeof(Member<float>) * tab.size(), hipMemcpyHostToDevice,
                        gr.mem[0]->s));
  return 0;
}

// One step of every member of the group: four launches.
int step_group(const Group& gr) {
  pf_handle* h0 = gr.mem[0];
  const Geo& g = h0->g;
  const long nmem = (long)gr.mem.size();
  int rc = xline_batch(g, h0->u, h0->tune, gr.table, nmem, h0->s);
  if (!rc) rc = ypass_batch<false>(g, gr.table, nmem, h0->s);
  if (!rc) rc = zline_batch(g, h0->u, h0->tune, h0->kt, gr.table, nmem, h0->s);
  if (!rc) rc = ypass_batch<true>(g, gr.table, nmem, h0->s);
  return rc;
}
}  // namespace

extern "C" int pf_ensemble_step(pf_handle* const* hs, int n, int nsteps) {
  if (n < 0 || nsteps < 0 || (n > 0 && !hs)) return fail(PF_EINVAL, "bad ensemble arguments");
  for (int i = 0; i < n; ++i) {
    if (!hs[i]) return fail(PF_EINVAL, "null handle in ensemble");
    for (int j = 0; j < i; ++j)
      if (hs[j] == hs[i]) return fail(PF_EINVAL, "a handle appears twice in the ensemble");
    if (hs[i]->dev != hs[0]->dev) return fail(PF_EINVAL, "ensemble handles are on different devices");
  }
  DeviceGuard dg(n > 0 ? hs[0]->dev : 0);
  // groups of two or more batchable members of one shape step with one launch
  // per stage; the rest step one by one on their own streams
  std::vector<std::vector<pf_handle*>> groups;
  std::vector<pf_handle*> single;
  for (int i = 0; i < n; ++i) {
    pf_handle* h = hs[i];
    if (!batchable(h)) { single.push_back(h); continue; }
    bool placed = false;
    for (auto& gr : groups)
      if (same_shape(gr[0], h)) { gr.push_back(h); placed = true; break; }
    if (!placed) groups.push_back({h});
  }
  std::vector<Group> batched;
  for (auto& gm : groups) {
    if (gm.size() < 2) single.push_back(gm[0]);
    else batched.push_back(Group{gm, nullptr});
  }
  // work earlier enqueued on the members' own streams comes first
  for (auto& gr : batched)
    for (pf_handle* h : gr.mem) PF_HIP(hipStreamSynchronize(h->s));
  int rc = 0;
  for (auto& gr : batched)
    if (!rc) rc = upload_table(gr);
  // the groups' steps go round-robin to their streams, so groups run concurrently
  for (int st = 0; st < nsteps && !rc; ++st)
    for (auto& gr : batched)
      if (!rc) rc = step_group(gr);
  // the tables must outlive the launches that read them
  for (auto& gr : batched) {
    if (hipStreamSynchronize(gr.mem[0]->s) != hipSuccess && !rc) rc = fail(PF_EHIP, "batched ensemble step failed");
    if (gr.table) (void)hipFree(gr.table);
  }
  if (rc) return rc;
  for (auto& gr : batched)
    for (pf_handle* h : gr.mem) {
      h->fint_ready = false;
      h->step += nsteps;
    }
  // members step without diagnostics: a captured one-step period each
  for (pf_handle* h : single) {
    if (nsteps >= kMinGraphReplays && h->graphs && !h->timing && h->gx_key != 0) {
      const int rc1 = capture_period(h, 0);
      if (rc1) return rc1;
    }
  }
  for (int st = 0; st < nsteps; ++st) {
    for (pf_handle* h : single) {
      if (h->gx && h->gx_key == 0 && !h->timing && nsteps >= kMinGraphReplays) {
        PF_HIP(hipGraphLaunch(h->gx, h->s));
        h->fint_ready = false;
      } else {
        const int rc1 = step_once(h, false, 0, false);
        if (rc1) return rc1;
      }
      h->step += 1;
    }
  }
  PF_HIP(hipGetLastError());
  for (pf_handle* h : single) {
    PF_HIP(hipStreamSynchronize(h->s));
    tflush(h);
  }
  return 0;
}

extern "C" int pf_step(pf_handle* h, int nsteps, int diag_every, pf_diag* out) {
  if (!h) return fail(PF_EINVAL, "null handle");
  DeviceGuard dg(h->dev);
  if (nsteps < 0) return fail(PF_EINVAL, "nsteps < 0");
  int rc = 0;
  if (out && (rc = ensure_diag(h, std::max(1, nsteps)))) return rc;
  std::vector<char> had(nsteps, 0);
  const int de = (out && diag_every > 0) ? diag_every : 0;
  const int P = de > 0 ? de : 1;
  int i = 0;
  while (i < nsteps) {
    const long stepno = h->step + 1;
    // replays of whole periods: the steps from i on follow the captured pattern
    // (the reduced forms carry f_int from the step before a reported one; the
    // faithful form recomputes it, so every step of de = 1 is alike)
    const bool carry = h->mode != PF_MODE_FP64_FAITHFUL;
    long K = 0;
    if (h->graphs && !h->timing) {
      if (de == 0 || (de == 1 && !carry)) K = nsteps - i;
      else if (de == 1) K = (i >= 1 && h->fint_ready) ? (nsteps - 1) - i : 0;
      else if (stepno % de == 1 && !h->fint_ready) K = (nsteps - i) / de;
    }
    if (K >= kMinGraphReplays) {
      if (h->gx_key != de && (rc = capture_period(h, de))) return rc;
    }
    if (K >= kMinGraphReplays && h->gx && h->gx_key == de) {
      if (de > 0) hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, h->s, h->dctr, i + P - 1);
      for (long k = 0; k < K; ++k) PF_HIP(hipGraphLaunch(h->gx, h->s));
      for (long k = 0; k < K; ++k)
        if (de > 0) had[i + k * P + P - 1] = 1;
      h->fint_ready = carry && de == 1;
      h->step += K * P;
      i += (int)(K * P);
      continue;
    }
    const bool diag = de > 0 && (stepno % de) == 0;
    const bool next_diag = de > 0 && ((stepno + 1) % de) == 0 && (i + 1) < nsteps;
    if ((rc = step_once(h, diag, i, next_diag))) return rc;
    had[i] = diag;
    h->step = stepno;
    ++i;
  }
  PF_HIP(hipGetLastError());
  std::vector<double> hd;
  if (out && nsteps > 0) {
    hd.resize(9 * (size_t)nsteps);
    PF_HIP(hipMemcpyAsync(hd.data(), h->ddiag, sizeof(double) * hd.size(), hipMemcpyDeviceToHost, h->s));
  }
  PF_HIP(hipStreamSynchronize(h->s));
  tflush(h);
  if (out) {
    bool bad = false;
    for (int i = 0; i < nsteps; ++i) {
      pf_diag& d = out[i];
      d.step = (int)(h->step - nsteps + 1 + i);
      d.valid = had[i];
      const double* r = hd.data() + 9 * (size_t)i;
      if (had[i]) {
        d.f_int = r[0]; d.f_ch = r[1]; d.f_eta = r[2]; d.e_tot = r[3];
        d.c_max = r[4]; d.c_min = r[5]; d.phi_min = r[6]; d.phi_3_max = r[7]; d.c_mean = r[8];
        if (!std::isfinite(d.e_tot) || !std::isfinite(d.c_mean)) bad = true;
      } else {
        d.f_int = d.f_ch = d.f_eta = d.e_tot = d.c_max = d.c_min = d.phi_min = d.phi_3_max = d.c_mean = NAN;
      }
    }
    if (bad) return fail(PF_ENONFINITE, "non-finite energy or mass in diagnostics");
  }
  return 0;
}

extern "C" int pf_diagnose(pf_handle* h, pf_diag* out) {
  if (!h || !out) return fail(PF_EINVAL, "null argument");
  DeviceGuard dg(h->dev);
  const Geo& g = h->g;
  hipStream_t s = h->s;
  int rc = ensure_diag(h, 1);
  if (rc) return rc;
  const long rows = (long)g.plane * g.F;
  const unsigned gr_nl = grid_for(g.P, NB);
  if (h->mode == PF_MODE_FP64_FAITHFUL) {
    const size_t shb = sizeof(double) * ((size_t)g.var * NB + NB);
    hipLaunchKernelGGL((k_nl_faithful<true>), dim3(gr_nl), dim3(NB), shb, s, g, h->u,
                       (const double*)h->fs.main, (double*)h->fdfdc.main, (double*)h->fpsi.main, h->fmob,
                       h->part_nl);
    h->n_part_nl = gr_nl;
    // f_int by Parseval on a transformed copy of the state (same quantity as loops 1-4)
    PF_HIP(hipMemcpyAsync(h->fgrad.main, h->fs.main, sizeof(double) * (size_t)g.P * g.F,
                          hipMemcpyDeviceToDevice, s));
    if ((rc = fwd3d<double>(h->fgrad.main, h->fgrad.nyq, g.nx, g.ny, g.nz, g.F, 1.0 / (double)(float)g.P, s)))
      return rc;
    hipLaunchKernelGGL((k_update<double, false, false, true, false>), dim3(grid_for(rows, 1)), dim3(NB), 0, s, g,
                       h->u, h->kt, h->fgrad, h->fgrad, h->fgrad, h->dc64, h->part_fi, rows);
    h->n_part_fi = grid_for(rows, 1);
    diag_final(h, 0, true, true, (const double*)h->fgrad.main, nullptr);
  } else {
    const bool f32 = h->mode == PF_MODE_FP32_MIXED;
    if (f32) {
      Bank<float> S{(c2<float>*)h->S_main, nullptr};
      hipLaunchKernelGGL((k_update<float, false, false, true, false, true>), dim3(grid_for(rows, 1)), dim3(NB), 0, s, g,
                         h->u, h->kt, S, S, S, h->dc64, h->part_fi, rows);
      const size_t shb = sizeof(float) * (size_t)g.var * NB + 16 + sizeof(double) * NB;
      if ((rc = real_state_reduced<float>(h))) return rc;
      hipLaunchKernelGGL((k_nl_reduced<float, true, false, false>), dim3(gr_nl), dim3(NB), shb, s, g, h->u,
                         (float*)h->X_main, (float*)nullptr, h->part_nl);
    } else {
      Bank<double> S{(c2<double>*)h->S_main, nullptr};
      hipLaunchKernelGGL((k_update<double, false, false, true, false, true>), dim3(grid_for(rows, 1)), dim3(NB), 0, s, g,
                         h->u, h->kt, S, S, S, h->dc64, h->part_fi, rows);
      const size_t shb = sizeof(double) * (size_t)g.var * NB + 16 + sizeof(double) * NB;
      if ((rc = real_state_reduced<double>(h))) return rc;
      hipLaunchKernelGGL((k_nl_reduced<double, true, false, false>), dim3(gr_nl), dim3(NB), shb, s, g, h->u,
                         (double*)h->X_main, (double*)nullptr, h->part_nl);
    }
    h->n_part_fi = grid_for(rows, 1);
    h->n_part_nl = gr_nl;
    h->fint_ready = true;
    diag_final(h, 0, true, true, f32 ? h->dc64 : (const double*)h->S_main, nullptr);
  }
  PF_HIP(hipGetLastError());
  double r[9];
  PF_HIP(hipMemcpyAsync(r, h->ddiag, sizeof(r), hipMemcpyDeviceToHost, s));
  PF_HIP(hipStreamSynchronize(s));
  out->step = (int)(h->step + 1);
  out->valid = 1;
  out->f_int = r[0]; out->f_ch = r[1]; out->f_eta = r[2]; out->e_tot = r[3];
  out->c_max = r[4]; out->c_min = r[5]; out->phi_min = r[6]; out->phi_3_max = r[7]; out->c_mean = r[8];
  if (!std::isfinite(out->e_tot) || !std::isfinite(out->c_mean))
    return fail(PF_ENONFINITE, "non-finite energy or mass in diagnostics");
  return 0;
}

extern "C" int pf_set_timing(pf_handle* h, int on) {
  if (!h) return fail(PF_EINVAL, "null handle");
  h->timing = on != 0;
  h->timing_mask = ~0u;
  for (int i = 0; i < NSTAGE; ++i) { h->st_ms[i] = 0.0; h->st_cnt[i] = 0; }
  return 0;
}

extern "C" int pf_set_timing_stages(pf_handle* h, unsigned mask) {
  if (!h) return fail(PF_EINVAL, "null handle");
  if (mask == 0) return fail(PF_EINVAL, "empty stage mask");
  h->timing = true;
  h->timing_mask = mask;
  for (int i = 0; i < NSTAGE; ++i) { h->st_ms[i] = 0.0; h->st_cnt[i] = 0; }
  return 0;
}

extern "C" int pf_stage_times(pf_handle* h, double* ms, long* count, int n) {
  if (!h) return fail(PF_EINVAL, "null handle");
  for (int i = 0; i < n && i < NSTAGE; ++i) {
    if (ms) ms[i] = h->st_ms[i];
    if (count) count[i] = h->st_cnt[i];
  }
  return NSTAGE;
}

extern "C" int pf_stage_name(int stage, char* buf, int len) {
  if (stage < 0 || stage >= NSTAGE || !buf || len < 1) return fail(PF_EINVAL, "bad stage");
  std::snprintf(buf, (size_t)len, "%s", kStageNames[stage]);
  return 0;
}

extern "C" double pf_stage_bytes(pf_handle* h, int stage) {
  if (!h || stage < 0 || stage >= NSTAGE) return 0.0;
  const Geo& g = h->g;
  const double s = (h->mode == PF_MODE_FP32_MIXED) ? 4.0 : 8.0;
  const double Br = (double)g.P * s;                       // one real field
  const double Bc = 2.0 * (double)(g.M + 1) * g.plane * s; // one half spectrum
  const double F = g.F;
  switch (stage) {
    case ST_NONLINEAR: return 2.0 * F * Br + (h->p.D_gb != 0.0 ? Br : 0.0);
    case ST_X_R2C: return F * (Br + Bc);
    case ST_Y_FWD: case ST_Z_FWD: case ST_Z_INV: case ST_Y_INV: return 2.0 * F * Bc;
    case ST_X_C2R: return F * (Bc + Br);
    case ST_UPDATE: return 4.0 * F * Bc;
    default: return 0.0;
  }
}

extern "C" void* pf_stream(pf_handle* h) { return h ? (void*)h->s : nullptr; }
