/*
 * CPU FP64 oracle: a from-scratch restatement of the reference timestep.
 *
 * TEST INFRASTRUCTURE ONLY.  It is the checker for the HIP product path and the
 * CPU baseline that bench.py times; the product never links or calls it.
 *
 * What each part restates (file:line in the reference tree):
 *   ora_ran2          src/ran_2.f90:1-40        (L'Ecuyer + Bays-Durham shuffle)
 *   ora_kspace_table  src/k_space.f90:13-59     (wavenumbers, Nyquist rules)
 *   ora_fft_r2c       src/f_trans.f90:11-12     (3-D R2C, e^{-i}, then / float(Nx*Ny*Nz))
 *   ora_fft_c2r       src/inv_trans.f90:10      (3-D C2R, unnormalised)
 *   ora_gamma         src/var_diff.f90:198-218
 *   ora_run           src/var_diff.f90:474-1118 (loops 1-14, summary energies, extrema)
 *   ora_pointwise     src/var_diff.f90:685-905  (loops 5-10 pointwise parts)
 *
 * The transform is this file's own mixed-radix (2,3,4,5) Stockham FFT; P3DFFT
 * and FFTW are not available, so their rounding order is not reproduced.  Each
 * 3-D transform runs x, then y, then z (forward) and z, y, x (inverse), which is
 * P3DFFT's order for the 'fft'/'tff' option strings.
 *
 * Floating-point expressions follow the reference's evaluation order where it is
 * visible in the source (left-to-right sums, (0.5*kappa)*(...) etc.).  The file is
 * compiled with -ffp-contract=off so that no FMA contraction happens.
 */
#include "pf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double r, i; } cx;

/* ------------------------------------------------------------------------- */
/* ran_2 (src/ran_2.f90:1-40).                                                */
/* ------------------------------------------------------------------------- */
enum {
  R2_IM1 = 2147483563, R2_IM2 = 2147483399, R2_IMM1 = R2_IM1 - 1,
  R2_IA1 = 40014, R2_IA2 = 40692, R2_IQ1 = 53668, R2_IQ2 = 52774,
  R2_IR1 = 12211, R2_IR2 = 3791, R2_NTAB = 32, R2_NDIV = 1 + R2_IMM1 / R2_NTAB
};

void ora_ran2_init(ora_ran2_state* st) {
  /* DATA statement of src/ran_2.f90:12 */
  st->idum2 = 123456789;
  st->iy = 0;
  for (int j = 0; j < R2_NTAB; ++j) st->iv[j] = 0;
}

double ora_ran2(ora_ran2_state* st, int* idum) {
  /* am = 1./im1 is a default-REAL (single precision) expression in the
   * reference's PARAMETER statement (src/ran_2.f90:5), so the constant is the
   * float32 reciprocal of float32(im1), widened to double. */
  const double am = (double)(1.0f / (float)R2_IM1);
  const double rnmx = 1.0 - 1.2e-7;
  int k;
  if (*idum <= 0) {
    *idum = (-*idum > 1) ? -*idum : 1;
    st->idum2 = *idum;
    for (int j = R2_NTAB + 8; j >= 1; --j) {
      k = *idum / R2_IQ1;
      *idum = R2_IA1 * (*idum - k * R2_IQ1) - k * R2_IR1;
      if (*idum < 0) *idum += R2_IM1;
      if (j <= R2_NTAB) st->iv[j - 1] = *idum;
    }
    st->iy = st->iv[0];
  }
  k = *idum / R2_IQ1;
  *idum = R2_IA1 * (*idum - k * R2_IQ1) - k * R2_IR1;
  if (*idum < 0) *idum += R2_IM1;
  k = st->idum2 / R2_IQ2;
  st->idum2 = R2_IA2 * (st->idum2 - k * R2_IQ2) - k * R2_IR2;
  if (st->idum2 < 0) st->idum2 += R2_IM2;
  int j = 1 + st->iy / R2_NDIV;
  st->iy = st->iv[j - 1] - st->idum2;
  st->iv[j - 1] = *idum;
  if (st->iy < 1) st->iy += R2_IMM1;
  double v = am * (double)st->iy;
  return v < rnmx ? v : rnmx;
}

/* ------------------------------------------------------------------------- */
/* FFT: mixed-radix Stockham, radices 4, 2, 3, 5.                             */
/* ------------------------------------------------------------------------- */
#define MAXN 2048
#define MAXR 40

typedef struct {
  int n, nr, rad[MAXR];
  cx* tw; /* tw[t] = exp(-2 pi i t / n), t in [0, n) */
} fplan;

static fplan* g_plans[MAXN + 1];

/* exp(-2 pi i t/n) with the argument reduced to the first octant so that
 * values at multiples of n/8 are exactly symmetric (W[n/4] = -i exactly). */
static cx twiddle(long t, long n) {
  t %= n;
  if (t < 0) t += n;
  long q = (4 * t) / n;          /* quadrant */
  long num = 4 * t - q * n;      /* remainder angle = 2pi*num/(4n), in [0, pi/2) */
  long double c, s;
  const long double two_pi = 6.283185307179586476925286766559L;
  if (2 * num <= n) {
    long double a = two_pi * (long double)num / (long double)(4 * n);
    c = cosl(a); s = sinl(a);
  } else {
    long double a = two_pi * (long double)(n - num) / (long double)(4 * n);
    c = sinl(a); s = cosl(a);
  }
  /* e^{-i x}: (c, -s); multiply by (-i)^q */
  double cr = (double)c, ci = -(double)s;
  cx w;
  switch (q) {
    case 0: w.r = cr; w.i = ci; break;
    case 1: w.r = ci; w.i = -cr; break;     /* (-i)(cr + i ci) = ci - i cr */
    case 2: w.r = -cr; w.i = -ci; break;
    default: w.r = -ci; w.i = cr; break;
  }
  return w;
}

void ora_twiddles(int n, double* w) {
  for (int t = 0; t < n; ++t) {
    cx v = twiddle(t, n);
    w[2 * t] = v.r; w[2 * t + 1] = v.i;
  }
}

static fplan* make_plan(int n) {
  fplan* p = (fplan*)calloc(1, sizeof(fplan));
  p->n = n;
  int m = n;
  const int order[4] = {4, 2, 3, 5};
  for (int o = 0; o < 4; ++o)
    while (m % order[o] == 0) { p->rad[p->nr++] = order[o]; m /= order[o]; }
  if (m != 1) { free(p); return NULL; }
  p->tw = (cx*)malloc(sizeof(cx) * (size_t)n);
  for (int t = 0; t < n; ++t) p->tw[t] = twiddle(t, n);
  return p;
}

static fplan* get_plan(int n) {
  if (n < 1 || n > MAXN) return NULL;
  fplan* p;
#pragma omp critical(ora_plan)
  {
    if (!g_plans[n]) g_plans[n] = make_plan(n);
    p = g_plans[n];
  }
  return p;
}

static inline cx cadd(cx a, cx b) { cx c = {a.r + b.r, a.i + b.i}; return c; }
static inline cx csub(cx a, cx b) { cx c = {a.r - b.r, a.i - b.i}; return c; }
static inline cx cmul(cx a, cx b) { cx c = {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; return c; }
static inline cx cconj(cx a) { cx c = {a.r, -a.i}; return c; }
/* multiply by -i*sg  (sg = +1 forward, -1 inverse) */
static inline cx mul_mi(cx a, double sg) { cx c = {sg * a.i, -sg * a.r}; return c; }

/* One line, in place.  a, b: n complex each (b is scratch). */
static void fft_line(const fplan* p, cx* a, cx* b, int inverse) {
  const int n = p->n;
  const double sg = inverse ? -1.0 : 1.0;
  cx* src = a; cx* dst = b;
  int Ls = 1;
  for (int s = 0; s < p->nr; ++s) {
    const int r = p->rad[s], nb = n / r, tstep = n / (Ls * r);
    for (int j = 0; j < nb; ++j) {
      const int k = j % Ls;
      cx v[5];
      for (int m = 0; m < r; ++m) {
        v[m] = src[j + m * nb];
        if (k && m) {
          cx w = p->tw[m * k * tstep];
          if (inverse) w = cconj(w);
          v[m] = cmul(v[m], w);
        }
      }
      cx y[5];
      if (r == 2) {
        y[0] = cadd(v[0], v[1]); y[1] = csub(v[0], v[1]);
      } else if (r == 4) {
        cx t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        cx t2 = cadd(v[1], v[3]), t3 = mul_mi(csub(v[1], v[3]), sg);
        y[0] = cadd(t0, t2); y[2] = csub(t0, t2);
        y[1] = cadd(t1, t3); y[3] = csub(t1, t3);
      } else if (r == 3) {
        const double s3 = 0.86602540378443864676;
        cx t1 = cadd(v[1], v[2]);
        cx t2 = {v[0].r - 0.5 * t1.r, v[0].i - 0.5 * t1.i};
        cx d = csub(v[1], v[2]);
        cx t3 = mul_mi(d, sg); t3.r *= s3; t3.i *= s3;
        y[0] = cadd(v[0], t1); y[1] = cadd(t2, t3); y[2] = csub(t2, t3);
      } else { /* r == 5 */
        const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
        const double s1 = 0.95105651629515357212, s2 = 0.58778525229247312917;
        cx a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
        cx a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
        cx p1 = {v[0].r + c1 * a1.r + c2 * a2.r, v[0].i + c1 * a1.i + c2 * a2.i};
        cx p2 = {v[0].r + c2 * a1.r + c1 * a2.r, v[0].i + c2 * a1.i + c1 * a2.i};
        cx q1 = {s1 * b1.r + s2 * b2.r, s1 * b1.i + s2 * b2.i};
        cx q2 = {s2 * b1.r - s1 * b2.r, s2 * b1.i - s1 * b2.i};
        cx m1 = mul_mi(q1, sg), m2 = mul_mi(q2, sg);
        y[0].r = v[0].r + a1.r + a2.r; y[0].i = v[0].i + a1.i + a2.i;
        y[1] = cadd(p1, m1); y[4] = csub(p1, m1);
        y[2] = cadd(p2, m2); y[3] = csub(p2, m2);
      }
      const int ob = (j - k) * r + k;
      for (int m = 0; m < r; ++m) dst[ob + m * Ls] = y[m];
    }
    cx* t = src; src = dst; dst = t;
    Ls *= r;
  }
  if (src != a) memcpy(a, src, sizeof(cx) * (size_t)n);
}

int ora_fft_c2c_line(int n, const double* in, double* out, int inverse) {
  fplan* p = get_plan(n);
  if (!p) return -1;
  cx* a = (cx*)malloc(sizeof(cx) * (size_t)n * 2);
  memcpy(a, in, sizeof(cx) * (size_t)n);
  fft_line(p, a, a + n, inverse);
  memcpy(out, a, sizeof(cx) * (size_t)n);
  free(a);
  return 0;
}

/* y and z passes over the half spectrum h (mx, ny, nz). */
static void fft_yz(cx* h, int mx, int ny, int nz, int inverse, int z_first) {
  fplan* py = get_plan(ny);
  fplan* pz = get_plan(nz);
  for (int pass = 0; pass < 2; ++pass) {
    int do_z = z_first ? (pass == 0) : (pass == 1);
    if (!do_z) {
#pragma omp parallel
      {
        cx* buf = (cx*)malloc(sizeof(cx) * (size_t)ny * 2);
#pragma omp for schedule(static)
        for (int z = 0; z < nz; ++z)
          for (int x = 0; x < mx; ++x) {
            cx* base = h + (size_t)z * mx * ny + x;
            for (int y = 0; y < ny; ++y) buf[y] = base[(size_t)y * mx];
            fft_line(py, buf, buf + ny, inverse);
            for (int y = 0; y < ny; ++y) base[(size_t)y * mx] = buf[y];
          }
        free(buf);
      }
    } else {
#pragma omp parallel
      {
        cx* buf = (cx*)malloc(sizeof(cx) * (size_t)nz * 2);
#pragma omp for schedule(static)
        for (int y = 0; y < ny; ++y)
          for (int x = 0; x < mx; ++x) {
            cx* base = h + (size_t)y * mx + x;
            const size_t st = (size_t)mx * ny;
            for (int z = 0; z < nz; ++z) buf[z] = base[z * st];
            fft_line(pz, buf, buf + nz, inverse);
            for (int z = 0; z < nz; ++z) base[z * st] = buf[z];
          }
        free(buf);
      }
    }
  }
}

/* f_trans: forward R2C (x, then y, then z), then divide by float(nx*ny*nz)
 * (src/f_trans.f90:11-12).  out: (nx/2+1, ny, nz) complex. */
int ora_fft_r2c(int nx, int ny, int nz, const double* in, double* outd, int scale) {
  if (nx % 2) return -1;
  fplan* px = get_plan(nx);
  if (!px || !get_plan(ny) || !get_plan(nz)) return -1;
  const int mx = nx / 2 + 1;
  cx* out = (cx*)outd;
#pragma omp parallel
  {
    cx* buf = (cx*)malloc(sizeof(cx) * (size_t)nx * 2);
#pragma omp for schedule(static)
    for (int l = 0; l < ny * nz; ++l) {
      const double* src = in + (size_t)l * nx;
      for (int x = 0; x < nx; ++x) { buf[x].r = src[x]; buf[x].i = 0.0; }
      fft_line(px, buf, buf + nx, 0);
      memcpy(out + (size_t)l * mx, buf, sizeof(cx) * (size_t)mx);
    }
    free(buf);
  }
  fft_yz(out, mx, ny, nz, 0, 0);
  if (scale) {
    /* float(Nx*Ny*Nz) is a default-REAL conversion; it is exact for every
     * grid with Nx*Ny*Nz <= 2^24 and for the power-of-two grids above it. */
    const double P = (double)(float)((long)nx * ny * nz);
    const size_t nh = (size_t)mx * ny * nz;
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < nh; ++t) { out[t].r = out[t].r / P; out[t].i = out[t].i / P; }
  }
  return 0;
}

/* inv_trans: inverse C2R (z, then y, then x), unnormalised (src/inv_trans.f90:10).
 * The imaginary parts of the kx = 0 and kx = nx/2 entries of each x line are
 * ignored, as a Hermitian (half-complex) line has no slot for them. */
int ora_fft_c2r(int nx, int ny, int nz, const double* ind, double* out) {
  if (nx % 2) return -1;
  fplan* px = get_plan(nx);
  if (!px || !get_plan(ny) || !get_plan(nz)) return -1;
  const int mx = nx / 2 + 1;
  const size_t nh = (size_t)mx * ny * nz;
  cx* h = (cx*)malloc(sizeof(cx) * nh);
  memcpy(h, ind, sizeof(cx) * nh);
  fft_yz(h, mx, ny, nz, 1, 1);
#pragma omp parallel
  {
    cx* buf = (cx*)malloc(sizeof(cx) * (size_t)nx * 2);
#pragma omp for schedule(static)
    for (int l = 0; l < ny * nz; ++l) {
      const cx* src = h + (size_t)l * mx;
      for (int k = 0; k < mx; ++k) buf[k] = src[k];
      for (int k = mx; k < nx; ++k) buf[k] = cconj(src[nx - k]);
      fft_line(px, buf, buf + nx, 1);
      double* dst = out + (size_t)l * nx;
      for (int x = 0; x < nx; ++x) dst[x] = buf[x].r;
    }
    free(buf);
  }
  free(h);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Wavenumbers (src/k_space.f90:13-59).                                       */
/* ------------------------------------------------------------------------- */
typedef struct { double a, ap; } kax; /* a: signed wavenumber; ap: Nyquist-zeroed */

static kax kaxis(int m, int n) {
  /* unit = 2*pi*i / float(n) with pi = atan(1)*4; lamda = float(m~) * unit */
  const double pi = atan(1.0) * 4.0;
  const double unit = (2.0 * pi) / (double)(float)n;
  int ms = (m <= n / 2) ? m : m - n;
  kax r;
  r.a = (double)(float)ms * unit;
  r.ap = (2 * m == n) ? 0.0 : r.a;
  return r;
}

/* Per-point combinations used by loops 11, 13 and 14
 * (src/var_diff.f90:932-935, 978-981, 1019). */
typedef struct { double s1, k2, k3, k4, kp2; } kpt;

static kpt kpoint(kax ax, kax ay, kax az) {
  kpt k;
  k.s1 = (ax.ap + ay.ap) + az.ap;                      /* kf_sum = i*s1          */
  k.k2 = (-(ax.a * ax.a) + -(ay.a * ay.a)) + -(az.a * az.a); /* k_sq (real)      */
  k.k3 = (-(ax.ap * (ax.a * ax.a)) + -(ay.ap * (ay.a * ay.a))) + -(az.ap * (az.a * az.a)); /* k_cube = i*k3 */
  double x4 = ((ax.a * ax.a) * ax.a) * ax.a;
  double y4 = ((ay.a * ay.a) * ay.a) * ay.a;
  double z4 = ((az.a * az.a) * az.a) * az.a;
  k.k4 = (x4 + y4) + z4;                               /* k_4 (real)             */
  k.kp2 = (ax.ap * ax.ap + ay.ap * ay.ap) + az.ap * az.ap; /* |kf|^2, Nyquist-zeroed */
  return k;
}

/* kf, kf_sq, kf_4 tables in the reference's layout (3, nz, ny, nx/2+1),
 * complex interleaved. */
void ora_kspace_table(int nx, int ny, int nz, double* kf, double* kf_sq, double* kf_4) {
  const int mx = nx / 2 + 1;
  for (int ix = 0; ix < mx; ++ix)
    for (int iy = 0; iy < ny; ++iy)
      for (int iz = 0; iz < nz; ++iz) {
        kax a[3] = {kaxis(ix, nx), kaxis(iy, ny), kaxis(iz, nz)};
        size_t o = 3 * ((size_t)iz + (size_t)nz * ((size_t)iy + (size_t)ny * ix));
        for (int d = 0; d < 3; ++d) {
          double v = a[d].a;
          kf[2 * (o + d)] = 0.0;    kf[2 * (o + d) + 1] = a[d].ap;
          kf_sq[2 * (o + d)] = -(v * v); kf_sq[2 * (o + d) + 1] = 0.0;
          kf_4[2 * (o + d)] = ((v * v) * v) * v; kf_4[2 * (o + d) + 1] = 0.0;
        }
      }
}

/* gamma(i,j), column-major (var x var), src/var_diff.f90:198-218. */
void ora_gamma(int var, double sigma_1, double sigma_2, double* g) {
  const int mat_var = var - 1;
  for (int j = 1; j <= var; ++j)
    for (int i = 1; i <= var; ++i) {
      double v = 0.0;
      if (i != j && i <= mat_var && j <= mat_var) v = sigma_1;
      if (j == var && i < var) v = sigma_2;
      if (i == var && j < var) v = sigma_2;
      g[(i - 1) + (size_t)var * (j - 1)] = v;
    }
}

/* ------------------------------------------------------------------------- */
/* Pointwise parts of loops 5-10 for one point.                               */
/* ------------------------------------------------------------------------- */
typedef struct { double f_eta, f_ch, df_dc, mob; } pt_out;

static pt_out point_terms(const ora_params* p, int var, const double* g,
                          const double* phi, double con, double* df_dphi) {
  pt_out o;
  const int mat_var = var - 1;
  /* loop 5 (src/var_diff.f90:697-714) */
  double fe = 0.0;
  for (int i = 0; i < var; ++i) {
    double fi = phi[i];
    double fi2 = fi * fi;
    fe = (fe - 0.5 * fi2) + 0.25 * (fi2 * fi2);
    for (int j = i + 1; j < var; ++j) {
      double fj = phi[j];
      fe = fe + ((g[i + (size_t)var * j] * fi2) * (fj * fj));
    }
  }
  o.f_eta = fe;
  /* loop 6 (src/var_diff.f90:735-752) */
  double num = 0.0, den = 0.0;
  for (int i = 0; i < var; ++i) {
    double f2 = phi[i] * phi[i];
    double f4 = f2 * f2;
    if (i < var - 1) num = num + f4;
    den = den + f4;
  }
  double phi_mat = num / den;
  double phi_ppt = 1.0 - phi_mat;
  double dm = con - p->con_0_mat, dp = con - p->con_0_ppt;
  double fch_mat = (phi_mat * p->am) * (dm * dm);
  double fch_ppt = (phi_ppt * p->ap) * (dp * dp);
  o.f_ch = fch_mat + fch_ppt;
  o.df_dc = ((2.0 * phi_mat) * p->am) * dm + ((2.0 * phi_ppt) * p->ap) * dp;
  /* loops 7-9 (src/var_diff.f90:779-858) */
  double fm = p->am * (dm * dm), fp = p->ap * (dp * dp);
  double fv = phi[var - 1];
  double fv2 = fv * fv, fv3 = fv2 * fv, fv4 = fv2 * fv2;
  double t2v = (4.0 * fv3) * (den - fv4);
  double den2 = den * den;
  for (int i = 0; i < var; ++i) {
    double fi = phi[i];
    double term1 = (fi * fi) * fi - fi;
    double term2 = 0.0;
    for (int j = 0; j < var; ++j) {
      if (j == i) continue;
      double fj = phi[j];
      term2 = term2 + ((2.0 * g[i + (size_t)var * j]) * fi) * (fj * fj);
    }
    double term4;
    if (i < mat_var) {
      double t1 = (4.0 * ((fi * fi) * fi)) * fv4;
      term4 = (t1 * (fm - fp)) / den2;
    } else {
      term4 = (t2v * (fp - fm)) / den2;
    }
    df_dphi[i] = (term1 + term2) + term4;
  }
  /* loop 10 mobility (src/var_diff.f90:893-898); D_bulk(var) = D_L */
  double ps = 0.0;
  for (int i = 0; i < var; ++i) ps = ps + phi[i] * phi[i];
  o.mob = p->D_L + ((2.0 * (1.0 - ps)) * p->D_gb);
  return o;
}

int ora_pointwise(const ora_params* p, int npts, int var, const double* phi,
                  const double* con, double* df_dc, double* df_dphi,
                  double* mob_c, double* f_ch, double* f_eta) {
  double* g = (double*)malloc(sizeof(double) * (size_t)var * var);
  ora_gamma(var, p->sigma_1, p->sigma_2, g);
  double sch = 0.0, seta = 0.0;
  for (int t = 0; t < npts; ++t) {
    pt_out o = point_terms(p, var, g, phi + (size_t)t * var, con[t], df_dphi + (size_t)t * var);
    df_dc[t] = o.df_dc;
    mob_c[t] = o.mob;
    sch += o.f_ch;
    seta += o.f_eta;
  }
  if (f_ch) *f_ch = sch;
  if (f_eta) *f_eta = seta;
  free(g);
  return 0;
}

/* Extrema (src/var_diff.f90:1076-1094), with the reference's initial values:
 * out = {c_max, c_min, phi_min, phi_3}. */
void ora_extrema(int npts, int var, const double* phi, const double* con, double* out) {
  double phi_min = 1.0, phi_3 = 0.0, c_max = 0.0, c_min = 1.0;
  for (int t = 0; t < npts; ++t) {
    const double* f = phi + (size_t)t * var;
    if (f[var - 1] > phi_3) phi_3 = f[var - 1];
    if (con[t] <= c_min) c_min = con[t];
    if (con[t] > c_max) c_max = con[t];
    double sq = 0.0;
    for (int i = 0; i < var; ++i) sq = sq + f[i] * f[i];
    if (sq < phi_min) phi_min = sq;
  }
  out[0] = c_max; out[1] = c_min; out[2] = phi_min; out[3] = phi_3;
}

/* ------------------------------------------------------------------------- */
/* One or more timesteps in the reference's order (src/var_diff.f90:556-1118). */
/* ------------------------------------------------------------------------- */
int ora_run(const ora_params* p, int nx, int ny, int nz, int var,
            double* phi, double* con, int nsteps, ora_diag* diag) {
  if (nx % 2 || var < 1) return -1;
  const int mx = nx / 2 + 1;
  const size_t P = (size_t)nx * ny * nz, H = (size_t)mx * ny * nz;
  double* g = (double*)malloc(sizeof(double) * (size_t)var * var);
  ora_gamma(var, p->sigma_1, p->sigma_2, g);

  kpt* kt = (kpt*)malloc(sizeof(kpt) * H);
  double* kap = (double*)malloc(sizeof(double) * 3 * H); /* Nyquist-zeroed a_x, a_y, a_z */
  for (int iz = 0; iz < nz; ++iz)
    for (int iy = 0; iy < ny; ++iy)
      for (int ix = 0; ix < mx; ++ix) {
        size_t o = (size_t)ix + (size_t)mx * ((size_t)iy + (size_t)ny * iz);
        kax ax = kaxis(ix, nx), ay = kaxis(iy, ny), az = kaxis(iz, nz);
        kt[o] = kpoint(ax, ay, az);
        kap[3 * o] = ax.ap; kap[3 * o + 1] = ay.ap; kap[3 * o + 2] = az.ap;
      }

  double* dummy = (double*)malloc(sizeof(double) * P);
  double* gradr = (double*)malloc(sizeof(double) * P);
  double* df_dc = (double*)malloc(sizeof(double) * P);
  double* mob_c = (double*)malloc(sizeof(double) * P);
  double* df_dphi = (double*)malloc(sizeof(double) * P * var);
  double* pf_ch = (double*)malloc(sizeof(double) * P);
  double* pf_eta = (double*)malloc(sizeof(double) * P);
  double* plane = (double*)malloc(sizeof(double) * (size_t)nz);
  cx* spec = (cx*)malloc(sizeof(cx) * H);
  cx* sgrad = (cx*)malloc(sizeof(cx) * H);
  cx* dft_con = (cx*)malloc(sizeof(cx) * H);
  cx* dft_dfdc = (cx*)malloc(sizeof(cx) * H);
  cx* dft_dummy = (cx*)malloc(sizeof(cx) * H);
  cx* spec2 = (cx*)malloc(sizeof(cx) * H);
  if (!dft_dummy || !spec2 || !df_dphi) return -2;

  /* extrema of the initial state: the reference's phi_min_glob starts at 1.0
   * and the other three are only printed after step 1 has set them. */
  double ext[4];
  ora_extrema((int)P, var, phi, con, ext);

  for (int step = 0; step < nsteps; ++step) {
    double f_int = 0.0;
    /* loops 1-2: gradient energy of each order parameter */
    for (int iv = 0; iv < var; ++iv) {
#pragma omp parallel for schedule(static)
      for (size_t t = 0; t < P; ++t) dummy[t] = phi[t * var + iv];
      ora_fft_r2c(nx, ny, nz, dummy, (double*)spec, 1);
      double part = 0.0;
      for (int d = 0; d < 3; ++d) {
#pragma omp parallel for schedule(static)
        for (size_t t = 0; t < H; ++t) {
          double a = kap[3 * t + d];
          sgrad[t].r = -(spec[t].i * a);
          sgrad[t].i = spec[t].r * a;
        }
        ora_fft_c2r(nx, ny, nz, (double*)sgrad, gradr);
        /* grad_x, grad_y, grad_z are kept per point in the reference and
         * summed as ((gx^2 + gy^2) + gz^2); accumulate the same way. */
        if (d == 0) memcpy(dummy, gradr, sizeof(double) * P);
        else if (d == 1) {
#pragma omp parallel for schedule(static)
          for (size_t t = 0; t < P; ++t) dummy[t] = dummy[t] * dummy[t] + gradr[t] * gradr[t];
        }
        else {
#pragma omp parallel for schedule(static)
          for (size_t t = 0; t < P; ++t) dummy[t] = dummy[t] + gradr[t] * gradr[t];
        }
      }
      (void)part;
#pragma omp parallel for schedule(static)
      for (int z = 0; z < nz; ++z) {
        double s = 0.0;
        for (size_t t = (size_t)z * nx * ny; t < (size_t)(z + 1) * nx * ny; ++t)
          s = s + (0.5 * p->grad_coeff_phi) * dummy[t];
        plane[z] = s;
      }
      for (int z = 0; z < nz; ++z) f_int = f_int + plane[z];
    }
    /* loops 3-4: gradient energy of c */
    ora_fft_r2c(nx, ny, nz, con, (double*)dft_con, 1);
    for (int d = 0; d < 3; ++d) {
#pragma omp parallel for schedule(static)
      for (size_t t = 0; t < H; ++t) {
        double a = kap[3 * t + d];
        sgrad[t].r = -(dft_con[t].i * a);
        sgrad[t].i = dft_con[t].r * a;
      }
      ora_fft_c2r(nx, ny, nz, (double*)sgrad, gradr);
      if (d == 0) memcpy(dummy, gradr, sizeof(double) * P);
      else if (d == 1) {
#pragma omp parallel for schedule(static)
        for (size_t t = 0; t < P; ++t) dummy[t] = dummy[t] * dummy[t] + gradr[t] * gradr[t];
      }
      else {
#pragma omp parallel for schedule(static)
        for (size_t t = 0; t < P; ++t) dummy[t] = dummy[t] + gradr[t] * gradr[t];
      }
    }
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z) {
      double s = 0.0;
      for (size_t t = (size_t)z * nx * ny; t < (size_t)(z + 1) * nx * ny; ++t)
        s = s + (0.5 * p->grad_coeff_c) * dummy[t];
      plane[z] = s;
    }
    for (int z = 0; z < nz; ++z) f_int = f_int + plane[z];

    /* loops 5-10 pointwise parts */
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < P; ++t) {
      pt_out o = point_terms(p, var, g, phi + t * var, con[t], df_dphi + t * var);
      df_dc[t] = o.df_dc;
      mob_c[t] = o.mob;
      pf_ch[t] = o.f_ch;
      pf_eta[t] = o.f_eta;
    }
    double f_ch = 0.0, f_eta = 0.0;
    for (size_t t = 0; t < P; ++t) { f_eta = f_eta + pf_eta[t]; f_ch = f_ch + pf_ch[t]; }

    if (diag) {
      ora_diag* d = diag + step;
      d->f_int = f_int; d->f_ch = f_ch; d->f_eta = f_eta;
      d->e_tot = (f_int + f_ch) + f_eta;
      d->c_max = ext[0]; d->c_min = ext[1]; d->phi_min = ext[2]; d->phi_3_max = ext[3];
    }

    /* loop 10 transforms (F(c) recomputed as in the reference) */
    ora_fft_r2c(nx, ny, nz, con, (double*)dft_con, 1);
    ora_fft_r2c(nx, ny, nz, df_dc, (double*)dft_dfdc, 1);
    const double D = p->D_L; /* D_mean = D_L (src/var_diff.f90:916) */
    /* loop 11: g^ = kf_sum*f^ - kappa_c*c^*k_cube */
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < H; ++t) {
      double s = kt[t].s1, q = kt[t].k3;
      cx f = dft_dfdc[t];
      cx kc = {p->grad_coeff_c * dft_con[t].r, p->grad_coeff_c * dft_con[t].i};
      dft_dummy[t].r = -(s * f.i) - (-(kc.i * q));
      dft_dummy[t].i = (s * f.r) - (kc.r * q);
    }
    ora_fft_c2r(nx, ny, nz, (double*)dft_dummy, dummy);
    /* loop 12 */
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < P; ++t) dummy[t] = dummy[t] * (mob_c[t] - D);
    ora_fft_r2c(nx, ny, nz, dummy, (double*)dft_dummy, 1);
    /* loop 13: CH update */
#pragma omp parallel for schedule(static)
    for (size_t t = 0; t < H; ++t) {
      const kpt k = kt[t];
      cx hh = dft_dummy[t];
      cx term11 = {-((hh.i * p->t_step) * k.s1), (hh.r * p->t_step) * k.s1};
      double term22 = 1.0 + ((D * p->t_step) * p->grad_coeff_c) * k.k4;
      double r33 = ((D * k.k2) * p->t_step);
      cx term33 = {r33 * dft_dfdc[t].r, r33 * dft_dfdc[t].i};
      cx num = {(dft_con[t].r + term11.r) + term33.r, (dft_con[t].i + term11.i) + term33.i};
      dft_con[t].r = num.r / term22;
      dft_con[t].i = num.i / term22;
    }
    ora_fft_c2r(nx, ny, nz, (double*)dft_con, con);
    /* loop 14: AC update per order parameter (uses phi from before this loop) */
    for (int iv = 0; iv < var; ++iv) {
#pragma omp parallel for schedule(static)
      for (size_t t = 0; t < P; ++t) dummy[t] = phi[t * var + iv];
      ora_fft_r2c(nx, ny, nz, dummy, (double*)spec, 1);
#pragma omp parallel for schedule(static)
      for (size_t t = 0; t < P; ++t) dummy[t] = df_dphi[t * var + iv];
      ora_fft_r2c(nx, ny, nz, dummy, (double*)spec2, 1);
      const double ldt = p->mob_phi * p->t_step;
      const double lkd = (p->mob_phi * p->grad_coeff_phi) * p->t_step;
#pragma omp parallel for schedule(static)
      for (size_t t = 0; t < H; ++t) {
        double term22 = 1.0 - lkd * kt[t].k2;
        cx term11 = {ldt * spec2[t].r, ldt * spec2[t].i};
        spec[t].r = (spec[t].r - term11.r) / term22;
        spec[t].i = (spec[t].i - term11.i) / term22;
      }
      ora_fft_c2r(nx, ny, nz, (double*)spec, dummy);
#pragma omp parallel for schedule(static)
      for (size_t t = 0; t < P; ++t) phi[t * var + iv] = dummy[t];
    }
    ora_extrema((int)P, var, phi, con, ext);
  }

  free(g); free(kt); free(kap); free(dummy); free(gradr); free(df_dc); free(mob_c);
  free(df_dphi); free(pf_ch); free(pf_eta); free(plane); free(spec); free(sgrad);
  free(dft_con); free(dft_dfdc); free(dft_dummy); free(spec2);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* Synthetic initial conditions (the reference has none of these; they are    */
/* defined once for the benchmark configurations and restated here).          */
/* ------------------------------------------------------------------------- */
int ora_ic_random(int nx, int ny, int nz, int var, int iseed, double c0,
                  double c_amp, double eta_amp, double* phi, double* con) {
  ora_ran2_state st;
  ora_ran2_init(&st);
  int idum = iseed;
  size_t t = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++t) {
        double u = ora_ran2(&st, &idum);
        con[t] = c0 + c_amp * (2.0 * u - 1.0);
        for (int v = 0; v < var; ++v) {
          double w = ora_ran2(&st, &idum);
          phi[t * var + v] = eta_amp * (2.0 * w - 1.0);
        }
      }
  return 0;
}

int ora_ic_voronoi(int nx, int ny, int nz, int var, int nseeds, int iseed,
                   double c0, double c_amp, double* phi, double* con) {
  ora_ran2_state st;
  ora_ran2_init(&st);
  int idum = iseed;
  double* sd = (double*)malloc(sizeof(double) * 3 * (size_t)nseeds);
  for (int s = 0; s < nseeds; ++s) {
    sd[3 * s] = ora_ran2(&st, &idum) * nx;
    sd[3 * s + 1] = ora_ran2(&st, &idum) * ny;
    sd[3 * s + 2] = ora_ran2(&st, &idum) * nz;
  }
  const int n[3] = {nx, ny, nz};
#pragma omp parallel for schedule(static)
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const double x[3] = {(double)i, (double)j, (double)k};
        int best = 0;
        double bd = 0.0;
        for (int s = 0; s < nseeds; ++s) {
          double d2 = 0.0;
          for (int d = 0; d < 3; ++d) {
            double dd = fabs(x[d] - sd[3 * s + d]);
            double alt = (double)n[d] - dd;
            if (alt < dd) dd = alt;
            d2 = d2 + dd * dd;
          }
          if (s == 0 || d2 < bd) { bd = d2; best = s; }
        }
        size_t t = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
        for (int v = 0; v < var; ++v) phi[t * var + v] = 0.0;
        phi[t * var + (best % var)] = 1.0;
      }
  size_t t = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i, ++t) {
        double u = ora_ran2(&st, &idum);
        con[t] = c0 + c_amp * (2.0 * u - 1.0);
      }
  free(sd);
  return 0;
}

void ora_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
