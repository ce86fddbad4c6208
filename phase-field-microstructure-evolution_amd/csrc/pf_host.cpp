// Host-side pieces of the C-ABI: error state, ran_2, k_space table, gamma and
// the synthetic initial conditions.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pf.h"
#include "pf_internal.hpp"

namespace pf {
namespace {
thread_local std::string g_err;
}
void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
}  // namespace pf

extern "C" const char* pf_last_error(void) { return pf::g_err.c_str(); }
extern "C" int pf_version(void) { return 1; }

// ---------------------------------------------------------------------------
// ran_2: L'Ecuyer's two-generator combination with a Bays-Durham shuffle
// (src/ran_2.f90:1-40).  The published constants of that generator are kept;
// am is the single-precision reciprocal of im1 widened to double, because the
// reference's PARAMETER statement evaluates 1./im1 in default REAL.
// ---------------------------------------------------------------------------
namespace {
constexpr int kIm1 = 2147483563, kIm2 = 2147483399, kImm1 = kIm1 - 1;
constexpr int kIa1 = 40014, kIa2 = 40692, kIq1 = 53668, kIq2 = 52774;
constexpr int kIr1 = 12211, kIr2 = 3791, kNtab = 32, kNdiv = 1 + kImm1 / kNtab;

inline int schrage(int x, int a, int q, int r, int m) {
  const int k = x / q;
  int y = a * (x - k * q) - k * r;
  if (y < 0) y += m;
  return y;
}
}  // namespace

extern "C" void pf_ran2_init(pf_ran2_state* st) {
  st->idum2 = 123456789;
  st->iy = 0;
  std::memset(st->iv, 0, sizeof(st->iv));
}

extern "C" double pf_ran2(pf_ran2_state* st, int* idum) {
  const double am = (double)(1.0f / (float)kIm1);
  const double rnmx = 1.0 - 1.2e-7;
  if (*idum <= 0) {
    *idum = std::max(-*idum, 1);
    st->idum2 = *idum;
    for (int j = kNtab + 7; j >= 0; --j) {
      *idum = schrage(*idum, kIa1, kIq1, kIr1, kIm1);
      if (j < kNtab) st->iv[j] = *idum;
    }
    st->iy = st->iv[0];
  }
  *idum = schrage(*idum, kIa1, kIq1, kIr1, kIm1);
  st->idum2 = schrage(st->idum2, kIa2, kIq2, kIr2, kIm2);
  const int j = st->iy / kNdiv;
  st->iy = st->iv[j] - st->idum2;
  st->iv[j] = *idum;
  if (st->iy < 1) st->iy += kImm1;
  return std::min(am * (double)st->iy, rnmx);
}

// ---------------------------------------------------------------------------
// Wavenumber table (src/k_space.f90:13-59): a_d = float(m~) * (2 pi / float(N_d)),
// m~ = m for m <= N/2 else m - N; kf = i a (0 at the Nyquist index), kf_sq = -a^2,
// kf_4 = ((a*a)*a)*a.  Layout (3, nz, ny, nx/2+1), complex interleaved.
// ---------------------------------------------------------------------------
namespace pf {
double kaxis(int m, int n, bool zero_nyquist) {
  const double pi = std::atan(1.0) * 4.0;
  const double unit = (2.0 * pi) / (double)(float)n;
  const int ms = (m <= n / 2) ? m : m - n;
  if (zero_nyquist && 2 * m == n) return 0.0;
  return (double)(float)ms * unit;
}
}  // namespace pf

extern "C" int pf_k_space(int nx, int ny, int nz, double* kf, double* kf_sq, double* kf_4) {
  if (nx < 2 || ny < 1 || nz < 1 || !kf || !kf_sq || !kf_4) return pf::fail(PF_EINVAL, "bad arguments");
  const int mx = nx / 2 + 1;
  const int n[3] = {nx, ny, nz};
  for (int ix = 0; ix < mx; ++ix)
    for (int iy = 0; iy < ny; ++iy)
      for (int iz = 0; iz < nz; ++iz) {
        const int m[3] = {ix, iy, iz};
        const size_t o = 3 * ((size_t)iz + (size_t)nz * ((size_t)iy + (size_t)ny * ix));
        for (int d = 0; d < 3; ++d) {
          const double a = pf::kaxis(m[d], n[d], false);
          kf[2 * (o + d)] = 0.0;
          kf[2 * (o + d) + 1] = pf::kaxis(m[d], n[d], true);
          kf_sq[2 * (o + d)] = -(a * a);
          kf_sq[2 * (o + d) + 1] = 0.0;
          kf_4[2 * (o + d)] = ((a * a) * a) * a;
          kf_4[2 * (o + d) + 1] = 0.0;
        }
      }
  return 0;
}

extern "C" int pf_gamma(int var, double sigma_1, double sigma_2, double* g) {
  if (var < 1 || !g) return pf::fail(PF_EINVAL, "bad arguments");
  for (int j = 0; j < var; ++j)
    for (int i = 0; i < var; ++i) {
      double v = 0.0;
      const bool im = i < var - 1, jm = j < var - 1;
      if (i != j && im && jm) v = sigma_1;
      if ((j == var - 1 && im) || (i == var - 1 && jm)) v = sigma_2;
      g[i + (size_t)var * j] = v;
    }
  return 0;
}

// ---------------------------------------------------------------------------
// Synthetic initial conditions.
// ---------------------------------------------------------------------------
extern "C" int pf_ic_random(int nx, int ny, int nz, int var, int iseed, double c0, double c_amp,
                            double eta_amp, double* phi, double* con) {
  if (nx < 1 || ny < 1 || nz < 1 || var < 1 || !phi || !con) return pf::fail(PF_EINVAL, "bad arguments");
  pf_ran2_state st;
  pf_ran2_init(&st);
  int idum = iseed;
  const size_t P = (size_t)nx * ny * nz;
  for (size_t t = 0; t < P; ++t) {
    con[t] = c0 + c_amp * (2.0 * pf_ran2(&st, &idum) - 1.0);
    for (int v = 0; v < var; ++v) phi[t * var + v] = eta_amp * (2.0 * pf_ran2(&st, &idum) - 1.0);
  }
  return 0;
}

extern "C" int pf_ic_voronoi(int nx, int ny, int nz, int var, int nseeds, int iseed, double c0,
                             double c_amp, double* phi, double* con) {
  if (nx < 1 || ny < 1 || nz < 1 || var < 1 || nseeds < 1 || !phi || !con)
    return pf::fail(PF_EINVAL, "bad arguments");
  pf_ran2_state st;
  pf_ran2_init(&st);
  int idum = iseed;
  std::vector<double> sx(nseeds), sy(nseeds), sz(nseeds);
  for (int s = 0; s < nseeds; ++s) {
    sx[s] = pf_ran2(&st, &idum) * nx;
    sy[s] = pf_ran2(&st, &idum) * ny;
    sz[s] = pf_ran2(&st, &idum) * nz;
  }
  auto wrap = [](double d, int n) {
    d = std::fabs(d);
    const double alt = (double)n - d;
    return alt < d ? alt : d;
  };
  auto work = [&](int z0, int z1) {
    for (int k = z0; k < z1; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          int best = 0;
          double bd = 0.0;
          for (int s = 0; s < nseeds; ++s) {
            const double dx = wrap((double)i - sx[s], nx), dy = wrap((double)j - sy[s], ny),
                         dz = wrap((double)k - sz[s], nz);
            const double d2 = (dx * dx + dy * dy) + dz * dz;
            if (s == 0 || d2 < bd) { bd = d2; best = s; }
          }
          const size_t t = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
          for (int v = 0; v < var; ++v) phi[t * var + v] = 0.0;
          phi[t * var + (best % var)] = 1.0;
        }
  };
  int nt = (int)std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
  nt = std::min(nt, nz);
  std::vector<std::thread> th;
  for (int w = 0; w < nt; ++w) th.emplace_back(work, (int)((long)nz * w / nt), (int)((long)nz * (w + 1) / nt));
  for (auto& t : th) t.join();
  const size_t P = (size_t)nx * ny * nz;
  for (size_t t = 0; t < P; ++t) con[t] = c0 + c_amp * (2.0 * pf_ran2(&st, &idum) - 1.0);
  return 0;
}
