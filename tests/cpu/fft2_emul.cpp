// CPU emulation of the register-resident Stockham schedule in pf_fft2.hpp:
// every thread of a line runs each pass's arithmetic on its registers, and the
// LDS exchanges between passes are replayed sequentially.  Output: one line per
// (N, plan variant) with the max error of both directions against a
// brute-force long-double DFT.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "../../phase-field-microstructure-evolution_amd/csrc/pf_fft2.hpp"


using namespace pf;
typedef std::complex<long double> cl;

// XF: exchange flags of the transform (XF_SWZ stores element e at lds_swz(e)).
template <int N, bool INV, int P, int V, int XF = 0>
void passes(std::vector<std::vector<double2>>& regs, const double2* tw) {
  constexpr int T = plan_T(N, V);
  for (int t = 0; t < T; ++t) rpass_compute<double, N, INV, P, V>(regs[t].data(), t, true, tw, 1);
  if constexpr (P + 1 < plan_np(N, V)) {
    std::vector<double2> lds(N + N / 16);
    using C0 = PassC<N, P, V>;
    using C1 = PassC<N, P + 1, V>;
    for (int t = 0; t < T; ++t)
      for (int q = 0; q < C0::Q; ++q)
        if (C0::valid(t, q))
          for (int m = 0; m < C0::r; ++m) lds[lds_swz<XF>(C0::out_idx(t, q, m))] = regs[t][q * C0::r + m];
    for (int t = 0; t < T; ++t)
      for (int q = 0; q < C1::Q; ++q)
        if (C1::valid(t, q))
          for (int m = 0; m < C1::r; ++m) regs[t][q * C1::r + m] = lds[lds_swz<XF>(C1::in_idx(t, q, m))];
    passes<N, INV, P + 1, V, XF>(regs, tw);
  }
}

template <int N, bool INV, int V, int XF = 0> double check() {
  std::vector<double> re(N), im(N);
  host_twiddles(N, re.data(), im.data());
  std::vector<double2> tw(N);
  for (int i = 0; i < N; ++i) { tw[i].x = re[i]; tw[i].y = im[i]; }
  std::vector<double2> x(N);
  unsigned s = 12345u + N;
  for (int i = 0; i < N; ++i) {
    s = s * 1103515245u + 12345u; x[i].x = ((s >> 8) % 20001) / 10000.0 - 1.0;
    s = s * 1103515245u + 12345u; x[i].y = ((s >> 8) % 20001) / 10000.0 - 1.0;
  }
  constexpr int T = plan_T(N, V), E = plan_E(N, V);
  static_assert(plan_ok(N, V), "plan does not factor N");
  std::vector<std::vector<double2>> regs(T, std::vector<double2>(E));
  for (int t = 0; t < T; ++t)
    for (int sl = 0; sl < kInSlots<N, V>; ++sl)
      if (rfft_in_valid<N, V>(t, sl)) regs[t][sl] = x[rfft_in_idx<N, V>(t, sl)];
  passes<N, INV, 0, V, XF>(regs, tw.data());
  std::vector<double2> y(N);
  std::vector<int> seen(N, 0);
  for (int t = 0; t < T; ++t)
    for (int sl = 0; sl < kOutSlots<N, V>; ++sl)
      if (rfft_out_valid<N, V>(t, sl)) { y[rfft_out_idx<N, V>(t, sl)] = regs[t][sl]; seen[rfft_out_idx<N, V>(t, sl)]++; }
  double err = 0.0;
  const long double pi2 = 6.283185307179586476925286766559L;
  for (int k = 0; k < N; ++k) {
    if (seen[k] != 1) return 1e300;
    cl acc = 0;
    for (int n = 0; n < N; ++n) {
      const long double a = (INV ? 1 : -1) * pi2 * (long double)((long)n * k % N) / N;
      acc += cl(x[n].x, x[n].y) * cl(cosl(a), sinl(a));
    }
    err = std::fmax(err, (double)std::abs(acc - cl(y[k].x, y[k].y)));
  }
  return err;
}


// 1 when, for every thread and valid slot of the last pass, the output element
// is the element the thread loaded as that pass's input (what lets callers
// write the transform back into its exchange buffer without a barrier).
template <int N, int V> int last_pass_in_place_emul() {
  using C = LastPass<N, V>;
  for (int t = 0; t < C::T; ++t)
    for (int q = 0; q < C::Q; ++q) {
      if (!C::valid(t, q)) continue;
      for (int m = 0; m < C::r; ++m)
        if (C::out_idx(t, q, m) != C::in_idx(t, q, m)) return 0;
    }
  return last_pass_in_place<N, V>() ? 1 : 0;
}

// columns: N, plan variant, passes, forward error, inverse error, in-place
// last pass, threads per line
#define RUNV(N, V)                                                                                   \
  std::printf("%d %d %d %.3e %.3e %d %d\n", N, V, plan_np(N, V), check<N, false, V>(), check<N, true, V>(), \
              last_pass_in_place_emul<N, V>(), plan_T(N, V));
#define RUN(N) RUNV(N, 0)
// exchanges through the padded x-line rows (lds_swz), reported as plan variant
// 9; the last column is 1 when lds_swz maps [0, N) one-to-one into [0, N + N/16)
template <int N> int swz_is_permutation() {
  std::vector<int> hit(N + N / 16, 0);
  for (int e = 0; e < N; ++e) {
    const int s = lds_swz<XF_SWZ>(e);
    if (s < 0 || s >= N + N / 16 || hit[s]++) return 0;
  }
  return 1;
}
#define RUNS(N)                                                                                       \
  std::printf("%d 9 %d %.3e %.3e %d %d\n", N, plan_np(N, 0), check<N, false, 0, XF_SWZ>(),            \
              check<N, true, 0, XF_SWZ>(), last_pass_in_place_emul<N, 0>(), swz_is_permutation<N>());
int main() {
  RUN(2) RUN(3) RUN(4) RUN(5) RUN(6) RUN(8) RUN(10) RUN(12) RUN(16) RUN(20) RUN(24) RUN(32) RUN(40)
  RUN(48) RUN(64) RUN(80) RUN(96) RUN(128) RUN(160) RUN(192) RUN(256) RUN(512)
  RUNV(64, 1) RUNV(128, 1) RUNV(160, 1) RUNV(192, 1) RUNV(256, 1) RUNV(512, 1)
  RUNV(64, 2) RUNV(128, 2) RUNV(160, 2) RUNV(192, 2) RUNV(256, 2) RUNV(512, 2)
  RUNV(64, 3) RUNV(128, 3) RUNV(256, 3)
  RUNS(4) RUNS(6) RUNS(8) RUNS(12) RUNS(16) RUNS(24) RUNS(32) RUNS(48) RUNS(64) RUNS(96) RUNS(128) RUNS(256)
  return 0;
}
