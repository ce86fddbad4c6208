// Register-resident mixed-radix Stockham FFT for gfx950.
//
// A line of length N is split over T = N / r0 threads.  Pass p (radix r_p,
// NB_p = N / r_p butterflies, LS_p = r_0 ... r_{p-1}) gives thread t the
// butterflies j = t + q*T (q < Q_p = ceil(NB_p / T)); butterfly j reads
// elements j + m*NB_p and writes (j - j%LS_p)*r_p + j%LS_p + m*LS_p.
// Pass 0 takes its inputs from registers the caller filled (typically straight
// from HBM) and the last pass leaves its outputs in registers with known element
// indices, so an N = r0*r1 transform makes a single LDS transpose.  Radices are
// 16, 12, 10, 8, 6, 5, 4, 3, 2 (largest first; powers of 8 use 8s; a few
// lengths have fixed sequences, plan_fixed), so every size this library
// supports needs at most three passes.
//
// The small DFTs are __host__ __device__ so that they, and a sequential
// emulation of the thread schedule, can be checked on the CPU.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

#include "pf_fft.hpp"

namespace pf {

// Host: twiddles exp(-2 pi i t/n), t < n, evaluated in long double on the first
// octant and mapped by exact quadrant symmetries, then rounded once to double.
inline void host_twiddles(int n, double* re, double* im) {
  const long double two_pi = 6.283185307179586476925286766559L;
  for (long t = 0; t < n; ++t) {
    const long q = (4 * t) / n;           // quadrant
    const long num = 4 * t - q * n;       // angle 2 pi num / (4n) in [0, pi/2)
    long double c, s;
    if (2 * num <= n) {
      const long double a = two_pi * (long double)num / (long double)(4L * n);
      c = cosl(a); s = sinl(a);
    } else {
      const long double a = two_pi * (long double)(n - num) / (long double)(4L * n);
      c = sinl(a); s = cosl(a);
    }
    const double cr = (double)c, ci = -(double)s;  // e^{-i a}
    double wr, wi;
    switch (q) {  // times (-i)^q
      case 0: wr = cr; wi = ci; break;
      case 1: wr = ci; wi = -cr; break;
      case 2: wr = -cr; wi = -ci; break;
      default: wr = -ci; wi = cr; break;
    }
    re[t] = wr; im[t] = wi;
  }
}

// ---------------------------------------------------------------------------
// Plans.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr bool is_pow8(int n) {
  while (n > 1 && n % 8 == 0) n /= 8;
  return n == 1;
}
// Fixed radix sequences for lengths whose greedy plan leaves threads idle in a
// later pass (NB_p not a multiple of T) and needs more registers: with these
// every pass of 48, 96 and 192 runs all T threads, 80 needs 10 slots per
// thread instead of 20 and 160 needs 12 instead of 20.  Measured on the FP32
// step: 192^3 1.83 -> 1.34 ms, 160^3 1.07 -> 0.87 ms.  0 = no fixed plan.
__host__ __device__ constexpr int plan_fixed(int N, int p) {
  switch (N) {
    case 48: { const int r[2] = {12, 4}; return p < 2 ? r[p] : 1; }
    case 80: { const int r[2] = {8, 10}; return p < 2 ? r[p] : 1; }
    case 96: { const int r[3] = {12, 4, 2}; return p < 3 ? r[p] : 1; }
    case 160: { const int r[3] = {10, 4, 4}; return p < 3 ? r[p] : 1; }
    case 192: { const int r[3] = {12, 4, 4}; return p < 3 ? r[p] : 1; }
    default: return 0;
  }
}
// Alternative radix sequences (plan variants V = 1, 2) for the lengths of the
// z-line transform, candidates of the tuner: V = 1 uses fewer threads per line
// (larger first radix), V = 2 more threads with fewer registers each.  0 = no
// alternative for that length (the variant is then the default plan).
__host__ __device__ constexpr int plan_alt(int N, int V, int p) {
  if (V == 1) {
    switch (N) {
      case 64: { const int r[2] = {16, 4}; return p < 2 ? r[p] : 1; }
      case 128: { const int r[2] = {8, 16}; return p < 2 ? r[p] : 1; }
      case 160: { const int r[2] = {16, 10}; return p < 2 ? r[p] : 1; }
      case 192: { const int r[2] = {16, 12}; return p < 2 ? r[p] : 1; }
      case 256: { const int r[3] = {16, 4, 4}; return p < 3 ? r[p] : 1; }
      case 512: { const int r[3] = {16, 8, 4}; return p < 3 ? r[p] : 1; }
      default: return 0;
    }
  }
  // V = 3: plans of up to 64 threads for one line done by a whole wave (the x
  // line's job that is alone in the block's last wave)
  if (V == 3) {
    switch (N) {
      case 64: { const int r[6] = {2, 2, 2, 2, 2, 2}; return p < 6 ? r[p] : 1; }
      case 128: { const int r[7] = {2, 2, 2, 2, 2, 2, 2}; return p < 7 ? r[p] : 1; }
      case 256: { const int r[4] = {4, 4, 4, 4}; return p < 4 ? r[p] : 1; }
      default: return 0;
    }
  }
  if (V == 2) {
    switch (N) {
      case 64: { const int r[2] = {4, 16}; return p < 2 ? r[p] : 1; }
      case 128: { const int r[3] = {8, 4, 4}; return p < 3 ? r[p] : 1; }
      case 160: { const int r[2] = {10, 16}; return p < 2 ? r[p] : 1; }
      case 192: { const int r[2] = {12, 16}; return p < 2 ? r[p] : 1; }
      case 256: { const int r[3] = {8, 8, 4}; return p < 3 ? r[p] : 1; }
      case 512: { const int r[3] = {16, 16, 2}; return p < 3 ? r[p] : 1; }
      default: return 0;
    }
  }
  return 0;
}
__host__ __device__ constexpr int plan_rad(int N, int p, int V = 0) {
  if (V > 0 && plan_alt(N, V, 0)) return plan_alt(N, V, p);
  if (plan_fixed(N, 0)) return plan_fixed(N, p);
  const int order[9] = {16, 12, 10, 8, 6, 5, 4, 3, 2};
  int n = N, idx = 0;
  if (is_pow8(N)) {
    while (n > 1) {
      if (idx == p) return 8;
      n /= 8;
      ++idx;
    }
    return 1;
  }
  while (n > 1) {
    int r = 1;
    for (int o = 0; o < 9; ++o)
      if (n % order[o] == 0) { r = order[o]; break; }
    if (r == 1) return 1;
    if (idx == p) return r;
    n /= r;
    ++idx;
  }
  return 1;
}
__host__ __device__ constexpr int plan_np(int N, int V = 0) {
  int p = 0;
  while (plan_rad(N, p, V) > 1) ++p;
  return p;
}
__host__ __device__ constexpr int plan_ls(int N, int p, int V = 0) {
  int ls = 1;
  for (int q = 0; q < p; ++q) ls *= plan_rad(N, q, V);
  return ls;
}
__host__ __device__ constexpr int plan_T(int N, int V = 0) { return N / plan_rad(N, 0, V); }
__host__ __device__ constexpr int plan_Q(int N, int p, int V = 0) {
  return (N / plan_rad(N, p, V) + plan_T(N, V) - 1) / plan_T(N, V);
}
// registers per thread (complex values)
__host__ __device__ constexpr int plan_E(int N, int V = 0) {
  int e = 0;
  for (int p = 0; p < plan_np(N, V); ++p) {
    const int v = plan_Q(N, p, V) * plan_rad(N, p, V);
    if (v > e) e = v;
  }
  return e;
}
__host__ __device__ constexpr bool plan_ok(int N, int V = 0) {
  int prod = 1;
  for (int p = 0; p < plan_np(N, V); ++p) prod *= plan_rad(N, p, V);
  return prod == N && N > 1;
}

// ---------------------------------------------------------------------------
// Small DFTs (forward e^{-i}; INV uses conjugate roots), in place on v[0..r).
// ---------------------------------------------------------------------------
template <typename R> __host__ __device__ __forceinline__ c2<R> hmk(R a, R b) { c2<R> r; r.x = a; r.y = b; return r; }
template <typename R> __host__ __device__ __forceinline__ c2<R> hadd(c2<R> a, c2<R> b) { return hmk<R>(a.x + b.x, a.y + b.y); }
template <typename R> __host__ __device__ __forceinline__ c2<R> hsub(c2<R> a, c2<R> b) { return hmk<R>(a.x - b.x, a.y - b.y); }
template <typename R> __host__ __device__ __forceinline__ c2<R> hmul(c2<R> a, c2<R> b) {
  return hmk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R> __host__ __device__ __forceinline__ c2<R> hmulc(c2<R> a, c2<R> b) {
  return hmk<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// a * exp(-i*theta) (forward) or a * exp(+i*theta) (inverse), given cos, sin of theta
template <typename R, bool INV>
__host__ __device__ __forceinline__ c2<R> rot(c2<R> a, R c, R s) {
  return INV ? hmk<R>(a.x * c - a.y * s, a.y * c + a.x * s) : hmk<R>(a.x * c + a.y * s, a.y * c - a.x * s);
}
template <typename R, bool INV> __host__ __device__ __forceinline__ c2<R> mmi(c2<R> a) {
  return INV ? hmk<R>(-a.y, a.x) : hmk<R>(a.y, -a.x);
}

template <typename R, bool INV>
__host__ __device__ __forceinline__ void d2(c2<R>& a, c2<R>& b) {
  const c2<R> t = a;
  a = hadd<R>(t, b);
  b = hsub<R>(t, b);
}
template <typename R, bool INV>
__host__ __device__ __forceinline__ void d3(c2<R>& a0, c2<R>& a1, c2<R>& a2) {
  const R s3 = R(0.86602540378443864676);
  const c2<R> t1 = hadd<R>(a1, a2);
  const c2<R> t2 = hmk<R>(a0.x - R(0.5) * t1.x, a0.y - R(0.5) * t1.y);
  const c2<R> d = mmi<R, INV>(hsub<R>(a1, a2));
  const c2<R> t3 = hmk<R>(s3 * d.x, s3 * d.y);
  a0 = hadd<R>(a0, t1);
  a1 = hadd<R>(t2, t3);
  a2 = hsub<R>(t2, t3);
}
template <typename R, bool INV>
__host__ __device__ __forceinline__ void d4(c2<R>& a0, c2<R>& a1, c2<R>& a2, c2<R>& a3) {
  const c2<R> t0 = hadd<R>(a0, a2), t1 = hsub<R>(a0, a2);
  const c2<R> t2 = hadd<R>(a1, a3), t3 = mmi<R, INV>(hsub<R>(a1, a3));
  a0 = hadd<R>(t0, t2);
  a2 = hsub<R>(t0, t2);
  a1 = hadd<R>(t1, t3);
  a3 = hsub<R>(t1, t3);
}
template <typename R, bool INV>
__host__ __device__ __forceinline__ void d5(c2<R>& a0, c2<R>& a1, c2<R>& a2, c2<R>& a3, c2<R>& a4) {
  const R c1 = R(0.30901699437494742410), c2v = R(-0.80901699437494742410);
  const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
  const c2<R> b1 = hadd<R>(a1, a4), e1 = hsub<R>(a1, a4);
  const c2<R> b2 = hadd<R>(a2, a3), e2 = hsub<R>(a2, a3);
  const c2<R> p1 = hmk<R>(a0.x + c1 * b1.x + c2v * b2.x, a0.y + c1 * b1.y + c2v * b2.y);
  const c2<R> p2 = hmk<R>(a0.x + c2v * b1.x + c1 * b2.x, a0.y + c2v * b1.y + c1 * b2.y);
  const c2<R> q1 = mmi<R, INV>(hmk<R>(s1 * e1.x + s2 * e2.x, s1 * e1.y + s2 * e2.y));
  const c2<R> q2 = mmi<R, INV>(hmk<R>(s2 * e1.x - s1 * e2.x, s2 * e1.y - s1 * e2.y));
  a0 = hmk<R>(a0.x + b1.x + b2.x, a0.y + b1.y + b2.y);
  a1 = hadd<R>(p1, q1);
  a4 = hsub<R>(p1, q1);
  a2 = hadd<R>(p2, q2);
  a3 = hsub<R>(p2, q2);
}

template <typename R, int r, bool INV> __host__ __device__ __forceinline__ void dftr(c2<R>* v) {
  if constexpr (r == 2) {
    d2<R, INV>(v[0], v[1]);
  } else if constexpr (r == 3) {
    d3<R, INV>(v[0], v[1], v[2]);
  } else if constexpr (r == 4) {
    d4<R, INV>(v[0], v[1], v[2], v[3]);
  } else if constexpr (r == 5) {
    d5<R, INV>(v[0], v[1], v[2], v[3], v[4]);
  } else if constexpr (r == 6) {
    // 6 = 2 x 3: X[k1 + 2 k2] = sum_n2 W6^{n2 k1} W3^{n2 k2} sum_n1 x[3 n1 + n2] W2^{n1 k1}
    const R h = R(0.5), s3 = R(0.86602540378443864676);
    c2<R> a0 = v[0], b0 = v[3], a1 = v[1], b1 = v[4], a2 = v[2], b2 = v[5];
    d2<R, INV>(a0, b0);
    d2<R, INV>(a1, b1);
    d2<R, INV>(a2, b2);
    b1 = rot<R, INV>(b1, h, s3);
    b2 = rot<R, INV>(b2, -h, s3);
    d3<R, INV>(a0, a1, a2);
    d3<R, INV>(b0, b1, b2);
    v[0] = a0; v[2] = a1; v[4] = a2;
    v[1] = b0; v[3] = b1; v[5] = b2;
  } else if constexpr (r == 8) {
    const R h = R(0.70710678118654752440);
    c2<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    c2<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    d4<R, INV>(e0, e1, e2, e3);
    d4<R, INV>(o0, o1, o2, o3);
    o1 = rot<R, INV>(o1, h, h);
    o2 = mmi<R, INV>(o2);
    o3 = rot<R, INV>(o3, -h, h);
    v[0] = hadd<R>(e0, o0); v[4] = hsub<R>(e0, o0);
    v[1] = hadd<R>(e1, o1); v[5] = hsub<R>(e1, o1);
    v[2] = hadd<R>(e2, o2); v[6] = hsub<R>(e2, o2);
    v[3] = hadd<R>(e3, o3); v[7] = hsub<R>(e3, o3);
  } else if constexpr (r == 10) {
    // 10 = 2 x 5
    const R c36 = R(0.80901699437494742410), s36 = R(0.58778525229247312917);
    const R c72 = R(0.30901699437494742410), s72 = R(0.95105651629515357212);
    c2<R> a[5], b[5];
#pragma unroll
    for (int n = 0; n < 5; ++n) {
      a[n] = v[n];
      b[n] = v[n + 5];
      d2<R, INV>(a[n], b[n]);
    }
    b[1] = rot<R, INV>(b[1], c36, s36);
    b[2] = rot<R, INV>(b[2], c72, s72);
    b[3] = rot<R, INV>(b[3], -c72, s72);
    b[4] = rot<R, INV>(b[4], -c36, s36);
    d5<R, INV>(a[0], a[1], a[2], a[3], a[4]);
    d5<R, INV>(b[0], b[1], b[2], b[3], b[4]);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      v[2 * k] = a[k];
      v[2 * k + 1] = b[k];
    }
  } else if constexpr (r == 12) {
    // 12 = 4 x 3: X[k1 + 4 k2] = sum_n2 W12^{n2 k1} W3^{n2 k2} sum_n1 x[3 n1 + n2] W4^{n1 k1}
    const R h = R(0.5), s3 = R(0.86602540378443864676);
    c2<R> a[3][4];
#pragma unroll
    for (int n2 = 0; n2 < 3; ++n2) {
      a[n2][0] = v[n2]; a[n2][1] = v[n2 + 3]; a[n2][2] = v[n2 + 6]; a[n2][3] = v[n2 + 9];
      d4<R, INV>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
    }
    a[1][1] = rot<R, INV>(a[1][1], s3, h);
    a[1][2] = rot<R, INV>(a[1][2], h, s3);
    a[1][3] = mmi<R, INV>(a[1][3]);
    a[2][1] = rot<R, INV>(a[2][1], h, s3);
    a[2][2] = rot<R, INV>(a[2][2], -h, s3);
    a[2][3] = hmk<R>(-a[2][3].x, -a[2][3].y);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      d3<R, INV>(a[0][k1], a[1][k1], a[2][k1]);
      v[k1] = a[0][k1]; v[k1 + 4] = a[1][k1]; v[k1 + 8] = a[2][k1];
    }
  } else if constexpr (r == 16) {
    // 16 = 4 x 4
    const R c1 = R(0.92387953251128675613), s1 = R(0.38268343236508977173);
    const R h = R(0.70710678118654752440);
    c2<R> a[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
      a[n2][0] = v[n2]; a[n2][1] = v[n2 + 4]; a[n2][2] = v[n2 + 8]; a[n2][3] = v[n2 + 12];
      d4<R, INV>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
    }
    a[1][1] = rot<R, INV>(a[1][1], c1, s1);
    a[1][2] = rot<R, INV>(a[1][2], h, h);
    a[1][3] = rot<R, INV>(a[1][3], s1, c1);
    a[2][1] = rot<R, INV>(a[2][1], h, h);
    a[2][2] = mmi<R, INV>(a[2][2]);
    a[2][3] = rot<R, INV>(a[2][3], -h, h);
    a[3][1] = rot<R, INV>(a[3][1], s1, c1);
    a[3][2] = rot<R, INV>(a[3][2], -h, h);
    a[3][3] = rot<R, INV>(a[3][3], -c1, -s1);
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
      d4<R, INV>(a[0][k1], a[1][k1], a[2][k1], a[3][k1]);
      v[k1] = a[0][k1]; v[k1 + 4] = a[1][k1]; v[k1 + 8] = a[2][k1]; v[k1 + 12] = a[3][k1];
    }
  }
}

// ---------------------------------------------------------------------------
// The register-resident transform.  v[q*r_p + m] <-> butterfly j = t + q*T,
// input m of the current pass (on entry, p = 0) / output m of the last pass (on
// exit).  lds: this line's exchange buffer, element e at lds[e * ES].
// Every per-pass quantity is a compile-time constant of PassC<N, P>.
// ---------------------------------------------------------------------------
template <int N, int P, int V = 0> struct PassC {
  static constexpr int r = plan_rad(N, P, V);
  static constexpr int NB = N / r;
  static constexpr int LS = plan_ls(N, P, V);
  static constexpr int Q = plan_Q(N, P, V);
  static constexpr int T = plan_T(N, V);
  static constexpr bool FULL = (NB % T) == 0;
  static __host__ __device__ __forceinline__ int in_idx(int t, int q, int m) { return t + q * T + m * NB; }
  static __host__ __device__ __forceinline__ int out_idx(int t, int q, int m) {
    const int j = t + q * T, k = j % LS;
    return (j - k) * r + k + m * LS;
  }
  static __host__ __device__ __forceinline__ bool valid(int t, int q) { return FULL || t + q * T < NB; }
};
template <int N, int V = 0> struct RPlan {
  static constexpr int NP = plan_np(N, V);
  static constexpr int T = plan_T(N, V);
  static constexpr int E = plan_E(N, V);
  static_assert(plan_ok(N, V), "unsupported transform length");
};

// One pass's arithmetic on registers: twiddles (p > 0) then radix-r DFTs.
// TWR: the pass has one butterfly per thread and takes its twiddles from twr[m - 1]
// (registers the caller filled with last_pass_twiddles) instead of the table.
template <typename R, int N, bool INV, int P, int V = 0, bool TWR = false>
__host__ __device__ __forceinline__ void rpass_compute(c2<R>* v, int t, bool act, const c2<R>* __restrict__ tw, int twm,
                                                       const c2<R>* twr = nullptr) {
  using C = PassC<N, P, V>;
  constexpr int r = C::r, NB = C::NB, LS = C::LS, Q = C::Q, T = C::T;
  static_assert(!TWR || (Q == 1 && LS > 1), "register twiddles: one butterfly per thread, a pass after the first");
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int j = t + q * T;
    if (!act || !((NB % T) == 0 || j < NB)) continue;
    if constexpr (LS > 1) {
      const int k = j % LS;
#pragma unroll
      for (int m = 1; m < r; ++m) {
        const c2<R> w = TWR ? twr[m - 1] : tw[(m * k * (N / (LS * r))) * twm];
        v[q * r + m] = INV ? hmulc<R>(v[q * r + m], w) : hmul<R>(v[q * r + m], w);
      }
    }
    dftr<R, r, INV>(v + q * r);
  }
}

// LDS element stride: compile-time ES, or ES == 0 for a runtime stride es.
template <int ES> __device__ __forceinline__ int lds_stride(int es) { return ES ? ES : es; }

// Exchange flags XF of a transform whose LDS line is private to one wave:
//   XF_WAVE: the T threads of each line are lanes of one wave, so an exchange
//            orders the wave's own LDS accesses (wave-scope fences) instead of
//            synchronising the block;
//   XF_SWZ:  element e of the line is stored at lds_swz(e) = e + e / 16 (one pad
//            element after every 16), which spreads the exchange accesses of
//            neighbouring threads, 16 elements apart in a radix-16 pass, over the
//            LDS banks; a line of N elements spans N + N / 16.  Added rather than
//            XORed, the pad keeps each access a per-thread base plus an immediate
//            (an XOR swizzle needed an address register per slot and made the
//            x-line kernel spill).
enum { XF_WAVE = 1, XF_SWZ = 2 };
template <int XF> __host__ __device__ constexpr int lds_swz(int e) {
  return (XF & XF_SWZ) ? e + (e >> 4) : e;
}
template <int XF> __device__ __forceinline__ void xsync() {
  if constexpr (XF & XF_WAVE) {
    // the scheduling barrier keeps the passes' code apart as a block barrier
    // would: interleaved across the exchange, the x-line kernel's passes needed
    // more than 128 VGPRs
    __builtin_amdgcn_sched_barrier(0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    __builtin_amdgcn_sched_barrier(0);
  } else {
    __syncthreads();
  }
}

template <typename R, int N, int P, int ES, int V = 0, int XF = 0>
__device__ __forceinline__ void rpass_store(const c2<R>* v, int t, bool act, c2<R>* lds, int es) {
  using C = PassC<N, P, V>;
  const int st = lds_stride<ES>(es);
#pragma unroll
  for (int q = 0; q < C::Q; ++q) {
    if (!act || !C::valid(t, q)) continue;
#pragma unroll
    for (int m = 0; m < C::r; ++m) lds[lds_swz<XF>(C::out_idx(t, q, m)) * st] = v[q * C::r + m];
  }
}
template <typename R, int N, int P, int ES, int V = 0, int XF = 0>
__device__ __forceinline__ void rpass_load(c2<R>* v, int t, bool act, const c2<R>* lds, int es) {
  using C = PassC<N, P, V>;
  const int st = lds_stride<ES>(es);
#pragma unroll
  for (int q = 0; q < C::Q; ++q) {
    if (!act || !C::valid(t, q)) continue;
#pragma unroll
    for (int m = 0; m < C::r; ++m) v[q * C::r + m] = lds[lds_swz<XF>(C::in_idx(t, q, m)) * st];
  }
}

// Passes P.. of the transform; each exchange is bracketed by block barriers, or
// by wave-scope fences when the line belongs to one wave (XF_WAVE).
// TWR: the last pass's twiddles come from twr (see rpass_compute).
template <typename R, int N, bool INV, int ES, int P, int V = 0, int XF = 0, bool TWR = false>
__device__ __forceinline__ void rfft_from(c2<R>* v, int t, bool act, c2<R>* lds, const c2<R>* __restrict__ tw,
                                          int twm, int es, const c2<R>* twr = nullptr) {
  rpass_compute<R, N, INV, P, V, TWR && (P + 1 == plan_np(N, V))>(v, t, act, tw, twm, twr);
  if constexpr (P + 1 < plan_np(N, V)) {
    xsync<XF>();
    rpass_store<R, N, P, ES, V, XF>(v, t, act, lds, es);
    xsync<XF>();
    rpass_load<R, N, P + 1, ES, V, XF>(v, t, act, lds, es);
    rfft_from<R, N, INV, ES, P + 1, V, XF, TWR>(v, t, act, lds, tw, twm, es, twr);
  }
}
// ES = 0: the LDS element stride is the runtime value es.  V: plan variant.
template <typename R, int N, bool INV, int ES, int V = 0, int XF = 0, bool TWR = false>
__device__ __forceinline__ void rfft(c2<R>* v, int t, bool act, c2<R>* lds, const c2<R>* __restrict__ tw,
                                     int twm = 1, int es = 0, const c2<R>* twr = nullptr) {
  rfft_from<R, N, INV, ES, 0, V, XF, TWR>(v, t, act, lds, tw, twm, es, twr);
}

// Element index of register slot s after the transform (last pass outputs) and
// before it (pass-0 inputs); s is a compile-time constant in unrolled loops.
template <int N, int V = 0> using LastPass = PassC<N, plan_np(N, V) - 1, V>;
template <int N, int V = 0> __host__ __device__ __forceinline__ int rfft_out_idx(int t, int s) {
  return LastPass<N, V>::out_idx(t, s / LastPass<N, V>::r, s % LastPass<N, V>::r);
}
template <int N, int V = 0> __host__ __device__ __forceinline__ int rfft_in_idx(int t, int s) {
  return PassC<N, 0, V>::in_idx(t, s / PassC<N, 0, V>::r, s % PassC<N, 0, V>::r);
}
template <int N, int V = 0> __host__ __device__ __forceinline__ bool rfft_in_valid(int t, int s) {
  return PassC<N, 0, V>::valid(t, s / PassC<N, 0, V>::r);
}
// In the last pass LS = N / r = NB, so j < NB gives k = j and out_idx(t, q, m) =
// j + m*NB = in_idx(t, q, m): each thread's outputs land on the elements it read
// as that pass's inputs, and callers may write them back without a barrier.
template <int N, int V = 0> constexpr bool last_pass_in_place() { return LastPass<N, V>::LS == LastPass<N, V>::NB; }
template <int N, int V = 0> __host__ __device__ __forceinline__ bool rfft_out_valid(int t, int s) {
  return LastPass<N, V>::valid(t, s / LastPass<N, V>::r);
}
// Number of register slots used at the input / output (variable templates, so
// that loops over them have compile-time trip counts and fully unroll).
template <int N, int V = 0> constexpr int kInSlots = PassC<N, 0, V>::Q * PassC<N, 0, V>::r;
template <int N, int V = 0> constexpr int kOutSlots = LastPass<N, V>::Q * LastPass<N, V>::r;
// True when the thread's output element set equals its input element set (so a
// forward transform's outputs can feed an inverse transform without exchange).
template <int N, int V = 0> __host__ __device__ constexpr bool rfft_io_same() {
  return plan_np(N, V) == 1 || (plan_np(N, V) == 2 && plan_rad(N, 0, V) == plan_rad(N, 1, V));
}

// The last pass's twiddles of thread t (one butterfly per thread) into twr[0 .. r-1).
template <typename R, int N, int V = 0>
__device__ __forceinline__ void last_pass_twiddles(c2<R>* twr, int t, const c2<R>* __restrict__ tw, int twm) {
  using C = LastPass<N, V>;
  static_assert(C::Q == 1 && C::LS > 1, "one butterfly per thread, a pass after the first");
  const int k = t % C::LS;
#pragma unroll
  for (int m = 1; m < C::r; ++m) twr[m - 1] = tw[(m * k * (N / (C::LS * C::r))) * twm];
}

}  // namespace pf
