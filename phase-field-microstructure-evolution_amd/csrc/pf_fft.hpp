// Device-side FFT building blocks for gfx950.
//
// Lines of length N are transformed in LDS with a mixed-radix Stockham
// autosort (radices 8, 4, 2, 3, 5, largest first).  Each pass loads r inputs
// per butterfly at stride N/r, applies the twiddles w^(m*k) (k = position in
// the current sub-transform of length Ls), does a radix-r DFT in registers and
// writes the outputs to (j-k)*r + k + m*Ls.  Every pass is unrolled at compile
// time, so index arithmetic by Ls is constant folding.
//
// Twiddles come from a table of exp(-2 pi i t / (N*TWM)), built on the host from
// long-double cos/sin with octant reduction (exact at multiples of pi/4); the
// inverse uses conjugates.  Forward = e^{-i}, inverse = e^{+i}, both unscaled.
#pragma once
#include <hip/hip_runtime.h>

namespace pf {

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };
template <typename R> using c2 = typename V2<R>::T;

template <typename R> __device__ __forceinline__ c2<R> mk(R a, R b) { c2<R> r; r.x = a; r.y = b; return r; }
template <typename R> __device__ __forceinline__ c2<R> cadd(c2<R> a, c2<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R> __device__ __forceinline__ c2<R> csub(c2<R> a, c2<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <typename R> __device__ __forceinline__ c2<R> cmul(c2<R> a, c2<R> b) {
  return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
template <typename R> __device__ __forceinline__ c2<R> cmulc(c2<R> a, c2<R> b) {
  return mk<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
template <typename R> __device__ __forceinline__ c2<R> cscale(c2<R> a, R s) { return mk<R>(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <typename R, bool INV> __device__ __forceinline__ c2<R> mul_mi(c2<R> a) {
  return INV ? mk<R>(-a.y, a.x) : mk<R>(a.y, -a.x);
}

// ---------------------------------------------------------------------------
// Radix plans.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int fft_radix(int N, int p) {
  int n = N, idx = 0;
  const int order[5] = {8, 4, 2, 3, 5};
  for (int o = 0; o < 5; ++o) {
    const int r = order[o];
    while (n % r == 0) {
      if (idx == p) return r;
      n /= r;
      ++idx;
    }
  }
  return 1;
}
__host__ __device__ constexpr bool fft_size_ok(int N) {
  int n = N;
  const int order[5] = {8, 4, 2, 3, 5};
  for (int o = 0; o < 5; ++o)
    while (n > 1 && n % order[o] == 0) n /= order[o];
  return n == 1;
}
// threads cooperating on one line = butterflies of the first pass
__host__ __device__ constexpr int fft_threads(int N) { return N / fft_radix(N, 0); }

// ---------------------------------------------------------------------------
// Small DFTs in registers.
// ---------------------------------------------------------------------------
template <typename R, bool INV> __device__ __forceinline__ void dft4(c2<R>& a0, c2<R>& a1, c2<R>& a2, c2<R>& a3) {
  c2<R> t0 = cadd<R>(a0, a2), t1 = csub<R>(a0, a2);
  c2<R> t2 = cadd<R>(a1, a3), t3 = mul_mi<R, INV>(csub<R>(a1, a3));
  a0 = cadd<R>(t0, t2); a2 = csub<R>(t0, t2);
  a1 = cadd<R>(t1, t3); a3 = csub<R>(t1, t3);
}

template <typename R, int r, bool INV> __device__ __forceinline__ void dft(c2<R>* v) {
  if constexpr (r == 2) {
    c2<R> t = v[0];
    v[0] = cadd<R>(t, v[1]);
    v[1] = csub<R>(t, v[1]);
  } else if constexpr (r == 4) {
    dft4<R, INV>(v[0], v[1], v[2], v[3]);
  } else if constexpr (r == 8) {
    const R h = R(0.70710678118654752440);
    c2<R> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    c2<R> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<R, INV>(e0, e1, e2, e3);
    dft4<R, INV>(o0, o1, o2, o3);
    // o_k *= w8^k, w8 = e^{-+ 2 pi i/8}
    c2<R> w1, w3;
    if (!INV) {
      w1 = mk<R>(h * (o1.x + o1.y), h * (o1.y - o1.x));
      w3 = mk<R>(h * (o3.y - o3.x), -h * (o3.x + o3.y));
    } else {
      w1 = mk<R>(h * (o1.x - o1.y), h * (o1.x + o1.y));
      w3 = mk<R>(-h * (o3.x + o3.y), h * (o3.x - o3.y));
    }
    c2<R> w2 = mul_mi<R, INV>(o2);
    v[0] = cadd<R>(e0, o0); v[4] = csub<R>(e0, o0);
    v[1] = cadd<R>(e1, w1); v[5] = csub<R>(e1, w1);
    v[2] = cadd<R>(e2, w2); v[6] = csub<R>(e2, w2);
    v[3] = cadd<R>(e3, w3); v[7] = csub<R>(e3, w3);
  } else if constexpr (r == 3) {
    const R s3 = R(0.86602540378443864676);
    c2<R> t1 = cadd<R>(v[1], v[2]);
    c2<R> t2 = mk<R>(v[0].x - R(0.5) * t1.x, v[0].y - R(0.5) * t1.y);
    c2<R> t3 = cscale<R>(mul_mi<R, INV>(csub<R>(v[1], v[2])), s3);
    v[0] = cadd<R>(v[0], t1);
    v[1] = cadd<R>(t2, t3);
    v[2] = csub<R>(t2, t3);
  } else if constexpr (r == 5) {
    const R c1 = R(0.30901699437494742410), c2v = R(-0.80901699437494742410);
    const R s1 = R(0.95105651629515357212), s2 = R(0.58778525229247312917);
    c2<R> a1 = cadd<R>(v[1], v[4]), b1 = csub<R>(v[1], v[4]);
    c2<R> a2 = cadd<R>(v[2], v[3]), b2 = csub<R>(v[2], v[3]);
    c2<R> p1 = mk<R>(v[0].x + c1 * a1.x + c2v * a2.x, v[0].y + c1 * a1.y + c2v * a2.y);
    c2<R> p2 = mk<R>(v[0].x + c2v * a1.x + c1 * a2.x, v[0].y + c2v * a1.y + c1 * a2.y);
    c2<R> q1 = mk<R>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y);
    c2<R> q2 = mk<R>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y);
    c2<R> m1 = mul_mi<R, INV>(q1), m2 = mul_mi<R, INV>(q2);
    v[0] = mk<R>(v[0].x + a1.x + a2.x, v[0].y + a1.y + a2.y);
    v[1] = cadd<R>(p1, m1); v[4] = csub<R>(p1, m1);
    v[2] = cadd<R>(p2, m2); v[3] = csub<R>(p2, m2);
  }
}

// ---------------------------------------------------------------------------
// One Stockham pass and the recursion over all passes.
//   line : LDS base of this line; element e lives at line[e*ES]
//   t    : thread index within the line's thread group, [0, fft_threads(N))
//   act  : false for threads with no line (they still take part in barriers)
//   tw   : table of exp(-2 pi i q/(N*TWM)), q in [0, N*TWM)
// Caller must __syncthreads() after filling the line; the function ends with a
// barrier.
// ---------------------------------------------------------------------------
template <typename R, int N, bool INV, int ES, int TWM, int P, int LS>
struct FftPass {
  static __device__ __forceinline__ void run(c2<R>* line, int t, bool act, const c2<R>* __restrict__ tw) {
    constexpr int r = fft_radix(N, P);
    if constexpr (r > 1) {
      constexpr int T = fft_threads(N);
      constexpr int NB = N / r;
      constexpr int Q = (NB + T - 1) / T;
      c2<R> v[Q][r];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int j = t + q * T;
        if (act && ((NB % T) == 0 || j < NB)) {
#pragma unroll
          for (int m = 0; m < r; ++m) v[q][m] = line[(j + m * NB) * ES];
          if constexpr (LS > 1) {
            const int k = j % LS;
#pragma unroll
            for (int m = 1; m < r; ++m) {
              const c2<R> w = tw[(m * k * (N / (LS * r))) * TWM];
              v[q][m] = INV ? cmulc<R>(v[q][m], w) : cmul<R>(v[q][m], w);
            }
          }
          dft<R, r, INV>(v[q]);
        }
      }
      __syncthreads();
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int j = t + q * T;
        if (act && ((NB % T) == 0 || j < NB)) {
          const int k = (LS > 1) ? (j % LS) : 0;
          const int ob = (j - k) * r + k;
#pragma unroll
          for (int m = 0; m < r; ++m) line[(ob + m * LS) * ES] = v[q][m];
        }
      }
      __syncthreads();
      FftPass<R, N, INV, ES, TWM, P + 1, LS * r>::run(line, t, act, tw);
    }
  }
};

template <typename R, int N, bool INV, int ES, int TWM = 1>
__device__ __forceinline__ void lds_fft(c2<R>* line, int t, bool act, const c2<R>* __restrict__ tw) {
  FftPass<R, N, INV, ES, TWM, 0, 1>::run(line, t, act, tw);
}

// ---------------------------------------------------------------------------
// Real <-> half-complex x-line transforms built on a complex FFT of length M=NX/2.
// z[n] = x[2n] + i x[2n+1];  Z = FFT_M(z);  W = exp(-2 pi i/NX).
//   R2C:  X_k = 1/2 (Z_k + conj Z_{M-k}) - i/2 W^k (Z_k - conj Z_{M-k}),  k = 0..M
//   C2R:  Z_k = (X_k + conj X_{M-k}) + i conj(W^k) (X_k - conj X_{M-k}), k = 0..M-1,
//         with the imaginary parts of X_0 and X_M dropped (Hermitian projection).
// ---------------------------------------------------------------------------
template <typename R>
__device__ __forceinline__ c2<R> r2c_post(c2<R> zk, c2<R> zmk, c2<R> wk, R half_scale) {
  const R ar = zk.x + zmk.x, ai = zk.y - zmk.y;
  const R br = zk.x - zmk.x, bi = zk.y + zmk.y;
  const R cr = wk.x * br - wk.y * bi, ci = wk.x * bi + wk.y * br;
  return mk<R>(half_scale * (ar + ci), half_scale * (ai - cr));
}
template <typename R>
__device__ __forceinline__ c2<R> c2r_pre(c2<R> xk, c2<R> xmk, c2<R> wk) {
  const R ar = xk.x + xmk.x, ai = xk.y - xmk.y;
  const R br = xk.x - xmk.x, bi = xk.y + xmk.y;
  const R cr = wk.x * br + wk.y * bi, ci = wk.x * bi - wk.y * br;
  return mk<R>(ar - ci, ai + cr);
}

}  // namespace pf
