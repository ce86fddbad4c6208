// Axis-pass kernels of the 3-D real<->half-complex transform.
//
// Layout of one field ("split-Nyquist half spectrum"):
//   main block  S[kz][ky][kx], kx in [0, M), M = NX/2, kx fastest   (M*NY*NZ complex)
//   Nyquist     Q[kz][ky]      for kx = M, ky fastest               (NY*NZ complex)
// The main block occupies exactly the bytes of the real field (NX*NY*NZ reals),
// so R2C and C2R run in place; the Nyquist plane lives in its own array.  With
// the odd kx = M column moved out, every y- and z-pass access is a run of whole
// cache lines.
//
// Batches of fields are contiguous: field f's main block starts at f*M*NY*NZ and
// its Nyquist plane at f*NY*NZ, so "lines" of the x pass are numbered
// L = y + NY*(z + NZ*f) across the batch.
#pragma once
#include "pf_fft.hpp"

namespace pf {

constexpr int kBlock = 256;

__host__ __device__ constexpr int lines_per_block(int N) {
  return (kBlock / fft_threads(N)) > 0 ? (kBlock / fft_threads(N)) : 1;
}
// columns per tile of a column pass: >= 16 for 128-byte runs, LDS tile <= 64 KiB
__host__ __device__ constexpr int col_tile(int N, int cbytes) {
  int tx = kBlock / fft_threads(N);
  if (tx < 16) tx = 16;
  while (tx > 1 && N * tx * cbytes > 65536) tx /= 2;
  return tx;
}

// ---------------------------------------------------------------------------
// x pass, forward: real lines (NX) -> half spectrum lines (M main + 1 Nyquist).
// data: in/out (real lines == main block); nyq: Nyquist entries, one per line.
// tw:   table of exp(-2 pi i q/NX), q in [0, NX).
// ---------------------------------------------------------------------------
template <typename R, int NX>
__global__ void __launch_bounds__(lines_per_block(NX / 2) * fft_threads(NX / 2))
k_x_r2c(c2<R>* __restrict__ data, c2<R>* __restrict__ nyq, long nlines,
        const c2<R>* __restrict__ tw, R scale) {
  constexpr int M = NX / 2;
  constexpr int T = fft_threads(M);
  constexpr int LPB = lines_per_block(M);
  constexpr int NT = LPB * T;
  __shared__ c2<R> sh[LPB * M];
  const long l0 = (long)blockIdx.x * LPB;
  const int nl = (int)((nlines - l0) < LPB ? (nlines - l0) : LPB);
  c2<R>* g = data + l0 * M;
  for (int i = threadIdx.x; i < nl * M; i += NT) sh[i] = g[i];
  __syncthreads();
  const int l = threadIdx.x / T, t = threadIdx.x % T;
  lds_fft<R, M, false, 1, 2>(sh + l * M, t, l < nl, tw);
  const R hs = R(0.5) * scale;
  for (int i = threadIdx.x; i < nl * M; i += NT) {
    const int ll = i / M, k = i - ll * M;
    const c2<R>* z = sh + ll * M;
    g[i] = r2c_post<R>(z[k], z[(M - k) % M], tw[k], hs);
  }
  if (threadIdx.x < nl) {
    const c2<R> z0 = sh[threadIdx.x * M];
    nyq[l0 + threadIdx.x] = mk<R>((z0.x - z0.y) * scale, R(0));
  }
}

// ---------------------------------------------------------------------------
// x pass, inverse: half spectrum lines -> real lines (unnormalised).
// ---------------------------------------------------------------------------
template <typename R, int NX>
__global__ void __launch_bounds__(lines_per_block(NX / 2) * fft_threads(NX / 2))
k_x_c2r(c2<R>* __restrict__ data, const c2<R>* __restrict__ nyq, long nlines,
        const c2<R>* __restrict__ tw) {
  constexpr int M = NX / 2;
  constexpr int T = fft_threads(M);
  constexpr int LPB = lines_per_block(M);
  constexpr int NT = LPB * T;
  __shared__ c2<R> sh[LPB * M];
  __shared__ R shn[LPB];
  const long l0 = (long)blockIdx.x * LPB;
  const int nl = (int)((nlines - l0) < LPB ? (nlines - l0) : LPB);
  c2<R>* g = data + l0 * M;
  for (int i = threadIdx.x; i < nl * M; i += NT) sh[i] = g[i];
  if (threadIdx.x < nl) shn[threadIdx.x] = nyq[l0 + threadIdx.x].x;
  __syncthreads();
  // pre-processing: each thread owns a pair (k, M-k)
  constexpr int NP = M / 2 + 1;
  for (int i = threadIdx.x; i < nl * NP; i += NT) {
    const int ll = i / NP, p = i - ll * NP;
    c2<R>* x = sh + ll * M;
    if (p == 0) {
      const R a0 = x[0].x, am = shn[ll];
      x[0] = mk<R>(a0 + am, a0 - am);
    } else {
      const int k1 = p, k2 = M - p;
      const c2<R> x1 = x[k1], x2 = x[k2];
      x[k1] = c2r_pre<R>(x1, x2, tw[k1]);
      if (k2 != k1) x[k2] = c2r_pre<R>(x2, x1, tw[k2]);
    }
  }
  __syncthreads();
  const int l = threadIdx.x / T, t = threadIdx.x % T;
  lds_fft<R, M, true, 1, 2>(sh + l * M, t, l < nl, tw);
  for (int i = threadIdx.x; i < nl * M; i += NT) g[i] = sh[i];
}

// ---------------------------------------------------------------------------
// Column pass: FFT along the rows of row-major panels (N rows x W columns),
// batch of panels at stride N*W.  Used for y (W = M, one panel per (z, field))
// and z (W = M*NY, one panel per field; W = NY for the Nyquist planes).
// ---------------------------------------------------------------------------
template <typename R, int N, bool INV>
__global__ void __launch_bounds__(col_tile(N, sizeof(c2<R>)) * fft_threads(N))
k_col_fft(c2<R>* __restrict__ data, long W, long tiles_per_panel,
          const c2<R>* __restrict__ tw) {
  constexpr int T = fft_threads(N);
  constexpr int TX = col_tile(N, sizeof(c2<R>));
  constexpr int NT = TX * T;
  __shared__ c2<R> sh[N * TX];
  const long b = blockIdx.x / tiles_per_panel;
  const long col0 = (blockIdx.x - b * tiles_per_panel) * TX;
  const int ncol = (int)((W - col0) < TX ? (W - col0) : TX);
  c2<R>* g = data + b * N * W + col0;
  for (int i = threadIdx.x; i < N * TX; i += NT) {
    const int e = i / TX, c = i % TX;
    if (c < ncol) sh[i] = g[e * W + c];
  }
  __syncthreads();
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  lds_fft<R, N, INV, TX, 1>(sh + c, t, c < ncol, tw);
  for (int i = threadIdx.x; i < N * TX; i += NT) {
    const int e = i / TX, cc = i % TX;
    if (cc < ncol) g[e * W + cc] = sh[i];
  }
}

// ---------------------------------------------------------------------------
// Row pass: FFT of contiguous complex lines of length N (Nyquist-plane y pass).
// ---------------------------------------------------------------------------
template <typename R, int N, bool INV>
__global__ void __launch_bounds__(lines_per_block(N) * fft_threads(N))
k_row_fft(c2<R>* __restrict__ data, long nlines, const c2<R>* __restrict__ tw) {
  constexpr int T = fft_threads(N);
  constexpr int LPB = lines_per_block(N);
  constexpr int NT = LPB * T;
  __shared__ c2<R> sh[LPB * N];
  const long l0 = (long)blockIdx.x * LPB;
  const int nl = (int)((nlines - l0) < LPB ? (nlines - l0) : LPB);
  c2<R>* g = data + l0 * N;
  for (int i = threadIdx.x; i < nl * N; i += NT) sh[i] = g[i];
  __syncthreads();
  const int l = threadIdx.x / T, t = threadIdx.x % T;
  lds_fft<R, N, INV, 1, 1>(sh + l * N, t, l < nl, tw);
  for (int i = threadIdx.x; i < nl * N; i += NT) g[i] = sh[i];
}

}  // namespace pf
