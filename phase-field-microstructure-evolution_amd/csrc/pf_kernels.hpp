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
//
// Transforms use the register-resident Stockham of pf_fft2.hpp: pass-0 inputs
// are loaded from HBM straight into registers and last-pass outputs are stored
// from registers; LDS is only used for the transposes between passes (and, in
// the x pass, for the real<->half-complex pairing of k and M-k).
#pragma once
#include "pf_fft2.hpp"

namespace pf {

constexpr int kBlock = 256;

// columns per tile of a column pass: a power of two from 16 to 64 (every row
// segment a whole number of 128-byte lines), block of <= ~256 threads, exchange
// tile <= 64 KiB.  (Lengths 96, 160, 192 have T = 6, 10, 12 threads per column;
// 256/T would give 42, 25, 21 columns and misaligned segments, which ran the
// 192^3 z-line at 1.8 TB/s.)
__host__ __device__ constexpr int col_tile(int N, int cbytes) {
  const int want = kBlock / plan_T(N);
  int tx = 16;
  while (tx * 2 <= want && tx < 64) tx *= 2;
  while (tx > 1 && N * tx * cbytes > 65536) tx /= 2;
  return tx;
}
// lines per block of a contiguous-line pass
__host__ __device__ constexpr int lines_per_block(int N) {
  return (kBlock / plan_T(N)) > 0 ? (kBlock / plan_T(N)) : 1;
}

// Map physical block ids to logical ones so that each XCD (blocks are dealt to
// the 8 XCDs round-robin) works on a contiguous range of tiles or lines.  The x
// line kernel's neighbouring lines then share that XCD's L2 for the Nyquist
// entries, and a column pass's neighbouring tiles (the 128-byte pieces of the
// same rows) are read by one XCD at about the same time: measured at 256^3 x 17,
// column passes 3-7% faster.  (The z-line kernel measured 1.05 against 1.34 ms
// with this remap; it uses xcd_runs below.)
__device__ __forceinline__ long xcd_block(long b, long nb) {
  if (nb % 8) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

// The same in runs of G: XCD x takes the blocks x*G .. x*G + G - 1 of every 8*G
// (unchanged when the grid is not a multiple of 8*G).  The z-line kernel uses
// runs of one main-block row (M / TX tiles), so each XCD reads whole 1 KiB rows
// of a z plane: measured at 256^3 x 17, 1.055 -> 1.003 ms per launch (runs of
// 8-32 tiles alike; 256 tiles 1.036, one contiguous range per XCD 1.34).
__device__ __forceinline__ long xcd_runs(long b, long nb, long G) {
  if (G < 1 || nb % (8 * G)) return b;
  const long x = b % 8, i = b / 8;
  return (i / G) * (8 * G) + x * G + i % G;
}

// ---------------------------------------------------------------------------
// Column pass: FFT along the rows of row-major panels (N rows x W columns),
// batch of panels at stride N*W.  Used for y (W = M, one panel per (z, field))
// and z (W = M*NY, one panel per field; W = NY for the Nyquist planes).
// ---------------------------------------------------------------------------
template <typename R, int N, bool INV>
__global__ void __launch_bounds__(col_tile(N, sizeof(c2<R>)) * plan_T(N))
k_col_fft(c2<R>* __restrict__ data, long W, long tiles_per_panel, const c2<R>* __restrict__ tw) {
  constexpr int E = plan_E(N);
  constexpr int TX = col_tile(N, sizeof(c2<R>));
  __shared__ c2<R> sh[(plan_np(N) > 1) ? N * TX : 1];
  const long bid = xcd_block(blockIdx.x, gridDim.x);
  const long b = bid / tiles_per_panel;
  const long col0 = (bid - b * tiles_per_panel) * TX;
  const int ncol = (int)((W - col0) < TX ? (W - col0) : TX);
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const bool act = c < ncol;
  c2<R>* g = data + b * N * W + col0 + c;
  c2<R> v[E];
#pragma unroll
  for (int s = 0; s < kInSlots<N>; ++s)
    if (act && rfft_in_valid<N>(t, s)) v[s] = ld_nt<R>(g + (long)rfft_in_idx<N>(t, s) * W);
  rfft<R, N, INV, TX>(v, t, act, sh + c, tw);
#pragma unroll
  for (int s = 0; s < kOutSlots<N>; ++s)
    if (act && rfft_out_valid<N>(t, s)) st_nt<R>(g + (long)rfft_out_idx<N>(t, s) * W, v[s]);
}

// ---------------------------------------------------------------------------
// Row pass: FFT of contiguous complex lines of length N (Nyquist-plane y pass).
// ---------------------------------------------------------------------------
template <typename R, int N, bool INV>
__global__ void __launch_bounds__(lines_per_block(N) * plan_T(N))
k_row_fft(c2<R>* __restrict__ data, long nlines, const c2<R>* __restrict__ tw) {
  constexpr int T = plan_T(N), E = plan_E(N);
  constexpr int LPB = lines_per_block(N);
  __shared__ c2<R> sh[(plan_np(N) > 1) ? LPB * N : 1];
  const int l = threadIdx.x / T, t = threadIdx.x % T;
  const long line = (long)blockIdx.x * LPB + l;
  const bool act = line < nlines;
  c2<R>* g = data + (act ? line : 0) * N;
  c2<R> v[E];
#pragma unroll
  for (int s = 0; s < kInSlots<N>; ++s)
    if (act && rfft_in_valid<N>(t, s)) v[s] = ld_nt<R>(g + rfft_in_idx<N>(t, s));
  rfft<R, N, INV, 1>(v, t, act, sh + ((plan_np(N) > 1) ? l * N : 0), tw);
#pragma unroll
  for (int s = 0; s < kOutSlots<N>; ++s)
    if (act && rfft_out_valid<N>(t, s)) st_nt<R>(g + rfft_out_idx<N>(t, s), v[s]);
}

// Lines per block of the x passes (k_x_r2c, k_x_c2r, k_x_c2r_sumsq).  FP64:
// blocks of one wave (e.g. 8 lines of 8 threads at NX = 256), a 16 KiB tile
// instead of 64 KiB, which held blocks to two per CU; measured for 17 fields at
// 256^3: R2C 0.90 -> 0.85 ms, C2R 1.19 -> 1.13 ms.  FP32: lines_per_block.
template <typename R> __host__ __device__ constexpr int x_lines(int M) {
  return sizeof(R) == 8 ? ((64 / plan_T(M)) > 0 ? 64 / plan_T(M) : 1) : lines_per_block(M);
}

// ---------------------------------------------------------------------------
// Real <-> half-complex x lines on a complex FFT of length M = NX/2.
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
// Z_k for C2R from the half line x (M entries) and the Nyquist real xn.
template <typename R, int M>
__device__ __forceinline__ c2<R> c2r_z(const c2<R>* x, R xn, int k, const c2<R>* __restrict__ tw) {
  if (k == 0) {
    const R a0 = x[0].x;
    return mk<R>(a0 + xn, a0 - xn);
  }
  return c2r_pre<R>(x[k], x[M - k], tw[k]);
}

// ---------------------------------------------------------------------------
// x pass, forward: real lines (NX) -> half spectrum lines (M main + 1 Nyquist).
// data: in/out (real lines == main block); nyq: Nyquist entries, one per line.
// tw:   table of exp(-2 pi i q/NX), q in [0, NX).
// ---------------------------------------------------------------------------
template <typename R, int NX>
__global__ void __launch_bounds__(x_lines<R>(NX / 2) * plan_T(NX / 2))
k_x_r2c(c2<R>* __restrict__ data, c2<R>* __restrict__ nyq, long nlines,
        const c2<R>* __restrict__ tw, R scale) {
  constexpr int M = NX / 2, T = plan_T(M), E = plan_E(M);
  constexpr int LPB = x_lines<R>(M), NT = LPB * T;
  __shared__ c2<R> sh[LPB * M];
  const long l0 = (long)blockIdx.x * LPB;
  const int nl = (int)((nlines - l0) < LPB ? (nlines - l0) : LPB);
  const int l = threadIdx.x / T, t = threadIdx.x % T;
  const bool act = l < nl;
  c2<R>* g = data + l0 * M;
  c2<R> v[E];
#pragma unroll
  for (int s = 0; s < kInSlots<M>; ++s)
    if (act && rfft_in_valid<M>(t, s)) v[s] = ld_nt<R>(g + (long)l * M + rfft_in_idx<M>(t, s));
  c2<R>* line = sh + l * M;
  rfft<R, M, false, 1>(v, t, act, line, tw, 2);
  __syncthreads();
#pragma unroll
  for (int s = 0; s < kOutSlots<M>; ++s)
    if (act && rfft_out_valid<M>(t, s)) line[rfft_out_idx<M>(t, s)] = v[s];
  __syncthreads();
  const R hs = R(0.5) * scale;
  for (int i = threadIdx.x; i < nl * M; i += NT) {
    const int ll = i / M, k = i - ll * M;
    const c2<R>* z = sh + ll * M;
    st_nt<R>(g + i, r2c_post<R>(z[k], z[(M - k) % M], tw[k], hs));
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
__global__ void __launch_bounds__(x_lines<R>(NX / 2) * plan_T(NX / 2))
k_x_c2r(c2<R>* __restrict__ data, const c2<R>* __restrict__ nyq, long nlines,
        const c2<R>* __restrict__ tw) {
  constexpr int M = NX / 2, T = plan_T(M), E = plan_E(M);
  constexpr int LPB = x_lines<R>(M);
  __shared__ c2<R> sh[(plan_np(M) > 1) ? LPB * M : 1];
  const int l = threadIdx.x / T, t = threadIdx.x % T;
  const long line = (long)blockIdx.x * LPB + l;
  const bool act = line < nlines;
  c2<R>* g = data + (act ? line : 0) * M;
  c2<R> v[E];
  // pre-processing from HBM straight into registers (X_k and X_{M-k} of the line)
#pragma unroll
  for (int s = 0; s < kInSlots<M>; ++s) {
    if (!(act && rfft_in_valid<M>(t, s))) continue;
    const int k = rfft_in_idx<M>(t, s);
    if (k == 0) {
      const R a0 = g[0].x, xn = nyq[line].x;
      v[s] = mk<R>(a0 + xn, a0 - xn);
    } else {
      v[s] = c2r_pre<R>(g[k], g[M - k], tw[k]);
    }
  }
  rfft<R, M, true, 1>(v, t, act, sh + ((plan_np(M) > 1) ? l * M : 0), tw, 2);
#pragma unroll
  for (int s = 0; s < kOutSlots<M>; ++s)
    if (act && rfft_out_valid<M>(t, s)) st_nt<R>(g + rfft_out_idx<M>(t, s), v[s]);
}

// ---------------------------------------------------------------------------
// x pass, inverse, reduced to weighted sums of squares instead of stored: for
// the gradient-energy sum of a batch of gradient fields (three per state field,
// lines_per_field lines each), w = w0 for state field 0 and w1 for the others.
// Block partials (fixed order) in part[blockIdx.x].
// ---------------------------------------------------------------------------
template <typename R, int NX>
__global__ void __launch_bounds__(x_lines<R>(NX / 2) * plan_T(NX / 2))
k_x_c2r_sumsq(const c2<R>* __restrict__ data, const c2<R>* __restrict__ nyq, long nlines, long lines_per_field,
              double w0, double w1, const c2<R>* __restrict__ tw, double* __restrict__ part) {
  constexpr int M = NX / 2, T = plan_T(M), E = plan_E(M);
  constexpr int LPB = x_lines<R>(M), NT = LPB * T;
  __shared__ c2<R> sh[(plan_np(M) > 1) ? LPB * M : 1];
  __shared__ double red[NT / 64 + 1];
  const int l = threadIdx.x / T, t = threadIdx.x % T;
  double acc = 0.0;
  const long ngroups = (nlines + LPB - 1) / LPB;
  for (long grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const long line = grp * LPB + l;
    const bool act = line < nlines;
    const c2<R>* g = data + (act ? line : 0) * M;
    c2<R> v[E];
#pragma unroll
    for (int s = 0; s < kInSlots<M>; ++s) {
      if (!(act && rfft_in_valid<M>(t, s))) continue;
      const int k = rfft_in_idx<M>(t, s);
      if (k == 0) {
        const R a0 = g[0].x, xn = nyq[line].x;
        v[s] = mk<R>(a0 + xn, a0 - xn);
      } else {
        v[s] = c2r_pre<R>(g[k], g[M - k], tw[k]);
      }
    }
    rfft<R, M, true, 1>(v, t, act, sh + ((plan_np(M) > 1) ? l * M : 0), tw, 2);
    double q = 0.0;
#pragma unroll
    for (int s = 0; s < kOutSlots<M>; ++s)
      if (act && rfft_out_valid<M>(t, s)) q += (double)v[s].x * v[s].x + (double)v[s].y * v[s].y;
    if (act) acc += ((line / lines_per_field) / 3 == 0 ? w0 : w1) * q;
    __syncthreads();  // the next group reuses the exchange tile
  }
  // block reduction in a fixed order
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int w = 0; w < (NT + 63) / 64; ++w) sum += red[w];
    part[blockIdx.x] = sum;
  }
}

}  // namespace pf
