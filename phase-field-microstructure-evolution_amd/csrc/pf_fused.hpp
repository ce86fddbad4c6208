// Fused kernels of the reduced timestep.
//
// Between steps the work bank W holds every field "x-half-line" state: the full
// spectrum transformed back along z and y but not x.  One step is four kernels:
//   k_xline  : per (y,z) line, all fields: C2R along x -> nonlinear terms ->
//              R2C along x (f_trans normalised), diagnostics of the old state
//   k_ypass  : forward y (main block columns + Nyquist-plane rows)
//   k_zline  : per z column tile: forward z -> semi-implicit update against the
//              carried spectrum S (written back) -> inverse z
//   k_ypass  : inverse y
// so each field's half spectrum crosses HBM 10 times per step (x: 2, y: 2+2,
// z: 4 including S), against 17 for separate transform and pointwise passes.
#pragma once
#include "pf_kernels.hpp"
#include "pf_step_kernels.hpp"

namespace pf {

struct UPar {
  double dt, kphi, kc, D, L, Dgb, am, ap, cm, cp, s1g, s2g, invP;
};

// One simulation of a batched ensemble (a device table of these drives the
// BATCH forms of the step kernels: one launch per stage covers every member of
// one grid shape): its carried spectrum S (column-tile layout, below), work
// bank W, FP64 k = 0 modes and parameters.
template <typename R> struct Member {
  c2<R>* Wm;
  c2<R>* Wn;
  c2<R>* S;
  double* dc64;
  UPar u;
};

// Deterministic block reductions for any blockDim (multiple of 64).
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}
template <bool MAX> __device__ __forceinline__ double wave_ext(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_down(v, o, 64);
    if (MAX ? (w > v) : (w < v)) v = w;
  }
  return v;
}
// red: >= 16 doubles of LDS.  Result valid in thread 0.
__device__ __forceinline__ double bsum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < nw; ++i) s += red[i];
  return s;
}
template <bool MAX> __device__ __forceinline__ double bext(double v, double* red) {
  v = wave_ext<MAX>(v);
  const int w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double s = v;
  if (threadIdx.x == 0) {
    s = red[0];
    for (int i = 1; i < nw; ++i)
      if (MAX ? (red[i] > s) : (red[i] < s)) s = red[i];
  }
  return s;
}

// Nonlinear terms for NPT = 1 point or NPT = 2 neighbouring points (x = 2k,
// 2k+1: the two halves of one complex LDS element per field), reduced O(var)
// form (gamma block structure of src/var_diff.f90:206-218).  v: field 0 = c,
// fields 1..V = eta, at a stride of fs elements (reals for one point, complex
// for two); the points' arithmetic is independent and interleaved.  Per matrix
// field the derivative
//   eta_i (a_i - 1) + 2 eta_i (s1g (sm - a_i) + s2g av) + cT4m eta_i a_i,  a_i = eta_i^2
// is evaluated as eta_i (alpha a_i + beta) with the point constants
//   alpha = 1 - 2 s1g + cT4m,  beta = 2 (s1g sm + s2g av) - 1,
// one multiply-add and two multiplies per field; the precipitate field's
// derivative likewise as eta_V (gamma av + delta).  The regrouping moves each
// derivative by a few ulp of beta (the derivative enters the update multiplied
// by dt L).  Extremes are tracked in the working precision (exact: they are
// values of the field); the two energy sums in FP64.  The matrix grains' part
// of the bulk energy needs no work per field beyond the sums the derivatives
// use: sum_i (a_i^2 / 4 - a_i / 2) = num / 4 - sm / 2 and
// sum_{i<j} a_i a_j = (sm^2 - num) / 2, with sm = sum a_i, num = sum a_i^2.
template <typename R, int NPT> struct Pts;
template <typename R> struct Pts<R, 1> {
  using T = R;
  static __device__ __forceinline__ R get(const R& v, int) { return v; }
  static __device__ __forceinline__ R make(const R* a) { return a[0]; }
};
template <typename R> struct Pts<R, 2> {
  using T = c2<R>;
  static __device__ __forceinline__ R get(const c2<R>& v, int h) { return h ? v.y : v.x; }
  static __device__ __forceinline__ c2<R> make(const R* a) { return mk<R>(a[0], a[1]); }
};

template <typename R, int NPT, bool DIAG>
__device__ __forceinline__ void nl_points(typename Pts<R, NPT>::T* v, long fs, int V, const UPar& u, R* mo,
                                          double& a_ch, double& a_eta, R& cmax, R& cmin, R& pmin, R& p3) {
  using P = Pts<R, NPT>;
  const R am = (R)u.am, ap = (R)u.ap, cm = (R)u.cm, cp = (R)u.cp, s1g = (R)u.s1g, s2g = (R)u.s2g;
  const int mat = V - 1;
  R con[NPT], sm[NPT], num[NPT], fv[NPT];
  {
    const typename P::T c0 = v[0], f0 = v[V * fs];
#pragma unroll
    for (int h = 0; h < NPT; ++h) {
      con[h] = P::get(c0, h);
      fv[h] = P::get(f0, h);
      sm[h] = num[h] = R(0);
    }
  }
#pragma unroll 2
  for (int i = 1; i <= mat; ++i) {
    const typename P::T f = v[i * fs];
#pragma unroll
    for (int h = 0; h < NPT; ++h) {
      const R fi = P::get(f, h), a = fi * fi;
      sm[h] += a;
      num[h] += a * a;
    }
  }
  R alpha[NPT], beta[NPT], gam[NPT], hm[NPT], hp[NPT], fm[NPT], fpp[NPT], av[NPT], out[NPT];
#pragma unroll
  for (int h = 0; h < NPT; ++h) {
    av[h] = fv[h] * fv[h];
    const R av2 = av[h] * av[h];
    const R d4 = num[h] + av2;
    const R inv_d4 = R(1) / d4;
    hm[h] = num[h] * inv_d4;
    hp[h] = R(1) - hm[h];
    const R dm = con[h] - cm, dp = con[h] - cp;
    fm[h] = am * (dm * dm);
    fpp[h] = ap * (dp * dp);
    out[h] = R(2) * (hm[h] * am * dm + hp[h] * ap * dp);
    const R cT4m = R(4) * (fm[h] - fpp[h]) * (av2 * inv_d4) * inv_d4;
    alpha[h] = (R(1) - R(2) * s1g) + cT4m;
    beta[h] = R(2) * (s1g * sm[h] + s2g * av[h]) - R(1);
    gam[h] = R(1) + R(4) * hm[h] * (fpp[h] - fm[h]) * inv_d4;
  }
  v[0] = P::make(out);
#pragma unroll 2
  for (int i = 1; i <= mat; ++i) {
    const typename P::T f = v[i * fs];
    R o[NPT];
#pragma unroll
    for (int h = 0; h < NPT; ++h) {
      const R fi = P::get(f, h);
      o[h] = fi * (fi * fi * alpha[h] + beta[h]);
    }
    v[i * fs] = P::make(o);
  }
#pragma unroll
  for (int h = 0; h < NPT; ++h) out[h] = fv[h] * (av[h] * gam[h] + (R(2) * s2g * sm[h] - R(1)));
  v[V * fs] = P::make(out);
#pragma unroll
  for (int h = 0; h < NPT; ++h) {
    const R ps = sm[h] + av[h];
    mo[h] = R(2) * (R(1) - ps) * (R)u.Dgb;
    if (DIAG) {
      const R pair = R(0.5) * fma(sm[h], sm[h], -num[h]);
      const R fe = (R(0.25) * num[h] - R(0.5) * sm[h]) + av[h] * (R(0.25) * av[h] - R(0.5)) + s1g * pair +
                   s2g * av[h] * sm[h];
      a_ch += (double)(hm[h] * fm[h] + hp[h] * fpp[h]);
      a_eta += (double)fe;
      if (fv[h] > p3) p3 = fv[h];
      if (con[h] <= cmin) cmin = con[h];
      if (con[h] > cmax) cmax = con[h];
      if (ps < pmin) pmin = ps;
    }
  }
}

// ---------------------------------------------------------------------------
// k_xline: LINES consecutive (y,z) lines of all F fields per iteration.
//   wm, wn : W bank (main block, Nyquist plane); in: x-half lines; out: R2C of
//            the nonlinear derivatives (df_dc, df_dphi_i), divided by P.
//   mobd   : (m - D_mean) per point when mob (input of the mobility chain).
//   XONLY  : mobility-chain variant for one field: C2R, multiply by mobd, R2C.
// One job = one (field, line) x transform, done by T = plan_T(M) threads.  The
// LDS tile is job-major: job j's M complex elements in row j, pitch MP.
//   * When T divides 64 and the plan has at most two passes (every length but
//     M = 80 and 96), a job's T threads are
//     consecutive lanes of one wave (t = tid % T), so the transforms' exchanges
//     stay inside the wave and need no block barrier (XF_WAVE).  Element k of a
//     row sits at lds_swz(k) = k + k / 16 (XF_SWZ) and the pitch MP is chosen
//     from a model of the LDS banks over every access pattern of the kernel
//     (exchanges, pairing k with M - k, rows by lanes): the strided exchange
//     accesses of the T lanes and of the 64 / T jobs of a wave then mostly fall
//     on distinct banks.
//   * Otherwise lanes run consecutive jobs (job = tid % njobs) with the odd
//     pitch MP = M + 1, and each exchange is bracketed by block barriers.
// Threads beyond the njobs*T that carry the transforms: with wave-local rows
// they sit in the last wave next to active lanes and skip the transform phases
// as a whole (one branch, no per-slot predicates, no barrier inside); otherwise
// they run the same schedule on a spare row njobs, so no slot is predicated.
// A last wave that holds a single job hands it to all its lanes (plan variant
// 3, where the length has one).
// Staging and the final store move whole rows, one per wave at a time.  For
// 256-point rows they also do the real <-> half-complex pairing, a lane taking
// elements k and M - k together (xline_partner, xline_pair_rows); shorter rows
// are paired by the transform threads as they load (C2R) and by the store's
// lanes (R2C), with the Nyquist values staged in LDS.  256-point FP32 rows
// (kXlineTwiddleRegs) skip the final store's row pass: each job's 16 threads
// pair their forward transform's outputs by DPP and store them (phase 5).
// The table W_NX^k, k < NX, is copied to LDS once per block: the pairing reads
// k < M, the x transforms (length M, stride 2) the rest.
// Dynamic LDS: (rows*MP + NX) complex (+ njobs + 1 reals for the Nyquist values
// of rows the staging does not pair) + 16 doubles, rows = njobs (+1 for the
// spare row without wave-local rows).
// ---------------------------------------------------------------------------
// Threads per x-line block: FP32 blocks may use 1024 threads, which a 512^3 grid
// with 37 fields needs (37 transforms x 16 threads); FP64 blocks stop at 512
// (640 for 512-point lines, kXlineMaxThreadsN).
template <typename R> constexpr int kXlineMaxThreads = sizeof(R) == 4 ? 1024 : 512;
// FP64 blocks of 512-point lines may use 640 threads (37 fields x 16), 3 waves
// per SIMD at most, which their registers allow.
template <typename R, int NX> constexpr int kXlineMaxThreadsN = (sizeof(R) == 8 && NX >= 512) ? 640 : kXlineMaxThreads<R>;
// Minimum waves per SIMD asked of the register allocator.  FP32: 5 (at most 102
// VGPRs; the 256-point kernel takes 92, so 5 three-wave blocks fit a CU's
// SIMDs with room for a sixth).  FP64: 3 (the 256-point kernel takes 141 VGPRs
// with no spills).
template <typename R> constexpr int kXlineWaves = sizeof(R) == 4 ? 5 : 3;
// 512-point FP32 lines keep their last transform pass's twiddles in registers
// (kXlineTwiddleRegs): at most 128 VGPRs, which 1024-thread blocks need anyway.
template <typename R, int NX> constexpr int kXlineWavesN = (sizeof(R) == 4 && NX >= 512) ? 4 : kXlineWaves<R>;
// Rows each wave keeps in flight when staging a line group.
constexpr int kXlineRows = 2;

// Wave-local rows for one- and two-pass plans whose T divides 64; the three-pass
// plan of 96 points spilled 15 VGPRs per lane that way and ran 19% slower at
// 192^3, so it keeps block barriers, as does 80 (T = 10).
__host__ __device__ constexpr bool xline_wave_local(int M) { return 64 % plan_T(M) == 0 && plan_np(M) <= 2; }
// FP64 rows of 256 points are not padded: a 16-byte access serves 16 lanes a
// cycle, which are one job's T = 16 threads in the exchanges and consecutive
// elements in the row-wise phases, so the rows need no skew; unpadded, 37 fields
// fit one block's LDS (148 KiB of rows).
__host__ __device__ constexpr bool xline_plain_rows(int M, int cbytes) { return cbytes == 16 && M >= 256; }
__host__ __device__ constexpr int xline_xf(int M, int cbytes) {
  return xline_wave_local(M) ? (XF_WAVE | (xline_plain_rows(M, cbytes) ? 0 : XF_SWZ)) : 0;
}
__host__ __device__ constexpr int xline_rows(int M, int njobs) { return njobs + (xline_wave_local(M) ? 0 : 1); }
// Row pitch in elements of cbytes bytes: the bank-model choices for FP32
// elements; FP64 rows take the smallest padded pitch (at 512 points they must
// fit 33 rows of 16-byte elements in LDS).
__host__ __device__ constexpr int xline_row_pitch(int M, int cbytes) {
  if (!xline_wave_local(M)) return M | 1;
  if (xline_plain_rows(M, cbytes)) return M;
  if (cbytes != 8) return M + M / 16 + 1;
  switch (M) {  // smallest of the bank model's best choices, >= M + M / 16
    case 24: return 26;
    case 32: return 34;
    case 48: return 52;
    case 64: return 72;
    case 96: return 104;
    case 128: return 136;
    case 256: return 272;
    default: return M + M / 16 + 1;
  }
}

// Plan variant 3 (radix 2 at 64 and 128 points, radix 4 at 256): one line done
// by a whole wave.  When njobs = 1 mod 64 / T, the block's last wave carries one
// job on T of its lanes; with this plan its 64 lanes share that job instead and
// the wave issues far fewer transform instructions (at 256^3 x 17, 17 jobs of 8
// threads: x-line kernel 0.577 -> 0.532 ms).
__host__ __device__ constexpr bool xline_has_wide(int M) {
  return xline_wave_local(M) && plan_T(M) < 64 && (M == 64 || M == 128 || M == 256);
}

// Phases 2 and 3 when the staging did not pair (xline_pair_rows false): the C2R
// pre-processing from the raw row straight into the transform's input
// registers, then the inverse transform in place, as xline_fft.
// The partner of slot s's element k = t + T s (pass 0 has one butterfly per
// thread, NB = T) is (M - k) mod M = u + T c, u = T - t, c = M / T - 1 - s; for
// k = 0 this reads element M instead, one past the row (inside the LDS tile,
// and not used).  In a swizzled row (e + e / 16) its address is a per-lane base,
// one for each residue of c mod 16 / T, plus a compile-time offset; computed
// per slot the swizzle took six or seven VALU instructions a slot.
template <typename R, int M, int V, int XF>
__device__ __forceinline__ void xline_c2r(c2<R>* v, c2<R>* col, const c2<R>* shw, const R* xn, int t) {
  constexpr int T = plan_T(M, V);
  static_assert(PassC<M, 0, V>::NB == T && PassC<M, 0, V>::Q == 1, "pass 0: element t + T s in slot s");
  constexpr int NR = T < 16 ? 16 / T : 1;
  const int u = T - t;
  int mb[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) mb[r] = u + ((u + T * r) >> 4);
#pragma unroll
  for (int s = 0; s < kInSlots<M, V>; ++s) {
    if (!rfft_in_valid<M, V>(t, s)) continue;
    const int k = rfft_in_idx<M, V>(t, s);
    const int c = M / T - 1 - s;
    const int mi = (XF & XF_SWZ) ? mb[c % NR] + T * c + (T < 16 ? c / NR : T * c / 16) : (M - k) % M;
    // k = 0 pairs with the Nyquist entry (its c2r_pre value is not used)
    const c2<R> pk = c2r_pre<R>(col[lds_swz<XF>(k)], col[mi], shw[k]);
    const R a0 = col[0].x;
    v[s] = (k == 0) ? mk<R>(a0 + *xn, a0 - *xn) : pk;
  }
  rfft<R, M, true, 1, V, XF>(v, t, true, col, shw, 2);
#pragma unroll
  for (int s = 0; s < kOutSlots<M, V>; ++s)
    if (rfft_out_valid<M, V>(t, s)) col[lds_swz<XF>(rfft_out_idx<M, V>(t, s))] = v[s];
}
// One job's x transform with plan V, in place in its LDS row (phases 3 and 5).
// No barrier before the write-back: the last Stockham pass leaves each
// thread's outputs at exactly the elements it loaded as that pass's inputs
// (out_idx = in_idx when j < NB = LS), so no thread overwrites an element
// another thread has yet to read.
template <typename R, int M, int V, int XF, bool INV, bool TWR = false>
__device__ __forceinline__ void xline_fft(c2<R>* v, c2<R>* col, const c2<R>* shw, int t,
                                          const c2<R>* twr = nullptr) {
#pragma unroll
  for (int s = 0; s < kInSlots<M, V>; ++s)
    if (rfft_in_valid<M, V>(t, s)) v[s] = col[lds_swz<XF>(rfft_in_idx<M, V>(t, s))];
  rfft<R, M, INV, 1, V, XF, TWR>(v, t, true, col, shw, 2, 0, twr);
#pragma unroll
  for (int s = 0; s < kOutSlots<M, V>; ++s)
    if (rfft_out_valid<M, V>(t, s)) col[lds_swz<XF>(rfft_out_idx<M, V>(t, s))] = v[s];
}

// The real <-> half-complex pairing couples element k of a row with element
// M - k.  A lane takes both: pair p < M / 2 is (p, M - p), and p = 0 stands for
// the two self-coupled entries, 0 (with the Nyquist value) and M / 2.  Each
// element is then read once and the twiddle of M - p follows from that of p,
// W_NX^(M-p) = -conj(W_NX^p) (bit for bit: the table is built by exact quadrant
// symmetries); W_NX^(M/2) = (-0, -1) exactly.
template <int M> __device__ __forceinline__ int xline_partner(int p) { return p ? M - p : M / 2; }
// Where the pairings happen: in the staging and the final store (each wave's
// lanes pair the row they move), or in the transform threads' loads and the
// store's lanes, from the raw rows (three LDS reads per element, the Nyquist
// values in LDS too).  256-point rows are paired by the moving waves (512^3 x
// 37: x line 15.4 -> 14.7 ms; the kernel is bound by LDS traffic there).  At
// 128 points and fewer, pairing in the staging made the x line slower (256^3 x
// 17: 0.532 -> 0.553 ms; 128^3 x 9: 0.061 -> 0.078 ms, half the lanes idle in a
// 64-point row's pairs) and pairing in the store alone moved it by less than
// the run-to-run spread.
__host__ __device__ constexpr bool xline_pair_rows(int M) { return M >= 256; }
// 256-point FP32 rows: the last transform pass (radix 16, one butterfly per
// thread) takes its 15 twiddles, the same for every line and both directions,
// from registers loaded once per block rather than from the LDS table.  Those
// blocks hold a CU's LDS alone (37 fields: 85 KB), whose traffic bounds them.
template <typename R, int M>
constexpr bool kXlineTwiddleRegs = xline_pair_rows(M) && sizeof(R) == 4 && LastPass<M>::Q == 1 && LastPass<M>::LS > 1;
template <typename R> __device__ __forceinline__ c2<R> xline_mirror_tw(c2<R> w, int p) {
  return p ? mk<R>(-w.x, w.y) : mk<R>(R(-0.0), R(-1));
}

// Lane t of a 16-lane row receives x from lane (16 - t) mod 16 of the same row:
// DPP row_mirror (lane 15 - t) then row_ror:1 (from lane t - 1), register moves
// with no LDS traffic.
__device__ __forceinline__ float dpp_partner16(float x) {
  const int m = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x140, 0xF, 0xF, false);
  const int r = __builtin_amdgcn_update_dpp(0, m, 0x121, 0xF, 0xF, false);
  return __builtin_bit_cast(float, r);
}

// BATCH: ngroups counts the line groups of every member in mt (member m owns
// groups m*gpm .. m*gpm + gpm - 1, gpm = plane / LINES), each taking its banks
// and parameters from its entry; wm, wn, u are unused; no diagnostics.
template <typename R> __device__ __forceinline__ const UPar& member_u(const UPar& u, const Member<R>* mt, long m,
                                                                    bool batch) {
  return batch ? mt[m].u : u;
}

template <typename R, int NX, int LINES, bool XONLY, bool BATCH = false>
__global__ void __launch_bounds__((kXlineMaxThreadsN<R, NX>)) __attribute__((amdgpu_waves_per_eu(kXlineWavesN<R, NX>)))
k_xline(Geo g, UPar u, c2<R>* __restrict__ wm, c2<R>* __restrict__ wn, R* __restrict__ mobd,
        const c2<R>* __restrict__ tw, double* __restrict__ part, long ngroups, int diag, int mob,
        const Member<R>* __restrict__ mt) {
  constexpr int M = NX / 2, T = plan_T(M);
  constexpr int MP = xline_row_pitch(M, sizeof(c2<R>)), XF = xline_xf(M, sizeof(c2<R>));
  constexpr bool WX = xline_wave_local(M), HAVEW = xline_has_wide(M);
  // the wide plan's variant (3) and threads; the T-thread plan stands in when
  // this length has no wide plan, and that branch is never taken then
  constexpr int WV = HAVEW ? 3 : 0, TW = plan_T(M, WV), JW = HAVEW ? 64 / T : 1;
  static_assert(last_pass_in_place<M>(), "the write-backs after the x transforms rely on it");
  extern __shared__ unsigned char smem[];
  const int F = XONLY ? 1 : g.F;
  const int njobs = F * LINES;
  c2<R>* sh = reinterpret_cast<c2<R>*>(smem);
  const int rows = xline_rows(M, njobs);
  constexpr bool PP = xline_pair_rows(M), TR = kXlineTwiddleRegs<R, M>;
  c2<R>* shw = sh + (long)rows * MP;  // W_NX^k, k < NX
  R* shn = reinterpret_cast<R*>(shw + NX);  // Nyquist values (without PP)
  double* red = reinterpret_cast<double*>(smem + (((sizeof(c2<R>) * ((size_t)rows * MP + NX) +
                                                    (PP ? 0 : sizeof(R) * (njobs + 1))) + 15) & ~(size_t)15));
  const int nt = blockDim.x, nw = nt >> 6;
  constexpr int PPL = (M / 2 + 63) / 64;  // element pairs per lane when a wave moves a row
  constexpr int KPL = (M + 63) / 64;      // row elements per lane when a wave moves a row
  for (int k = threadIdx.x; k < NX; k += nt) shw[k] = tw[k];
  __syncthreads();  // the staging of the first line group reads the table
  c2<R> twr[TR ? LastPass<M>::r - 1 : 1];
  if constexpr (TR) last_pass_twiddles<R, M>(twr, threadIdx.x % T, shw, 2);
  const R scale = (R)u.invP;
  double a_ch = 0.0, a_eta = 0.0;
  R cmax = 0, cmin = 1, pmin = 1, p3 = 0;
  c2<R> v[plan_E(M) > plan_E(M, WV) ? plan_E(M) : plan_E(M, WV)];

  const long lb = xcd_block(blockIdx.x, gridDim.x);
  const long chunk = (ngroups + gridDim.x - 1) / gridDim.x;
  const long lg_end = ((lb + 1) * chunk < ngroups) ? (lb + 1) * chunk : ngroups;
  for (long lg = lb * chunk; lg < lg_end; ++lg) {
    // Per-thread indices are recomputed each iteration from an opaque copy of
    // the thread id: hoisted, the addresses derived from them stayed live across
    // the loop (and the transforms' twiddles, read from global memory, too),
    // 126 VGPRs at 256^3 x 17; recomputed, with the twiddles from LDS, 90.
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const int lane = tid & 63, wv = tid >> 6;
    const int job = WX ? tid / T : tid % njobs, tq = WX ? tid % T : tid / njobs;
    // threads beyond the njobs*T run the same schedule on a spare row, so the
    // transform phases carry no per-slot predicates
    const bool act = WX ? job < njobs : tq < T;
    const int t = act ? tq : tq % T;
    __builtin_assume(t >= 0 && t < T);
    const int row = act ? job : njobs;
    c2<R>* col = sh + row * MP;
    // the lone job of the last wave (njobs = 1 mod 64 / T) goes to the whole
    // wave; the condition is uniform over the wave (a scalar branch)
    const bool wide = HAVEW && njobs % JW == 1 &&
                      (__builtin_amdgcn_readfirstlane(tid) >> 6) == (njobs - 1) / JW;
    long L0 = lg * LINES, m = 0;
    c2<R>* wmm = wm;
    c2<R>* wnn = wn;
    if constexpr (BATCH) {
      const long gpm = g.plane / LINES;
      m = lg / gpm;
      L0 = (lg - m * gpm) * LINES;
      wmm = mt[m].Wm;
      wnn = mt[m].Wn;
    }
    const UPar& uu = member_u<R>(u, mt, m, BATCH);
    if constexpr (PP) {
      // 1. half lines of every field -> LDS rows, with the C2R pre-processing:
      //    wave w stages rows w, w + nw, ... (coalesced along k, and along M - k
      //    in descending order), kXlineRows rows per batch in flight
      for (int jb0 = wv; jb0 < njobs; jb0 += kXlineRows * nw) {
        c2<R> ta[kXlineRows][PPL], tb[kXlineRows][PPL];
        R xn[kXlineRows];
#pragma unroll
        for (int r = 0; r < kXlineRows; ++r) {
          const int jb = jb0 + r * nw;
          if (jb < njobs) {
            const int f = jb / LINES, l = jb - f * LINES;
            const c2<R>* src = wmm + (long)f * g.hmain + (L0 + l) * M;
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
              const int p = lane + 64 * j;
              if ((M / 2) % 64 == 0 || p < M / 2) {
                ta[r][j] = ld_nt<R>(src + p);
                tb[r][j] = ld_nt<R>(src + xline_partner<M>(p));
              }
            }
            xn[r] = (lane == 0) ? wnn[(long)f * g.plane + L0 + l].x : R(0);
          }
        }
#pragma unroll
        for (int r = 0; r < kXlineRows; ++r) {
          const int jb = jb0 + r * nw;
          if (jb < njobs) {
            c2<R>* rw = sh + jb * MP;
#pragma unroll
            for (int j = 0; j < PPL; ++j) {
              const int p = lane + 64 * j;
              if ((M / 2) % 64 == 0 || p < M / 2) {
                const c2<R> a = ta[r][j], b = tb[r][j], w = shw[p];
                rw[lds_swz<XF>(p)] = p ? c2r_pre<R>(a, b, w) : mk<R>(a.x + xn[r], a.x - xn[r]);
                rw[lds_swz<XF>(xline_partner<M>(p))] = c2r_pre<R>(b, p ? a : b, xline_mirror_tw<R>(w, p));
              }
            }
          }
        }
      }
    } else {
      // 1. half lines of every field -> LDS rows: wave w stages rows w, w + nw,
      //    ... (coalesced along k), kXlineRows rows per batch in flight
      for (int jb0 = wv; jb0 < njobs; jb0 += kXlineRows * nw) {
        c2<R> tmp[kXlineRows][KPL];
#pragma unroll
        for (int r = 0; r < kXlineRows; ++r) {
          const int jb = jb0 + r * nw;
          if (jb < njobs) {
            const int f = jb / LINES, l = jb - f * LINES;
            const c2<R>* src = wmm + (long)f * g.hmain + (L0 + l) * M;
#pragma unroll
            for (int kk = 0; kk < KPL; ++kk)
              if (M % 64 == 0 || lane + 64 * kk < M) tmp[r][kk] = ld_nt<R>(src + lane + 64 * kk);
          }
        }
#pragma unroll
        for (int r = 0; r < kXlineRows; ++r) {
          const int jb = jb0 + r * nw;
          if (jb < njobs) {
#pragma unroll
            for (int kk = 0; kk < KPL; ++kk)
              if (M % 64 == 0 || lane + 64 * kk < M) sh[jb * MP + lds_swz<XF>(lane + 64 * kk)] = tmp[r][kk];
          }
        }
      }
      for (int i = tid; i < njobs; i += nt) {
        const int f = i / LINES, l = i - f * LINES;
        shn[i] = wnn[(long)f * g.plane + L0 + l].x;
      }
    }
    __syncthreads();
    // 2-3. (C2R pre-processing and) inverse x transform of each job; real
    //      pairs z[n] = (x[2n], x[2n+1])
    if (wide) {
      if (TW == 64 || lane < TW) {
        if constexpr (PP) xline_fft<R, M, WV, XF, true>(v, sh + (njobs - 1) * MP, shw, lane);
        else xline_c2r<R, M, WV, XF>(v, sh + (njobs - 1) * MP, shw, shn + njobs - 1, lane);
      }
    } else if (!WX || act) {
      if constexpr (PP) xline_fft<R, M, 0, XF, true, TR>(v, col, shw, t, twr);
      else xline_c2r<R, M, 0, XF>(v, col, shw, shn + row, t);
    }
    __syncthreads();
    // 4. pointwise: point x of line l is component x%2 of element x/2 of row
    //    f*LINES + l.  When the group has more points than the block has threads
    //    (256 points, 192 threads at 256^3 x 17), each thread takes the two
    //    points of one element: one pass instead of two for the first wave
    //    (measured 0.645 -> 0.616 ms); otherwise one point per thread (pairs
    //    measured 4% slower at 128^3, where the points fit one pass anyway).
    if (!XONLY && LINES * NX > nt) {
      constexpr long fs = (long)LINES * MP;  // field stride in complex elements
      for (int p = tid; p < LINES * M; p += nt) {
        const int l = p / M, k = p - l * M;
        c2<R>* vv = sh + (long)l * MP + lds_swz<XF>(k);
        const long gp = (L0 + l) * NX + 2 * k;  // point index within a field
        R mo[2];
        if (diag) nl_points<R, 2, true>(vv, fs, g.var, uu, mo, a_ch, a_eta, cmax, cmin, pmin, p3);
        else nl_points<R, 2, false>(vv, fs, g.var, uu, mo, a_ch, a_eta, cmax, cmin, pmin, p3);
        if (mob) {
          mobd[gp] = mo[0];
          mobd[gp + 1] = mo[1];
        }
      }
    } else {
      R* rl = reinterpret_cast<R*>(sh);
      constexpr long fs = 2L * LINES * MP;  // field stride in reals
      for (int p = tid; p < LINES * NX; p += nt) {
        const int l = p / NX, x = p - l * NX;
        R* vv = rl + 2L * ((long)l * MP + lds_swz<XF>(x >> 1)) + (x & 1);
        const long gp = (L0 + l) * NX + x;
        if (XONLY) {
          vv[0] = vv[0] * mobd[gp];
        } else {
          R mo[1];
          if (diag) nl_points<R, 1, true>(vv, fs, g.var, uu, mo, a_ch, a_eta, cmax, cmin, pmin, p3);
          else nl_points<R, 1, false>(vv, fs, g.var, uu, mo, a_ch, a_eta, cmax, cmin, pmin, p3);
          if (mob) mobd[gp] = mo[0];
        }
      }
    }
    __syncthreads();
    // 5. forward x transform of each job.  With TR (256-point FP32 rows, 16
    //    threads a job, one 16-lane DPP row each) a job's threads also do its R2C
    //    post-processing and store: after the last pass thread t holds
    //    Z[t + 16 s] in slot s, whose partner Z[M - t - 16 s] is slot 15 - s of
    //    thread 16 - t (for t = 0: its own slot (16 - s) mod 16), so the pairs
    //    come by DPP instead of a write-back and a second read of the row (512^3
    //    x 37: x line 13.5-13.8 -> 13.0-13.1 ms).  The lone wide job, if any,
    //    still goes through its row and phase 6.
    const bool wide_any = HAVEW && njobs % JW == 1;
    if (wide) {
      if (TW == 64 || lane < TW) xline_fft<R, M, WV, XF, false>(v, sh + (njobs - 1) * MP, shw, lane);
    } else if (!WX || act) {
      if constexpr (TR) {
        static_assert(T == 16 && kInSlots<M> == 16 && kOutSlots<M> == 16, "16 threads, 16 slots a job");
#pragma unroll
        for (int s = 0; s < 16; ++s) v[s] = col[lds_swz<XF>(rfft_in_idx<M>(t, s))];
        rfft<R, M, false, 1, 0, XF, true>(v, t, true, col, shw, 2, 0, twr);
        const int f = job / LINES, l = job - f * LINES;
        c2<R>* dst = wmm + (long)f * g.hmain + (L0 + l) * M;
        const R hs = R(0.5) * scale;
#pragma unroll
        for (int s = 0; s < 16; ++s) {
          const int k = rfft_out_idx<M>(t, s);  // t + 16 s
          const c2<R> src = v[15 - s];
          c2<R> pr = mk<R>(dpp_partner16(src.x), dpp_partner16(src.y));
          if (t == 0) pr = v[(16 - s) % 16];
          st_nt<R>(dst + k, r2c_post<R>(v[s], pr, shw[k], hs));
        }
        if (t == 0) wnn[(long)f * g.plane + L0 + l] = mk<R>((v[0].x - v[0].y) * scale, R(0));
      } else {
        xline_fft<R, M, 0, XF, false, TR>(v, col, shw, t, twr);
      }
    }
    __syncthreads();
    // 6. R2C post-processing and store, one row per wave: coalesced along k, or
    //    with the pairs (p, M - p) of 1 per lane, along k and M - k (with TR only
    //    the wide job's row)
    const R hs = R(0.5) * scale;
    for (int jb = wv; jb < njobs; jb += nw) {
      if (TR && !(wide_any && jb == njobs - 1)) continue;
      const int f = jb / LINES, l = jb - f * LINES;
      c2<R>* dst = wmm + (long)f * g.hmain + (L0 + l) * M;
      const c2<R>* rw = sh + jb * MP;
      if constexpr (PP) {
#pragma unroll
        for (int j = 0; j < PPL; ++j) {
          const int p = lane + 64 * j;
          if ((M / 2) % 64 == 0 || p < M / 2) {
            const int q = xline_partner<M>(p);
            const c2<R> a = rw[lds_swz<XF>(p)], b = rw[lds_swz<XF>(q)], w = shw[p];
            st_nt<R>(dst + p, r2c_post<R>(a, p ? b : a, w, hs));
            st_nt<R>(dst + q, r2c_post<R>(b, p ? a : b, xline_mirror_tw<R>(w, p), hs));
          }
        }
      } else {
#pragma unroll
        for (int kk = 0; kk < KPL; ++kk) {
          const int k = lane + 64 * kk;
          if (M % 64 == 0 || k < M)
            st_nt<R>(dst + k, r2c_post<R>(rw[lds_swz<XF>(k)], rw[lds_swz<XF>((M - k) % M)], shw[k], hs));
        }
      }
    }
    for (int i = tid; i < njobs; i += nt) {
      if (TR && !(wide_any && i == njobs - 1)) continue;
      const int f = i / LINES, l = i - f * LINES;
      const c2<R> z0 = sh[i * MP];
      wnn[(long)f * g.plane + L0 + l] = mk<R>((z0.x - z0.y) * scale, R(0));
    }
    __syncthreads();
  }
  const int tid = threadIdx.x;
  if (!XONLY && diag) {
    const double s0 = bsum(a_ch, red), s1 = bsum(a_eta, red);
    const double e0 = bext<true>((double)cmax, red), e1 = bext<false>((double)cmin, red);
    const double e2 = bext<false>((double)pmin, red), e3 = bext<true>((double)p3, red);
    if (tid == 0) {
      double* o = part + (long)blockIdx.x * 6;
      o[0] = s0; o[1] = s1; o[2] = e0; o[3] = e1; o[4] = e2; o[5] = e3;
    }
  }
}

// ---------------------------------------------------------------------------
// k_ypass: y transform of the main block (column tiles of W = M, panels (z, f))
// and of the Nyquist planes (contiguous lines of length NY), one launch.
// ---------------------------------------------------------------------------
// BATCH: the W banks of every member in mt, ppm panels (and Nyquist lines) per
// member (fields x nz); main and nyq are unused.
template <typename R, int NY, bool INV, bool BATCH = false>
__global__ void __launch_bounds__(col_tile(NY, sizeof(c2<R>)) * plan_T(NY))
k_ypass(c2<R>* __restrict__ main, c2<R>* __restrict__ nyq, long M, long npanels, long nlines,
        const c2<R>* __restrict__ tw, const Member<R>* __restrict__ mt, long ppm) {
  constexpr int T = plan_T(NY), E = plan_E(NY);
  constexpr int TX = col_tile(NY, sizeof(c2<R>));
  __shared__ c2<R> sh[(plan_np(NY) > 1) ? NY * TX : 1];
  const long tpp = (M + TX - 1) / TX;
  const long nmain = tpp * npanels;
  const long b = xcd_block(blockIdx.x, gridDim.x);
  c2<R> v[E];
  if (b < nmain) {
    const long pl = b / tpp;
    long pnl = pl;
    c2<R>* mb = main;
    if constexpr (BATCH) {
      const long m = pl / ppm;
      pnl = pl - m * ppm;
      mb = mt[m].Wm;
    }
    const long col0 = (b - pl * tpp) * TX;
    const int ncol = (int)((M - col0) < TX ? (M - col0) : TX);
    const int c = threadIdx.x % TX, t = threadIdx.x / TX;
    const bool act = c < ncol;
    c2<R>* gp = mb + pnl * NY * M + col0 + c;
#pragma unroll
    for (int s = 0; s < kInSlots<NY>; ++s)
      if (act && rfft_in_valid<NY>(t, s)) v[s] = ld_nt<R>(gp + (long)rfft_in_idx<NY>(t, s) * M);
    rfft<R, NY, INV, TX>(v, t, act, sh + c, tw);
#pragma unroll
    for (int s = 0; s < kOutSlots<NY>; ++s)
      if (act && rfft_out_valid<NY>(t, s)) st_nt<R>(gp + (long)rfft_out_idx<NY>(t, s) * M, v[s]);
  } else {
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    const long line = (b - nmain) * TX + l;
    const bool act = line < nlines;
    long gl = line;
    c2<R>* nb = nyq;
    if constexpr (BATCH) {
      const long ln = act ? line : 0;
      const long m = ln / ppm;
      gl = ln - m * ppm;
      nb = mt[m].Wn;
    }
    c2<R>* gp = nb + (act ? gl : 0) * NY;
#pragma unroll
    for (int s = 0; s < kInSlots<NY>; ++s)
      if (act && rfft_in_valid<NY>(t, s)) v[s] = ld_nt<R>(gp + rfft_in_idx<NY>(t, s));
    rfft<R, NY, INV, 1>(v, t, act, sh + ((plan_np(NY) > 1) ? l * NY : 0), tw);
#pragma unroll
    for (int s = 0; s < kOutSlots<NY>; ++s)
      if (act && rfft_out_valid<NY>(t, s)) st_nt<R>(gp + rfft_out_idx<NY>(t, s), v[s]);
  }
}

// ---------------------------------------------------------------------------
// k_zline: forward z -> update -> inverse z on column tiles of each field.
// ROLE_REG : c (D_gb = 0 form) and eta updates; W -> S, W
// ROLE_GHAT: c only; W_c -> f^ (stored to X) and g^ = kf_sum f^ - kappa_c c^ k_cube,
//            inverse z of g^ -> Y
// ROLE_CMOB: c only; h^ from Y (forward z) with f^ from X -> S_c, W_c
// The spectrum values a thread holds after the forward transform are updated in
// registers (S read and written at the same rows); when the transform's output
// rows differ from its input rows an LDS exchange precedes the inverse.
// S is stored in the kernel's own tile order: tile tt of field f (tt < tm: main
// block columns tt*TX .., else Nyquist-plane columns) is NZ rows of TX complex,
// contiguous, at S.main + ((f*(tm+tn) + tt)*NZ + z)*TX + column (Geo::stx, stm,
// stpf describe it).  Every tile of S is then one contiguous 32 KiB (FP32,
// NZ = 256) block rather than NZ row segments one z plane apart.  At 256^3 x 17
// a copy of the pass's data movement alone went 0.99 -> 0.87 ms; the pass
// itself, with the buffer addressing below, 0.962 -> 0.929 ms.
// The wavenumbers and the update coefficients are evaluated in FP64 in both
// precisions: in FP32 a rounded coefficient A = 1/(1 + dt M kappa k^4) would
// bias every step the same way, and over 1000 steps of a slowly decaying mode
// that bias grows linearly (measured 1.5e-4 max error in c at 256^3 x 17,
// against rounding of the products alone, which averages out).
// ---------------------------------------------------------------------------
enum { ROLE_REG = 0, ROLE_GHAT = 1, ROLE_CMOB = 2 };

// 1/x in FP64 for x in [1, 2^60): hardware reciprocal estimate and two Newton
// steps (each doubles the correct bits), within a few ulp of the quotient.
__device__ __forceinline__ double recip_fast(double x) {
  double r = __builtin_amdgcn_rcp(x);
  r = r * (2.0 - x * r);
  return r * (2.0 - x * r);
}

// Minimum waves per SIMD asked of the register allocator: 3 for FP32 plans with
// up to 16 complex values per thread, 2 for plans holding more (E > 16) and for
// FP64.  At 256 the FP32 update takes 165 VGPRs (3 waves) with no spills;
// capped at 4 waves it spilled 6-25 registers per lane and ran 16% slower.
template <typename R, bool PARSEVAL, int E = 16> constexpr int kZlineWaves = (sizeof(R) == 4 && E <= 16) ? 3 : 2;
// 512-point FP32 columns with the 32-thread plans (V = 1, 2: 512-thread blocks,
// 64 KiB of LDS): capped at 128 VGPRs (4 waves per SIMD) two blocks share a CU,
// one loading while the other transforms.  They spill 11-15 VGPRs per lane that
// way and still take 16.3-16.5 ms per update pass at 512^3 x 37, against
// 21.0-22.3 ms uncapped (one block per CU) and 20.9 ms for the default plan's
// 1 024-thread blocks.
template <typename R, int NZ, int V, bool PARSEVAL>
constexpr int kZlineWavesN = (sizeof(R) == 4 && NZ >= 512 && V > 0) ? 4 : kZlineWaves<R, PARSEVAL, plan_E(NZ, V)>;

// The update coefficients of one parameter set, FP64.
struct ZCo {
  double Dk, Ddt, Lkd, Ldt, dt, kc, kphi;
};
__device__ __forceinline__ ZCo zco(const UPar& u) {
  return ZCo{(u.D * u.dt) * u.kc, u.D * u.dt, (u.L * u.kphi) * u.dt, u.L * u.dt, u.dt, u.kc, u.kphi};
}

// BATCH (ROLE_REG only): the fields f0 .. f0 + nf - 1 run over every member in
// mt, g.F per member, each with its own S, W, k = 0 modes and parameters; S, W
// and dc64 are unused.
template <typename R, int NZ, int ROLE, bool DC64, bool PARSEVAL, bool BATCH = false, int V = 0>
__global__ void __launch_bounds__(col_tile(NZ, sizeof(c2<R>)) * plan_T(NZ, V))
__attribute__((amdgpu_waves_per_eu(kZlineWavesN<R, NZ, V, PARSEVAL>)))
k_zline(Geo g, UPar u, KTab kt, Bank<R> S, Bank<R> W, Bank<R> X, Bank<R> Y,
        double* __restrict__ dc64, double* __restrict__ part, const c2<R>* __restrict__ tw, int f0, int nf,
        long tm, long tn, long run_tiles, const Member<R>* __restrict__ mt) {
  static_assert(!BATCH || ROLE == ROLE_REG, "batched members step without the mobility chain");
  constexpr int TX = col_tile(NZ, sizeof(c2<R>));
  constexpr int NI = kInSlots<NZ, V>, NO = kOutSlots<NZ, V>, E = plan_E(NZ, V);
  constexpr unsigned CB = sizeof(c2<R>);
  static_assert(last_pass_in_place<NZ, V>(), "slot rows of the last pass are t + a per-slot constant");
  // the exchange tile (up to 64 KiB), which after the tile loop holds the
  // block reduction's 16 doubles
  constexpr int SHN = (plan_np(NZ, V) > 1) ? NZ * TX : 1;
  __shared__ c2<R> sh[SHN * sizeof(c2<R>) >= 16 * sizeof(double) ? SHN : 16 * sizeof(double) / sizeof(c2<R>)];
  double* const red = reinterpret_cast<double*>(sh);
  const long tpf = tm + tn;
  const long total = (long)nf * tpf;
  const int c = threadIdx.x % TX, t = threadIdx.x / TX;
  const ZCo z0c = zco(u);
  double acc = 0.0;
  c2<R> v[E];
  c2<R> sv[NO];
  c2<R> xv[(ROLE == ROLE_CMOB) ? NO : 1];
  const long tile0 = xcd_runs(blockIdx.x, gridDim.x, run_tiles);
  for (long tile = tile0; tile < total; tile += gridDim.x) {
    int f = f0 + (int)(tile / tpf);
    Bank<R> Sb = S, Wb = W;
    double* dcb = dc64;
    ZCo zc = z0c;
    if constexpr (BATCH) {
      const long m = f / g.F;
      f -= (int)(m * g.F);
      Sb = Bank<R>{mt[m].S, nullptr};
      Wb = Bank<R>{mt[m].Wm, mt[m].Wn};
      dcb = mt[m].dc64;
      zc = zco(mt[m].u);
    }
    const long tt = tile % tpf;
    const bool isn = tt >= tm;
    const long Wd = isn ? (long)g.ny : (long)g.M * g.ny;
    const long col0 = (isn ? tt - tm : tt) * TX;
    const int ncol = (int)((Wd - col0) < TX ? (Wd - col0) : TX);
    const bool act = c < ncol;
    // Buffer descriptors on the tile's first column (block-uniform): the
    // natural-layout banks W, X, Y (z rows Wd columns apart; X, Y hold field c
    // only) and the tile of S.  A thread's row of slot s is its own t plus a
    // per-slot constant (first pass: in_idx; last pass, in place: out_idx), so
    // every access is one lane offset plus a uniform one.
    const unsigned rowb = (unsigned)Wd * CB, nb = (unsigned)NZ * rowb;
    const BufRes rw = buf_res((isn ? Wb.nyq : Wb.main) + (isn ? (long)f * g.plane : (long)f * g.hmain) + col0, nb);
    const BufRes rx = buf_res((isn ? X.nyq : X.main) + col0, nb);
    const BufRes ry = buf_res((isn ? Y.nyq : Y.main) + col0, nb);
    const BufRes rs = buf_res(Sb.main + ((long)f * tpf + tt) * ((long)NZ * TX), (unsigned)(NZ * TX) * CB);
    const unsigned lw = ((unsigned)t * (unsigned)Wd + (unsigned)c) * CB;  // natural layout
    const unsigned ls = (unsigned)threadIdx.x * CB;                        // S tile: row t, column c
    // 1. all loads of the tile at once: transform input, spectrum (and f^) at the output rows
#pragma unroll
    for (int s = 0; s < NI; ++s)
      if (act && rfft_in_valid<NZ, V>(t, s))
        v[s] = buf_ld<R>((ROLE == ROLE_CMOB) ? ry : rw, lw, (unsigned)rfft_in_idx<NZ, V>(0, s) * rowb);
#pragma unroll
    for (int s = 0; s < NO; ++s)
      if (act && rfft_out_valid<NZ, V>(t, s)) {
        sv[s] = buf_ld<R>(rs, ls, (unsigned)(rfft_out_idx<NZ, V>(0, s) * TX) * CB);
        if constexpr (ROLE == ROLE_CMOB) xv[s] = buf_ld<R, 0>(rx, lw, (unsigned)rfft_out_idx<NZ, V>(0, s) * rowb);
      }
    const long col = col0 + c;
    const int kx = isn ? g.M : (int)(col % g.M);
    const int ky = isn ? (int)col : (int)(col / g.M);
    const KCol<double> kc = kcol<double>(kt, kx, ky);
    const double kap = (f == 0) ? zc.kc : zc.kphi;
    // The z wavenumbers of a thread's rows are the same for every tile; an
    // opaque copy of t keeps the compiler from hoisting all of them (2 doubles
    // per slot) out of the tile loop into registers, recomputing them instead.
    int tz = t;
    asm volatile("" : "+v"(tz));
    // 2. forward z, update in registers
    rfft<R, NZ, false, TX, V>(v, t, act, sh + c, tw);
#pragma unroll
    for (int s = 0; s < NO; ++s) {
      if (!(act && rfft_out_valid<NZ, V>(t, s))) continue;
      const int e = rfft_out_idx<NZ, V>(t, s);
      const KCoefR<double> k = kcoef_z(kc, rfft_out_idx<NZ, V>(tz, s), NZ, kt.uz);
      const c2<R> w = v[s];
      c2<R> sn;
      if constexpr (ROLE == ROLE_GHAT) {
        const c2<R> ch = sv[s];
        buf_st<R, 0>(rx, lw, (unsigned)rfft_out_idx<NZ, V>(0, s) * rowb, w);  // f^
        sn = mk<R>((R)(-(k.s1 * w.y) + zc.kc * ch.y * k.k3), (R)(k.s1 * w.x - zc.kc * ch.x * k.k3));
      } else {
        const c2<R> s0 = sv[s];
        // A = 1/den multiplies the whole state each step; B (and Cc) the step's
        // increment
        const double den = (f == 0) ? 1.0 + zc.Dk * k.k4 : 1.0 - zc.Lkd * k.k2;
        if constexpr (sizeof(R) == 8) {
          const double A = 1.0 / den;
          const double B = (f == 0) ? (zc.Ddt * k.k2) * A : -zc.Ldt * A;
          if constexpr (ROLE == ROLE_CMOB) {
            const c2<R> fh = xv[s];
            const double Cc = (zc.dt * k.s1) * A;
            sn = mk<R>(A * s0.x + B * fh.x - Cc * w.y, A * s0.y + B * fh.y + Cc * w.x);
          } else {
            sn = mk<R>(A * s0.x + B * w.x, A * s0.y + B * w.y);
          }
        } else {
          // FP32 state: A carried as an unevaluated sum hi + lo of two floats, so
          // the only per-step error on the state is the rounding of the products
          const double A = recip_fast(den);
          const float Ah = (float)A, Al = (float)(A - (double)Ah);
          const float B = (float)((f == 0) ? (zc.Ddt * k.k2) * A : -zc.Ldt * A);
          float ix = B * (float)w.x, iy = B * (float)w.y;
          if constexpr (ROLE == ROLE_CMOB) {
            const c2<R> fh = xv[s];
            const float Cc = (float)((zc.dt * k.s1) * A);
            ix = B * fh.x - Cc * w.y;
            iy = B * fh.y + Cc * w.x;
          }
          sn = mk<R>(fmaf(Ah, s0.x, fmaf(Al, s0.x, ix)), fmaf(Ah, s0.y, fmaf(Al, s0.y, iy)));
        }
        if (DC64 && kx == 0 && ky == 0 && e == 0) {
          if (f != 0) dcb[f] = dcb[f] - zc.Ldt * (double)w.x;
          sn = mk<R>((R)dcb[f], R(0));
        }
        buf_st<R>(rs, ls, (unsigned)(rfft_out_idx<NZ, V>(0, s) * TX) * CB, sn);

      }
      v[s] = sn;
    }
    if (PARSEVAL && ROLE != ROLE_GHAT) {
      // Parseval terms of the new spectrum, in a pass of their own after the
      // update (fewer registers live at once): |kf'|^2 = (x,y part) + azp^2,
      // the column part applied once per tile
      R pa1 = R(0), pa2 = R(0);
#pragma unroll
      for (int s = 0; s < NO; ++s) {
        if (!(act && rfft_out_valid<NZ, V>(t, s))) continue;
        const R m2 = v[s].x * v[s].x + v[s].y * v[s].y;
        const int e = rfft_out_idx<NZ, V>(t, s);
        const int ez = rfft_out_idx<NZ, V>(tz, s);
        const R zp = (2 * ez == NZ) ? R(0) : (R)az_of(ez, NZ, kt.uz);  // a'_z: a_z zeroed at Nyquist
        pa1 += m2;
        pa2 += zp * zp * m2;
      }
      const double wgt = (kx == 0 || kx == g.M) ? 1.0 : 2.0;
      acc += (0.5 * kap) * wgt * (kc.kp2 * (double)pa1 + (double)pa2);
    }
    // 3. inverse z (exchange first when the output rows are not the input rows)
    if constexpr (!rfft_io_same<NZ, V>()) {
      __syncthreads();
#pragma unroll
      for (int s = 0; s < NO; ++s)
        if (act && rfft_out_valid<NZ, V>(t, s)) sh[rfft_out_idx<NZ, V>(t, s) * TX + c] = v[s];
      __syncthreads();
#pragma unroll
      for (int s = 0; s < NI; ++s)
        if (act && rfft_in_valid<NZ, V>(t, s)) v[s] = sh[rfft_in_idx<NZ, V>(t, s) * TX + c];
    }
    rfft<R, NZ, true, TX, V>(v, t, act, sh + c, tw);
#pragma unroll
    for (int s = 0; s < NO; ++s)
      if (act && rfft_out_valid<NZ, V>(t, s))
        buf_st<R>((ROLE == ROLE_GHAT) ? ry : rw, lw, (unsigned)rfft_out_idx<NZ, V>(0, s) * rowb, v[s]);
    __syncthreads();
  }
  if (PARSEVAL) {
    const double s = bsum(acc, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s * (double)g.P;
  }
}

}  // namespace pf
