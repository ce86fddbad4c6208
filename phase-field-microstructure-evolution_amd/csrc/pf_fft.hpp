// Complex value types and helpers shared by the device code.
//
// Transform convention everywhere: forward = e^{-i}, inverse = e^{+i}, both
// unscaled at this level.  Twiddles come from tables of exp(-2 pi i t / n) built
// on the host from long-double cos/sin with octant reduction (exact at
// multiples of pi/4); inverse transforms use conjugates.
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

// Streaming global accesses: the fused passes read and write each element of
// a bank once, so these carry the non-temporal hint (no retention in L2/MALL).
template <typename R> __device__ __forceinline__ c2<R> ld_nt(const c2<R>* p) {
  typedef R N __attribute__((ext_vector_type(2)));
  const N v = __builtin_nontemporal_load(reinterpret_cast<const N*>(p));
  return mk<R>(v.x, v.y);
}
template <typename R> __device__ __forceinline__ void st_nt(c2<R>* p, c2<R> a) {
  typedef R N __attribute__((ext_vector_type(2)));
  N v;
  v.x = a.x;
  v.y = a.y;
  __builtin_nontemporal_store(v, reinterpret_cast<N*>(p));
}

// Raw buffer accesses for the column passes: a block-uniform base address held
// in a buffer descriptor (scalar registers), a 32-bit per-lane byte offset (one
// vector register) and a block-uniform byte offset (scalar register or
// immediate).  A 64-bit pointer per register slot would instead hold two vector
// registers for the whole transform.  The descriptor's range (bytes from the
// base) bounds every access: outside it loads read 0 and stores are dropped.
// Word 3 of the descriptor is the CDNA raw-buffer format (no swizzle, no
// typed format); aux bit 1 is the non-temporal hint.
typedef __amdgpu_buffer_rsrc_t BufRes;
__device__ __forceinline__ BufRes buf_res(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
constexpr int kBufNT = 2;
template <typename R, int AUX = kBufNT> __device__ __forceinline__ c2<R> buf_ld(BufRes r, unsigned voff, unsigned soff) {
  if constexpr (sizeof(R) == 4) {
    typedef unsigned U __attribute__((ext_vector_type(2)));
    const U v = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, AUX);
    // elements copied to scalars first: __builtin_bit_cast applied to a vector
    // element (v.y) reads this compiler's element 0
    const unsigned lo = v.x, hi = v.y;
    return mk<R>(__builtin_bit_cast(float, lo), __builtin_bit_cast(float, hi));
  } else {
    typedef unsigned U __attribute__((ext_vector_type(4)));
    const U v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, AUX);
    const unsigned long long lo = (unsigned long long)v.x | ((unsigned long long)v.y << 32);
    const unsigned long long hi = (unsigned long long)v.z | ((unsigned long long)v.w << 32);
    return mk<R>(__builtin_bit_cast(double, lo), __builtin_bit_cast(double, hi));
  }
}
template <typename R, int AUX = kBufNT>
__device__ __forceinline__ void buf_st(BufRes r, unsigned voff, unsigned soff, c2<R> a) {
  if constexpr (sizeof(R) == 4) {
    typedef unsigned U __attribute__((ext_vector_type(2)));
    const float ax = a.x, ay = a.y;
    U v;
    v.x = __builtin_bit_cast(unsigned, ax);
    v.y = __builtin_bit_cast(unsigned, ay);
    __builtin_amdgcn_raw_buffer_store_b64(v, r, voff, soff, AUX);
  } else {
    typedef unsigned U __attribute__((ext_vector_type(4)));
    const double ax = a.x, ay = a.y;
    const unsigned long long lo = __builtin_bit_cast(unsigned long long, ax);
    const unsigned long long hi = __builtin_bit_cast(unsigned long long, ay);
    U v;
    v.x = (unsigned)lo;
    v.y = (unsigned)(lo >> 32);
    v.z = (unsigned)hi;
    v.w = (unsigned)(hi >> 32);
    __builtin_amdgcn_raw_buffer_store_b128(v, r, voff, soff, AUX);
  }
}

}  // namespace pf
