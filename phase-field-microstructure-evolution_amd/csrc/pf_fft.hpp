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

}  // namespace pf
