// Host-side declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

#include "pf_fft.hpp"

namespace pf {

// Error state behind pf_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define PF_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return ::pf::fail(-3, std::string(#call " failed: ") + hipGetErrorString(e_) +  \
                                 " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

// Grid sizes with compiled transform kernels (every axis must be one of them).
bool grid_size_supported(int n);
const char* supported_sizes_text();

// Device table of exp(-2 pi i q/n), q in [0, n), in precision R (cached per device).
template <typename R> const c2<R>* twiddle_table(int n);

// Whole-axis transforms over a batch of fields in the split-Nyquist layout.
//   data: batch * (nx/2)*ny*nz complex == batch * nx*ny*nz reals (in place)
//   nyq : batch * ny*nz complex
template <typename R>
int fwd3d(c2<R>* data, c2<R>* nyq, int nx, int ny, int nz, int batch, R scale, hipStream_t s);
template <typename R>
int inv3d(c2<R>* data, c2<R>* nyq, int nx, int ny, int nz, int batch, hipStream_t s);

// Individual passes (used by the step pipelines).
template <typename R>
int x_r2c(int nx, c2<R>* data, c2<R>* nyq, long nlines, R scale, hipStream_t s);
template <typename R>
int x_c2r(int nx, c2<R>* data, const c2<R>* nyq, long nlines, hipStream_t s);
template <typename R>
int col_pass(int n, bool inv, c2<R>* data, long W, long npanels, hipStream_t s);
// Inverse x pass of a batch of gradient fields reduced to
// sum_lines w * sum_points out^2 (w = w0 for the lines of gradient fields 0-2,
// w1 after), one partial per block (at most max_blocks) in part.
template <typename R>
int x_c2r_sumsq(int nx, const c2<R>* data, const c2<R>* nyq, long nlines, long lines_per_field, double w0,
                double w1, double* part, long max_blocks, long* nblocks, hipStream_t s);
template <typename R>
int row_pass(int n, bool inv, c2<R>* data, long nlines, hipStream_t s);

// Copies between device memory and the caller's (pageable) host arrays on
// stream s; both return once the copy has finished.  Copies of 128 MiB and more
// go through two pinned bounce buffers of 64 MiB (process-wide, made on first
// use): the GPU moves one chunk while host threads move the neighbouring chunk
// between the other buffer and the caller's array.
int to_host(void* dst, const void* src, size_t bytes, hipStream_t s);
int to_device(void* dst, const void* src, size_t bytes, hipStream_t s);

// Dense half-spectrum layouts at the C-ABI.
enum { LAYOUT_XFAST = 0, LAYOUT_ZFAST = 1 };
template <typename R>
int pack_spec(const c2<R>* main, const c2<R>* nyq, c2<R>* dense, int nx, int ny, int nz,
              int batch, int layout, hipStream_t s);
template <typename R>
int unpack_spec(const c2<R>* dense, c2<R>* main, c2<R>* nyq, int nx, int ny, int nz,
                int batch, int layout, hipStream_t s);

}  // namespace pf
