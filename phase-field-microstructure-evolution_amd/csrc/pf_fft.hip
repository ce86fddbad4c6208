// 3-D real<->half-complex transforms on gfx950: twiddle tables, size dispatch,
// whole-transform drivers, layout packing and the transform entry points of the
// C-ABI (pf_fft_r2c / pf_fft_c2r on device buffers, pf_f_trans / pf_inv_trans
// on host buffers with the reference's f_trans / inv_trans semantics).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/pf.h"
#include "pf_internal.hpp"
#include "pf_kernels.hpp"

namespace pf {

#define PF_GRID_SIZES(X) \
  X(8) X(12) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(160) X(192) X(256) X(512)

bool grid_size_supported(int n) {
  switch (n) {
#define PF_CASE(N) case N:
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    return true;
    default: return false;
  }
}
const char* supported_sizes_text() { return "8 12 16 24 32 48 64 96 128 160 192 256 512"; }

// ---------------------------------------------------------------------------
// Host transfers through pinned bounce buffers.
// ---------------------------------------------------------------------------

namespace {
constexpr size_t kBounceBytes = size_t(64) << 20;
constexpr size_t kBounceMin = size_t(128) << 20;
std::mutex g_bounce_mu;
void* g_bounce[2] = {nullptr, nullptr};

// memcpy split over up to 8 host threads
void par_copy(void* dst, const void* src, size_t n) {
  const int nth = (int)std::max(1u, std::min<unsigned>(8, std::thread::hardware_concurrency()));
  if (nth == 1 || n < (size_t(4) << 20)) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t per = (n + nth - 1) / nth;
  std::vector<std::thread> th;
  for (int w = 0; w < nth; ++w) {
    const size_t a = std::min(n, per * w), b = std::min(n, a + per);
    if (a < b) th.emplace_back([=] { std::memcpy((char*)dst + a, (const char*)src + a, b - a); });
  }
  for (auto& t : th) t.join();
}

// Under g_bounce_mu.  Pinned memory is allocated once and kept until exit.
bool bounce_ready() {
  for (int i = 0; i < 2; ++i)
    if (!g_bounce[i] && hipHostMalloc(&g_bounce[i], kBounceBytes, hipHostMallocDefault) != hipSuccess) {
      g_bounce[i] = nullptr;
      (void)hipGetLastError();
      return false;
    }
  return true;
}

struct EventPair {
  hipEvent_t e[2] = {nullptr, nullptr};
  bool ok = true;
  EventPair() {
    for (auto& x : e)
      if (hipEventCreateWithFlags(&x, hipEventDisableTiming) != hipSuccess) ok = false;
  }
  ~EventPair() {
    for (auto& x : e)
      if (x) (void)hipEventDestroy(x);
  }
};
}  // namespace

namespace {
// Under g_bounce_mu.
int to_host_locked(void* dst, const void* src, size_t bytes, hipStream_t s) {
  EventPair ev;
  if (bytes < kBounceMin || !ev.ok || !bounce_ready()) {
    PF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    PF_HIP(hipStreamSynchronize(s));
    return 0;
  }
  const size_t nch = (bytes + kBounceBytes - 1) / kBounceBytes;
  // chunk k goes into buffer k & 1; chunk k - 2 was moved out of it on the host
  // in the previous iteration
  for (size_t k = 0; k <= nch; ++k) {
    if (k < nch) {
      const size_t off = k * kBounceBytes, len = std::min(kBounceBytes, bytes - off);
      PF_HIP(hipMemcpyAsync(g_bounce[k & 1], (const char*)src + off, len, hipMemcpyDeviceToHost, s));
      PF_HIP(hipEventRecord(ev.e[k & 1], s));
    }
    if (k > 0) {
      const size_t j = k - 1, off = j * kBounceBytes, len = std::min(kBounceBytes, bytes - off);
      PF_HIP(hipEventSynchronize(ev.e[j & 1]));
      par_copy((char*)dst + off, g_bounce[j & 1], len);
    }
  }
  return 0;
}

int to_device_locked(void* dst, const void* src, size_t bytes, hipStream_t s) {
  EventPair ev;
  if (bytes < kBounceMin || !ev.ok || !bounce_ready()) {
    PF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    PF_HIP(hipStreamSynchronize(s));
    return 0;
  }
  const size_t nch = (bytes + kBounceBytes - 1) / kBounceBytes;
  for (size_t k = 0; k < nch; ++k) {
    const size_t off = k * kBounceBytes, len = std::min(kBounceBytes, bytes - off);
    // buffer k & 1 is free once the copy of chunk k - 2 out of it has finished
    if (k >= 2) PF_HIP(hipEventSynchronize(ev.e[k & 1]));
    par_copy(g_bounce[k & 1], (const char*)src + off, len);
    PF_HIP(hipMemcpyAsync((char*)dst + off, g_bounce[k & 1], len, hipMemcpyHostToDevice, s));
    PF_HIP(hipEventRecord(ev.e[k & 1], s));
  }
  // the bounce buffers are shared: they must be drained before the lock goes
  PF_HIP(hipStreamSynchronize(s));
  return 0;
}
}  // namespace

// The bounce buffers are released with the lock: on an error part-way through,
// the stream is drained first so that no copy into or out of them is still in
// flight when the next transfer takes them.
int to_host(void* dst, const void* src, size_t bytes, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_bounce_mu);
  const int rc = to_host_locked(dst, src, bytes, s);
  if (rc) (void)hipStreamSynchronize(s);
  return rc;
}

int to_device(void* dst, const void* src, size_t bytes, hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_bounce_mu);
  const int rc = to_device_locked(dst, src, bytes, s);
  if (rc) (void)hipStreamSynchronize(s);
  return rc;
}

// ---------------------------------------------------------------------------
// Twiddle tables.
// ---------------------------------------------------------------------------

// Every length's table W_n^t = exp(-2 pi i t/n), t < n, back to back in the
// order of PF_GRID_SIZES, one array per precision in static device memory.
// Each device's copy is filled once (a host-to-device copy, no allocation), on
// the first use of any table on that device.
namespace {
#define PF_SUM(N) +N
constexpr int kTwTotal = 0 PF_GRID_SIZES(PF_SUM);
#undef PF_SUM
__device__ __align__(256) double2 g_tw64[kTwTotal];
__device__ __align__(256) float2 g_tw32[kTwTotal];

int tw_offset(int n) {
  int off = 0;
#define PF_OFF(N) \
  if (n == N) return off; \
  off += N;
  PF_GRID_SIZES(PF_OFF)
#undef PF_OFF
  return -1;
}

std::mutex g_tw_mu;
std::map<int, std::pair<double2*, float2*>> g_tw;  // device -> its copies of g_tw64, g_tw32

// Under g_tw_mu: the device's table addresses, filling them on first use.
bool tw_ready(int dev, std::pair<double2*, float2*>* out) {
  auto it = g_tw.find(dev);
  if (it != g_tw.end()) { *out = it->second; return true; }
  void *p64 = nullptr, *p32 = nullptr;
  if (hipGetSymbolAddress(&p64, HIP_SYMBOL(g_tw64)) != hipSuccess ||
      hipGetSymbolAddress(&p32, HIP_SYMBOL(g_tw32)) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  std::vector<double2> h64(kTwTotal);
  std::vector<float2> h32(kTwTotal);
  int off = 0;
  auto fill = [&](int n) {
    std::vector<double> re(n), im(n);
    host_twiddles(n, re.data(), im.data());
    for (int t = 0; t < n; ++t) {
      h64[off + t].x = re[t]; h64[off + t].y = im[t];
      h32[off + t].x = (float)re[t]; h32[off + t].y = (float)im[t];
    }
    off += n;
  };
#define PF_FILL(N) fill(N);
  PF_GRID_SIZES(PF_FILL)
#undef PF_FILL
  if (hipMemcpy(p64, h64.data(), sizeof(double2) * kTwTotal, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(p32, h32.data(), sizeof(float2) * kTwTotal, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  *out = g_tw[dev] = std::make_pair((double2*)p64, (float2*)p32);
  return true;
}
}  // namespace

template <typename R> const c2<R>* twiddle_table(int n) {
  int dev = 0;
  const int off = tw_offset(n);
  if (off < 0 || hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::pair<double2*, float2*> t;
  {
    std::lock_guard<std::mutex> lk(g_tw_mu);
    if (!tw_ready(dev, &t)) return nullptr;
  }
  if constexpr (sizeof(R) == 8) return t.first + off;
  else return t.second + off;
}
template const c2<float>* twiddle_table<float>(int);
template const c2<double>* twiddle_table<double>(int);

// ---------------------------------------------------------------------------
// Pass launchers with compile-time sizes.
// ---------------------------------------------------------------------------
static inline unsigned ceil_div(long a, long b) { return (unsigned)((a + b - 1) / b); }

template <typename R>
int x_r2c(int nx, c2<R>* data, c2<R>* nyq, long nlines, R scale, hipStream_t s) {
  const c2<R>* tw = twiddle_table<R>(nx);
  if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");
  switch (nx) {
#define PF_CASE(N)                                                                      \
  case N: {                                                                             \
    constexpr int LPB = x_lines<R>(N / 2);                                              \
    hipLaunchKernelGGL((k_x_r2c<R, N>), dim3(ceil_div(nlines, LPB)),                    \
                       dim3(LPB * plan_T(N / 2)), 0, s, data, nyq, nlines, tw, scale); \
    break;                                                                              \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported nx " + std::to_string(nx));
  }
  return 0;
}

template <typename R>
int x_c2r(int nx, c2<R>* data, const c2<R>* nyq, long nlines, hipStream_t s) {
  const c2<R>* tw = twiddle_table<R>(nx);
  if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");
  switch (nx) {
#define PF_CASE(N)                                                                   \
  case N: {                                                                          \
    constexpr int LPB = x_lines<R>(N / 2);                                           \
    hipLaunchKernelGGL((k_x_c2r<R, N>), dim3(ceil_div(nlines, LPB)),                 \
                       dim3(LPB * plan_T(N / 2)), 0, s, data, nyq, nlines, tw); \
    break;                                                                           \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported nx " + std::to_string(nx));
  }
  return 0;
}

template <typename R>
int x_c2r_sumsq(int nx, const c2<R>* data, const c2<R>* nyq, long nlines, long lines_per_field, double w0,
                double w1, double* part, long max_blocks, long* nblocks, hipStream_t s) {
  const c2<R>* tw = twiddle_table<R>(nx);
  if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");
  switch (nx) {
#define PF_CASE(N)                                                                                   \
  case N: {                                                                                          \
    constexpr int LPB = x_lines<R>(N / 2);                                                           \
    const long grid = std::max<long>(1, std::min<long>((nlines + LPB - 1) / LPB, max_blocks));       \
    hipLaunchKernelGGL((k_x_c2r_sumsq<R, N>), dim3((unsigned)grid), dim3(LPB * plan_T(N / 2)), 0, s, \
                       data, nyq, nlines, lines_per_field, w0, w1, tw, part);                        \
    *nblocks = grid;                                                                                 \
    break;                                                                                           \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported nx " + std::to_string(nx));
  }
  return 0;
}

template <typename R>
int col_pass(int n, bool inv, c2<R>* data, long W, long npanels, hipStream_t s) {
  const c2<R>* tw = twiddle_table<R>(n);
  if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");
  switch (n) {
#define PF_CASE(N)                                                                          \
  case N: {                                                                                 \
    constexpr int TX = col_tile(N, sizeof(c2<R>));                                          \
    const long tpp = (W + TX - 1) / TX;                                                     \
    dim3 grid((unsigned)(tpp * npanels)), block(TX * plan_T(N));                       \
    if (inv)                                                                                \
      hipLaunchKernelGGL((k_col_fft<R, N, true>), grid, block, 0, s, data, W, tpp, tw);    \
    else                                                                                    \
      hipLaunchKernelGGL((k_col_fft<R, N, false>), grid, block, 0, s, data, W, tpp, tw);   \
    break;                                                                                  \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported axis length " + std::to_string(n));
  }
  return 0;
}

template <typename R>
int row_pass(int n, bool inv, c2<R>* data, long nlines, hipStream_t s) {
  const c2<R>* tw = twiddle_table<R>(n);
  if (!tw) return fail(PF_EHIP, "twiddle table set-up failed");
  switch (n) {
#define PF_CASE(N)                                                                      \
  case N: {                                                                             \
    constexpr int LPB = lines_per_block(N);                                             \
    dim3 grid(ceil_div(nlines, LPB)), block(LPB * plan_T(N));                      \
    if (inv)                                                                            \
      hipLaunchKernelGGL((k_row_fft<R, N, true>), grid, block, 0, s, data, nlines, tw);  \
    else                                                                                \
      hipLaunchKernelGGL((k_row_fft<R, N, false>), grid, block, 0, s, data, nlines, tw); \
    break;                                                                              \
  }
    PF_GRID_SIZES(PF_CASE)
#undef PF_CASE
    default: return fail(PF_EINVAL, "unsupported axis length " + std::to_string(n));
  }
  return 0;
}

#define PF_INST(R)                                                                \
  template int x_r2c<R>(int, c2<R>*, c2<R>*, long, R, hipStream_t);               \
  template int x_c2r<R>(int, c2<R>*, const c2<R>*, long, hipStream_t);            \
  template int x_c2r_sumsq<R>(int, const c2<R>*, const c2<R>*, long, long, double, double, double*, long, \
                              long*, hipStream_t);                                 \
  template int col_pass<R>(int, bool, c2<R>*, long, long, hipStream_t);           \
  template int row_pass<R>(int, bool, c2<R>*, long, hipStream_t);
PF_INST(float)
PF_INST(double)
#undef PF_INST

// ---------------------------------------------------------------------------
// Whole 3-D transforms (P3DFFT order: forward x, y, z; inverse z, y, x).
// ---------------------------------------------------------------------------
template <typename R>
int fwd3d(c2<R>* data, c2<R>* nyq, int nx, int ny, int nz, int batch, R scale, hipStream_t s) {
  const long M = nx / 2;
  int rc;
  if ((rc = x_r2c<R>(nx, data, nyq, (long)ny * nz * batch, scale, s))) return rc;
  if ((rc = col_pass<R>(ny, false, data, M, (long)nz * batch, s))) return rc;
  if ((rc = row_pass<R>(ny, false, nyq, (long)nz * batch, s))) return rc;
  if ((rc = col_pass<R>(nz, false, data, M * ny, batch, s))) return rc;
  if ((rc = col_pass<R>(nz, false, nyq, ny, batch, s))) return rc;
  return 0;
}

template <typename R>
int inv3d(c2<R>* data, c2<R>* nyq, int nx, int ny, int nz, int batch, hipStream_t s) {
  const long M = nx / 2;
  int rc;
  if ((rc = col_pass<R>(nz, true, data, M * ny, batch, s))) return rc;
  if ((rc = col_pass<R>(nz, true, nyq, ny, batch, s))) return rc;
  if ((rc = col_pass<R>(ny, true, data, M, (long)nz * batch, s))) return rc;
  if ((rc = row_pass<R>(ny, true, nyq, (long)nz * batch, s))) return rc;
  if ((rc = x_c2r<R>(nx, data, nyq, (long)ny * nz * batch, s))) return rc;
  return 0;
}
template int fwd3d<float>(c2<float>*, c2<float>*, int, int, int, int, float, hipStream_t);
template int fwd3d<double>(c2<double>*, c2<double>*, int, int, int, int, double, hipStream_t);
template int inv3d<float>(c2<float>*, c2<float>*, int, int, int, int, hipStream_t);
template int inv3d<double>(c2<double>*, c2<double>*, int, int, int, int, hipStream_t);

// ---------------------------------------------------------------------------
// Split <-> dense half-spectrum layouts.
//   XFAST: kx + (M+1)*(ky + ny*(kz + nz*f))     (numpy rfftn of a (nz,ny,nx) array)
//   ZFAST: kz + nz*(ky + ny*(kx + (M+1)*f))     (P3DFFT STRIDE1 Z-pencil, as f_trans)
// ---------------------------------------------------------------------------
template <typename R, bool PACK>
__global__ void k_layout(const c2<R>* __restrict__ a, c2<R>* __restrict__ b, c2<R>* __restrict__ nq,
                         int M, int ny, int nz, long total, int layout) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const long mx = M + 1;
  long kx, ky, kz, f;
  if (layout == LAYOUT_XFAST) {
    kx = i % mx; long r = i / mx; ky = r % ny; r /= ny; kz = r % nz; f = r / nz;
  } else {
    kz = i % nz; long r = i / nz; ky = r % ny; r /= ny; kx = r % mx; f = r / mx;
  }
  const long plane = (long)ny * nz;
  if (PACK) {
    // a = main (split), nq = nyquist, b = dense
    const c2<R> v = (kx < M) ? a[kx + M * (ky + ny * (kz + nz * f))] : nq[ky + ny * kz + plane * f];
    b[i] = v;
  } else {
    // a = dense, b = main, nq = nyquist
    const c2<R> v = a[i];
    if (kx < M) b[kx + M * (ky + ny * (kz + nz * f))] = v;
    else nq[ky + ny * kz + plane * f] = v;
  }
}

template <typename R>
int pack_spec(const c2<R>* main, const c2<R>* nyq, c2<R>* dense, int nx, int ny, int nz,
              int batch, int layout, hipStream_t s) {
  const long total = (long)(nx / 2 + 1) * ny * nz * batch;
  hipLaunchKernelGGL((k_layout<R, true>), dim3(ceil_div(total, 256)), dim3(256), 0, s, main, dense,
                     const_cast<c2<R>*>(nyq), nx / 2, ny, nz, total, layout);
  return 0;
}
template <typename R>
int unpack_spec(const c2<R>* dense, c2<R>* main, c2<R>* nyq, int nx, int ny, int nz, int batch,
                int layout, hipStream_t s) {
  const long total = (long)(nx / 2 + 1) * ny * nz * batch;
  hipLaunchKernelGGL((k_layout<R, false>), dim3(ceil_div(total, 256)), dim3(256), 0, s, dense, main,
                     nyq, nx / 2, ny, nz, total, layout);
  return 0;
}
template int pack_spec<float>(const c2<float>*, const c2<float>*, c2<float>*, int, int, int, int, int, hipStream_t);
template int pack_spec<double>(const c2<double>*, const c2<double>*, c2<double>*, int, int, int, int, int, hipStream_t);
template int unpack_spec<float>(const c2<float>*, c2<float>*, c2<float>*, int, int, int, int, int, hipStream_t);
template int unpack_spec<double>(const c2<double>*, c2<double>*, c2<double>*, int, int, int, int, int, hipStream_t);

}  // namespace pf

// ===========================================================================
// C-ABI transform entry points.
// ===========================================================================
namespace {
using namespace pf;

int check_dims(int nx, int ny, int nz, int batch) {
  if (!grid_size_supported(nx) || !grid_size_supported(ny) || !grid_size_supported(nz))
    return fail(PF_EINVAL, std::string("grid ") + std::to_string(nx) + "x" + std::to_string(ny) +
                               "x" + std::to_string(nz) + " not supported; each axis must be one of " +
                               supported_sizes_text());
  if (batch < 1) return fail(PF_EINVAL, "batch must be >= 1");
  return 0;
}

// Scratch of pf_fft_r2c / pf_fft_c2r: one buffer per device, shared by every
// caller and stream.  A call holds the device's lock from taking the buffer
// until its work is enqueued; that work first waits (on the caller's stream)
// for the event the previous user recorded after its last use, and records the
// event again afterwards, so uses on different streams or threads run one after
// another on the GPU.  A buffer that must grow is freed only after that event
// has completed.
struct Workspace {
  std::mutex mu;
  void* p = nullptr;
  size_t bytes = 0;
  hipEvent_t done = nullptr;
};
std::mutex g_ws_mu;
std::map<int, std::unique_ptr<Workspace>> g_ws;

Workspace* device_workspace() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto& w = g_ws[dev];
  if (!w) w.reset(new Workspace);
  return w.get();
}

// Holds the device workspace (locked) for one call on stream s.
struct WorkspaceLease {
  Workspace* w = nullptr;
  std::unique_lock<std::mutex> lk;
  hipStream_t s;
  void* p = nullptr;
  std::string err;
  WorkspaceLease(size_t bytes, hipStream_t s_) : s(s_) {
    w = device_workspace();
    if (!w) { err = "no current device"; return; }
    lk = std::unique_lock<std::mutex>(w->mu);
    if (!w->done && hipEventCreateWithFlags(&w->done, hipEventDisableTiming) != hipSuccess) {
      w->done = nullptr;
      err = "event creation failed";
      return;
    }
    if (w->bytes < bytes) {
      if (w->p) {
        if (hipEventSynchronize(w->done) != hipSuccess) { err = "waiting for the previous use failed"; return; }
        (void)hipFree(w->p);
      }
      w->p = nullptr;
      w->bytes = 0;
      if (hipMalloc(&w->p, bytes) != hipSuccess) {
        w->p = nullptr;
        (void)hipGetLastError();
        err = "workspace allocation of " + std::to_string(bytes) + " bytes failed";
        return;
      }
      w->bytes = bytes;
    } else if (hipStreamWaitEvent(s, w->done, 0) != hipSuccess) {
      err = "hipStreamWaitEvent failed";
      return;
    }
    p = w->p;
  }
  // Marks the end of this call's use (after its last enqueued access).
  ~WorkspaceLease() {
    if (p) (void)hipEventRecord(w->done, s);
  }
};

// Bytes of scratch a batch of transforms needs: the split-layout copy (main
// block + Nyquist planes).
template <typename R> size_t r2c_scratch_bytes(int nx, int ny, int nz, int batch) {
  return sizeof(R) * (size_t)nx * ny * nz * batch + sizeof(c2<R>) * (size_t)ny * nz * batch;
}

// The forward transform with its scratch at ws.
template <typename R>
int r2c_core(int nx, int ny, int nz, int batch, const void* in, void* out, int layout, double scale,
             void* wsp, hipStream_t s) {
  const size_t P = (size_t)nx * ny * nz;
  char* ws = (char*)wsp;
  c2<R>* main = (c2<R>*)ws;
  c2<R>* nyq = (c2<R>*)(ws + sizeof(R) * P * batch);
  PF_HIP(hipMemcpyAsync(main, in, sizeof(R) * P * batch, hipMemcpyDeviceToDevice, s));
  int rc = fwd3d<R>(main, nyq, nx, ny, nz, batch, (R)scale, s);
  if (rc) return rc;
  rc = pack_spec<R>(main, nyq, (c2<R>*)out, nx, ny, nz, batch, layout, s);
  if (rc) return rc;
  PF_HIP(hipGetLastError());
  return 0;
}

template <typename R>
int r2c_dev(int nx, int ny, int nz, int batch, const void* in, void* out, int layout, double scale,
            hipStream_t s) {
  WorkspaceLease lease(r2c_scratch_bytes<R>(nx, ny, nz, batch), s);
  if (!lease.p) return fail(PF_ENOMEM, lease.err);
  return r2c_core<R>(nx, ny, nz, batch, in, out, layout, scale, lease.p, s);
}

template <typename R>
int c2r_dev(int nx, int ny, int nz, int batch, const void* in, void* out, int layout, hipStream_t s) {
  const size_t P = (size_t)nx * ny * nz, plane = (size_t)ny * nz;
  WorkspaceLease lease(sizeof(R) * P * batch + sizeof(c2<R>) * plane * batch, s);
  if (!lease.p) return fail(PF_ENOMEM, lease.err);
  char* ws = (char*)lease.p;
  c2<R>* main = (c2<R>*)ws;
  c2<R>* nyq = (c2<R>*)(ws + sizeof(R) * P * batch);
  int rc = unpack_spec<R>((const c2<R>*)in, main, nyq, nx, ny, nz, batch, layout, s);
  if (rc) return rc;
  rc = inv3d<R>(main, nyq, nx, ny, nz, batch, s);
  if (rc) return rc;
  PF_HIP(hipMemcpyAsync(out, main, sizeof(R) * P * batch, hipMemcpyDeviceToDevice, s));
  PF_HIP(hipGetLastError());
  return 0;
}
}  // namespace

extern "C" int pf_fft_r2c(int nx, int ny, int nz, int batch, int prec, const void* d_in, void* d_out,
                          int layout, double scale, void* stream) {
  int rc = check_dims(nx, ny, nz, batch);
  if (rc) return rc;
  if (!d_in || !d_out) return fail(PF_EINVAL, "null buffer");
  if (layout != PF_LAYOUT_XFAST && layout != PF_LAYOUT_ZFAST) return fail(PF_EINVAL, "bad layout");
  hipStream_t s = (hipStream_t)stream;
  if (prec == PF_PREC_FP64) return r2c_dev<double>(nx, ny, nz, batch, d_in, d_out, layout, scale, s);
  if (prec == PF_PREC_FP32) return r2c_dev<float>(nx, ny, nz, batch, d_in, d_out, layout, scale, s);
  return fail(PF_EINVAL, "bad precision");
}

extern "C" int pf_fft_c2r(int nx, int ny, int nz, int batch, int prec, const void* d_in, void* d_out,
                          int layout, void* stream) {
  int rc = check_dims(nx, ny, nz, batch);
  if (rc) return rc;
  if (!d_in || !d_out) return fail(PF_EINVAL, "null buffer");
  if (layout != PF_LAYOUT_XFAST && layout != PF_LAYOUT_ZFAST) return fail(PF_EINVAL, "bad layout");
  hipStream_t s = (hipStream_t)stream;
  if (prec == PF_PREC_FP64) return c2r_dev<double>(nx, ny, nz, batch, d_in, d_out, layout, s);
  if (prec == PF_PREC_FP32) return c2r_dev<float>(nx, ny, nz, batch, d_in, d_out, layout, s);
  return fail(PF_EINVAL, "bad precision");
}

// ---------------------------------------------------------------------------
// pf_fft_r2c_noalloc: pf_fft_r2c with its scratch in static device memory.
// scratch is this file's array; in host code its name only stands for the
// symbol, so every use goes through the address hipGetSymbolAddress returns for
// the current device (each device holds its own copy).
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
std::mutex g_scratch_mu;
std::map<int, void*> g_scratch;  // device -> the address of its copy of scratch

// Set-up, once per device with that device current: the device address of
// scratch (and, through twiddle_table, the static twiddle tables).
void* scratch_ptr() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch.find(dev);
  if (it != g_scratch.end()) return it->second;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    (void)hipGetLastError();
    return nullptr;
  }
  g_scratch[dev] = ptr;
  return ptr;
}
}  // namespace

extern "C" int pf_fft_r2c_noalloc(int nx, int ny, int nz, int batch, int prec, const void* d_in, void* d_out,
                                  int layout, double scale, void* stream) {
  int rc = check_dims(nx, ny, nz, batch);
  if (rc) return rc;
  if (!d_in || !d_out) return fail(PF_EINVAL, "null buffer");
  if (layout != PF_LAYOUT_XFAST && layout != PF_LAYOUT_ZFAST) return fail(PF_EINVAL, "bad layout");
  if (prec != PF_PREC_FP64 && prec != PF_PREC_FP32) return fail(PF_EINVAL, "bad precision");
  const size_t need = prec == PF_PREC_FP64 ? r2c_scratch_bytes<double>(nx, ny, nz, batch)
                                           : r2c_scratch_bytes<float>(nx, ny, nz, batch);
  if (need > kNoAllocScratchBytes)
    return fail(PF_EINVAL, "pf_fft_r2c_noalloc: " + std::to_string(batch) + " field(s) of " + std::to_string(nx) +
                               "x" + std::to_string(ny) + "x" + std::to_string(nz) + " need " +
                               std::to_string(need) + " bytes of scratch; the limit is " +
                               std::to_string(kNoAllocScratchBytes));
  void* ws = scratch_ptr();
  if (!ws) return fail(PF_EHIP, "pf_fft_r2c_noalloc: no device address for the static scratch");
  hipStream_t s = (hipStream_t)stream;
  if (prec == PF_PREC_FP64) return r2c_core<double>(nx, ny, nz, batch, d_in, d_out, layout, scale, ws, s);
  return r2c_core<float>(nx, ny, nz, batch, d_in, d_out, layout, scale, ws, s);
}

// Device buffers of pf_f_trans / pf_inv_trans, per device, grown as needed and
// kept until the process ends.  These calls are synchronous: the lock is held
// for the whole call, so calls from several threads take turns.
namespace {
struct HostCallBuffers {
  std::mutex mu;
  void* in = nullptr;
  void* out = nullptr;
  size_t in_bytes = 0, out_bytes = 0;
};
std::mutex g_hcb_mu;
std::map<int, std::unique_ptr<HostCallBuffers>> g_hcb;

HostCallBuffers* host_call_buffers() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(g_hcb_mu);
  auto& b = g_hcb[dev];
  if (!b) b.reset(new HostCallBuffers);
  return b.get();
}

int grow(void** p, size_t* have, size_t want) {
  if (*have >= want) return 0;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  *have = 0;
  if (hipMalloc(p, want) != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(PF_ENOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
  }
  *have = want;
  return 0;
}
}  // namespace

// Host-buffer drop-ins for f_trans (src/f_trans.f90:1-15) and inv_trans
// (src/inv_trans.f90:1-13): FP64, the spectrum in the reference's Z-pencil order.
extern "C" int pf_f_trans(const double* in, double* out, int nx, int ny, int nz) {
  int rc = check_dims(nx, ny, nz, 1);
  if (rc) return rc;
  if (!in || !out) return fail(PF_EINVAL, "null buffer");
  const size_t P = (size_t)nx * ny * nz, H = (size_t)(nx / 2 + 1) * ny * nz;
  HostCallBuffers* b = host_call_buffers();
  if (!b) return fail(PF_EHIP, "no current device");
  std::lock_guard<std::mutex> lk(b->mu);
  if ((rc = grow(&b->in, &b->in_bytes, sizeof(double) * P))) return rc;
  if ((rc = grow(&b->out, &b->out_bytes, sizeof(double) * 2 * H))) return rc;
  rc = to_device(b->in, in, sizeof(double) * P, nullptr);
  const double scale = 1.0 / (double)(float)((long)nx * ny * nz);
  if (!rc) rc = pf_fft_r2c(nx, ny, nz, 1, PF_PREC_FP64, b->in, b->out, PF_LAYOUT_ZFAST, scale, nullptr);
  if (!rc) rc = to_host(out, b->out, sizeof(double) * 2 * H, nullptr);
  return rc;
}

extern "C" int pf_inv_trans(const double* in, double* out, int nx, int ny, int nz) {
  int rc = check_dims(nx, ny, nz, 1);
  if (rc) return rc;
  if (!in || !out) return fail(PF_EINVAL, "null buffer");
  const size_t P = (size_t)nx * ny * nz, H = (size_t)(nx / 2 + 1) * ny * nz;
  HostCallBuffers* b = host_call_buffers();
  if (!b) return fail(PF_EHIP, "no current device");
  std::lock_guard<std::mutex> lk(b->mu);
  if ((rc = grow(&b->in, &b->in_bytes, sizeof(double) * 2 * H))) return rc;
  if ((rc = grow(&b->out, &b->out_bytes, sizeof(double) * P))) return rc;
  rc = to_device(b->in, in, sizeof(double) * 2 * H, nullptr);
  if (!rc) rc = pf_fft_c2r(nx, ny, nz, 1, PF_PREC_FP64, b->in, b->out, PF_LAYOUT_ZFAST, nullptr);
  if (!rc) rc = to_host(out, b->out, sizeof(double) * P, nullptr);
  return rc;
}
