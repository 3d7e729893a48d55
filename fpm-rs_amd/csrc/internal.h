// internal.h — shared declarations of libfpmhip (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/fpmhip.h"

namespace fpmhip {

// ---- errors: thread-local message + stable parameter name -------------------
struct Failure {
  int code;
};
[[noreturn]] void fail(int code, const std::string& message, const char* name = "");
void clear_error();
int capture_bad_alloc();
#define HIP_CHECK(expr)                                                                        \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      ::fpmhip::fail(e_ == hipErrorOutOfMemory ? FPMHIP_OUT_OF_MEMORY : FPMHIP_DEVICE_ERROR,   \
                     std::string(#expr) + ": " + hipGetErrorString(e_));                       \
  } while (0)

// ---- device memory pool ---------------------------------------------------------
// Cache of device blocks for an owner that allocates and frees many buffers
// of recurring sizes on one device (the server's jobs: about 40 buffers each,
// whose hipMalloc / hipFree pairs otherwise cost ~40 us and sit between the
// server's sweep launches). A DeviceBuffer allocated on a thread inside a
// PoolScope takes its block from the pool and gives it back on release.
// Blocks come in size classes: multiples of 1/8 of the request's power of
// two (at least 256 B), so a block is at most 1/8 larger than asked for.
// A block given back is pending and is handed out again only after
// recycle(), which the owner calls where no device work queued before that
// point can still touch it (every stream the blocks are used on
// synchronised): that replaces hipFree's implicit device synchronisation.
// Idle blocks above `max_cached` bytes are freed at recycle(); a failed
// hipMalloc frees every idle block and tries once more.
class DevicePool {
 public:
  explicit DevicePool(size_t max_cached) : max_cached_(max_cached) {}
  ~DevicePool();  // frees every idle and pending block
  DevicePool(const DevicePool&) = delete;
  DevicePool& operator=(const DevicePool&) = delete;
  void* acquire(size_t bytes, size_t* cls);
  void give_back(void* p, size_t cls);
  void recycle();
  uint64_t reuses() const { return reuses_; }
  uint64_t allocations() const { return allocations_; }

 private:
  void free_idle_locked(size_t keep);
  std::mutex mu_;
  std::map<size_t, std::vector<void*>> idle_;  // size class -> blocks
  std::vector<std::pair<void*, size_t>> pending_;
  size_t idle_bytes_ = 0;
  size_t max_cached_;
  std::atomic<uint64_t> reuses_{0}, allocations_{0};
};
// Size class of a request (the block size a pool hands out for it).
size_t pool_size_class(size_t bytes);
// Pool of the DeviceBuffers allocated on this thread (null: hipMalloc).
DevicePool*& thread_pool();
struct PoolScope {
  DevicePool* saved;
  explicit PoolScope(DevicePool* p) : saved(thread_pool()) { thread_pool() = p; }
  ~PoolScope() { thread_pool() = saved; }
  PoolScope(const PoolScope&) = delete;
  PoolScope& operator=(const PoolScope&) = delete;
};

// ---- device buffer ------------------------------------------------------------
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DevicePool* pool = nullptr;  // the block's pool (null: from hipMalloc)
  size_t cls = 0;              // its size class in that pool
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { allocate(n); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes), pool(o.pool), cls(o.cls) {
    o.ptr = nullptr;
    o.bytes = 0;
    o.pool = nullptr;
    o.cls = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      pool = o.pool;
      cls = o.cls;
      o.ptr = nullptr;
      o.bytes = 0;
      o.pool = nullptr;
      o.cls = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { release(); }
  void allocate(size_t n) {
    release();
    if (n == 0) return;
    if (DevicePool* p = thread_pool()) {
      ptr = p->acquire(n, &cls);
      pool = p;
    } else {
      HIP_CHECK(hipMalloc(&ptr, n));
    }
    bytes = n;
  }
  void release() {
    if (ptr) {
      if (pool)
        pool->give_back(ptr, cls);
      else
        (void)hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
    pool = nullptr;
    cls = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr);
  }
};

// ---- FFT plans ------------------------------------------------------------------
constexpr int kMaxStages = 24;
struct LinePlan {
  uint32_t n;
  uint32_t nstages;
  uint32_t radix[kMaxStages];
};

// One direction of a 1-D transform of length n: radices + device unit roots.
struct Fft1 {
  LinePlan plan{};
  DeviceBuffer tw_f64;  // double2[n]
  DeviceBuffer tw_f32;  // float2[n]
};
struct Fft1Pair {
  Fft1 fwd, inv;
};
std::unique_ptr<Fft1Pair> make_fft1(uint32_t n);
// exp(sign * 2 pi i j / n) with quarter turns exact (host, long double)
void unit_root(uint64_t j, uint64_t n, int sign, double* re, double* im);
std::vector<uint32_t> factorize(uint32_t n);

struct Fft2 {
  uint32_t h = 0, w = 0;
  std::unique_ptr<Fft1Pair> rows, cols;  // rows: length w, cols: length h
};
std::unique_ptr<Fft2> make_fft2(uint32_t h, uint32_t w);

// In-place 2-D transform of `count` consecutive h x w arrays on `stream`:
// forward scaled by 1/(h w) (division), inverse unscaled.
void fft2_f64(const Fft2& f, double2* data, uint32_t count, bool inverse, hipStream_t stream);
void fft2_f32(const Fft2& f, float2* data, uint32_t count, bool inverse, hipStream_t stream);

// ---- library objects ------------------------------------------------------------
struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  hipDeviceProp_t props{};
  void* scratch = nullptr;  // device address of this device's forward-stack scratch array (noalloc.hip)
};
// Device address of the allocation-free forward stack's scratch array on the
// current device (noalloc.hip), looked up once per context.
void* noalloc_scratch_address(int device);
// Drops the FFT plans of resident-buffer transforms made for a context
// (resident.hip), when it is destroyed.
void release_resident_plans(const struct ::fpmhip_context* c);

struct Axis {
  uint32_t lower, upper;
  double wl, wu;
};

struct Model {
  Context* ctx = nullptr;
  uint32_t nh = 0, nw = 0, rh = 0, rw = 0;
  uint64_t S = 0, F = 0;
  std::vector<uint64_t> crop_r, crop_c;
  std::vector<double> off_r, off_c;
  std::vector<bool> zero_offset;
  std::vector<Axis> ax_r, ax_c;
  std::vector<double> pupil;  // host copy, interleaved complex
  std::vector<uint8_t> support;
  uint64_t support_count = 0;
  bool has_gain = false;
  std::vector<double> gain;
  std::vector<double> background;  // empty, N or F*N
  bool mux = false;
  std::vector<uint64_t> row_ptr, msrc;
  std::vector<double> mw;
  bool has_k = false;
  std::vector<double> kx, ky;
  DeviceBuffer d_pupil64, d_pupil32, d_support, d_background;
  DeviceBuffer d_src;                          // SrcParam[S]
  DeviceBuffer d_gains;                        // double[F] (ones when absent)
  DeviceBuffer d_mux_w, d_mux_src, d_mux_ptr;  // CSR multiplexing (double, uint32, uint32)
  std::unique_ptr<Fft2> fft_low, fft_high;
  uint64_t N() const { return (uint64_t)nh * nw; }
  uint64_t M() const { return (uint64_t)rh * rw; }
};

struct Measurements {
  Context* ctx = nullptr;
  const Model* model = nullptr;
  uint64_t F = 0, N = 0;
  bool is_u16 = false;
  std::vector<double> weight;
  std::vector<uint8_t> masks;  // host copy (empty = none)
  std::vector<uint16_t> counts_host;
  std::vector<double> frames_host;
  DeviceBuffer d_frames;  // uint16_t[F*N] or double[F*N]
  DeviceBuffer d_masks;   // uint8_t[F*N] or empty
  DeviceBuffer d_weights; // double[F]
  double value(uint64_t f, uint64_t p) const {
    return is_u16 ? (double)counts_host[f * N + p] : frames_host[f * N + p];
  }
};

Axis interpolation_axis(uint64_t start, uint64_t length, uint64_t bound, double offset);

}  // namespace fpmhip

struct fpmhip_context : fpmhip::Context {};
struct fpmhip_model : fpmhip::Model {};
struct fpmhip_measurements : fpmhip::Measurements {};
