// internal.h — shared declarations of libfpmhip (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
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

// ---- device buffer ------------------------------------------------------------
struct DeviceBuffer {
  void* ptr = nullptr;
  size_t bytes = 0;
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) { allocate(n); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr(o.ptr), bytes(o.bytes) {
    o.ptr = nullptr;
    o.bytes = 0;
  }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      release();
      ptr = o.ptr;
      bytes = o.bytes;
      o.ptr = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DeviceBuffer() { release(); }
  void allocate(size_t n) {
    release();
    if (n == 0) return;
    HIP_CHECK(hipMalloc(&ptr, n));
    bytes = n;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
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
