// resident.hip — the ResidentBackend surface of fpm-rs's backend seam
// (src/backend/traits.rs:77-106) on the device: typed complex128 / f64
// buffers in HBM, length-checked uploads and downloads, and an in-place 2-D
// FFT of a resident buffer with fpm-rs's normalisation (forward / (h w),
// inverse unscaled; src/backend/cpu.rs:162-172). Errors follow CpuBackend's
// (cpu.rs:238-285): a buffer of the other kind, or from another context, is
// InvalidParameter named "complex buffer" / "real buffer"; a transfer whose
// length differs from the buffer's is InvalidShape.
#include <map>
#include <mutex>
#include <utility>

#include "session.h"

struct fpmhip_buffer {
  fpmhip_context* ctx = nullptr;
  int32_t kind = 0;  // fpmhip_buffer_kind
  uint64_t len = 0;  // elements (complex values or reals)
  fpmhip::DeviceBuffer mem;
};

using namespace fpmhip;

namespace {

template <class F>
int guarded(F&& body) {
  clear_error();
  try {
    body();
    return FPMHIP_OK;
  } catch (const Failure& f) {
    return f.code;
  } catch (const std::bad_alloc&) {
    return capture_bad_alloc();
  }
}

size_t element_bytes(int32_t kind) { return kind == FPMHIP_BUFFER_COMPLEX ? sizeof(double2) : sizeof(double); }

// cpu_complex / cpu_real (cpu.rs:238-275): the buffer must be this
// context's and of the expected kind
void expect(const fpmhip_context* c, const fpmhip_buffer* b, int32_t kind) {
  const char* name = kind == FPMHIP_BUFFER_COMPLEX ? "complex buffer" : "real buffer";
  if (!b || b->ctx != c || b->kind != kind) fail(FPMHIP_INVALID_PARAMETER, "buffer was not allocated by this backend", name);
}

// validate_transfer_length (cpu.rs:278-285)
void expect_length(uint64_t destination, uint64_t source) {
  if (destination != source)
    fail(FPMHIP_INVALID_SHAPE, "buffer transfer length " + std::to_string(source) +
                                   " does not match destination length " + std::to_string(destination));
}

// 2-D FFT plans of resident transforms, per context and shape
std::map<std::pair<const fpmhip_context*, std::pair<uint64_t, uint64_t>>, std::unique_ptr<Fft2>> g_plans;
std::mutex g_plans_mu;

const Fft2& plan_for(const fpmhip_context* c, uint64_t h, uint64_t w) {
  std::lock_guard<std::mutex> lock(g_plans_mu);
  auto& p = g_plans[{c, {h, w}}];
  if (!p) p = make_fft2((uint32_t)h, (uint32_t)w);
  return *p;
}

}  // namespace

namespace fpmhip {
void release_resident_plans(const fpmhip_context* c) {
  std::lock_guard<std::mutex> lock(g_plans_mu);
  for (auto it = g_plans.begin(); it != g_plans.end();)
    it = it->first.first == c ? g_plans.erase(it) : std::next(it);
}
}  // namespace fpmhip

extern "C" {

int fpmhip_context_capabilities(const fpmhip_context* c, fpmhip_capabilities* out) {
  return guarded([&] {
    if (!c || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "capabilities");
    out->preferred_memory = FPMHIP_MEMORY_DEVICE;
    out->resident_buffers = 1;
  });
}

int fpmhip_buffer_allocate(fpmhip_context* c, int32_t kind, uint64_t len, fpmhip_buffer** out) {
  return guarded([&] {
    if (!c || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "buffer");
    if (kind != FPMHIP_BUFFER_COMPLEX && kind != FPMHIP_BUFFER_REAL)
      fail(FPMHIP_INVALID_PARAMETER, "unknown buffer kind", "kind");
    HIP_CHECK(hipSetDevice(c->device));
    auto b = std::make_unique<fpmhip_buffer>();
    b->ctx = c;
    b->kind = kind;
    b->len = len;
    b->mem.allocate(len * element_bytes(kind));
    if (len) HIP_CHECK(hipMemsetAsync(b->mem.ptr, 0, b->mem.bytes, c->stream));  // zeroed, as CpuBackend's
    HIP_CHECK(hipStreamSynchronize(c->stream));
    *out = b.release();
  });
}

int fpmhip_buffer_free(fpmhip_buffer* b) {
  return guarded([&] {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    delete b;
  });
}

int fpmhip_buffer_info(const fpmhip_buffer* b, int32_t* kind, uint64_t* len, int32_t* location, void** device) {
  return guarded([&] {
    if (!b) fail(FPMHIP_INVALID_PARAMETER, "null buffer", "buffer");
    if (kind) *kind = b->kind;
    if (len) *len = b->len;
    if (location) *location = FPMHIP_MEMORY_DEVICE;
    if (device) *device = b->mem.ptr;
  });
}

int fpmhip_buffer_upload(fpmhip_context* c, fpmhip_buffer* b, int32_t kind, const double* src, uint64_t len) {
  return guarded([&] {
    if (!c) fail(FPMHIP_INVALID_PARAMETER, "null context", "context");
    expect(c, b, kind);
    expect_length(b->len, len);
    if (len && !src) fail(FPMHIP_INVALID_PARAMETER, "null source", "source");
    HIP_CHECK(hipSetDevice(c->device));
    if (len)
      HIP_CHECK(hipMemcpyAsync(b->mem.ptr, src, len * element_bytes(kind), hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int fpmhip_buffer_download(fpmhip_context* c, const fpmhip_buffer* b, int32_t kind, double* dst, uint64_t len) {
  return guarded([&] {
    if (!c) fail(FPMHIP_INVALID_PARAMETER, "null context", "context");
    expect(c, b, kind);
    expect_length(len, b->len);
    if (len && !dst) fail(FPMHIP_INVALID_PARAMETER, "null destination", "destination");
    HIP_CHECK(hipSetDevice(c->device));
    if (len)
      HIP_CHECK(hipMemcpyAsync(dst, b->mem.ptr, len * element_bytes(kind), hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int fpmhip_fft2_resident(fpmhip_context* c, fpmhip_buffer* b, uint64_t h, uint64_t w, int32_t direction) {
  return guarded([&] {
    if (!c) fail(FPMHIP_INVALID_PARAMETER, "null context", "context");
    expect(c, b, FPMHIP_BUFFER_COMPLEX);
    if (h == 0 || w == 0) fail(FPMHIP_INVALID_SHAPE, "FFT dimensions must be non-zero");
    if (h > 0xFFFFFFFFull || w > 0xFFFFFFFFull || h * w != b->len)
      fail(FPMHIP_INVALID_SHAPE, "FFT shape " + std::to_string(h) + "x" + std::to_string(w) +
                                     " does not match buffer length " + std::to_string(b->len));
    HIP_CHECK(hipSetDevice(c->device));
    fft2_f64(plan_for(c, h, w), b->mem.as<double2>(), 1, direction == FPMHIP_INVERSE, c->stream);
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
