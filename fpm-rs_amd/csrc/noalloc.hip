// noalloc.hip — the forward stack for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid allocation): fpmhip_forward_stack_noalloc takes its inputs and its
// output in the caller's device memory and keeps the per-source fields in a
// file-scope device array of fixed size (`scratch`, kScratchBytes) instead
// of a hipMalloc'd buffer. The array exists once per GPU the library runs on
// and lives until the process ends; its device address is looked up once
// per device when a context is created (a host reference to the array's
// name is not a device address). Every call on a GPU shares that GPU's
// array, so calls that overlap on one GPU race; calls on one context's
// stream are ordered.
//
// The arithmetic is fpmhip_forward_stack's (the same kernels over the same
// per-source values), in chunks of frames whose fields fit the array, so the
// output is identical to it.
#include "session.h"

namespace fpmhip {

namespace {

constexpr size_t kScratchBytes = 16u << 20;  // 16 MiB: 64 sources of 128^2, 16 of 256^2, 4 of 512^2
constexpr size_t kScratchElems = kScratchBytes / sizeof(double2);

__device__ __align__(256) double2 scratch[kScratchElems];

}  // namespace

void* noalloc_scratch_address(int device) {
  // called with `device` current (context creation)
  (void)device;
  void* ptr = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) fail(FPMHIP_DEVICE_ERROR, "the forward-stack scratch array has no device address");
  return ptr;
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// Frames [f0, f1) of the forward stack of a model without multiplexing, the
// fields of their sources in `fields` (at least (f1 - f0) N complex128).
void launch_forward_stack_frames(const Model& m, const double2* spec, const double2* pupil, double2* fields,
                                 double* out, uint64_t f0, uint64_t f1, hipStream_t stream);

}  // namespace fpmhip

using namespace fpmhip;

extern "C" int fpmhip_forward_stack_noalloc(fpmhip_context* c, const fpmhip_model* m, const double* spectrum,
                                            const double* pupil, double* out) {
  clear_error();
  try {
    if (!c || !m || !spectrum || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "forward_stack");
    if (!c->scratch) fail(FPMHIP_DEVICE_ERROR, "context has no forward-stack scratch array");
    if (m->ctx->device != c->device) fail(FPMHIP_INVALID_PARAMETER, "model and context are on different devices",
                                          "forward_stack");
    if (m->mux)
      fail(FPMHIP_UNSUPPORTED, "the allocation-free forward stack takes models without multiplexing");
    for (uint64_t f = 0; f < m->F; ++f) {
      const double g = m->has_gain ? m->gain[f] : 1.0;
      if (!std::isfinite(g) || g <= 0.0) fail(FPMHIP_INVALID_MODEL, "frame has invalid gain");
    }
    const uint64_t N = m->N();
    const uint64_t per_chunk = kScratchElems / N;  // frames (= sources) whose fields fit the array
    if (per_chunk == 0)
      fail(FPMHIP_INVALID_SHAPE, "image of " + std::to_string(N) + " pixels exceeds the allocation-free forward "
                                 "stack's scratch of " + std::to_string(kScratchBytes) + " bytes (" +
                                 std::to_string(kScratchElems) + " complex128 values)");
    HIP_CHECK(hipSetDevice(c->device));
    const double2* pp = pupil ? reinterpret_cast<const double2*>(pupil) : m->d_pupil64.as<double2>();
    double2* fields = static_cast<double2*>(c->scratch);
    for (uint64_t f0 = 0; f0 < m->F; f0 += per_chunk)
      launch_forward_stack_frames(*m, reinterpret_cast<const double2*>(spectrum), pp, fields, out, f0,
                                  std::min<uint64_t>(m->F, f0 + per_chunk), c->stream);
    return FPMHIP_OK;
  } catch (const Failure& f) {
    return f.code;
  } catch (const std::bad_alloc&) {
    return capture_bad_alloc();
  }
}
