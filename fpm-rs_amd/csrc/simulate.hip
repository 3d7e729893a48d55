// simulate.hip — fpm-rs's detector model and the device part of its
// simulator: CameraModel::validate / measure_frame (src/simulation/camera.rs:
// 145-287) and Simulator::simulate's object spectrum, forward stack, missing
// frames and camera pass (src/simulation/simulator.rs:107-226).
//
// The camera is one elementwise pass over the F x N stack (read and write
// 8 B per pixel, HBM-bound). Its deterministic steps follow measure_frame in
// order and with its operation order (max(value, 0) * sensitivity * photons
// + dark; electrons * gain + offset; clamp to [0, max count]; round half
// away from zero; bad pixels last, clamped and rounded alike), with FMA
// contraction off so that the arithmetic is IEEE-identical to a scalar
// restatement.
//
// Noise comes from a counter-based generator, so that each pixel's draws
// depend on (seed, frame, pixel) only and not on the launch shape: Philox4x64
// with 10 rounds (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011), counter (p, f, block, 0), key (seed, 0); each
// block gives four 64-bit words and each word one uniform double
// (x >> 11) * 2^-53 in [0, 1). Read noise uses the first block's words 0 and
// 1 (Box-Muller, cosine branch); shot noise uses blocks 1, 2, ...: inversion
// by sequential search for means below 10, and above that the transformed
// rejection method with squeeze of Hoermann ("The transformed rejection
// method for generating Poisson random variables", Insurance: Mathematics
// and Economics 12, 1993), two words per trial.
#pragma clang fp contract(off)

#include <algorithm>
#include <limits>

#include "session.h"

namespace fpmhip {

namespace {

constexpr int kCamBlock = 256;

struct CameraParams {
  double ppp, gain, offset, read_noise, dark, max_count;
  int32_t shot, quantize;
  uint64_t seed;
};

// ---- Philox4x64-10 ----
struct U64x4 {
  uint64_t v[4];
};

__device__ inline U64x4 philox4x64_10(U64x4 c, uint64_t k0, uint64_t k1) {
  constexpr uint64_t kM0 = 0xD2E7470EE14C6C93ull, kM1 = 0xCA5A826395121157ull;
  constexpr uint64_t kW0 = 0x9E3779B97F4A7C15ull, kW1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += kW0;
      k1 += kW1;
    }
    const uint64_t hi0 = __umul64hi(kM0, c.v[0]), lo0 = kM0 * c.v[0];
    const uint64_t hi1 = __umul64hi(kM1, c.v[2]), lo1 = kM1 * c.v[2];
    c = U64x4{{hi1 ^ c.v[1] ^ k0, lo1, hi0 ^ c.v[3] ^ k1, lo0}};
  }
  return c;
}

__device__ inline double to_unit(uint64_t x) { return (double)(x >> 11) * 0x1.0p-53; }

// The uniform stream of one pixel: words of consecutive Philox blocks.
struct PixelStream {
  uint64_t p, f, k0;
  uint64_t block = 0;
  int word = 4;
  U64x4 buf{};
  __device__ void seek(uint64_t b) {
    block = b;
    word = 4;
  }
  __device__ double next() {
    if (word == 4) {
      buf = philox4x64_10(U64x4{{p, f, block, 0}}, k0, 0);
      ++block;
      word = 0;
    }
    return to_unit(buf.v[word++]);
  }
};

// Poisson(lambda), lambda > 0
__device__ double poisson(double lambda, PixelStream& rs) {
  rs.seek(1);
  if (lambda < 10.0) {
    // inversion: the smallest k with F(k) > u; 1000 terms exceed the mass
    // of any mean below 10 to within rounding
    const double u = rs.next();
    double p = exp(-lambda), s = p;
    int k = 0;
    while (u >= s && k < 1000) {
      ++k;
      p = p * lambda / (double)k;
      s = s + p;
    }
    return (double)k;
  }
  const double slam = sqrt(lambda), loglam = log(lambda);
  const double b = 0.931 + 2.53 * slam;
  const double a = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  const double vr = 0.9277 - 3.6224 / (b - 2.0);
  // each trial accepts with probability above 0.9; 256 trials bound the loop
  for (int trial = 0; trial < 256; ++trial) {
    const double U = rs.next() - 0.5;
    const double V = rs.next();
    const double us = 0.5 - fabs(U);
    const double k = floor((2.0 * a / us + b) * U + lambda + 0.43);
    if (us >= 0.07 && V <= vr) return k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -lambda + k * loglam - lgamma(k + 1.0)) return k;
  }
  return floor(lambda);
}

// Normal(0, 1) from the first block's words 0 and 1
__device__ double standard_normal(PixelStream& rs) {
  rs.seek(0);
  const double u1 = 1.0 - rs.next();  // (0, 1]
  const double u2 = rs.next();
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

__host__ __device__ inline double clamp_count(double x, double hi) { return x < 0.0 ? 0.0 : (x > hi ? hi : x); }

// One pass over frames [0, F) of N pixels, in place. missing (uint8[F] or
// null): frames replaced by the optical background first; bad (uint8[N] or
// null): pixels set to bad_value after every other effect.
__global__ void __launch_bounds__(kCamBlock)
    k_camera(double* data, uint64_t N, uint64_t F, CameraParams cp, const double* sensitivity, const uint8_t* missing,
             const double* bg, uint64_t bg_len, const uint8_t* bad, double bad_value, int32_t apply_camera) {
  const uint64_t total = N * F;
  for (uint64_t i = (uint64_t)blockIdx.x * kCamBlock + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * kCamBlock) {
    const uint64_t f = i / N, p = i - f * N;
    double value = data[i];
    if (missing && missing[f]) value = bg_len == 0 ? 0.0 : bg[bg_len == N ? p : i];
    if (apply_camera) {
      if (bad && bad[p]) {
        value = bad_value;
      } else {
        PixelStream rs{p, f, cp.seed};
        const double s = sensitivity ? sensitivity[p] : 1.0;
        const double expected = fmax(value, 0.0) * s * cp.ppp + cp.dark;
        double electrons = (cp.shot && expected > 0.0) ? poisson(expected, rs) : expected;
        if (cp.read_noise > 0.0) electrons = electrons + (0.0 + cp.read_noise * standard_normal(rs));
        double counts = electrons * cp.gain + cp.offset;
        counts = clamp_count(counts, cp.max_count);
        if (cp.quantize) counts = round(counts);
        value = counts;
      }
    }
    data[i] = value;
  }
}

uint32_t camera_grid(uint64_t total) {
  const uint64_t blocks = (total + kCamBlock - 1) / kCamBlock;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(blocks, 256u * 64u));
}

// CameraModel::maximum_count (camera.rs:320-327)
double maximum_count(const fpmhip_camera& c) {
  const double digital = c.has_bit_depth ? std::ldexp(1.0, c.bit_depth) - 1.0 : std::numeric_limits<double>::infinity();
  const double sat = c.has_saturation ? c.saturation_counts : std::numeric_limits<double>::infinity();
  return std::min(sat, digital);
}

// CameraModel::validate (camera.rs:145-223) and validate_for_frame (:226-245)
void validate_camera(const fpmhip_camera& c, uint64_t frame_len) {
  const char* positive = "must be finite and positive";
  if (!std::isfinite(c.photons_per_pixel) || c.photons_per_pixel <= 0.0)
    fail(FPMHIP_INVALID_PARAMETER, positive, "photons_per_pixel");
  if (!std::isfinite(c.gain_counts_per_electron) || c.gain_counts_per_electron <= 0.0)
    fail(FPMHIP_INVALID_PARAMETER, positive, "gain_counts_per_electron");
  if (!std::isfinite(c.offset_counts) || !std::isfinite(c.read_noise_electrons) || c.read_noise_electrons < 0.0 ||
      !std::isfinite(c.dark_current_electrons) || c.dark_current_electrons < 0.0)
    fail(FPMHIP_INVALID_PARAMETER, "offset must be finite; read noise and dark current must be non-negative",
         "camera noise/offset");
  if (c.has_bit_depth && (c.bit_depth <= 0 || c.bit_depth > 32))
    fail(FPMHIP_INVALID_PARAMETER, "must be between 1 and 32", "bit_depth");
  if (c.has_saturation && (!std::isfinite(c.saturation_counts) || c.saturation_counts <= 0.0))
    fail(FPMHIP_INVALID_PARAMETER, positive, "saturation_counts");
  if (c.has_bad_pixel_value && (!std::isfinite(c.bad_pixel_value_counts) || c.bad_pixel_value_counts < 0.0))
    fail(FPMHIP_INVALID_PARAMETER, "must be finite and non-negative", "bad_pixel_value_counts");
  if (c.bad_pixel_count && !c.bad_pixels) fail(FPMHIP_INVALID_PARAMETER, "null bad-pixel list", "bad_pixels");
  if (c.bad_pixel_count && !c.has_bad_pixel_value && !std::isfinite(maximum_count(c)))
    fail(FPMHIP_INVALID_PARAMETER, "a finite bad-pixel value or camera maximum is required", "bad_pixels");
  if (c.pixel_sensitivity && frame_len) {
    for (uint64_t p = 0; p < frame_len; ++p)
      if (!std::isfinite(c.pixel_sensitivity[p]) || c.pixel_sensitivity[p] < 0.0)
        fail(FPMHIP_INVALID_PARAMETER, "values must be finite and non-negative", "pixel_sensitivity");
  }
  std::vector<uint64_t> sorted(c.bad_pixels, c.bad_pixels + c.bad_pixel_count);
  std::sort(sorted.begin(), sorted.end());
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    fail(FPMHIP_INVALID_PARAMETER, "indices must be unique", "bad_pixels");
  if (frame_len) {
    // the sensitivity's length is the caller's contract: the array is read
    // as frame_len values, and a NULL array means none
    if (!sorted.empty() && sorted.back() >= frame_len)
      fail(FPMHIP_INVALID_PARAMETER, "indices must be below the frame size " + std::to_string(frame_len),
           "bad_pixels");
  }
}

// The camera pass over a device stack `data` [F*N], with optional missing
// frames replaced by the model's background first.
void camera_pass(const Context& ctx, double* data, uint64_t N, uint64_t F, const fpmhip_camera* cam, uint64_t seed,
                 const uint8_t* d_missing, const Model* bg_model) {
  CameraParams cp{};
  DeviceBuffer d_sens, d_bad;
  double bad_value = 0.0;
  if (cam) {
    cp.ppp = cam->photons_per_pixel;
    cp.gain = cam->gain_counts_per_electron;
    cp.offset = cam->offset_counts;
    cp.read_noise = cam->read_noise_electrons;
    cp.dark = cam->dark_current_electrons;
    cp.max_count = maximum_count(*cam);
    cp.shot = cam->shot_noise != 0;
    cp.quantize = cam->quantize != 0;
    cp.seed = seed;
    if (cam->pixel_sensitivity) {
      d_sens.allocate(N * sizeof(double));
      HIP_CHECK(hipMemcpyAsync(d_sens.ptr, cam->pixel_sensitivity, N * sizeof(double), hipMemcpyHostToDevice,
                               ctx.stream));
    }
    if (cam->bad_pixel_count) {
      std::vector<uint8_t> mask(N, 0);
      for (uint64_t i = 0; i < cam->bad_pixel_count; ++i) mask[cam->bad_pixels[i]] = 1;
      d_bad.allocate(N);
      HIP_CHECK(hipMemcpyAsync(d_bad.ptr, mask.data(), N, hipMemcpyHostToDevice, ctx.stream));
      // camera.rs:277-286: the value (or the camera maximum) clamped, then rounded
      bad_value = clamp_count(cam->has_bad_pixel_value ? cam->bad_pixel_value_counts : cp.max_count, cp.max_count);
      if (cp.quantize) bad_value = std::round(bad_value);
    }
    // the mask upload reads a host vector: finish it before the vector goes
    HIP_CHECK(hipStreamSynchronize(ctx.stream));
  }
  const uint64_t total = N * F;
  if (total == 0) return;
  hipLaunchKernelGGL(k_camera, dim3(camera_grid(total)), dim3(kCamBlock), 0, ctx.stream, data, N, F, cp,
                     d_sens.as<double>(), d_missing, bg_model ? bg_model->d_background.as<double>() : nullptr,
                     bg_model ? (uint64_t)bg_model->background.size() : 0, d_bad.as<uint8_t>(), bad_value,
                     (int32_t)(cam != nullptr));
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(ctx.stream));
}

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

}  // namespace

}  // namespace fpmhip

using namespace fpmhip;

extern "C" {

int fpmhip_camera_validate(const fpmhip_camera* camera, uint64_t frame_len) {
  return guarded([&] {
    if (!camera) fail(FPMHIP_INVALID_PARAMETER, "null camera", "camera");
    validate_camera(*camera, frame_len);
  });
}

int fpmhip_camera_measure(fpmhip_context* c, const fpmhip_camera* camera, double* frames, uint64_t frame_count,
                          uint64_t frame_len, uint64_t seed) {
  return guarded([&] {
    if (!c || !camera) fail(FPMHIP_INVALID_PARAMETER, "null argument", "camera");
    if (frame_len == 0) fail(FPMHIP_INVALID_SHAPE, "camera frames must have at least one pixel");
    validate_camera(*camera, frame_len);
    if (frame_count && !frames) fail(FPMHIP_INVALID_PARAMETER, "null frames", "frames");
    if (frame_count > std::numeric_limits<uint64_t>::max() / frame_len / sizeof(double))
      fail(FPMHIP_INVALID_SHAPE, "camera stack size overflows");
    HIP_CHECK(hipSetDevice(c->device));
    const uint64_t total = frame_count * frame_len;
    if (total == 0) return;
    DeviceBuffer d(total * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(d.ptr, frames, d.bytes, hipMemcpyHostToDevice, c->stream));
    camera_pass(*c, d.as<double>(), frame_len, frame_count, camera, seed, nullptr, nullptr);
    HIP_CHECK(hipMemcpyAsync(frames, d.ptr, d.bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int fpmhip_simulate(fpmhip_context* c, const fpmhip_model* m, const double* object, const fpmhip_camera* camera,
                    const uint64_t* missing_frames, uint64_t missing_count, uint64_t seed, double* out) {
  return guarded([&] {
    if (!c || !m || !object || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "simulate");
    if (m->ctx->device != c->device)
      fail(FPMHIP_INVALID_PARAMETER, "model and context are on different devices", "simulate");
    const uint64_t N = m->N(), M = m->M(), F = m->F;
    if (camera) validate_camera(*camera, N);
    // apply_illumination_acquisition_errors' missing-frame check (simulator.rs:242-251)
    if (missing_count && !missing_frames) fail(FPMHIP_INVALID_PARAMETER, "null missing-frame list", "missing_frames");
    std::vector<uint8_t> missing(F, 0);
    for (uint64_t i = 0; i < missing_count; ++i) {
      const uint64_t f = missing_frames[i];
      if (f >= F || missing[f])
        fail(FPMHIP_INVALID_PARAMETER, "must contain unique frame indices below " + std::to_string(F),
             "missing_frames");
      missing[f] = 1;
    }
    for (uint64_t f = 0; f < F; ++f) {
      const double g = m->has_gain ? m->gain[f] : 1.0;
      if (!std::isfinite(g) || g <= 0.0) fail(FPMHIP_INVALID_MODEL, "frame has invalid gain");
    }
    HIP_CHECK(hipSetDevice(c->device));
    DeviceBuffer obj(M * sizeof(double2)), spec(M * sizeof(double2)), fields(m->S * N * sizeof(double2)),
        stack(F * N * sizeof(double)), d_missing;
    HIP_CHECK(hipMemcpyAsync(obj.ptr, object, obj.bytes, hipMemcpyHostToDevice, c->stream));
    // object_spectrum (simulator.rs:209-226): forward FFT / R, then fftshift
    fft2_f64(*m->fft_high, obj.as<double2>(), 1, false, c->stream);
    launch_shift(obj.as<double2>(), spec.as<double2>(), m->rh, m->rw, m->rh / 2, m->rw / 2, c->stream);
    launch_forward_stack(*m, spec.as<double2>(), m->d_pupil64.as<double2>(), fields.as<double2>(),
                         stack.as<double>(), c->stream);
    if (missing_count) {
      d_missing.allocate(F);
      HIP_CHECK(hipMemcpyAsync(d_missing.ptr, missing.data(), F, hipMemcpyHostToDevice, c->stream));
    }
    if (camera || missing_count)
      camera_pass(*c, stack.as<double>(), N, F, camera, seed, d_missing.as<uint8_t>(), m);
    HIP_CHECK(hipMemcpyAsync(out, stack.ptr, stack.bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
