// diagnostics.hip — build_diagnostics (src/reconstruction/runner.rs:298-433)
// and the SnrWeighted frame scores (src/reconstruction/schedule.rs:106-128)
// on the device.
//
// Diagnostics of a state are one forward prediction of every frame (the
// forward-stack kernels, with the state's frame gains and background) and
// one reduction per frame: a 256-thread workgroup per frame walks the
// frame's pixels once, reading the prediction and the resident measurement
// (and mask), and keeps the sums of IntensityStats (single.rs:25-61), of the
// comparison accumulator (compare.rs:187-226) and of the runner's amplitude
// objective (runner.rs:361-373), reduced in a fixed order (deterministic for
// a given image size). The host finishes each frame's record with the
// reference's closing formulas. HBM traffic: the S N complex128 fields
// written and read twice by the prediction, then 8 B of prediction and 2 or
// 8 B of measurement per pixel (plus 8 B of residual image when requested).
#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>

#include "session.h"

namespace fpmhip {

namespace {

constexpr int kDiagBlock = 256;
constexpr int kRec = 16;  // doubles per frame record

// record fields
enum : int {
  kLoss = 0, kValid, kRefSum, kEstSum, kL1, kR2, kRSum, kRMax, kRef2,
  kRawSum, kRawSq, kRawMin, kRawMax, kRawZero, kRawBad, kPairBad
};

__device__ inline double frame_value(const void* frames, int is_u16, uint64_t i) {
  return is_u16 ? (double)static_cast<const uint16_t*>(frames)[i] : static_cast<const double*>(frames)[i];
}

__device__ inline bool finite(double x) { return isfinite(x); }

__global__ void __launch_bounds__(kDiagBlock)
    k_frame_diagnostics(const double* __restrict__ pred, const void* __restrict__ frames, int is_u16,
                        const uint8_t* __restrict__ masks, uint64_t N, double* __restrict__ residual,
                        double* __restrict__ records) {
  __shared__ double sh[kRec][kDiagBlock];
  const uint64_t f = blockIdx.x;
  const int t = threadIdx.x;
  double a[kRec];
#pragma unroll
  for (int k = 0; k < kRec; ++k) a[k] = 0.0;
  a[kRawMin] = INFINITY;
  a[kRawMax] = -INFINITY;
  for (uint64_t p = t; p < N; p += kDiagBlock) {
    const uint64_t i = f * N + p;
    const double y = frame_value(frames, is_u16, i);
    const double e = pred[i];
    const bool valid = !masks || masks[i] != 0;
    // IntensityStats of the measured frame, every pixel
    a[kRawSum] += y;
    a[kRawSq] += y * y;
    a[kRawMin] = fmin(a[kRawMin], y);
    a[kRawMax] = fmax(a[kRawMax], y);
    a[kRawZero] += (y == 0.0) ? 1.0 : 0.0;
    a[kRawBad] += finite(y) ? 0.0 : 1.0;
    if (valid) {
      const double r = sqrt(fmax(e, 0.0)) - sqrt(fmax(y, 0.0));
      a[kLoss] += r * r;
      a[kValid] += 1.0;
      const double res = e - y;
      a[kRefSum] += y;
      a[kEstSum] += e;
      a[kL1] += fabs(res);
      a[kR2] += res * res;
      a[kRSum] += res;
      a[kRMax] = fmax(a[kRMax], fabs(res));
      a[kRef2] += y * y;
      a[kPairBad] += (finite(y) && finite(e)) ? 0.0 : 1.0;
    }
    if (residual) residual[i] = valid ? e - y : 0.0;
  }
#pragma unroll
  for (int k = 0; k < kRec; ++k) sh[k][t] = a[k];
  __syncthreads();
  for (int w = kDiagBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
#pragma unroll
      for (int k = 0; k < kRec; ++k) {
        const double x = sh[k][t], y = sh[k][t + w];
        sh[k][t] = k == kRMax || k == kRawMax ? fmax(x, y) : (k == kRawMin ? fmin(x, y) : x + y);
      }
    }
    __syncthreads();
  }
  if (t < kRec) records[f * kRec + t] = sh[t][0];
}

// per frame: sum over unmasked pixels of max(y - background, 0), and their count
__global__ void __launch_bounds__(kDiagBlock)
    k_snr_sums(const void* __restrict__ frames, int is_u16, const uint8_t* __restrict__ masks,
               const double* __restrict__ bg, uint64_t bg_len, uint64_t N, double* __restrict__ out) {
  __shared__ double s_sum[kDiagBlock], s_cnt[kDiagBlock];
  const uint64_t f = blockIdx.x;
  const int t = threadIdx.x;
  double sum = 0.0, cnt = 0.0;
  for (uint64_t p = t; p < N; p += kDiagBlock) {
    const uint64_t i = f * N + p;
    if (masks && masks[i] == 0) continue;
    const double b = bg_len == 0 ? 0.0 : bg[bg_len == N ? p : i];
    sum += fmax(frame_value(frames, is_u16, i) - b, 0.0);
    cnt += 1.0;
  }
  s_sum[t] = sum;
  s_cnt[t] = cnt;
  __syncthreads();
  for (int w = kDiagBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      s_sum[t] += s_sum[t + w];
      s_cnt[t] += s_cnt[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[2 * f] = s_sum[0];
    out[2 * f + 1] = s_cnt[0];
  }
}

// complex::amplitude / complex::phase (norm = hypot, arg = atan2)
__global__ void k_amplitude_phase(const double2* __restrict__ v, uint64_t n, double* __restrict__ amp,
                                  double* __restrict__ ph) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 z = v[i];
  if (amp) amp[i] = hypot(z.x, z.y);
  if (ph) ph[i] = atan2(z.y, z.x);
}

uint32_t blocks(uint64_t n) { return (uint32_t)((n + kDiagBlock - 1) / kDiagBlock); }

void amplitude_phase(const Context& c, const double2* v, uint64_t n, double* amp_host, double* ph_host) {
  if (!amp_host && !ph_host) return;
  DeviceBuffer amp(amp_host ? n * sizeof(double) : 0), ph(ph_host ? n * sizeof(double) : 0);
  hipLaunchKernelGGL(k_amplitude_phase, dim3(blocks(n)), dim3(kDiagBlock), 0, c.stream, v, n, amp.as<double>(),
                     ph.as<double>());
  HIP_CHECK(hipGetLastError());
  if (amp_host) HIP_CHECK(hipMemcpyAsync(amp_host, amp.ptr, amp.bytes, hipMemcpyDeviceToHost, c.stream));
  if (ph_host) HIP_CHECK(hipMemcpyAsync(ph_host, ph.ptr, ph.bytes, hipMemcpyDeviceToHost, c.stream));
  HIP_CHECK(hipStreamSynchronize(c.stream));
}

// The state's pupil in FP64 on the device: its FP64 master, the complex128
// pupil, or the complex64 pupil widened into `tmp`.
const double2* pupil64_device(const fpmhip_session& s, DeviceBuffer& tmp) {
  if (s.b.pupil64.ptr) return s.b.pupil64.as<double2>();
  if (s.precision == FPMHIP_PLAIN_C128) return s.b.pupil.as<double2>();
  const uint64_t N = s.model->N();
  tmp.allocate(N * sizeof(double2));
  launch_pupil_to_f64(s.b.pupil.ptr, true, tmp.as<double2>(), N, s.ctx->stream);
  return tmp.as<double2>();
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

std::vector<double> snr_scores(const fpmhip_session& s) {
  const Model& m = *s.model;
  const Measurements& me = *s.meas;
  const uint64_t F = m.F, N = m.N();
  std::vector<double> score(F, -std::numeric_limits<double>::infinity());
  if (F == 0) return score;
  DeviceBuffer sums(F * 2 * sizeof(double));
  hipLaunchKernelGGL(k_snr_sums, dim3((uint32_t)F), dim3(kDiagBlock), 0, s.ctx->stream, me.d_frames.ptr,
                     (int)me.is_u16, me.d_masks.as<uint8_t>(), m.d_background.as<double>(),
                     (uint64_t)m.background.size(), N, sums.as<double>());
  HIP_CHECK(hipGetLastError());
  std::vector<double> h(F * 2);
  HIP_CHECK(hipMemcpyAsync(h.data(), sums.ptr, sums.bytes, hipMemcpyDeviceToHost, s.ctx->stream));
  HIP_CHECK(hipStreamSynchronize(s.ctx->stream));
  for (uint64_t f = 0; f < F; ++f) {
    // frame_snr_score (schedule.rs:106-128): zero weight or no valid pixel last
    if (me.weight[f] == 0.0 || h[2 * f + 1] == 0.0) continue;
    score[f] = me.weight[f] * std::sqrt(h[2 * f] / h[2 * f + 1]);
  }
  return score;
}

}  // namespace fpmhip

using namespace fpmhip;

extern "C" int fpmhip_session_diagnostics(fpmhip_session* s, const fpmhip_diagnostics* out) {
  return guarded([&] {
    if (!s || !out) fail(FPMHIP_INVALID_PARAMETER, "null argument", "diagnostics");
    if (!s->initialized) fail(FPMHIP_INVALID_MODEL, "session state is not initialized");
    HIP_CHECK(hipSetDevice(s->ctx->device));
    const Model& m = *s->model;
    const Measurements& me = *s->meas;
    const uint64_t F = m.F, N = m.N(), M = m.M();
    hipStream_t st = s->ctx->stream;
    DeviceBuffer ptmp;
    const double2* pupil = pupil64_device(*s, ptmp);
    const bool frames = out->per_frame_objective || out->raw_frame_statistics || out->frame_summaries ||
                        out->residual_images;
    if (frames && F) {
      DeviceBuffer fields(m.S * N * sizeof(double2)), pred(F * N * sizeof(double)), records(F * kRec * sizeof(double)),
          residual(out->residual_images ? F * N * sizeof(double) : 0);
      // the prediction of model_with_state_calibration (runner.rs:436-455)
      launch_forward_stack_with(m, s->b.spec.as<double2>(), pupil, fields.as<double2>(), pred.as<double>(),
                                s->b.gains.as<double>(), s->b.bg.as<double>(), s->bg_len, st);
      hipLaunchKernelGGL(k_frame_diagnostics, dim3((uint32_t)F), dim3(kDiagBlock), 0, st, pred.as<double>(),
                         me.d_frames.ptr, (int)me.is_u16, me.d_masks.as<uint8_t>(), N, residual.as<double>(),
                         records.as<double>());
      HIP_CHECK(hipGetLastError());
      std::vector<double> r(F * kRec);
      HIP_CHECK(hipMemcpyAsync(r.data(), records.ptr, records.bytes, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      // the reference's checks, in its order: raw statistics, then per frame the comparison
      if (out->raw_frame_statistics)
        for (uint64_t f = 0; f < F; ++f)
          if (r[f * kRec + kRawBad] != 0.0)
            fail(FPMHIP_NUMERICAL, "intensity statistics contain a non-finite value");
      if (out->frame_summaries)
        for (uint64_t f = 0; f < F; ++f) {
          if (r[f * kRec + kPairBad] != 0.0)
            fail(FPMHIP_NUMERICAL, "intensity comparison contains a non-finite value");
          if (r[f * kRec + kValid] == 0.0) fail(FPMHIP_INVALID_PARAMETER, "must select at least one pixel", "valid_mask");
        }
      for (uint64_t f = 0; f < F; ++f) {
        const double* a = &r[f * kRec];
        if (out->per_frame_objective) out->per_frame_objective[f] = a[kValid] == 0.0 ? 0.0 : a[kLoss] / a[kValid];
        if (out->raw_frame_statistics) {
          const double n = (double)N, mean = a[kRawSum] / n;
          fpmhip_intensity_stats& o = out->raw_frame_statistics[f];
          o.mean = mean;
          o.std = std::sqrt(std::max(a[kRawSq] / n - mean * mean, 0.0));
          o.min = a[kRawMin];
          o.max = a[kRawMax];
          o.sum = a[kRawSum];
          o.saturated_pixels = 0;  // the runner passes no saturation value
          o.zero_pixels = (uint64_t)a[kRawZero];
        }
        if (out->frame_summaries) {
          const double n = a[kValid], mean = a[kRSum] / n;
          fpmhip_frame_comparison& o = out->frame_summaries[f];
          o.reference_sum = a[kRefSum];
          o.estimate_sum = a[kEstSum];
          o.residual_l1 = a[kL1];
          o.residual_l2 = std::sqrt(a[kR2]);
          o.residual_mean = mean;
          o.residual_std = std::sqrt(std::max(a[kR2] / n - mean * mean, 0.0));
          o.residual_max_abs = a[kRMax];
          o.normalized_l2 = std::sqrt(a[kR2]) / (std::sqrt(a[kRef2]) + std::numeric_limits<double>::epsilon());
        }
      }
      if (out->residual_images) {
        HIP_CHECK(hipMemcpyAsync(out->residual_images, residual.ptr, residual.bytes, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
      }
    }
    if (out->object_amplitude || out->object_phase) {
      launch_object(*s, s->b.work_m.as<double2>());
      amplitude_phase(*s->ctx, s->b.work_m.as<double2>(), M, out->object_amplitude, out->object_phase);
    }
    if (out->pupil_amplitude || out->pupil_phase) amplitude_phase(*s->ctx, pupil, N, out->pupil_amplitude, out->pupil_phase);
  });
}
