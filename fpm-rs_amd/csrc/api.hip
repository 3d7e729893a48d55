// api.hip — device-facing C ABI of libfpmhip: contexts, the Backend::fft2
// primitive, model and measurement upload, the forward stack and the
// reconstruction session (ReconstructionState + Runner loop).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>

#include "fused.h"
#include "gauge.h"
#include "session.h"

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

void require(bool ok, const char* name, const char* reason) {
  if (!ok) fail(FPMHIP_INVALID_PARAMETER, reason, name);
}

template <class T>
void upload(DeviceBuffer& b, const T* src, size_t n, hipStream_t st) {
  b.allocate(n * sizeof(T));
  if (n) HIP_CHECK(hipMemcpyAsync(b.ptr, src, n * sizeof(T), hipMemcpyHostToDevice, st));
}

void set_device(const Context* c) { HIP_CHECK(hipSetDevice(c->device)); }

inline bool off_zero(double r, double c) { return std::fabs(r) <= 1e-12 && std::fabs(c) <= 1e-12; }

// ---- session helpers ----------------------------------------------------------------

void download_pupil64(fpmhip_session& s, double* out) {
  const Model& m = *s.model;
  const uint64_t N = m.N();
  if (s.b.pupil64.ptr) {
    HIP_CHECK(hipMemcpyAsync(out, s.b.pupil64.ptr, N * sizeof(double2), hipMemcpyDeviceToHost, s.ctx->stream));
  } else if (s.precision == FPMHIP_PLAIN_C128) {
    HIP_CHECK(hipMemcpyAsync(out, s.b.pupil.ptr, N * sizeof(double2), hipMemcpyDeviceToHost, s.ctx->stream));
  } else {
    launch_pupil_to_f64(s.b.pupil.ptr, true, s.b.work_m.as<double2>(), N, s.ctx->stream);
    HIP_CHECK(hipMemcpyAsync(out, s.b.work_m.ptr, N * sizeof(double2), hipMemcpyDeviceToHost, s.ctx->stream));
  }
  HIP_CHECK(hipStreamSynchronize(s.ctx->stream));
}

void upload_pupil64(fpmhip_session& s, const double* src) {
  const uint64_t N = s.model->N();
  if (s.precision == FPMHIP_PLAIN_C128) {
    HIP_CHECK(hipMemcpyAsync(s.b.pupil.ptr, src, N * sizeof(double2), hipMemcpyHostToDevice, s.ctx->stream));
  } else {
    HIP_CHECK(hipMemcpyAsync(s.b.work_m.ptr, src, N * sizeof(double2), hipMemcpyHostToDevice, s.ctx->stream));
    set_session_pupil(s, s.b.work_m.as<double2>());
  }
  HIP_CHECK(hipStreamSynchronize(s.ctx->stream));
}

void refresh_max_pupil(fpmhip_session& s) {
  launch_max_abs2(s, s.b.pupil.ptr, s.precision == FPMHIP_PLAIN_C128, s.model->N(), 0);
}

// canonicalize_state (runner.rs:122, 220) of an EPRY session: the device
// gauge (canonicalize_object_pupil, gauge.rs:20-296) on the session stream,
// which also refreshes max |P|^2; its failure code lands in flag[2].
void canonicalize_session(fpmhip_session& s) {
  if (s.model->support_count == 0) fail(FPMHIP_INVALID_MODEL, "compiled reference pupil support is empty");
  launch_gauge(s.b.gauge.as<GaugeTile>(), 1, s.ctx->stream);
}

// The gauge descriptor of a session that canonicalizes, uploaded.
void prepare_gauge(fpmhip_session& s) {
  const GaugeTile t = gauge_tile(s);
  s.b.gauge.allocate(sizeof(GaugeTile));
  HIP_CHECK(hipMemcpyAsync(s.b.gauge.ptr, &t, sizeof(t), hipMemcpyHostToDevice, s.ctx->stream));
  HIP_CHECK(hipStreamSynchronize(s.ctx->stream));  // `t` is on the host stack
}

void check_flags(fpmhip_session& s) {
  int flags[4];
  HIP_CHECK(hipMemcpyAsync(flags, s.b.flag.ptr, sizeof(flags), hipMemcpyDeviceToHost, s.ctx->stream));
  HIP_CHECK(hipStreamSynchronize(s.ctx->stream));
  if (flags[0]) fail(FPMHIP_NUMERICAL, "mPIE momentum produced a non-finite object or velocity");
  if (flags[1]) fail(FPMHIP_INVALID_MEASUREMENTS, "no positive-weight, unmasked measurements are available");
  if (flags[2]) {
    int code = FPMHIP_NUMERICAL;
    const char* msg = gauge_failure(flags[2], &code);
    fail(code, msg ? msg : "object/pupil gauge normalization failed");
  }
  if (flags[3]) fail(FPMHIP_INVALID_MODEL, "state frame has invalid gain");
}

void init_aux(fpmhip_session& s) {
  if (!s.cfg.momentum) return;
  const uint64_t M = s.model->M();
  if (s.precision == FPMHIP_PLAIN_C128) {
    launch_fill_zero(s.b.vel.as<double2>(), M, s.ctx->stream);
    HIP_CHECK(hipMemcpyAsync(s.b.anc.ptr, s.b.spec.ptr, M * sizeof(double2), hipMemcpyDeviceToDevice, s.ctx->stream));
  } else {
    launch_fill_zero(s.b.vel.as<float2>(), M, s.ctx->stream);
    launch_convert_pupil(s.b.spec.as<double2>(), s.b.anc.ptr, true, M, s.ctx->stream);
  }
  s.counter = 0;
}

// Philox4x64-10 (Salmon et al., SC 2011) of one counter block, host side
void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B97F4A7C15ull;
      k1 += 0xBB67AE8584CAA73Bull;
    }
    const unsigned __int128 p0 = (unsigned __int128)0xD2E7470EE14C6C93ull * c[0];
    const unsigned __int128 p1 = (unsigned __int128)0xCA5A826395121157ull * c[2];
    const uint64_t n0 = (uint64_t)(p1 >> 64) ^ c[1] ^ k0, n2 = (uint64_t)(p0 >> 64) ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = (uint64_t)p1;
    c[2] = n2;
    c[3] = (uint64_t)p0;
  }
}

// FrameSchedule::order_for_problem (schedule.rs:35-104) for zero-based iteration `iteration`
std::vector<uint64_t> schedule_order(const fpmhip_session& s, const fpmhip_frame_schedule& sched, uint64_t iteration) {
  const Model& m = *s.model;
  std::vector<uint64_t> order(m.F);
  std::iota(order.begin(), order.end(), 0);
  const int schedule = sched.kind;
  if (schedule == FPMHIP_SCHEDULE_SEQUENTIAL) return order;
  if (schedule == FPMHIP_SCHEDULE_RANDOM_SHUFFLE) {
    // a per-iteration seeded permutation (Fisher-Yates from the top)
    const uint64_t key = sched.seed + iteration * 0x9e3779b97f4a7c15ull;
    for (uint64_t i = m.F; i-- > 1;) {
      uint64_t c[4] = {i, 0, 0, 0};
      philox4x64_10(c, key, 1);
      const double u = (double)(c[0] >> 11) * 0x1.0p-53;
      const uint64_t j = std::min<uint64_t>(i, (uint64_t)(u * (double)(i + 1)));
      std::swap(order[i], order[j]);
    }
    return order;
  }
  if (schedule == FPMHIP_SCHEDULE_SNR_WEIGHTED) {  // schedule.rs:84-104, scores reduced on the device
    const std::vector<double> score = snr_scores(s);
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return score[a] > score[b]; });
    return order;
  }
  if (!m.has_k) fail(FPMHIP_INVALID_MODEL, "schedule requires k-vectors");
  std::vector<double> rad(m.F), ang(m.F);
  for (uint64_t f = 0; f < m.F; ++f) {  // frame_vector (schedule.rs:130-146)
    double kx = 0.0, ky = 0.0;
    if (m.mux) {
      double tw = 0.0;
      for (uint64_t e = m.row_ptr[f]; e < m.row_ptr[f + 1]; ++e) {
        kx += m.mw[e] * m.kx[m.msrc[e]];
        ky += m.mw[e] * m.ky[m.msrc[e]];
        tw += m.mw[e];
      }
      kx /= tw;
      ky /= tw;
    } else {
      kx = m.kx[f];
      ky = m.ky[f];
    }
    rad[f] = std::hypot(kx, ky);
    double a = std::fmod(std::atan2(ky, kx), 6.283185307179586);
    ang[f] = a < 0.0 ? a + 6.283185307179586 : a;
  }
  if (schedule == FPMHIP_SCHEDULE_BRIGHTFIELD_FIRST)
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return rad[a] < rad[b]; });
  else if (schedule == FPMHIP_SCHEDULE_SPIRAL_OUT)
    std::stable_sort(order.begin(), order.end(),
                     [&](uint64_t a, uint64_t b) { return rad[a] != rad[b] ? rad[a] < rad[b] : ang[a] < ang[b]; });
  else
    fail(FPMHIP_UNSUPPORTED, "unknown frame schedule");
  return order;
}

// Frame gains and background of the reconstruction state, as initialised
// from the model (state.rs:280-355): host gains, device gains, device
// background (length 0, N or F * N).
void load_calibration(fpmhip_session& s) {
  const Model& m = *s.model;
  s.has_gain = m.has_gain;
  s.gain = m.gain;
  // the model's device gains hold ones where it has none
  if (m.F)
    HIP_CHECK(hipMemcpyAsync(s.b.gains.ptr, m.d_gains.ptr, m.F * sizeof(double), hipMemcpyDeviceToDevice,
                             s.ctx->stream));
  s.bg_len = m.background.size();
  if (s.bg_len)
    HIP_CHECK(hipMemcpyAsync(s.b.bg.ptr, m.d_background.ptr, s.bg_len * sizeof(double), hipMemcpyDeviceToDevice,
                             s.ctx->stream));
}

// update_gain then update_background for one frame (common.rs:144-168,
// 390-489), on the device from the frame's intrinsic prediction: the state's
// gain of the frame and its background are updated in place, with no wait
// on the host. Whether the background step applies (the frame has unmasked
// pixels) is known on the host; the first step makes the background
// per-frame (update_background expands a shared or absent background).
void calibrate(fpmhip_session& s, uint64_t f, uint32_t K, const double* mode_w_dev) {
  const Model& m = *s.model;
  const Measurements& me = *s.meas;
  if (s.cfg.gain_update) launch_calibration(s, f, K, mode_w_dev, 0);
  if (s.cfg.bg_update) {
    bool any = me.masks.empty();
    for (uint64_t p = 0; !any && p < m.N(); ++p) any = me.masks[f * m.N() + p] != 0;
    if (any) {
      launch_calibration(s, f, K, mode_w_dev, 1);
      const uint64_t total = m.N() * m.F;
      if (s.bg_len != total) {
        launch_background_expand(s);
        s.bg_len = total;
      }
      launch_background_shift(s, f, s.cfg.bg_min, s.cfg.bg_max);
    }
  }
}

// The host mirror of the state's frame gains after a sweep that recovered
// them on the device (the stream has been synchronised).
void download_recovered_gains(fpmhip_session& s) {
  const Model& m = *s.model;
  double present = 0.0;
  HIP_CHECK(hipMemcpy(&present, s.b.scal.as<double>() + kScalGainsPresent, sizeof(double), hipMemcpyDeviceToHost));
  if (present == 0.0 && !s.has_gain) return;
  s.has_gain = true;
  s.gain.resize(m.F);
  HIP_CHECK(hipMemcpy(s.gain.data(), s.b.gains.ptr, m.F * sizeof(double), hipMemcpyDeviceToHost));
}

// The launches of one iteration of projection_update over `order`
// (common.rs:57-306), frame by frame on the session stream (checked by
// sweep_checks beforehand, so nothing here fails mid-way). Host-side state
// they advance: the mPIE counter and the background length.
void issue_sweep(fpmhip_session& s, const uint64_t* order, uint64_t n) {
  const Model& m = *s.model;
  const Measurements& me = *s.meas;
  const uint64_t N = m.N();
  launch_flag_reset(s);
  if (!s.cfg.has_pupil_step) refresh_max_pupil(s);
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t f = order[i];
    const double w = me.weight[f];
    if (w == 0.0) continue;
    uint64_t e0 = f, e1 = f + 1;
    if (m.mux) {
      e0 = m.row_ptr[f];
      e1 = m.row_ptr[f + 1];
    }
    const uint32_t K = (uint32_t)(e1 - e0);
    double sw_sum = 0.0;
    if (m.mux)
      for (uint64_t e = e0; e < e1; ++e) sw_sum += m.mw[e];
    else
      sw_sum = 1.0;
    if (s.cfg.has_pupil_step) refresh_max_pupil(s);
    for (uint32_t k = 0; k < K; ++k) launch_extract(s, m.mux ? m.msrc[e0 + k] : f, k);
    if (s.precision == FPMHIP_PLAIN_C128)
      fft2_f64(*m.fft_low, s.b.field.as<double2>(), K, true, s.ctx->stream);
    else
      fft2_f32(*m.fft_low, s.b.field.as<float2>(), K, true, s.ctx->stream);
    if (s.cfg.gain_update || s.cfg.bg_update) calibrate(s, f, K, m.mux ? m.d_mux_w.as<double>() + e0 : nullptr);
    launch_project(s, f, K, m.mux ? m.d_mux_w.as<double>() + e0 : nullptr, sw_sum, i);
    for (uint32_t k = 0; k < K; ++k) {
      const uint64_t src = m.mux ? m.msrc[e0 + k] : f;
      const double swk = m.mux ? m.mw[e0 + k] : 1.0;
      if (k > 0) launch_difference_frame(s, k, f);
      const size_t vb = s.vbytes();
      if (s.precision == FPMHIP_PLAIN_C128)
        fft2_f64(*m.fft_low, reinterpret_cast<double2*>(s.b.field.as<char>() + k * N * vb), 1, false, s.ctx->stream);
      else
        fft2_f32(*m.fft_low, reinterpret_cast<float2*>(s.b.field.as<char>() + k * N * vb), 1, false, s.ctx->stream);
      if (s.cfg.has_pupil_step)
        launch_max_abs2(s, s.b.patch.as<char>() + k * N * vb, s.precision == FPMHIP_PLAIN_C128, N, 1);
      const double nsw = swk / sw_sum;
      launch_update(s, k, nsw, k == 0);
      launch_insert(s, src, k, w * s.cfg.object_step * nsw);
    }
    if (s.cfg.has_pupil_step) launch_pupil_step(s, w * s.cfg.pupil_step);
    if (s.cfg.momentum) {  // apply_object_momentum (common.rs:308-353)
      s.counter += 1;
      if (s.counter >= s.cfg.interval) {
        launch_momentum(s);
        s.counter = 0;
      }
    }
  }
  if (s.cfg.gauge) canonicalize_session(s);
}

// The checks of a sweep that the reference makes while it runs, made up
// front from host state: frame range, the state's gains (unless recovered
// on the device, which checks them), the mPIE counter.
void sweep_checks(const fpmhip_session& s, const uint64_t* order, uint64_t n) {
  const Model& m = *s.model;
  const Measurements& me = *s.meas;
  uint64_t counter = s.counter;
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t f = order[i];
    if (f >= m.F) fail(FPMHIP_FRAME_OUT_OF_RANGE, "frame index out of range");
    if (me.weight[f] == 0.0) continue;
    if (!s.cfg.gain_update) {
      const double gain = s.has_gain ? s.gain[f] : 1.0;
      if (!std::isfinite(gain) || gain <= 0.0) fail(FPMHIP_INVALID_MODEL, "state frame has invalid gain");
    }
    if (s.cfg.momentum) {
      counter += 1;
      if (counter >= s.cfg.interval) {
        if (counter != s.cfg.interval)
          fail(FPMHIP_INVALID_MODEL, "mPIE frame counter exceeds the configured momentum interval");
        counter = 0;
      }
    }
  }
}

// One iteration (projection_update over `order`, then canonicalize_state),
// as one replay of a HIP graph of its launches: a sweep's launches depend on
// the order, the mPIE counter and the background length at its start, and
// the graph of each such start is captured once per session and replayed
// after (the session keeps at most kSweepGraphs of them). Objectives are
// reduced on the host afterwards, in order.
void sweep(fpmhip_session& s, const uint64_t* order, uint64_t n, double* per_frame, double* objective) {
  const Model& m = *s.model;
  const Measurements& me = *s.meas;
  sweep_checks(s, order, n);
  SweepGraph* g = nullptr;
  for (SweepGraph& c : s.graphs)
    if (c.counter0 == s.counter && c.bg0 == s.bg_len && c.order.size() == n &&
        std::equal(c.order.begin(), c.order.end(), order))
      g = &c;
  if (!g) {
    if (s.graphs.size() >= kSweepGraphs) s.graphs.erase(s.graphs.begin());
    s.graphs.emplace_back();
    g = &s.graphs.back();
    g->order.assign(order, order + n);
    g->counter0 = s.counter;
    g->bg0 = s.bg_len;
    hipGraph_t graph = nullptr;
    HIP_CHECK(hipStreamBeginCapture(s.ctx->stream, hipStreamCaptureModeThreadLocal));
    try {
      issue_sweep(s, order, n);
    } catch (...) {
      (void)hipStreamEndCapture(s.ctx->stream, &graph);
      if (graph) (void)hipGraphDestroy(graph);
      s.counter = g->counter0;
      s.bg_len = g->bg0;
      s.graphs.pop_back();
      throw;
    }
    HIP_CHECK(hipStreamEndCapture(s.ctx->stream, &graph));
    const hipError_t e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    g->counter1 = s.counter;
    g->bg1 = s.bg_len;
    s.counter = g->counter0;
    s.bg_len = g->bg0;
    if (e != hipSuccess) {
      s.graphs.pop_back();
      HIP_CHECK(e);
    }
  }
  HIP_CHECK(hipGraphLaunch(g->exec, s.ctx->stream));
  s.counter = g->counter1;
  s.bg_len = g->bg1;
  std::vector<double> partials(n * s.nblk * 2);
  if (n)
    HIP_CHECK(hipMemcpyAsync(partials.data(), s.b.partials.ptr, partials.size() * sizeof(double),
                             hipMemcpyDeviceToHost, s.ctx->stream));
  check_flags(s);
  if (s.cfg.gain_update) download_recovered_gains(s);
  double osum = 0.0, wsum = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    const double fw = me.weight[order[i]];
    double loss = 0.0;
    if (fw != 0.0) {
      double ls = 0.0, lc = 0.0;
      for (uint32_t b = 0; b < s.nblk; ++b) {
        ls += partials[(i * s.nblk + b) * 2];
        lc += partials[(i * s.nblk + b) * 2 + 1];
      }
      if (lc == 0.0) fail(FPMHIP_INVALID_MEASUREMENTS, "frame has no unmasked pixels");
      loss = ls / lc;
    }
    if (per_frame) per_frame[i] = loss;
    osum += fw * loss;  // StepSummary::push_frame (metrics.rs:25-31)
    wsum += fw;
  }
  (void)m;
  if (objective) *objective = wsum > 0.0 ? osum / wsum : std::numeric_limits<double>::quiet_NaN();
}

}  // namespace

namespace fpmhip {

// The gauge descriptor of a session: its spectrum and pupil, the model's
// reference pupil and support, status in the session's flag[2].
GaugeTile gauge_tile(const fpmhip_session& s) {
  const Model& m = *s.model;
  GaugeTile t{};
  t.spec = s.b.spec.as<double2>();
  t.pupil = s.b.pupil.ptr;
  t.ref = m.d_pupil64.as<double2>();
  t.support = m.d_support.as<uint8_t>();
  t.maxp_out = s.b.scal.as<double>();
  t.status = s.b.flag.as<int32_t>() + 2;
  t.tol = 64.0 * std::numeric_limits<double>::epsilon() * (double)m.support_count;
  t.nh = m.nh;
  t.nw = m.nw;
  t.rh = m.rh;
  t.rw = m.rw;
  t.pupil_f32 = s.precision == FPMHIP_MIXED_F32 ? 1 : 0;
  t.pupil64 = s.b.pupil64.as<double2>();
  bool row_exact = true, col_exact = true;
  for (uint64_t k = 0; k < m.S; ++k) {
    row_exact = row_exact && std::fabs(m.off_r[k]) <= 1e-12;
    col_exact = col_exact && std::fabs(m.off_c[k]) <= 1e-12;
  }
  t.row_exact = row_exact ? 1 : 0;
  t.col_exact = col_exact ? 1 : 0;
  return t;
}


uint64_t momentum_span_capacity(const Model& m) { return 1 + (uint64_t)m.rh * ((m.rw + 7 + 63) / 64); }

void upload_momentum_spans(fpmhip_session& s, const double* pupil) {
  if (!s.b.mspan.ptr) return;
  const Model& m = *s.model;
  const int nh = (int)m.nh, nw = (int)m.nw;
  // columns [lo, hi) of each pupil row that are non-zero in complex64
  std::vector<int> lo(nh, nw), hi(nh, 0);
  for (int r = 0; r < nh; ++r)
    for (int c = 0; c < nw; ++c) {
      const uint64_t p = (uint64_t)r * nw + c;
      if ((float)pupil[2 * p] != 0.0f || (float)pupil[2 * p + 1] != 0.0f) {
        lo[r] = std::min(lo[r], c);
        hi[r] = c + 1;
      }
    }
  // window row a (0..nh) receives the taps of pupil rows a - 1 and a, reaching
  // one column further right
  std::vector<int> wlo(nh + 1, nw + 1), whi(nh + 1, 0);
  for (int a = 0; a <= nh; ++a)
    for (int r = a > 0 ? a - 1 : 0; r <= a && r < nh; ++r)
      if (hi[r] > lo[r]) {
        wlo[a] = std::min(wlo[a], lo[r]);
        whi[a] = std::max(whi[a], std::min(hi[r] + 1, nw + 1));
      }
  // per object row, the hull of the columns any source's window reaches
  std::vector<int64_t> ylo(m.rh, std::numeric_limits<int64_t>::max()), yhi(m.rh, 0);
  for (uint64_t src = 0; src < m.S; ++src)
    for (int a = 0; a <= nh; ++a) {
      if (whi[a] <= wlo[a]) continue;
      const uint64_t y = (uint64_t)m.ax_r[src].lower + a;
      if (y >= m.rh) continue;
      const int64_t c0 = (int64_t)m.ax_c[src].lower + wlo[a];
      const int64_t c1 = std::min<int64_t>((int64_t)m.ax_c[src].lower + whi[a], m.rw);
      ylo[y] = std::min(ylo[y], c0);
      yhi[y] = std::max(yhi[y], c1);
    }
  std::vector<uint32_t> spans(1, 0);
  for (uint32_t y = 0; y < m.rh; ++y)
    for (int64_t c = ylo[y] & ~int64_t{7}; c < yhi[y]; c += 64) {  // chunks start on 128-byte lines
      const uint32_t len = (uint32_t)std::min<int64_t>(64, yhi[y] - c);
      spans.push_back((uint32_t)((uint64_t)y * m.rw + (uint64_t)c) | ((len - 1) << 25));
    }
  spans[0] = (uint32_t)(spans.size() - 1);
  if (spans.size() > momentum_span_capacity(m)) fail(FPMHIP_DEVICE_ERROR, "momentum span table overflow");
  HIP_CHECK(hipMemcpyAsync(s.b.mspan.ptr, spans.data(), spans.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           s.ctx->stream));
  HIP_CHECK(hipStreamSynchronize(s.ctx->stream));  // `spans` is pageable host memory
}

void session_initialize_launch(fpmhip_session& s) {
  for (uint64_t f = 0; f < s.model->F; ++f) {
    const double g = s.model->has_gain ? s.model->gain[f] : 1.0;
    if (!std::isfinite(g) || g <= 0.0) fail(FPMHIP_INVALID_MODEL, "frame has invalid gain");
  }
  s.initialized = false;
  launch_flag_reset(s);
  launch_init(s);
  set_session_pupil(s, s.model->d_pupil64.as<double2>());
  if (!s.extent_from_model) {  // the model's pupil: its extents and lists do not change
    launch_pupil_extent(s);
    upload_momentum_spans(s, s.model->pupil.data());
    s.extent_from_model = true;
  }
  load_calibration(s);
  init_aux(s);
  refresh_max_pupil(s);
}

void session_initialize_check(fpmhip_session& s) {
  if (s.cfg.gauge) canonicalize_session(s);
  check_flags(s);  // waits for the launches
  s.initialized = true;
}

}  // namespace fpmhip

extern "C" {

int fpmhip_context_create(int32_t device, fpmhip_context** out) {
  return guarded([&] {
    require(out != nullptr, "context", "output is null");
    int count = 0;
    HIP_CHECK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) fail(FPMHIP_DEVICE_ERROR, "no such HIP device");
    auto c = std::make_unique<fpmhip_context>();
    c->device = device;
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipGetDeviceProperties(&c->props, device));
    HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->scratch = noalloc_scratch_address(device);
    *out = c.release();
  });
}

int fpmhip_context_destroy(fpmhip_context* c) {
  return guarded([&] {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) {
      // work queued on the stream finishes before the stream goes away
      (void)hipStreamSynchronize(c->stream);
      (void)hipStreamDestroy(c->stream);
    }
    release_resident_plans(c);
    delete c;
  });
}

int fpmhip_context_synchronize(fpmhip_context* c) {
  return guarded([&] {
    require(c != nullptr, "context", "context is null");
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int fpmhip_fft2_host(fpmhip_context* c, double* values, uint64_t h, uint64_t w, int32_t direction) {
  return guarded([&] {
    require(c && values, "values", "null argument");
    if (h == 0 || w == 0) fail(FPMHIP_INVALID_SHAPE, "FFT dimensions must be non-zero");
    set_device(c);
    auto plan = make_fft2((uint32_t)h, (uint32_t)w);
    DeviceBuffer d(h * w * sizeof(double2));
    HIP_CHECK(hipMemcpyAsync(d.ptr, values, d.bytes, hipMemcpyHostToDevice, c->stream));
    fft2_f64(*plan, d.as<double2>(), 1, direction == FPMHIP_INVERSE, c->stream);
    HIP_CHECK(hipMemcpyAsync(values, d.ptr, d.bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

// ImagePlaneModel::validate (image_plane_fpm.rs:841-972) + upload
int fpmhip_model_create(fpmhip_context* c, const fpmhip_model_desc* d, fpmhip_model** out) {
  return guarded([&] {
    require(c && d && out, "model", "null argument");
    set_device(c);
    auto m = std::make_unique<fpmhip_model>();
    m->ctx = c;
    if (d->image_h == 0 || d->image_w == 0 || d->source_count == 0)
      fail(FPMHIP_INVALID_MODEL, "source k-vectors and crops; counts must be equal and non-zero");
    if (d->recon_h < d->image_h || d->recon_w < d->image_w)
      fail(FPMHIP_INVALID_MODEL, "reconstruction shape must contain a low-resolution crop");
    if (d->image_h > 4096 || d->image_w > 4096 || d->recon_h > 65536 || d->recon_w > 65536)
      fail(FPMHIP_UNSUPPORTED, "grid is larger than this build supports");
    m->nh = (uint32_t)d->image_h;
    m->nw = (uint32_t)d->image_w;
    m->rh = (uint32_t)d->recon_h;
    m->rw = (uint32_t)d->recon_w;
    m->S = d->source_count;
    m->mux = d->mux_row_ptr != nullptr;
    m->F = m->mux ? d->frame_count : d->source_count;
    if (m->F == 0) fail(FPMHIP_INVALID_MODEL, "multiplexing matrix must contain at least one measured frame");
    const uint64_t N = m->N();
    std::vector<SrcParam> sp(m->S);
    for (uint64_t s = 0; s < m->S; ++s) {
      const uint64_t r0 = d->crop_start[2 * s], c0 = d->crop_start[2 * s + 1];
      if (r0 + m->nh > m->rh || c0 + m->nw > m->rw)
        fail(FPMHIP_INVALID_MODEL, "crop lies outside reconstruction shape");
      const double orow = d->offset ? d->offset[2 * s] : 0.0, ocol = d->offset ? d->offset[2 * s + 1] : 0.0;
      m->crop_r.push_back(r0);
      m->crop_c.push_back(c0);
      m->off_r.push_back(orow);
      m->off_c.push_back(ocol);
      const Axis ar = interpolation_axis(r0, m->nh, m->rh, orow);
      const Axis ac = interpolation_axis(c0, m->nw, m->rw, ocol);
      sp[s] = SrcParam{r0, c0, ar, ac, off_zero(orow, ocol) ? 1 : 0, 0};
      m->ax_r.push_back(ar);
      m->ax_c.push_back(ac);
    }
    m->pupil.assign(d->pupil, d->pupil + 2 * N);
    for (double v : m->pupil)
      if (!std::isfinite(v)) fail(FPMHIP_INVALID_MODEL, "pupil support or numeric values are invalid");
    m->support.assign(d->support, d->support + N);
    for (uint8_t v : m->support) {
      if (v > 1) fail(FPMHIP_INVALID_MODEL, "pupil support values must be exactly zero or one");
      m->support_count += v;
    }
    std::vector<double> gains(m->F, 1.0);
    if (d->frame_gain) {
      m->has_gain = true;
      m->gain.assign(d->frame_gain, d->frame_gain + m->F);
      for (double g : m->gain)
        if (!std::isfinite(g) || g < 0.0) fail(FPMHIP_INVALID_MODEL, "frame gains must be finite and non-negative");
      gains = m->gain;
    }
    if (d->background && d->background_len) {
      if (d->background_len != N && d->background_len != N * m->F)
        fail(FPMHIP_INVALID_MODEL, "background must be one image or one image per frame");
      m->background.assign(d->background, d->background + d->background_len);
      for (double v : m->background)
        if (!std::isfinite(v)) fail(FPMHIP_INVALID_MODEL, "background values must be finite");
    }
    if (m->mux) {
      m->row_ptr.assign(d->mux_row_ptr, d->mux_row_ptr + m->F + 1);
      const uint64_t nnz = m->row_ptr[m->F];
      m->msrc.assign(d->mux_source, d->mux_source + nnz);
      m->mw.assign(d->mux_weight, d->mux_weight + nnz);
      for (uint64_t f = 0; f < m->F; ++f) {
        std::vector<bool> seen(m->S, false);
        bool bad = m->row_ptr[f + 1] <= m->row_ptr[f];
        for (uint64_t e = m->row_ptr[f]; e < m->row_ptr[f + 1] && !bad; ++e) {
          const uint64_t s = m->msrc[e];
          bad = s >= m->S || !std::isfinite(m->mw[e]) || m->mw[e] < 0.0 || seen[s];
          if (s < m->S) seen[s] = true;
        }
        if (bad)
          fail(FPMHIP_INVALID_MODEL, "multiplexing row is empty or contains a duplicate/invalid source or weight");
      }
      std::vector<uint32_t> rp(m->row_ptr.begin(), m->row_ptr.end()), ms(m->msrc.begin(), m->msrc.end());
      upload(m->d_mux_ptr, rp.data(), rp.size(), c->stream);
      upload(m->d_mux_src, ms.data(), ms.size(), c->stream);
      upload(m->d_mux_w, m->mw.data(), m->mw.size(), c->stream);
    }
    if (d->k_vectors) {
      m->has_k = true;
      for (uint64_t s = 0; s < m->S; ++s) {
        m->kx.push_back(d->k_vectors[2 * s]);
        m->ky.push_back(d->k_vectors[2 * s + 1]);
      }
    }
    upload(m->d_src, sp.data(), sp.size(), c->stream);
    upload(m->d_gains, gains.data(), gains.size(), c->stream);
    upload(m->d_pupil64, m->pupil.data(), m->pupil.size(), c->stream);
    upload(m->d_support, m->support.data(), m->support.size(), c->stream);
    if (!m->background.empty()) upload(m->d_background, m->background.data(), m->background.size(), c->stream);
    m->fft_low = make_fft2(m->nh, m->nw);
    m->fft_high = make_fft2(m->rh, m->rw);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    *out = m.release();
  });
}

int fpmhip_model_destroy(fpmhip_model* m) {
  return guarded([&] {
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    delete m;
  });
}

// MeasurementStack::from_vec validation (src/measurements/stack.rs:60-105)
int fpmhip_measurements_create(fpmhip_context* c, const fpmhip_model* m, const uint16_t* counts,
                               const double* frames, const double* weights, const uint8_t* masks,
                               fpmhip_measurements** out) {
  return guarded([&] {
    require(c && m && out, "measurements", "null argument");
    if ((counts == nullptr) == (frames == nullptr))
      fail(FPMHIP_INVALID_MEASUREMENTS, "exactly one of uint16 counts and f64 frames is required");
    set_device(c);
    auto me = std::make_unique<fpmhip_measurements>();
    me->ctx = c;
    me->model = m;
    me->F = m->F;
    me->N = m->N();
    const uint64_t FN = me->F * me->N;
    me->weight.assign(me->F, 1.0);
    if (weights)
      for (uint64_t f = 0; f < me->F; ++f) {
        if (!std::isfinite(weights[f]) || weights[f] < 0.0)
          fail(FPMHIP_INVALID_MEASUREMENTS, "frame weights must be finite and non-negative");
        me->weight[f] = weights[f];
      }
    if (counts) {
      me->is_u16 = true;
      me->counts_host.assign(counts, counts + FN);
      upload(me->d_frames, counts, FN, c->stream);
    } else {
      for (uint64_t i = 0; i < FN; ++i)
        if (!std::isfinite(frames[i])) fail(FPMHIP_INVALID_MEASUREMENTS, "measurements must be finite");
      me->frames_host.assign(frames, frames + FN);
      upload(me->d_frames, frames, FN, c->stream);
    }
    if (masks) {
      me->masks.assign(masks, masks + FN);
      for (uint64_t f = 0; f < me->F; ++f) {
        uint64_t valid = 0;
        for (uint64_t p = 0; p < me->N; ++p) valid += masks[f * me->N + p] != 0;
        if (valid == 0 && me->weight[f] > 0.0)
          fail(FPMHIP_INVALID_MEASUREMENTS, "a positive-weight frame has no unmasked pixels");
      }
      upload(me->d_masks, masks, FN, c->stream);
    }
    upload(me->d_weights, me->weight.data(), me->F, c->stream);
    HIP_CHECK(hipStreamSynchronize(c->stream));
    *out = me.release();
  });
}

int fpmhip_measurements_destroy(fpmhip_measurements* me) {
  return guarded([&] {
    if (!me) return;
    (void)hipSetDevice(me->ctx->device);
    delete me;
  });
}

int fpmhip_forward_stack(fpmhip_context* c, const fpmhip_model* m, const double* spectrum, const double* pupil,
                         double* out) {
  return guarded([&] {
    require(c && m && spectrum && out, "forward_stack", "null argument");
    set_device(c);
    for (uint64_t f = 0; f < m->F; ++f) {
      const double g = m->has_gain ? m->gain[f] : 1.0;
      if (!std::isfinite(g) || g <= 0.0) fail(FPMHIP_INVALID_MODEL, "frame has invalid gain");
    }
    const uint64_t N = m->N(), M = m->M();
    DeviceBuffer ds(M * sizeof(double2)), dp, fields(m->S * N * sizeof(double2)), dout(m->F * N * sizeof(double));
    HIP_CHECK(hipMemcpyAsync(ds.ptr, spectrum, ds.bytes, hipMemcpyHostToDevice, c->stream));
    const double2* pp = m->d_pupil64.as<double2>();
    if (pupil) {
      dp.allocate(N * sizeof(double2));
      HIP_CHECK(hipMemcpyAsync(dp.ptr, pupil, dp.bytes, hipMemcpyHostToDevice, c->stream));
      pp = dp.as<double2>();
    }
    launch_forward_stack(*m, ds.as<double2>(), pp, fields.as<double2>(), dout.as<double>(), c->stream);
    HIP_CHECK(hipMemcpyAsync(out, dout.ptr, dout.bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  });
}

int fpmhip_session_create(const fpmhip_model* m, const fpmhip_measurements* me, int32_t precision,
                          const fpmhip_algorithm* alg, fpmhip_session** out) {
  return guarded([&] {
    require(m && me && alg && out, "session", "null argument");
    if (me->model != m && (me->F != m->F || me->N != m->N()))
      fail(FPMHIP_INVALID_MEASUREMENTS, "measurements do not match the model");
    if (precision != FPMHIP_PLAIN_C128 && precision != FPMHIP_MIXED_F32)
      fail(FPMHIP_INVALID_PARAMETER, "unknown precision", "precision");
    if (int rc = fpmhip_algorithm_validate(alg)) throw Failure{rc};
    set_device(m->ctx);
    auto s = std::make_unique<fpmhip_session>();
    s->ctx = m->ctx;
    s->model = m;
    s->meas = me;
    s->precision = precision;
    s->alg = *alg;
    s->cfg = make_config(*alg);
    s->has_gain = m->has_gain;
    s->gain = m->gain;
    uint32_t kmax = 1;
    if (m->mux)
      for (uint64_t f = 0; f < m->F; ++f) kmax = std::max<uint32_t>(kmax, (uint32_t)(m->row_ptr[f + 1] - m->row_ptr[f]));
    s->kmax = kmax;
    const uint64_t N = m->N(), M = m->M();
    s->nblk = (uint32_t)((N + 255) / 256);
    const size_t vb = s->vbytes();
    SessionBuffers& b = s->b;
    b.spec.allocate(M * sizeof(double2));
    if (s->cfg.momentum) {
      b.vel.allocate(M * vb);
      b.anc.allocate(M * vb);
    }
    b.pupil.allocate(N * vb);
    b.patch.allocate(kmax * N * vb);
    b.field.allocate(kmax * N * vb);
    b.pred.allocate(N * sizeof(double));
    b.proj.allocate(N * vb);
    b.U.allocate(N * vb);
    b.grad.allocate(N * vb);
    b.scal.allocate(8 * sizeof(double));
    b.partials.allocate(std::max<uint64_t>(m->F, 1) * s->nblk * 2 * sizeof(double));
    b.flag.allocate(4 * sizeof(int));
    b.work_m.allocate(std::max(M, N) * sizeof(double2));
    b.avg.allocate(N * sizeof(double));
    b.aw.allocate(N * sizeof(double));
    // background of the state: per frame once recovery may write it
    b.bg.allocate(std::max<uint64_t>(s->cfg.bg_update ? N * m->F : m->background.size(), 1) * sizeof(double));
    b.calib.allocate((size_t)s->nblk * 2 * sizeof(double));
    b.gains.allocate(std::max<uint64_t>(m->F, 1) * sizeof(double));
    load_calibration(*s);
    if (precision == FPMHIP_MIXED_F32 && m->nh == m->nw && m->nh > kFusedLdsMaxN && m->nh <= kFusedMaxN) {
      b.group.allocate(group_state_bytes());
      HIP_CHECK(hipMemsetAsync(b.group.ptr, 0, b.group.bytes, m->ctx->stream));
    }
    if (precision == FPMHIP_MIXED_F32 && s->cfg.has_pupil_step) {
      b.pupil64.allocate(N * sizeof(double2));
      b.bounds.allocate((size_t)((m->rh + kBoundBlock - 1) / kBoundBlock) * ((m->rw + kBoundBlock - 1) / kBoundBlock) *
                        sizeof(uint64_t));
    }
    set_session_pupil(*s, m->d_pupil64.as<double2>());
    if (precision == FPMHIP_MIXED_F32 && s->cfg.momentum && m->M() < (1ull << 25))
      b.mspan.allocate(momentum_span_capacity(*m) * sizeof(uint32_t));
    if (precision == FPMHIP_MIXED_F32) {
      b.extent.allocate(m->nh * sizeof(uint32_t));
      if (m->nh == m->nw && m->nh <= kFusedMaxN) b.dest.allocate(dest_table_words(m->nh) * sizeof(uint32_t));
      launch_pupil_extent(*s);
      s->extent_from_model = true;
    }
    upload_momentum_spans(*s, m->pupil.data());
    launch_flag_reset(*s);
    if (s->cfg.gauge) prepare_gauge(*s);
    HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
    *out = s.release();
  });
}

int fpmhip_session_destroy(fpmhip_session* s) {
  return guarded([&] {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    delete s;
  });
}


int fpmhip_session_initialize(fpmhip_session* s) {
  return guarded([&] {
    require(s != nullptr, "session", "session is null");
    set_device(s->ctx);
    session_initialize_launch(*s);
    session_initialize_check(*s);
  });
}

int fpmhip_session_set_state(fpmhip_session* s, const double* spectrum, const double* pupil, const double* velocity,
                             const double* anchor, uint64_t counter, const double* frame_gains,
                             const double* background, uint64_t background_len) {
  return guarded([&] {
    require(s && spectrum && pupil, "state", "null argument");
    if (background) {
      const uint64_t N = s->model->N(), F = s->model->F;
      if (background_len != N && background_len != N * F)
        fail(FPMHIP_INVALID_SHAPE, "background must hold one image or one image per frame");
    }
    set_device(s->ctx);
    const uint64_t M = s->model->M();
    HIP_CHECK(hipMemcpyAsync(s->b.spec.ptr, spectrum, M * sizeof(double2), hipMemcpyHostToDevice, s->ctx->stream));
    upload_pupil64(*s, pupil);
    launch_pupil_extent(*s);
    upload_momentum_spans(*s, pupil);
    s->extent_from_model = false;
    if (frame_gains) {
      for (uint64_t f = 0; f < s->model->F; ++f)
        if (!std::isfinite(frame_gains[f]) || frame_gains[f] <= 0.0)
          fail(FPMHIP_INVALID_MODEL, "frame gains must be finite and positive");
      s->has_gain = true;
      s->gain.assign(frame_gains, frame_gains + s->model->F);
      HIP_CHECK(hipMemcpyAsync(s->b.gains.ptr, s->gain.data(), s->gain.size() * sizeof(double), hipMemcpyHostToDevice,
                               s->ctx->stream));
    }
    if (background) {
      if (s->b.bg.bytes < background_len * sizeof(double)) {
        HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
        s->b.bg.allocate(background_len * sizeof(double));
      }
      HIP_CHECK(hipMemcpyAsync(s->b.bg.ptr, background, background_len * sizeof(double), hipMemcpyHostToDevice,
                               s->ctx->stream));
      s->bg_len = background_len;
    }
    if (s->cfg.momentum) {
      if (velocity && anchor) {
        if (counter >= s->cfg.interval)
          fail(FPMHIP_INVALID_MODEL, "mPIE frame counter exceeds the configured momentum interval");
        if (s->precision == FPMHIP_PLAIN_C128) {
          HIP_CHECK(hipMemcpyAsync(s->b.vel.ptr, velocity, M * sizeof(double2), hipMemcpyHostToDevice, s->ctx->stream));
          HIP_CHECK(hipMemcpyAsync(s->b.anc.ptr, anchor, M * sizeof(double2), hipMemcpyHostToDevice, s->ctx->stream));
        } else {
          // complex64 storage: through the FP64 scratch, rounded on the device
          for (int k = 0; k < 2; ++k) {
            HIP_CHECK(hipMemcpyAsync(s->b.work_m.ptr, k ? anchor : velocity, M * sizeof(double2),
                                     hipMemcpyHostToDevice, s->ctx->stream));
            launch_convert_pupil(s->b.work_m.as<double2>(), k ? s->b.anc.ptr : s->b.vel.ptr, true, M, s->ctx->stream);
          }
        }
        s->counter = counter;
      } else {
        init_aux(*s);
      }
    }
    refresh_max_pupil(*s);
    launch_flag_reset(*s);
    if (s->cfg.gauge) canonicalize_session(*s);
    check_flags(*s);
    s->initialized = true;
  });
}

// ReconstructionState::from_object (state.rs:356-383): the state of
// fpmhip_session_initialize with the object spectrum replaced by
// fftshift(FFT(object) / M) of the given object, and the algorithm's
// auxiliary state and gauge made from that spectrum.
int fpmhip_session_initialize_from_object(fpmhip_session* s, const double* object) {
  return guarded([&] {
    require(s != nullptr && object != nullptr, "session", "null argument");
    set_device(s->ctx);
    const Model& m = *s->model;
    session_initialize_launch(*s);
    HIP_CHECK(hipMemcpyAsync(s->b.work_m.ptr, object, m.M() * sizeof(double2), hipMemcpyHostToDevice, s->ctx->stream));
    fft2_f64(*m.fft_high, s->b.work_m.as<double2>(), 1, false, s->ctx->stream);
    launch_shift(s->b.work_m.as<double2>(), s->b.spec.as<double2>(), m.rh, m.rw, m.rh / 2, m.rw / 2, s->ctx->stream);
    init_aux(*s);
    session_initialize_check(*s);
  });
}

int fpmhip_session_sweep(fpmhip_session* s, const uint64_t* order, uint64_t order_len, double* per_frame,
                         double* objective) {
  return guarded([&] {
    require(s != nullptr, "session", "session is null");
    if (!s->initialized) fail(FPMHIP_INVALID_MODEL, "session state is not initialized");
    set_device(s->ctx);
    std::vector<uint64_t> seq;
    if (!order) {
      seq.resize(s->model->F);
      std::iota(seq.begin(), seq.end(), 0);
      order = seq.data();
      order_len = seq.size();
    }
    if (order_len > s->model->F) {
      // partial buffers hold one slot per frame of an iteration
      fail(FPMHIP_INVALID_PARAMETER, "order is longer than the frame count", "order");
    }
    sweep(*s, order, order_len, per_frame, objective);
  });
}

int fpmhip_session_schedule(fpmhip_session* s, int32_t schedule, uint64_t* order) {
  return guarded([&] {
    require(s && order, "schedule", "null argument");
    fpmhip_frame_schedule sched{schedule, 0, 0};
    std::vector<uint64_t> o = schedule_order(*s, sched, 0);
    std::copy(o.begin(), o.end(), order);
  });
}

int fpmhip_session_run(fpmhip_session* s, uint64_t iterations, int32_t schedule, double* trace) {
  if (!s) return guarded([&] { require(false, "session", "session is null"); });
  const fpmhip_frame_schedule sched{schedule, 0, 0};
  return fpmhip_session_run_schedule(s, iterations, &sched, 0, trace);
}

int fpmhip_session_schedule_at(fpmhip_session* s, const fpmhip_frame_schedule* schedule, uint64_t iteration,
                               uint64_t* order) {
  return guarded([&] {
    require(s && schedule && order, "schedule", "null argument");
    set_device(s->ctx);
    std::vector<uint64_t> o = schedule_order(*s, *schedule, iteration);
    std::copy(o.begin(), o.end(), order);
  });
}

int fpmhip_session_run_schedule(fpmhip_session* s, uint64_t iterations, const fpmhip_frame_schedule* schedule,
                                uint64_t first_iteration, double* trace) {
  return guarded([&] {
    require(s != nullptr && schedule != nullptr, "session", "null argument");
    set_device(s->ctx);
    if (!s->initialized) {
      if (int rc = fpmhip_session_initialize(s)) throw Failure{rc};
    }
    // orders that do not depend on the iteration are computed once
    const bool per_iteration = schedule->kind == FPMHIP_SCHEDULE_RANDOM_SHUFFLE;
    std::vector<uint64_t> order = schedule_order(*s, *schedule, first_iteration);
    for (uint64_t it = 0; it < iterations; ++it) {
      if (per_iteration && it) order = schedule_order(*s, *schedule, first_iteration + it);
      double obj = 0.0;
      sweep(*s, order.data(), order.size(), nullptr, &obj);
      if (trace) trace[it] = obj;
    }
  });
}

int fpmhip_session_get_spectrum(fpmhip_session* s, double* out) {
  return guarded([&] {
    require(s && out, "spectrum", "null argument");
    set_device(s->ctx);
    HIP_CHECK(hipMemcpyAsync(out, s->b.spec.ptr, s->model->M() * sizeof(double2), hipMemcpyDeviceToHost, s->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  });
}

int fpmhip_session_get_pupil(fpmhip_session* s, double* out) {
  return guarded([&] {
    require(s && out, "pupil", "null argument");
    set_device(s->ctx);
    download_pupil64(*s, out);
  });
}

int fpmhip_session_get_object(fpmhip_session* s, double* out) {
  return guarded([&] {
    require(s && out, "object", "null argument");
    set_device(s->ctx);
    launch_object(*s, s->b.work_m.as<double2>());
    HIP_CHECK(hipMemcpyAsync(out, s->b.work_m.ptr, s->model->M() * sizeof(double2), hipMemcpyDeviceToHost, s->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
  });
}

int fpmhip_session_get_momentum(fpmhip_session* s, double* velocity, double* anchor, uint64_t* counter) {
  return guarded([&] {
    require(s != nullptr, "session", "session is null");
    if (!s->cfg.momentum) fail(FPMHIP_UNSUPPORTED, "session has no mPIE momentum state");
    set_device(s->ctx);
    const uint64_t M = s->model->M();
    const size_t bytes = M * sizeof(double2);
    for (int k = 0; k < 2; ++k) {
      double* out = k ? anchor : velocity;
      if (!out) continue;
      const DeviceBuffer& src = k ? s->b.anc : s->b.vel;
      if (s->precision == FPMHIP_PLAIN_C128) {
        HIP_CHECK(hipMemcpyAsync(out, src.ptr, bytes, hipMemcpyDeviceToHost, s->ctx->stream));
      } else {
        launch_pupil_to_f64(src.ptr, true, s->b.work_m.as<double2>(), M, s->ctx->stream);
        HIP_CHECK(hipMemcpyAsync(out, s->b.work_m.ptr, bytes, hipMemcpyDeviceToHost, s->ctx->stream));
      }
    }
    HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    if (counter) *counter = s->counter;
  });
}

int fpmhip_session_get_calibration(fpmhip_session* s, double* frame_gains, double* background,
                                   uint64_t* background_len) {
  return guarded([&] {
    require(s != nullptr, "session", "session is null");
    const Model& m = *s->model;
    if (frame_gains)
      for (uint64_t f = 0; f < m.F; ++f) frame_gains[f] = s->has_gain ? s->gain[f] : 1.0;
    set_device(s->ctx);
    if (background && s->bg_len) {
      HIP_CHECK(hipMemcpyAsync(background, s->b.bg.ptr, s->bg_len * sizeof(double), hipMemcpyDeviceToHost,
                               s->ctx->stream));
      HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
    }
    if (background_len) *background_len = s->bg_len;
  });
}

int fpmhip_session_calibration_present(const fpmhip_session* s, int32_t* frame_gains, uint64_t* background_len) {
  return guarded([&] {
    require(s != nullptr, "session", "session is null");
    if (frame_gains) *frame_gains = s->has_gain ? 1 : 0;
    if (background_len) *background_len = s->bg_len;
  });
}

int fpmhip_session_clear_calibration(fpmhip_session* s, int32_t frame_gains, int32_t background) {
  return guarded([&] {
    require(s != nullptr, "session", "session is null");
    if (frame_gains && s->has_gain) {
      set_device(s->ctx);
      s->has_gain = false;
      s->gain.assign(s->model->F, 1.0);
      if (s->model->F) {
        HIP_CHECK(hipMemcpyAsync(s->b.gains.ptr, s->gain.data(), s->gain.size() * sizeof(double),
                                 hipMemcpyHostToDevice, s->ctx->stream));
        HIP_CHECK(hipStreamSynchronize(s->ctx->stream));
      }
    }
    if (background) s->bg_len = 0;
  });
}

int fpmhip_session_shape(const fpmhip_session* s, uint64_t* nh, uint64_t* nw, uint64_t* rh, uint64_t* rw,
                         uint64_t* F) {
  return guarded([&] {
    require(s != nullptr, "session", "session is null");
    if (nh) *nh = s->model->nh;
    if (nw) *nw = s->model->nw;
    if (rh) *rh = s->model->rh;
    if (rw) *rw = s->model->rw;
    if (F) *F = s->model->F;
  });
}

}  // extern "C"
