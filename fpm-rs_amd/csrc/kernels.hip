// kernels.hip — per-stage HIP kernels of one LED update (projection_update,
// src/algorithms/common.rs:57-306), state initialisation (state.rs:280-355),
// mPIE momentum (common.rs:355-387), the result transform (result.rs:397-413)
// and the forward stack (forward.rs:245-372).
//
// Templated on the session precision R: PLAIN_C128 keeps every buffer in
// complex128; MIXED_F32 keeps patches, fields, pupil and updates in complex64
// and runs the FFTs in FP32, while the object spectrum, every pointwise
// formula and all reductions stay FP64. Pointwise formulas follow the
// reference term by term (num-complex operation order; no FMA contraction).
#include "session.h"

namespace fpmhip {

namespace {

constexpr int kBlock = 256;

__device__ inline double meas_value(const void* frames, int is_u16, uint64_t idx) {
  return is_u16 ? (double)static_cast<const uint16_t*>(frames)[idx] : static_cast<const double*>(frames)[idx];
}

__device__ inline double bg_value(const double* bg, uint64_t bg_len, uint64_t N, uint64_t f, uint64_t p) {
  if (bg_len == 0) return 0.0;
  return bg[bg_len == N ? p : f * N + p];
}

// point_loss (src/algorithms/objective.rs:47-71)
__device__ inline double point_loss(double predicted, double measured, int loss) {
  if (loss == FPMHIP_LOSS_INTENSITY_MSE) {
    double r = predicted - measured;
    return r * r;
  }
  if (loss == FPMHIP_LOSS_POISSON_NLL) {
    double mean = fmax(predicted, 1e-12);
    return mean - fmax(measured, 0.0) * log(mean);
  }
  double r = sqrt(fmax(predicted, 0.0)) - sqrt(fmax(measured, 0.0));
  if (loss == FPMHIP_LOSS_HUBER_AMPLITUDE) {
    double a = fabs(r);
    return a <= 1.0 ? 0.5 * r * r : a - 0.5;
  }
  return r * r;
}

template <int BLOCK>
__device__ inline double block_sum(double v, double* sh) {
  // fixed-order tree reduction: deterministic for a given launch shape
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = sh[t] + sh[t + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

using D = cx<double>;

// MIXED_F32 sessions (float2 buffers) transform the field change psi' - psi.
template <class V>
constexpr bool kMixed = false;
template <>
constexpr bool kMixed<float2> = true;

// bilinear extract (crop.rs:129-167) of one source, times pupil, ifftshift
// (forward.rs:538-549) into the field buffer.
template <class V>
__global__ void k_extract(const double2* __restrict__ spec, uint32_t rw, const SrcParam* __restrict__ src,
                          uint64_t source, uint32_t h, uint32_t w, const V* __restrict__ pupil,
                          V* __restrict__ patch, V* __restrict__ field) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= h * w) return;
  const SrcParam sp = src[source];
  const uint32_t r = p / w, c = p % w;
  double2 v;
  if (sp.zero) {
    v = spec[(sp.sr + r) * rw + sp.sc + c];
  } else {
    const uint64_t lr = sp.ar.lower + r, ur = sp.ar.upper + r, lc = sp.ac.lower + c, uc = sp.ac.upper + c;
    const double wll = sp.ar.wl * sp.ac.wl, wlu = sp.ar.wl * sp.ac.wu, wul = sp.ar.wu * sp.ac.wl,
                 wuu = sp.ar.wu * sp.ac.wu;
    v = D::mulr(spec[lr * rw + lc], wll);
    v = D::add(v, D::mulr(spec[lr * rw + uc], wlu));
    v = D::add(v, D::mulr(spec[ur * rw + lc], wul));
    v = D::add(v, D::mulr(spec[ur * rw + uc], wuu));
  }
  const V pv = from_d2<V>(v);
  patch[p] = pv;
  const double2 e = D::mul(to_d2(pv), to_d2(pupil[p]));
  field[((r + (h + 1) / 2) % h) * w + (c + (w + 1) / 2) % w] = from_d2<V>(e);
}

// predicted intensity, objective terms and amplitude ratio (common.rs:106-196);
// the first mode's difference field (:219-228) is written in place.
// The frame's gain is read from the state's device gains (EPRY recovery may
// have just updated it); an invalid gain raises flag[3] (common.rs:139-143).
template <class V>
__global__ void __launch_bounds__(kBlock) k_project(V* __restrict__ fields, uint32_t K, const double* __restrict__ mode_w,
                                                    double sw_sum, uint32_t N, const void* __restrict__ frames,
                                                    int is_u16, uint64_t frame, const uint8_t* __restrict__ masks,
                                                    const double* __restrict__ gain_p, const double* __restrict__ bg,
                                                    uint64_t bg_len, double eps, int loss, double* __restrict__ pred_out,
                                                    V* __restrict__ proj_out, double* __restrict__ partials,
                                                    int* __restrict__ flag) {
  __shared__ double sh[kBlock];
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const double gain = *gain_p;
  if (p == 0 && !(isfinite(gain) && gain > 0.0)) flag[3] = 1;
  double lsum = 0.0, lcnt = 0.0;
  if (p < N) {
    double pred = 0.0;
    for (uint32_t k = 0; k < K; ++k) {
      const double2 f = to_d2(fields[(uint64_t)k * N + p]);
      pred += (mode_w ? mode_w[k] : 1.0) * D::norm2(f);
    }
    const bool masked = masks && masks[frame * N + p] == 0;
    double2 proj = make_double2(1.0, 0.0);
    if (!masked) {
      const double b = bg_value(bg, bg_len, N, frame, p);
      const double y = meas_value(frames, is_u16, frame * N + p);
      lsum = point_loss(gain * pred + b, y, loss);
      lcnt = 1.0;
      const double target = fmax((y - b) / gain, 0.0);
      proj = pred > eps ? make_double2(sqrt(target / pred), 0.0) : make_double2(0.0, sqrt(target / sw_sum));
    }
    pred_out[p] = pred;
    if (K > 1) proj_out[p] = from_d2<V>(proj);
    const double2 f0 = to_d2(fields[p]);
    double2 d = masked ? f0 : (pred > eps ? D::mulr(f0, proj.x) : make_double2(proj.y, 0.0));
    // MIXED_F32 keeps psi' - psi: FFT(psi' - psi)/N equals FFT(psi')/N - patch P
    // by linearity and avoids cancelling two nearly equal FP32 spectra.
    if (kMixed<V>) d = D::sub(d, f0);
    fields[p] = from_d2<V>(d);
  }
  const double s = block_sum<kBlock>(lsum, sh);
  const double c = block_sum<kBlock>(lcnt, sh);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = s;
    partials[2 * blockIdx.x + 1] = c;
  }
}

// Pixel sums of the calibration updates over the unmasked pixels of one
// frame, from the intrinsic prediction sum_k w_k |psi_k|^2 of the current
// fields (common.rs:390-460): mode 0 (update_gain) sums pred * max(y - bg, 0)
// and pred^2; mode 1 (update_background) sums y - gain * pred - bg and the
// pixel count. One partial pair per block, summed on the host in block order.
template <class V>
__global__ void __launch_bounds__(kBlock)
    k_calibration_sums(const V* __restrict__ fields, uint32_t K, const double* __restrict__ mode_w, uint32_t N,
                       const void* __restrict__ frames, int is_u16, uint64_t frame, const uint8_t* __restrict__ masks,
                       const double* __restrict__ bg, uint64_t bg_len, const double* __restrict__ gain_p, int mode,
                       double* __restrict__ partials) {
  __shared__ double sh[kBlock];
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const double gain = *gain_p;
  double a = 0.0, b = 0.0;
  if (p < N && !(masks && masks[frame * N + p] == 0)) {
    double pred = 0.0;
    for (uint32_t k = 0; k < K; ++k) pred += (mode_w ? mode_w[k] : 1.0) * D::norm2(to_d2(fields[(uint64_t)k * N + p]));
    const double bgv = bg_value(bg, bg_len, N, frame, p);
    const double y = meas_value(frames, is_u16, frame * N + p);
    if (mode == 0) {
      a = pred * fmax(y - bgv, 0.0);
      b = pred * pred;
    } else {
      a = y - gain * pred - bgv;
      b = 1.0;
    }
  }
  const double sa = block_sum<kBlock>(a, sh);
  const double sb = block_sum<kBlock>(b, sh);
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = sa;
    partials[2 * blockIdx.x + 1] = sb;
  }
}

// clamp as f64::clamp: NaN stays NaN
__device__ inline double clamp_nan(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

// The calibration step of one frame from the block partials of
// k_calibration_sums, summed in block order by one thread (common.rs:390-489):
// mode 0 (update_gain): when sum pred^2 > eps, gain <- clamp((1 - step) gain +
// step clamp(sum pred max(y - bg, 0) / sum pred^2)), written to the state's
// gains, and present[0] = 1 (the state now holds frame gains); mode 1
// (update_background): out[0] = step * residual sum / unmasked pixel count.
__global__ void k_calibration_apply(const double* __restrict__ partials, uint32_t nblk, int mode,
                                    double* __restrict__ gain, double eps, double step, double lo, double hi,
                                    double* __restrict__ out, double* __restrict__ present) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double a = 0.0, b = 0.0;
  for (uint32_t k = 0; k < nblk; ++k) {
    a += partials[2 * k];
    b += partials[2 * k + 1];
  }
  if (mode == 0) {
    if (b > eps) {
      const double est = clamp_nan(a / b, lo, hi);
      *gain = clamp_nan((1.0 - step) * *gain + step * est, lo, hi);
      *present = 1.0;
    }
  } else {
    *out = step * a / b;  // launched only for frames with unmasked pixels (b > 0)
  }
}

__global__ void k_background_shift(double* __restrict__ bg, uint64_t N, const double* __restrict__ correction,
                                   double lo, double hi) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < N) bg[p] = clamp_nan(bg[p] + *correction, lo, hi);
}

// in place: entries [0, N) hold the shared background (or nothing)
__global__ void k_background_expand(double* __restrict__ bg, uint64_t src_len, uint64_t N, uint64_t total) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total && i >= src_len) bg[i] = src_len == 0 ? 0.0 : bg[i % N];
}

// difference field of a further multiplexed mode (common.rs:219-228)
template <class V>
__global__ void k_difference(V* __restrict__ field, const V* __restrict__ proj, const double* __restrict__ pred,
                             uint32_t N, const uint8_t* __restrict__ mask, double eps) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  const double2 f = to_d2(field[p]);
  const double2 pr = to_d2(proj[p]);
  const bool masked = mask && mask[p] == 0;
  double2 d = masked ? f : (pred[p] > eps ? D::mulr(f, pr.x) : make_double2(pr.y, 0.0));
  if (kMixed<V>) d = D::sub(d, f);
  field[p] = from_d2<V>(d);
}

// fftshift of the projected spectrum, exit-wave difference, object update
// U = conj(P) diff / denominator and the EPRY pupil gradient (common.rs:235-285)
template <class V>
__global__ void k_update(const V* __restrict__ fftdiff, const V* __restrict__ patch, const V* __restrict__ pupil,
                         uint32_t h, uint32_t w, int denom, double alpha, double eps,
                         const double* __restrict__ scal, double nsw, int pupil_step, int first_mode,
                         V* __restrict__ U, V* __restrict__ grad) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= h * w) return;
  const uint32_t r = p / w, c = p % w;
  const double2 ps = to_d2(fftdiff[((r + (h + 1) / 2) % h) * w + (c + (w + 1) / 2) % w]);
  const double2 pa = to_d2(patch[p]);
  const double2 P = to_d2(pupil[p]);
  const double2 diff = kMixed<V> ? ps : D::sub(ps, D::mul(pa, P));
  const double local = D::norm2(P);
  const double maxp = fmax(scal[0], eps);
  double den;
  if (denom == FPMHIP_DENOMINATOR_LOCAL)
    den = local + eps;
  else if (denom == FPMHIP_DENOMINATOR_GLOBAL)
    den = maxp + eps;
  else
    den = (1.0 - alpha) * local + alpha * maxp + eps;
  U[p] = from_d2<V>(D::divr(D::mul(D::conj(P), diff), den));
  if (pupil_step) {
    const double maxo = fmax(scal[1], eps);
    const double2 g = D::divr(D::mul(D::scale(nsw, D::conj(pa)), diff), maxo + eps);
    const double2 prev = first_mode ? make_double2(0.0, 0.0) : to_d2(grad[p]);
    grad[p] = from_d2<V>(D::add(prev, g));
  }
}

// exact adjoint of the bilinear extract (crop.rs:188-256), written as a
// gather over the destination window: each destination pixel adds its taps
// in the reference's order (source row-major; taps LL, LU, UL, UU), so no
// two threads write the same element.
template <class V>
__global__ void k_insert(double2* __restrict__ spec, uint32_t rw, const SrcParam* __restrict__ src, uint64_t source,
                         uint32_t h, uint32_t w, const V* __restrict__ U, double scale) {
  const SrcParam sp = src[source];
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (sp.zero) {
    if (q >= h * w) return;
    const uint32_t r = q / w, c = q % w;
    double2& d = spec[(sp.sr + r) * rw + sp.sc + c];
    d = D::add(d, D::scale(scale, to_d2(U[q])));
    return;
  }
  const bool fr = sp.ar.upper != sp.ar.lower, fc = sp.ac.upper != sp.ac.lower;
  const uint32_t wh = h + (fr ? 1 : 0), ww = w + (fc ? 1 : 0);
  if (q >= wh * ww) return;
  const int a = (int)(q / ww), b = (int)(q % ww);
  double2& dst = spec[(sp.ar.lower + a) * (uint64_t)rw + sp.ac.lower + b];
  double2 acc = dst;
  const double wll = sp.ar.wl * sp.ac.wl, wlu = sp.ar.wl * sp.ac.wu, wul = sp.ar.wu * sp.ac.wl,
               wuu = sp.ar.wu * sp.ac.wu;
  auto tap = [&](int sr_, int sc_, double wt) {
    if (sr_ >= 0 && sr_ < (int)h && sc_ >= 0 && sc_ < (int)w) {
      const double2 v = D::scale(scale, to_d2(U[sr_ * w + sc_]));
      acc = D::add(acc, D::mulr(v, wt));
    }
  };
  if (fr && fc) {
    tap(a - 1, b - 1, wuu);
    tap(a - 1, b, wul);
    tap(a, b - 1, wlu);
    tap(a, b, wll);
  } else if (fr) {
    tap(a - 1, b, wul);
    tap(a, b, wll);
  } else if (fc) {
    tap(a, b - 1, wlu);
    tap(a, b, wll);
  } else {
    tap(a, b, wll);
  }
  dst = acc;
}

// max |v|^2 over n values by one 1024-thread block (max is order independent)
template <class V>
__device__ inline void block_max_abs2(const V* __restrict__ v, uint64_t n, double* __restrict__ out, double* sh) {
  double m = 0.0;
  for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) m = fmax(m, D::norm2(to_d2(v[i])));
  sh[threadIdx.x] = m;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

template <class V>
__global__ void __launch_bounds__(1024) k_max_abs2(const V* __restrict__ v, uint64_t n, double* __restrict__ out) {
  __shared__ double sh[1024];
  block_max_abs2(v, n, out, sh);
}

// P += w_f beta g; support projection (common.rs:287-300, pupil.rs:171-182).
// With an FP64 master (MIXED_F32 pupil recovery), the step accumulates into
// the master and the pupil becomes its rounded copy.
template <class V>
__global__ void k_pupil_step(V* __restrict__ pupil, double2* __restrict__ master, const V* __restrict__ grad,
                             double step, const uint8_t* __restrict__ support, int constrain, uint32_t N) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  double2 P = D::add(master ? master[p] : to_d2(pupil[p]), D::scale(step, to_d2(grad[p])));
  if (constrain && support[p] == 0) P = make_double2(0.0, 0.0);
  if (master) master[p] = P;
  pupil[p] = from_d2<V>(P);
}

// apply_momentum_event (common.rs:355-387)
// The velocity and anchor are stored as W (complex128, or complex64 in
// MIXED_F32 sessions); the recurrence is evaluated in FP64 either way.
template <class W>
__global__ void k_momentum(double2* __restrict__ obj, W* __restrict__ vel, W* __restrict__ anc, uint64_t n,
                           double friction, double feedback, int* __restrict__ flag) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double2 o = obj[i];
  const double2 nv = D::add(D::scale(friction, to_d2(vel[i])), D::sub(o, to_d2(anc[i])));
  const double2 no = feedback == 0.0 ? o : D::add(o, D::scale(feedback, nv));
  if (!isfinite(nv.x) || !isfinite(nv.y) || !isfinite(no.x) || !isfinite(no.y)) flag[0] = 1;
  vel[i] = from_d2<W>(nv);
  obj[i] = no;
  anc[i] = from_d2<W>(no);
}

// Clears the session's flags and its per-sweep calibration marker.
__global__ void k_flag_reset(int* flag, double* gains_present) {
  if (threadIdx.x < 4) flag[threadIdx.x] = 0;
  if (threadIdx.x == 0) *gains_present = 0.0;
}

// weighted mean measured amplitude per low-resolution pixel (state.rs:288-313)
// Per-pixel partial sums of the weighted mean amplitude over the frames of
// chunk y ([y * chunk, (y + 1) * chunk), in order): part[(y * N + p)]
// = (sum w_f sqrt(max((y - bg) / g, 0)), sum w_f) over the chunk's
// positive-weight, unmasked frames (state.rs:288-313).
__device__ inline void average_part(const void* __restrict__ frames, int is_u16, const double* __restrict__ weights,
                                    const uint8_t* __restrict__ masks, const double* __restrict__ gains,
                                    const double* __restrict__ bg, uint64_t bg_len, uint64_t F, uint32_t N,
                                    uint64_t chunk, uint32_t y, uint32_t p, double2* __restrict__ part) {
  const uint64_t fa = (uint64_t)y * chunk, fb = fa + chunk < F ? fa + chunk : F;
  double a = 0.0, w = 0.0;
  // frames in groups of G: the group's loads are issued before the
  // in-order accumulation uses them
  constexpr uint64_t G = 8;
  for (uint64_t f0 = fa; f0 < fb; f0 += G) {
    double yv[G];
    uint8_t keep[G];
#pragma unroll
    for (uint64_t k = 0; k < G; ++k) {
      const uint64_t f = f0 + k;
      yv[k] = f < fb ? meas_value(frames, is_u16, f * N + p) : 0.0;
      keep[k] = f < fb && (!masks || masks[f * N + p] != 0);
    }
#pragma unroll
    for (uint64_t k = 0; k < G; ++k) {
      const uint64_t f = f0 + k;
      if (f >= fb || !keep[k]) continue;
      const double fw = weights[f];
      if (fw == 0.0) continue;
      const double b = bg_value(bg, bg_len, N, f, p);
      const double amp = sqrt(fmax((yv[k] - b) / gains[f], 0.0));
      a += fw * amp;
      w += fw;
    }
  }
  part[(uint64_t)y * N + p] = make_double2(a, w);
}

__global__ void k_init_average_part(const void* __restrict__ frames, int is_u16, const double* __restrict__ weights,
                                    const uint8_t* __restrict__ masks, const double* __restrict__ gains,
                                    const double* __restrict__ bg, uint64_t bg_len, uint64_t F, uint32_t N,
                                    uint64_t chunk, double2* __restrict__ part) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  average_part(frames, is_u16, weights, masks, gains, bg, bg_len, F, N, chunk, blockIdx.y, p, part);
}

// The chunk partials summed in chunk order.
__device__ inline void average_sum(const double2* __restrict__ part, uint32_t chunks, uint32_t N, uint32_t p,
                                   double* __restrict__ avg, double* __restrict__ aw) {
  double a = 0.0, w = 0.0;
  for (uint32_t c = 0; c < chunks; ++c) {
    const double2 v = part[(uint64_t)c * N + p];
    a += v.x;
    w += v.y;
  }
  avg[p] = a;
  aw[p] = w;
}

__global__ void k_init_average_sum(const double2* __restrict__ part, uint32_t chunks, uint32_t N,
                                   double* __restrict__ avg, double* __restrict__ aw) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= N) return;
  average_sum(part, chunks, N, p, avg, aw);
}

// fallback amplitude (state.rs:314-326) into *fallback_out, by one
// 1024-thread block
__device__ inline void block_fallback(const double* __restrict__ avg, const double* __restrict__ aw, uint32_t N,
                                      double* __restrict__ fallback_out, int* __restrict__ flag, double* sa,
                                      double* sw) {
  double ta = 0.0, tw = 0.0;
  for (uint32_t p = threadIdx.x; p < N; p += blockDim.x) {
    ta += avg[p];
    tw += aw[p];
  }
  sa[threadIdx.x] = ta;
  sw[threadIdx.x] = tw;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sa[threadIdx.x] += sa[threadIdx.x + s];
      sw[threadIdx.x] += sw[threadIdx.x + s];
    }
    __syncthreads();
  }
  const double total_w = sw[0];
  if (threadIdx.x == 0) {
    *fallback_out = total_w > 0.0 ? sa[0] / total_w : 0.0;
    if (total_w == 0.0) flag[1] = 1;
  }
}

__global__ void __launch_bounds__(1024) k_init_fallback(const double* __restrict__ avg, const double* __restrict__ aw,
                                                        uint32_t N, double* __restrict__ fallback_out,
                                                        int* __restrict__ flag) {
  __shared__ double sa[1024], sw[1024];
  block_fallback(avg, aw, N, fallback_out, flag, sa, sw);
}

// nearest-neighbour upsample of the mean amplitude (state.rs:327-336)
__device__ inline double2 object_value(const double* __restrict__ avg, const double* __restrict__ aw, uint32_t nh,
                                       uint32_t nw, uint32_t rh, uint32_t rw, double fallback, uint64_t i) {
  const uint32_t r = (uint32_t)(i / rw), c = (uint32_t)(i % rw);
  const uint32_t lr = (uint32_t)((uint64_t)r * nh / rh), lc = (uint32_t)((uint64_t)c * nw / rw);
  const uint32_t lp = lr * nw + lc;
  const double v = aw[lp] > 0.0 ? avg[lp] / aw[lp] : fallback;
  return make_double2(v, 0.0);
}

__global__ void k_init_object(const double* __restrict__ avg, const double* __restrict__ aw, uint32_t nh, uint32_t nw,
                              uint32_t rh, uint32_t rw, const double* __restrict__ fallback,
                              double2* __restrict__ object) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)rh * rw) return;
  object[i] = object_value(avg, aw, nh, nw, rh, rw, *fallback, i);
}

// destination of element i under a cyclic shift by (rs, cs) of an h x w array
__device__ inline uint64_t shifted(uint64_t i, uint32_t h, uint32_t w, uint32_t rs, uint32_t cs) {
  const uint32_t r = (uint32_t)(i / w), c = (uint32_t)(i % w);
  return (uint64_t)((r + rs) % h) * w + (c + cs) % w;
}

__global__ void k_shift(const double2* __restrict__ src, double2* __restrict__ dst, uint32_t h, uint32_t w,
                        uint32_t rs, uint32_t cs) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)h * w) return;
  dst[shifted(i, h, w, rs, cs)] = src[i];
}

// ---- batched forms over InitTile arrays (grid y or z = tile) ----

__global__ void k_binit_reset(const InitTile* __restrict__ tiles) {
  const InitTile& T = tiles[blockIdx.x];
  if (threadIdx.x < 4) T.flag[threadIdx.x] = 0;
  if (threadIdx.x == 0 && T.counter) *T.counter = 0;
}

__global__ void k_binit_average_part(const InitTile* __restrict__ tiles) {
  const InitTile& T = tiles[blockIdx.z];
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (blockIdx.y >= T.chunks || p >= T.N) return;
  average_part(T.frames, (int)T.is_u16, T.weights, T.masks, T.model_gains, T.model_bg, T.bg_len, T.F, T.N, T.chunk,
               blockIdx.y, p, T.part);
}

__global__ void k_binit_average_sum(const InitTile* __restrict__ tiles) {
  const InitTile& T = tiles[blockIdx.y];
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= T.N) return;
  average_sum(T.part, T.chunks, T.N, p, T.avg, T.aw);
}

__global__ void __launch_bounds__(1024) k_binit_fallback(const InitTile* __restrict__ tiles) {
  __shared__ double sa[1024], sw[1024];
  const InitTile& T = tiles[blockIdx.x];
  block_fallback(T.avg, T.aw, T.N, T.scal + 2, T.flag, sa, sw);
}

__global__ void k_binit_object(const InitTile* __restrict__ tiles) {
  const InitTile& T = tiles[blockIdx.y];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)T.rh * T.rw) return;
  T.slot[i] = object_value(T.avg, T.aw, T.nh, T.nw, T.rh, T.rw, T.scal[2], i);
}

// pupil (complex64), gains and background of the state from the model
__global__ void k_binit_pupil(const InitTile* __restrict__ tiles) {
  const InitTile& T = tiles[blockIdx.y];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < T.N) T.pupil[i] = from_d2<float2>(T.model_pupil[i]);
  if (i < T.N && T.pupil64) T.pupil64[i] = T.model_pupil[i];
  if (i < T.F) T.gains[i] = T.model_gains[i];
  if (i < T.bg_len) T.bg[i] = T.model_bg[i];
}

__global__ void __launch_bounds__(1024) k_binit_max_pupil(const InitTile* __restrict__ tiles) {
  __shared__ double sh[1024];
  const InitTile& T = tiles[blockIdx.x];
  block_max_abs2(T.pupil, T.N, T.scal, sh);
}

__global__ void k_binit_finish(const InitTile* __restrict__ tiles) {
  const InitTile& T = tiles[blockIdx.y];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)T.rh * T.rw) return;
  const double2 v = T.slot[i];
  const uint64_t j = shifted(i, T.rh, T.rw, T.rh / 2, T.rw / 2);
  T.spec[j] = v;
  if (T.vel) {
    T.anc[j] = from_d2<float2>(v);
    T.vel[j] = make_float2(0.0f, 0.0f);
  }
}

__global__ void k_binit_flags(const InitTile* __restrict__ tiles, uint32_t count, int* __restrict__ flags) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  flags[2 * t] = tiles[t].flag[0];
  flags[2 * t + 1] = tiles[t].flag[1];
}

__global__ void k_bobject_shift(const InitTile* __restrict__ tiles) {
  const InitTile& T = tiles[blockIdx.y];
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (uint64_t)T.rh * T.rw) return;
  T.slot[shifted(i, T.rh, T.rw, (T.rh + 1) / 2, (T.rw + 1) / 2)] = T.spec[i];
}

template <class V>
__global__ void k_fill_zero(V* dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = V{};
}

template <class V>
__global__ void k_convert_from_d2(const double2* __restrict__ src, V* __restrict__ dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = from_d2<V>(src[i]);
}
template <class V>
__global__ void k_convert_to_d2(const V* __restrict__ src, double2* __restrict__ dst, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = to_d2(src[i]);
}

// forward stack: every source's field (extract x pupil, ifftshift) ...
__global__ void k_extract_all(const double2* __restrict__ spec, uint32_t rw, const SrcParam* __restrict__ src,
                              uint32_t h, uint32_t w, const double2* __restrict__ pupil,
                              double2* __restrict__ fields) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t s = blockIdx.y;
  if (p >= h * w) return;
  const SrcParam sp = src[s];
  const uint32_t r = p / w, c = p % w;
  double2 v;
  if (sp.zero) {
    v = spec[(sp.sr + r) * rw + sp.sc + c];
  } else {
    const uint64_t lr = sp.ar.lower + r, ur = sp.ar.upper + r, lc = sp.ac.lower + c, uc = sp.ac.upper + c;
    const double wll = sp.ar.wl * sp.ac.wl, wlu = sp.ar.wl * sp.ac.wu, wul = sp.ar.wu * sp.ac.wl,
                 wuu = sp.ar.wu * sp.ac.wu;
    v = D::mulr(spec[lr * rw + lc], wll);
    v = D::add(v, D::mulr(spec[lr * rw + uc], wlu));
    v = D::add(v, D::mulr(spec[ur * rw + lc], wul));
    v = D::add(v, D::mulr(spec[ur * rw + uc], wuu));
  }
  fields[s * h * w + ((r + (h + 1) / 2) % h) * w + (c + (w + 1) / 2) % w] = D::mul(v, pupil[p]);
}

// ... then |field|^2 per frame (incoherent sum when multiplexed), x gain + background
__global__ void k_intensity(const double2* __restrict__ fields, uint32_t N, const uint32_t* __restrict__ mux_ptr,
                            const uint32_t* __restrict__ mux_src, const double* __restrict__ mux_w,
                            const double* __restrict__ gains, const double* __restrict__ bg, uint64_t bg_len,
                            double* __restrict__ out) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t f = blockIdx.y;
  if (p >= N) return;
  double v;
  if (mux_ptr) {
    v = 0.0;
    for (uint32_t e = mux_ptr[f]; e < mux_ptr[f + 1]; ++e) v += mux_w[e] * D::norm2(fields[(uint64_t)mux_src[e] * N + p]);
  } else {
    v = D::norm2(fields[f * N + p]);
  }
  out[f * N + p] = gains[f] * v + bg_value(bg, bg_len, N, f, p);
}

inline uint32_t blocks_for(uint64_t n, uint32_t b = kBlock) { return (uint32_t)((n + b - 1) / b); }

}  // namespace

// ---- launchers ------------------------------------------------------------------

#define DISPATCH_V(s, EXPR)                          \
  do {                                               \
    if ((s).precision == FPMHIP_PLAIN_C128) {        \
      using V = double2;                             \
      EXPR;                                          \
    } else {                                         \
      using V = float2;                              \
      EXPR;                                          \
    }                                                \
  } while (0)

void launch_extract(const fpmhip_session& s, uint64_t source, uint32_t mode) {
  const Model& m = *s.model;
  const uint64_t N = m.N();
  DISPATCH_V(s, hipLaunchKernelGGL(k_extract<V>, dim3(blocks_for(N)), dim3(kBlock), 0, s.ctx->stream,
                                   s.b.spec.as<double2>(), m.rw, m.d_src.as<SrcParam>(), source, m.nh,
                                   m.nw, s.b.pupil.as<V>(), s.b.patch.as<V>() + mode * N,
                                   s.b.field.as<V>() + mode * N));
  HIP_CHECK(hipGetLastError());
}

void launch_project(const fpmhip_session& s, uint64_t frame, uint32_t K, const double* mode_w_dev, double sw_sum,
                    uint64_t slot) {
  const Model& m = *s.model;
  const Measurements& me = *s.meas;
  const uint64_t N = m.N();
  DISPATCH_V(s, hipLaunchKernelGGL(k_project<V>, dim3(s.nblk), dim3(kBlock), 0, s.ctx->stream, s.b.field.as<V>(),
                                   K, mode_w_dev, sw_sum, (uint32_t)N, me.d_frames.ptr, me.is_u16 ? 1 : 0, frame,
                                   me.d_masks.as<uint8_t>(), s.b.gains.as<double>() + frame, s.b.bg.as<double>(),
                                   s.bg_len, s.cfg.epsilon, s.cfg.loss, s.b.pred.as<double>(), s.b.proj.as<V>(),
                                   s.b.partials.as<double>() + slot * s.nblk * 2, s.b.flag.as<int>()));
  HIP_CHECK(hipGetLastError());
}

void launch_calibration(const fpmhip_session& s, uint64_t frame, uint32_t K, const double* mode_w_dev, int mode) {
  const Model& m = *s.model;
  const Measurements& me = *s.meas;
  double* gain = s.b.gains.as<double>() + frame;
  DISPATCH_V(s, hipLaunchKernelGGL(k_calibration_sums<V>, dim3(s.nblk), dim3(kBlock), 0, s.ctx->stream,
                                   s.b.field.as<V>(), K, mode_w_dev, (uint32_t)m.N(), me.d_frames.ptr,
                                   me.is_u16 ? 1 : 0, frame, me.d_masks.as<uint8_t>(), s.b.bg.as<double>(),
                                   s.bg_len, gain, mode, s.b.calib.as<double>()));
  HIP_CHECK(hipGetLastError());
  const Config& c = s.cfg;
  double* scal = s.b.scal.as<double>();
  hipLaunchKernelGGL(k_calibration_apply, dim3(1), dim3(64), 0, s.ctx->stream, s.b.calib.as<double>(), s.nblk, mode,
                     gain, c.epsilon, mode == 0 ? c.gain_step : c.bg_step, mode == 0 ? c.gain_min : c.bg_min,
                     mode == 0 ? c.gain_max : c.bg_max, scal + kScalBackgroundStep, scal + kScalGainsPresent);
  HIP_CHECK(hipGetLastError());
}

void launch_background_shift(const fpmhip_session& s, uint64_t frame, double lo, double hi) {
  const uint64_t N = s.model->N();
  hipLaunchKernelGGL(k_background_shift, dim3(blocks_for(N)), dim3(kBlock), 0, s.ctx->stream,
                     s.b.bg.as<double>() + frame * N, N, s.b.scal.as<double>() + kScalBackgroundStep, lo, hi);
  HIP_CHECK(hipGetLastError());
}

void launch_background_expand(const fpmhip_session& s) {
  const uint64_t N = s.model->N(), total = N * s.model->F;
  hipLaunchKernelGGL(k_background_expand, dim3(blocks_for(total)), dim3(kBlock), 0, s.ctx->stream,
                     s.b.bg.as<double>(), s.bg_len, N, total);
  HIP_CHECK(hipGetLastError());
}

void launch_update(const fpmhip_session& s, uint32_t mode, double nsw, bool first_mode) {
  const Model& m = *s.model;
  const uint64_t N = m.N();
  DISPATCH_V(s, hipLaunchKernelGGL(k_update<V>, dim3(blocks_for(N)), dim3(kBlock), 0, s.ctx->stream,
                                   s.b.field.as<V>() + mode * N, s.b.patch.as<V>() + mode * N, s.b.pupil.as<V>(),
                                   m.nh, m.nw, s.cfg.denom, s.cfg.alpha, s.cfg.epsilon, s.b.scal.as<double>(), nsw,
                                   s.cfg.has_pupil_step ? 1 : 0, first_mode ? 1 : 0, s.b.U.as<V>(),
                                   s.b.grad.as<V>()));
  HIP_CHECK(hipGetLastError());
}

void launch_insert(const fpmhip_session& s, uint64_t source, uint32_t mode, double scale) {
  (void)mode;
  const Model& m = *s.model;
  const uint64_t window = (uint64_t)(m.nh + 1) * (m.nw + 1);
  DISPATCH_V(s, hipLaunchKernelGGL(k_insert<V>, dim3(blocks_for(window)), dim3(kBlock), 0, s.ctx->stream,
                                   s.b.spec.as<double2>(), m.rw, m.d_src.as<SrcParam>(), source, m.nh,
                                   m.nw, s.b.U.as<V>(), scale));
  HIP_CHECK(hipGetLastError());
}

void launch_max_abs2(const fpmhip_session& s, const void* v, bool f64v, uint64_t n, uint32_t slot) {
  if (f64v)
    hipLaunchKernelGGL(k_max_abs2<double2>, dim3(1), dim3(1024), 0, s.ctx->stream, static_cast<const double2*>(v),
                       n, s.b.scal.as<double>() + slot);
  else
    hipLaunchKernelGGL(k_max_abs2<float2>, dim3(1), dim3(1024), 0, s.ctx->stream, static_cast<const float2*>(v), n,
                       s.b.scal.as<double>() + slot);
  HIP_CHECK(hipGetLastError());
}

void launch_pupil_step(const fpmhip_session& s, double step) {
  const Model& m = *s.model;
  const uint64_t N = m.N();
  DISPATCH_V(s, hipLaunchKernelGGL(k_pupil_step<V>, dim3(blocks_for(N)), dim3(kBlock), 0, s.ctx->stream,
                                   s.b.pupil.as<V>(), s.b.pupil64.as<double2>(), s.b.grad.as<V>(), step,
                                   m.d_support.as<uint8_t>(),
                                   s.cfg.constrain ? 1 : 0, (uint32_t)N));
  HIP_CHECK(hipGetLastError());
}

void launch_momentum(const fpmhip_session& s) {
  const uint64_t M = s.model->M();
  DISPATCH_V(s, hipLaunchKernelGGL(k_momentum<V>, dim3(blocks_for(M)), dim3(kBlock), 0, s.ctx->stream,
                                   s.b.spec.as<double2>(), s.b.vel.as<V>(), s.b.anc.as<V>(), M, s.cfg.friction,
                                   s.cfg.feedback, s.b.flag.as<int>()));
  HIP_CHECK(hipGetLastError());
}

void set_session_pupil(const fpmhip_session& s, const double2* src) {
  const uint64_t N = s.model->N();
  if (s.b.pupil64.ptr)
    HIP_CHECK(hipMemcpyAsync(s.b.pupil64.ptr, src, N * sizeof(double2), hipMemcpyDeviceToDevice, s.ctx->stream));
  launch_convert_pupil(src, s.b.pupil.ptr, s.precision == FPMHIP_MIXED_F32, N, s.ctx->stream);
}

void launch_flag_reset(const fpmhip_session& s) {
  hipLaunchKernelGGL(k_flag_reset, dim3(1), dim3(64), 0, s.ctx->stream, s.b.flag.as<int>(),
                     s.b.scal.as<double>() + kScalGainsPresent);
  HIP_CHECK(hipGetLastError());
}

// frames in up to 8 ordered chunks (partials in the init scratch work_m,
// max(M, N) complex128 slots), so that the pass fills the device
uint32_t init_chunks(const fpmhip_session& s) {
  const Model& m = *s.model;
  const uint64_t slots = s.b.work_m.bytes / sizeof(double2);
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({8, slots / m.N(), m.F ? m.F : 1}));
}

void launch_init(const fpmhip_session& s) {
  const Model& m = *s.model;
  const Measurements& me = *s.meas;
  const uint64_t N = m.N();
  hipStream_t st = s.ctx->stream;
  const uint32_t chunks = init_chunks(s);
  const uint64_t chunk = (m.F + chunks - 1) / chunks;
  hipLaunchKernelGGL(k_init_average_part, dim3(blocks_for(N), chunks), dim3(kBlock), 0, st, me.d_frames.ptr,
                     me.is_u16 ? 1 : 0, me.d_weights.as<double>(), me.d_masks.as<uint8_t>(), m.d_gains.as<double>(),
                     m.d_background.as<double>(), (uint64_t)m.background.size(), m.F, (uint32_t)N, chunk,
                     s.b.work_m.as<double2>());
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_init_average_sum, dim3(blocks_for(N)), dim3(kBlock), 0, st, s.b.work_m.as<double2>(), chunks,
                     (uint32_t)N, s.b.avg.as<double>(), s.b.aw.as<double>());
  HIP_CHECK(hipGetLastError());
  // scal[2]: the fallback amplitude
  hipLaunchKernelGGL(k_init_fallback, dim3(1), dim3(1024), 0, st, s.b.avg.as<double>(), s.b.aw.as<double>(),
                     (uint32_t)N, s.b.scal.as<double>() + 2, s.b.flag.as<int>());
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_init_object, dim3(blocks_for(m.M())), dim3(kBlock), 0, st, s.b.avg.as<double>(),
                     s.b.aw.as<double>(), m.nh, m.nw, m.rh, m.rw, s.b.scal.as<double>() + 2, s.b.work_m.as<double2>());
  HIP_CHECK(hipGetLastError());
  fft2_f64(*m.fft_high, s.b.work_m.as<double2>(), 1, false, st);
  launch_shift(s.b.work_m.as<double2>(), s.b.spec.as<double2>(), m.rh, m.rw, m.rh / 2, m.rw / 2, st);
}

void launch_init_batch_pre(const InitTile* tiles, uint32_t count, uint32_t nmax, uint32_t mmax, uint64_t cmax,
                           hipStream_t st) {
  if (!count) return;
  hipLaunchKernelGGL(k_binit_reset, dim3(count), dim3(64), 0, st, tiles);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_binit_average_part, dim3(blocks_for(nmax), 8, count), dim3(kBlock), 0, st, tiles);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_binit_average_sum, dim3(blocks_for(nmax), count), dim3(kBlock), 0, st, tiles);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_binit_fallback, dim3(count), dim3(1024), 0, st, tiles);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_binit_object, dim3(blocks_for(mmax), count), dim3(kBlock), 0, st, tiles);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_binit_pupil, dim3(blocks_for(cmax), count), dim3(kBlock), 0, st, tiles);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_binit_max_pupil, dim3(count), dim3(1024), 0, st, tiles);
  HIP_CHECK(hipGetLastError());
}

void launch_init_batch_post(const InitTile* tiles, uint32_t count, uint32_t mmax, hipStream_t st) {
  if (!count) return;
  hipLaunchKernelGGL(k_binit_finish, dim3(blocks_for(mmax), count), dim3(kBlock), 0, st, tiles);
  HIP_CHECK(hipGetLastError());
}

void launch_init_batch_flags(const InitTile* tiles, uint32_t count, int* flags, hipStream_t st) {
  if (!count) return;
  hipLaunchKernelGGL(k_binit_flags, dim3(blocks_for(count)), dim3(kBlock), 0, st, tiles, count, flags);
  HIP_CHECK(hipGetLastError());
}

void launch_object_batch_shift(const InitTile* tiles, uint32_t count, uint32_t mmax, hipStream_t st) {
  if (!count) return;
  hipLaunchKernelGGL(k_bobject_shift, dim3(blocks_for(mmax), count), dim3(kBlock), 0, st, tiles);
  HIP_CHECK(hipGetLastError());
}

void launch_object(const fpmhip_session& s, double2* out_dev) {
  const Model& m = *s.model;
  launch_shift(s.b.spec.as<double2>(), out_dev, m.rh, m.rw, (m.rh + 1) / 2, (m.rw + 1) / 2, s.ctx->stream);
  fft2_f64(*m.fft_high, out_dev, 1, true, s.ctx->stream);
}

void launch_shift(const double2* src, double2* dst, uint32_t h, uint32_t w, uint32_t rs, uint32_t cs,
                  hipStream_t stream) {
  hipLaunchKernelGGL(k_shift, dim3(blocks_for((uint64_t)h * w)), dim3(kBlock), 0, stream, src, dst, h, w, rs, cs);
  HIP_CHECK(hipGetLastError());
}

void launch_convert_pupil(const double2* src, void* dst, bool to_f32, uint64_t n, hipStream_t stream) {
  if (to_f32)
    hipLaunchKernelGGL(k_convert_from_d2<float2>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, src,
                       static_cast<float2*>(dst), n);
  else
    hipLaunchKernelGGL(k_convert_from_d2<double2>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, src,
                       static_cast<double2*>(dst), n);
  HIP_CHECK(hipGetLastError());
}

// Per-row support of a complex64 pupil: extent[r] = lo | hi << 16, where
// columns [lo, hi) hold every non-zero value of row r (lo = w, hi = 0 for a
// row of zeros).
__global__ void k_pupil_extent(const float2* __restrict__ pupil, uint32_t h, uint32_t w, uint32_t* __restrict__ extent) {
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= h) return;
  uint32_t lo = w, hi = 0;
  for (uint32_t c = 0; c < w; ++c) {
    const float2 p = pupil[(uint64_t)r * w + c];
    if (p.x != 0.0f || p.y != 0.0f) {
      lo = c < lo ? c : lo;
      hi = c + 1;
    }
  }
  extent[r] = lo | (hi << 16);
}

// Live-pixel lists of the fused sweep for a square n x n tile.
// Destinations of the adjoint insert: window row a (0..n) can receive a
// non-zero update only in columns [lo, hi + 1) (clamped to n + 1), where
// [lo, hi) is the union of the support extents of pupil rows a - 1 and a:
// the bilinear taps of a destination reach back one row and one column, and
// the update U is zero outside the support. Support pixels: row r, columns
// of its extent (the gather reads nothing elsewhere). Layout: dest[0] and
// dest[1] the two counts, then the destinations a | b << 16 and the support
// strips r | c << 10 | mask << 20, both row-major.
constexpr uint32_t kDestMaxRows = 513;
__global__ void k_dest_table(const uint32_t* __restrict__ extent, uint32_t n, uint32_t* __restrict__ dest) {
  __shared__ uint32_t lo_s[kDestMaxRows], len_s[kDestMaxRows], start_s[kDestMaxRows + 1];
  __shared__ uint32_t top_s[kDestMaxRows], cnt_s[kDestMaxRows], cstart_s[kDestMaxRows];
  const uint32_t rows = n + 1;
  for (uint32_t a = threadIdx.x; a < rows; a += blockDim.x) {
    uint32_t lo = rows, hi = 0;
    for (uint32_t r = (a > 0 ? a - 1 : 0); r <= a && r < n; ++r) {
      const uint32_t e = extent[r], elo = e & 0xFFFFu, ehi = e >> 16;
      if (ehi > elo) {
        lo = elo < lo ? elo : lo;
        hi = ehi > hi ? ehi : hi;
      }
    }
    if (hi > lo) hi = hi + 1 < rows ? hi + 1 : rows;
    lo_s[a] = lo;
    len_s[a] = hi > lo ? hi - lo : 0;
  }
  // strips: bands of kGatherStrip rows, each over the columns from its
  // first to its last support pixel (consecutive list entries are adjacent
  // columns of one band, so a wave's loads are contiguous)
  const uint32_t bands = (n + kGatherStrip - 1) / kGatherStrip;
  for (uint32_t k = threadIdx.x; k < bands; k += blockDim.x) {
    uint32_t lo = n, hi = 0;
    for (uint32_t r = k * kGatherStrip; r < (k + 1) * kGatherStrip && r < n; ++r) {
      const uint32_t e = extent[r], elo = e & 0xFFFFu, ehi = e >> 16;
      if (ehi > elo) {
        lo = elo < lo ? elo : lo;
        hi = ehi > hi ? ehi : hi;
      }
    }
    top_s[k] = lo;
    cnt_s[k] = hi > lo ? hi - lo : 0;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0, sacc = 0;
    for (uint32_t a = 0; a < rows; ++a) {
      start_s[a] = acc;
      acc += len_s[a];
    }
    for (uint32_t k = 0; k < bands; ++k) {
      cstart_s[k] = sacc;
      sacc += cnt_s[k];
    }
    cstart_s[bands] = sacc;
    start_s[rows] = acc;
    dest[0] = acc;
    dest[1] = sacc;
  }
  __syncthreads();
  uint32_t* dl = dest + 2;
  uint32_t* sl = dest + dest_strips_offset(n);
  uint32_t* rs = dest + dest_row_starts_offset(n);
  for (uint32_t a = threadIdx.x; a <= rows; a += blockDim.x) rs[a] = start_s[a];
  uint32_t* bs = dest + dest_band_starts_offset(n);
  for (uint32_t k = threadIdx.x; k <= bands; k += blockDim.x) bs[k] = cstart_s[k];
  for (uint32_t a = 0; a < rows; ++a)
    for (uint32_t j = threadIdx.x; j < len_s[a]; j += blockDim.x) dl[start_s[a] + j] = a | ((lo_s[a] + j) << 16);
  for (uint32_t k = 0; k < bands; ++k) {
    const uint32_t r0 = k * kGatherStrip;
    for (uint32_t j = threadIdx.x; j < cnt_s[k]; j += blockDim.x) {
      const uint32_t c = top_s[k] + j;
      uint32_t mask = 0;
      for (uint32_t i = 0; i < kGatherStrip && r0 + i < n; ++i) {
        const uint32_t e = extent[r0 + i];
        if (c >= (e & 0xFFFFu) && c < (e >> 16)) mask |= 1u << i;
      }
      sl[cstart_s[k] + j] = r0 | (c << 10) | (mask << 20);
    }
  }
}

void launch_pupil_extent(const fpmhip_session& s) {
  const Model& m = *s.model;
  if (s.precision != FPMHIP_MIXED_F32 || !s.b.extent.ptr) return;
  hipLaunchKernelGGL(k_pupil_extent, dim3((m.nh + 63) / 64), dim3(64), 0, s.ctx->stream, s.b.pupil.as<float2>(),
                     (uint32_t)m.nh, (uint32_t)m.nw, s.b.extent.as<uint32_t>());
  HIP_CHECK(hipGetLastError());
  if (s.b.dest.ptr && m.nh + 1 <= kDestMaxRows) {
    hipLaunchKernelGGL(k_dest_table, dim3(1), dim3(256), 0, s.ctx->stream, s.b.extent.as<uint32_t>(), (uint32_t)m.nh,
                       s.b.dest.as<uint32_t>());
    HIP_CHECK(hipGetLastError());
  }
}

void launch_pupil_to_f64(const void* src, bool from_f32, double2* dst, uint64_t n, hipStream_t stream) {
  if (from_f32)
    hipLaunchKernelGGL(k_convert_to_d2<float2>, dim3(blocks_for(n)), dim3(kBlock), 0, stream,
                       static_cast<const float2*>(src), dst, n);
  else
    hipLaunchKernelGGL(k_convert_to_d2<double2>, dim3(blocks_for(n)), dim3(kBlock), 0, stream,
                       static_cast<const double2*>(src), dst, n);
  HIP_CHECK(hipGetLastError());
}

void launch_fill_zero(double2* dst, uint64_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_fill_zero<double2>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, dst, n);
  HIP_CHECK(hipGetLastError());
}

void launch_fill_zero(float2* dst, uint64_t n, hipStream_t stream) {
  hipLaunchKernelGGL(k_fill_zero<float2>, dim3(blocks_for(n)), dim3(kBlock), 0, stream, dst, n);
  HIP_CHECK(hipGetLastError());
}

void launch_forward_stack_with(const Model& m, const double2* spec, const double2* pupil, double2* fields,
                               double* out, const double* gains, const double* bg, uint64_t bg_len,
                               hipStream_t stream) {
  const uint64_t N = m.N();
  hipLaunchKernelGGL(k_extract_all, dim3(blocks_for(N), (uint32_t)m.S), dim3(kBlock), 0, stream, spec, m.rw,
                     m.d_src.as<SrcParam>(), m.nh, m.nw, pupil, fields);
  HIP_CHECK(hipGetLastError());
  fft2_f64(*m.fft_low, fields, (uint32_t)m.S, true, stream);
  hipLaunchKernelGGL(k_intensity, dim3(blocks_for(N), (uint32_t)m.F), dim3(kBlock), 0, stream, fields,
                     (uint32_t)N, m.mux ? m.d_mux_ptr.as<uint32_t>() : nullptr, m.d_mux_src.as<uint32_t>(),
                     m.d_mux_w.as<double>(), gains, bg, bg_len, out);
  HIP_CHECK(hipGetLastError());
}

void launch_forward_stack(const Model& m, const double2* spec, const double2* pupil, double2* fields, double* out,
                          hipStream_t stream) {
  launch_forward_stack_with(m, spec, pupil, fields, out, m.d_gains.as<double>(), m.d_background.as<double>(),
                            (uint64_t)m.background.size(), stream);
}

void launch_forward_stack_frames(const Model& m, const double2* spec, const double2* pupil, double2* fields,
                                 double* out, uint64_t f0, uint64_t f1, hipStream_t stream) {
  const uint64_t N = m.N(), count = f1 - f0, bg_len = m.background.size();
  // without multiplexing frame f is source f: the frames' sources are
  // consecutive, and a per-frame background starts at frame f0
  const double* bg = bg_len == m.F * N && bg_len != N ? m.d_background.as<double>() + f0 * N
                                                      : m.d_background.as<double>();
  hipLaunchKernelGGL(k_extract_all, dim3(blocks_for(N), (uint32_t)count), dim3(kBlock), 0, stream, spec, m.rw,
                     m.d_src.as<SrcParam>() + f0, m.nh, m.nw, pupil, fields);
  HIP_CHECK(hipGetLastError());
  fft2_f64(*m.fft_low, fields, (uint32_t)count, true, stream);
  hipLaunchKernelGGL(k_intensity, dim3(blocks_for(N), (uint32_t)count), dim3(kBlock), 0, stream, fields, (uint32_t)N,
                     nullptr, nullptr, nullptr, m.d_gains.as<double>() + f0, bg, bg_len, out + f0 * N);
  HIP_CHECK(hipGetLastError());
}

// difference of mode `mode` of frame `frame` (multiplexed frames only)
void launch_difference_frame(const fpmhip_session& s, uint32_t mode, uint64_t frame) {
  const Model& m = *s.model;
  const uint64_t N = m.N();
  const uint8_t* mask = s.meas->d_masks.ptr ? s.meas->d_masks.as<uint8_t>() + frame * N : nullptr;
  DISPATCH_V(s, hipLaunchKernelGGL(k_difference<V>, dim3(blocks_for(N)), dim3(kBlock), 0, s.ctx->stream,
                                   s.b.field.as<V>() + mode * N, s.b.proj.as<V>(), s.b.pred.as<double>(),
                                   (uint32_t)N, mask, s.cfg.epsilon));
  HIP_CHECK(hipGetLastError());
}

}  // namespace fpmhip
