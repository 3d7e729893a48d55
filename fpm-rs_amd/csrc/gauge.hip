// gauge.hip — EPRY's object/pupil gauge canonicalisation on the device
// (canonicalize_object_pupil, src/algorithms/gauge.rs:20-296), one
// 1024-thread workgroup per tile, so that the per-iteration gauge needs no
// host round trip of the spectrum and pupil:
//
//   relative phase slopes of the pupil against the compiled reference pupil
//   on the axes where every source offset is exact (gauge.rs:46-131), and the
//   affine phase ramp that removes them from pupil and spectrum (:132-156)
//   pupil energy match s = sqrt(E_ref / E) and piston -arg(sum conj(P_ref) P)
//   (:158-240): P *= s e^{-i phi}, spectrum *= 1 / (s e^{-i phi})
//   object phase anchor: the DC value, or the strongest value when DC is
//   weak, rotated to the real axis (:242-296)
//
// Arithmetic is FP64 with num-complex's formulas (the library is built with
// -ffp-contract=off); the pupil is read in its stored precision and written
// back once. Sums are fixed-order workgroup reductions; "strongest" picks the
// first index of the largest weight, as the reference's strict comparison
// does. Failures leave a code in the tile's status word (GaugeStatus).
#include <cfloat>

#include "gauge.h"

namespace fpmhip {

namespace {

constexpr int kThreads = 1024;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double cabs(double2 a) { return hypot(a.x, a.y); }
__device__ __forceinline__ double carg(double2 a) { return atan2(a.y, a.x); }
__device__ __forceinline__ double2 cpolar(double r, double t) { return make_double2(r * cos(t), r * sin(t)); }
__device__ __forceinline__ bool cfinite(double2 a) { return isfinite(a.x) && isfinite(a.y); }

// Partial reduction state of one thread: a complex sum, a real sum and the
// (weight, index) of the first largest weight seen.
struct Acc {
  double2 sum;
  double total;
  double best;
  uint32_t idx;
};
__device__ __forceinline__ Acc acc_zero() { return {make_double2(0.0, 0.0), 0.0, 0.0, 0xFFFFFFFFu}; }
__device__ __forceinline__ void acc_pick(Acc& a, double w, uint32_t i) {
  // strictly larger, or equal at a smaller index (the first occurrence)
  if (w > a.best || (w == a.best && w > 0.0 && i < a.idx)) {
    a.best = w;
    a.idx = i;
  }
}
__device__ __forceinline__ void acc_merge(Acc& a, const Acc& b) {
  a.sum = make_double2(a.sum.x + b.sum.x, a.sum.y + b.sum.y);
  a.total += b.total;
  acc_pick(a, b.best, b.idx);
}

// Workgroup reduction of Acc (fixed order); the result is valid in every thread.
__device__ Acc block_reduce(Acc a, Acc* red) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    Acc b;
    b.sum.x = __shfl_xor(a.sum.x, off, 64);
    b.sum.y = __shfl_xor(a.sum.y, off, 64);
    b.total = __shfl_xor(a.total, off, 64);
    b.best = __shfl_xor(a.best, off, 64);
    b.idx = (uint32_t)__shfl_xor((int)a.idx, off, 64);
    acc_merge(a, b);
  }
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  Acc r = red[0];
  for (int w = 1; w < kWaves; ++w) acc_merge(r, red[w]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ double2 load_pupil(const GaugeTile& t, uint32_t p) {
  if (t.pupil64) return t.pupil64[p];
  if (t.pupil_f32) {
    const float2 v = static_cast<const float2*>(t.pupil)[p];
    return make_double2((double)v.x, (double)v.y);
  }
  return static_cast<const double2*>(t.pupil)[p];
}

// The ramp factor of pupil pixel (r, c) or object pixel (r, c) (sign +1 for
// the object, -1 for the pupil; apply_affine_correction, gauge.rs:132-156).
__device__ __forceinline__ double2 ramp(double rs, double cs, uint32_t r, uint32_t c, uint32_t h, uint32_t w,
                                        double sign) {
  const double y = (double)r - (double)(h / 2), x = (double)c - (double)(w / 2);
  const double ph = sign > 0 ? rs * y + cs * x : -(rs * y + cs * x);
  return cpolar(1.0, ph);
}

// Relative phase slope of the pupil along one axis (relative_phase_slope,
// gauge.rs:75-131): the pairs (cur, next) of supported neighbours.
__device__ double phase_slope(const GaugeTile& t, bool row_axis, Acc* red) {
  const uint32_t h = t.nh, w = t.nw;
  const uint32_t rl = row_axis ? h - 1 : h, cl = row_axis ? w : w - 1;
  Acc a = acc_zero();
  for (uint32_t k = threadIdx.x; k < rl * cl; k += kThreads) {
    const uint32_t r = k / cl, c = k % cl;
    const uint32_t cur = r * w + c, nxt = row_axis ? (r + 1) * w + c : cur + 1;
    if (!t.support[cur] || !t.support[nxt]) continue;
    const double2 ce = cmul(load_pupil(t, nxt), cconj(load_pupil(t, cur)));
    const double2 re = cmul(t.ref[nxt], cconj(t.ref[cur]));
    const double2 contribution = cmul(ce, cconj(re));
    const double weight = cabs(contribution);
    if (!isfinite(weight)) continue;
    a.sum = make_double2(a.sum.x + contribution.x, a.sum.y + contribution.y);
    a.total += weight;
    acc_pick(a, weight, cur);
  }
  a = block_reduce(a, red);
  if (a.total == 0.0) return 0.0;
  if (cabs(a.sum) > t.tol * a.total) return carg(a.sum);
  if (a.best > 0.0) {
    const uint32_t cur = a.idx, nxt = row_axis ? cur + w : cur + 1;
    const double2 ce = cmul(load_pupil(t, nxt), cconj(load_pupil(t, cur)));
    const double2 re = cmul(t.ref[nxt], cconj(t.ref[cur]));
    return carg(cmul(ce, cconj(re)));
  }
  return 0.0;
}

struct Shared {
  Acc red[kWaves];
  double2 dc;
  int stop;
};

__global__ void __launch_bounds__(kThreads) k_gauge(const GaugeTile* __restrict__ tiles) {
  __shared__ Shared sh;
  const GaugeTile t = tiles[blockIdx.x];
  const uint32_t N = t.nh * t.nw, M = t.rh * t.rw;
  const double tol = t.tol;
  auto failed = [&](int code) {
    if (threadIdx.x == 0 && *t.status == 0) *t.status = code;
  };
  // 1. ramp slopes on the exact axes (both from the pupil before any change)
  const double rs = t.row_exact ? phase_slope(t, true, sh.red) : 0.0;
  const double cs = t.col_exact ? phase_slope(t, false, sh.red) : 0.0;
  const bool ramped = fabs(rs) > tol || fabs(cs) > tol;
  auto pupil_at = [&](uint32_t p) {
    const double2 v = load_pupil(t, p);
    return ramped ? cmul(v, ramp(rs, cs, p / t.nw, p % t.nw, t.nh, t.nw, -1.0)) : v;
  };
  // 2. energies and the overlap with the reference pupil over the support
  Acc e = acc_zero(), o = acc_zero();
  for (uint32_t p = threadIdx.x; p < N; p += kThreads) {
    if (!t.support[p]) continue;
    const double2 P = pupil_at(p), R = t.ref[p];
    e.sum.x += R.x * R.x + R.y * R.y;  // reference energy
    e.total += P.x * P.x + P.y * P.y;  // current energy
    const double2 contribution = cmul(cconj(R), P);
    const double weight = cabs(contribution);
    o.sum = make_double2(o.sum.x + contribution.x, o.sum.y + contribution.y);
    o.total += weight;
    acc_pick(o, weight, p);
  }
  e = block_reduce(e, sh.red);
  o = block_reduce(o, sh.red);
  const double eref = e.sum.x, ecur = e.total;
  if (!isfinite(eref) || eref <= 0.0) return failed(kGaugeBadReference);
  if (!isfinite(ecur) || ecur <= 0.0) return failed(kGaugeBadEnergy);
  const double computed = sqrt(eref / ecur);
  if (!isfinite(computed) || computed <= 0.0) return failed(kGaugeBadScale);
  const double scale = fabs(computed - 1.0) <= tol ? 1.0 : computed;
  double phase;
  if (o.total > 0.0 && cabs(o.sum) > tol * o.total)
    phase = carg(o.sum);
  else if (o.best > 0.0)
    phase = carg(cmul(cconj(t.ref[o.idx]), pupil_at(o.idx)));
  else
    return failed(kGaugeNoAnchor);
  if (!isfinite(phase)) return failed(kGaugeBadPhase);
  if (fabs(phase) <= tol) phase = 0.0;
  const double2 pf = cpolar(scale, -phase);
  const double n2 = pf.x * pf.x + pf.y * pf.y;
  const double2 of = make_double2((1.0 * pf.x + 0.0 * pf.y) / n2, (0.0 * pf.x - 1.0 * pf.y) / n2);
  const bool unit = pf.x == 1.0 && pf.y == 0.0;
  // 3. object phase anchor over the spectrum as scaled by step 2
  auto spec_at = [&](uint32_t i) {
    double2 v = t.spec[i];
    if (ramped) v = cmul(v, ramp(rs, cs, i / t.rw, i % t.rw, t.rh, t.rw, 1.0));
    return unit ? v : cmul(v, of);
  };
  const uint32_t dci = (t.rh / 2) * t.rw + t.rw / 2;
  Acc s = acc_zero();
  int bad = 0;
  for (uint32_t i = threadIdx.x; i < M; i += kThreads) {
    const double2 v = spec_at(i);
    bad |= !cfinite(v);
    acc_pick(s, cabs(v), i);
    if (i == dci) sh.dc = v;
  }
  s.total = (double)bad;
  s = block_reduce(s, sh.red);
  if (s.total != 0.0) return failed(kGaugeBadSpectrum);
  double2 corr = make_double2(1.0, 0.0);
  bool rotate = false;
  if (s.best > 0.0) {  // smag == 0: no object phase step (gauge.rs:257)
    const double2 dc = sh.dc;
    const double2 anchor = cabs(dc) > tol * s.best ? dc : spec_at(s.idx);
    const double oph = carg(anchor);
    if (!isfinite(oph)) return failed(kGaugeBadObjectPhase);
    if (fabs(oph) > tol) {
      corr = cpolar(1.0, -oph);
      rotate = true;
    }
  }
  // 4. write back what changed
  if (ramped || !unit || rotate)
    for (uint32_t i = threadIdx.x; i < M; i += kThreads) {
      const double2 v = spec_at(i);
      t.spec[i] = rotate ? cmul(v, corr) : v;
    }
  double maxp = 0.0;
  for (uint32_t p = threadIdx.x; p < N; p += kThreads) {
    double2 P = pupil_at(p);
    if (!unit) P = cmul(P, pf);
    if (t.pupil64 && (ramped || !unit)) t.pupil64[p] = P;
    if (t.pupil_f32) {
      const float2 f = make_float2((float)P.x, (float)P.y);
      if (ramped || !unit) static_cast<float2*>(t.pupil)[p] = f;
      maxp = fmax(maxp, (double)f.x * (double)f.x + (double)f.y * (double)f.y);
    } else {
      if (ramped || !unit) static_cast<double2*>(t.pupil)[p] = P;
      maxp = fmax(maxp, P.x * P.x + P.y * P.y);
    }
  }
  if (t.maxp_out) {
    Acc m = acc_zero();
    m.total = maxp;  // max, not a sum: reduced below with fmax
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m.total = fmax(m.total, __shfl_xor(m.total, off, 64));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh.red[threadIdx.x >> 6].total = m.total;
    __syncthreads();
    if (threadIdx.x == 0) {
      double r = 0.0;
      for (int w = 0; w < kWaves; ++w) r = fmax(r, sh.red[w].total);
      *t.maxp_out = r;
    }
  }
}

}  // namespace

void launch_gauge(const GaugeTile* tiles_dev, uint32_t count, hipStream_t stream) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_gauge, dim3(count), dim3(kThreads), 0, stream, tiles_dev);
  HIP_CHECK(hipGetLastError());
}

const char* gauge_failure(int32_t code, int* status) {
  *status = FPMHIP_NUMERICAL;
  switch (code) {
    case kGaugeBadReference:
      *status = FPMHIP_INVALID_MODEL;
      return "compiled reference pupil has invalid supported energy";
    case kGaugeBadEnergy:
      return "object/pupil gauge normalization requires positive finite recovered pupil energy";
    case kGaugeBadScale:
      return "object/pupil gauge normalization produced an invalid pupil scale";
    case kGaugeNoAnchor:
      return "object/pupil gauge normalization could not anchor recovered pupil phase";
    case kGaugeBadPhase:
      return "object/pupil gauge normalization produced a non-finite pupil phase";
    case kGaugeBadSpectrum:
      return "object/pupil gauge normalization requires a finite object spectrum";
    case kGaugeBadObjectPhase:
      return "object/pupil gauge normalization produced a non-finite object phase";
    default:
      return nullptr;
  }
}

}  // namespace fpmhip
