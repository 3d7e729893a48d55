// fpm_oracle.cpp — CPU FP64 restatement of fpm-rs's image-plane forward model
// and per-LED reconstruction updates.
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
// path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load it, and never as the thing that is measured or shipped.
//
// Every function cites the hgrecco/fpm-rs file:line it restates. Arithmetic
// follows num-complex 0.4.6's formulas term by term (Mul, Div by complex and
// by f64, norm = hypot, arg = atan2, from_polar) and is compiled with
// -ffp-contract=off, as Rust never contracts a*b+c into an FMA. The FFT is the
// DFT with fpm-rs's normalization (forward x 1/N, inverse unscaled,
// src/backend/cpu.rs:162-172) computed by a Stockham mixed-radix algorithm of
// this file; rustfft 6.4.1's internal rounding is not reproduced (it is not
// vendored in the reference, and parity is pinned to the DFT definition).
//
// Pinning (what checks each part; the tests live in tests/test_oracle.py):
//   fft2            DFT definition by brute force; numpy.fft; round trip
//                   (tests/forward_model_tests.rs:13); constant object -> 4.0
//                   (tests/forward_model_tests.rs:76)
//   shifts          numpy.fft.fftshift/ifftshift, odd and even sizes
//   extract/insert  linear-ramp closed form (forward_model_tests.rs:117),
//                   cos(pi nu/H) attenuation (:155), adjoint identity (:219),
//                   extract after insert (:104)
//   compile         fractional offsets 0.25/0.5 and shift sign
//                   (tests/geometry_tests.rs:474), smooth/minimum sizes by
//                   brute force, circular pupil closed form
//   initialize      gain 4 / background 9 / amplitude 2 (reconstruction_tests.rs:2713)
//   momentum        golden vector (src/algorithms/common.rs:513-535)
//   update          AP with unit pupil and integer crops equals the textbook
//                   amplitude-replacement step written independently in numpy;
//                   mPIE feedback 0 == FPIE bitwise (reconstruction_tests.rs:655);
//                   loss decrease (:66, :638); mPIE < 0.95 FPIE and first to
//                   5.4e-5 (:992)
//   gauge           energy / overlap / DC invariants (reconstruction_tests.rs:2674),
//                   prediction invariance (src/algorithms/gauge.rs:436,465)
//   point_loss      the reference's scalar values (tests/diagnostics_tests.rs:34-58)
//                   and the Huber / clamping branches by closed form
//   schedules       BrightfieldFirst / SpiralOut orders (tests/schedule_tests.rs:15-22),
//                   SnrWeighted with a masked outlier (:37-48)
//   EPRY pupil      pupil phase error falls from the defocus + astigmatism
//                   start (tests/reconstruction_tests.rs:2923)
//   aberrations     DC phase = spherical, axis values of astigmatism / coma,
//                   apodization, defocus quadratic in |k| with its sign
//   forward stack   constant object (forward_model_tests.rs:76), intensity =
//                   |IFFT|^2 written with numpy
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../include/fpmhip.h"

namespace {

thread_local std::string g_error;
thread_local std::string g_error_name;

struct Failure {
  int code;
};

[[noreturn]] void fail(int code, const std::string& message, const char* name = "") {
  g_error = message;
  g_error_name = name;
  throw Failure{code};
}

// ---------------------------------------------------------------------------
// complex arithmetic with num-complex 0.4.6 semantics
struct C {
  double re, im;
};
inline C cadd(C a, C b) { return {a.re + b.re, a.im + b.im}; }
inline C csub(C a, C b) { return {a.re - b.re, a.im - b.im}; }
inline C cmul(C a, C b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline C cscale(double s, C a) { return {s * a.re, s * a.im}; }      // f64 * Complex
inline C cmulr(C a, double s) { return {a.re * s, a.im * s}; }       // Complex * f64
inline C cdivr(C a, double s) { return {a.re / s, a.im / s}; }       // Complex / f64
inline C cconj(C a) { return {a.re, -a.im}; }
inline double cnorm_sqr(C a) { return a.re * a.re + a.im * a.im; }
inline double cnorm(C a) { return std::hypot(a.re, a.im); }
inline double carg(C a) { return std::atan2(a.im, a.re); }
inline C cpolar(double r, double t) { return {r * std::cos(t), r * std::sin(t)}; }
inline C cdiv(C a, C b) {  // Complex / Complex
  double n = b.re * b.re + b.im * b.im;
  double re = a.re * b.re + a.im * b.im;
  double im = a.im * b.re - a.re * b.im;
  return {re / n, im / n};
}
inline bool cfinite(C a) { return std::isfinite(a.re) && std::isfinite(a.im); }

// ---------------------------------------------------------------------------
// FFT: Stockham autosort, decimation in frequency, radices 4, 2, 3, 5, 7 and
// a direct DFT butterfly for any other prime factor.

// exp(sign * 2 pi i j / n) with quarter-turn points exact.
C unit_root(uint64_t j, uint64_t n, int sign) {
  j %= n;
  if ((4 * j) % n == 0) {
    uint64_t q = (4 * j) / n;
    static const C quarter[4] = {{1.0, 0.0}, {0.0, 1.0}, {-1.0, 0.0}, {0.0, -1.0}};
    C v = quarter[q];
    return {v.re, sign * v.im};
  }
  long double theta = 2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
  return {(double)cosl(theta), (double)(sign * sinl(theta))};
}

struct Plan1 {
  uint64_t n = 0;
  std::vector<uint64_t> radices;
  std::vector<C> fwd, inv;  // unit roots W_n^j, j in [0, n)
};

std::vector<uint64_t> factorize(uint64_t n) {
  std::vector<uint64_t> r;
  while (n % 4 == 0) { r.push_back(4); n /= 4; }
  while (n % 2 == 0) { r.push_back(2); n /= 2; }
  for (uint64_t p = 3; p * p <= n; p += 2)
    while (n % p == 0) { r.push_back(p); n /= p; }
  if (n > 1) r.push_back(n);
  return r;
}

Plan1 make_plan(uint64_t n) {
  Plan1 p;
  p.n = n;
  p.radices = factorize(n);
  p.fwd.resize(n);
  p.inv.resize(n);
  for (uint64_t j = 0; j < n; ++j) {
    p.fwd[j] = unit_root(j, n, -1);
    p.inv[j] = unit_root(j, n, +1);
  }
  return p;
}

// In-place 1-D transform of x[0..n) (contiguous), unscaled; `work` has n entries.
void fft1(const Plan1& plan, C* x, C* work, bool inverse) {
  const uint64_t N = plan.n;
  if (N == 1) return;
  const std::vector<C>& W = inverse ? plan.inv : plan.fwd;
  const double sgn = inverse ? 1.0 : -1.0;
  C* src = x;
  C* dst = work;
  uint64_t ncur = N, s = 1;
  C a[64], b[64];
  std::vector<C> big_a, big_b;
  for (uint64_t p : plan.radices) {
    const uint64_t m = ncur / p;
    C* pa = a;
    C* pb = b;
    if (p > 64) {
      big_a.resize(p);
      big_b.resize(p);
      pa = big_a.data();
      pb = big_b.data();
    }
    for (uint64_t pp = 0; pp < m; ++pp) {
      for (uint64_t q = 0; q < s; ++q) {
        for (uint64_t r = 0; r < p; ++r) pa[r] = src[q + s * (pp + r * m)];
        if (p == 2) {
          pb[0] = cadd(pa[0], pa[1]);
          pb[1] = csub(pa[0], pa[1]);
        } else if (p == 4) {
          C t0 = cadd(pa[0], pa[2]);
          C t1 = csub(pa[0], pa[2]);
          C t2 = cadd(pa[1], pa[3]);
          C d = csub(pa[1], pa[3]);
          // d * (sgn * i): forward -> (d.im, -d.re); inverse -> (-d.im, d.re)
          C t3 = inverse ? C{-d.im, d.re} : C{d.im, -d.re};
          pb[0] = cadd(t0, t2);
          pb[1] = cadd(t1, t3);
          pb[2] = csub(t0, t2);
          pb[3] = csub(t1, t3);
        } else {
          const uint64_t step = N / p;  // W_p^j = W_N^{j N / p}
          for (uint64_t k = 0; k < p; ++k) {
            C acc = pa[0];
            for (uint64_t r = 1; r < p; ++r) acc = cadd(acc, cmul(pa[r], W[((r * k) % p) * step]));
            pb[k] = acc;
          }
        }
        dst[q + s * (p * pp)] = pb[0];
        for (uint64_t k = 1; k < p; ++k) {
          // twiddle W_ncur^{pp k} = W_N^{pp k s}
          C w = W[(pp * k * s) % N];
          dst[q + s * (p * pp + k)] = cmul(pb[k], w);
        }
      }
    }
    (void)sgn;
    std::swap(src, dst);
    ncur = m;
    s *= p;
  }
  if (src != x) std::memcpy(x, src, sizeof(C) * N);
}

struct Plan2 {
  uint64_t h = 0, w = 0;
  Plan1 row, col;
};

// CpuBackend::fft2 (src/backend/cpu.rs:130-174): row transforms, then column
// transforms through a gathered column, then x 1/N for the forward direction.
void fft2(const Plan2& plan, C* v, bool inverse) {
  const uint64_t h = plan.h, w = plan.w;
  std::vector<C> work(std::max(h, w)), column(h);
  for (uint64_t r = 0; r < h; ++r) fft1(plan.row, v + r * w, work.data(), inverse);
  for (uint64_t c = 0; c < w; ++c) {
    for (uint64_t r = 0; r < h; ++r) column[r] = v[r * w + c];
    fft1(plan.col, column.data(), work.data(), inverse);
    for (uint64_t r = 0; r < h; ++r) v[r * w + c] = column[r];
  }
  if (!inverse) {
    const double norm = (double)(h * w);
    for (uint64_t i = 0; i < h * w; ++i) v[i] = cdivr(v[i], norm);
  }
}

Plan2 make_plan2(uint64_t h, uint64_t w) {
  Plan2 p;
  p.h = h;
  p.w = w;
  p.row = make_plan(w);
  p.col = make_plan(h);
  return p;
}

// shift_copy / fftshift_copy / ifftshift_copy (src/model/forward.rs:530-569)
void shift_copy(const C* src, C* dst, uint64_t h, uint64_t w, uint64_t rs, uint64_t cs) {
  for (uint64_t r = 0; r < h; ++r)
    for (uint64_t c = 0; c < w; ++c) dst[((r + rs) % h) * w + (c + cs) % w] = src[r * w + c];
}
void fftshift(const C* s, C* d, uint64_t h, uint64_t w) { shift_copy(s, d, h, w, h / 2, w / 2); }
void ifftshift(const C* s, C* d, uint64_t h, uint64_t w) {
  shift_copy(s, d, h, w, (h + 1) / 2, (w + 1) / 2);
}

// ---------------------------------------------------------------------------
// bilinear crop and its adjoint (src/model/crop.rs:129-313)

struct Axis {
  uint64_t lower, upper;
  double wl, wu;
};

// interpolation_axis (src/model/crop.rs:266-313)
Axis interpolation_axis(uint64_t start, uint64_t length, uint64_t bound, double offset) {
  if (length == 0 || bound == 0 || !std::isfinite(offset))
    fail(FPMHIP_INVALID_MODEL, "subpixel crop dimensions and offset must be finite and non-zero");
  double nearest = std::round(offset);
  double integer_offset, fraction;
  if (std::fabs(offset - nearest) <= 1e-12) {
    integer_offset = nearest;
    fraction = 0.0;
  } else {
    double fl = std::floor(offset);
    integer_offset = fl;
    fraction = offset - fl;
  }
  int64_t lower = (int64_t)start + (int64_t)integer_offset;
  if (lower < 0) fail(FPMHIP_INVALID_MODEL, "subpixel crop starts outside the grid");
  uint64_t lower_end = (uint64_t)lower + length;
  bool needs_upper = fraction > 0.0;
  if (lower_end > bound || (needs_upper && lower_end >= bound))
    fail(FPMHIP_INVALID_MODEL, "subpixel crop interpolation stencil lies outside the grid");
  return {(uint64_t)lower, (uint64_t)lower + (needs_upper ? 1u : 0u), 1.0 - fraction, fraction};
}

bool offset_is_zero(double r, double c) { return std::fabs(r) <= 1e-12 && std::fabs(c) <= 1e-12; }

// extract_subpixel_standard (src/model/crop.rs:129-167), integer path :79-102
void extract(const C* spec, uint64_t R_h, uint64_t R_w, uint64_t sr, uint64_t sc, uint64_t h,
             uint64_t w, double orow, double ocol, C* dst) {
  if (offset_is_zero(orow, ocol)) {
    if (sr + h > R_h || sc + w > R_w) fail(FPMHIP_INVALID_MODEL, "crop lies outside reconstruction shape");
    for (uint64_t r = 0; r < h; ++r)
      for (uint64_t c = 0; c < w; ++c) dst[r * w + c] = spec[(sr + r) * R_w + sc + c];
    return;
  }
  Axis rows = interpolation_axis(sr, h, R_h, orow);
  Axis cols = interpolation_axis(sc, w, R_w, ocol);
  const double wll = rows.wl * cols.wl, wlu = rows.wl * cols.wu, wul = rows.wu * cols.wl,
               wuu = rows.wu * cols.wu;
  for (uint64_t r = 0; r < h; ++r) {
    uint64_t lr = rows.lower + r, ur = rows.upper + r;
    for (uint64_t c = 0; c < w; ++c) {
      uint64_t lc = cols.lower + c, uc = cols.upper + c;
      C v = cmulr(spec[lr * R_w + lc], wll);
      v = cadd(v, cmulr(spec[lr * R_w + uc], wlu));
      v = cadd(v, cmulr(spec[ur * R_w + lc], wul));
      v = cadd(v, cmulr(spec[ur * R_w + uc], wuu));
      dst[r * w + c] = v;
    }
  }
}

// insert_subpixel_adjoint_slice (src/model/crop.rs:188-256)
void insert_adjoint(C* spec, uint64_t R_h, uint64_t R_w, uint64_t sr, uint64_t sc, uint64_t h,
                    uint64_t w, const C* update, double scale, double orow, double ocol) {
  if (!std::isfinite(scale)) fail(FPMHIP_INVALID_PARAMETER, "must be finite", "scale");
  if (offset_is_zero(orow, ocol)) {
    if (sr + h > R_h || sc + w > R_w) fail(FPMHIP_INVALID_MODEL, "crop lies outside reconstruction shape");
    for (uint64_t r = 0; r < h; ++r)
      for (uint64_t c = 0; c < w; ++c) {
        C& d = spec[(sr + r) * R_w + sc + c];
        d = cadd(d, cscale(scale, update[r * w + c]));
      }
    return;
  }
  Axis rows = interpolation_axis(sr, h, R_h, orow);
  Axis cols = interpolation_axis(sc, w, R_w, ocol);
  const double wll = rows.wl * cols.wl, wlu = rows.wl * cols.wu, wul = rows.wu * cols.wl,
               wuu = rows.wu * cols.wu;
  for (uint64_t r = 0; r < h; ++r) {
    uint64_t lr = rows.lower + r, ur = rows.upper + r;
    for (uint64_t c = 0; c < w; ++c) {
      uint64_t lc = cols.lower + c, uc = cols.upper + c;
      C value = cscale(scale, update[r * w + c]);
      C& a = spec[lr * R_w + lc];
      a = cadd(a, cmulr(value, wll));
      C& b = spec[lr * R_w + uc];
      b = cadd(b, cmulr(value, wlu));
      C& cc = spec[ur * R_w + lc];
      cc = cadd(cc, cmulr(value, wul));
      C& d = spec[ur * R_w + uc];
      d = cadd(d, cmulr(value, wuu));
    }
  }
}

// ---------------------------------------------------------------------------
// compiled model (src/model/image_plane_fpm.rs:58-73)

struct Model {
  uint64_t nh, nw, rh, rw, S, F;
  std::vector<uint64_t> crop_r, crop_c;
  std::vector<double> off_r, off_c;
  std::vector<C> pupil;
  std::vector<uint8_t> support;
  bool has_gain = false;
  std::vector<double> gain;
  std::vector<double> background;  // empty, N or F*N
  bool mux = false;
  std::vector<uint64_t> row_ptr, msrc;
  std::vector<double> mw;
  bool has_k = false;
  std::vector<double> kx, ky;
  uint64_t N() const { return nh * nw; }
  uint64_t M() const { return rh * rw; }
};

Model load_model(const fpmhip_model_desc* d) {
  if (!d) fail(FPMHIP_INVALID_PARAMETER, "model descriptor is null", "model");
  Model m;
  m.nh = d->image_h;
  m.nw = d->image_w;
  m.rh = d->recon_h;
  m.rw = d->recon_w;
  m.S = d->source_count;
  m.F = d->mux_row_ptr ? d->frame_count : d->source_count;
  if (m.nh == 0 || m.nw == 0 || m.S == 0) fail(FPMHIP_INVALID_MODEL, "empty model");
  if (m.rh < m.nh || m.rw < m.nw)
    fail(FPMHIP_INVALID_MODEL, "reconstruction shape must contain a low-resolution crop");
  const uint64_t N = m.N();
  m.crop_r.resize(m.S);
  m.crop_c.resize(m.S);
  m.off_r.assign(m.S, 0.0);
  m.off_c.assign(m.S, 0.0);
  for (uint64_t s = 0; s < m.S; ++s) {
    m.crop_r[s] = d->crop_start[2 * s];
    m.crop_c[s] = d->crop_start[2 * s + 1];
    if (m.crop_r[s] + m.nh > m.rh || m.crop_c[s] + m.nw > m.rw)
      fail(FPMHIP_INVALID_MODEL, "crop lies outside reconstruction shape");
    if (d->offset) {
      m.off_r[s] = d->offset[2 * s];
      m.off_c[s] = d->offset[2 * s + 1];
      interpolation_axis(m.crop_r[s], m.nh, m.rh, m.off_r[s]);
      interpolation_axis(m.crop_c[s], m.nw, m.rw, m.off_c[s]);
    }
  }
  m.pupil.resize(N);
  m.support.resize(N);
  for (uint64_t i = 0; i < N; ++i) {
    m.pupil[i] = {d->pupil[2 * i], d->pupil[2 * i + 1]};
    if (!cfinite(m.pupil[i])) fail(FPMHIP_INVALID_MODEL, "pupil support or numeric values are invalid");
    m.support[i] = d->support[i];
    if (m.support[i] > 1) fail(FPMHIP_INVALID_MODEL, "pupil support values must be exactly zero or one");
  }
  if (d->frame_gain) {
    m.has_gain = true;
    m.gain.assign(d->frame_gain, d->frame_gain + m.F);
    for (double g : m.gain)
      if (!std::isfinite(g) || g < 0.0) fail(FPMHIP_INVALID_MODEL, "frame gains must be finite and non-negative");
  }
  if (d->background && d->background_len) {
    if (d->background_len != N && d->background_len != N * m.F)
      fail(FPMHIP_INVALID_MODEL, "background must be one image or one image per frame");
    m.background.assign(d->background, d->background + d->background_len);
  }
  if (d->mux_row_ptr) {
    m.mux = true;
    m.row_ptr.assign(d->mux_row_ptr, d->mux_row_ptr + m.F + 1);
    uint64_t nnz = m.row_ptr[m.F];
    m.msrc.assign(d->mux_source, d->mux_source + nnz);
    m.mw.assign(d->mux_weight, d->mux_weight + nnz);
    for (uint64_t f = 0; f < m.F; ++f) {
      if (m.row_ptr[f + 1] <= m.row_ptr[f]) fail(FPMHIP_INVALID_MODEL, "multiplexing row is empty");
      for (uint64_t e = m.row_ptr[f]; e < m.row_ptr[f + 1]; ++e)
        if (m.msrc[e] >= m.S || !std::isfinite(m.mw[e]) || m.mw[e] < 0.0)
          fail(FPMHIP_INVALID_MODEL, "multiplexing row contains an invalid source or weight");
    }
  }
  if (d->k_vectors) {
    m.has_k = true;
    m.kx.resize(m.S);
    m.ky.resize(m.S);
    for (uint64_t s = 0; s < m.S; ++s) {
      m.kx[s] = d->k_vectors[2 * s];
      m.ky[s] = d->k_vectors[2 * s + 1];
    }
  }
  return m;
}

// background_value (src/algorithms/common.rs:491-504, image_plane_fpm.rs:816-837)
inline double bg_value(const std::vector<double>& bg, uint64_t N, uint64_t f, uint64_t p) {
  if (bg.empty()) return 0.0;
  return bg[bg.size() == N ? p : f * N + p];
}

struct Meas {
  uint64_t F, N;
  const double* frames;
  std::vector<double> weight;
  const uint8_t* masks;
};

Meas load_meas(const Model& m, const double* frames, const double* weights, const uint8_t* masks) {
  Meas s;
  s.F = m.F;
  s.N = m.N();
  s.frames = frames;
  s.masks = masks;
  s.weight.assign(m.F, 1.0);
  if (weights)
    for (uint64_t f = 0; f < m.F; ++f) {
      if (!std::isfinite(weights[f]) || weights[f] < 0.0)
        fail(FPMHIP_INVALID_MEASUREMENTS, "frame weights must be finite and non-negative");
      s.weight[f] = weights[f];
    }
  for (uint64_t i = 0; i < s.F * s.N; ++i)
    if (!std::isfinite(frames[i])) fail(FPMHIP_INVALID_MEASUREMENTS, "measurements must be finite");
  return s;
}

// ---------------------------------------------------------------------------
// state (src/reconstruction/state.rs:159-172) and algorithm configuration

struct State {
  std::vector<C> spectrum, pupil;
  bool has_gain = false;
  std::vector<double> gain;
  std::vector<double> background;
  bool has_aux = false;
  std::vector<C> velocity, anchor;
  uint64_t counter = 0;
};

struct Config {
  double object_step;
  bool has_pupil_step;
  double pupil_step;
  double epsilon;
  int loss;
  int denom;
  double alpha;
  bool constrain;
  bool gain_update;
  double gain_step, gain_min, gain_max;
  bool bg_update;
  double bg_step, bg_min, bg_max;
  bool momentum;
  uint64_t interval;
  double friction, feedback;
};

// point_loss (src/algorithms/objective.rs:47-71)
double point_loss(double predicted, double measured, int loss) {
  switch (loss) {
    case FPMHIP_LOSS_INTENSITY_MSE: {
      double r = predicted - measured;
      return r * r;
    }
    case FPMHIP_LOSS_POISSON_NLL: {
      double mean = std::max(predicted, 1e-12);
      return mean - std::max(measured, 0.0) * std::log(mean);
    }
    case FPMHIP_LOSS_HUBER_AMPLITUDE: {
      double r = std::sqrt(std::max(predicted, 0.0)) - std::sqrt(std::max(measured, 0.0));
      double a = std::fabs(r);
      return a <= 1.0 ? 0.5 * r * r : a - 0.5;
    }
    default: {
      double r = std::sqrt(std::max(predicted, 0.0)) - std::sqrt(std::max(measured, 0.0));
      return r * r;
    }
  }
}

Config make_config(const fpmhip_algorithm* a) {
  Config c{};
  c.object_step = a->object_step;
  c.epsilon = a->epsilon;
  c.loss = a->loss_type;
  c.constrain = true;
  switch (a->kind) {
    case FPMHIP_ALTERNATING_PROJECTION:  // alternating_projection.rs:108-132
      c.denom = FPMHIP_DENOMINATOR_LOCAL;
      break;
    case FPMHIP_FPIE:  // fpie.rs:111-131
      c.denom = FPMHIP_DENOMINATOR_RPIE;
      c.alpha = a->stability;
      break;
    case FPMHIP_MPIE:  // mpie.rs:298-327
      c.denom = FPMHIP_DENOMINATOR_RPIE;
      c.alpha = a->stability;
      c.momentum = true;
      c.interval = a->momentum_interval;
      c.friction = a->momentum_friction;
      c.feedback = a->momentum_feedback;
      break;
    case FPMHIP_EPRY:  // epry.rs:251-287
      c.denom = FPMHIP_DENOMINATOR_GLOBAL;
      c.has_pupil_step = a->recover_pupil != 0;
      c.pupil_step = a->pupil_step;
      c.constrain = a->constrain_pupil != 0;
      c.gain_update = a->recover_gain != 0;
      c.gain_step = a->gain_step;
      c.gain_min = a->minimum_gain;
      c.gain_max = a->maximum_gain;
      c.bg_update = a->recover_background != 0;
      c.bg_step = a->background_step;
      c.bg_min = a->minimum_background;
      c.bg_max = a->maximum_background;
      break;
    default:
      fail(FPMHIP_INVALID_PARAMETER, "unknown algorithm kind", "kind");
  }
  return c;
}

// apply_momentum_event (src/algorithms/common.rs:355-387)
void momentum_event(C* object, C* velocity, C* anchor, uint64_t n, double friction, double feedback) {
  for (uint64_t i = 0; i < n; ++i) {
    C nv = cadd(cscale(friction, velocity[i]), csub(object[i], anchor[i]));
    C no = feedback == 0.0 ? object[i] : cadd(object[i], cscale(feedback, nv));
    if (!cfinite(nv) || !cfinite(no))
      fail(FPMHIP_NUMERICAL, "mPIE momentum produced a non-finite object or velocity");
    velocity[i] = nv;
    object[i] = no;
    anchor[i] = no;
  }
}

struct Scratch {
  std::vector<C> patch, exit, field, projected, projected_spectrum, difference, pupil_gradient;
  std::vector<double> prediction;
  std::vector<C> mfields, mpatches;
  std::vector<double> moffr, moffc;
};

// projection_update (src/algorithms/common.rs:57-306) for one frame; returns
// the frame objective through *loss_out and whether the frame was counted.
void update_frame(const Model& m, const Meas& meas, const Config& cfg, const Plan2& plow, State& st,
                  Scratch& sc, uint64_t frame, double* loss_out, double* weight_out) {
  const uint64_t N = m.N(), nh = m.nh, nw = m.nw;
  const double frame_weight = meas.weight[frame];
  *weight_out = frame_weight;
  if (frame_weight == 0.0) {
    *loss_out = 0.0;
    return;
  }
  // sources of the frame (multiplexing row or the identity)
  std::vector<std::pair<uint64_t, double>> sources;
  if (m.mux)
    for (uint64_t e = m.row_ptr[frame]; e < m.row_ptr[frame + 1]; ++e) sources.push_back({m.msrc[e], m.mw[e]});
  else
    sources.push_back({frame, 1.0});
  double source_weight_sum = 0.0;
  for (auto& s : sources) source_weight_sum += s.second;
  const uint64_t K = sources.size();
  sc.mfields.resize(K * N);
  sc.mpatches.resize(K * N);
  sc.moffr.resize(K);
  sc.moffc.resize(K);
  std::fill(sc.prediction.begin(), sc.prediction.end(), 0.0);
  for (uint64_t k = 0; k < K; ++k) {
    uint64_t s = sources[k].first;
    double sw = sources[k].second;
    sc.moffr[k] = m.off_r[s];
    sc.moffc[k] = m.off_c[s];
    extract(st.spectrum.data(), m.rh, m.rw, m.crop_r[s], m.crop_c[s], nh, nw, m.off_r[s], m.off_c[s],
            sc.patch.data());
    std::copy(sc.patch.begin(), sc.patch.end(), sc.mpatches.begin() + k * N);
    for (uint64_t p = 0; p < N; ++p) sc.exit[p] = cmul(sc.patch[p], st.pupil[p]);
    ifftshift(sc.exit.data(), sc.field.data(), nh, nw);
    fft2(plow, sc.field.data(), true);
    std::copy(sc.field.begin(), sc.field.end(), sc.mfields.begin() + k * N);
    for (uint64_t p = 0; p < N; ++p) sc.prediction[p] += sw * cnorm_sqr(sc.field[p]);
  }
  const double* measured = meas.frames + frame * N;
  const uint8_t* mask = meas.masks ? meas.masks + frame * N : nullptr;
  double gain = st.has_gain ? st.gain[frame] : 1.0;
  if (!std::isfinite(gain) || gain <= 0.0) fail(FPMHIP_INVALID_MODEL, "state frame has invalid gain");
  if (cfg.gain_update) {  // update_gain (common.rs:390-422)
    double num = 0.0, den = 0.0;
    for (uint64_t p = 0; p < N; ++p) {
      if (mask && mask[p] == 0) continue;
      double bg = bg_value(st.background, N, frame, p);
      double pred = sc.prediction[p];
      num += pred * std::max(measured[p] - bg, 0.0);
      den += pred * pred;
    }
    if (den > cfg.epsilon) {
      double est = std::clamp(num / den, cfg.gain_min, cfg.gain_max);
      gain = std::clamp((1.0 - cfg.gain_step) * gain + cfg.gain_step * est, cfg.gain_min, cfg.gain_max);
      if (!st.has_gain) {
        st.has_gain = true;
        st.gain.assign(m.F, 1.0);
      }
      st.gain[frame] = gain;
    }
  }
  if (cfg.bg_update) {  // update_background (common.rs:425-489)
    double rsum = 0.0;
    uint64_t rcount = 0;
    for (uint64_t p = 0; p < N; ++p) {
      if (mask && mask[p] == 0) continue;
      rsum += measured[p] - gain * sc.prediction[p] - bg_value(st.background, N, frame, p);
      rcount += 1;
    }
    if (rcount > 0) {
      double correction = cfg.bg_step * rsum / (double)rcount;
      std::vector<double> bg;
      if (st.background.empty())
        bg.assign(N * m.F, 0.0);
      else if (st.background.size() == N) {
        bg.reserve(N * m.F);
        for (uint64_t f = 0; f < m.F; ++f) bg.insert(bg.end(), st.background.begin(), st.background.end());
      } else
        bg = st.background;
      for (uint64_t p = 0; p < N; ++p) {
        double& v = bg[frame * N + p];
        v = std::clamp(v + correction, cfg.bg_min, cfg.bg_max);
      }
      st.background.swap(bg);
    }
  }
  double frame_loss = 0.0;
  uint64_t valid = 0;
  for (uint64_t p = 0; p < N; ++p) {
    double intrinsic = sc.prediction[p];
    if (mask && mask[p] == 0) {
      sc.projected[p] = {1.0, 0.0};
      continue;
    }
    valid += 1;
    double bg = bg_value(st.background, N, frame, p);
    double predicted = gain * intrinsic + bg;
    frame_loss += point_loss(predicted, measured[p], cfg.loss);
    double target = std::max((measured[p] - bg) / gain, 0.0);
    sc.projected[p] = intrinsic > cfg.epsilon ? C{std::sqrt(target / intrinsic), 0.0}
                                              : C{0.0, std::sqrt(target / source_weight_sum)};
  }
  if (valid == 0) fail(FPMHIP_INVALID_MEASUREMENTS, "frame has no unmasked pixels");
  *loss_out = frame_loss / (double)valid;

  double max_pupil_power = 0.0;
  for (uint64_t p = 0; p < N; ++p) max_pupil_power = std::max(max_pupil_power, cnorm_sqr(st.pupil[p]));
  max_pupil_power = std::max(max_pupil_power, cfg.epsilon);
  if (cfg.has_pupil_step) std::fill(sc.pupil_gradient.begin(), sc.pupil_gradient.end(), C{0.0, 0.0});
  for (uint64_t k = 0; k < K; ++k) {
    uint64_t s = sources[k].first;
    double sw = sources[k].second;
    std::copy(sc.mpatches.begin() + k * N, sc.mpatches.begin() + (k + 1) * N, sc.patch.begin());
    std::copy(sc.mfields.begin() + k * N, sc.mfields.begin() + (k + 1) * N, sc.field.begin());
    for (uint64_t p = 0; p < N; ++p) {
      C proj = sc.projected[p];
      if (mask && mask[p] == 0)
        sc.difference[p] = sc.field[p];
      else if (sc.prediction[p] > cfg.epsilon)
        sc.difference[p] = cmulr(sc.field[p], proj.re);
      else
        sc.difference[p] = {proj.im, 0.0};
    }
    fft2(plow, sc.difference.data(), false);
    fftshift(sc.difference.data(), sc.projected_spectrum.data(), nh, nw);
    for (uint64_t p = 0; p < N; ++p)
      sc.difference[p] = csub(sc.projected_spectrum[p], cmul(sc.patch[p], st.pupil[p]));
    const double nsw = sw / source_weight_sum;
    for (uint64_t p = 0; p < N; ++p) {
      C pup = st.pupil[p];
      double local = cnorm_sqr(pup);
      double den;
      if (cfg.denom == FPMHIP_DENOMINATOR_LOCAL)
        den = local + cfg.epsilon;
      else if (cfg.denom == FPMHIP_DENOMINATOR_GLOBAL)
        den = max_pupil_power + cfg.epsilon;
      else
        den = (1.0 - cfg.alpha) * local + cfg.alpha * max_pupil_power + cfg.epsilon;
      sc.exit[p] = cdivr(cmul(cconj(pup), sc.difference[p]), den);
    }
    insert_adjoint(st.spectrum.data(), m.rh, m.rw, m.crop_r[s], m.crop_c[s], nh, nw, sc.exit.data(),
                   frame_weight * cfg.object_step * nsw, sc.moffr[k], sc.moffc[k]);
    if (cfg.has_pupil_step) {
      double max_obj = 0.0;
      for (uint64_t p = 0; p < N; ++p) max_obj = std::max(max_obj, cnorm_sqr(sc.patch[p]));
      max_obj = std::max(max_obj, cfg.epsilon);
      for (uint64_t p = 0; p < N; ++p)
        sc.pupil_gradient[p] = cadd(sc.pupil_gradient[p],
                                    cdivr(cmul(cscale(nsw, cconj(sc.patch[p])), sc.difference[p]),
                                          max_obj + cfg.epsilon));
    }
  }
  if (cfg.has_pupil_step) {
    for (uint64_t p = 0; p < N; ++p)
      st.pupil[p] = cadd(st.pupil[p], cscale(frame_weight * cfg.pupil_step, sc.pupil_gradient[p]));
    if (cfg.constrain)
      for (uint64_t p = 0; p < N; ++p)
        if (m.support[p] == 0) st.pupil[p] = {0.0, 0.0};
  }
  if (cfg.momentum) {  // apply_object_momentum (common.rs:308-353)
    st.counter += 1;
    if (st.counter >= cfg.interval) {
      if (st.counter != cfg.interval)
        fail(FPMHIP_INVALID_MODEL, "mPIE frame counter exceeds the configured momentum interval");
      momentum_event(st.spectrum.data(), st.velocity.data(), st.anchor.data(), m.M(), cfg.friction,
                     cfg.feedback);
      st.counter = 0;
    }
  }
}

// ReconstructionState::initialize_with_backend (src/reconstruction/state.rs:280-355)
void initialize(const Model& m, const Meas& meas, const Plan2& phigh, State& st) {
  const uint64_t N = m.N();
  std::vector<double> average(N, 0.0), aweight(N, 0.0);
  double total_amp = 0.0, total_w = 0.0;
  for (uint64_t f = 0; f < m.F; ++f) {
    double fw = meas.weight[f];
    if (fw == 0.0) continue;
    const double* frame = meas.frames + f * N;
    const uint8_t* mask = meas.masks ? meas.masks + f * N : nullptr;
    double gain = m.has_gain ? m.gain[f] : 1.0;
    for (uint64_t p = 0; p < N; ++p) {
      if (mask && mask[p] == 0) continue;
      double bg = bg_value(m.background, N, f, p);
      double amp = std::sqrt(std::max((frame[p] - bg) / gain, 0.0));
      average[p] += fw * amp;
      aweight[p] += fw;
      total_amp += fw * amp;
      total_w += fw;
    }
  }
  if (total_w == 0.0)
    fail(FPMHIP_INVALID_MEASUREMENTS, "no positive-weight, unmasked measurements are available");
  double fallback = total_amp / total_w;
  for (uint64_t p = 0; p < N; ++p) average[p] = aweight[p] > 0.0 ? average[p] / aweight[p] : fallback;
  std::vector<C> object(m.M());
  for (uint64_t r = 0; r < m.rh; ++r) {
    uint64_t lr = r * m.nh / m.rh;
    for (uint64_t c = 0; c < m.rw; ++c) {
      uint64_t lc = c * m.nw / m.rw;
      object[r * m.rw + c] = {average[lr * m.nw + lc], 0.0};
    }
  }
  fft2(phigh, object.data(), false);
  st.spectrum.resize(m.M());
  fftshift(object.data(), st.spectrum.data(), m.rh, m.rw);
  st.pupil = m.pupil;
  st.has_gain = m.has_gain;
  st.gain = m.gain;
  st.background = m.background;
}

// canonicalize_object_pupil (src/algorithms/gauge.rs:20-282)
double relative_phase_slope(const State& st, const Model& m, bool row_axis, double tol) {
  const uint64_t h = m.nh, w = m.nw;
  C resultant{0, 0}, strongest{0, 0};
  double total = 0.0, strongest_w = 0.0;
  uint64_t rl = row_axis ? (h ? h - 1 : 0) : h, cl = row_axis ? w : (w ? w - 1 : 0);
  for (uint64_t r = 0; r < rl; ++r)
    for (uint64_t c = 0; c < cl; ++c) {
      uint64_t cur = r * w + c, nxt = row_axis ? (r + 1) * w + c : r * w + c + 1;
      if (m.support[cur] == 0 || m.support[nxt] == 0) continue;
      C ce = cmul(st.pupil[nxt], cconj(st.pupil[cur]));
      C re = cmul(m.pupil[nxt], cconj(m.pupil[cur]));
      C contribution = cmul(ce, cconj(re));
      double weight = cnorm(contribution);
      if (!std::isfinite(weight)) continue;
      resultant = cadd(resultant, contribution);
      total += weight;
      if (weight > strongest_w) {
        strongest = contribution;
        strongest_w = weight;
      }
    }
  if (total == 0.0) return 0.0;
  if (cnorm(resultant) > tol * total) return carg(resultant);
  if (strongest_w > 0.0) return carg(strongest);
  return 0.0;
}

void canonicalize(const Model& m, State& st) {
  uint64_t support_count = 0;
  for (uint8_t s : m.support) support_count += s != 0;
  if (support_count == 0) fail(FPMHIP_INVALID_MODEL, "compiled reference pupil support is empty");
  const double tol = 64.0 * std::numeric_limits<double>::epsilon() * (double)support_count;
  bool row_exact = true, col_exact = true;
  for (uint64_t s = 0; s < m.S; ++s) {
    row_exact = row_exact && offset_is_zero(m.off_r[s], 0.0);
    col_exact = col_exact && offset_is_zero(0.0, m.off_c[s]);
  }
  double row_slope = row_exact ? relative_phase_slope(st, m, true, tol) : 0.0;
  double col_slope = col_exact ? relative_phase_slope(st, m, false, tol) : 0.0;
  if (std::fabs(row_slope) > tol || std::fabs(col_slope) > tol) {  // apply_affine_correction :132-156
    const uint64_t pc0 = m.nh / 2, pc1 = m.nw / 2;
    for (uint64_t r = 0; r < m.nh; ++r) {
      double y = (double)r - (double)pc0;
      for (uint64_t c = 0; c < m.nw; ++c) {
        double x = (double)c - (double)pc1;
        double ph = -(row_slope * y + col_slope * x);
        st.pupil[r * m.nw + c] = cmul(st.pupil[r * m.nw + c], cpolar(1.0, ph));
      }
    }
    const uint64_t oc0 = m.rh / 2, oc1 = m.rw / 2;
    for (uint64_t r = 0; r < m.rh; ++r) {
      double y = (double)r - (double)oc0;
      for (uint64_t c = 0; c < m.rw; ++c) {
        double x = (double)c - (double)oc1;
        double ph = row_slope * y + col_slope * x;
        st.spectrum[r * m.rw + c] = cmul(st.spectrum[r * m.rw + c], cpolar(1.0, ph));
      }
    }
  }
  // normalize_pupil_scale_and_phase (gauge.rs:158-240)
  double eref = 0.0, ecur = 0.0;
  for (uint64_t p = 0; p < m.N(); ++p)
    if (m.support[p]) {
      eref += cnorm_sqr(m.pupil[p]);
      ecur += cnorm_sqr(st.pupil[p]);
    }
  if (!std::isfinite(eref) || eref <= 0.0)
    fail(FPMHIP_INVALID_MODEL, "compiled reference pupil has invalid supported energy");
  if (!std::isfinite(ecur) || ecur <= 0.0)
    fail(FPMHIP_NUMERICAL, "object/pupil gauge normalization requires positive finite recovered pupil energy");
  double computed = std::sqrt(eref / ecur);
  if (!std::isfinite(computed) || computed <= 0.0)
    fail(FPMHIP_NUMERICAL, "object/pupil gauge normalization produced an invalid pupil scale");
  double scale = std::fabs(computed - 1.0) <= tol ? 1.0 : computed;
  C overlap{0, 0}, strongest{0, 0};
  double ow = 0.0, sw = 0.0;
  for (uint64_t p = 0; p < m.N(); ++p) {
    if (!m.support[p]) continue;
    C contribution = cmul(cconj(m.pupil[p]), st.pupil[p]);
    double weight = cnorm(contribution);
    overlap = cadd(overlap, contribution);
    ow += weight;
    if (weight > sw) {
      strongest = contribution;
      sw = weight;
    }
  }
  double phase;
  if (ow > 0.0 && cnorm(overlap) > tol * ow)
    phase = carg(overlap);
  else if (sw > 0.0)
    phase = carg(strongest);
  else
    fail(FPMHIP_NUMERICAL, "object/pupil gauge normalization could not anchor recovered pupil phase");
  if (!std::isfinite(phase)) fail(FPMHIP_NUMERICAL, "object/pupil gauge normalization produced a non-finite pupil phase");
  if (std::fabs(phase) <= tol) phase = 0.0;
  C pf = cpolar(scale, -phase);
  C of = cdiv(C{1.0, 0.0}, pf);
  if (!(pf.re == 1.0 && pf.im == 0.0)) {
    for (auto& v : st.pupil) v = cmul(v, pf);
    for (auto& v : st.spectrum) v = cmul(v, of);
  }
  // normalize_object_phase (gauge.rs:242-282)
  C strongest_o{0, 0};
  double smag = 0.0;
  for (const C& v : st.spectrum) {
    if (!cfinite(v)) fail(FPMHIP_NUMERICAL, "object/pupil gauge normalization requires a finite object spectrum");
    double mag = cnorm(v);
    if (mag > smag) {
      strongest_o = v;
      smag = mag;
    }
  }
  if (smag == 0.0) return;
  C dc = st.spectrum[(m.rh / 2) * m.rw + m.rw / 2];
  C anchor = cnorm(dc) > tol * smag ? dc : strongest_o;
  double oph = carg(anchor);
  if (!std::isfinite(oph)) fail(FPMHIP_NUMERICAL, "object/pupil gauge normalization produced a non-finite object phase");
  if (std::fabs(oph) <= tol) return;
  C corr = cpolar(1.0, -oph);
  for (auto& v : st.spectrum) v = cmul(v, corr);
}

// FrameSchedule::order / order_for_problem (src/reconstruction/schedule.rs:35-146)
void frame_vector(const Model& m, uint64_t f, double* kx, double* ky) {
  if (m.mux) {
    double vx = 0.0, vy = 0.0, tw = 0.0;
    for (uint64_t e = m.row_ptr[f]; e < m.row_ptr[f + 1]; ++e) {
      vx += m.mw[e] * m.kx[m.msrc[e]];
      vy += m.mw[e] * m.ky[m.msrc[e]];
      tw += m.mw[e];
    }
    *kx = vx / tw;
    *ky = vy / tw;
  } else {
    *kx = m.kx[f];
    *ky = m.ky[f];
  }
}

double rem_euclid_tau(double v) {
  const double tau = 6.283185307179586;
  double r = std::fmod(v, tau);
  return r < 0.0 ? r + tau : r;
}

std::vector<uint64_t> schedule_order(const Model& m, const Meas& meas, int schedule) {
  std::vector<uint64_t> order(m.F);
  for (uint64_t f = 0; f < m.F; ++f) order[f] = f;
  if (schedule == FPMHIP_SCHEDULE_SEQUENTIAL) return order;
  if (schedule == FPMHIP_SCHEDULE_SNR_WEIGHTED) {
    std::vector<double> score(m.F);
    const uint64_t N = m.N();
    for (uint64_t f = 0; f < m.F; ++f) {
      if (meas.weight[f] == 0.0) {
        score[f] = -std::numeric_limits<double>::infinity();
        continue;
      }
      double sum = 0.0;
      uint64_t valid = 0;
      for (uint64_t p = 0; p < N; ++p)
        if (!meas.masks || meas.masks[f * N + p] != 0) {
          sum += std::max(meas.frames[f * N + p] - bg_value(m.background, N, f, p), 0.0);
          valid += 1;
        }
      score[f] = valid == 0 ? -std::numeric_limits<double>::infinity()
                            : meas.weight[f] * std::sqrt(sum / (double)valid);
    }
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return score[a] > score[b]; });
    return order;
  }
  if (!m.has_k) fail(FPMHIP_INVALID_MODEL, "schedule requires k-vectors");
  std::vector<double> rad(m.F), ang(m.F);
  for (uint64_t f = 0; f < m.F; ++f) {
    double kx, ky;
    frame_vector(m, f, &kx, &ky);
    rad[f] = std::hypot(kx, ky);
    ang[f] = rem_euclid_tau(std::atan2(ky, kx));
  }
  if (schedule == FPMHIP_SCHEDULE_BRIGHTFIELD_FIRST)
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return rad[a] < rad[b]; });
  else if (schedule == FPMHIP_SCHEDULE_SPIRAL_OUT)
    std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
      if (rad[a] != rad[b]) return rad[a] < rad[b];
      return ang[a] < ang[b];
    });
  else
    fail(FPMHIP_UNSUPPORTED, "schedule is not supported");
  return order;
}

// ForwardModel::forward_intensity_stack_into (src/model/forward.rs:245-372)
void forward_stack(const Model& m, const C* spectrum, const C* pupil, double* out) {
  const uint64_t N = m.N();
  Plan2 plow = make_plan2(m.nh, m.nw);
  std::vector<C> patch(N), exit(N), field(N);
  for (uint64_t f = 0; f < m.F; ++f) {
    double* dst = out + f * N;
    double gain = m.has_gain ? m.gain[f] : 1.0;
    if (!std::isfinite(gain) || gain <= 0.0) fail(FPMHIP_INVALID_MODEL, "frame has invalid gain");
    std::fill(dst, dst + N, 0.0);
    auto source_field = [&](uint64_t s) {
      extract(spectrum, m.rh, m.rw, m.crop_r[s], m.crop_c[s], m.nh, m.nw, m.off_r[s], m.off_c[s], patch.data());
      for (uint64_t p = 0; p < N; ++p) exit[p] = cmul(patch[p], pupil[p]);
      ifftshift(exit.data(), field.data(), m.nh, m.nw);
      fft2(plow, field.data(), true);
    };
    if (m.mux) {
      for (uint64_t e = m.row_ptr[f]; e < m.row_ptr[f + 1]; ++e) {
        source_field(m.msrc[e]);
        for (uint64_t p = 0; p < N; ++p) dst[p] += m.mw[e] * cnorm_sqr(field[p]);
      }
    } else {
      source_field(f);
      for (uint64_t p = 0; p < N; ++p) dst[p] = cnorm_sqr(field[p]);
    }
    for (uint64_t p = 0; p < N; ++p) dst[p] = gain * dst[p] + bg_value(m.background, N, f, p);
  }
}

// ---------------------------------------------------------------------------
// ImagePlaneModel::from_experiment for a planar LED array
// (src/model/image_plane_fpm.rs:167-349, 750-797, 975-1198; pupil.rs:78-124;
//  led_array.rs:87-113, 292-313; illumination.rs:385-421; optics.rs:128-204)

struct GridShift {
  int64_t integer;
  double offset;
  int64_t lower, upper;
};

GridShift checked_grid_shift(double value) {  // image_plane_fpm.rs:975-1024
  if (!std::isfinite(value)) fail(FPMHIP_INVALID_MODEL, "illumination shift must be finite");
  double rounded = std::round(value);
  int64_t integer = (int64_t)rounded;
  double offset = value - rounded;
  double nearest = std::round(offset);
  GridShift g{integer, offset, 0, 0};
  if (std::fabs(offset - nearest) <= 1e-12) {
    g.lower = g.upper = integer + (int64_t)nearest;
  } else if (offset > 0.0) {
    g.lower = integer;
    g.upper = integer + 1;
  } else {
    g.lower = integer - 1;
    g.upper = integer;
  }
  return g;
}

bool axis_contains(uint64_t n, uint64_t R, int64_t dmin, int64_t dmax) {  // :1089-1114
  int64_t base = (int64_t)(R / 2) - (int64_t)(n / 2);
  int64_t lower = base + dmin;
  int64_t upper = base + (int64_t)(n - 1) + dmax;
  return lower >= 0 && upper < (int64_t)R;
}

uint64_t next_smooth(uint64_t minimum) {  // :1176-1198
  uint64_t best = 0;
  struct V {
    static void visit(uint64_t value, int idx, uint64_t minimum, uint64_t& best) {
      if (value >= minimum) {
        if (best == 0 || value < best) best = value;
        return;
      }
      static const uint64_t primes[4] = {2, 3, 5, 7};
      for (int i = idx; i < 4; ++i) {
        uint64_t next = value * primes[i];
        if (best == 0 || next < best) visit(next, i, minimum, best);
      }
    }
  };
  V::visit(1, 0, minimum, best);
  return best;
}

uint64_t gcd(uint64_t a, uint64_t b) {
  while (b) {
    uint64_t r = a % b;
    a = b;
    b = r;
  }
  return a;
}

void validate_optics(const fpmhip_optics* o) {  // optics.rs:128-193
  const char* names[6] = {"wavelength_vacuum_m", "objective_na", "magnification", "camera_pixel_size",
                          "illumination_refractive_index", "objective_medium_refractive_index"};
  double vals[6] = {o->wavelength_vacuum_m, o->objective_na, o->magnification, o->camera_pixel_size,
                    o->illumination_refractive_index, o->objective_medium_refractive_index};
  for (int i = 0; i < 6; ++i)
    if (!std::isfinite(vals[i]) || vals[i] <= 0.0) fail(FPMHIP_INVALID_PARAMETER, "must be finite and positive", names[i]);
  if (o->objective_na > o->objective_medium_refractive_index)
    fail(FPMHIP_INVALID_PARAMETER, "cannot exceed the medium refractive index", "objective_na");
  double px = o->camera_pixel_size / o->magnification;
  double limit = o->wavelength_vacuum_m / (2.0 * o->objective_na);
  if (px >= limit) fail(FPMHIP_INVALID_PARAMETER, "object-plane pitch must be below the coherent sampling limit", "camera_pixel_size");
}

struct Compiled {
  uint64_t S, rh, rw;
  std::vector<uint64_t> crop;
  std::vector<double> offset, k;
  std::vector<C> pupil;
  std::vector<uint8_t> support;
};

Compiled compile_planar(const fpmhip_optics* o, const fpmhip_planar_array* a, uint64_t nh, uint64_t nw,
                        int policy, uint64_t eh, uint64_t ew) {
  validate_optics(o);
  if (nh == 0 || nw == 0) fail(FPMHIP_INVALID_SHAPE, "image shape must be non-zero");
  if (a->rows == 0 || a->columns == 0) fail(FPMHIP_INVALID_PARAMETER, "rows and columns must be non-zero", "shape");
  Compiled out;
  // PlanarLedArray::resolve (led_array.rs:292-313) + ArrayPose::transform_point (:87-113)
  std::vector<double> kv;
  const double k0 = 6.283185307179586 * o->illumination_refractive_index / o->wavelength_vacuum_m;
  for (uint64_t r = 0; r < a->rows; ++r)
    for (uint64_t c = 0; c < a->columns; ++c) {
      double local[3] = {((double)c - a->reference_column) * a->pitch_x_m,
                         ((double)r - a->reference_row) * a->pitch_y_m, 0.0};
      double sx = std::sin(a->rotation_rad[0]), cx = std::cos(a->rotation_rad[0]);
      double ax[3] = {local[0], cx * local[1] - sx * local[2], sx * local[1] + cx * local[2]};
      double sy = std::sin(a->rotation_rad[1]), cy = std::cos(a->rotation_rad[1]);
      double ay[3] = {cy * ax[0] + sy * ax[2], ax[1], -sy * ax[0] + cy * ax[2]};
      double sz = std::sin(a->rotation_rad[2]), cz = std::cos(a->rotation_rad[2]);
      double rot[3] = {cz * ay[0] - sz * ay[1], sz * ay[0] + cz * ay[1], ay[2]};
      double p[3] = {rot[0] + a->translation_m[0], rot[1] + a->translation_m[1], rot[2] + a->translation_m[2]};
      // ResolvedSources::from_positions (illumination.rs:385-421)
      double norm = std::hypot(std::hypot(p[0], p[1]), p[2]);
      if (!std::isfinite(norm) || norm <= 0.0)
        fail(FPMHIP_INVALID_PARAMETER, "no source may coincide with the sample origin", "positions_m");
      double d0 = -p[0] / norm, d1 = -p[1] / norm;
      kv.push_back(k0 * d0);
      kv.push_back(k0 * d1);
    }
  const uint64_t S = kv.size() / 2;
  out.S = S;
  const double px = o->camera_pixel_size / o->magnification;
  const double dkx = 6.283185307179586 / ((double)nw * px);
  const double dky = 6.283185307179586 / ((double)nh * px);
  // crop_displacement_bounds (:255-285)
  int64_t minr = INT64_MAX, maxr = INT64_MIN, minc = INT64_MAX, maxc = INT64_MIN;
  for (uint64_t s = 0; s < S; ++s) {
    GridShift gr = checked_grid_shift(kv[2 * s + 1] / dky);
    GridShift gc = checked_grid_shift(kv[2 * s] / dkx);
    minr = std::min(minr, gr.lower);
    maxr = std::max(maxr, gr.upper);
    minc = std::min(minc, gc.lower);
    maxc = std::max(maxc, gc.upper);
  }
  auto fits = [&](uint64_t h, uint64_t w) {
    return axis_contains(nh, h, minr, maxr) && axis_contains(nw, w, minc, maxc);
  };
  // resolve_reconstruction_shape (:287-348)
  uint64_t rh, rw;
  if (policy == FPMHIP_SHAPE_EXACT) {
    if (eh < nh || ew < nw) fail(FPMHIP_INVALID_SHAPE, "image shape must fit reconstruction shape");
    if (eh * nw != ew * nh) fail(FPMHIP_INVALID_SHAPE, "reconstruction must use the same scale factor in both dimensions");
    if (!fits(eh, ew)) fail(FPMHIP_INVALID_SHAPE, "reconstruction shape does not contain every illumination crop and subpixel interpolation stencil");
    rh = eh;
    rw = ew;
  } else {
    uint64_t div = gcd(nh, nw), ah = nh / div, aw = nw / div;
    uint64_t maxm = (uint64_t)INT64_MAX / std::max(ah, aw);
    uint64_t mult = div;
    if (!fits(ah * mult, aw * mult)) {  // minimum_fitting_multiplier (:1116-1155)
      uint64_t lower = mult, upper = mult;
      for (;;) {
        upper = std::min(upper * 2, maxm);
        if (upper == lower) fail(FPMHIP_INVALID_SHAPE, "illumination crops require a reconstruction shape outside the supported index range");
        if (fits(ah * upper, aw * upper)) break;
        lower = upper;
      }
      while (lower + 1 < upper) {
        uint64_t mid = lower + (upper - lower) / 2;
        if (fits(ah * mid, aw * mid))
          upper = mid;
        else
          lower = mid;
      }
      mult = upper;
    }
    if (policy == FPMHIP_SHAPE_SMOOTH)
      mult = next_smooth(mult);
    else if (policy == FPMHIP_SHAPE_POWER_OF_TWO) {
      uint64_t p = 1;
      while (p < mult) p <<= 1;
      mult = p;
    } else if (policy != FPMHIP_SHAPE_MINIMUM)
      fail(FPMHIP_INVALID_PARAMETER, "unknown reconstruction shape policy", "reconstruction_shape");
    rh = ah * mult;
    rw = aw * mult;
  }
  out.rh = rh;
  out.rw = rw;
  // Pupil::circular (pupil.rs:78-124) on the low-resolution grid
  const double cutoff = 6.283185307179586 * o->objective_na / o->wavelength_vacuum_m;
  const double medium_k = 6.283185307179586 * o->objective_medium_refractive_index / o->wavelength_vacuum_m;
  out.pupil.resize(nh * nw);
  out.support.resize(nh * nw);
  for (uint64_t r = 0; r < nh; ++r) {
    double ky = ((double)r - (double)(nh / 2)) * dky;
    for (uint64_t c = 0; c < nw; ++c) {
      double kx = ((double)c - (double)(nw / 2)) * dkx;
      double radius = std::hypot(kx, ky);
      bool inside = radius <= cutoff;
      out.support[r * nw + c] = inside ? 1 : 0;
      if (!inside) {
        out.pupil[r * nw + c] = {0.0, 0.0};
        continue;
      }
      double rho = cutoff > 0.0 ? radius / cutoff : 0.0;
      double theta = std::atan2(ky, kx);
      double phase = 0.0;
      if (o->has_defocus) phase -= o->defocus_distance * (kx * kx + ky * ky) / (2.0 * medium_k);
      double amplitude = 1.0;
      if (o->has_aberration) {
        phase += o->astigmatism * rho * rho * std::cos(2.0 * theta);
        phase += o->coma * (3.0 * std::pow(rho, 3) - 2.0 * rho) * std::cos(theta);
        phase += o->spherical * (6.0 * std::pow(rho, 4) - 6.0 * rho * rho + 1.0);
        amplitude = std::exp(-o->edge_apodization * rho * rho);
      }
      out.pupil[r * nw + c] = cpolar(amplitude, phase);
    }
  }
  // crop_for_k_vector (:750-797)
  out.crop.resize(2 * S);
  out.offset.resize(2 * S);
  out.k = kv;
  for (uint64_t s = 0; s < S; ++s) {
    GridShift gr = checked_grid_shift(kv[2 * s + 1] / dky);
    GridShift gc = checked_grid_shift(kv[2 * s] / dkx);
    int64_t sr = (int64_t)(rh / 2) - (int64_t)(nh / 2) + gr.integer;
    int64_t scl = (int64_t)(rw / 2) - (int64_t)(nw / 2) + gc.integer;
    if (sr < 0 || scl < 0) fail(FPMHIP_INVALID_MODEL, "illumination vector produces a crop outside the reconstruction grid");
    out.crop[2 * s] = (uint64_t)sr;
    out.crop[2 * s + 1] = (uint64_t)scl;
    out.offset[2 * s] = gr.offset;
    out.offset[2 * s + 1] = gc.offset;
    interpolation_axis((uint64_t)sr, nh, rh, gr.offset);
    interpolation_axis((uint64_t)scl, nw, rw, gc.offset);
  }
  return out;
}

// ---------------------------------------------------------------------------
// Camera (src/simulation/camera.rs:248-327). The deterministic steps restate
// measure_frame; the noise is not the reference's (rand's StdRng / ChaCha12
// stream and rand_distr's samplers are not vendored in the reference, so its
// draws cannot be reproduced): it restates the product's own generator —
// Philox4x64-10 (Salmon et al., SC 2011) with counter (pixel, frame, block, 0)
// and key (seed, 0), uniforms (x >> 11) 2^-53, Box-Muller for read noise,
// inversion below mean 10 and Hoermann's PTRS (1993) above — so that device
// draws can be checked draw for draw. Pins: Philox against numpy's Philox
// bit generator; the samplers' moments and distribution against the exact
// Poisson / normal laws; the deterministic steps against the reference's own
// camera tests (tests/simulation_tests.rs:77-122,208-222,554-573).

struct Philox {
  uint64_t v[4];
};

Philox philox4x64_10(Philox c, uint64_t k0, uint64_t k1) {
  const uint64_t m0 = 0xD2E7470EE14C6C93ull, m1 = 0xCA5A826395121157ull;
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B97F4A7C15ull;
      k1 += 0xBB67AE8584CAA73Bull;
    }
    const unsigned __int128 p0 = (unsigned __int128)m0 * c.v[0];
    const unsigned __int128 p1 = (unsigned __int128)m1 * c.v[2];
    const uint64_t hi0 = (uint64_t)(p0 >> 64), lo0 = (uint64_t)p0;
    const uint64_t hi1 = (uint64_t)(p1 >> 64), lo1 = (uint64_t)p1;
    c = Philox{{hi1 ^ c.v[1] ^ k0, lo1, hi0 ^ c.v[3] ^ k1, lo0}};
  }
  return c;
}

struct Uniforms {
  uint64_t pixel, frame, seed, block = 0;
  int used = 4;
  Philox words{};
  void at_block(uint64_t b) {
    block = b;
    used = 4;
  }
  double next() {
    if (used == 4) {
      words = philox4x64_10(Philox{{pixel, frame, block, 0}}, seed, 0);
      ++block;
      used = 0;
    }
    return (double)(words.v[used++] >> 11) * 0x1.0p-53;
  }
};

double draw_poisson(double mean, Uniforms& u) {
  u.at_block(1);
  if (mean < 10.0) {
    const double x = u.next();
    double term = std::exp(-mean), cdf = term;
    int k = 0;
    while (x >= cdf && k < 1000) {
      ++k;
      term = term * mean / (double)k;
      cdf = cdf + term;
    }
    return (double)k;
  }
  const double root = std::sqrt(mean), lmean = std::log(mean);
  const double b = 0.931 + 2.53 * root;
  const double a = -0.059 + 0.02483 * b;
  const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
  const double v_r = 0.9277 - 3.6224 / (b - 2.0);
  for (int t = 0; t < 256; ++t) {
    const double U = u.next() - 0.5;
    const double V = u.next();
    const double us = 0.5 - std::fabs(U);
    const double k = std::floor((2.0 * a / us + b) * U + mean + 0.43);
    if (us >= 0.07 && V <= v_r) return k;
    if (k < 0.0 || (us < 0.013 && V > us)) continue;
    if (std::log(V) + std::log(inv_alpha) - std::log(a / (us * us) + b) <= -mean + k * lmean - std::lgamma(k + 1.0))
      return k;
  }
  return std::floor(mean);
}

double draw_normal(Uniforms& u) {
  u.at_block(0);
  const double u1 = 1.0 - u.next();
  const double u2 = u.next();
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

// CameraModel::maximum_count (camera.rs:320-327)
double camera_maximum(const fpmhip_camera& c) {
  const double inf = std::numeric_limits<double>::infinity();
  const double digital = c.has_bit_depth ? std::pow(2.0, c.bit_depth) - 1.0 : inf;
  return std::min(c.has_saturation ? c.saturation_counts : inf, digital);
}

// f64::clamp(0, hi): NaN passes through
double clamp0(double x, double hi) { return x < 0.0 ? 0.0 : (x > hi ? hi : x); }

// CameraModel::measure_frame (camera.rs:248-287) on frame `frame` of n pixels
void camera_measure_frame(const fpmhip_camera& c, double* v, uint64_t n, uint64_t frame, uint64_t seed) {
  const double top = camera_maximum(c);
  for (uint64_t p = 0; p < n; ++p) {
    Uniforms u{p, frame, seed};
    const double sensitivity = c.pixel_sensitivity ? c.pixel_sensitivity[p] : 1.0;
    const double expected = std::fmax(v[p], 0.0) * sensitivity * c.photons_per_pixel + c.dark_current_electrons;
    double electrons = (c.shot_noise && expected > 0.0) ? draw_poisson(expected, u) : expected;
    if (c.read_noise_electrons > 0.0) electrons += 0.0 + c.read_noise_electrons * draw_normal(u);
    double counts = electrons * c.gain_counts_per_electron + c.offset_counts;
    counts = clamp0(counts, top);
    if (c.quantize) counts = std::round(counts);
    v[p] = counts;
  }
  for (uint64_t i = 0; i < c.bad_pixel_count; ++i) {
    double value = clamp0(c.has_bad_pixel_value ? c.bad_pixel_value_counts : top, top);
    if (c.quantize) value = std::round(value);
    v[c.bad_pixels[i]] = value;
  }
}

template <class F>
int guarded(F&& body) {
  g_error.clear();
  g_error_name.clear();
  try {
    body();
    return FPMHIP_OK;
  } catch (const Failure& f) {
    return f.code;
  } catch (const std::bad_alloc&) {
    g_error = "host allocation failed";
    return FPMHIP_OUT_OF_MEMORY;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// C interface of the oracle (loaded with ctypes by tests/ and bench.py's
// cpu_baseline leg).

extern "C" {

typedef struct orc_resume {
  const double* spectrum;  // [M] complex
  const double* pupil;     // [N] complex
  const double* velocity;  // [M] complex or NULL
  const double* anchor;    // [M] complex or NULL
  uint64_t counter;
  const double* frame_gains;  // [F] or NULL
  const double* background;   // or NULL
  uint64_t background_len;
} orc_resume;

typedef struct orc_outputs {
  double* spectrum;         // [M] complex
  double* pupil;            // [N] complex
  double* object;           // [M] complex: state_object (result.rs:397-413)
  double* trace;            // [iterations]
  double* per_frame;        // [iterations * F] objective by frame index
  double* velocity;         // [M] complex
  double* anchor;           // [M] complex
  uint64_t* counter;
  double* frame_gains;      // [F]
  double* background;       // [F*N]
} orc_outputs;

const char* orc_last_error(void) { return g_error.c_str(); }
const char* orc_last_error_name(void) { return g_error_name.c_str(); }

// Philox4x64-10 of one counter block under a key (4 words out).
int orc_philox(const uint64_t* counter, const uint64_t* key, uint64_t* out) {
  return guarded([&] {
    Philox r = philox4x64_10(Philox{{counter[0], counter[1], counter[2], counter[3]}}, key[0], key[1]);
    std::copy(r.v, r.v + 4, out);
  });
}

// The camera's samplers on pixels 0..n-1 of frame 0: Poisson(mean) or, with
// mean < 0, standard normal draws.
int orc_camera_draws(double mean, uint64_t n, uint64_t seed, double* out) {
  return guarded([&] {
    for (uint64_t p = 0; p < n; ++p) {
      Uniforms u{p, 0, seed};
      out[p] = mean < 0.0 ? draw_normal(u) : draw_poisson(mean, u);
    }
  });
}

// CameraModel::measure_frame over `frames` frames of n pixels, in place.
// The camera is expected valid (the product validates).
int orc_camera_measure(const fpmhip_camera* camera, double* values, uint64_t frames, uint64_t n, uint64_t seed) {
  return guarded([&] {
    for (uint64_t f = 0; f < frames; ++f) camera_measure_frame(*camera, values + f * n, n, f, seed);
  });
}

// stats (src/metrics/intensity/single.rs:25-61): out = mean, std, min, max,
// sum, saturated_pixels, zero_pixels.
int orc_intensity_stats(const double* v, uint64_t n, int32_t has_saturation, double saturation, double* out) {
  return guarded([&] {
    if (n == 0) fail(FPMHIP_INVALID_SHAPE, "intensity statistics require non-empty values");
    double sum = 0.0, sq = 0.0, lo = std::numeric_limits<double>::infinity(), hi = -lo;
    uint64_t sat = 0, zero = 0;
    for (uint64_t i = 0; i < n; ++i) {
      const double x = v[i];
      if (!std::isfinite(x)) fail(FPMHIP_NUMERICAL, "intensity statistics contain a non-finite value");
      sum += x;
      sq += x * x;
      lo = std::min(lo, x);
      hi = std::max(hi, x);
      zero += x == 0.0;
      sat += has_saturation && x >= saturation;
    }
    const double mean = sum / (double)n;
    out[0] = mean;
    out[1] = std::sqrt(std::max(sq / (double)n - mean * mean, 0.0));
    out[2] = lo;
    out[3] = hi;
    out[4] = sum;
    out[5] = (double)sat;
    out[6] = (double)zero;
  });
}

// compare_intensity_u8_masked (src/metrics/intensity/compare.rs:134-226):
// out = reference_sum, estimate_sum, residual_l1, residual_l2, residual_mean,
// residual_std, residual_max_abs, normalized_l2, saturated_pixels.
int orc_compare_intensity(const double* reference, const double* estimate, const uint8_t* mask, uint64_t n,
                          int32_t has_saturation, double saturation, double* out) {
  return guarded([&] {
    if (n == 0) fail(FPMHIP_INVALID_SHAPE, "intensity comparison inputs have lengths 0 and 0");
    double rs = 0.0, es = 0.0, l1 = 0.0, r2 = 0.0, rsum = 0.0, rmax = 0.0, ref2 = 0.0;
    uint64_t sat = 0, count = 0;
    for (uint64_t i = 0; i < n; ++i) {
      if (mask && mask[i] == 0) continue;
      const double y = reference[i], e = estimate[i];
      if (!std::isfinite(y) || !std::isfinite(e))
        fail(FPMHIP_NUMERICAL, "intensity comparison contains a non-finite value");
      const double r = e - y;
      rs += y;
      es += e;
      l1 += std::fabs(r);
      r2 += r * r;
      rsum += r;
      rmax = std::max(rmax, std::fabs(r));
      ref2 += y * y;
      sat += has_saturation && y >= saturation;
      ++count;
    }
    if (count == 0) fail(FPMHIP_INVALID_PARAMETER, "must select at least one pixel", "valid_mask");
    const double c = (double)count, mean = rsum / c;
    out[0] = rs;
    out[1] = es;
    out[2] = l1;
    out[3] = std::sqrt(r2);
    out[4] = mean;
    out[5] = std::sqrt(std::max(r2 / c - mean * mean, 0.0));
    out[6] = rmax;
    out[7] = std::sqrt(r2) / (std::sqrt(ref2) + std::numeric_limits<double>::epsilon());
    out[8] = (double)sat;
  });
}

// The runner's per-frame objective of a state (runner.rs:361-373): mean of
// (sqrt(max(predicted, 0)) - sqrt(max(measured, 0)))^2 over unmasked pixels.
int orc_frame_objective(const double* predicted, const double* measured, const uint8_t* mask, uint64_t n,
                        double* out) {
  return guarded([&] {
    double sum = 0.0;
    uint64_t valid = 0;
    for (uint64_t i = 0; i < n; ++i)
      if (!mask || mask[i] != 0) {
        const double r = std::sqrt(std::max(predicted[i], 0.0)) - std::sqrt(std::max(measured[i], 0.0));
        sum += r * r;
        ++valid;
      }
    *out = valid == 0 ? 0.0 : sum / (double)valid;
  });
}

// point_loss (src/algorithms/objective.rs:47-71) of n value pairs.
int orc_point_loss(const double* predicted, const double* measured, uint64_t n, int32_t loss, double* out) {
  return guarded([&] {
    for (uint64_t i = 0; i < n; ++i) out[i] = point_loss(predicted[i], measured[i], loss);
  });
}

int orc_fft2(double* values, uint64_t h, uint64_t w, int32_t direction) {
  return guarded([&] {
    if (h == 0 || w == 0) fail(FPMHIP_INVALID_SHAPE, "FFT dimensions must be non-zero");
    Plan2 p = make_plan2(h, w);
    fft2(p, reinterpret_cast<C*>(values), direction == FPMHIP_INVERSE);
  });
}

int orc_shift(const double* src, double* dst, uint64_t h, uint64_t w, int32_t inverse) {
  return guarded([&] {
    if (inverse)
      ifftshift(reinterpret_cast<const C*>(src), reinterpret_cast<C*>(dst), h, w);
    else
      fftshift(reinterpret_cast<const C*>(src), reinterpret_cast<C*>(dst), h, w);
  });
}

int orc_extract(const double* spectrum, uint64_t rh, uint64_t rw, uint64_t sr, uint64_t sc, uint64_t h,
                uint64_t w, double orow, double ocol, double* out) {
  return guarded([&] {
    extract(reinterpret_cast<const C*>(spectrum), rh, rw, sr, sc, h, w, orow, ocol, reinterpret_cast<C*>(out));
  });
}

int orc_insert_adjoint(double* spectrum, uint64_t rh, uint64_t rw, uint64_t sr, uint64_t sc, uint64_t h,
                       uint64_t w, const double* update, double scale, double orow, double ocol) {
  return guarded([&] {
    insert_adjoint(reinterpret_cast<C*>(spectrum), rh, rw, sr, sc, h, w, reinterpret_cast<const C*>(update),
                   scale, orow, ocol);
  });
}

int orc_momentum_event(double* object, double* velocity, double* anchor, uint64_t n, double friction,
                       double feedback) {
  return guarded([&] {
    momentum_event(reinterpret_cast<C*>(object), reinterpret_cast<C*>(velocity), reinterpret_cast<C*>(anchor), n,
                   friction, feedback);
  });
}

int orc_compile_planar(const fpmhip_optics* optics, const fpmhip_planar_array* array, uint64_t image_h,
                       uint64_t image_w, int32_t policy, uint64_t exact_h, uint64_t exact_w,
                       fpmhip_compiled_shape* shape, uint64_t* crop_start, double* offset, double* k_vectors,
                       double* pupil, uint8_t* support, double* frame_gain) {
  return guarded([&] {
    Compiled c = compile_planar(optics, array, image_h, image_w, policy, exact_h, exact_w);
    shape->source_count = c.S;
    shape->recon_h = c.rh;
    shape->recon_w = c.rw;
    if (crop_start) std::copy(c.crop.begin(), c.crop.end(), crop_start);
    if (offset) std::copy(c.offset.begin(), c.offset.end(), offset);
    if (k_vectors) std::copy(c.k.begin(), c.k.end(), k_vectors);
    if (pupil) std::memcpy(pupil, c.pupil.data(), sizeof(C) * c.pupil.size());
    if (support) std::copy(c.support.begin(), c.support.end(), support);
    if (frame_gain) std::fill(frame_gain, frame_gain + c.S, 1.0);
  });
}

int orc_forward_stack(const fpmhip_model_desc* desc, const double* spectrum, const double* pupil, double* out) {
  return guarded([&] {
    Model m = load_model(desc);
    const C* pp = pupil ? reinterpret_cast<const C*>(pupil) : m.pupil.data();
    forward_stack(m, reinterpret_cast<const C*>(spectrum), pp, out);
  });
}

int orc_initialize(const fpmhip_model_desc* desc, const double* frames, const double* weights,
                   const uint8_t* masks, double* spectrum_out) {
  return guarded([&] {
    Model m = load_model(desc);
    Meas meas = load_meas(m, frames, weights, masks);
    State st;
    Plan2 ph = make_plan2(m.rh, m.rw);
    initialize(m, meas, ph, st);
    std::memcpy(spectrum_out, st.spectrum.data(), sizeof(C) * m.M());
  });
}

int orc_canonicalize(const fpmhip_model_desc* desc, double* spectrum, double* pupil) {
  return guarded([&] {
    Model m = load_model(desc);
    State st;
    st.spectrum.assign(reinterpret_cast<C*>(spectrum), reinterpret_cast<C*>(spectrum) + m.M());
    st.pupil.assign(reinterpret_cast<C*>(pupil), reinterpret_cast<C*>(pupil) + m.N());
    canonicalize(m, st);
    std::memcpy(spectrum, st.spectrum.data(), sizeof(C) * m.M());
    std::memcpy(pupil, st.pupil.data(), sizeof(C) * m.N());
  });
}

int orc_schedule(const fpmhip_model_desc* desc, const double* frames, const double* weights,
                 const uint8_t* masks, int32_t schedule, uint64_t* order) {
  return guarded([&] {
    Model m = load_model(desc);
    Meas meas = load_meas(m, frames, weights, masks);
    std::vector<uint64_t> o = schedule_order(m, meas, schedule);
    std::copy(o.begin(), o.end(), order);
  });
}

// Runner::run (src/reconstruction/runner.rs:76-285) with the algorithm's
// initialize (state.rs:280, mpie.rs:166) or a resumed state, `iterations`
// sweeps in schedule order (or explicit orders [iterations][F]), canonicalize
// after init and after each iteration (EPRY with pupil recovery), then the
// result object.
int orc_run(const fpmhip_model_desc* desc, const double* frames, const double* weights, const uint8_t* masks,
            const fpmhip_algorithm* algorithm, uint64_t iterations, int32_t schedule, const uint64_t* orders,
            const orc_resume* resume, orc_outputs* out) {
  return guarded([&] {
    Model m = load_model(desc);
    Meas meas = load_meas(m, frames, weights, masks);
    Config cfg = make_config(algorithm);
    const uint64_t N = m.N(), M = m.M();
    Plan2 plow = make_plan2(m.nh, m.nw), phigh = make_plan2(m.rh, m.rw);
    State st;
    if (resume) {
      st.spectrum.assign(reinterpret_cast<const C*>(resume->spectrum), reinterpret_cast<const C*>(resume->spectrum) + M);
      st.pupil.assign(reinterpret_cast<const C*>(resume->pupil), reinterpret_cast<const C*>(resume->pupil) + N);
      st.has_gain = resume->frame_gains != nullptr || m.has_gain;
      if (resume->frame_gains)
        st.gain.assign(resume->frame_gains, resume->frame_gains + m.F);
      else
        st.gain = m.gain;
      if (resume->background)
        st.background.assign(resume->background, resume->background + resume->background_len);
      else
        st.background = m.background;
    } else {
      initialize(m, meas, phigh, st);
    }
    if (cfg.momentum) {  // Mpie::prepare_auxiliary (mpie.rs:152-172)
      st.has_aux = true;
      if (resume && resume->velocity) {
        st.velocity.assign(reinterpret_cast<const C*>(resume->velocity), reinterpret_cast<const C*>(resume->velocity) + M);
        st.anchor.assign(reinterpret_cast<const C*>(resume->anchor), reinterpret_cast<const C*>(resume->anchor) + M);
        st.counter = resume->counter;
      } else {
        st.velocity.assign(M, C{0.0, 0.0});
        st.anchor = st.spectrum;
        st.counter = 0;
      }
    }
    const bool gauge = algorithm->kind == FPMHIP_EPRY && algorithm->recover_pupil;
    if (gauge) canonicalize(m, st);
    Scratch sc;
    sc.patch.resize(N);
    sc.exit.resize(N);
    sc.field.resize(N);
    sc.projected.resize(N);
    sc.projected_spectrum.resize(N);
    sc.difference.resize(N);
    sc.pupil_gradient.resize(N);
    sc.prediction.resize(N);
    std::vector<uint64_t> order = schedule_order(m, meas, schedule);
    for (uint64_t it = 0; it < iterations; ++it) {
      const uint64_t* ord = orders ? orders + it * m.F : order.data();
      double objective_sum = 0.0, weight_sum = 0.0;
      for (uint64_t i = 0; i < m.F; ++i) {
        uint64_t f = ord[i];
        if (f >= m.F) fail(FPMHIP_FRAME_OUT_OF_RANGE, "frame index out of range");
        double loss, w;
        update_frame(m, meas, cfg, plow, st, sc, f, &loss, &w);
        objective_sum += w * loss;  // StepSummary::push_frame (metrics.rs:25-31)
        weight_sum += w;
        if (out->per_frame) out->per_frame[it * m.F + f] = loss;
      }
      if (gauge) canonicalize(m, st);
      if (out->trace)
        out->trace[it] = weight_sum > 0.0 ? objective_sum / weight_sum : std::numeric_limits<double>::quiet_NaN();
    }
    if (out->spectrum) std::memcpy(out->spectrum, st.spectrum.data(), sizeof(C) * M);
    if (out->pupil) std::memcpy(out->pupil, st.pupil.data(), sizeof(C) * N);
    if (out->object) {
      std::vector<C> obj(M);
      ifftshift(st.spectrum.data(), obj.data(), m.rh, m.rw);
      fft2(phigh, obj.data(), true);
      std::memcpy(out->object, obj.data(), sizeof(C) * M);
    }
    if (cfg.momentum) {
      if (out->velocity) std::memcpy(out->velocity, st.velocity.data(), sizeof(C) * M);
      if (out->anchor) std::memcpy(out->anchor, st.anchor.data(), sizeof(C) * M);
      if (out->counter) *out->counter = st.counter;
    }
    if (out->frame_gains) {
      for (uint64_t f = 0; f < m.F; ++f) out->frame_gains[f] = st.has_gain ? st.gain[f] : 1.0;
    }
    if (out->background) {
      for (uint64_t f = 0; f < m.F; ++f)
        for (uint64_t p = 0; p < N; ++p) out->background[f * N + p] = bg_value(st.background, N, f, p);
    }
  });
}

}  // extern "C"
