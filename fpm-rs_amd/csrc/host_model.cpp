// host_model.cpp — host-side (no device) parts of the C ABI: error state,
// algorithm defaults/validation, and ImagePlaneModel::from_experiment for a
// planar LED array. Built into libfpmhip with the HIP sources.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>

#include "session.h"

namespace fpmhip {

namespace {
thread_local std::string t_error;
thread_local std::string t_error_name;
constexpr double kTau = 6.283185307179586;
}  // namespace

void fail(int code, const std::string& message, const char* name) {
  t_error = message;
  t_error_name = name ? name : "";
  throw Failure{code};
}

void clear_error() {
  t_error.clear();
  t_error_name.clear();
}

int capture_bad_alloc() {
  t_error = "host allocation failed";
  t_error_name.clear();
  return FPMHIP_OUT_OF_MEMORY;
}

// interpolation_axis (src/model/crop.rs:266-313)
Axis interpolation_axis(uint64_t start, uint64_t length, uint64_t bound, double offset) {
  if (length == 0 || bound == 0 || !std::isfinite(offset))
    fail(FPMHIP_INVALID_MODEL, "subpixel crop dimensions and offset must be finite and non-zero");
  const double nearest = std::round(offset);
  double whole, fraction;
  if (std::fabs(offset - nearest) <= 1e-12) {
    whole = nearest;
    fraction = 0.0;
  } else {
    whole = std::floor(offset);
    fraction = offset - whole;
  }
  const int64_t lower = (int64_t)start + (int64_t)whole;
  if (lower < 0) fail(FPMHIP_INVALID_MODEL, "subpixel crop starts outside the grid");
  const uint64_t end = (uint64_t)lower + length;
  const bool upper = fraction > 0.0;
  if (end > bound || (upper && end >= bound))
    fail(FPMHIP_INVALID_MODEL, "subpixel crop interpolation stencil lies outside the grid");
  return Axis{(uint32_t)lower, (uint32_t)lower + (upper ? 1u : 0u), 1.0 - fraction, fraction};
}

Config make_config(const fpmhip_algorithm& a) {
  Config c;
  c.object_step = a.object_step;
  c.epsilon = a.epsilon;
  c.loss = a.loss_type;
  switch (a.kind) {
    case FPMHIP_ALTERNATING_PROJECTION:  // alternating_projection.rs:108-132
      c.denom = FPMHIP_DENOMINATOR_LOCAL;
      break;
    case FPMHIP_FPIE:  // fpie.rs:111-131
      c.denom = FPMHIP_DENOMINATOR_RPIE;
      c.alpha = a.stability;
      break;
    case FPMHIP_MPIE:  // mpie.rs:298-327
      c.denom = FPMHIP_DENOMINATOR_RPIE;
      c.alpha = a.stability;
      c.momentum = true;
      c.interval = a.momentum_interval;
      c.friction = a.momentum_friction;
      c.feedback = a.momentum_feedback;
      break;
    case FPMHIP_EPRY:  // epry.rs:240-287
      c.denom = FPMHIP_DENOMINATOR_GLOBAL;
      c.has_pupil_step = a.recover_pupil != 0;
      c.pupil_step = a.pupil_step;
      c.constrain = a.constrain_pupil != 0;
      c.gauge = a.recover_pupil != 0;
      c.gain_update = a.recover_gain != 0;  // GainUpdateConfiguration (common.rs:390-422)
      c.gain_step = a.gain_step;
      c.gain_min = a.minimum_gain;
      c.gain_max = a.maximum_gain;
      c.bg_update = a.recover_background != 0;  // BackgroundUpdateConfiguration (common.rs:425-489)
      c.bg_step = a.background_step;
      c.bg_min = a.minimum_background;
      c.bg_max = a.maximum_background;
      break;
    default:
      fail(FPMHIP_INVALID_PARAMETER, "unknown algorithm kind", "kind");
  }
  return c;
}

}  // namespace fpmhip

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

struct Shift {
  int64_t integer;
  double offset;
  int64_t lower, upper;
};

// checked_grid_shift (src/model/image_plane_fpm.rs:975-1024)
Shift grid_shift(double value) {
  if (!std::isfinite(value)) fail(FPMHIP_INVALID_MODEL, "illumination shift must be finite");
  const double rounded = std::round(value);
  Shift s{(int64_t)rounded, value - rounded, 0, 0};
  const double nearest = std::round(s.offset);
  if (std::fabs(s.offset - nearest) <= 1e-12)
    s.lower = s.upper = s.integer + (int64_t)nearest;
  else if (s.offset > 0.0) {
    s.lower = s.integer;
    s.upper = s.integer + 1;
  } else {
    s.lower = s.integer - 1;
    s.upper = s.integer;
  }
  return s;
}

// smallest 2/3/5/7-smooth integer >= minimum (image_plane_fpm.rs:1176-1198)
uint64_t smooth_at_least(uint64_t minimum) {
  uint64_t best = std::numeric_limits<uint64_t>::max();
  for (uint64_t a = 1; a < best && a < (1ull << 40); a *= 2)
    for (uint64_t b = a; b < best; b *= 3)
      for (uint64_t c = b; c < best; c *= 5)
        for (uint64_t d = c; d < best; d *= 7)
          if (d >= minimum) {
            best = d;
            break;
          }
  return best;
}

}  // namespace

extern "C" {

const char* fpmhip_last_error(void) { return t_error.c_str(); }
const char* fpmhip_last_error_name(void) { return t_error_name.c_str(); }
const char* fpmhip_version(void) { return "fpmhip 0.1.0 (gfx950)"; }

int fpmhip_algorithm_default(int32_t kind, fpmhip_algorithm* out) {
  return guarded([&] {
    require(out != nullptr, "algorithm", "output is null");
    fpmhip_algorithm a{};
    a.kind = kind;
    a.loss_type = FPMHIP_LOSS_AMPLITUDE_MSE;
    a.batch_size = 1;
    a.epsilon = 1e-10;
    a.constrain_pupil = 1;
    switch (kind) {
      case FPMHIP_ALTERNATING_PROJECTION:
        a.iterations = 50;
        a.object_step = 1.0;
        break;
      case FPMHIP_FPIE:
        a.iterations = 50;
        a.object_step = 0.8;
        a.stability = 0.1;
        break;
      case FPMHIP_MPIE:
        a.iterations = 50;
        a.object_step = 0.2;
        a.stability = 0.05;
        a.momentum_interval = 30;
        a.momentum_friction = 0.9;
        a.momentum_feedback = 0.9;
        break;
      case FPMHIP_EPRY:
        a.iterations = 100;
        a.object_step = 0.8;
        a.pupil_step = 0.1;
        a.recover_pupil = 1;
        a.constrain_pupil = 1;
        a.gain_step = 0.2;
        a.minimum_gain = 1e-6;
        a.maximum_gain = 1e6;
        a.background_step = 0.2;
        a.minimum_background = 0.0;
        a.maximum_background = 1e12;
        break;
      default:
        fail(FPMHIP_INVALID_PARAMETER, "unknown algorithm kind", "kind");
    }
    *out = a;
  });
}

int fpmhip_algorithm_validate(const fpmhip_algorithm* a) {
  return guarded([&] {
    require(a != nullptr, "algorithm", "algorithm is null");
    const double fin = 0.0;
    (void)fin;
    switch (a->kind) {
      case FPMHIP_ALTERNATING_PROJECTION:  // alternating_projection.rs:108-132
        require(std::isfinite(a->object_step) && a->object_step > 0.0, "object_step", "must be finite and positive");
        require(std::isfinite(a->epsilon) && a->epsilon > 0.0, "epsilon", "must be finite and positive");
        require(a->batch_size != 0, "batch_size", "must be greater than zero");
        break;
      case FPMHIP_FPIE:  // fpie.rs:111-131
        require(std::isfinite(a->object_step) && a->object_step > 0.0, "object_step", "must be finite and positive");
        require(std::isfinite(a->stability) && a->stability >= 0.0 && a->stability <= 1.0, "stability",
                "must be finite and between zero and one");
        require(std::isfinite(a->epsilon) && a->epsilon > 0.0 && a->batch_size != 0, "epsilon/batch_size",
                "epsilon must be positive and batch size non-zero");
        break;
      case FPMHIP_MPIE:  // mpie.rs:229-272
        require(std::isfinite(a->object_step) && a->object_step > 0.0, "object_step", "must be finite and positive");
        require(std::isfinite(a->stability) && a->stability >= 0.0 && a->stability <= 1.0, "stability",
                "must be finite and between zero and one");
        require(a->momentum_interval != 0, "momentum_interval", "must be greater than zero");
        require(std::isfinite(a->momentum_friction) && a->momentum_friction >= 0.0 && a->momentum_friction < 1.0,
                "momentum_friction", "must be finite, at least zero, and less than one");
        require(std::isfinite(a->momentum_feedback) && a->momentum_feedback >= 0.0 && a->momentum_feedback <= 1.0,
                "momentum_feedback", "must be finite and between zero and one");
        require(a->batch_size != 0, "batch_size", "must be greater than zero");
        require(std::isfinite(a->epsilon) && a->epsilon > 0.0, "epsilon", "must be finite and positive");
        break;
      case FPMHIP_EPRY:  // epry.rs:187-238
        require(std::isfinite(a->object_step) && a->object_step > 0.0, "object_step", "must be finite and positive");
        require(std::isfinite(a->pupil_step) && a->pupil_step >= 0.0, "pupil_step", "must be finite and non-negative");
        require(std::isfinite(a->epsilon) && a->epsilon > 0.0 && a->batch_size != 0, "epsilon/batch_size",
                "epsilon must be positive and batch size non-zero");
        require(std::isfinite(a->gain_step) && a->gain_step >= 0.0 && a->gain_step <= 1.0, "gain_step",
                "must be finite and between zero and one");
        require(std::isfinite(a->minimum_gain) && std::isfinite(a->maximum_gain) && a->minimum_gain > 0.0 &&
                    a->maximum_gain > a->minimum_gain,
                "gain_bounds", "must be finite, positive, and strictly increasing");
        require(std::isfinite(a->background_step) && a->background_step >= 0.0 && a->background_step <= 1.0,
                "background_step", "must be finite and between zero and one");
        require(std::isfinite(a->minimum_background) && std::isfinite(a->maximum_background) &&
                    a->maximum_background > a->minimum_background,
                "background_bounds", "must be finite and strictly increasing");
        break;
      default:
        fail(FPMHIP_INVALID_PARAMETER, "unknown algorithm kind", "kind");
    }
    require(a->loss_type >= 0 && a->loss_type <= 3, "loss_type", "unknown loss type");
  });
}

// ImagePlaneModel::from_experiment (image_plane_fpm.rs:167-253) for a planar
// LED array: positions (led_array.rs:292-313, pose :87-113) -> directions and
// k-vectors (illumination.rs:385-421) -> grid bounds (:255-285) -> shape
// policy (:287-348, 1116-1198) -> circular pupil (pupil.rs:78-124) -> crops
// and fractional offsets (:750-797).
int fpmhip_compile_planar(const fpmhip_optics* o, const fpmhip_planar_array* a, uint64_t nh, uint64_t nw,
                          int32_t policy, uint64_t exact_h, uint64_t exact_w, fpmhip_compiled_shape* shape,
                          uint64_t* crop_start, double* offset, double* k_vectors, double* pupil,
                          uint8_t* support, double* frame_gain) {
  return guarded([&] {
    require(o && a && shape, "compile", "null argument");
    // Optics::validate (optics.rs:128-193)
    const std::pair<const char*, double> positive[] = {
        {"wavelength_vacuum_m", o->wavelength_vacuum_m},
        {"objective_na", o->objective_na},
        {"magnification", o->magnification},
        {"camera_pixel_size", o->camera_pixel_size},
        {"illumination_refractive_index", o->illumination_refractive_index},
        {"objective_medium_refractive_index", o->objective_medium_refractive_index}};
    for (auto& kv : positive) require(std::isfinite(kv.second) && kv.second > 0.0, kv.first, "must be finite and positive");
    require(o->objective_na <= o->objective_medium_refractive_index, "objective_na",
            "cannot exceed the medium refractive index");
    const double px = o->camera_pixel_size / o->magnification;
    require(px < o->wavelength_vacuum_m / (2.0 * o->objective_na), "camera_pixel_size",
            "object-plane pitch must be below the coherent sampling limit");
    if (nh == 0 || nw == 0) fail(FPMHIP_INVALID_SHAPE, "image shape must be non-zero");
    require(a->rows > 0 && a->columns > 0, "shape", "rows and columns must be non-zero");
    require(std::isfinite(a->pitch_x_m) && std::isfinite(a->pitch_y_m) && a->pitch_x_m > 0.0 && a->pitch_y_m > 0.0,
            "pitch_m", "pitch_x and pitch_y must be finite and positive");

    const uint64_t S = a->rows * a->columns;
    std::vector<double> k(2 * S);
    const double k0 = kTau * o->illumination_refractive_index / o->wavelength_vacuum_m;
    const double sx = std::sin(a->rotation_rad[0]), cxr = std::cos(a->rotation_rad[0]);
    const double sy = std::sin(a->rotation_rad[1]), cyr = std::cos(a->rotation_rad[1]);
    const double sz = std::sin(a->rotation_rad[2]), czr = std::cos(a->rotation_rad[2]);
    for (uint64_t r = 0; r < a->rows; ++r)
      for (uint64_t c = 0; c < a->columns; ++c) {
        const double lx = ((double)c - a->reference_column) * a->pitch_x_m;
        const double ly = ((double)r - a->reference_row) * a->pitch_y_m;
        const double lz = 0.0;
        const double x1 = lx, y1 = cxr * ly - sx * lz, z1 = sx * ly + cxr * lz;
        const double x2 = cyr * x1 + sy * z1, y2 = y1, z2 = -sy * x1 + cyr * z1;
        const double x3 = czr * x2 - sz * y2, y3 = sz * x2 + czr * y2, z3 = z2;
        const double px_ = x3 + a->translation_m[0], py_ = y3 + a->translation_m[1], pz_ = z3 + a->translation_m[2];
        const double norm = std::hypot(std::hypot(px_, py_), pz_);
        require(std::isfinite(norm) && norm > 0.0, "positions_m", "no source may coincide with the sample origin");
        const uint64_t s = r * a->columns + c;
        k[2 * s] = k0 * (-px_ / norm);
        k[2 * s + 1] = k0 * (-py_ / norm);
      }
    const double dkx = kTau / ((double)nw * px), dky = kTau / ((double)nh * px);
    int64_t rmin = INT64_MAX, rmax = INT64_MIN, cmin = INT64_MAX, cmax = INT64_MIN;
    for (uint64_t s = 0; s < S; ++s) {
      const Shift gr = grid_shift(k[2 * s + 1] / dky), gc = grid_shift(k[2 * s] / dkx);
      rmin = std::min(rmin, gr.lower);
      rmax = std::max(rmax, gr.upper);
      cmin = std::min(cmin, gc.lower);
      cmax = std::max(cmax, gc.upper);
    }
    auto axis_ok = [](uint64_t n, uint64_t R, int64_t lo, int64_t hi) {
      const int64_t base = (int64_t)(R / 2) - (int64_t)(n / 2);
      return base + lo >= 0 && base + (int64_t)(n - 1) + hi < (int64_t)R;
    };
    auto fits = [&](uint64_t h, uint64_t w) { return axis_ok(nh, h, rmin, rmax) && axis_ok(nw, w, cmin, cmax); };
    uint64_t rh, rw;
    if (policy == FPMHIP_SHAPE_EXACT) {
      if (exact_h < nh || exact_w < nw) fail(FPMHIP_INVALID_SHAPE, "image shape must fit reconstruction shape");
      if (exact_h * nw != exact_w * nh)
        fail(FPMHIP_INVALID_SHAPE, "reconstruction must use the same scale factor in both dimensions");
      if (!fits(exact_h, exact_w))
        fail(FPMHIP_INVALID_SHAPE,
             "reconstruction shape does not contain every illumination crop and subpixel interpolation stencil");
      rh = exact_h;
      rw = exact_w;
    } else {
      uint64_t g = nh, b = nw;
      while (b) {
        uint64_t t = g % b;
        g = b;
        b = t;
      }
      const uint64_t ah = nh / g, aw = nw / g;
      // the minimum multiplier (monotone in the multiplier: binary search)
      uint64_t lo = g, hi = g;
      if (!fits(ah * g, aw * g)) {
        while (!fits(ah * hi, aw * hi)) {
          lo = hi;
          hi *= 2;
          if (hi > (1ull << 40)) fail(FPMHIP_INVALID_SHAPE, "illumination crops require an unsupported grid");
        }
        while (lo + 1 < hi) {
          const uint64_t mid = lo + (hi - lo) / 2;
          if (fits(ah * mid, aw * mid))
            hi = mid;
          else
            lo = mid;
        }
      }
      uint64_t mult = hi;
      if (policy == FPMHIP_SHAPE_SMOOTH)
        mult = smooth_at_least(mult);
      else if (policy == FPMHIP_SHAPE_POWER_OF_TWO) {
        uint64_t p = 1;
        while (p < mult) p <<= 1;
        mult = p;
      } else if (policy != FPMHIP_SHAPE_MINIMUM)
        fail(FPMHIP_INVALID_PARAMETER, "unknown reconstruction shape policy", "reconstruction_shape");
      rh = ah * mult;
      rw = aw * mult;
    }
    shape->source_count = S;
    shape->recon_h = rh;
    shape->recon_w = rw;
    if (!crop_start) return;
    const double cutoff = kTau * o->objective_na / o->wavelength_vacuum_m;
    const double medium_k = kTau * o->objective_medium_refractive_index / o->wavelength_vacuum_m;
    for (uint64_t r = 0; r < nh; ++r) {
      const double ky = ((double)r - (double)(nh / 2)) * dky;
      for (uint64_t c = 0; c < nw; ++c) {
        const double kx = ((double)c - (double)(nw / 2)) * dkx;
        const double radius = std::hypot(kx, ky);
        const uint64_t i = r * nw + c;
        const bool inside = radius <= cutoff;
        if (support) support[i] = inside ? 1 : 0;
        double re = 0.0, im = 0.0;
        if (inside) {
          const double rho = cutoff > 0.0 ? radius / cutoff : 0.0;
          const double theta = std::atan2(ky, kx);
          double phase = 0.0, amplitude = 1.0;
          if (o->has_defocus) phase -= o->defocus_distance * (kx * kx + ky * ky) / (2.0 * medium_k);
          if (o->has_aberration) {
            phase += o->astigmatism * rho * rho * std::cos(2.0 * theta);
            phase += o->coma * (3.0 * std::pow(rho, 3) - 2.0 * rho) * std::cos(theta);
            phase += o->spherical * (6.0 * std::pow(rho, 4) - 6.0 * rho * rho + 1.0);
            amplitude = std::exp(-o->edge_apodization * rho * rho);
          }
          re = amplitude * std::cos(phase);
          im = amplitude * std::sin(phase);
        }
        if (pupil) {
          pupil[2 * i] = re;
          pupil[2 * i + 1] = im;
        }
      }
    }
    for (uint64_t s = 0; s < S; ++s) {
      const Shift gr = grid_shift(k[2 * s + 1] / dky), gc = grid_shift(k[2 * s] / dkx);
      const int64_t r0 = (int64_t)(rh / 2) - (int64_t)(nh / 2) + gr.integer;
      const int64_t c0 = (int64_t)(rw / 2) - (int64_t)(nw / 2) + gc.integer;
      if (r0 < 0 || c0 < 0) fail(FPMHIP_INVALID_MODEL, "illumination vector produces a crop outside the grid");
      interpolation_axis((uint64_t)r0, nh, rh, gr.offset);
      interpolation_axis((uint64_t)c0, nw, rw, gc.offset);
      crop_start[2 * s] = (uint64_t)r0;
      crop_start[2 * s + 1] = (uint64_t)c0;
      if (offset) {
        offset[2 * s] = gr.offset;
        offset[2 * s + 1] = gc.offset;
      }
    }
    if (k_vectors) std::copy(k.begin(), k.end(), k_vectors);
    if (frame_gain) std::fill(frame_gain, frame_gain + S, 1.0);
  });
}

}  // extern "C"
