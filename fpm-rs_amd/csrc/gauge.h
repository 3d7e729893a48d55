// gauge.h — EPRY object/pupil gauge on the device (gauge.hip).
#pragma once

#include "internal.h"

struct fpmhip_session;

namespace fpmhip {

// One tile's gauge inputs (canonicalize_object_pupil, src/algorithms/gauge.rs:20-296).
struct GaugeTile {
  double2* spec;           // [rh * rw] centred object spectrum (FP64)
  void* pupil;             // [nh * nw] recovered pupil, complex64 (pupil_f32) or complex128
  double2* pupil64;        // [nh * nw] its FP64 master (read, and written with the pupil), or null
  const double2* ref;      // [nh * nw] compiled reference pupil (FP64)
  const uint8_t* support;  // [nh * nw] compiled support
  double* maxp_out;        // max |P|^2 of the written pupil (the step denominator), or null
  int32_t* status;         // failure code (GaugeStatus) when still 0
  double tol;              // 64 eps |support| (gauge.rs:28)
  uint32_t nh, nw, rh, rw;
  int32_t pupil_f32;
  int32_t row_exact, col_exact;  // every source offset is zero on that axis (gauge.rs:46-67)
};

enum GaugeStatus : int32_t {
  kGaugeBadReference = 16,
  kGaugeBadEnergy,
  kGaugeBadScale,
  kGaugeNoAnchor,
  kGaugeBadPhase,
  kGaugeBadSpectrum,
  kGaugeBadObjectPhase,
};

// The gauge descriptor of a session (status: the session's flag[2], max
// |P|^2 into its scal[0]); defined with the session code (api.hip).
GaugeTile gauge_tile(const fpmhip_session& s);

// One workgroup per tile, on `stream`.
void launch_gauge(const GaugeTile* tiles_dev, uint32_t count, hipStream_t stream);
// Message and fpmhip_status of a gauge failure code (null: not a gauge code).
const char* gauge_failure(int32_t code, int* status);

}  // namespace fpmhip
