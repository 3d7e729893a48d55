// session.h — reconstruction session state and the launchers of the
// per-stage kernels (kernels.hip). Internal to libfpmhip.
#pragma once

#include "cx.h"
#include "internal.h"

namespace fpmhip {

// Per-source crop parameters, uploaded once per model.
struct SrcParam {
  uint64_t sr, sc;   // crop origin
  Axis ar, ac;       // interpolation axes (crop.rs:266-313)
  int32_t zero;      // FourierOffset::is_zero (crop.rs:28)
  int32_t pad;
};


// Internal configuration (UpdateConfiguration, common.rs:24-55).
struct Config {
  double object_step = 1.0;
  bool has_pupil_step = false;
  double pupil_step = 0.0;
  double epsilon = 1e-10;
  int loss = 0;
  int denom = 0;
  double alpha = 0.0;
  bool constrain = true;
  bool momentum = false;
  uint64_t interval = 0;
  double friction = 0.0, feedback = 0.0;
  bool gauge = false;  // canonicalize_state after init and each iteration (EPRY)
  bool gain_update = false, bg_update = false;  // EPRY frame-gain / background recovery
  double gain_step = 0.0, gain_min = 0.0, gain_max = 0.0;
  double bg_step = 0.0, bg_min = 0.0, bg_max = 0.0;
};
Config make_config(const fpmhip_algorithm& a);

// Slots of SessionBuffers::scal written by the calibration updates.
constexpr int kScalBackgroundStep = 4;
constexpr int kScalGainsPresent = 5;

// Rows per strip of the fused sweep's gather list.
constexpr uint32_t kGatherStrip = 5;
// Layout of SessionBuffers::dest for an n x n tile, in uint32 words: header
// (2), insert destinations ((n+1)^2 slots), support strips ((n / kGatherStrip
// + 1) n slots), the first destination index of each window row (n + 2),
// then the first strip index of each band of kGatherStrip rows (n /
// kGatherStrip + 2).
__host__ __device__ inline uint64_t dest_strips_offset(uint32_t n) { return 2 + (uint64_t)(n + 1) * (n + 1); }
__host__ __device__ inline uint64_t dest_row_starts_offset(uint32_t n) {
  return dest_strips_offset(n) + (uint64_t)(n / kGatherStrip + 1) * n;
}
__host__ __device__ inline uint64_t dest_band_starts_offset(uint32_t n) { return dest_row_starts_offset(n) + n + 2; }
__host__ __device__ inline uint64_t dest_table_words(uint32_t n) {
  return dest_band_starts_offset(n) + n / kGatherStrip + 2;
}

struct SessionBuffers {
  DeviceBuffer spec;      // double2[M] centred object spectrum (FP64 accumulation)
  DeviceBuffer vel, anc;  // V[M] mPIE velocity/anchor (momentum arithmetic is FP64 either way)
  DeviceBuffer pupil;     // V[N]
  DeviceBuffer pupil64;   // double2[N] FP64 master of the recovered pupil (MIXED_F32 EPRY pupil recovery):
                          // every pupil step accumulates into it and `pupil` is its complex64 copy
  DeviceBuffer patch;     // V[K*N]
  DeviceBuffer bounds;    // uint64[ceil(R_h / 8) * ceil(R_w / 8)]: upper bounds of |O|^2 per 8 x 8 block of the
                          // spectrum, as FP64 bit patterns (batched EPRY: prunes the whole-patch maximum)
  DeviceBuffer field;     // V[K*N]
  DeviceBuffer pred;      // double[N]
  DeviceBuffer proj;      // V[N]
  DeviceBuffer U;         // V[N]
  DeviceBuffer grad;      // V[N]
  DeviceBuffer scal;      // double[8]: [0] max|P|^2, [1] max|patch|^2, [2] initial fallback amplitude,
                          // [kScalBackgroundStep], [kScalGainsPresent] (EPRY calibration)
  DeviceBuffer partials;  // double[F * nblk * 2]
  DeviceBuffer flag;      // int[4]
  DeviceBuffer work_m;    // double2[M] (init / result scratch)
  DeviceBuffer avg, aw;   // double[N] (init)
  DeviceBuffer group;     // cooperative-group state of the fused sweep (large MIXED_F32 tiles)
  DeviceBuffer extent;    // uint32[n_h] per-row support of the complex64 pupil (MIXED_F32)
  DeviceBuffer dest;      // uint32[dest_table_words(n)] live-pixel lists of the fused sweep (square
                          // MIXED_F32 tiles): [0] destinations of the adjoint insert, [1] support
                          // strips; then the destinations a | b << 16 (row-major), the strips
                          // r | c << 10 | mask << 20 (kGatherStrip rows r.. of column c, bit j: pixel
                          // r + j is in the support; band of rows by band, columns in order within a
                          // band) and the index of each window row's first destination
  DeviceBuffer mspan;     // uint32 momentum chunks of the fused sweep (MIXED_F32 mPIE): [0] count, then
                          // start | (length - 1) << 25, up to 64 object pixels of one row each, covering
                          // every pixel an LED update can change
  DeviceBuffer bg;        // double[F * N] background of the reconstruction state (bg_len used)
  DeviceBuffer gains;     // double[F] frame gains of the state (1 where the model has none)
  DeviceBuffer calib;     // double[nblk * 2] block partial sums of the calibration updates
  DeviceBuffer gauge;     // GaugeTile of the EPRY gauge (sessions that canonicalize)
};

}  // namespace fpmhip

namespace fpmhip {
// A captured per-stage sweep: its frame order and the mPIE counter and
// background length before and after it.
struct SweepGraph {
  std::vector<uint64_t> order;
  uint64_t counter0 = 0, bg0 = 0, counter1 = 0, bg1 = 0;
  hipGraphExec_t exec = nullptr;
  SweepGraph() = default;
  SweepGraph(const SweepGraph&) = delete;
  SweepGraph& operator=(const SweepGraph&) = delete;
  SweepGraph(SweepGraph&& o) noexcept { *this = std::move(o); }
  SweepGraph& operator=(SweepGraph&& o) noexcept {
    if (this != &o) {
      if (exec) (void)hipGraphExecDestroy(exec);
      order = std::move(o.order);
      counter0 = o.counter0, bg0 = o.bg0, counter1 = o.counter1, bg1 = o.bg1;
      exec = o.exec;
      o.exec = nullptr;
    }
    return *this;
  }
  ~SweepGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
  }
};
constexpr size_t kSweepGraphs = 64;
}  // namespace fpmhip

struct fpmhip_session {
  fpmhip::Context* ctx = nullptr;
  const fpmhip::Model* model = nullptr;
  const fpmhip::Measurements* meas = nullptr;
  int precision = FPMHIP_PLAIN_C128;
  fpmhip_algorithm alg{};
  fpmhip::Config cfg;
  bool initialized = false;
  uint64_t counter = 0;
  uint32_t kmax = 1;   // largest multiplexing row
  uint32_t nblk = 0;   // blocks of the projection kernel
  bool has_gain = false;
  std::vector<double> gain;
  uint64_t bg_len = 0;  // background of the state: 0 (none), N (shared) or F * N (per frame)
  bool extent_from_model = false;  // b.extent / b.dest describe the model's pupil
  fpmhip::SessionBuffers b;
  std::vector<fpmhip::SweepGraph> graphs;  // captured per-stage sweeps (api.hip)
  size_t vbytes() const { return precision == FPMHIP_PLAIN_C128 ? sizeof(double2) : sizeof(float2); }
};

namespace fpmhip {

// ---- launchers (kernels.hip) ------------------------------------------------
// All take the precision of the session (double2 for PLAIN_C128, float2 for MIXED_F32).
void launch_extract(const fpmhip_session& s, uint64_t source, uint32_t mode);
// Projection of `frame` with its gain from the state's device gains.
void launch_project(const fpmhip_session& s, uint64_t frame, uint32_t K, const double* mode_w_dev,
                    double sw_sum, uint64_t slot);
// One calibration update of `frame` on the device, from the current fields
// over its unmasked pixels: mode 0 (update_gain) sums pred * max(y - bg, 0)
// and pred^2 and updates the state's gain of the frame (scal[kScalGainsPresent]
// = 1 once any gain was updated); mode 1 (update_background) sums the residual
// y - gain pred - bg and the pixel count into scal[kScalBackgroundStep]
// (launch only for frames with unmasked pixels).
void launch_calibration(const fpmhip_session& s, uint64_t frame, uint32_t K, const double* mode_w_dev, int mode);
// Background state of the session: bg[frame * N + p] = clamp(bg +
// scal[kScalBackgroundStep], lo, hi).
void launch_background_shift(const fpmhip_session& s, uint64_t frame, double lo, double hi);
// Expands a shared (length N) or absent background state to per-frame F * N.
void launch_background_expand(const fpmhip_session& s);
void launch_difference_frame(const fpmhip_session& s, uint32_t mode, uint64_t frame);
void launch_update(const fpmhip_session& s, uint32_t mode, double nsw, bool first_mode);
void launch_insert(const fpmhip_session& s, uint64_t source, uint32_t mode, double scale);
void launch_max_abs2(const fpmhip_session& s, const void* v, bool f64v, uint64_t n, uint32_t slot);
void launch_pupil_step(const fpmhip_session& s, double step);
void launch_momentum(const fpmhip_session& s);
void launch_flag_reset(const fpmhip_session& s);
void launch_init(const fpmhip_session& s);
// ReconstructionState::initialize_with_backend in two parts: every device
// launch (no host synchronisation), then the host-side checks and the EPRY
// gauge (api.hip). A batch launches all of its tiles before checking any.
void session_initialize_launch(fpmhip_session& s);
void session_initialize_check(fpmhip_session& s);
void launch_object(const fpmhip_session& s, double2* out_dev);
void launch_shift(const double2* src, double2* dst, uint32_t h, uint32_t w, uint32_t rs, uint32_t cs,
                  hipStream_t stream);
void launch_convert_pupil(const double2* src, void* dst, bool to_f32, uint64_t n, hipStream_t stream);
// The session's pupil from `src` (FP64, device): its FP64 master when it has
// one, and the pupil in the session's precision.
void set_session_pupil(const fpmhip_session& s, const double2* src);
// Per-row support extents of a MIXED_F32 session's pupil (for the fused sweep).
void launch_pupil_extent(const fpmhip_session& s);
void launch_pupil_to_f64(const void* src, bool from_f32, double2* dst, uint64_t n, hipStream_t stream);
void launch_fill_zero(double2* dst, uint64_t n, hipStream_t stream);
void launch_fill_zero(float2* dst, uint64_t n, hipStream_t stream);
void launch_forward_stack(const Model& m, const double2* spec, const double2* pupil, double2* fields,
                          double* out, hipStream_t stream);
// The forward stack with explicit frame gains [F] and background (length
// bg_len: 0, N or F N) in place of the model's.
void launch_forward_stack_with(const Model& m, const double2* spec, const double2* pupil, double2* fields,
                               double* out, const double* gains, const double* bg, uint64_t bg_len,
                               hipStream_t stream);
// SnrWeighted frame scores (schedule.rs:106-128) from the resident
// measurements and the model's background, reduced on the device.
std::vector<double> snr_scores(const fpmhip_session& s);

// Capacity (uint32 entries) of a session's momentum chunk table, and its
// upload for a given pupil (interleaved complex128; the table follows the
// pupil's complex64 support and the model's source windows).
uint64_t momentum_span_capacity(const Model& m);
void upload_momentum_spans(fpmhip_session& s, const double* pupil);

// ---- batched initialisation and result transform of MIXED_F32 tiles ----
// One entry per tile, in a device array: every pass below is one launch for
// all tiles (grid y or z = tile), computing what session_initialize_launch
// and launch_object compute for each tile, element for element.
struct InitTile {
  const void* frames;
  const double* weights;
  const uint8_t* masks;
  const double* model_gains;   // F (ones where the model has none)
  const double* model_bg;      // bg_len
  const double2* model_pupil;  // N, complex128
  double2* part;               // chunk partials: the session's init scratch
  double* avg;
  double* aw;
  double* scal;
  int* flag;
  uint64_t* counter;  // the batch's per-tile mPIE counter
  double2* slot;      // the tile's R-array in the batch scratch
  double2* spec;
  float2* pupil;
  double2* pupil64;            // FP64 pupil master, or null
  float2* vel;  // null without momentum
  float2* anc;
  double* gains;
  double* bg;
  uint64_t F, chunk, bg_len;
  uint32_t N, nh, nw, rh, rw, chunks, is_u16, pad;
};
// Chunking of the initial mean amplitude for a session (also used by
// launch_init, so that both sum the frames in the same order).
uint32_t init_chunks(const fpmhip_session& s);
// Reset, mean amplitude, fallback and nearest-neighbour object into each
// tile's slot (forward transform not included); pupil, gains and background
// copies and the pupil maximum. `nmax`, `mmax`: largest N and R of the tiles;
// `cmax`: largest of N, F and the background length.
void launch_init_batch_pre(const InitTile* tiles, uint32_t count, uint32_t nmax, uint32_t mmax, uint64_t cmax,
                           hipStream_t stream);
// Each slot (transformed) shifted into its tile's spectrum, with the mPIE
// anchor (= spectrum, complex64) and zero velocity.
void launch_init_batch_post(const InitTile* tiles, uint32_t count, uint32_t mmax, hipStream_t stream);
// flags[2 t], flags[2 t + 1] = tile t's flag[0], flag[1].
void launch_init_batch_flags(const InitTile* tiles, uint32_t count, int* flags, hipStream_t stream);
// Each tile's spectrum, inverse-shifted into its slot (the inverse transform
// of the slots completes the object, as launch_object does).
void launch_object_batch_shift(const InitTile* tiles, uint32_t count, uint32_t mmax, hipStream_t stream);

}  // namespace fpmhip
