/*
 * fpmhip.h — C ABI of the MI355X (gfx950) implementation of fpm-rs's
 * image-plane forward model and per-LED reconstruction updates
 * (AlternatingProjection, Fpie, Mpie, Epry).
 *
 * Conventions shared with fpm-rs (reference citations are path:line in
 * hgrecco/fpm-rs):
 *   - arrays are C-contiguous row-major (README.md:59-75, src/array_layout.rs);
 *   - complex arrays are interleaved (re, im) pairs of double, the memory
 *     layout of num_complex::Complex64;
 *   - object spectra are centred (fftshift by floor(n/2)); the forward 2-D FFT
 *     is scaled by 1/N and the inverse is unscaled (src/backend/cpu.rs:162-172);
 *   - every function returns an fpmhip_status; on failure a thread-local
 *     message is available from fpmhip_last_error() and, for
 *     FPMHIP_INVALID_PARAMETER, the stable parameter name from
 *     fpmhip_last_error_name() (src/error.rs:14-174).
 *
 * Ownership: handles are created and destroyed explicitly. A model or a
 * measurement set may be shared by several sessions; a session must not be
 * used from two threads at once (a session owns its scratch, as a
 * ReconstructionState does, src/reconstruction/state.rs:159-172).
 * Host pointers passed in are read during the call only.
 */
#ifndef FPMHIP_H
#define FPMHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes: one per fpm_rs::Error variant on this path (src/error.rs:14-174) ---- */
typedef enum fpmhip_status {
  FPMHIP_OK = 0,
  FPMHIP_INVALID_SHAPE = 1,         /* Error::InvalidShape */
  FPMHIP_INVALID_PARAMETER = 2,     /* Error::InvalidParameter{name, reason} */
  FPMHIP_INVALID_MODEL = 3,         /* Error::InvalidModel */
  FPMHIP_INVALID_MEASUREMENTS = 4,  /* Error::InvalidMeasurements */
  FPMHIP_LENGTH_MISMATCH = 5,       /* Error::LengthMismatch */
  FPMHIP_SHAPE_OVERFLOW = 6,        /* Error::ShapeOverflow */
  FPMHIP_NON_STANDARD_LAYOUT = 7,   /* Error::NonStandardLayout */
  FPMHIP_FRAME_OUT_OF_RANGE = 8,    /* Error::FrameOutOfRange */
  FPMHIP_NUMERICAL = 9,             /* Error::Numerical */
  FPMHIP_UNSUPPORTED = 10,          /* Error::Unsupported */
  FPMHIP_DEVICE_ERROR = 11,         /* HIP runtime failure (no reference counterpart) */
  FPMHIP_OUT_OF_MEMORY = 12         /* device allocation failure (no reference counterpart) */
} fpmhip_status;

/* Message of the last failure on the calling thread ("" after success). */
const char* fpmhip_last_error(void);
/* Stable parameter name of the last FPMHIP_INVALID_PARAMETER ("" otherwise),
 * e.g. "object_step", "momentum_friction" (tests/reconstruction_tests.rs:907). */
const char* fpmhip_last_error_name(void);
/* Library version string. */
const char* fpmhip_version(void);

/* ---- enums ---- */
typedef enum fpmhip_direction { FPMHIP_FORWARD = 0, FPMHIP_INVERSE = 1 } fpmhip_direction;

/* Algorithms whose step is projection_update (src/algorithms/common.rs:57-306). */
typedef enum fpmhip_algorithm_kind {
  FPMHIP_ALTERNATING_PROJECTION = 0, /* src/algorithms/alternating_projection.rs:37 */
  FPMHIP_FPIE = 1,                   /* src/algorithms/fpie.rs:36 */
  FPMHIP_MPIE = 2,                   /* src/algorithms/mpie.rs:58 */
  FPMHIP_EPRY = 3                    /* src/algorithms/epry.rs:55 */
} fpmhip_algorithm_kind;

/* ObjectDenominator (src/algorithms/common.rs:18-22). */
typedef enum fpmhip_denominator {
  FPMHIP_DENOMINATOR_LOCAL = 0,
  FPMHIP_DENOMINATOR_RPIE = 1,
  FPMHIP_DENOMINATOR_GLOBAL = 2
} fpmhip_denominator;

/* LossType (src/algorithms/objective.rs:9-19). */
typedef enum fpmhip_loss {
  FPMHIP_LOSS_AMPLITUDE_MSE = 0,
  FPMHIP_LOSS_INTENSITY_MSE = 1,
  FPMHIP_LOSS_POISSON_NLL = 2,
  FPMHIP_LOSS_HUBER_AMPLITUDE = 3
} fpmhip_loss;

/* ReconstructionShape (src/model/image_plane_fpm.rs:287-348). */
typedef enum fpmhip_shape_policy {
  FPMHIP_SHAPE_EXACT = 0,
  FPMHIP_SHAPE_MINIMUM = 1,
  FPMHIP_SHAPE_SMOOTH = 2,
  FPMHIP_SHAPE_POWER_OF_TWO = 3
} fpmhip_shape_policy;

/* Arithmetic of a session.
 *   PLAIN_C128: every buffer and operation in complex128/f64, the reference's
 *               arithmetic, one LED step at a time.
 *   MIXED_F32:  FFTs and pointwise updates in complex64; object-spectrum
 *               accumulation, step-size denominators (max|P|^2, max|patch|^2)
 *               and objective reductions in f64. */
typedef enum fpmhip_precision {
  FPMHIP_PLAIN_C128 = 0,
  FPMHIP_MIXED_F32 = 1
} fpmhip_precision;

/* FrameSchedule (src/reconstruction/schedule.rs:10-27). RandomShuffle's
 * permutation for (seed, iteration) is this library's own (a Fisher-Yates
 * shuffle driven by Philox4x64-10 under key (seed + iteration *
 * 0x9e3779b97f4a7c15, 1), counter (i, 0, 0, 0)): seeded per iteration as the
 * reference's is, but not rand's StdRng permutation. */
typedef enum fpmhip_schedule {
  FPMHIP_SCHEDULE_SEQUENTIAL = 0,
  FPMHIP_SCHEDULE_BRIGHTFIELD_FIRST = 1,
  FPMHIP_SCHEDULE_SPIRAL_OUT = 2,
  FPMHIP_SCHEDULE_SNR_WEIGHTED = 3,
  FPMHIP_SCHEDULE_RANDOM_SHUFFLE = 4
} fpmhip_schedule;

/* A FrameSchedule value: kind plus RandomShuffle's seed. */
typedef struct fpmhip_frame_schedule {
  int32_t kind;
  int32_t pad_;
  uint64_t seed;
} fpmhip_frame_schedule;

/* ---- physical inputs of ImagePlaneModel::from_experiment ---- */

/* Optics (src/experiment/optics.rs:44-60); has_defocus/has_aberration mark
 * the Option fields. */
typedef struct fpmhip_optics {
  double wavelength_vacuum_m;
  double objective_na;
  double magnification;
  double camera_pixel_size;
  double illumination_refractive_index;
  double objective_medium_refractive_index;
  int32_t has_defocus;
  int32_t has_aberration;
  double defocus_distance;
  double astigmatism, coma, spherical, edge_apodization; /* PupilAberration */
} fpmhip_optics;

/* PlanarLedArray (src/experiment/led_array.rs:120-313) with an ArrayPose:
 * shape (rows, columns), pitch (x, y), reference index (column, row),
 * translation (x, y, z) and active extrinsic XYZ rotation in radians. */
typedef struct fpmhip_planar_array {
  uint64_t rows, columns;
  double pitch_x_m, pitch_y_m;
  double reference_column, reference_row;
  double translation_m[3];
  double rotation_rad[3];
} fpmhip_planar_array;

/* ---- compiled model (ImagePlaneModel fields, src/model/image_plane_fpm.rs:58-73) ---- */
typedef struct fpmhip_model_desc {
  uint64_t image_h, image_w;        /* low-resolution grid n_h x n_w */
  uint64_t recon_h, recon_w;        /* high-resolution grid R_h x R_w */
  uint64_t source_count;            /* S individual sources */
  uint64_t frame_count;             /* F measured frames (== S without multiplexing) */
  const uint64_t* crop_start;       /* [S][2] (row, column) crop origins */
  const double* offset;             /* [S][2] (row, column) fractional offsets, or NULL */
  const double* pupil;              /* [N] complex128, interleaved */
  const uint8_t* support;           /* [N] 0/1 */
  const double* frame_gain;         /* [F] or NULL */
  const double* background;         /* [N] or [F*N] or NULL */
  uint64_t background_len;          /* 0, N or F*N */
  const uint64_t* mux_row_ptr;      /* [F+1] CSR row pointers, or NULL (identity) */
  const uint64_t* mux_source;       /* [nnz] source per entry */
  const double* mux_weight;         /* [nnz] incoherent weight per entry */
  const double* k_vectors;          /* [S][2] (kx, ky) rad/m, or NULL; used by schedules */
} fpmhip_model_desc;

/* Sizes of a compile: fill image shape and policy, get S and R back. */
typedef struct fpmhip_compiled_shape {
  uint64_t source_count;
  uint64_t recon_h, recon_w;
} fpmhip_compiled_shape;

/* ImagePlaneModel::from_experiment (src/model/image_plane_fpm.rs:167-253) for a
 * planar LED array, in two calls: first with all outputs NULL to get the
 * shape, then with caller-allocated outputs. `exact_h/exact_w` are read for
 * FPMHIP_SHAPE_EXACT only. Outputs: crop_start [S][2], offset [S][2],
 * k_vectors [S][2], pupil [N] complex, support [N], frame_gain [S] (unit
 * gains). Host-only computation; no device needed. */
int fpmhip_compile_planar(const fpmhip_optics* optics, const fpmhip_planar_array* array,
                          uint64_t image_h, uint64_t image_w, int32_t policy,
                          uint64_t exact_h, uint64_t exact_w, fpmhip_compiled_shape* shape,
                          uint64_t* crop_start, double* offset, double* k_vectors,
                          double* pupil, uint8_t* support, double* frame_gain);

/* ---- algorithm configuration (UpdateConfiguration, src/algorithms/common.rs:24-55) ---- */
typedef struct fpmhip_algorithm {
  int32_t kind;               /* fpmhip_algorithm_kind */
  int32_t loss_type;          /* fpmhip_loss */
  int32_t recover_pupil;      /* EPRY recover_pupil */
  int32_t constrain_pupil;    /* EPRY constrain_pupil_support */
  int32_t recover_gain;       /* EPRY recover_frame_gains */
  int32_t recover_background; /* EPRY recover_background */
  uint64_t iterations;
  uint64_t batch_size;
  double object_step;
  double pupil_step;          /* EPRY */
  double stability;           /* FPIE/mPIE rPIE alpha */
  double epsilon;
  uint64_t momentum_interval; /* mPIE */
  double momentum_friction;
  double momentum_feedback;
  double gain_step, minimum_gain, maximum_gain;
  double background_step, minimum_background, maximum_background;
} fpmhip_algorithm;

/* Fills the reference defaults of `kind` (alternating_projection.rs:51-60,
 * fpie.rs:53-62, mpie.rs:81-95, epry.rs:91-111). */
int fpmhip_algorithm_default(int32_t kind, fpmhip_algorithm* out);
/* Validates as ReconstructionAlgorithm::validate does, with the same stable
 * parameter names (mpie.rs:229-272, epry.rs:187-238, fpie.rs:111-131,
 * alternating_projection.rs:108-132). */
int fpmhip_algorithm_validate(const fpmhip_algorithm* algorithm);

/* ---- opaque handles ---- */
typedef struct fpmhip_context fpmhip_context;
typedef struct fpmhip_model fpmhip_model;
typedef struct fpmhip_measurements fpmhip_measurements;
typedef struct fpmhip_session fpmhip_session;
typedef struct fpmhip_batch fpmhip_batch;
typedef struct fpmhip_server fpmhip_server;

/* One HIP stream per context on `device`. */
int fpmhip_context_create(int32_t device, fpmhip_context** out);
int fpmhip_context_destroy(fpmhip_context* context);
int fpmhip_context_synchronize(fpmhip_context* context);

/* ---- FFT primitive: Backend::fft2 (src/backend/traits.rs:123, cpu.rs:130-174) ----
 * In-place 2-D complex128 FFT of a host array, forward scaled by 1/(h*w),
 * inverse unscaled. Any h, w >= 1 (mixed radix). */
int fpmhip_fft2_host(fpmhip_context* context, double* values, uint64_t h, uint64_t w,
                     int32_t direction);

/* ---- ResidentBackend: typed device buffers (src/backend/traits.rs:17-106) ---- */
/* MemoryLocation (traits.rs:17-24). */
typedef enum fpmhip_memory_location { FPMHIP_MEMORY_HOST = 0, FPMHIP_MEMORY_DEVICE = 1 } fpmhip_memory_location;
/* BackendCapabilities (traits.rs:27-33): device memory, resident buffers. */
typedef struct fpmhip_capabilities {
  int32_t preferred_memory; /* fpmhip_memory_location */
  int32_t resident_buffers;
} fpmhip_capabilities;
int fpmhip_context_capabilities(const fpmhip_context* context, fpmhip_capabilities* out);

/* ComplexBuffer / RealBuffer (traits.rs:44-73): `len` complex128 or f64
 * values in device memory, zeroed at allocation (CpuBackend zeroes,
 * cpu.rs:178-189). */
typedef enum fpmhip_buffer_kind { FPMHIP_BUFFER_COMPLEX = 0, FPMHIP_BUFFER_REAL = 1 } fpmhip_buffer_kind;
typedef struct fpmhip_buffer fpmhip_buffer;
/* allocate_complex / allocate_real (traits.rs:79-81). */
int fpmhip_buffer_allocate(fpmhip_context* context, int32_t kind, uint64_t len, fpmhip_buffer** out);
int fpmhip_buffer_free(fpmhip_buffer* buffer);
/* len() and location() (traits.rs:46-53), and the buffer's device address
 * (for the allocation-free entry points; any output may be NULL). */
int fpmhip_buffer_info(const fpmhip_buffer* buffer, int32_t* kind, uint64_t* len, int32_t* location,
                       void** device_address);
/* upload_complex / upload_real and download_complex / download_real
 * (traits.rs:82-95): `kind` names the typed operation; a buffer of the other
 * kind or from another context is FPMHIP_INVALID_PARAMETER ("complex buffer"
 * / "real buffer"), and `len` (complex values or reals) different from the
 * buffer's is FPMHIP_INVALID_SHAPE (cpu.rs:238-285). */
int fpmhip_buffer_upload(fpmhip_context* context, fpmhip_buffer* buffer, int32_t kind, const double* source,
                         uint64_t len);
int fpmhip_buffer_download(fpmhip_context* context, const fpmhip_buffer* buffer, int32_t kind,
                           double* destination, uint64_t len);
/* fft2_resident (traits.rs:97-103): in-place 2-D FFT of a complex buffer of
 * h * w values, forward / (h w), inverse unscaled. */
int fpmhip_fft2_resident(fpmhip_context* context, fpmhip_buffer* buffer, uint64_t h, uint64_t w,
                         int32_t direction);

/* ---- model and measurement upload ---- */
/* Validates like ImagePlaneModel::validate (image_plane_fpm.rs:841-972) and
 * uploads pupil, support, crops, offsets, gains, background, multiplexing. */
int fpmhip_model_create(fpmhip_context* context, const fpmhip_model_desc* desc,
                        fpmhip_model** out);
int fpmhip_model_destroy(fpmhip_model* model);

/* MeasurementStack (src/measurements/stack.rs:47-105): F frames of N pixels.
 * Exactly one of `counts_u16` (exact detector counts) and `frames_f64` is
 * non-NULL. `weights` [F] (>= 0; NULL = all 1), `masks` [F*N] 0/1 (NULL = none). */
int fpmhip_measurements_create(fpmhip_context* context, const fpmhip_model* model,
                               const uint16_t* counts_u16, const double* frames_f64,
                               const double* weights, const uint8_t* masks,
                               fpmhip_measurements** out);
int fpmhip_measurements_destroy(fpmhip_measurements* measurements);

/* ---- forward model: ForwardModel::forward_intensity_stack_into (src/model/forward.rs:300-372) ----
 * spectrum [R] complex128 centred, pupil [N] complex128 (NULL = model pupil),
 * out [F*N] f64 = gain * sum_s w_s |IFFT(ifftshift(patch_s * P))|^2 + background. */
int fpmhip_forward_stack(fpmhip_context* context, const fpmhip_model* model,
                         const double* spectrum, const double* pupil, double* out);

/* Allocation-free variant of fpmhip_forward_stack. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation). `spectrum` [R] and `pupil` [N] (NULL = model pupil) are
 * complex128 and `out` [F*N] f64 in the caller's DEVICE memory; the call
 * enqueues the work on the context's stream and returns (results are
 * complete once the stream is synchronised, e.g. fpmhip_context_synchronize)
 * and makes no device allocation. The per-source fields live in a fixed
 * 16 MiB device array of the library, one per GPU, held until the process
 * ends and shared by every call on that GPU: calls that overlap on one GPU
 * race (calls on one context's stream are ordered). Models with multiplexed
 * frames are Unsupported; an image above 1 Mi pixels (the array's 2^20
 * complex128 values) is InvalidShape. The output equals fpmhip_forward_stack's. */
int fpmhip_forward_stack_noalloc(fpmhip_context* context, const fpmhip_model* model,
                                 const double* spectrum, const double* pupil, double* out);

/* ---- simulation: CameraModel (src/simulation/camera.rs:7-34) and Simulator::simulate
 * (src/simulation/simulator.rs:107-207) ----
 * Each Option of CameraModel is a has_ flag plus its value. pixel_sensitivity
 * [frame_len] and bad_pixels [bad_pixel_count] are host arrays (NULL / 0 = none).
 * Noise: shot noise is Poisson(expected electrons) and read noise
 * Normal(0, read_noise_electrons), drawn on the device from a counter-based
 * Philox4x64-10 stream: pixel p of frame f takes the blocks of counter
 * (p, f, block, 0) under key (seed, 0), so a seed gives the same counts on any
 * grid shape. The reference's StdRng stream is not reproduced (same
 * distributions, other draws). */
typedef struct fpmhip_camera {
  double photons_per_pixel;        /* > 0 */
  double gain_counts_per_electron; /* > 0 */
  double offset_counts;
  double read_noise_electrons;     /* >= 0 */
  double dark_current_electrons;   /* >= 0 */
  int32_t shot_noise;
  int32_t quantize;
  int32_t has_bit_depth;
  int32_t bit_depth;               /* 1..=32 when present */
  int32_t has_saturation;
  int32_t has_bad_pixel_value;
  double saturation_counts;        /* finite, > 0 when present */
  double bad_pixel_value_counts;   /* finite, >= 0 when present */
  const double* pixel_sensitivity;
  const uint64_t* bad_pixels;
  uint64_t bad_pixel_count;
} fpmhip_camera;

/* CameraModel::validate (camera.rs:145-223), and validate_for_frame
 * (camera.rs:226-245) when frame_len > 0. InvalidParameter named as the
 * reference names it ("photons_per_pixel", "camera noise/offset", "bit_depth", ...). */
int fpmhip_camera_validate(const fpmhip_camera* camera, uint64_t frame_len);

/* CameraModel::measure_frame (camera.rs:248-287) on frame_count frames of
 * frame_len pixels, in place (host memory); frame f draws counter (p, f, ...). */
int fpmhip_camera_measure(fpmhip_context* context, const fpmhip_camera* camera, double* frames,
                          uint64_t frame_count, uint64_t frame_len, uint64_t seed);

/* The device part of Simulator::simulate (simulator.rs:138-189) for a true
 * model whose acquisition errors (source permutation, frame-gain variation)
 * the caller has already applied to it: object [R] complex128 (row-major
 * field) -> centred spectrum (forward FFT / R, fftshift; simulator.rs:209-226)
 * -> forward intensity stack -> frames listed in missing_frames [missing_count]
 * (unique, < F) replaced by the model's optical background -> camera (NULL =
 * direct optical intensities) -> out [F*N] f64. */
int fpmhip_simulate(fpmhip_context* context, const fpmhip_model* true_model, const double* object,
                    const fpmhip_camera* camera, const uint64_t* missing_frames, uint64_t missing_count,
                    uint64_t seed, double* out);

/* ---- reconstruction session: ReconstructionState + one algorithm ---- */
int fpmhip_session_create(const fpmhip_model* model, const fpmhip_measurements* measurements,
                          int32_t precision, const fpmhip_algorithm* algorithm,
                          fpmhip_session** out);
int fpmhip_session_destroy(fpmhip_session* session);
/* ReconstructionState::initialize_with_backend (src/reconstruction/state.rs:280-355)
 * followed by the algorithm's auxiliary state (mpie.rs:152-172) and, for
 * EPRY with pupil recovery, canonicalize_state (runner.rs:122). */
int fpmhip_session_initialize(fpmhip_session* session);
/* ReconstructionState::from_object (src/reconstruction/state.rs:356-383):
 * fpmhip_session_initialize, then the object spectrum replaced by
 * fftshift(FFT(object) / M) of `object` [R] complex128, and the algorithm's
 * auxiliary state (mPIE velocity 0, anchor = spectrum) and EPRY gauge made
 * from it. */
int fpmhip_session_initialize_from_object(fpmhip_session* session, const double* object);
/* Restores state (ReconstructionState::from_checkpoint_with_backend, state.rs:398-419):
 * spectrum [R] and pupil [N] complex128; mPIE velocity/anchor [R] (NULL = zero
 * velocity, anchor = spectrum) and counter; gains [F] / background (NULL = model). */
int fpmhip_session_set_state(fpmhip_session* session, const double* spectrum,
                             const double* pupil, const double* velocity, const double* anchor,
                             uint64_t counter, const double* frame_gains, const double* background,
                             uint64_t background_len);
/* One iteration: projection_update over `order` (F_order frame indices; NULL =
 * sequential 0..F-1), then canonicalize_state (runner.rs:156-227). Writes the
 * per-frame objective in order position (may be NULL) and the iteration
 * objective Σ w_f L_f / Σ w_f (NaN when Σ w_f = 0; metrics.rs:33). */
int fpmhip_session_sweep(fpmhip_session* session, const uint64_t* order, uint64_t order_len,
                         double* per_frame_objective, double* objective);
/* Frame order of FrameSchedule::order_for_problem (schedule.rs:35-104) for the
 * session's model and measurements, written to `order` [F]. */
int fpmhip_session_schedule(fpmhip_session* session, int32_t schedule, uint64_t* order);
/* Runner::run body (runner.rs:76-285) for `iterations` sweeps in `schedule`
 * order; trace [iterations] of objectives (may be NULL). */
int fpmhip_session_run(fpmhip_session* session, uint64_t iterations, int32_t schedule,
                       double* trace);
/* FrameSchedule::order_for_problem for zero-based iteration `iteration`
 * (schedule.rs:35-104; only RandomShuffle depends on it). SnrWeighted scores
 * the frames on the device from the resident measurements. */
int fpmhip_session_schedule_at(fpmhip_session* session, const fpmhip_frame_schedule* schedule,
                               uint64_t iteration, uint64_t* order);
/* The Runner's iteration loop (runner.rs:156-163) for iterations
 * first_iteration .. first_iteration + iterations - 1 (zero-based), each in
 * its own schedule order; trace [iterations] (may be NULL). */
int fpmhip_session_run_schedule(fpmhip_session* session, uint64_t iterations,
                                const fpmhip_frame_schedule* schedule, uint64_t first_iteration,
                                double* trace);

/* ---- diagnostics: build_diagnostics (src/reconstruction/runner.rs:298-433) ----
 * Of the current state, computed on the device: the forward prediction of
 * every frame with the state's frame gains and background, then per frame
 * one reduction over its pixels. IntensityStats (src/metrics/intensity/
 * single.rs:7-61) of the measured frames and IntensityComparisonMetrics
 * (src/metrics/intensity/compare.rs:88-110,134-226; reference = measured,
 * estimate = predicted, unmasked pixels) without a saturation threshold,
 * as the runner computes them. */
typedef struct fpmhip_intensity_stats {
  double mean, std, min, max, sum;
  uint64_t saturated_pixels, zero_pixels;
} fpmhip_intensity_stats;

typedef struct fpmhip_frame_comparison {
  double reference_sum, estimate_sum, residual_l1, residual_l2, residual_mean, residual_std,
      residual_max_abs, normalized_l2;
} fpmhip_frame_comparison;

/* Each non-NULL output is a DiagnosticRequest (src/diagnostics/context.rs:7-24):
 * per_frame_objective [F] (PerFrameError without a step: mean squared
 * amplitude residual over unmasked pixels, 0 for a fully masked frame),
 * raw_frame_statistics [F] (RawFrameStats), frame_summaries [F]
 * (FrameSummaries), residual_images [F*N] (ResidualImages: predicted -
 * measured, 0 where masked), object_amplitude / object_phase [M]
 * (ObjectAmplitude / ObjectPhase of state_object), pupil_amplitude /
 * pupil_phase [N] (Pupil). Errors as the reference's: a non-finite measured
 * value in RawFrameStats or a non-finite pair in FrameSummaries is
 * Numerical; a frame with no unmasked pixel in FrameSummaries is
 * InvalidParameter "valid_mask". */
typedef struct fpmhip_diagnostics {
  double* per_frame_objective;
  fpmhip_intensity_stats* raw_frame_statistics;
  fpmhip_frame_comparison* frame_summaries;
  double* residual_images;
  double* object_amplitude;
  double* object_phase;
  double* pupil_amplitude;
  double* pupil_phase;
} fpmhip_diagnostics;

int fpmhip_session_diagnostics(fpmhip_session* session, const fpmhip_diagnostics* out);
/* Downloads (complex arrays interleaved). object = state_object (result.rs:397-413). */
int fpmhip_session_get_spectrum(fpmhip_session* session, double* out);
int fpmhip_session_get_pupil(fpmhip_session* session, double* out);
int fpmhip_session_get_object(fpmhip_session* session, double* out);
int fpmhip_session_get_momentum(fpmhip_session* session, double* velocity, double* anchor,
                                uint64_t* counter);
int fpmhip_session_get_calibration(fpmhip_session* session, double* frame_gains,
                                   double* background, uint64_t* background_len);
/* Whether the state holds frame gains (ReconstructionState::frame_gains() is
 * Some, state.rs:216) and its background length (0 = None, state.rs:221). */
int fpmhip_session_calibration_present(const fpmhip_session* session, int32_t* frame_gains,
                                       uint64_t* background_len);
/* Drops the state's frame gains (all frames then use gain 1) and/or its
 * background (0), as a checkpoint with `frame_gains` / `background` None
 * restores them (state.rs:411-412). */
int fpmhip_session_clear_calibration(fpmhip_session* session, int32_t frame_gains, int32_t background);
int fpmhip_session_shape(const fpmhip_session* session, uint64_t* image_h, uint64_t* image_w,
                         uint64_t* recon_h, uint64_t* recon_w, uint64_t* frame_count);

/* ---- batched form: many independent tiles (each its own model, pupil and
 * measurements, uneven sizes) advanced together, one launch per sweep for
 * all tiles, captured in a HIP graph. Sessions must be MIXED_F32, sequential
 * schedule, square 32/64/128/256 images without multiplexing, and all use
 * the same algorithm kind (AP, FPIE or mPIE). ---- */
int fpmhip_batch_create(fpmhip_context* context, fpmhip_session* const* sessions, uint64_t count,
                        fpmhip_batch** out);
int fpmhip_batch_destroy(fpmhip_batch* batch);
/* Initialise every tile, run `iterations` sweeps (graph replays) and finish
 * each tile's result; trace [count][iterations] objectives (may be NULL). */
int fpmhip_batch_run(fpmhip_batch* batch, uint64_t iterations, double* trace);
/* Device time of the sweep launches of the last fpmhip_batch_run, measured
 * with HIP events on the batch stream: mean and max per sweep (ms), count. */
int fpmhip_batch_timing(const fpmhip_batch* batch, double* sweep_ms_mean, double* sweep_ms_max,
                        uint64_t* sweeps);
/* Every tile's result object (state_object, result.rs:397-413), in session
 * order, concatenated into `out` (sum of the tiles' R_h * R_w complex128
 * values, interleaved): fpmhip_session_get_object for all tiles, computed by
 * one launch for all tiles and one batched FFT per distinct R, then copied. */
int fpmhip_batch_objects(fpmhip_batch* batch, double* out);
/* The same objects left in device memory (HBM): objects[count] receives one
 * device pointer per tile, in session order, to its R_h x R_w row-major
 * complex128 object (interleaved). Returns once they are complete; valid
 * until the next run or objects call on this batch, or its destruction. */
int fpmhip_batch_objects_device(fpmhip_batch* batch, const double** objects);
/* The same into a page-locked host buffer owned by the batch (allocated on
 * first use, freed with the batch), for full-rate PCIe copies: *out points
 * at the concatenated objects and stays valid until the next call on this
 * batch or its destruction. */
int fpmhip_batch_objects_pinned(fpmhip_batch* batch, const double** out);

/* Schedule of the batch's sweep launch (no counterpart in the reference; the
 * reference runs one tile at a time on the host). */
typedef struct fpmhip_batch_schedule {
  uint32_t group_members; /* workgroups per 256^2 tile: 0 (default) cooperative groups of 8 plus
                             an equal share of the co-resident room left over, up to 32; 8..32
                             that many; 1 none. Groups (512^2 tiles: 32) are used only when the
                             whole grid can be resident at once */
  uint32_t largest_first; /* 1: launch the largest tiles first; 0: session order */
  uint32_t radix[4];      /* first-stage FFT radix for 32, 64, 128, 256 pixel tiles (0: default;
                             32: 8|4, 64: 8|4|16, 128: 8|16, 256: 16) */
  uint32_t sweeps_per_launch; /* sweeps one launch runs (1..1000; 0: default, 50): fewer launches
                                 and host synchronisations per run */
} fpmhip_batch_schedule;
int fpmhip_batch_get_schedule(const fpmhip_batch* batch, fpmhip_batch_schedule* out);
/* Rebuilds the captured launch with `schedule`. */
int fpmhip_batch_set_schedule(fpmhip_batch* batch, const fpmhip_batch_schedule* schedule);

typedef struct fpmhip_batch_tuning {
  uint32_t candidates; /* schedules timed */
  double first_ms;     /* mean sweep time of the schedule in force before tuning */
  double best_ms;      /* mean sweep time of the schedule chosen */
} fpmhip_batch_tuning;
/* Autotuner: times candidate schedules on the batch's own tiles (each from
 * the tiles' initial state, one untimed sweep then `trial_sweeps` timed
 * sweeps) by coordinate descent over the group size, the radix of each tile
 * size present and the launch order, and keeps the fastest (a change must
 * gain 2%). Tile state afterwards is that of the trials; fpmhip_batch_run
 * re-initialises every tile. */
int fpmhip_batch_tune(fpmhip_batch* batch, uint32_t trial_sweeps, fpmhip_batch_tuning* report);

/* ---- serving form: continuous batching of reconstruction jobs ----
 * A worker thread owns the device. Submitted jobs wait in a queue and are
 * admitted while tile slots and the HBM budget allow; every admitted job is
 * one tile whose state stays resident until it finishes. Each tick is one
 * fused sweep launch over all resident tiles, each at its own iteration;
 * a tile runs a quota of sweeps per tick inversely proportional to its
 * per-sweep work (LEDs x pixels; at least one), so small tiles are not held
 * to the pace of the largest resident one. Finished tiles leave and queued
 * jobs enter between ticks. A job's result is identical whether it ran
 * alone or among others. Jobs must satisfy the batched form's constraints
 * (MIXED_F32, square 32/64/128/256 images, no multiplexing, not EPRY). */
typedef struct fpmhip_server_config {
  uint64_t max_tiles;        /* resident tiles at most; admission also keeps the resident tiles'
                                workgroups (8 per 256^2 tile, a cooperative group) within what
                                the device holds at once */
  uint64_t hbm_budget_bytes; /* admission limit for resident job state */
} fpmhip_server_config;

typedef struct fpmhip_server_statistics {
  uint64_t submitted, admitted, completed, failed;
  uint64_t ticks;            /* fused sweep launches */
  uint64_t resident_tiles;   /* at the time of the call */
  uint64_t resident_bytes;
  double busy_seconds;       /* wall time the worker spent with resident tiles */
  uint64_t pool_reuses;      /* job buffers served from the server's device memory pool */
  uint64_t pool_allocations; /* job buffers the pool had to allocate (hipMalloc) */
  uint64_t peak_tick_workgroups; /* largest grid of one tick (at most the device's co-resident
                                    k_sweep workgroups) */
} fpmhip_server_statistics;

int fpmhip_server_create(fpmhip_context* context, const fpmhip_server_config* config,
                         fpmhip_server** out);
/* Stops the worker after the current tick; unfinished jobs are dropped. */
int fpmhip_server_destroy(fpmhip_server* server);
/* Copies the job's inputs (model descriptor arrays, counts or frames,
 * optional weights) and queues it; returns immediately with its id. */
int fpmhip_server_submit(fpmhip_server* server, const fpmhip_model_desc* model,
                         const uint16_t* counts_u16, const double* frames_f64, const double* weights,
                         const fpmhip_algorithm* algorithm, uint64_t* job_id);
/* state: 0 queued, 1 resident, 2 done, 3 failed. */
int fpmhip_server_poll(fpmhip_server* server, uint64_t job_id, int32_t* state);
/* Blocks until the job is done or failed (timeout_s <= 0: no limit). */
int fpmhip_server_wait(fpmhip_server* server, uint64_t job_id, double timeout_s);
/* Results of a finished job: object [R] complex128, spectrum [R] complex128,
 * trace [iterations] (each may be NULL); latency = completion - submission
 * (s), queue_s = admission - submission (s). Frees the job's record. */
int fpmhip_server_result(fpmhip_server* server, uint64_t job_id, double* object, double* spectrum,
                         double* trace, double* latency_s, double* queue_s);
int fpmhip_server_stats(fpmhip_server* server, fpmhip_server_statistics* out);

#ifdef __cplusplus
}
#endif
#endif /* FPMHIP_H */
