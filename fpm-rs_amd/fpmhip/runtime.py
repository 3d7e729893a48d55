"""ctypes binding of libfpmhip.so (include/fpmhip.h) and the host-side mirror
of fpm-rs's Python surface: compile_model, simulate's forward stack, and
`Algorithm.run(problem)` (python/src/reconstruction.rs:1107 run_algorithm).

There is no CPU fallback: a missing library or device raises.
"""
from __future__ import annotations

import atexit
import ctypes as ct
import os
import weakref
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi
from .model import ImagePlaneModel, MeasurementStack, ReconstructionProblem

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfpmhip.so")

EXPORTS = [
    "fpmhip_last_error", "fpmhip_last_error_name", "fpmhip_version", "fpmhip_compile_planar",
    "fpmhip_algorithm_default", "fpmhip_algorithm_validate", "fpmhip_context_create", "fpmhip_context_destroy",
    "fpmhip_context_synchronize", "fpmhip_fft2_host", "fpmhip_model_create", "fpmhip_model_destroy",
    "fpmhip_measurements_create", "fpmhip_measurements_destroy", "fpmhip_forward_stack",
    "fpmhip_forward_stack_noalloc", "fpmhip_context_capabilities", "fpmhip_buffer_allocate", "fpmhip_buffer_free",
    "fpmhip_buffer_info", "fpmhip_buffer_upload", "fpmhip_buffer_download", "fpmhip_fft2_resident",
    "fpmhip_session_initialize_from_object", "fpmhip_session_create",
    "fpmhip_session_destroy", "fpmhip_session_initialize", "fpmhip_session_set_state", "fpmhip_session_sweep",
    "fpmhip_session_schedule", "fpmhip_session_run", "fpmhip_session_get_spectrum", "fpmhip_session_get_pupil",
    "fpmhip_session_get_object", "fpmhip_session_get_momentum", "fpmhip_session_get_calibration",
    "fpmhip_session_calibration_present", "fpmhip_session_clear_calibration",
    "fpmhip_session_shape", "fpmhip_batch_create", "fpmhip_batch_destroy", "fpmhip_batch_run",
    "fpmhip_batch_timing", "fpmhip_batch_objects", "fpmhip_batch_objects_pinned", "fpmhip_batch_objects_device", "fpmhip_batch_get_schedule", "fpmhip_batch_set_schedule", "fpmhip_batch_tune",
    "fpmhip_server_create", "fpmhip_server_destroy", "fpmhip_server_submit",
    "fpmhip_server_poll", "fpmhip_server_wait", "fpmhip_server_result", "fpmhip_server_stats",
    "fpmhip_camera_validate", "fpmhip_camera_measure", "fpmhip_simulate", "fpmhip_session_schedule_at",
    "fpmhip_session_run_schedule", "fpmhip_session_diagnostics",
]


class FpmError(RuntimeError):
    """An fpm_rs::Error variant returned through the C ABI."""

    def __init__(self, code, message, name=""):
        self.code = code
        self.kind = abi.STATUS_NAMES.get(code, str(code))
        self.message = message
        self.name = name
        super().__init__(f"{self.kind}: {message}" + (f" [{name}]" if name else ""))


_lib = None

# Ordered teardown. Every live handle is registered here; at interpreter exit
# (or on close_all()) they are destroyed dependents first: servers, batches,
# sessions, measurements, models, then contexts, while the HIP runtime is
# still up. After that, __del__ does nothing, so that no destroy call runs in
# an arbitrary order during the interpreter's own finalization.
_LEVELS = ("server", "batch", "session", "measurements", "model", "buffer", "context")
_live = {k: weakref.WeakSet() for k in _LEVELS}
_closed = False


def _register(level, obj):
    _live[level].add(obj)


def close_all():
    """Destroys every live handle of this process, dependents first (also
    run at exit). Handles closed this way are unusable afterwards."""
    global _closed, _default_context
    for level in _LEVELS:
        for obj in list(_live[level]):
            try:
                obj.close()
            except Exception:
                pass
    _default_context = None
    _closed = True


atexit.register(close_all)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libfpmhip.so is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ct.CDLL(LIB_PATH)
        for name in EXPORTS:
            getattr(L, name).restype = ct.c_int
        L.fpmhip_last_error.restype = ct.c_char_p
        L.fpmhip_last_error_name.restype = ct.c_char_p
        L.fpmhip_version.restype = ct.c_char_p
        _lib = L
    return _lib


def check(code):
    if code != 0:
        L = lib()
        raise FpmError(code, L.fpmhip_last_error().decode(), L.fpmhip_last_error_name().decode())


def _cd(a):
    return a.ctypes.data_as(abi.P_dbl)


def compile_model(optics, array, image_shape, reconstruction_shape="smooth") -> ImagePlaneModel:
    """fpm_rs.compile_model (python/src/model.rs:163-172; default 'smooth'),
    computed by libfpmhip's host code (no device needed)."""
    if isinstance(reconstruction_shape, str):
        policy, eh, ew = abi.SHAPE[reconstruction_shape], 0, 0
    else:
        policy = abi.SHAPE["exact"]
        eh, ew = reconstruction_shape
    oc, ac = optics.to_c(), array.to_c()
    shape = abi.CompiledShape()
    L = lib()
    args = (ct.byref(oc), ct.byref(ac), ct.c_uint64(image_shape[0]), ct.c_uint64(image_shape[1]),
            ct.c_int32(policy), ct.c_uint64(eh), ct.c_uint64(ew), ct.byref(shape))
    check(L.fpmhip_compile_planar(*args, None, None, None, None, None, None))
    S = shape.source_count
    crop = np.zeros((S, 2), np.uint64)
    off = np.zeros((S, 2), np.float64)
    kv = np.zeros((S, 2), np.float64)
    pupil = np.zeros(image_shape, np.complex128)
    support = np.zeros(image_shape, np.uint8)
    gains = np.zeros(S, np.float64)
    check(L.fpmhip_compile_planar(*args, abi.ptr(crop, ct.c_uint64), _cd(off), _cd(kv), _cd(pupil),
                                  abi.ptr(support, ct.c_uint8), _cd(gains)))
    return ImagePlaneModel(tuple(image_shape), (shape.recon_h, shape.recon_w), crop, off, pupil, support,
                           frame_gain=gains, k_vectors=kv)


def algorithm_default(kind) -> abi.Algorithm:
    a = abi.Algorithm()
    check(lib().fpmhip_algorithm_default(ct.c_int32(kind), ct.byref(a)))
    return a


def validate(algorithm: abi.Algorithm):
    check(lib().fpmhip_algorithm_validate(ct.byref(algorithm)))


class Context:
    """One HIP stream on one device (fpmhip_context)."""

    def __init__(self, device=0):
        self.handle = ct.c_void_p()
        check(lib().fpmhip_context_create(ct.c_int32(device), ct.byref(self.handle)))
        _register("context", self)

    def close(self):
        if self.handle:
            lib().fpmhip_context_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        if not _closed:
            try:
                self.close()
            except Exception:
                pass

    def synchronize(self):
        check(lib().fpmhip_context_synchronize(self.handle))

    def fft2(self, values, direction="forward"):
        v = np.array(values, dtype=np.complex128, order="C", copy=True)
        check(lib().fpmhip_fft2_host(self.handle, _cd(v), ct.c_uint64(v.shape[0]), ct.c_uint64(v.shape[1]),
                                     ct.c_int32(abi.FORWARD if direction == "forward" else abi.INVERSE)))
        return v


_default_context = None


def default_context() -> Context:
    global _default_context
    if _default_context is None:
        _default_context = Context(0)
    return _default_context


class DeviceModel:
    def __init__(self, ctx: Context, model: ImagePlaneModel):
        self.ctx = ctx
        self.model = model
        self.handle = ct.c_void_p()
        d = model.desc()
        check(lib().fpmhip_model_create(ctx.handle, ct.byref(d), ct.byref(self.handle)))
        _register("model", self)

    def close(self):
        if self.handle:
            lib().fpmhip_model_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        if not _closed:
            try:
                self.close()
            except Exception:
                pass


class DeviceMeasurements:
    def __init__(self, ctx: Context, dmodel: DeviceModel, meas: MeasurementStack):
        self.ctx = ctx
        self.dmodel = dmodel
        self.handle = ct.c_void_p()
        counts = meas.counts
        frames = meas.frames
        w = meas.weights
        m = meas.masks
        check(lib().fpmhip_measurements_create(
            ctx.handle, dmodel.handle, None if counts is None else abi.ptr(counts, ct.c_uint16),
            None if frames is None else _cd(frames), None if w is None else _cd(w),
            None if m is None else abi.ptr(m, ct.c_uint8), ct.byref(self.handle)))
        _register("measurements", self)

    def close(self):
        if self.handle:
            lib().fpmhip_measurements_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        if not _closed:
            try:
                self.close()
            except Exception:
                pass


def forward_stack(model: ImagePlaneModel, spectrum, pupil=None, context: Optional[Context] = None,
                  dmodel: Optional[DeviceModel] = None):
    """ForwardModel::forward_intensity_stack_into on the device: (F, n_h, n_w) f64."""
    ctx = context or default_context()
    dm = dmodel or DeviceModel(ctx, model)
    s = np.ascontiguousarray(spectrum, dtype=np.complex128)
    p = None if pupil is None else np.ascontiguousarray(pupil, dtype=np.complex128)
    out = np.zeros((model.frame_count,) + tuple(model.image_shape), np.float64)
    check(lib().fpmhip_forward_stack(ctx.handle, dm.handle, _cd(s), None if p is None else _cd(p), _cd(out)))
    return out


class Capabilities(ct.Structure):
    _fields_ = [("preferred_memory", ct.c_int32), ("resident_buffers", ct.c_int32)]


def capabilities(context: Optional[Context] = None) -> dict:
    """Backend::capabilities (src/backend/traits.rs:27-33)."""
    ctx = context or default_context()
    c = Capabilities()
    check(lib().fpmhip_context_capabilities(ctx.handle, ct.byref(c)))
    return {"preferred_memory": "device" if c.preferred_memory == 1 else "host",
            "resident_buffers": bool(c.resident_buffers)}


COMPLEX, REAL = 0, 1


class ResidentBuffer:
    """A ComplexBuffer / RealBuffer of the ResidentBackend seam
    (src/backend/traits.rs:44-106) in device memory: `len` complex128 or
    f64 values, zeroed at allocation."""

    def __init__(self, kind, length, context: Optional[Context] = None):
        self.ctx = context or default_context()
        self.kind = COMPLEX if kind in ("complex", COMPLEX) else REAL
        self.handle = ct.c_void_p()
        check(lib().fpmhip_buffer_allocate(self.ctx.handle, ct.c_int32(self.kind), ct.c_uint64(length),
                                           ct.byref(self.handle)))
        _register("buffer", self)

    def close(self):
        if self.handle:
            lib().fpmhip_buffer_free(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        if not _closed:
            try:
                self.close()
            except Exception:
                pass

    def info(self) -> dict:
        k, n, loc, dev = ct.c_int32(), ct.c_uint64(), ct.c_int32(), ct.c_void_p()
        check(lib().fpmhip_buffer_info(self.handle, ct.byref(k), ct.byref(n), ct.byref(loc), ct.byref(dev)))
        return {"kind": "complex" if k.value == COMPLEX else "real", "len": n.value,
                "location": "device" if loc.value == 1 else "host", "device_address": dev.value or 0}

    def __len__(self):
        return self.info()["len"]

    @property
    def device_address(self) -> int:
        return self.info()["device_address"]

    def upload(self, values, kind=None):
        """upload_complex / upload_real (traits.rs:82-93); `kind` defaults to the buffer's."""
        k = self.kind if kind is None else (COMPLEX if kind in ("complex", COMPLEX) else REAL)
        a = np.ascontiguousarray(values, dtype=np.complex128 if k == COMPLEX else np.float64).ravel()
        check(lib().fpmhip_buffer_upload(self.ctx.handle, self.handle, ct.c_int32(k), _cd(a.view(np.float64)),
                                         ct.c_uint64(a.size)))

    def download(self, length=None, kind=None):
        """download_complex / download_real (traits.rs:86-95)."""
        k = self.kind if kind is None else (COMPLEX if kind in ("complex", COMPLEX) else REAL)
        n = len(self) if length is None else length
        out = np.zeros(n, np.complex128 if k == COMPLEX else np.float64)
        check(lib().fpmhip_buffer_download(self.ctx.handle, self.handle, ct.c_int32(k), _cd(out.view(np.float64)),
                                           ct.c_uint64(n)))
        return out

    def fft2(self, shape, direction="forward"):
        """fft2_resident (traits.rs:97-103), in place."""
        check(lib().fpmhip_fft2_resident(self.ctx.handle, self.handle, ct.c_uint64(shape[0]), ct.c_uint64(shape[1]),
                                         ct.c_int32(abi.FORWARD if direction == "forward" else abi.INVERSE)))


def forward_stack_noalloc(context: Context, dmodel: DeviceModel, spectrum_dev: int, pupil_dev: int, out_dev: int):
    """fpmhip_forward_stack_noalloc: the forward stack over caller-owned
    device buffers (addresses as ints; pupil_dev 0 = the model pupil), with
    no device allocation. Enqueued on the context's stream: synchronise the
    context before reading `out_dev`."""
    check(lib().fpmhip_forward_stack_noalloc(context.handle, dmodel.handle, ct.c_void_p(spectrum_dev),
                                             ct.c_void_p(pupil_dev or None), ct.c_void_p(out_dev)))


class Session:
    """A ReconstructionState on the device plus one algorithm."""

    def __init__(self, problem: ReconstructionProblem, algorithm, precision="c128", context=None,
                 dmodel=None, dmeas=None):
        self.ctx = context or default_context()
        self.problem = problem
        self.dmodel = dmodel or DeviceModel(self.ctx, problem.model)
        self.dmeas = dmeas or DeviceMeasurements(self.ctx, self.dmodel, problem.measurements)
        self.alg = algorithm if isinstance(algorithm, abi.Algorithm) else algorithm.to_c()
        self.precision = abi.PLAIN_C128 if precision in ("c128", "plain", abi.PLAIN_C128) else abi.MIXED_F32
        self.handle = ct.c_void_p()
        check(lib().fpmhip_session_create(self.dmodel.handle, self.dmeas.handle, ct.c_int32(self.precision),
                                          ct.byref(self.alg), ct.byref(self.handle)))
        m = problem.model
        self.R = tuple(m.reconstruction_shape)
        self.N = tuple(m.image_shape)
        self.F = m.frame_count
        _register("session", self)

    def close(self):
        if self.handle:
            lib().fpmhip_session_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        if not _closed:
            try:
                self.close()
            except Exception:
                pass

    def initialize(self):
        check(lib().fpmhip_session_initialize(self.handle))

    def initialize_from_object(self, obj):
        """ReconstructionState::from_object (state.rs:356-383)."""
        o = np.ascontiguousarray(obj, dtype=np.complex128)
        check(lib().fpmhip_session_initialize_from_object(self.handle, _cd(o)))

    def set_state(self, spectrum, pupil, velocity=None, anchor=None, counter=0, frame_gains=None, background=None):
        """Checkpoint restore (ReconstructionState::from_checkpoint, state.rs:398-419);
        background: one image or one per frame."""
        keep = []

        def c(a, dt=np.complex128):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return _cd(a)

        bg = None if background is None else np.ascontiguousarray(background, dtype=np.float64).ravel()
        check(lib().fpmhip_session_set_state(self.handle, c(spectrum), c(pupil), c(velocity), c(anchor),
                                             ct.c_uint64(counter), c(frame_gains, np.float64), c(bg, np.float64),
                                             ct.c_uint64(0 if bg is None else bg.size)))

    def sweep(self, order=None):
        per = np.zeros(self.F if order is None else len(order), np.float64)
        obj = ct.c_double(0.0)
        o = None if order is None else np.ascontiguousarray(order, dtype=np.uint64)
        check(lib().fpmhip_session_sweep(self.handle, None if o is None else abi.ptr(o, ct.c_uint64),
                                         ct.c_uint64(0 if o is None else o.size), _cd(per), ct.byref(obj)))
        return obj.value, per

    def schedule(self, name="sequential", iteration=0, seed=0):
        """FrameSchedule::order_for_problem (schedule.rs:35-104) for a
        zero-based iteration; `seed` is RandomShuffle's."""
        order = np.zeros(self.F, np.uint64)
        sched = abi.FrameSchedule(abi.SCHEDULE[name], 0, seed)
        check(lib().fpmhip_session_schedule_at(self.handle, ct.byref(sched), ct.c_uint64(iteration),
                                               abi.ptr(order, ct.c_uint64)))
        return order

    def run(self, iterations, schedule="sequential", seed=0, first_iteration=0):
        """`iterations` sweeps, iteration i (zero-based, from first_iteration)
        in its schedule order (runner.rs:156-163); returns the objective trace."""
        trace = np.zeros(iterations, np.float64)
        sched = abi.FrameSchedule(abi.SCHEDULE[schedule], 0, seed)
        check(lib().fpmhip_session_run_schedule(self.handle, ct.c_uint64(iterations), ct.byref(sched),
                                                ct.c_uint64(first_iteration), _cd(trace)))
        return trace

    def diagnostics(self, requests, iteration=None):
        """build_diagnostics (runner.rs:298-433) of the current state for a
        set of DiagnosticRequest names (diagnostics/context.rs:7-24):
        "per_frame_error", "raw_frame_stats", "frame_summaries",
        "residual_images", "object_amplitude", "object_phase", "pupil".
        Computed on the device; returns a Diagnostics."""
        req = set(requests)
        unknown = req - set(DIAGNOSTIC_REQUESTS)
        if unknown:
            raise ValueError(f"unknown diagnostic requests {sorted(unknown)}")
        F, n_img, M = self.F, self.N[0] * self.N[1], self.R[0] * self.R[1]
        d = abi.Diagnostics()
        arrays = {}

        def want(name, key, shape, ctype=None):
            if name in req:
                a = np.zeros(shape, np.float64) if ctype is None else (ctype * F)()
                arrays[key] = a
                setattr(d, key, _cd(a) if ctype is None else ct.cast(a, ct.POINTER(ctype)))

        want("per_frame_error", "per_frame_objective", F)
        want("raw_frame_stats", "raw_frame_statistics", F, abi.IntensityStats)
        want("frame_summaries", "frame_summaries", F, abi.FrameComparison)
        want("residual_images", "residual_images", (F,) + self.N)
        want("object_amplitude", "object_amplitude", self.R)
        want("object_phase", "object_phase", self.R)
        want("pupil", "pupil_amplitude", self.N)
        want("pupil", "pupil_phase", self.N)
        check(lib().fpmhip_session_diagnostics(self.handle, ct.byref(d)))
        out = Diagnostics()
        out.per_frame_objective = arrays.get("per_frame_objective")
        if "raw_frame_statistics" in arrays:
            out.raw_frame_statistics = [
                {"frame_index": f, "metrics": {k: getattr(r, k) for k, _ in abi.IntensityStats._fields_}}
                for f, r in enumerate(arrays["raw_frame_statistics"])]
        if "frame_summaries" in arrays:
            out.frame_diagnostics = [
                {"iteration": iteration, "frame_index": f, "illumination_index": f,
                 "metrics": dict({k: getattr(r, k) for k, _ in abi.FrameComparison._fields_}, saturated_pixels=None)}
                for f, r in enumerate(arrays["frame_summaries"])]
        out.residual_images = arrays.get("residual_images")
        out.object_amplitude = arrays.get("object_amplitude")
        out.object_phase = arrays.get("object_phase")
        out.pupil_amplitude = arrays.get("pupil_amplitude")
        out.pupil_phase = arrays.get("pupil_phase")
        return out

    def spectrum(self):
        out = np.zeros(self.R, np.complex128)
        check(lib().fpmhip_session_get_spectrum(self.handle, _cd(out)))
        return out

    def pupil(self):
        out = np.zeros(self.N, np.complex128)
        check(lib().fpmhip_session_get_pupil(self.handle, _cd(out)))
        return out

    def object(self):
        out = np.zeros(self.R, np.complex128)
        check(lib().fpmhip_session_get_object(self.handle, _cd(out)))
        return out

    def calibration(self):
        """Frame gains (F,) and background of the state: None, one image
        (N,) or one per frame (F, N) (EPRY recovers both)."""
        gains = np.zeros(self.F, np.float64)
        n_img = self.N[0] * self.N[1]
        bg = np.zeros(self.F * n_img, np.float64)
        n = ct.c_uint64(0)
        check(lib().fpmhip_session_get_calibration(self.handle, _cd(gains), _cd(bg), ct.byref(n)))
        if n.value == 0:
            return gains, None
        return gains, (bg[:n_img].reshape(self.N) if n.value == n_img else bg.reshape((self.F,) + self.N))

    def calibration_present(self):
        """(state holds frame gains, background length: 0, N or F*N)."""
        g = ct.c_int32(0)
        n = ct.c_uint64(0)
        check(lib().fpmhip_session_calibration_present(self.handle, ct.byref(g), ct.byref(n)))
        return bool(g.value), n.value

    def clear_calibration(self, frame_gains=True, background=True):
        check(lib().fpmhip_session_clear_calibration(self.handle, ct.c_int32(int(bool(frame_gains))),
                                                     ct.c_int32(int(bool(background)))))

    def momentum(self):
        v = np.zeros(self.R, np.complex128)
        a = np.zeros(self.R, np.complex128)
        c = ct.c_uint64(0)
        check(lib().fpmhip_session_get_momentum(self.handle, _cd(v), _cd(a), ct.byref(c)))
        return v, a, c.value


DIAGNOSTIC_REQUESTS = ("per_frame_error", "raw_frame_stats", "frame_summaries", "residual_images",
                       "object_amplitude", "object_phase", "pupil")


@dataclass
class Diagnostics:
    """Diagnostics (src/diagnostics/context.rs:26-48); None = not requested."""

    objective: Optional[float] = None
    per_frame_objective: Optional[np.ndarray] = None
    raw_frame_statistics: Optional[list] = None
    frame_diagnostics: Optional[list] = None
    object_amplitude: Optional[np.ndarray] = None
    object_phase: Optional[np.ndarray] = None
    pupil_amplitude: Optional[np.ndarray] = None
    pupil_phase: Optional[np.ndarray] = None
    residual_images: Optional[np.ndarray] = None


@dataclass
class ReconstructionResult:
    """fpm_rs ReconstructionResult fields (src/reconstruction/result.rs:40-76)."""

    object: np.ndarray
    amplitude: np.ndarray
    phase: np.ndarray
    object_spectrum: np.ndarray
    recovered_pupil: np.ndarray
    trace: np.ndarray
    extra: dict = field(default_factory=dict)


def mpie_parameters(alg: abi.Algorithm) -> dict:
    """The mPIE parameters a checkpoint's auxiliary state records (state.rs:34-53)."""
    return dict(object_step=alg.object_step, stability=alg.stability, epsilon=alg.epsilon,
                momentum_interval=int(alg.momentum_interval), momentum_friction=alg.momentum_friction,
                momentum_feedback=alg.momentum_feedback)


def _restore(s: Session, alg: abi.Algorithm, problem: ReconstructionProblem, ck):
    """ReconstructionState::from_checkpoint (state.rs:398-419) and the mPIE
    auxiliary check of prepare_auxiliary (algorithms/mpie.rs:166-222)."""
    from .checkpoint import MpieAuxiliary

    ck.validate_for_problem(problem)
    if ck.illumination_corrections is not None or ck.physical_illumination_calibration is not None:
        raise FpmError(abi.STATUS_CODES["Unsupported"],
                       "checkpoints with illumination corrections or physical calibration resume only in "
                       "fpm-rs's calibration solvers")
    aux = ck.algorithm_auxiliary
    v = a = None
    counter = 0
    if alg.kind == abi.MPIE and aux is not None:
        if not isinstance(aux, MpieAuxiliary):
            raise FpmError(abi.STATUS_CODES["InvalidModel"],
                           "mPIE cannot resume auxiliary state owned by another algorithm")
        mine = mpie_parameters(alg)
        for name in ("object_step", "stability", "epsilon", "momentum_friction", "momentum_feedback"):
            cur, stored = float(mine[name]), float(getattr(aux, name))
            if np.float64(cur).view(np.uint64) != np.float64(stored).view(np.uint64):
                raise FpmError(abi.STATUS_CODES["InvalidParameter"],
                               f"value {cur} differs from checkpointed mPIE value {stored}", name)
        if mine["momentum_interval"] != aux.momentum_interval:
            raise FpmError(abi.STATUS_CODES["InvalidParameter"],
                           f"value {mine['momentum_interval']} differs from checkpointed mPIE value "
                           f"{aux.momentum_interval}", "momentum_interval")
        v, a, counter = aux.velocity, aux.anchor, aux.effective_frames_since_momentum
    s.set_state(ck.object_spectrum, ck.pupil_values, velocity=v, anchor=a, counter=counter,
                frame_gains=ck.frame_gains, background=ck.background)
    s.clear_calibration(frame_gains=ck.frame_gains is None, background=ck.background is None)


def run_algorithm(algorithm, problem: ReconstructionProblem, precision="c128", schedule="sequential",
                  context=None, callbacks=None, resume_from=None, schedule_seed=0) -> ReconstructionResult:
    """Runner::run (src/reconstruction/runner.rs:76-285) on one device session.

    `resume_from` restores a ReconstructionCheckpoint and continues from its
    completed iterations up to `algorithm.iterations` (runner.rs:94-121);
    `callbacks` takes CheckpointEvery (callbacks/checkpoint.rs). The trace
    holds the checkpoint's records followed by the new ones, with elapsed
    seconds continuing from the last restored record. `schedule_seed` is
    RandomShuffle's seed; iteration i (zero-based) is ordered for i, also
    after a resume."""
    import time

    from .checkpoint import CheckpointEvery, ReconstructionCheckpoint

    alg = algorithm.to_c()
    validate(alg)
    callbacks = list(callbacks or [])
    for cb in callbacks:
        if not isinstance(cb, CheckpointEvery):
            raise FpmError(abi.STATUS_CODES["Unsupported"], f"unsupported callback {type(cb).__name__}")
    started = time.perf_counter()
    s = Session(problem, alg, precision=precision, context=context)
    if resume_from is not None:
        _restore(s, alg, problem, resume_from)
        records = list(resume_from.trace)
        metrics = list(resume_from.algorithm_metrics)
        first = int(resume_from.completed_iterations)
        carried_aux = None if alg.kind == abi.MPIE else resume_from.algorithm_auxiliary
    else:
        s.initialize()
        records, metrics, first, carried_aux = [], [], 0, None
    previous = records[-1][2] if records else 0.0
    for cb in callbacks:
        cb.on_start()
    total = int(algorithm.iterations)
    mparams = mpie_parameters(alg) if alg.kind == abi.MPIE else None
    if callbacks:
        for it in range(first, total):
            obj = float(s.run(1, schedule, schedule_seed, first_iteration=it)[0])
            records.append((it + 1, obj, previous + time.perf_counter() - started))
            for cb in callbacks:
                if cb.due(it + 1):
                    ck = ReconstructionCheckpoint.capture(it + 1, s, records, mparams, metrics, carried_aux)
                    cb.write(ck)
    elif total > first:
        objs = s.run(total - first, schedule, schedule_seed, first_iteration=first)
        end = previous + time.perf_counter() - started
        # one timestamp for the whole device run: the records share it
        records += [(first + k + 1, float(o), end) for k, o in enumerate(objs)]
    trace = np.array([r[1] for r in records], np.float64)
    obj = s.object()
    res = ReconstructionResult(obj, np.abs(obj), np.angle(obj), s.spectrum(), s.pupil(), trace)
    res.extra.update(trace_records=records)
    if alg.kind == abi.MPIE:
        v, a, c = s.momentum()
        res.extra.update(velocity=v, anchor=a, counter=c)
    if alg.kind == abi.EPRY and (alg.recover_gain or alg.recover_background):
        gains, bg = s.calibration()
        res.extra.update(recovered_frame_gains=gains, recovered_background=bg)
    return res


class Batch:
    """Many MIXED_F32 sessions (tiles) advanced together: one fused sweep
    launch per iteration for all tiles, replayed from a HIP graph."""

    def __init__(self, sessions, context=None):
        self.ctx = context or sessions[0].ctx
        self.sessions = list(sessions)
        arr = (ct.c_void_p * len(self.sessions))(*[s.handle.value for s in self.sessions])
        self.handle = ct.c_void_p()
        check(lib().fpmhip_batch_create(self.ctx.handle, arr, ct.c_uint64(len(self.sessions)),
                                        ct.byref(self.handle)))
        _register("batch", self)

    def close(self):
        if self.handle:
            lib().fpmhip_batch_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        if not _closed:
            try:
                self.close()
            except Exception:
                pass

    def run(self, iterations):
        """Initialise every tile and run `iterations` sweeps; returns the
        (tiles, iterations) objective trace."""
        trace = np.zeros((len(self.sessions), iterations), np.float64)
        check(lib().fpmhip_batch_run(self.handle, ct.c_uint64(iterations), _cd(trace)))
        return trace

    def objects(self, pinned=False):
        """Every tile's result object (one launch queue, one wait). With
        pinned=True the arrays are views of the batch's page-locked result
        buffer, valid until the next objects() call or the batch's release
        (copy them to keep them)."""
        shapes = [tuple(s.problem.model.reconstruction_shape) for s in self.sessions]
        total = sum(h * w for h, w in shapes)
        if pinned:
            p = ct.POINTER(ct.c_double)()
            check(lib().fpmhip_batch_objects_pinned(self.handle, ct.byref(p)))
            flat = np.ctypeslib.as_array(p, shape=(2 * total,)).view(np.complex128)
        else:
            flat = np.zeros(total, np.complex128)
            check(lib().fpmhip_batch_objects(self.handle, _cd(flat)))
        out, off = [], 0
        for h, w in shapes:
            out.append(flat[off:off + h * w].reshape(h, w))
            off += h * w
        return out

    def objects_device(self):
        """Every tile's result object left in device memory: a list of
        (device address, (R_h, R_w)) per tile, complex128 row-major, valid
        until the next run() or objects call on this batch."""
        ptrs = (ct.c_void_p * len(self.sessions))()
        check(lib().fpmhip_batch_objects_device(self.handle, ptrs))
        return [(int(p or 0), tuple(s.problem.model.reconstruction_shape)) for p, s in zip(ptrs, self.sessions)]

    def timing(self):
        mean, mx, n = ct.c_double(), ct.c_double(), ct.c_uint64()
        check(lib().fpmhip_batch_timing(self.handle, ct.byref(mean), ct.byref(mx), ct.byref(n)))
        return mean.value, mx.value, n.value

    def schedule(self) -> dict:
        """The sweep launch's schedule: group_members (workgroups per 256^2
        tile: 0 automatic, 1 none, 8..32 that many), largest_first, radix
        (first-stage FFT radix for 32, 64, 128,
        256 pixel tiles; 0 = default), sweeps_per_launch (0 = default)."""
        sch = BatchSchedule()
        check(lib().fpmhip_batch_get_schedule(self.handle, ct.byref(sch)))
        return {"group_members": sch.group_members, "largest_first": sch.largest_first, "radix": list(sch.radix),
                "sweeps_per_launch": sch.sweeps_per_launch}

    def set_schedule(self, group_members=None, largest_first=None, radix=None, sweeps_per_launch=None):
        cur = self.schedule()
        sch = BatchSchedule()
        sch.group_members = cur["group_members"] if group_members is None else group_members
        sch.largest_first = cur["largest_first"] if largest_first is None else int(largest_first)
        for i, r in enumerate(cur["radix"] if radix is None else radix):
            sch.radix[i] = r
        sch.sweeps_per_launch = cur["sweeps_per_launch"] if sweeps_per_launch is None else sweeps_per_launch
        check(lib().fpmhip_batch_set_schedule(self.handle, ct.byref(sch)))

    def tune(self, trial_sweeps=2) -> dict:
        """Autotune the schedule on this batch's tiles (fpmhip_batch_tune)."""
        rep = BatchTuning()
        check(lib().fpmhip_batch_tune(self.handle, ct.c_uint32(trial_sweeps), ct.byref(rep)))
        return {"candidates": rep.candidates, "first_ms": rep.first_ms, "best_ms": rep.best_ms,
                "schedule": self.schedule()}


class BatchSchedule(ct.Structure):
    _fields_ = [("group_members", ct.c_uint32), ("largest_first", ct.c_uint32), ("radix", ct.c_uint32 * 4),
                ("sweeps_per_launch", ct.c_uint32)]


class BatchTuning(ct.Structure):
    _fields_ = [("candidates", ct.c_uint32), ("first_ms", ct.c_double), ("best_ms", ct.c_double)]


class ServerConfig(ct.Structure):
    _fields_ = [("max_tiles", ct.c_uint64), ("hbm_budget_bytes", ct.c_uint64)]


class ServerStatistics(ct.Structure):
    _fields_ = [("submitted", ct.c_uint64), ("admitted", ct.c_uint64), ("completed", ct.c_uint64),
                ("failed", ct.c_uint64), ("ticks", ct.c_uint64), ("resident_tiles", ct.c_uint64),
                ("resident_bytes", ct.c_uint64), ("busy_seconds", ct.c_double),
                ("pool_reuses", ct.c_uint64), ("pool_allocations", ct.c_uint64),
                ("peak_tick_workgroups", ct.c_uint64)]


class Server:
    """Continuous-batching server (fpmhip_server): jobs are admitted by HBM
    footprint and tile slots and advanced one fused sweep per tick."""

    def __init__(self, max_tiles=256, hbm_budget_bytes=64 << 30, context=None):
        self.ctx = context or default_context()
        cfg = ServerConfig(max_tiles, hbm_budget_bytes)
        self.handle = ct.c_void_p()
        check(lib().fpmhip_server_create(self.ctx.handle, ct.byref(cfg), ct.byref(self.handle)))
        self._shapes = {}
        _register("server", self)

    def close(self):
        """Stops the worker thread (joined) and frees the server's device state."""
        if self.handle:
            lib().fpmhip_server_destroy(self.handle)
            self.handle = ct.c_void_p()

    def __del__(self):
        if not _closed:
            try:
                self.close()
            except Exception:
                pass

    def submit(self, problem: ReconstructionProblem, algorithm) -> int:
        """Queue one tile job; returns its id."""
        m = problem.model
        d = m.desc()
        meas = problem.measurements
        alg = algorithm if isinstance(algorithm, abi.Algorithm) else algorithm.to_c()
        if not isinstance(algorithm, abi.Algorithm):
            alg.iterations = int(algorithm.iterations)
        jid = ct.c_uint64(0)
        check(lib().fpmhip_server_submit(
            self.handle, ct.byref(d), None if meas.counts is None else abi.ptr(meas.counts, ct.c_uint16),
            None if meas.frames is None else _cd(meas.frames), None if meas.weights is None else _cd(meas.weights),
            ct.byref(alg), ct.byref(jid)))
        self._shapes[jid.value] = (tuple(m.reconstruction_shape), int(alg.iterations))
        return jid.value

    def poll(self, job_id) -> int:
        st = ct.c_int32(0)
        check(lib().fpmhip_server_poll(self.handle, ct.c_uint64(job_id), ct.byref(st)))
        return st.value

    def wait(self, job_id, timeout_s=0.0):
        check(lib().fpmhip_server_wait(self.handle, ct.c_uint64(job_id), ct.c_double(timeout_s)))

    def result(self, job_id):
        """(object, spectrum, trace, latency_s, queue_s) of a finished job."""
        R, iters = self._shapes.pop(job_id)
        obj = np.zeros(R, np.complex128)
        spec = np.zeros(R, np.complex128)
        trace = np.zeros(iters, np.float64)
        lat, q = ct.c_double(), ct.c_double()
        check(lib().fpmhip_server_result(self.handle, ct.c_uint64(job_id), _cd(obj), _cd(spec), _cd(trace),
                                         ct.byref(lat), ct.byref(q)))
        return obj, spec, trace, lat.value, q.value

    def stats(self) -> dict:
        st = ServerStatistics()
        check(lib().fpmhip_server_stats(self.handle, ct.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStatistics._fields_}
