"""Resumable reconstruction checkpoints in fpm-rs's JSON format version 2.

`ReconstructionCheckpoint` mirrors src/reconstruction/checkpoint.rs: the
same fields, validation rules, error kinds and messages, and the same JSON
layout, so that a checkpoint written here resumes in fpm-rs and one written
by fpm-rs resumes on the device:

- the top-level object of `Serialize for ReconstructionCheckpoint`
  (checkpoint.rs:406-440): format_version, completed_iterations,
  object_spectrum, pupil, illumination_corrections, frame_gains, background,
  physical_illumination_calibration, calibrated_model, algorithm_auxiliary,
  trace;
- 2-D arrays as `Array2Data` (src/array_serde.rs:8-14): height, width and
  the row-major data;
- complex values as num-complex 0.4's serde form, a two-element sequence
  [re, im];
- the pupil as values and support arrays (src/model/pupil.rs:185-226);
- the mPIE auxiliary state as the externally tagged enum variant
  {"Mpie": {...}} with the fields of MpieAuxiliaryState
  (src/reconstruction/state.rs:34-53);
- the trace as iteration records and algorithm metrics
  (src/reconstruction/trace.rs:6-41).

The physical illumination calibration, calibrated model and illumination
corrections belong to solvers outside this package; a checkpoint carrying
them loads and validates, but resuming it here is Unsupported.
"""
from __future__ import annotations

import json
import math
import os
import tempfile
from dataclasses import dataclass, field

import numpy as np

from . import abi

FORMAT_VERSION = 2  # CHECKPOINT_FORMAT_VERSION, checkpoint.rs:26


def _error(kind: str, message: str, name: str = ""):
    from .runtime import FpmError

    return FpmError(abi.STATUS_CODES[kind], message, name)


def _serde(message: str):
    from .runtime import FpmError

    class SerializationError(FpmError):
        """fpm_rs::Error::Serialization (src/error.rs:173): the JSON text is
        not a checkpoint of this format."""

        def __init__(self, msg):
            RuntimeError.__init__(self, f"Serialization: {msg}")
            self.code, self.kind, self.message, self.name = None, "Serialization", msg, ""

    return SerializationError(message)


def _reject_constant(token):
    raise _serde(f"invalid number {token}")


def _uint(v, what: str) -> int:
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        raise _serde(f"{what}: expected a non-negative integer, got {v!r}")
    return v


def _float(v, what: str) -> float:
    if isinstance(v, bool) or not isinstance(v, (int, float)):
        raise _serde(f"{what}: expected a number, got {v!r}")
    return float(v)


def _fields(obj, required, what: str, optional=(), deny_unknown=True) -> dict:
    if not isinstance(obj, dict):
        raise _serde(f"{what}: expected an object")
    if deny_unknown:
        unknown = set(obj) - set(required) - set(optional)
        if unknown:
            raise _serde(f"{what}: unknown field `{sorted(unknown)[0]}`")
    missing = [k for k in required if k not in obj]
    if missing:
        raise _serde(f"{what}: missing field `{missing[0]}`")
    return obj


def _float_list(v, what: str):
    if v is None:
        return None
    if not isinstance(v, list):
        raise _serde(f"{what}: expected a sequence")
    return [_float(x, what) for x in v]


def _complex_list(a: np.ndarray) -> list:
    flat = np.ascontiguousarray(a, dtype=np.complex128).ravel()
    return [[float(v.real), float(v.imag)] for v in flat]


def _complex_array(data, what: str) -> np.ndarray:
    """num-complex's serde form: each value a [re, im] sequence."""
    if not isinstance(data, list):
        raise _serde(f"{what}: expected a sequence")
    out = np.empty(len(data), np.complex128)
    for k, v in enumerate(data):
        if not isinstance(v, list) or len(v) != 2:
            raise _serde(f"{what}: a complex value is a [re, im] pair")
        out[k] = complex(_float(v[0], what), _float(v[1], what))
    return out


def _array2(obj, complex_values: bool, what: str) -> np.ndarray:
    """Array2Data (array_serde.rs:8-39): height, width, row-major data."""
    _fields(obj, ("height", "width", "data"), what)
    h, w = _uint(obj["height"], what), _uint(obj["width"], what)
    if complex_values:
        data = _complex_array(obj["data"], what)
    else:
        if not isinstance(obj["data"], list):
            raise _serde(f"{what}: expected a sequence")
        vals = [_uint(x, what) for x in obj["data"]]
        if any(x > 255 for x in vals):
            raise _serde(f"{what}: u8 value out of range")
        data = np.asarray(vals, np.uint8)
    if data.size != h * w:  # Error::LengthMismatch, surfaced through serde
        raise _serde(f"{what}: length mismatch: {data.size} values for shape ({h}, {w})")
    return data.reshape(h, w)


@dataclass
class MpieAuxiliary:
    """MpieAuxiliaryState (state.rs:34-53)."""

    velocity: np.ndarray  # flat, centred row-major spectrum order
    anchor: np.ndarray
    effective_frames_since_momentum: int
    object_step: float
    stability: float
    epsilon: float
    momentum_interval: int
    momentum_friction: float
    momentum_feedback: float

    def to_json(self) -> dict:
        return {"Mpie": {
            "velocity": _complex_list(self.velocity),
            "anchor": _complex_list(self.anchor),
            "effective_frames_since_momentum": int(self.effective_frames_since_momentum),
            "object_step": float(self.object_step),
            "stability": float(self.stability),
            "epsilon": float(self.epsilon),
            "momentum_interval": int(self.momentum_interval),
            "momentum_friction": float(self.momentum_friction),
            "momentum_feedback": float(self.momentum_feedback),
        }}


@dataclass
class ReconstructionCheckpoint:
    """src/reconstruction/checkpoint.rs:38-52."""

    completed_iterations: int
    object_spectrum: np.ndarray  # (R_h, R_w) complex128, centred
    pupil_values: np.ndarray  # (n_h, n_w) complex128
    pupil_support: np.ndarray  # (n_h, n_w) uint8
    frame_gains: list | None = None
    background: list | None = None
    algorithm_auxiliary: object = None  # MpieAuxiliary, or another solver's state kept as JSON
    trace: list = field(default_factory=list)  # [(iteration, objective, elapsed_seconds)]
    algorithm_metrics: list = field(default_factory=list)  # [(iteration, namespace, metric, value)]
    illumination_corrections: list | None = None
    physical_illumination_calibration: object = None
    calibrated_model: object = None
    format_version: int = FORMAT_VERSION

    # ---- ReconstructionCheckpoint::capture (checkpoint.rs:56-74) -----------
    @classmethod
    def capture(cls, completed_iterations: int, session, trace_records, mpie_parameters=None,
                algorithm_metrics=(), carried_auxiliary=None):
        """Clones a device session's resumable state and the trace of its
        `completed_iterations` complete passes. `mpie_parameters` (a dict of
        the MpieAuxiliary parameter fields) adds the session's momentum
        state; otherwise `carried_auxiliary` (another solver's state restored
        from a checkpoint) is kept as it was."""
        has_gains, bg_len = session.calibration_present()
        gains, bg = session.calibration()
        aux = carried_auxiliary
        if mpie_parameters is not None:
            v, a, c = session.momentum()
            aux = MpieAuxiliary(v.ravel().copy(), a.ravel().copy(), int(c), **mpie_parameters)
        return cls(int(completed_iterations), session.spectrum(), session.pupil(),
                   np.asarray(session.problem.model.support, np.uint8).copy(),
                   [float(x) for x in gains] if has_gains else None,
                   [float(x) for x in np.ravel(bg)] if bg_len else None,
                   aux, [tuple(r) for r in trace_records], [tuple(m) for m in algorithm_metrics])

    # ---- serde --------------------------------------------------------------
    def to_json(self) -> dict:
        h, w = self.object_spectrum.shape
        ph, pw = self.pupil_values.shape
        aux = self.algorithm_auxiliary
        return {
            "format_version": int(self.format_version),
            "completed_iterations": int(self.completed_iterations),
            "object_spectrum": {"height": h, "width": w, "data": _complex_list(self.object_spectrum)},
            "pupil": {"values": {"height": ph, "width": pw, "data": _complex_list(self.pupil_values)},
                      "support": {"height": ph, "width": pw,
                                  "data": [int(x) for x in np.asarray(self.pupil_support).ravel()]}},
            "illumination_corrections": None if self.illumination_corrections is None
            else [[float(r), float(c)] for r, c in self.illumination_corrections],
            "frame_gains": None if self.frame_gains is None else [float(x) for x in self.frame_gains],
            "background": None if self.background is None else [float(x) for x in self.background],
            "physical_illumination_calibration": self.physical_illumination_calibration,
            "calibrated_model": self.calibrated_model,
            "algorithm_auxiliary": aux.to_json() if isinstance(aux, MpieAuxiliary) else aux,
            "trace": {
                "iterations": [{"iteration": int(i), "objective": float(o), "elapsed_seconds": float(e)}
                               for i, o, e in self.trace],
                "algorithm_metrics": [{"iteration": int(i), "namespace": n, "metric": m, "value": float(v)}
                                      for i, n, m, v in self.algorithm_metrics],
            },
        }

    @classmethod
    def from_json(cls, obj: dict) -> "ReconstructionCheckpoint":
        """Deserialize for ReconstructionCheckpoint (checkpoint.rs:467-486)
        followed by validate(): unknown fields are rejected and
        algorithm_auxiliary may be absent; structural failures are
        Serialization errors, as serde_json reports them."""
        _fields(obj, ("format_version", "completed_iterations", "object_spectrum", "pupil",
                      "illumination_corrections", "frame_gains", "background", "physical_illumination_calibration",
                      "calibrated_model", "trace"), "checkpoint", optional=("algorithm_auxiliary",))
        pupil = _fields(obj["pupil"], ("values", "support"), "pupil")
        values = _array2(pupil["values"], True, "pupil values")
        support = _array2(pupil["support"], False, "pupil support")
        # Pupil::new (src/model/pupil.rs:32-62), its errors surfaced through serde
        if values.shape != support.shape:
            raise _serde(f"pupil support shape {support.shape} does not match value shape {values.shape}")
        if 0 in values.shape:
            raise _serde(f"pupil dimensions must be non-zero, got {values.shape}")
        if not np.isfinite(values).all():
            raise _serde("pupil contains non-finite complex values")
        if (support > 1).any():
            raise _serde("pupil support values must be exactly zero or one")
        aux = obj.get("algorithm_auxiliary")
        if isinstance(aux, dict) and len(aux) == 1 and "Mpie" in aux:
            m = _fields(aux["Mpie"], ("velocity", "anchor", "effective_frames_since_momentum", "object_step",
                                      "stability", "epsilon", "momentum_interval", "momentum_friction",
                                      "momentum_feedback"), "Mpie", deny_unknown=False)
            aux = MpieAuxiliary(_complex_array(m["velocity"], "velocity"), _complex_array(m["anchor"], "anchor"),
                                _uint(m["effective_frames_since_momentum"], "effective_frames_since_momentum"),
                                _float(m["object_step"], "object_step"), _float(m["stability"], "stability"),
                                _float(m["epsilon"], "epsilon"), _uint(m["momentum_interval"], "momentum_interval"),
                                _float(m["momentum_friction"], "momentum_friction"),
                                _float(m["momentum_feedback"], "momentum_feedback"))
        elif aux is not None and not (isinstance(aux, dict) and len(aux) == 1 and next(iter(aux)) in
                                      ("Admm", "AdaptiveAlternatingProjection")):
            raise _serde("algorithm_auxiliary: unknown variant")
        tr = _fields(obj["trace"], ("iterations", "algorithm_metrics"), "trace")
        if not isinstance(tr["iterations"], list) or not isinstance(tr["algorithm_metrics"], list):
            raise _serde("trace: expected sequences")
        records = []
        for r in tr["iterations"]:
            _fields(r, ("iteration", "objective", "elapsed_seconds"), "iteration record")
            records.append((_uint(r["iteration"], "iteration"), _float(r["objective"], "objective"),
                            _float(r["elapsed_seconds"], "elapsed_seconds")))
        metrics = []
        for r in tr["algorithm_metrics"]:
            _fields(r, ("iteration", "namespace", "metric", "value"), "metric record")
            if not isinstance(r["namespace"], str) or not isinstance(r["metric"], str):
                raise _serde("metric record: namespace and metric are strings")
            metrics.append((_uint(r["iteration"], "iteration"), r["namespace"], r["metric"],
                            _float(r["value"], "value")))
        corr = obj["illumination_corrections"]
        if corr is not None:
            if not isinstance(corr, list) or any(not isinstance(p, list) or len(p) != 2 for p in corr):
                raise _serde("illumination_corrections: expected (row, column) pairs")
            corr = [(_float(p[0], "correction"), _float(p[1], "correction")) for p in corr]
        ck = cls(
            completed_iterations=_uint(obj["completed_iterations"], "completed_iterations"),
            object_spectrum=_array2(obj["object_spectrum"], True, "object_spectrum"),
            pupil_values=values,
            pupil_support=support,
            frame_gains=_float_list(obj["frame_gains"], "frame_gains"),
            background=_float_list(obj["background"], "background"),
            algorithm_auxiliary=aux,
            trace=records,
            algorithm_metrics=metrics,
            illumination_corrections=corr,
            physical_illumination_calibration=obj["physical_illumination_calibration"],
            calibrated_model=obj["calibrated_model"],
            format_version=_uint(obj["format_version"], "format_version"),
        )
        ck.validate()
        return ck

    # ---- files (checkpoint.rs:137-156) ------------------------------------
    def save(self, path) -> None:
        """Validates and writes compact JSON, atomically (temporary file in
        the same directory, then rename)."""
        self.validate()
        path = os.fspath(path)
        directory = os.path.dirname(os.path.abspath(path))
        fd, tmp = tempfile.mkstemp(prefix=".checkpoint-", suffix=".json", dir=directory)
        try:
            with os.fdopen(fd, "w") as f:
                json.dump(self.to_json(), f, separators=(",", ":"), allow_nan=False)
            os.replace(tmp, path)
        except BaseException:
            if os.path.exists(tmp):
                os.unlink(tmp)
            raise

    @staticmethod
    def load(path) -> "ReconstructionCheckpoint":
        with open(path) as f:
            try:
                obj = json.load(f, parse_constant=_reject_constant)  # serde_json has no NaN/Infinity literals
            except json.JSONDecodeError as e:
                raise _serde(str(e)) from None
        return ReconstructionCheckpoint.from_json(obj)

    @staticmethod
    def load_for_problem(path, problem) -> "ReconstructionCheckpoint":
        ck = ReconstructionCheckpoint.load(path)
        ck.validate_for_problem(problem)
        return ck

    # ---- validation (checkpoint.rs:158-404) -------------------------------
    def validate(self) -> None:
        if self.format_version != FORMAT_VERSION:
            raise _error("InvalidParameter", f"expected {FORMAT_VERSION}, got {self.format_version}",
                         "checkpoint format_version")
        if self.pupil_support.shape != self.pupil_values.shape:
            raise _error("InvalidShape", "checkpoint pupil support and values have different lengths")
        if not (np.isfinite(self.object_spectrum).all() and np.isfinite(self.pupil_values).all()):
            raise _error("InvalidModel", "checkpoint contains non-finite complex values")
        if self.illumination_corrections is not None and any(
                not (math.isfinite(r) and math.isfinite(c)) for r, c in self.illumination_corrections):
            raise _error("InvalidModel", "checkpoint illumination corrections contain non-finite values")
        if self.frame_gains is not None and any(not math.isfinite(g) or g <= 0.0 for g in self.frame_gains):
            raise _error("InvalidModel", "checkpoint frame gains must be finite and positive")
        if self.background is not None and any(not math.isfinite(b) for b in self.background):
            raise _error("InvalidModel", "checkpoint background contains non-finite values")
        if (self.physical_illumination_calibration is None) != (self.calibrated_model is None):
            raise _error("InvalidModel",
                         "checkpoint physical calibration and calibrated model must be present together")
        aux = self.algorithm_auxiliary
        if isinstance(aux, MpieAuxiliary) and (
                aux.velocity.size != aux.anchor.size
                or not (np.isfinite(aux.velocity).all() and np.isfinite(aux.anchor).all())
                or not math.isfinite(aux.object_step) or aux.object_step <= 0.0
                or not math.isfinite(aux.stability) or not 0.0 <= aux.stability <= 1.0
                or not math.isfinite(aux.epsilon) or aux.epsilon <= 0.0
                or aux.momentum_interval == 0
                or aux.effective_frames_since_momentum >= aux.momentum_interval
                or not math.isfinite(aux.momentum_friction) or not 0.0 <= aux.momentum_friction < 1.0
                or not math.isfinite(aux.momentum_feedback) or not 0.0 <= aux.momentum_feedback <= 1.0):
            raise _error("InvalidModel", "checkpoint algorithm auxiliary state is inconsistent or non-finite")
        recs = self.trace
        if (len(recs) != self.completed_iterations
                or any(i != k + 1 or not math.isfinite(o) or not math.isfinite(e) or e < 0.0
                       for k, (i, o, e) in enumerate(recs))
                or any(b[2] < a[2] for a, b in zip(recs, recs[1:]))):
            raise _error("InvalidModel", "checkpoint trace is incomplete, non-finite, or non-monotonic")
        if any(i == 0 or i > self.completed_iterations or not n or not m or not math.isfinite(v)
               for i, n, m, v in self.algorithm_metrics):
            raise _error("InvalidModel", "checkpoint algorithm metrics are invalid")

    def validate_for_problem(self, problem) -> None:
        self.validate()
        model = problem.model
        if tuple(self.object_spectrum.shape) != tuple(model.reconstruction_shape):
            raise _error("InvalidShape", f"checkpoint spectrum shape {tuple(self.object_spectrum.shape)} differs "
                                         f"from reconstruction shape {tuple(model.reconstruction_shape)}")
        if tuple(self.pupil_values.shape) != tuple(model.image_shape):
            raise _error("InvalidShape", "checkpoint pupil does not match the model image shape")
        if not np.array_equal(np.asarray(self.pupil_support, np.uint8), np.asarray(model.support, np.uint8)):
            raise _error("InvalidModel", "checkpoint pupil support differs from the reconstruction problem")
        if self.illumination_corrections is not None and len(self.illumination_corrections) != model.source_count:
            raise _error("InvalidModel", "checkpoint illumination correction count does not match the model")
        if self.frame_gains is not None and len(self.frame_gains) != model.frame_count:
            raise _error("InvalidModel", "checkpoint frame gain count does not match the model")
        N = int(np.prod(model.image_shape))
        if self.background is not None and len(self.background) not in (N, N * model.frame_count):
            raise _error("InvalidModel", "checkpoint background dimensions do not match the model")
        M = int(np.prod(model.reconstruction_shape))
        aux = self.algorithm_auxiliary
        if isinstance(aux, MpieAuxiliary) and (aux.velocity.size != M or aux.anchor.size != M):
            raise _error("InvalidModel", "checkpoint algorithm auxiliary dimensions do not match the model")


class CheckpointEvery:
    """callbacks/checkpoint.rs: writes `checkpoint_{iteration:05}.json` into
    `directory` after every `every`-th complete iteration (0 counts as 1),
    creating the directory when the run starts."""

    def __init__(self, every: int, directory):
        self.every = max(int(every), 1)
        self.directory = os.fspath(directory)

    def on_start(self) -> None:
        os.makedirs(self.directory, exist_ok=True)

    def due(self, iteration: int) -> bool:
        return iteration % self.every == 0

    def path(self, iteration: int) -> str:
        return os.path.join(self.directory, f"checkpoint_{iteration:05d}.json")

    def write(self, checkpoint: ReconstructionCheckpoint) -> None:
        checkpoint.save(self.path(checkpoint.completed_iterations))
