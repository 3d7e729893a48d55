"""fpm_rs::simulation on the device: CameraModel (src/simulation/camera.rs),
IlluminationAcquisitionErrors (src/simulation/illumination_acquisition_errors.rs)
and Simulator::simulate (src/simulation/simulator.rs:107-341).

The host side mirrors the reference's builders and checks: the true model's
acquisition errors (source permutation, frame-gain variation) are edits of
the compiled model made here before it is uploaded, and the camera's known
linear response is compiled into the reconstruction model here. The object
spectrum, the forward stack, missing frames and the detector pass run in
libfpmhip (fpmhip_simulate / fpmhip_camera_measure).

Random streams: camera noise is drawn on the device from a counter-based
Philox4x64-10 stream keyed by the seed, per (frame, pixel); the frame-gain
variation is drawn on the host from numpy's Philox generator keyed by
(seed, 1). A seed reproduces a simulation exactly, but not the reference's
StdRng draws: the distributions are the reference's, the samples are not.
"""
import copy
import ctypes as ct
from dataclasses import dataclass, field, replace
from typing import List, Optional

import numpy as np

from . import abi
from .model import ImagePlaneModel, MeasurementStack
from .runtime import DeviceModel, FpmError, _cd, check, default_context, lib

_GAIN_STREAM = 1


def _invalid(message, name):
    return FpmError(abi.STATUS_CODES["InvalidParameter"], message, name)


@dataclass
class CameraModel:
    """Detector pipeline from optical intensity to digital counts
    (camera.rs:7-53; defaults are CameraModel::default)."""

    photons_per_pixel: float = 1000.0
    gain_counts_per_electron: float = 1.0
    offset_counts: float = 0.0
    read_noise_electrons: float = 0.0
    dark_current_electrons: float = 0.0
    shot_noise: bool = False
    pixel_sensitivity: Optional[np.ndarray] = None
    bit_depth: Optional[int] = 16
    saturation_counts: Optional[float] = None
    quantize: bool = True
    bad_pixels: List[int] = field(default_factory=list)
    bad_pixel_value_counts: Optional[float] = None

    @classmethod
    def new(cls):
        return cls()

    @classmethod
    def ideal(cls):
        """Unit response without noise, clipping or quantization (camera.rs:62-77)."""
        return cls(photons_per_pixel=1.0, bit_depth=None, quantize=False)

    # builders (camera.rs:79-141): each returns a modified copy
    def with_photons_per_pixel(self, value):
        return replace(self, photons_per_pixel=float(value))

    def gain(self, counts_per_electron):
        return replace(self, gain_counts_per_electron=float(counts_per_electron))

    def with_offset_counts(self, value):
        return replace(self, offset_counts=float(value))

    def with_read_noise_electrons(self, value):
        return replace(self, read_noise_electrons=float(value))

    def with_dark_current_electrons(self, value):
        return replace(self, dark_current_electrons=float(value))

    def with_shot_noise(self, enabled):
        return replace(self, shot_noise=bool(enabled))

    def with_pixel_sensitivity(self, values):
        return replace(self, pixel_sensitivity=np.ascontiguousarray(values, dtype=np.float64))

    def with_bit_depth(self, bits):
        return replace(self, bit_depth=int(bits))

    def saturation(self, counts):
        return replace(self, saturation_counts=float(counts))

    def with_quantize(self, enabled):
        return replace(self, quantize=bool(enabled))

    def with_bad_pixels(self, indices, value_counts):
        return replace(self, bad_pixels=list(indices), bad_pixel_value_counts=float(value_counts))

    def to_c(self):
        """(abi.Camera, arrays it points into)."""
        c = abi.Camera()
        c.photons_per_pixel = self.photons_per_pixel
        c.gain_counts_per_electron = self.gain_counts_per_electron
        c.offset_counts = self.offset_counts
        c.read_noise_electrons = self.read_noise_electrons
        c.dark_current_electrons = self.dark_current_electrons
        c.shot_noise = int(bool(self.shot_noise))
        c.quantize = int(bool(self.quantize))
        c.has_bit_depth = int(self.bit_depth is not None)
        c.bit_depth = int(self.bit_depth) if self.bit_depth is not None else 0
        c.has_saturation = int(self.saturation_counts is not None)
        c.saturation_counts = self.saturation_counts if self.saturation_counts is not None else 0.0
        c.has_bad_pixel_value = int(self.bad_pixel_value_counts is not None)
        c.bad_pixel_value_counts = self.bad_pixel_value_counts if self.bad_pixel_value_counts is not None else 0.0
        keep = []
        if self.pixel_sensitivity is not None:
            s = np.ascontiguousarray(self.pixel_sensitivity, dtype=np.float64).ravel()
            keep.append(s)
            c.pixel_sensitivity = _cd(s)
        if any(int(i) < 0 for i in self.bad_pixels):
            raise _invalid("indices must be non-negative", "bad_pixels")
        bp = np.ascontiguousarray(np.asarray(self.bad_pixels, dtype=np.uint64))
        keep.append(bp)
        c.bad_pixels = bp.ctypes.data_as(abi.P_u64) if bp.size else None
        c.bad_pixel_count = bp.size
        return c, keep

    def validate(self):
        """CameraModel::validate (camera.rs:145-223), frame-independent."""
        c, keep = self.to_c()
        check(lib().fpmhip_camera_validate(ct.byref(c), ct.c_uint64(0)))
        if self.pixel_sensitivity is not None:
            # validate() checks the values whatever the frame length
            s = np.asarray(self.pixel_sensitivity, dtype=np.float64)
            if not np.all(np.isfinite(s)) or np.any(s < 0.0):
                raise _invalid("values must be finite and non-negative", "pixel_sensitivity")

    def validate_for_frame(self, frame_len):
        """CameraModel::validate_for_frame (camera.rs:226-245)."""
        self.validate()
        if self.pixel_sensitivity is not None and np.asarray(self.pixel_sensitivity).size != frame_len:
            raise _invalid(f"must contain {frame_len} values", "pixel_sensitivity")
        c, keep = self.to_c()
        check(lib().fpmhip_camera_validate(ct.byref(c), ct.c_uint64(frame_len)))

    def maximum_count(self):
        """CameraModel::maximum_count (camera.rs:320-327)."""
        digital = float(2 ** self.bit_depth - 1) if self.bit_depth is not None else float("inf")
        sat = self.saturation_counts if self.saturation_counts is not None else float("inf")
        return min(sat, digital)

    def measure_frame(self, intensity, seed, context=None):
        """CameraModel::measure_frame (camera.rs:248-287) of one frame, in
        place on a float64 array (also returned); the noise stream is the
        seed's frame 0."""
        return self.measure_frames(intensity, seed, frame_len=np.asarray(intensity).size, context=context)

    def measure_frames(self, frames, seed, frame_len=None, context=None):
        """measure_frame over consecutive frames of `frame_len` pixels (the
        last axes of a (F, h, w) stack by default), in place."""
        if not (isinstance(frames, np.ndarray) and frames.dtype == np.float64 and frames.flags["C_CONTIGUOUS"]):
            raise TypeError("frames must be a C-contiguous float64 numpy array (measured in place)")
        if frame_len is None:
            frame_len = int(np.prod(frames.shape[1:])) if frames.ndim > 1 else frames.size
        if frame_len == 0 or frames.size % frame_len:
            raise FpmError(abi.STATUS_CODES["InvalidShape"], "frames are not a whole number of camera frames")
        self.validate_for_frame(frame_len)
        c, keep = self.to_c()
        ctx = context or default_context()
        check(lib().fpmhip_camera_measure(ctx.handle, ct.byref(c), _cd(frames), ct.c_uint64(frames.size // frame_len),
                                          ct.c_uint64(frame_len), ct.c_uint64(seed)))
        return frames

    def compile_reconstruction_model(self, model: ImagePlaneModel) -> ImagePlaneModel:
        """CameraModel::compile_reconstruction_model (camera.rs:289-318): a
        copy of `model` with frame gains x photons x gain and background
        scaled and offset by the camera's additive counts."""
        self.validate_for_frame(model.image_len)
        out = copy.deepcopy(model)
        scale = self.photons_per_pixel * self.gain_counts_per_electron
        if model.frame_gain is not None:
            out.frame_gain = np.asarray(model.frame_gain, dtype=np.float64) * scale
        else:
            out.frame_gain = np.full(model.frame_count, scale)
        additive = self.dark_current_electrons * self.gain_counts_per_electron + self.offset_counts
        if model.background is not None:
            out.background = np.asarray(model.background, dtype=np.float64) * scale + additive
        elif additive != 0.0:
            out.background = np.full(model.image_len, additive)
        return out


@dataclass
class IlluminationAcquisitionErrors:
    """Non-geometric illumination errors (illumination_acquisition_errors.rs:3-14)."""

    frame_gain_relative_std: float = 0.0
    missing_frames: List[int] = field(default_factory=list)
    source_permutation: Optional[List[int]] = None

    @classmethod
    def new(cls):
        return cls()

    def with_frame_gain_relative_std(self, value):
        return replace(self, frame_gain_relative_std=float(value))

    def with_missing_frames(self, indices):
        return replace(self, missing_frames=list(indices))

    def with_source_permutation(self, permutation):
        return replace(self, source_permutation=list(permutation))


@dataclass
class SimulationParameters:
    """result.rs:29-40."""

    ideal: bool
    frame_count: int
    image_shape: tuple
    missing_frames: List[int]


@dataclass
class SimulationResult:
    """result.rs:8-27."""

    measurements: MeasurementStack
    ground_truth_object: np.ndarray
    true_model: ImagePlaneModel
    reconstruction_model: ImagePlaneModel
    camera: Optional[CameraModel]
    illumination_acquisition_errors: Optional[IlluminationAcquisitionErrors]
    parameters: SimulationParameters
    random_seed: int


def apply_illumination_acquisition_errors(model: ImagePlaneModel, errors: IlluminationAcquisitionErrors, seed: int):
    """apply_illumination_acquisition_errors (simulator.rs:228-341) on a copy
    of the true model: checks, the source permutation of k-vectors, crops,
    offsets and (without multiplexing) gains and per-frame background, then
    gains x max(Normal(1, std), 0.01) per frame."""
    std = errors.frame_gain_relative_std
    if not np.isfinite(std) or std < 0.0:
        raise _invalid("must be finite and non-negative", "frame_gain_relative_std")
    F, S = model.frame_count, model.source_count
    missing = sorted(int(f) for f in errors.missing_frames)
    if any(f < 0 or f >= F for f in missing) or any(a == b for a, b in zip(missing, missing[1:])):
        raise _invalid(f"must contain unique frame indices below {F}", "missing_frames")
    out = copy.deepcopy(model)
    perm = errors.source_permutation
    if perm is not None:
        perm = [int(i) for i in perm]
        if len(perm) != S or sorted(perm) != list(range(S)):
            raise _invalid(f"must be a permutation of 0..{S}", "source_permutation")
        idx = np.asarray(perm)
        if out.k_vectors is not None:
            out.k_vectors = np.ascontiguousarray(np.asarray(out.k_vectors).reshape(-1, 2)[idx])
        out.crop_start = np.ascontiguousarray(np.asarray(out.crop_start).reshape(-1, 2)[idx])
        out.offset = np.ascontiguousarray(np.asarray(out.offset).reshape(-1, 2)[idx])
        multiplexed = out.multiplexing is not None
        if not multiplexed and out.frame_gain is not None:
            out.frame_gain = np.ascontiguousarray(np.asarray(out.frame_gain)[idx])
        N = out.image_len
        if not multiplexed and out.background is not None and np.asarray(out.background).size == F * N:
            out.background = np.ascontiguousarray(np.asarray(out.background).reshape(F, N)[idx].ravel())
    if std > 0.0:
        rng = np.random.Generator(np.random.Philox(key=[int(seed) & ((1 << 64) - 1), _GAIN_STREAM]))
        gains = np.ones(F) if out.frame_gain is None else np.array(out.frame_gain, dtype=np.float64)
        draws = rng.normal(1.0, std, F)
        out.frame_gain = gains * np.maximum(draws, 0.01)
    return out


class Simulator:
    """Builder of deterministic synthetic acquisitions (simulator.rs:20-105)."""

    def __init__(self, true_model: ImagePlaneModel, ideal=False):
        self.true_model = true_model
        self._reconstruction_model = None
        self._object = None
        self._camera = None
        self._errors = None
        self._seed = 0
        self._ideal = ideal

    @classmethod
    def new(cls, true_model):
        return cls(true_model)

    @classmethod
    def ideal(cls, model):
        """Optical intensities without camera or acquisition errors."""
        return cls(model, ideal=True)

    def object(self, field_):
        """The high-resolution complex ground-truth field (R_h, R_w)."""
        self._object = np.ascontiguousarray(field_, dtype=np.complex128)
        return self

    def camera(self, camera: CameraModel):
        self._camera = camera
        self._ideal = False
        return self

    def illumination_acquisition_errors(self, errors: IlluminationAcquisitionErrors):
        self._errors = errors
        self._ideal = False
        return self

    def reconstruction_model(self, model: ImagePlaneModel):
        self._reconstruction_model = model
        return self

    def seed(self, seed: int):
        self._seed = int(seed)
        return self

    def simulate(self, context=None) -> SimulationResult:
        """Simulator::simulate (simulator.rs:107-207)."""
        tm = self.true_model
        if self._object is None:
            raise _invalid("a ground-truth object is required", "object")
        if tuple(self._object.shape) != tuple(tm.reconstruction_shape):
            raise FpmError(abi.STATUS_CODES["InvalidShape"],
                           f"object shape {tuple(self._object.shape)} differs from reconstruction shape "
                           f"{tuple(tm.reconstruction_shape)}")
        rm = self._reconstruction_model if self._reconstruction_model is not None else copy.deepcopy(tm)
        if (tuple(rm.image_shape) != tuple(tm.image_shape)
                or tuple(rm.reconstruction_shape) != tuple(tm.reconstruction_shape)
                or rm.frame_count != tm.frame_count):
            raise FpmError(abi.STATUS_CODES["InvalidModel"],
                           "true and reconstruction models must have matching image, object, and frame dimensions")
        if self._camera is not None:
            rm = self._camera.compile_reconstruction_model(rm)
        missing = list(self._errors.missing_frames) if self._errors is not None else []
        true_model = copy.deepcopy(tm)
        if self._errors is not None:
            true_model = apply_illumination_acquisition_errors(true_model, self._errors, self._seed)
        ctx = context or default_context()
        dm = DeviceModel(ctx, true_model)  # ImagePlaneModel::validate on upload
        try:
            F, N = true_model.frame_count, true_model.image_len
            out = np.zeros((F,) + tuple(true_model.image_shape), np.float64)
            cam, keep = self._camera.to_c() if self._camera is not None else (None, None)
            mf = np.ascontiguousarray(np.asarray(missing, dtype=np.uint64))
            check(lib().fpmhip_simulate(ctx.handle, dm.handle, _cd(self._object),
                                        ct.byref(cam) if cam is not None else None,
                                        mf.ctypes.data_as(abi.P_u64) if mf.size else None, ct.c_uint64(mf.size),
                                        ct.c_uint64(self._seed & ((1 << 64) - 1)), _cd(out)))
        finally:
            dm.close()
        weights = None
        if missing:
            weights = np.ones(F)
            weights[missing] = 0.0
        return SimulationResult(
            measurements=MeasurementStack(frames=out, weights=weights),
            ground_truth_object=self._object.copy(),
            true_model=true_model,
            reconstruction_model=rm,
            camera=self._camera,
            illumination_acquisition_errors=self._errors,
            parameters=SimulationParameters(ideal=self._ideal, frame_count=F, image_shape=tuple(true_model.image_shape),
                                            missing_frames=missing),
            random_seed=self._seed,
        )
