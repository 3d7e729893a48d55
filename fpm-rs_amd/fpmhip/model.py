"""Host-side data of the reference interface: optics, LED array, compiled
ImagePlaneModel arrays, MeasurementStack and ReconstructionProblem.

Mirrors fpm-rs's Python surface (python/src/model.rs:163-172 compile_model,
python/src/reconstruction.rs ReconstructionProblem) with numpy storage; the
arithmetic of `compile_model` runs in the C++ host code of libfpmhip
(fpmhip_compile_planar).
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import abi


@dataclass
class Optics:
    """fpm_rs.Optics (src/experiment/optics.rs:44-60)."""

    wavelength_vacuum_m: float = 532e-9
    objective_na: float = 0.10
    magnification: float = 4.0
    camera_pixel_size: float = 6.5e-6
    illumination_refractive_index: float = 1.0
    objective_medium_refractive_index: float = 1.0
    defocus_distance: Optional[float] = None
    pupil_aberration: Optional[tuple] = None  # (astigmatism, coma, spherical, edge_apodization)

    def to_c(self) -> abi.Optics:
        o = abi.Optics()
        o.wavelength_vacuum_m = self.wavelength_vacuum_m
        o.objective_na = self.objective_na
        o.magnification = self.magnification
        o.camera_pixel_size = self.camera_pixel_size
        o.illumination_refractive_index = self.illumination_refractive_index
        o.objective_medium_refractive_index = self.objective_medium_refractive_index
        o.has_defocus = int(self.defocus_distance is not None)
        o.defocus_distance = self.defocus_distance or 0.0
        o.has_aberration = int(self.pupil_aberration is not None)
        if self.pupil_aberration is not None:
            o.astigmatism, o.coma, o.spherical, o.edge_apodization = self.pupil_aberration
        return o


@dataclass
class PlanarLedArray:
    """fpm_rs.PlanarLedArray (src/experiment/led_array.rs:120-313): shape
    (rows, columns), pitch (x, y), reference index (column, row), pose."""

    shape: tuple = (3, 3)
    pitch_m: tuple = (4e-3, 4e-3)
    reference_index: Optional[tuple] = None
    translation_m: tuple = (0.0, 0.0, -90e-3)
    rotation_rad: tuple = (0.0, 0.0, 0.0)

    def to_c(self) -> abi.PlanarArray:
        a = abi.PlanarArray()
        a.rows, a.columns = self.shape
        a.pitch_x_m, a.pitch_y_m = self.pitch_m
        ref = self.reference_index
        if ref is None:
            ref = ((self.shape[1] - 1) / 2.0, (self.shape[0] - 1) / 2.0)
        a.reference_column, a.reference_row = ref
        a.translation_m = (ct.c_double * 3)(*self.translation_m)
        a.rotation_rad = (ct.c_double * 3)(*self.rotation_rad)
        return a


@dataclass
class ImagePlaneModel:
    """Compiled model arrays (src/model/image_plane_fpm.rs:58-73)."""

    image_shape: tuple
    reconstruction_shape: tuple
    crop_start: np.ndarray  # (S, 2) uint64 (row, column)
    offset: np.ndarray  # (S, 2) float64 (row, column)
    pupil: np.ndarray  # (n_h, n_w) complex128
    support: np.ndarray  # (n_h, n_w) uint8
    frame_gain: Optional[np.ndarray] = None  # (F,)
    background: Optional[np.ndarray] = None  # (N,) or (F*N,)
    multiplexing: Optional[list] = None  # F rows of [(source, weight), ...]
    k_vectors: Optional[np.ndarray] = None  # (S, 2) (kx, ky)
    _keep: list = field(default_factory=list, repr=False)

    @property
    def source_count(self) -> int:
        return int(self.crop_start.shape[0])

    @property
    def frame_count(self) -> int:
        return len(self.multiplexing) if self.multiplexing is not None else self.source_count

    @property
    def image_len(self) -> int:
        return int(self.image_shape[0] * self.image_shape[1])

    def with_frame_gains(self, gains):
        self.frame_gain = None if gains is None else np.ascontiguousarray(gains, dtype=np.float64)
        return self

    def with_background(self, background):
        self.background = None if background is None else np.ascontiguousarray(background, dtype=np.float64).ravel()
        return self

    def with_multiplexing(self, rows):
        self.multiplexing = rows
        return self

    def with_subpixel_offsets(self, offsets):
        self.offset = np.ascontiguousarray(offsets, dtype=np.float64).reshape(-1, 2)
        return self

    def desc(self) -> abi.ModelDesc:
        """C descriptor; numpy buffers are kept alive on the model."""
        d = abi.ModelDesc()
        d.image_h, d.image_w = self.image_shape
        d.recon_h, d.recon_w = self.reconstruction_shape
        d.source_count = self.source_count
        d.frame_count = self.frame_count
        keep = []

        def hold(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a

        crop = hold(self.crop_start, np.uint64)
        d.crop_start = abi.ptr(crop, ct.c_uint64)
        off = hold(self.offset, np.float64)
        d.offset = abi.ptr(off, ct.c_double)
        pup = hold(self.pupil, np.complex128)
        d.pupil = pup.ctypes.data_as(abi.P_dbl)
        sup = hold(self.support, np.uint8)
        d.support = abi.ptr(sup, ct.c_uint8)
        if self.frame_gain is not None:
            g = hold(self.frame_gain, np.float64)
            d.frame_gain = abi.ptr(g, ct.c_double)
        if self.background is not None:
            b = hold(self.background, np.float64)
            d.background = abi.ptr(b, ct.c_double)
            d.background_len = b.size
        if self.multiplexing is not None:
            row_ptr = [0]
            src, wt = [], []
            for row in self.multiplexing:
                for s, w in row:
                    src.append(s)
                    wt.append(w)
                row_ptr.append(len(src))
            rp = hold(np.array(row_ptr), np.uint64)
            sa = hold(np.array(src), np.uint64)
            wa = hold(np.array(wt), np.float64)
            d.mux_row_ptr = abi.ptr(rp, ct.c_uint64)
            d.mux_source = abi.ptr(sa, ct.c_uint64)
            d.mux_weight = abi.ptr(wa, ct.c_double)
        if self.k_vectors is not None:
            k = hold(self.k_vectors, np.float64)
            d.k_vectors = abi.ptr(k, ct.c_double)
        self._keep = keep
        return d


@dataclass
class MeasurementStack:
    """fpm_rs.MeasurementStack (src/measurements/stack.rs:47-105): F frames of
    n_h x n_w. `counts` (uint16) keeps exact detector counts; `frames` is f64."""

    frames: Optional[np.ndarray] = None  # (F, n_h, n_w) float64
    counts: Optional[np.ndarray] = None  # (F, n_h, n_w) uint16
    weights: Optional[np.ndarray] = None  # (F,)
    masks: Optional[np.ndarray] = None  # (F, n_h, n_w) uint8

    def __post_init__(self):
        if (self.frames is None) == (self.counts is None):
            raise ValueError("exactly one of frames and counts is required")
        if self.frames is not None:
            self.frames = np.ascontiguousarray(self.frames, dtype=np.float64)
        else:
            self.counts = np.ascontiguousarray(self.counts, dtype=np.uint16)
        if self.weights is not None:
            self.weights = np.ascontiguousarray(self.weights, dtype=np.float64)
        if self.masks is not None:
            self.masks = np.ascontiguousarray(self.masks, dtype=np.uint8)

    @property
    def frame_count(self) -> int:
        return int((self.frames if self.frames is not None else self.counts).shape[0])

    def as_f64(self) -> np.ndarray:
        return self.frames if self.frames is not None else self.counts.astype(np.float64)

    def set_frame_weight(self, frame: int, weight: float):
        if self.weights is None:
            self.weights = np.ones(self.frame_count)
        self.weights[frame] = weight
        return self


@dataclass
class ReconstructionProblem:
    """fpm_rs.ReconstructionProblem (src/reconstruction/problem.rs:19-70)."""

    measurements: MeasurementStack
    model: ImagePlaneModel
    name: Optional[str] = None
