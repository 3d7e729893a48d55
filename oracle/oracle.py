"""ctypes wrapper of the CPU FP64 oracle (libfpm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
from __future__ import annotations

import ctypes as ct
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "fpm-rs_amd")
if PKG not in sys.path:
    sys.path.insert(0, PKG)

from fpmhip import abi  # noqa: E402  (struct layouts of include/fpmhip.h)
from fpmhip.model import ImagePlaneModel, MeasurementStack  # noqa: E402

LIB_PATH = os.path.join(HERE, "libfpm_oracle.so")


class OracleError(RuntimeError):
    def __init__(self, code, message, name):
        super().__init__(f"{abi.STATUS_NAMES.get(code, code)}: {message}" + (f" [{name}]" if name else ""))
        self.code = code
        self.kind = abi.STATUS_NAMES.get(code, str(code))
        self.message = message
        self.name = name


class OrcResume(ct.Structure):
    _fields_ = [
        ("spectrum", abi.P_dbl),
        ("pupil", abi.P_dbl),
        ("velocity", abi.P_dbl),
        ("anchor", abi.P_dbl),
        ("counter", ct.c_uint64),
        ("frame_gains", abi.P_dbl),
        ("background", abi.P_dbl),
        ("background_len", ct.c_uint64),
    ]


class OrcOutputs(ct.Structure):
    _fields_ = [
        ("spectrum", abi.P_dbl),
        ("pupil", abi.P_dbl),
        ("object", abi.P_dbl),
        ("trace", abi.P_dbl),
        ("per_frame", abi.P_dbl),
        ("velocity", abi.P_dbl),
        ("anchor", abi.P_dbl),
        ("counter", abi.P_u64),
        ("frame_gains", abi.P_dbl),
        ("background", abi.P_dbl),
    ]


_lib = None


def host_isa_level():
    """'v4' when this host's CPU has the AVX-512 subsets of x86-64-v4, else 'v3'."""
    try:
        flags = set()
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
        return "v4" if {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags else "v3"
    except OSError:
        return "v3"


def use_widest_isa():
    """Loads the x86-64-v4 build of the oracle when the host supports it
    (before the first call into the oracle); returns the level in use."""
    global LIB_PATH
    level = host_isa_level()
    v4 = os.path.join(HERE, "libfpm_oracle_v4.so")
    if _lib is None and level == "v4" and os.path.exists(v4):
        LIB_PATH = v4
    return "v4" if LIB_PATH == v4 else "v3"


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        _lib = ct.CDLL(LIB_PATH)
        _lib.orc_last_error.restype = ct.c_char_p
        _lib.orc_last_error_name.restype = ct.c_char_p
    return _lib


def _check(code):
    if code != 0:
        L = lib()
        raise OracleError(code, L.orc_last_error().decode(), L.orc_last_error_name().decode())


def _cd(a):
    return a.ctypes.data_as(abi.P_dbl)


def point_loss(predicted, measured, loss):
    """point_loss (objective.rs:47-71) element-wise; loss is an abi LOSS code."""
    pr = np.ascontiguousarray(predicted, dtype=np.float64).ravel()
    me = np.ascontiguousarray(measured, dtype=np.float64).ravel()
    out = np.zeros_like(pr)
    _check(lib().orc_point_loss(_cd(pr), _cd(me), ct.c_uint64(pr.size), ct.c_int32(loss), _cd(out)))
    return out


def fft2(values, direction="forward"):
    v = np.array(values, dtype=np.complex128, order="C", copy=True)
    h, w = v.shape
    _check(lib().orc_fft2(_cd(v), ct.c_uint64(h), ct.c_uint64(w), ct.c_int32(0 if direction == "forward" else 1)))
    return v


def shift(values, inverse=False):
    v = np.ascontiguousarray(values, dtype=np.complex128)
    out = np.empty_like(v)
    h, w = v.shape
    _check(lib().orc_shift(_cd(v), _cd(out), ct.c_uint64(h), ct.c_uint64(w), ct.c_int32(int(inverse))))
    return out


def extract(spectrum, start, shape, offset=(0.0, 0.0)):
    s = np.ascontiguousarray(spectrum, dtype=np.complex128)
    out = np.empty(shape, dtype=np.complex128)
    _check(lib().orc_extract(_cd(s), ct.c_uint64(s.shape[0]), ct.c_uint64(s.shape[1]), ct.c_uint64(start[0]),
                             ct.c_uint64(start[1]), ct.c_uint64(shape[0]), ct.c_uint64(shape[1]),
                             ct.c_double(offset[0]), ct.c_double(offset[1]), _cd(out)))
    return out


def insert_adjoint(spectrum, start, update, scale=1.0, offset=(0.0, 0.0)):
    s = np.array(spectrum, dtype=np.complex128, order="C", copy=True)
    u = np.ascontiguousarray(update, dtype=np.complex128)
    _check(lib().orc_insert_adjoint(_cd(s), ct.c_uint64(s.shape[0]), ct.c_uint64(s.shape[1]),
                                    ct.c_uint64(start[0]), ct.c_uint64(start[1]), ct.c_uint64(u.shape[0]),
                                    ct.c_uint64(u.shape[1]), _cd(u), ct.c_double(scale), ct.c_double(offset[0]),
                                    ct.c_double(offset[1])))
    return s


def momentum_event(obj, velocity, anchor, friction, feedback):
    o = np.array(obj, dtype=np.complex128, copy=True)
    v = np.array(velocity, dtype=np.complex128, copy=True)
    a = np.array(anchor, dtype=np.complex128, copy=True)
    _check(lib().orc_momentum_event(_cd(o), _cd(v), _cd(a), ct.c_uint64(o.size), ct.c_double(friction),
                                    ct.c_double(feedback)))
    return o, v, a


def compile_model(optics, array, image_shape, reconstruction_shape="smooth"):
    """ImagePlaneModel::from_experiment for a planar LED array, in FP64 on the CPU."""
    if isinstance(reconstruction_shape, str):
        policy = abi.SHAPE[reconstruction_shape]
        eh = ew = 0
    else:
        policy = abi.SHAPE["exact"]
        eh, ew = reconstruction_shape
    oc, ac = optics.to_c(), array.to_c()
    shape = abi.CompiledShape()
    L = lib()
    args = (ct.byref(oc), ct.byref(ac), ct.c_uint64(image_shape[0]), ct.c_uint64(image_shape[1]),
            ct.c_int32(policy), ct.c_uint64(eh), ct.c_uint64(ew), ct.byref(shape))
    _check(L.orc_compile_planar(*args, None, None, None, None, None, None))
    S = shape.source_count
    crop = np.zeros((S, 2), np.uint64)
    off = np.zeros((S, 2), np.float64)
    kv = np.zeros((S, 2), np.float64)
    pupil = np.zeros(image_shape, np.complex128)
    support = np.zeros(image_shape, np.uint8)
    gains = np.zeros(S, np.float64)
    _check(L.orc_compile_planar(*args, abi.ptr(crop, ct.c_uint64), _cd(off), _cd(kv), _cd(pupil),
                                abi.ptr(support, ct.c_uint8), _cd(gains)))
    return ImagePlaneModel(tuple(image_shape), (shape.recon_h, shape.recon_w), crop, off, pupil, support,
                           frame_gain=gains, k_vectors=kv)


def philox(counter, key):
    """Philox4x64-10 of one counter block: 4 uint64 words."""
    c = np.ascontiguousarray(counter, dtype=np.uint64)
    k = np.ascontiguousarray(key, dtype=np.uint64)
    out = np.zeros(4, np.uint64)
    _check(lib().orc_philox(c.ctypes.data_as(abi.P_u64), k.ctypes.data_as(abi.P_u64), out.ctypes.data_as(abi.P_u64)))
    return out


def camera_draws(mean, n, seed):
    """The camera samplers on pixels 0..n-1 of frame 0: Poisson(mean), or
    standard normal draws for mean < 0."""
    out = np.zeros(n, np.float64)
    _check(lib().orc_camera_draws(ct.c_double(mean), ct.c_uint64(n), ct.c_uint64(seed), _cd(out)))
    return out


def _camera_struct(cam):
    """abi.Camera from an object with CameraModel's fields (camera.rs:7-34)."""
    c = abi.Camera()
    c.photons_per_pixel = cam.photons_per_pixel
    c.gain_counts_per_electron = cam.gain_counts_per_electron
    c.offset_counts = cam.offset_counts
    c.read_noise_electrons = cam.read_noise_electrons
    c.dark_current_electrons = cam.dark_current_electrons
    c.shot_noise = int(bool(cam.shot_noise))
    c.quantize = int(bool(cam.quantize))
    c.has_bit_depth = int(cam.bit_depth is not None)
    c.bit_depth = int(cam.bit_depth or 0)
    c.has_saturation = int(cam.saturation_counts is not None)
    c.saturation_counts = cam.saturation_counts or 0.0
    c.has_bad_pixel_value = int(cam.bad_pixel_value_counts is not None)
    c.bad_pixel_value_counts = cam.bad_pixel_value_counts or 0.0
    keep = []
    if cam.pixel_sensitivity is not None:
        s = np.ascontiguousarray(cam.pixel_sensitivity, dtype=np.float64).ravel()
        keep.append(s)
        c.pixel_sensitivity = _cd(s)
    bp = np.ascontiguousarray(np.asarray(cam.bad_pixels, dtype=np.uint64))
    keep.append(bp)
    c.bad_pixels = bp.ctypes.data_as(abi.P_u64) if bp.size else None
    c.bad_pixel_count = bp.size
    return c, keep


def camera_measure(cam, frames, seed, frame_len=None):
    """CameraModel::measure_frame (camera.rs:248-287) over consecutive frames
    of frame_len pixels (default: the last two axes); returns a new array."""
    v = np.array(frames, dtype=np.float64, order="C", copy=True)
    n = frame_len or (int(np.prod(v.shape[1:])) if v.ndim > 1 else v.size)
    c, keep = _camera_struct(cam)
    _check(lib().orc_camera_measure(ct.byref(c), _cd(v), ct.c_uint64(v.size // n), ct.c_uint64(n), ct.c_uint64(seed)))
    return v


def forward_stack(model: ImagePlaneModel, spectrum, pupil=None):
    d = model.desc()
    s = np.ascontiguousarray(spectrum, dtype=np.complex128)
    p = None if pupil is None else np.ascontiguousarray(pupil, dtype=np.complex128)
    out = np.zeros((model.frame_count,) + tuple(model.image_shape), np.float64)
    _check(lib().orc_forward_stack(ct.byref(d), _cd(s), None if p is None else _cd(p), _cd(out)))
    return out


def _meas_args(meas: MeasurementStack):
    frames = np.ascontiguousarray(meas.as_f64())
    w = meas.weights
    m = meas.masks
    return frames, w, m


def initialize(model, meas):
    d = model.desc()
    frames, w, m = _meas_args(meas)
    out = np.zeros(model.reconstruction_shape, np.complex128)
    _check(lib().orc_initialize(ct.byref(d), _cd(frames), None if w is None else _cd(w),
                                None if m is None else abi.ptr(m, ct.c_uint8), _cd(out)))
    return out


def canonicalize(model, spectrum, pupil):
    d = model.desc()
    s = np.array(spectrum, dtype=np.complex128, copy=True)
    p = np.array(pupil, dtype=np.complex128, copy=True)
    _check(lib().orc_canonicalize(ct.byref(d), _cd(s), _cd(p)))
    return s, p


def schedule(model, meas, name="sequential"):
    d = model.desc()
    frames, w, m = _meas_args(meas)
    order = np.zeros(model.frame_count, np.uint64)
    _check(lib().orc_schedule(ct.byref(d), _cd(frames), None if w is None else _cd(w),
                              None if m is None else abi.ptr(m, ct.c_uint8), ct.c_int32(abi.SCHEDULE[name]),
                              abi.ptr(order, ct.c_uint64)))
    return order


def run(model, meas, algorithm: abi.Algorithm, iterations, schedule_name="sequential", orders=None, resume=None):
    """Runner::run with the oracle; returns a dict of numpy arrays."""
    d = model.desc()
    frames, w, m = _meas_args(meas)
    R = tuple(model.reconstruction_shape)
    N = tuple(model.image_shape)
    F = model.frame_count
    res = {
        "object_spectrum": np.zeros(R, np.complex128),
        "recovered_pupil": np.zeros(N, np.complex128),
        "object": np.zeros(R, np.complex128),
        "trace": np.zeros(iterations, np.float64),
        "per_frame": np.zeros((iterations, F), np.float64),
        "velocity": np.zeros(R, np.complex128),
        "anchor": np.zeros(R, np.complex128),
        "frame_gains": np.zeros(F, np.float64),
        "background": np.zeros((F,) + N, np.float64),
    }
    counter = ct.c_uint64(0)
    out = OrcOutputs(_cd(res["object_spectrum"]), _cd(res["recovered_pupil"]), _cd(res["object"]),
                     _cd(res["trace"]), _cd(res["per_frame"]), _cd(res["velocity"]), _cd(res["anchor"]),
                     ct.pointer(counter), _cd(res["frame_gains"]), _cd(res["background"]))
    keep = []
    rp = None
    if resume is not None:
        def c(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.complex128)
            keep.append(a)
            return _cd(a)

        def r(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).ravel()
            keep.append(a)
            return _cd(a)

        bg = resume.get("background")
        rp = OrcResume(c(resume["spectrum"]), c(resume["pupil"]), c(resume.get("velocity")),
                       c(resume.get("anchor")), resume.get("counter", 0), r(resume.get("frame_gains")), r(bg),
                       0 if bg is None else np.asarray(bg).size)
    ord_arr = None
    if orders is not None:
        ord_arr = np.ascontiguousarray(orders, dtype=np.uint64).reshape(iterations, F)
    _check(lib().orc_run(ct.byref(d), _cd(frames), None if w is None else _cd(w),
                         None if m is None else abi.ptr(m, ct.c_uint8), ct.byref(algorithm),
                         ct.c_uint64(iterations), ct.c_int32(abi.SCHEDULE[schedule_name]),
                         None if ord_arr is None else abi.ptr(ord_arr, ct.c_uint64),
                         None if rp is None else ct.byref(rp), ct.byref(out)))
    res["counter"] = counter.value
    res["amplitude"] = np.abs(res["object"])
    res["phase"] = np.angle(res["object"])
    return res


def intensity_stats(values, saturation=None):
    """stats (src/metrics/intensity/single.rs:25-61) as a dict."""
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    out = np.zeros(7)
    _check(lib().orc_intensity_stats(_cd(v), ct.c_uint64(v.size), ct.c_int32(saturation is not None),
                                     ct.c_double(saturation or 0.0), _cd(out)))
    keys = ("mean", "std", "min", "max", "sum", "saturated_pixels", "zero_pixels")
    return {k: (int(x) if k.endswith("pixels") else float(x)) for k, x in zip(keys, out)}


def compare_intensity(reference, estimate, mask=None, saturation=None):
    """compare_intensity_u8_masked (src/metrics/intensity/compare.rs:134-226) as a dict."""
    r = np.ascontiguousarray(reference, dtype=np.float64).ravel()
    e = np.ascontiguousarray(estimate, dtype=np.float64).ravel()
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    out = np.zeros(9)
    _check(lib().orc_compare_intensity(_cd(r), _cd(e), None if m is None else abi.ptr(m, ct.c_uint8),
                                       ct.c_uint64(r.size), ct.c_int32(saturation is not None),
                                       ct.c_double(saturation or 0.0), _cd(out)))
    keys = ("reference_sum", "estimate_sum", "residual_l1", "residual_l2", "residual_mean", "residual_std",
            "residual_max_abs", "normalized_l2")
    d = {k: float(x) for k, x in zip(keys, out)}
    d["saturated_pixels"] = int(out[8]) if saturation is not None else None
    return d


def frame_objective(predicted, measured, mask=None):
    """The runner's per-frame amplitude objective of a state (runner.rs:361-373)."""
    p = np.ascontiguousarray(predicted, dtype=np.float64).ravel()
    y = np.ascontiguousarray(measured, dtype=np.float64).ravel()
    m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8).ravel()
    out = ct.c_double(0.0)
    _check(lib().orc_frame_objective(_cd(p), _cd(y), None if m is None else abi.ptr(m, ct.c_uint8),
                                     ct.c_uint64(p.size), ct.byref(out)))
    return out.value
