"""ctypes mirror of the structs and enums declared in include/fpmhip.h."""
import ctypes as ct

c_u64 = ct.c_uint64
c_i32 = ct.c_int32
c_dbl = ct.c_double
P_dbl = ct.POINTER(ct.c_double)
P_u64 = ct.POINTER(ct.c_uint64)
P_u8 = ct.POINTER(ct.c_uint8)
P_u16 = ct.POINTER(ct.c_uint16)

# fpmhip_status
OK = 0
STATUS_NAMES = {
    0: "Ok",
    1: "InvalidShape",
    2: "InvalidParameter",
    3: "InvalidModel",
    4: "InvalidMeasurements",
    5: "LengthMismatch",
    6: "ShapeOverflow",
    7: "NonStandardLayout",
    8: "FrameOutOfRange",
    9: "Numerical",
    10: "Unsupported",
    11: "DeviceError",
    12: "OutOfMemory",
}
STATUS_CODES = {name: code for code, name in STATUS_NAMES.items()}

FORWARD, INVERSE = 0, 1
AP, FPIE, MPIE, EPRY = 0, 1, 2, 3
LOSS = {"amplitude_mse": 0, "intensity_mse": 1, "poisson_nll": 2, "huber_amplitude": 3}
SHAPE = {"exact": 0, "minimum": 1, "smooth": 2, "power_of_two": 3}
PLAIN_C128, MIXED_F32 = 0, 1
SCHEDULE = {"sequential": 0, "brightfield_first": 1, "spiral_out": 2, "snr_weighted": 3, "random_shuffle": 4}


class Optics(ct.Structure):
    _fields_ = [
        ("wavelength_vacuum_m", c_dbl),
        ("objective_na", c_dbl),
        ("magnification", c_dbl),
        ("camera_pixel_size", c_dbl),
        ("illumination_refractive_index", c_dbl),
        ("objective_medium_refractive_index", c_dbl),
        ("has_defocus", c_i32),
        ("has_aberration", c_i32),
        ("defocus_distance", c_dbl),
        ("astigmatism", c_dbl),
        ("coma", c_dbl),
        ("spherical", c_dbl),
        ("edge_apodization", c_dbl),
    ]


class PlanarArray(ct.Structure):
    _fields_ = [
        ("rows", c_u64),
        ("columns", c_u64),
        ("pitch_x_m", c_dbl),
        ("pitch_y_m", c_dbl),
        ("reference_column", c_dbl),
        ("reference_row", c_dbl),
        ("translation_m", c_dbl * 3),
        ("rotation_rad", c_dbl * 3),
    ]


class ModelDesc(ct.Structure):
    _fields_ = [
        ("image_h", c_u64),
        ("image_w", c_u64),
        ("recon_h", c_u64),
        ("recon_w", c_u64),
        ("source_count", c_u64),
        ("frame_count", c_u64),
        ("crop_start", P_u64),
        ("offset", P_dbl),
        ("pupil", P_dbl),
        ("support", P_u8),
        ("frame_gain", P_dbl),
        ("background", P_dbl),
        ("background_len", c_u64),
        ("mux_row_ptr", P_u64),
        ("mux_source", P_u64),
        ("mux_weight", P_dbl),
        ("k_vectors", P_dbl),
    ]


class CompiledShape(ct.Structure):
    _fields_ = [("source_count", c_u64), ("recon_h", c_u64), ("recon_w", c_u64)]


class Algorithm(ct.Structure):
    _fields_ = [
        ("kind", c_i32),
        ("loss_type", c_i32),
        ("recover_pupil", c_i32),
        ("constrain_pupil", c_i32),
        ("recover_gain", c_i32),
        ("recover_background", c_i32),
        ("iterations", c_u64),
        ("batch_size", c_u64),
        ("object_step", c_dbl),
        ("pupil_step", c_dbl),
        ("stability", c_dbl),
        ("epsilon", c_dbl),
        ("momentum_interval", c_u64),
        ("momentum_friction", c_dbl),
        ("momentum_feedback", c_dbl),
        ("gain_step", c_dbl),
        ("minimum_gain", c_dbl),
        ("maximum_gain", c_dbl),
        ("background_step", c_dbl),
        ("minimum_background", c_dbl),
        ("maximum_background", c_dbl),
    ]


class Camera(ct.Structure):
    _fields_ = [
        ("photons_per_pixel", c_dbl),
        ("gain_counts_per_electron", c_dbl),
        ("offset_counts", c_dbl),
        ("read_noise_electrons", c_dbl),
        ("dark_current_electrons", c_dbl),
        ("shot_noise", c_i32),
        ("quantize", c_i32),
        ("has_bit_depth", c_i32),
        ("bit_depth", c_i32),
        ("has_saturation", c_i32),
        ("has_bad_pixel_value", c_i32),
        ("saturation_counts", c_dbl),
        ("bad_pixel_value_counts", c_dbl),
        ("pixel_sensitivity", P_dbl),
        ("bad_pixels", ct.POINTER(c_u64)),
        ("bad_pixel_count", c_u64),
    ]


class FrameSchedule(ct.Structure):
    _fields_ = [("kind", c_i32), ("pad_", c_i32), ("seed", c_u64)]


class IntensityStats(ct.Structure):
    _fields_ = [("mean", c_dbl), ("std", c_dbl), ("min", c_dbl), ("max", c_dbl), ("sum", c_dbl),
                ("saturated_pixels", c_u64), ("zero_pixels", c_u64)]


class FrameComparison(ct.Structure):
    _fields_ = [("reference_sum", c_dbl), ("estimate_sum", c_dbl), ("residual_l1", c_dbl), ("residual_l2", c_dbl),
                ("residual_mean", c_dbl), ("residual_std", c_dbl), ("residual_max_abs", c_dbl),
                ("normalized_l2", c_dbl)]


class Diagnostics(ct.Structure):
    _fields_ = [("per_frame_objective", P_dbl), ("raw_frame_statistics", ct.POINTER(IntensityStats)),
                ("frame_summaries", ct.POINTER(FrameComparison)), ("residual_images", P_dbl),
                ("object_amplitude", P_dbl), ("object_phase", P_dbl), ("pupil_amplitude", P_dbl),
                ("pupil_phase", P_dbl)]


def ptr(array, ctype):
    """Pointer to a C-contiguous numpy array's data (None for None)."""
    if array is None:
        return None
    if not array.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return array.ctypes.data_as(ct.POINTER(ctype))
