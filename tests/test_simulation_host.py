"""Host side of the simulation surface (CPU, no device calls): CameraModel
validation through the C ABI (camera.rs:145-245), compile_reconstruction_model
(camera.rs:289-318), the acquisition-error edits of the true model and the
simulator's argument checks (simulator.rs:107-341)."""
import numpy as np
import pytest

from fpmhip.runtime import FpmError
from fpmhip.simulation import (CameraModel, IlluminationAcquisitionErrors, Simulator,
                               apply_illumination_acquisition_errors)

import refobjects as ro


def bad(fn, name, kind="InvalidParameter"):
    with pytest.raises(FpmError) as e:
        fn()
    assert e.value.kind == kind and e.value.name == name, (e.value.kind, e.value.name, e.value.message)
    return e.value


@pytest.mark.parametrize("camera,name", [
    (CameraModel.new().with_photons_per_pixel(0.0), "photons_per_pixel"),
    (CameraModel.new().with_photons_per_pixel(float("nan")), "photons_per_pixel"),
    (CameraModel.new().gain(-1.0), "gain_counts_per_electron"),
    (CameraModel.new().with_offset_counts(float("inf")), "camera noise/offset"),
    (CameraModel.new().with_read_noise_electrons(-0.1), "camera noise/offset"),
    (CameraModel.new().with_dark_current_electrons(-1.0), "camera noise/offset"),
    (CameraModel.new().with_bit_depth(0), "bit_depth"),
    (CameraModel.new().with_bit_depth(33), "bit_depth"),
    (CameraModel.new().saturation(0.0), "saturation_counts"),
    (CameraModel.new().with_bad_pixels([1], -2.0), "bad_pixel_value_counts"),
    (CameraModel.new().with_pixel_sensitivity([1.0, -0.5]), "pixel_sensitivity"),
    (CameraModel.new().with_bad_pixels([3, 1, 3], 0.0), "bad_pixels"),
])
def test_camera_validate_rejects(camera, name):
    bad(camera.validate, name)


def test_bad_pixels_need_a_finite_value_or_maximum():
    # camera.rs:189-198: no value, no bit depth, no saturation
    cam = CameraModel.ideal()
    cam.bad_pixels = [0]
    bad(cam.validate, "bad_pixels")
    cam.bit_depth = 8
    cam.validate()


def test_validate_for_frame_checks_lengths_and_bounds():
    # simulation_tests.rs:236-250
    bad(lambda: CameraModel.ideal().with_pixel_sensitivity([1.0]).validate_for_frame(2), "pixel_sensitivity")
    e = bad(lambda: CameraModel.ideal().with_bad_pixels([2], 0.0).validate_for_frame(2), "bad_pixels")
    assert "below the frame size 2" in e.message
    CameraModel.ideal().with_bad_pixels([1], 0.0).validate_for_frame(2)


def test_maximum_count():
    assert CameraModel.new().maximum_count() == 65535.0
    assert CameraModel.new().with_bit_depth(8).saturation(200.0).maximum_count() == 200.0
    assert CameraModel.ideal().maximum_count() == float("inf")


def test_compile_reconstruction_model_gains_and_background():
    # camera.rs:289-318; simulation_tests.rs:576-611 (background 2 -> 10)
    m = ro.direct_model()
    cam = CameraModel.new().with_photons_per_pixel(250.0).gain(2.0)
    out = cam.compile_reconstruction_model(m)
    assert np.array_equal(out.frame_gain, np.full(5, 500.0)) and out.background is None
    assert m.frame_gain is None
    m.with_frame_gains([1.0, 2.0, 1.0, 1.0, 0.5])
    assert np.array_equal(cam.compile_reconstruction_model(m).frame_gain, [500.0, 1000.0, 500.0, 500.0, 250.0])
    m2 = ro.direct_model().with_background(np.full(64, 2.0))
    out2 = CameraModel.ideal().with_dark_current_electrons(3.0).with_offset_counts(5.0).compile_reconstruction_model(m2)
    assert np.array_equal(out2.background, np.full(64, 10.0))
    assert np.array_equal(m2.background, np.full(64, 2.0))
    out3 = CameraModel.ideal().with_offset_counts(4.0).compile_reconstruction_model(ro.direct_model())
    assert np.array_equal(out3.background, np.full(64, 4.0))


def test_acquisition_error_checks():
    # simulation_tests.rs:520-551
    m = ro.direct_model()
    bad(lambda: apply_illumination_acquisition_errors(
        m, IlluminationAcquisitionErrors.new().with_frame_gain_relative_std(-0.1), 0), "frame_gain_relative_std")
    bad(lambda: apply_illumination_acquisition_errors(
        m, IlluminationAcquisitionErrors.new().with_missing_frames([1, 1]), 0), "missing_frames")
    bad(lambda: apply_illumination_acquisition_errors(
        m, IlluminationAcquisitionErrors.new().with_missing_frames([5]), 0), "missing_frames")
    bad(lambda: apply_illumination_acquisition_errors(
        m, IlluminationAcquisitionErrors.new().with_source_permutation([0, 1, 2, 3, 3]), 0), "source_permutation")


def test_source_permutation_reorders_true_sources_only():
    # simulation_tests.rs:462-484 (the model edit; the device run is in the GPU tests)
    m = ro.direct_model().with_frame_gains([1.0, 2.0, 3.0, 4.0, 5.0])
    order = [4, 3, 2, 1, 0]
    out = apply_illumination_acquisition_errors(m, IlluminationAcquisitionErrors.new().with_source_permutation(order), 0)
    for frame, source in enumerate(order):
        assert np.array_equal(out.k_vectors[frame], m.k_vectors[source])
        assert np.array_equal(out.crop_start[frame], m.crop_start[source])
        assert out.frame_gain[frame] == m.frame_gain[source]
    assert np.array_equal(m.k_vectors[0], [-3.0, 0.0])


def test_frame_gain_variation_multiplies_existing_gains_and_is_seeded():
    # simulation_tests.rs:486-518
    err = IlluminationAcquisitionErrors.new().with_frame_gain_relative_std(0.2)
    w = apply_illumination_acquisition_errors(ro.direct_model().with_frame_gains(np.full(5, 2.0)), err, 19)
    u = apply_illumination_acquisition_errors(ro.direct_model(), err, 19)
    assert np.allclose(w.frame_gain, 2.0 * u.frame_gain, rtol=0, atol=1e-12)
    assert np.all(u.frame_gain >= 0.01) and not np.allclose(u.frame_gain, 1.0)
    assert not np.array_equal(u.frame_gain, apply_illumination_acquisition_errors(ro.direct_model(), err, 20).frame_gain)


def test_simulator_argument_checks():
    m = ro.direct_model()
    bad(lambda: Simulator.new(m).simulate(), "object")
    bad(lambda: Simulator.new(m).object(np.ones((8, 8), complex)).simulate(), "", "InvalidShape")
    other = ro.direct_model()
    other.image_shape = (4, 4)
    bad(lambda: Simulator.new(m).object(np.ones((16, 16), complex)).reconstruction_model(other).simulate(), "",
        "InvalidModel")
