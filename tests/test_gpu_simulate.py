"""The simulation surface on the device (fpmhip_camera_measure /
fpmhip_simulate) against the oracle, and the reference's simulation tests
(tests/simulation_tests.rs) restated. The camera's deterministic steps must
match the oracle exactly; with noise the device draws the same Philox
stream as the oracle's restatement, so counts agree draw for draw (a device
log / lgamma / cos can differ from the host's by an ulp: quantized counts
may then differ by one at a rounding or acceptance boundary, allowed in at
most 1e-5 of the pixels; unquantized values agree to 1e-12 relative)."""
import sys

import numpy as np
import pytest

import oracle
from fpmhip import AlternatingProjection, ReconstructionProblem, runtime
from fpmhip import synthetic as syn
from fpmhip.runtime import FpmError
from fpmhip.simulation import CameraModel, IlluminationAcquisitionErrors, Simulator

import refobjects as ro

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def assert_counts_match(dev, ref, quantized):
    if quantized:
        diff = dev != ref
        assert diff.mean() <= 1e-5 and (not diff.any() or np.abs(dev - ref)[diff].max() <= 1.0), diff.sum()
    else:
        assert np.allclose(dev, ref, rtol=1e-12, atol=1e-9)


CAMERAS = {
    "deterministic": CameraModel.new().with_photons_per_pixel(700.0).gain(1.7).with_offset_counts(3.0)
    .with_dark_current_electrons(2.5).with_bit_depth(12).saturation(3000.0),
    "shot": CameraModel.new().with_photons_per_pixel(400.0).with_shot_noise(True),
    "shot_low": CameraModel.new().with_photons_per_pixel(6.0).with_shot_noise(True).with_bit_depth(8),
    "shot_read_float": CameraModel.ideal().with_photons_per_pixel(90.0).with_shot_noise(True)
    .with_read_noise_electrons(2.0).gain(1.3).with_offset_counts(4.0),
    "everything": CameraModel.new().with_photons_per_pixel(1500.0).with_shot_noise(True).with_read_noise_electrons(3.0)
    .with_dark_current_electrons(4.0).gain(0.8).with_offset_counts(10.0).with_bit_depth(10).saturation(900.0)
    .with_pixel_sensitivity(np.linspace(0.5, 1.5, 96 * 80)).with_bad_pixels([0, 17, 96 * 80 - 1], 50.4),
}


@pytest.mark.parametrize("name", list(CAMERAS))
def test_camera_measure_matches_oracle(ctx, name):
    cam = CAMERAS[name]
    rng = np.random.default_rng(5)
    stack = rng.uniform(-0.05, 1.2, size=(3, 96, 80))
    say("camera", name, stack.shape)
    dev = cam.measure_frames(stack.copy(), seed=77)
    ref = oracle.camera_measure(cam, stack, 77)
    assert_counts_match(dev, ref, cam.quantize)
    if cam.shot_noise:
        # the frame index enters the stream: equal frames give different draws
        same = cam.measure_frames(np.repeat(stack[:1], 2, axis=0).copy(), seed=77)
        assert not np.array_equal(same[0], same[1])


def test_reference_camera_unit_cases(ctx):
    # simulation_tests.rs:101-123, 208-222
    assert list(CameraModel.ideal().measure_frame(np.array([0.0, 0.5, 2.0]), 1)) == [0.0, 0.5, 2.0]
    cam = (CameraModel.ideal().with_photons_per_pixel(10.0).with_pixel_sensitivity([1.0, 0.5])
           .with_dark_current_electrons(3.0).gain(2.0).with_offset_counts(5.0))
    assert list(cam.measure_frame(np.array([2.0, 2.0]), 2)) == [51.0, 31.0]
    cam = (CameraModel.ideal().with_photons_per_pixel(10.0).gain(2.0).with_offset_counts(1.0).with_bit_depth(4)
           .saturation(10.0).with_quantize(True).with_bad_pixels([1], 7.4))
    assert list(cam.measure_frame(np.array([-1.0, 0.12, 0.55, 1.0]), 7)) == [1.0, 7.0, 10.0, 10.0]


def test_camera_noise_is_seeded(ctx):
    # simulation_tests.rs:125-141
    cam = CameraModel.ideal().with_photons_per_pixel(50.0).with_shot_noise(True).with_read_noise_electrons(1.5)
    a = cam.measure_frame(np.ones(16), 3)
    b = cam.measure_frame(np.ones(16), 3)
    assert np.array_equal(a, b) and np.any(a != 50.0)
    assert not np.array_equal(a, cam.measure_frame(np.ones(16), 4))


@pytest.mark.parametrize("kind", ["shot", "read", "shot+read"])
def test_camera_noise_moments(ctx, kind):
    # simulation_tests.rs:143-205, the reference's bounds on 65536 pixels
    frame = np.ones(65_536)
    if kind == "shot":
        cam = CameraModel.ideal().with_photons_per_pixel(100.0).with_shot_noise(True)
        mean, var, tm, tv = 100.0, 100.0, 0.5, 4.0
    elif kind == "read":
        frame = np.zeros(65_536)
        cam = (CameraModel.ideal().with_dark_current_electrons(50.0).with_read_noise_electrons(4.0).gain(2.0)
               .with_offset_counts(7.0))
        mean, var, tm, tv = 107.0, 64.0, 0.35, 3.0
    else:
        cam = (CameraModel.ideal().with_photons_per_pixel(80.0).with_shot_noise(True).with_read_noise_electrons(3.0)
               .gain(1.5).with_offset_counts(11.0))
        mean, var, tm, tv = 1.5 * 80 + 11, 1.5**2 * (80 + 9), 0.6, 5.0
    out = cam.measure_frame(frame, 6)
    say(kind, out.mean(), out.var(ddof=1))
    assert abs(out.mean() - mean) < tm and abs(out.var(ddof=1) - var) < tv


def test_ideal_simulation_is_the_forward_stack(ctx):
    # simulation_tests.rs:14-40
    model = ro.direct_model()
    obj = ro.phase_disk((16, 16), 4.0, 0.7)
    sim = Simulator.ideal(model).object(obj).seed(7).simulate()
    I = sim.measurements.frames
    ref = oracle.forward_stack(model, ro.spectrum_of(obj))
    assert I.shape == (5, 8, 8) and rel(I, ref) <= 1e-13
    assert np.all(np.isfinite(I)) and np.all(I >= 0)
    assert np.array_equal(sim.true_model.pupil, model.pupil) and np.array_equal(sim.reconstruction_model.pupil,
                                                                                 model.pupil)
    assert sim.parameters.ideal and sim.measurements.weights is None


def test_simulation_quantizes_and_saturates(ctx):
    # simulation_tests.rs:76-98
    sim = (Simulator.new(ro.direct_model()).object(np.full((16, 16), 10.0 + 0j))
           .camera(CameraModel.new().with_photons_per_pixel(1000.0).with_bit_depth(8).saturation(200.0)).simulate())
    v = sim.measurements.frames
    assert np.all(v == np.round(v)) and v.max() <= 200.0 and np.any(v == 200.0)


def test_poisson_gaussian_simulation_is_reproducible_and_matches_oracle(ctx):
    # simulation_tests.rs:49-74
    model = ro.direct_model()
    obj = ro.resolution_target((16, 16))
    cam = CameraModel.new().with_photons_per_pixel(200.0).with_shot_noise(True).with_read_noise_electrons(1.5) \
        .with_bit_depth(12)
    a = Simulator.new(model).object(obj).camera(cam).seed(1234).simulate()
    b = Simulator.new(model).object(obj).camera(cam).seed(1234).simulate()
    assert np.array_equal(a.measurements.frames, b.measurements.frames)
    ref = oracle.camera_measure(cam, oracle.forward_stack(model, ro.spectrum_of(obj)), 1234)
    assert_counts_match(a.measurements.frames, ref, True)
    assert np.array_equal(a.reconstruction_model.frame_gain, np.full(5, 200.0))


def test_missing_frames_become_zero_weight_background_frames(ctx):
    # simulation_tests.rs:437-460, and simulator.rs:177-184 with an optical background
    sim = (Simulator.new(ro.direct_model()).object(np.ones((16, 16), complex))
           .illumination_acquisition_errors(IlluminationAcquisitionErrors.new().with_missing_frames([1, 3])).simulate())
    assert sim.parameters.missing_frames == [1, 3]
    for f in (1, 3):
        assert sim.measurements.weights[f] == 0.0 and np.all(sim.measurements.frames[f] == 0.0)
    assert np.all(sim.measurements.weights[[0, 2, 4]] == 1.0) and np.all(sim.measurements.frames[0] > 0)
    bg = np.arange(5 * 64, dtype=np.float64).reshape(5, 64) * 0.01
    model = ro.direct_model().with_background(bg.ravel())
    sim = (Simulator.new(model).object(np.ones((16, 16), complex))
           .illumination_acquisition_errors(IlluminationAcquisitionErrors.new().with_missing_frames([2])).simulate())
    assert np.array_equal(sim.measurements.frames[2].ravel(), bg[2])
    ReconstructionProblem(sim.measurements, sim.reconstruction_model)


def test_source_permutation_reorders_true_sources_only(ctx):
    # simulation_tests.rs:462-484
    model = ro.direct_model()
    obj = ro.phase_disk((16, 16), 4.0, 0.5)
    order = [4, 3, 2, 1, 0]
    base = Simulator.new(model).object(obj).simulate()
    sim = (Simulator.new(model).object(obj)
           .illumination_acquisition_errors(IlluminationAcquisitionErrors.new().with_source_permutation(order))
           .simulate())
    for frame, source in enumerate(order):
        assert np.array_equal(sim.true_model.k_vectors[frame], model.k_vectors[source])
        assert np.array_equal(sim.measurements.frames[frame], base.measurements.frames[source])
    assert np.array_equal(sim.reconstruction_model.k_vectors, model.k_vectors)
    assert sim.illumination_acquisition_errors is not None


def test_frame_gain_variation_scales_frames(ctx):
    # simulation_tests.rs:486-518
    err = IlluminationAcquisitionErrors.new().with_frame_gain_relative_std(0.2)
    obj = np.ones((16, 16), complex)
    w = Simulator.new(ro.direct_model().with_frame_gains(np.full(5, 2.0))).object(obj) \
        .illumination_acquisition_errors(err).seed(19).simulate()
    u = Simulator.new(ro.direct_model()).object(obj).illumination_acquisition_errors(err).seed(19).simulate()
    assert np.allclose(w.true_model.frame_gain, 2.0 * u.true_model.frame_gain, rtol=0, atol=1e-12)
    assert rel(w.measurements.frames, 2.0 * u.measurements.frames) <= 1e-14
    assert np.array_equal(w.reconstruction_model.frame_gain, np.full(5, 2.0))


def test_dark_current_and_fixed_bad_pixels(ctx):
    # simulation_tests.rs:553-573
    sim = (Simulator.new(ro.direct_model()).object(np.ones((16, 16), complex))
           .camera(CameraModel.new().with_photons_per_pixel(10.0).with_dark_current_electrons(7.0)
                   .with_bad_pixels([0, 5], 123.0).with_bit_depth(16)).simulate())
    for f in range(5):
        v = sim.measurements.frames[f].ravel()
        assert v[0] == 123.0 and v[5] == 123.0 and v[1] == 17.0


def test_optical_background_stays_in_the_model_and_offset_in_the_camera(ctx):
    # simulation_tests.rs:575-611
    tm = ro.direct_model().with_background(np.full(64, 2.0))
    obj = np.ones((16, 16), complex)
    optical = Simulator.ideal(tm).object(obj).simulate()
    measured = (Simulator.new(tm).reconstruction_model(ro.direct_model().with_background(np.full(64, 2.0)))
                .object(obj).camera(CameraModel.ideal().with_dark_current_electrons(3.0).with_offset_counts(5.0))
                .simulate())
    assert np.abs(measured.measurements.frames - optical.measurements.frames - 8.0).max() < 1e-12
    assert np.array_equal(measured.true_model.background, np.full(64, 2.0))
    assert np.array_equal(measured.reconstruction_model.background, np.full(64, 10.0))


def test_source_gains_encode_vignetting(ctx):
    # simulation_tests.rs:374-435
    rm = ro.direct_model().with_frame_gains(np.full(5, 2.0))
    tm = ro.direct_model().with_frame_gains([1.0, 1.0, 2.0, 1.0, 1.0])
    obj = ro.mixed_test_pattern((16, 16))
    baseline = Simulator.ideal(rm).object(obj).simulate()
    ordinary = Simulator.new(tm).reconstruction_model(rm).object(obj).simulate()
    for f in (0, 1, 3, 4):
        assert np.abs(ordinary.measurements.frames[f] - 0.5 * baseline.measurements.frames[f]).max() < 1e-12
    assert np.array_equal(ordinary.reconstruction_model.frame_gain, np.full(5, 2.0))


def test_multiplexed_simulation_is_an_incoherent_weighted_sum(ctx):
    # simulation_tests.rs:632-676
    obj = ro.mixed_test_pattern((16, 16))
    base = Simulator.ideal(ro.direct_model()).object(obj).simulate().measurements.frames
    mm = ro.direct_model().with_multiplexing([[(0, 0.25), (2, 0.75)], [(1, 0.5), (4, 0.5)]])
    mux = Simulator.ideal(mm).object(obj).simulate()
    assert mux.measurements.frames.shape == (2, 8, 8)
    assert np.abs(mux.measurements.frames[0] - (0.25 * base[0] + 0.75 * base[2])).max() < 1e-12
    assert np.abs(mux.measurements.frames[1] - (0.5 * base[1] + 0.5 * base[4])).max() < 1e-12
    prob = ReconstructionProblem(mux.measurements, mux.reconstruction_model)
    s = runtime.Session(prob, AlternatingProjection(iterations=2))
    try:
        s.initialize()
        assert len(s.run(2)) == 2
    finally:
        s.close()


def test_simulate_is_forward_stack_then_camera(ctx):
    """The bench workload's geometry (225 LEDs, 128^2 -> 375^2) with a
    Poisson 16-bit camera: the simulation equals the device forward stack
    measured by the device camera with the same seed, and the oracle's."""
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(15), (128, 128))
    obj = syn.random_field_object(model.reconstruction_shape, 7)
    cam = CameraModel.new().with_photons_per_pixel(3.0e4).with_shot_noise(True)
    say("simulate 225 x 128^2", model.reconstruction_shape)
    sim = Simulator.new(model).object(obj).camera(cam).seed(99).simulate()
    I = runtime.forward_stack(model, syn.object_spectrum(obj))
    assert_counts_match(sim.measurements.frames, cam.measure_frames(I.copy(), 99), True)
    ref = oracle.camera_measure(cam, oracle.forward_stack(model, ro.spectrum_of(obj)), 99)
    assert_counts_match(sim.measurements.frames, ref, True)
    assert sim.measurements.frames.max() <= 65535.0


def test_simulate_errors_are_typed(ctx):
    m = ro.direct_model()
    obj = np.ones((16, 16), complex)
    with pytest.raises(FpmError) as e:
        Simulator.new(m).object(obj).camera(CameraModel.new().with_bad_pixels([64], 1.0)).simulate()
    assert e.value.name == "bad_pixels"
    with pytest.raises(FpmError) as e:
        Simulator.new(m).object(obj).camera(CameraModel.new().with_pixel_sensitivity(np.ones(63))).simulate()
    assert e.value.name == "pixel_sensitivity"
