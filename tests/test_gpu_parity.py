"""GPU parity: the HIP path through the C ABI against the FP64 oracle on the
same inputs. PLAIN_C128 must match within 1e-10 relative L2 (object
spectrum, pupil, object) and MIXED_F32 within the stated FP32 tolerance."""
import math
import sys

import numpy as np
import pytest

import oracle
import refobjects as ro
from fpmhip import AlternatingProjection, Epry, Fpie, MeasurementStack, Mpie, ReconstructionProblem, runtime
from fpmhip import synthetic as syn

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_both(model, meas, alg, iters, precision="c128", schedule="sequential", orders=None):
    say("case", type(alg).__name__, model.image_shape, model.reconstruction_shape, iters, precision)
    ref = oracle.run(model, meas, alg.to_c(), iters, schedule_name=schedule)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision=precision)
    s.initialize()
    trace = s.run(iters, schedule)
    got = {"object_spectrum": s.spectrum(), "recovered_pupil": s.pupil(), "object": s.object(), "trace": trace}
    if isinstance(alg, Mpie):
        v, a, c = s.momentum()
        got.update(velocity=v, anchor=a, counter=c)
    return got, ref


def assert_parity(got, ref, tol, trace_tol=None):
    for key in ("object_spectrum", "recovered_pupil", "object"):
        e = rel(got[key], ref[key])
        say(key, e)
        assert e <= tol, (key, e)
    tt = trace_tol if trace_tol is not None else max(tol, 1e-12)
    te = np.abs(got["trace"] - ref["trace"]) / np.maximum(np.abs(ref["trace"]), 1e-300)
    say("trace", te.max() if te.size else 0)
    assert te.size == 0 or te.max() <= tt


@pytest.mark.parametrize("shape", [(7, 9), (8, 8), (128, 128), (42, 42), (375, 375), (64, 96)])
def test_fft2_matches_oracle(ctx, shape):
    rng = np.random.default_rng(5)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    for d in ("forward", "inverse"):
        g = ctx.fft2(x, d)
        r = oracle.fft2(x, d)
        assert rel(g, r) <= 1e-15, (d, rel(g, r))


def test_fft_round_trip_reference_case(ctx):
    """tests/forward_model_tests.rs:13 through the device primitive."""
    idx = np.arange(63)
    x = (idx / 13.0 + 1j * (idx % 5)).reshape(7, 9)
    np.testing.assert_allclose(ctx.fft2(ctx.fft2(x, "forward"), "inverse"), x, rtol=0, atol=1e-10)


def test_forward_stack_matches_oracle(ctx):
    model = ro.direct_model().with_subpixel_offsets([[0, 0], [0.3, -0.2], [0, 0], [-0.4, 0.1], [0, 0]])
    model.with_multiplexing([[(0, 0.7), (1, 0.3)], [(2, 0.4), (3, 0.6)], [(4, 1.0)]])
    model.with_frame_gains(np.array([2.0, 3.0, 0.5])).with_background(np.linspace(0, 1, 64))
    rng = np.random.default_rng(4)
    spec = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    P = np.exp(1j * rng.standard_normal((8, 8)))
    assert rel(runtime.forward_stack(model, spec, P), oracle.forward_stack(model, spec, P)) <= 1e-14


def test_forward_stack_constant_object(ctx):
    """tests/forward_model_tests.rs:76 on the device."""
    model = ro.direct_model()
    I = runtime.forward_stack(model, ro.spectrum_of(np.full((16, 16), 2.0 + 0j)))
    assert np.all(np.abs(I[2] - 4.0) < 1e-10)


def test_forward_stack_config2_model(ctx):
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(15), (128, 128))
    spec = syn.object_spectrum(syn.random_field_object(model.reconstruction_shape, 7))
    assert rel(runtime.forward_stack(model, spec), oracle.forward_stack(model, spec)) <= 1e-13


def _direct_problem(obj):
    model = ro.direct_model()
    return model, MeasurementStack(frames=oracle.forward_stack(model, ro.spectrum_of(obj)))


@pytest.mark.parametrize("alg", [AlternatingProjection(iterations=8), Fpie(iterations=8),
                                 Mpie(iterations=8, momentum_interval=3), Epry(iterations=4, pupil_step=0.05)])
def test_direct_model_plain_c128(ctx, alg):
    model, meas = _direct_problem(ro.phase_disk((16, 16), 4.5, 0.8) * ro.resolution_target((16, 16)))
    got, ref = run_both(model, meas, alg, alg.iterations)
    assert_parity(got, ref, 1e-10)
    if isinstance(alg, Mpie):
        assert got["counter"] == ref["counter"]
        assert rel(got["velocity"], ref["velocity"]) <= 1e-10


def test_config1_cpu_reference_case_ap20(ctx):
    """3x3 LEDs, 532 nm, NA 0.1, 4x, 6.5 um; 32 -> 42 ('smooth'); amplitude
    U[0.5,1], phase U[-pi/2, pi/2], seed 1234, sigma 1 px; ideal; AP 20."""
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(3), (32, 32))
    assert model.reconstruction_shape == (42, 42)
    obj = syn.uniform_object(model.reconstruction_shape, 1234)
    meas = MeasurementStack(frames=runtime.forward_stack(model, syn.object_spectrum(obj)))
    got, ref = run_both(model, meas, AlternatingProjection(iterations=20), 20)
    assert_parity(got, ref, 1e-10)


def _config2(n=128, k=15, peak=1e4, seed=7):
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(k), (n, n))
    obj = syn.random_field_object(model.reconstruction_shape, seed)
    I = runtime.forward_stack(model, syn.object_spectrum(obj))
    counts, ppp = syn.poisson_counts(I, peak, seed)
    model.with_frame_gains(np.full(model.frame_count, ppp))
    return model, MeasurementStack(counts=counts)


def test_config2_plain_c128_mpie(ctx):
    """225 LEDs, 128^2 -> 375^2, uint16 Poisson counts at 1e4 peak photons;
    mPIE with interval 30 (the default) and friction/feedback 0.5, 3 sweeps."""
    model, meas = _config2()
    got, ref = run_both(model, meas, Mpie(iterations=3, momentum_friction=0.5, momentum_feedback=0.5), 3)
    assert_parity(got, ref, 1e-10)


def test_config2_reduced_default_mpie_diverges_identically(ctx):
    """fpm-rs's default mPIE (interval 30, 0.9/0.9) diverges on a 225-LED
    array; the device follows the same trajectory (64^2, 6 sweeps)."""
    model, meas = _config2(n=64)
    got, ref = run_both(model, meas, Mpie(iterations=6), 6)
    assert_parity(got, ref, 1e-10)
    assert ref["trace"][-1] > ref["trace"][0]


def test_schedules_and_explicit_order(ctx):
    model, meas = _config2(n=32, k=7)
    for sched in ("brightfield_first", "spiral_out", "snr_weighted"):
        got, ref = run_both(model, meas, Fpie(iterations=2), 2, schedule=sched)
        assert_parity(got, ref, 1e-10)
        s = runtime.Session(ReconstructionProblem(meas, model), Fpie())
        assert np.array_equal(s.schedule(sched), oracle.schedule(model, meas, sched))


def test_multiplexed_masked_weighted(ctx):
    model = ro.direct_model().with_subpixel_offsets([[0, 0], [0.3, -0.2], [0, 0], [-0.4, 0.1], [0, 0]])
    model.with_multiplexing([[(0, 0.6), (1, 0.4)], [(2, 0.3), (3, 0.2), (4, 0.5)]])
    obj = ro.mixed_test_pattern((16, 16))
    I = oracle.forward_stack(model, ro.spectrum_of(obj))
    mask = np.ones(I.shape, np.uint8)
    mask[0, 2, 3] = 0
    mask[1, 5, 5] = 0
    meas = MeasurementStack(frames=I, weights=np.array([1.0, 0.7]), masks=mask)
    for alg in (AlternatingProjection(iterations=4), Mpie(iterations=4, momentum_interval=3),
                Epry(iterations=3, pupil_step=0.05)):
        got, ref = run_both(model, meas, alg, alg.iterations)
        assert_parity(got, ref, 1e-10)


def test_zero_weight_frame_skipped(ctx):
    model, meas = _direct_problem(ro.mixed_test_pattern((16, 16)))
    meas.set_frame_weight(1, 0.0)
    got, ref = run_both(model, meas, Mpie(iterations=3, momentum_interval=3), 3)
    assert_parity(got, ref, 1e-10)
    assert got["counter"] == ref["counter"]


def test_mpie_resume_exact_on_device(ctx):
    """tests/reconstruction_tests.rs:773 through set_state: 2 + 2 == 4."""
    model, meas = _direct_problem(ro.mixed_test_pattern((16, 16)))
    prob = ReconstructionProblem(meas, model)
    alg = Mpie(momentum_interval=7)
    a = runtime.Session(prob, alg)
    a.initialize()
    a.run(2)
    v, an, c = a.momentum()
    assert c == 3
    b = runtime.Session(prob, alg)
    b.set_state(a.spectrum(), a.pupil(), v, an, c)
    b.run(2)
    full = runtime.Session(prob, alg)
    full.initialize()
    full.run(4)
    assert np.abs(b.spectrum() - full.spectrum()).max() <= 1e-14


def test_initialization_undoes_gain_and_background(ctx):
    """tests/reconstruction_tests.rs:2713 on the device."""
    model = ro.direct_model().with_frame_gains(np.full(5, 4.0)).with_background(np.full(64, 9.0))
    I = oracle.forward_stack(model, ro.spectrum_of(np.full((16, 16), 2.0 + 0j)))
    r = AlternatingProjection(iterations=0).run(ReconstructionProblem(MeasurementStack(frames=I), model))
    np.testing.assert_allclose(r.amplitude, 2.0, rtol=0, atol=1e-10)


def test_errors_are_typed(ctx):
    model, meas = _direct_problem(ro.mixed_test_pattern((16, 16)))
    mask = np.ones((5, 8, 8), np.uint8)
    mask[3] = 0
    with pytest.raises(runtime.FpmError) as e:
        runtime.Session(ReconstructionProblem(MeasurementStack(frames=meas.frames, masks=mask), model), Fpie())
    assert e.value.kind == "InvalidMeasurements"
    bad = ro.direct_model().with_frame_gains(np.array([1, 1, 0.0, 1, 1]))
    with pytest.raises(runtime.FpmError) as e:
        s = runtime.Session(ReconstructionProblem(meas, bad), Fpie())
        s.initialize()
    assert e.value.kind == "InvalidModel"


@pytest.mark.parametrize("n", [64, 128])
def test_mixed_f32_per_stage_matches_fp64_oracle(ctx, n):
    """MIXED_F32 per-stage path: FP32 FFTs and storage, FP64 object spectrum
    and reductions. Amplitude/phase relative L2 vs the FP64 oracle after 5
    mPIE sweeps (225 LEDs, converging momentum 0.5/0.5) within 1e-5."""
    model, meas = _config2(n=n)
    alg = Mpie(iterations=5, momentum_friction=0.5, momentum_feedback=0.5)
    got, ref = run_both(model, meas, alg, 5, precision="f32")
    ea = rel(np.abs(got["object"]), np.abs(ref["object"]))
    ep = rel(np.angle(got["object"] * np.conj(ref["object"])) + np.angle(ref["object"]), np.angle(ref["object"]))
    say("mixed amplitude relL2", ea, "phase relL2", ep)
    assert ea <= 1e-5 and ep <= 1e-5


@pytest.mark.parametrize("precision,tol", [("c128", 1e-10), ("f32", 1e-5)])
def test_epry_gain_and_background_recovery_matches_oracle(ctx, precision, tol):
    """EPRY frame-gain and background recovery (common.rs:144-168, 390-489) on
    the device against the oracle: 81 LEDs, 64^2, per-frame gains and a
    per-frame background in the data, both recovered from a model without
    them; spectrum, pupil, gains and background."""
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(9), (64, 64))
    obj = syn.random_field_object(model.reconstruction_shape, 3)
    F, N = model.frame_count, model.image_len
    rng = np.random.default_rng(3)
    true = runtime.compile_model(syn.base_optics(), syn.planar_array(9), (64, 64))
    true.with_frame_gains(rng.uniform(0.7, 1.3, F) * 50.0)
    true.with_background(np.repeat(rng.uniform(0.0, 4.0, F), N))
    I = oracle.forward_stack(true, syn.object_spectrum(obj))
    model.with_frame_gains(np.full(F, 50.0))
    meas = MeasurementStack(frames=I)
    alg = Epry(iterations=3, recover_frame_gains=True, gain_step=0.5, recover_background=True,
               background_step=0.5, minimum_background=-10.0, maximum_background=10.0)
    ref = oracle.run(model, meas, alg.to_c(), 3)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision=precision)
    s.initialize()
    s.run(3)
    gains, bg = s.calibration()
    for got, key in ((s.spectrum(), "object_spectrum"), (s.pupil(), "recovered_pupil"), (gains, "frame_gains"),
                     (bg, "background")):
        e = rel(np.asarray(got).ravel(), ref[key].ravel())
        say(precision, key, e)
        assert e <= tol, (key, e)
    assert bg.shape == (F,) + tuple(model.image_shape)


def test_epry_recovery_reference_cases_on_device(ctx):
    """The reference's gain (:3037) and background (:3081) recovery cases,
    run through run_algorithm on the device."""
    true_gains = np.array([0.5, 1.0, 1.5, 2.0, 0.75])
    meas = MeasurementStack(frames=oracle.forward_stack(ro.direct_model().with_frame_gains(true_gains),
                                                        ro.spectrum_of(np.full((16, 16), 2.0 + 0j))))
    r = Epry(iterations=2, object_step=1e-6, recover_pupil=False, recover_frame_gains=True,
             gain_step=1.0).run(ReconstructionProblem(meas, ro.direct_model()))
    g = r.extra["recovered_frame_gains"]
    a = np.dot(g, true_gains) / np.dot(g, g)
    err = math.sqrt(np.sum((a * g - true_gains) ** 2) / np.sum(true_gains ** 2))
    init = math.sqrt(np.sum((true_gains.mean() - true_gains) ** 2) / np.sum(true_gains ** 2))
    assert err < 0.1 * init
    true_bg = np.array([0.0, 1.0, 2.0, 3.0, 0.5])
    meas = MeasurementStack(frames=oracle.forward_stack(ro.direct_model().with_background(np.repeat(true_bg, 64)),
                                                        ro.spectrum_of(np.full((16, 16), 2.0 + 0j))))
    r = Epry(iterations=1, object_step=1e-6, recover_pupil=False, recover_background=True, background_step=1.0,
             minimum_background=-10.0, maximum_background=10.0).run(ReconstructionProblem(meas, ro.direct_model()))
    means = r.extra["recovered_background"].reshape(5, 64).mean(axis=1)
    assert np.allclose(means - means.mean(), true_bg - true_bg.mean(), atol=1e-8, rtol=0)


def test_state_background_and_gains_restore(ctx):
    """set_state with per-frame gains and background: the device continues
    from them like the oracle's resume (FPIE, 2 sweeps, 1e-10)."""
    model, meas = _direct_problem(ro.mixed_test_pattern((16, 16)))
    F = model.frame_count
    rng = np.random.default_rng(5)
    gains = rng.uniform(0.8, 1.2, F)
    bg = rng.uniform(0.0, 0.1, (F, 8, 8))
    spec = ro.spectrum_of(ro.mixed_test_pattern((16, 16)))
    alg = Fpie(iterations=2)
    ref = oracle.run(model, meas, alg.to_c(), 2, resume={"spectrum": spec, "pupil": model.pupil,
                                                          "frame_gains": gains, "background": bg.ravel()})
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision="c128")
    s.set_state(spec, model.pupil, frame_gains=gains, background=bg)
    s.run(2)
    e = rel(s.spectrum(), ref["object_spectrum"])
    say("restored calibration", e)
    assert e <= 1e-10
    g2, b2 = s.calibration()
    assert np.array_equal(g2, gains) and np.array_equal(b2, bg)


def test_config2_plain_c128_mpie_full_length(ctx):
    """The measured workload's plain path at its full length: 225 LEDs, 128^2
    -> 375^2, uint16 counts, mPIE 50 sweeps (375 momentum events, interval
    30, friction/feedback 0.5) within 1e-10 of the oracle, velocity and
    counter included."""
    model, meas = _config2()
    alg = Mpie(iterations=50, momentum_friction=0.5, momentum_feedback=0.5)
    got, ref = run_both(model, meas, alg, 50)
    assert_parity(got, ref, 1e-10)
    assert got["counter"] == ref["counter"]
    assert rel(got["velocity"], ref["velocity"]) <= 1e-10


@pytest.mark.parametrize("loss", ["intensity_mse", "poisson_nll", "huber_amplitude"])
def test_loss_types_match_oracle(ctx, loss):
    """The non-default LossTypes (objective.rs:47-71) on the device: the
    per-stage FP64 path within 1e-10 (object, pupil and the objective trace,
    which is where the loss enters), and the fused batch (FP32 per-pixel
    terms, FP64 frame sums) within 1e-5. 49 LEDs, 64^2, mPIE 3 sweeps."""
    model, meas = _config2(n=64, k=7, peak=4e3)
    alg = Mpie(iterations=3, momentum_interval=30, momentum_friction=0.5, momentum_feedback=0.5, loss_type=loss)
    got, ref = run_both(model, meas, alg, 3)
    assert_parity(got, ref, 1e-10)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision="f32")
    b = runtime.Batch([s])
    trace = b.run(3)[0]
    te = np.abs(trace - ref["trace"]) / np.abs(ref["trace"])
    e = rel(s.spectrum(), ref["object_spectrum"])
    say(loss, "fused spectrum", e, "trace", te.max())
    assert e <= 1e-5 and te.max() <= 1e-5
    b.close()
