"""Diagnostics of a reconstruction state on the device (build_diagnostics,
src/reconstruction/runner.rs:298-433) against the oracle's metrics of the
oracle's prediction from the same state, and the frame schedules
(src/reconstruction/schedule.rs) on the device. Sums are reduced in a fixed
tree order on the device and sequentially in the oracle: 1e-12 relative;
counts, minima and maxima exact."""
import copy
import sys

import numpy as np
import pytest

import oracle
from fpmhip import Epry, Fpie, MeasurementStack, Mpie, ReconstructionProblem, runtime
from fpmhip import synthetic as syn
from fpmhip.runtime import FpmError

import refobjects as ro

pytestmark = pytest.mark.gpu

ALL = ("per_frame_error", "raw_frame_stats", "frame_summaries", "residual_images", "object_amplitude", "object_phase",
       "pupil")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problem(n=32, k=7, seed=3):
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(k), (n, n))
    obj = syn.random_field_object(model.reconstruction_shape, seed)
    model.with_frame_gains(np.linspace(0.8, 1.3, model.frame_count))
    model.with_background(np.full(model.image_len, 0.02))
    I = runtime.forward_stack(model, syn.object_spectrum(obj))
    counts, ppp = syn.poisson_counts(I, 3e3, seed)
    model.with_frame_gains(model.frame_gain * ppp).with_background(np.full(model.image_len, 0.02 * ppp))
    masks = np.ones(counts.shape, np.uint8)
    masks[1, :5, :7] = 0
    masks[4, 10:, 3] = 0
    return model, MeasurementStack(counts=counts, masks=masks)


def close(a, b, tol=1e-12):
    return abs(a - b) <= tol * max(abs(b), 1e-300)


def check_against_oracle(s, model, meas, d):
    gains, bg = s.calibration()
    m2 = copy.deepcopy(model)
    m2.with_frame_gains(gains)
    m2.with_background(None if bg is None else np.asarray(bg).ravel())
    pred = oracle.forward_stack(m2, s.spectrum(), s.pupil())
    y = meas.as_f64()
    F = model.frame_count
    for f in range(F):
        mk = meas.masks[f] if meas.masks is not None else None
        want = oracle.frame_objective(pred[f], y[f], mk)
        assert close(d.per_frame_objective[f], want), (f, d.per_frame_objective[f], want)
        st = oracle.intensity_stats(y[f])
        got = d.raw_frame_statistics[f]["metrics"]
        for k in ("mean", "std", "sum"):
            assert close(got[k], st[k]), (f, k)
        for k in ("min", "max", "zero_pixels", "saturated_pixels"):
            assert got[k] == st[k], (f, k)
        cm = oracle.compare_intensity(y[f], pred[f], mk)
        got = d.frame_diagnostics[f]["metrics"]
        for k in ("reference_sum", "estimate_sum", "residual_l1", "residual_l2", "residual_std", "normalized_l2"):
            assert close(got[k], cm[k]), (f, k, got[k], cm[k])
        assert abs(got["residual_mean"] - cm["residual_mean"]) <= 1e-12 * max(cm["residual_l1"], 1e-300)
        assert close(got["residual_max_abs"], cm["residual_max_abs"], 1e-11)
        r = np.where(mk.astype(bool) if mk is not None else True, pred[f] - y[f], 0.0)
        assert np.abs(d.residual_images[f] - r).max() <= 1e-11 * np.abs(r).max()
    o = s.object()
    assert np.abs(d.object_amplitude - np.abs(o)).max() <= 1e-15 * np.abs(o).max()
    assert np.abs(np.angle(np.exp(1j * (d.object_phase - np.angle(o))))).max() <= 1e-13
    P = s.pupil()
    assert np.array_equal(d.pupil_amplitude == 0, np.abs(P) == 0)
    assert np.abs(d.pupil_amplitude - np.abs(P)).max() <= 1e-15 * max(np.abs(P).max(), 1e-300)


@pytest.mark.parametrize("precision", ["c128", "f32"])
def test_diagnostics_match_oracle_after_mpie(ctx, precision):
    model, meas = problem()
    s = runtime.Session(ReconstructionProblem(meas, model), Mpie(iterations=3, momentum_interval=20),
                        precision=precision)
    try:
        s.initialize()
        s.run(3)
        say("diagnostics", precision)
        d = s.diagnostics(ALL, iteration=3)
        check_against_oracle(s, model, meas, d)
        assert d.frame_diagnostics[2]["iteration"] == 3 and d.frame_diagnostics[2]["frame_index"] == 2
        only = s.diagnostics({"per_frame_error"})
        assert np.array_equal(only.per_frame_objective, d.per_frame_objective)
        assert only.raw_frame_statistics is None and only.residual_images is None and only.pupil_phase is None
    finally:
        s.close()


def test_diagnostics_use_the_state_calibration(ctx):
    """EPRY recovering gains and background: the prediction uses the
    state's calibration (model_with_state_calibration, runner.rs:436-455)."""
    model, meas = problem(seed=5)
    alg = Epry(iterations=2, recover_frame_gains=True, recover_background=True, pupil_step=0.05)
    s = runtime.Session(ReconstructionProblem(meas, model), alg)
    try:
        s.initialize()
        s.run(2)
        gains, bg = s.calibration()
        assert not np.allclose(gains, model.frame_gain)
        check_against_oracle(s, model, meas, s.diagnostics(ALL))
    finally:
        s.close()


def test_fully_masked_frame(ctx):
    model, meas = problem()
    meas.masks[3] = 0
    meas.set_frame_weight(3, 0.0)  # a positive-weight frame needs an unmasked pixel
    s = runtime.Session(ReconstructionProblem(meas, model), Fpie(iterations=1))
    try:
        s.initialize()
        s.run(1)
        d = s.diagnostics({"per_frame_error", "raw_frame_stats"})
        assert d.per_frame_objective[3] == 0.0 and d.per_frame_objective[2] > 0.0
        with pytest.raises(FpmError) as e:
            s.diagnostics({"frame_summaries"})
        assert e.value.kind == "InvalidParameter" and e.value.name == "valid_mask"
    finally:
        s.close()


def test_snr_schedule_on_device_matches_reference(ctx):
    """tests/schedule_tests.rs:37-48 with the scores reduced on the device."""
    model = ro.direct_model()
    frames = np.stack([np.full((8, 8), v) for v in (4.0, 9.0, 100.0, 16.0, 1.0)])
    frames[1, 0, 0] = 1e12
    masks = np.ones((5, 8, 8), np.uint8)
    masks[1, 0, 0] = 0
    masks[2] = 0
    meas = MeasurementStack(frames=frames, weights=np.array([1.0, 1.0, 0.0, 0.5, 1.0]), masks=masks)
    s = runtime.Session(ReconstructionProblem(meas, model), Fpie())
    try:
        assert s.schedule("snr_weighted").tolist() == [1, 0, 3, 4, 2]
        assert s.schedule("spiral_out").tolist() == [2, 4, 3, 0, 1]
        assert s.schedule("brightfield_first")[0] == 2
    finally:
        s.close()


def test_random_shuffle_is_seeded_per_iteration_and_runs_like_the_oracle(ctx):
    """tests/schedule_tests.rs:24-35, and a shuffled run against the oracle
    given the same explicit orders."""
    model, meas = problem(n=32, k=5)
    prob = ReconstructionProblem(meas, model)
    s = runtime.Session(prob, Fpie(iterations=3))
    try:
        a = s.schedule("random_shuffle", iteration=3, seed=42)
        assert np.array_equal(a, s.schedule("random_shuffle", iteration=3, seed=42))
        assert not np.array_equal(a, s.schedule("random_shuffle", iteration=4, seed=42))
        assert not np.array_equal(a, s.schedule("random_shuffle", iteration=3, seed=43))
        assert sorted(a.tolist()) == list(range(model.frame_count))
        orders = [s.schedule("random_shuffle", iteration=i, seed=42) for i in range(3)]
        s.initialize()
        trace = s.run(3, "random_shuffle", seed=42)
        ref = oracle.run(model, meas, Fpie(iterations=3).to_c(), 3, orders=np.stack(orders))
        assert np.linalg.norm(s.spectrum() - ref["object_spectrum"]) <= 1e-10 * np.linalg.norm(ref["object_spectrum"])
        assert np.abs(trace - ref["trace"]).max() <= 1e-10 * np.abs(ref["trace"]).max()
    finally:
        s.close()
    # resumed halves: iterations 0-1, then 2 from first_iteration=2, equal the full run
    s1 = runtime.Session(prob, Fpie(iterations=3))
    s2 = runtime.Session(prob, Fpie(iterations=3))
    try:
        s1.initialize()
        s1.run(3, "random_shuffle", seed=42)
        s2.initialize()
        s2.run(2, "random_shuffle", seed=42)
        s2.run(1, "random_shuffle", seed=42, first_iteration=2)
        assert np.array_equal(s1.spectrum(), s2.spectrum())
    finally:
        s1.close()
        s2.close()
