"""Checkpoint resume on the device (python/tests/test_callbacks_and_checkpoints.py
and tests/callback_tests.rs:339 of fpm-rs): checkpoints written by
CheckpointEvery during a device run, reloaded from JSON and resumed on a
fresh session, match the uninterrupted run; an oracle checkpoint resumes on
the device to the oracle's own uninterrupted result; mPIE rejects a resume
whose parameters differ from the checkpoint's."""
import json
import sys

import numpy as np
import pytest

import oracle
import refobjects as ro
from fpmhip import (AlternatingProjection, CheckpointEvery, Epry, Fpie, MeasurementStack, Mpie,
                    ReconstructionCheckpoint, ReconstructionProblem, runtime)
from fpmhip import synthetic as syn
from fpmhip.checkpoint import MpieAuxiliary

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem():
    model = ro.direct_model()
    meas = MeasurementStack(frames=oracle.forward_stack(model, ro.spectrum_of(ro.resolution_target((16, 16)))))
    return ReconstructionProblem(meas, model)


def _close(a, b, tol):
    assert a.shape == b.shape
    assert np.abs(a - b).max() <= tol, np.abs(a - b).max()


def test_checkpoint_every_writes_and_ap_resumes(ctx, tmp_path):
    """test_callbacks_and_checkpoints.py:10-54 and callback_tests.rs:339."""
    say("case AP checkpoint every 1, resume 2 -> 4")
    prob = _problem()
    d = tmp_path / "checkpoints"
    AlternatingProjection(iterations=2).run(prob, callbacks=[CheckpointEvery(1, d)])
    assert (d / "checkpoint_00001.json").is_file() and (d / "checkpoint_00002.json").is_file()
    ck = ReconstructionCheckpoint.load(d / "checkpoint_00002.json")
    ck.save(tmp_path / "copied.json")
    restored = ReconstructionCheckpoint.load(tmp_path / "copied.json")
    assert restored.format_version == 2 and restored.completed_iterations == 2 and len(restored.trace) == 2
    resumed = AlternatingProjection(iterations=4).run(prob, resume_from=restored)
    full = AlternatingProjection(iterations=4).run(prob)
    assert len(resumed.trace) == 4
    assert [r[0] for r in resumed.extra["trace_records"]] == [1, 2, 3, 4]
    _close(resumed.object_spectrum, full.object_spectrum, 1e-14)
    assert np.array_equal(resumed.recovered_pupil, full.recovered_pupil)
    np.testing.assert_allclose(resumed.trace, full.trace, rtol=1e-12, atol=0)
    elapsed = [r[2] for r in resumed.extra["trace_records"]]
    assert elapsed == sorted(elapsed) and elapsed[1] == restored.trace[1][2]


def test_epry_resume_preserves_canonical_arrays(ctx, tmp_path):
    """test_callbacks_and_checkpoints.py:57-82 (1e-12)."""
    say("case EPRY resume 1 -> 2")
    prob = _problem()
    alg1, alg2 = Epry(iterations=1, pupil_step=0.05), Epry(iterations=2, pupil_step=0.05)
    alg1.run(prob, callbacks=[CheckpointEvery(1, tmp_path)])
    ck = ReconstructionCheckpoint.load(tmp_path / "checkpoint_00001.json")
    resumed = alg2.run(prob, resume_from=ck)
    full = alg2.run(prob)
    _close(resumed.object_spectrum, full.object_spectrum, 1e-12)
    _close(resumed.recovered_pupil, full.recovered_pupil, 1e-12)


def test_mpie_resume_preserves_partial_momentum_interval(ctx, tmp_path):
    """test_callbacks_and_checkpoints.py:85-108: interval 7, checkpoint after
    1 pass of 5 frames, resumed to 3 passes (1e-12)."""
    say("case mPIE interval 7 resume 1 -> 3")
    prob = _problem()
    Mpie(iterations=1, momentum_interval=7).run(prob, callbacks=[CheckpointEvery(1, tmp_path)])
    ck = ReconstructionCheckpoint.load(tmp_path / "checkpoint_00001.json")
    assert isinstance(ck.algorithm_auxiliary, MpieAuxiliary)
    assert ck.algorithm_auxiliary.effective_frames_since_momentum == 5
    resumed = Mpie(iterations=3, momentum_interval=7).run(prob, resume_from=ck)
    full = Mpie(iterations=3, momentum_interval=7).run(prob)
    _close(resumed.object_spectrum, full.object_spectrum, 1e-12)
    assert resumed.extra["counter"] == full.extra["counter"] == 1  # 15 frames, events at 7 and 14


def test_oracle_checkpoint_resumes_on_device(ctx, tmp_path):
    """An oracle checkpoint (2 passes) written as JSON resumes on the device
    to 4 passes and matches the oracle's uninterrupted 4 (1e-10 relative)."""
    say("case oracle checkpoint -> device resume, FPIE and mPIE")
    prob = _problem()
    for alg in (Fpie(iterations=4), Mpie(iterations=4, momentum_interval=3)):
        r = oracle.run(prob.model, prob.measurements, alg.to_c(), 2)
        aux = None
        if isinstance(alg, Mpie):
            aux = MpieAuxiliary(r["velocity"].ravel(), r["anchor"].ravel(), r["counter"],
                                **runtime.mpie_parameters(alg.to_c()))
        recs = [(k + 1, float(o), 0.5 * (k + 1)) for k, o in enumerate(r["trace"])]
        ReconstructionCheckpoint(2, r["object_spectrum"], r["recovered_pupil"],
                                 np.asarray(prob.model.support, np.uint8).copy(), None, None, aux,
                                 recs).save(tmp_path / "o.json")
        resumed = alg.run(prob, resume_from=ReconstructionCheckpoint.load(tmp_path / "o.json"))
        full = oracle.run(prob.model, prob.measurements, alg.to_c(), 4)
        err = np.linalg.norm(resumed.object_spectrum - full["object_spectrum"]) / np.linalg.norm(full["object_spectrum"])
        assert err <= 1e-10, err
        np.testing.assert_allclose(resumed.trace, full["trace"], rtol=1e-10, atol=0)


def test_device_checkpoint_resumes_in_oracle(ctx, tmp_path):
    """The reverse direction: a device checkpoint of 2 mPIE passes resumed by
    the oracle matches the oracle's uninterrupted 4 (1e-10 relative)."""
    say("case device checkpoint -> oracle resume")
    prob = _problem()
    alg = Mpie(iterations=2, momentum_interval=3)
    alg.run(prob, callbacks=[CheckpointEvery(2, tmp_path)])
    ck = ReconstructionCheckpoint.load(tmp_path / "checkpoint_00002.json")
    a = ck.algorithm_auxiliary
    R = ck.object_spectrum.shape
    b = oracle.run(prob.model, prob.measurements, alg.to_c(), 2,
                   resume={"spectrum": ck.object_spectrum, "pupil": ck.pupil_values,
                           "velocity": a.velocity.reshape(R), "anchor": a.anchor.reshape(R),
                           "counter": a.effective_frames_since_momentum,
                           "frame_gains": None if ck.frame_gains is None else np.asarray(ck.frame_gains)})
    full = oracle.run(prob.model, prob.measurements, alg.to_c(), 4)
    err = np.linalg.norm(b["object_spectrum"] - full["object_spectrum"]) / np.linalg.norm(full["object_spectrum"])
    assert err <= 1e-10, err


def test_mpie_resume_rejects_changed_parameters(ctx, tmp_path):
    """algorithms/mpie.rs:166-222: parameters compared bit for bit; an
    auxiliary state of another algorithm is refused; a checkpoint without one
    starts fresh momentum."""
    say("case mPIE parameter checks")
    prob = _problem()
    Mpie(iterations=1, momentum_interval=7).run(prob, callbacks=[CheckpointEvery(1, tmp_path)])
    ck = ReconstructionCheckpoint.load(tmp_path / "checkpoint_00001.json")
    for kw, name in ((dict(object_step=0.2000000001), "object_step"), (dict(momentum_interval=8), "momentum_interval"),
                     (dict(momentum_feedback=0.5), "momentum_feedback")):
        cfg = dict(iterations=2, momentum_interval=7)
        cfg.update(kw)
        with pytest.raises(runtime.FpmError) as e:
            Mpie(**cfg).run(prob, resume_from=ck)
        assert e.value.kind == "InvalidParameter" and e.value.name == name
    foreign = ReconstructionCheckpoint.from_json(ck.to_json())
    foreign.algorithm_auxiliary = {"Admm": {}}
    with pytest.raises(runtime.FpmError) as e:
        Mpie(iterations=2, momentum_interval=7).run(prob, resume_from=foreign)
    assert e.value.kind == "InvalidModel"
    # AP checkpoint (no auxiliary state) resumed by mPIE: zero velocity, anchor = spectrum
    AlternatingProjection(iterations=1).run(prob, callbacks=[CheckpointEvery(1, tmp_path / "ap")])
    ap = ReconstructionCheckpoint.load(tmp_path / "ap" / "checkpoint_00001.json")
    got = Mpie(iterations=2, momentum_interval=7).run(prob, resume_from=ap)
    ref = oracle.run(prob.model, prob.measurements, Mpie(momentum_interval=7).to_c(), 1,
                     resume={"spectrum": ap.object_spectrum, "pupil": ap.pupil_values})
    err = np.linalg.norm(got.object_spectrum - ref["object_spectrum"]) / np.linalg.norm(ref["object_spectrum"])
    assert err <= 1e-10, err


def test_epry_calibration_round_trips_through_checkpoint(ctx, tmp_path):
    """EPRY recovering gains and background: the checkpoint holds both
    (frame_gains Some, background F*N), and resuming reproduces the
    uninterrupted run (1e-12)."""
    say("case EPRY gains + background resume")
    prob = _problem()
    cfg = dict(pupil_step=0.05, recover_frame_gains=True, recover_background=True)
    Epry(iterations=2, **cfg).run(prob, callbacks=[CheckpointEvery(2, tmp_path)])
    ck = ReconstructionCheckpoint.load(tmp_path / "checkpoint_00002.json")
    assert ck.frame_gains is not None and len(ck.frame_gains) == prob.model.frame_count
    assert ck.background is not None and len(ck.background) == prob.model.frame_count * prob.model.image_len
    resumed = Epry(iterations=3, **cfg).run(prob, resume_from=ck)
    full = Epry(iterations=3, **cfg).run(prob)
    _close(resumed.object_spectrum, full.object_spectrum, 1e-12)
    np.testing.assert_allclose(resumed.extra["recovered_frame_gains"], full.extra["recovered_frame_gains"],
                               rtol=1e-12, atol=0)
    np.testing.assert_allclose(resumed.extra["recovered_background"], full.extra["recovered_background"],
                               rtol=1e-12, atol=1e-12)


def test_null_calibration_in_checkpoint_clears_model_values(ctx, tmp_path):
    """A checkpoint with frame_gains / background null restores a state
    without them (state.rs:411-412): the device then uses gain 1 and no
    background even where the model has both, like the oracle given unit
    gains and a zero background."""
    say("case null calibration restore")
    base = _problem()
    model = ro.direct_model().with_frame_gains(np.full(5, 2.0)).with_background(np.full(64, 0.5))
    prob = ReconstructionProblem(base.measurements, model)
    r = oracle.run(model, base.measurements, Fpie().to_c(), 1)
    ck = ReconstructionCheckpoint(1, r["object_spectrum"], r["recovered_pupil"],
                                  np.asarray(model.support, np.uint8).copy(), None, None, None,
                                  [(1, float(r["trace"][0]), 0.1)])
    got = Fpie(iterations=3).run(prob, resume_from=ck)
    ref = oracle.run(model, base.measurements, Fpie().to_c(), 2,
                     resume={"spectrum": ck.object_spectrum, "pupil": ck.pupil_values,
                             "frame_gains": np.ones(5), "background": np.zeros(64)})
    err = np.linalg.norm(got.object_spectrum - ref["object_spectrum"]) / np.linalg.norm(ref["object_spectrum"])
    assert err <= 1e-10, err


def test_precision_form_mpie_resume(ctx, tmp_path):
    """MIXED_F32 mPIE (complex64 velocity and anchor) at 64^2 with 25 LEDs:
    a checkpoint after 2 passes resumed to 4 matches the uninterrupted 4."""
    say("case f32 mPIE resume 64^2")
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(5), (64, 64))
    obj = syn.random_field_object(model.reconstruction_shape, 3)
    I = runtime.forward_stack(model, syn.object_spectrum(obj))
    prob = ReconstructionProblem(MeasurementStack(frames=I), model)
    alg2, alg4 = Mpie(iterations=2, momentum_interval=7), Mpie(iterations=4, momentum_interval=7)
    alg2.run(prob, precision="f32", callbacks=[CheckpointEvery(2, tmp_path)])
    ck = ReconstructionCheckpoint.load(tmp_path / "checkpoint_00002.json")
    resumed = alg4.run(prob, precision="f32", resume_from=ck)
    full = alg4.run(prob, precision="f32")
    err = np.linalg.norm(resumed.object_spectrum - full.object_spectrum) / np.linalg.norm(full.object_spectrum)
    say("f32 resume relative difference", err)
    assert err <= 1e-12, err
    text = json.dumps(ck.to_json())
    assert '"Mpie"' in text
