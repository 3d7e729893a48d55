"""Checkpoint format version 2 (src/reconstruction/checkpoint.rs) on the CPU:
the JSON layout of fpm-rs's serde representation, save/load round trips,
the corruption and mismatch cases of tests/callback_tests.rs:379-440, and a
resume through a JSON file with the oracle (2 + 2 == 4 iterations).

The layout checks follow the serde definitions the format is made of:
checkpoint.rs:430-465 (field order), array_serde.rs:8-14 (Array2Data),
model/pupil.rs:185-201, state.rs:34-53 and :86-95 (externally tagged enum),
trace.rs:6-41, and num-complex 0.4.6's Serialize, which writes a complex
value as the tuple (re, im), i.e. a JSON [re, im] array."""
import json
import os

import numpy as np
import pytest

import oracle
import refobjects as ro
from fpmhip import AlternatingProjection, Epry, MeasurementStack, Mpie, ReconstructionProblem
from fpmhip.checkpoint import FORMAT_VERSION, CheckpointEvery, MpieAuxiliary, ReconstructionCheckpoint
from fpmhip.runtime import FpmError, mpie_parameters

FIELD_ORDER = ["format_version", "completed_iterations", "object_spectrum", "pupil", "illumination_corrections",
               "frame_gains", "background", "physical_illumination_calibration", "calibrated_model",
               "algorithm_auxiliary", "trace"]
MPIE_FIELDS = ["velocity", "anchor", "effective_frames_since_momentum", "object_step", "stability", "epsilon",
               "momentum_interval", "momentum_friction", "momentum_feedback"]


def _problem():
    """callback_tests.rs:129: the direct model and an ideal simulation of a
    16 x 16 resolution target."""
    model = ro.direct_model()
    meas = MeasurementStack(frames=oracle.forward_stack(model, ro.spectrum_of(ro.resolution_target((16, 16)))))
    return ReconstructionProblem(meas, model)


def _oracle_checkpoint(problem, alg, iterations, resume=None):
    """ReconstructionCheckpoint::capture of an oracle run of `iterations`
    passes (continuing `resume`, another checkpoint, when given)."""
    kw = {}
    first, records = 0, []
    if resume is not None:
        kw["resume"] = resume_dict(resume)
        first, records = resume.completed_iterations, list(resume.trace)
    r = oracle.run(problem.model, problem.measurements, alg.to_c(), iterations - first, **kw)
    aux = None
    if isinstance(alg, Mpie):
        aux = MpieAuxiliary(r["velocity"].ravel(), r["anchor"].ravel(), r["counter"], **mpie_parameters(alg.to_c()))
    start = records[-1][2] if records else 0.0
    records += [(first + k + 1, float(o), start + 0.25 * (k + 1)) for k, o in enumerate(r["trace"])]
    ck = ReconstructionCheckpoint(iterations, r["object_spectrum"], r["recovered_pupil"],
                                  np.asarray(problem.model.support, np.uint8).copy(), None, None, aux, records)
    return r, ck


def resume_dict(ck):
    d = {"spectrum": ck.object_spectrum, "pupil": ck.pupil_values}
    if isinstance(ck.algorithm_auxiliary, MpieAuxiliary):
        R = ck.object_spectrum.shape
        d.update(velocity=ck.algorithm_auxiliary.velocity.reshape(R), anchor=ck.algorithm_auxiliary.anchor.reshape(R),
                 counter=ck.algorithm_auxiliary.effective_frames_since_momentum)
    if ck.frame_gains is not None:
        d["frame_gains"] = np.asarray(ck.frame_gains)
    if ck.background is not None:
        d["background"] = np.asarray(ck.background)
    return d


def _write(path, obj):
    with open(path, "w") as f:
        json.dump(obj, f)


def test_json_layout_is_the_serde_representation(tmp_path):
    prob = _problem()
    r, ck = _oracle_checkpoint(prob, Mpie(momentum_interval=7), 2)
    ck.save(tmp_path / "c.json")
    text = open(tmp_path / "c.json").read()
    obj = json.loads(text)
    assert list(obj) == FIELD_ORDER
    assert obj["format_version"] == 2 and obj["completed_iterations"] == 2
    spec = obj["object_spectrum"]
    assert list(spec) == ["height", "width", "data"] and (spec["height"], spec["width"]) == (16, 16)
    # row-major [re, im] pairs: element (2, 5) sits at 2 * 16 + 5
    assert spec["data"][2 * 16 + 5] == [r["object_spectrum"][2, 5].real, r["object_spectrum"][2, 5].imag]
    assert list(obj["pupil"]) == ["values", "support"]
    assert obj["pupil"]["support"]["data"] == [int(x) for x in np.asarray(prob.model.support).ravel()]
    for k in ("illumination_corrections", "frame_gains", "background", "physical_illumination_calibration",
              "calibrated_model"):
        assert obj[k] is None
    assert list(obj["algorithm_auxiliary"]) == ["Mpie"]
    m = obj["algorithm_auxiliary"]["Mpie"]
    assert list(m) == MPIE_FIELDS
    assert m["effective_frames_since_momentum"] == r["counter"] == 3  # 2 x 5 frames, interval 7
    assert (m["momentum_interval"], m["object_step"], m["stability"]) == (7, 0.2, 0.05)
    assert m["velocity"][17] == [r["velocity"].ravel()[17].real, r["velocity"].ravel()[17].imag]
    assert list(obj["trace"]) == ["iterations", "algorithm_metrics"]
    assert [list(rec) for rec in obj["trace"]["iterations"]] == [["iteration", "objective", "elapsed_seconds"]] * 2
    assert [rec["iteration"] for rec in obj["trace"]["iterations"]] == [1, 2]
    assert obj["trace"]["algorithm_metrics"] == []
    # compact, like serde_json::to_writer
    assert "\n" not in text and ", " not in text


def test_save_load_round_trip_is_bitwise(tmp_path):
    prob = _problem()
    _, ck = _oracle_checkpoint(prob, Mpie(momentum_interval=7), 3)
    ck.frame_gains = [1.0, 0.5, 2.0 / 3.0, 1e-300, 7.0]
    ck.background = list(np.linspace(-1.0 / 3.0, 1e9, 64))
    ck.save(tmp_path / "a.json")
    back = ReconstructionCheckpoint.load_for_problem(tmp_path / "a.json", prob)
    assert back.format_version == FORMAT_VERSION and back.completed_iterations == 3
    for a, b in ((ck.object_spectrum, back.object_spectrum), (ck.pupil_values, back.pupil_values),
                 (ck.algorithm_auxiliary.velocity, back.algorithm_auxiliary.velocity),
                 (ck.algorithm_auxiliary.anchor, back.algorithm_auxiliary.anchor)):
        assert a.shape == b.shape and a.tobytes() == b.tobytes()
    assert back.frame_gains == ck.frame_gains and back.background == ck.background
    assert back.trace == ck.trace
    back.save(tmp_path / "b.json")
    assert open(tmp_path / "a.json").read() == open(tmp_path / "b.json").read()
    assert not [p for p in os.listdir(tmp_path) if p.startswith(".checkpoint-")]  # no temporary left


def test_resume_through_json_matches_uninterrupted_oracle(tmp_path):
    """callback_tests.rs:339 with mPIE (interval 7, 2 + 2 == 4, 1e-14) and
    AP, through a checkpoint file."""
    prob = _problem()
    for alg in (Mpie(momentum_interval=7), AlternatingProjection()):
        _, ck = _oracle_checkpoint(prob, alg, 2)
        ck.save(tmp_path / "two.json")
        back = ReconstructionCheckpoint.load_for_problem(tmp_path / "two.json", prob)
        resumed, ck4 = _oracle_checkpoint(prob, alg, 4, resume=back)
        full = oracle.run(prob.model, prob.measurements, alg.to_c(), 4)
        assert np.abs(resumed["object_spectrum"] - full["object_spectrum"]).max() <= 1e-14
        assert np.array_equal(resumed["recovered_pupil"], full["recovered_pupil"])
        assert [r[0] for r in ck4.trace] == [1, 2, 3, 4]
        np.testing.assert_allclose([r[1] for r in ck4.trace], full["trace"], rtol=1e-12, atol=0)


def test_corruption_and_problem_mismatch_are_rejected(tmp_path):
    """callback_tests.rs:379-440."""
    prob = _problem()
    _, ck = _oracle_checkpoint(prob, AlternatingProjection(), 0)
    base = ck.to_json()

    bad = json.loads(json.dumps(base))
    bad["completed_iterations"] = 1  # trace holds no record for it
    _write(tmp_path / "invalid.json", bad)
    with pytest.raises(FpmError) as e:
        ReconstructionCheckpoint.load(tmp_path / "invalid.json")
    assert e.value.kind == "InvalidModel"

    ck.save(tmp_path / "valid.json")
    ReconstructionCheckpoint.load_for_problem(tmp_path / "valid.json", prob)

    sup = json.loads(json.dumps(base))
    first_on = int(np.flatnonzero(np.asarray(prob.model.support).ravel())[0])
    sup["pupil"]["support"]["data"][first_on] = 0
    _write(tmp_path / "support.json", sup)
    loaded = ReconstructionCheckpoint.load(tmp_path / "support.json")
    with pytest.raises(FpmError) as e:
        loaded.validate_for_problem(prob)
    assert e.value.kind == "InvalidModel"

    ver = json.loads(json.dumps(base))
    ver["format_version"] = 999
    _write(tmp_path / "corrupted.json", ver)
    with pytest.raises(FpmError) as e:
        ReconstructionCheckpoint.load(tmp_path / "corrupted.json")
    assert e.value.kind == "InvalidParameter" and e.value.name == "checkpoint format_version"

    inc = json.loads(json.dumps(base))
    inc["object_spectrum"] = {"height": 2, "width": 2, "data": [[0.0, 0.0]] * 4}
    _write(tmp_path / "incompatible.json", inc)
    ReconstructionCheckpoint.load(tmp_path / "incompatible.json")
    with pytest.raises(FpmError) as e:
        ReconstructionCheckpoint.load_for_problem(tmp_path / "incompatible.json", prob)
    assert e.value.kind == "InvalidShape"
    # the same values in {"re", "im"} form are not num-complex's representation
    inc["object_spectrum"]["data"] = [{"re": 0.0, "im": 0.0}] * 4
    _write(tmp_path / "objects.json", inc)
    with pytest.raises(FpmError) as e:
        ReconstructionCheckpoint.load(tmp_path / "objects.json")
    assert e.value.kind == "Serialization"


@pytest.mark.parametrize("mutate,kind", [
    (lambda o: o.update(extra=1), "Serialization"),  # deny_unknown_fields
    (lambda o: o.pop("trace"), "Serialization"),
    (lambda o: o["object_spectrum"]["data"].pop(), "Serialization"),  # LengthMismatch through serde
    (lambda o: o["pupil"]["support"]["data"].__setitem__(0, 2), "Serialization"),  # Pupil::new
    (lambda o: o.update(completed_iterations=-1), "Serialization"),
    (lambda o: o.update(frame_gains=[1.0, 1.0, 0.0, 1.0, 1.0]), "InvalidModel"),
    (lambda o: o.update(physical_illumination_calibration={}), "InvalidModel"),
    (lambda o: o["trace"]["iterations"][1].update(elapsed_seconds=0.1), "InvalidModel"),  # non-monotonic
    (lambda o: o["trace"]["iterations"][0].update(iteration=2), "InvalidModel"),
    (lambda o: o["trace"]["algorithm_metrics"].append(
        {"iteration": 3, "namespace": "a", "metric": "b", "value": 1.0}), "InvalidModel"),
    (lambda o: o["algorithm_auxiliary"]["Mpie"].update(effective_frames_since_momentum=7), "InvalidModel"),
    (lambda o: o["algorithm_auxiliary"]["Mpie"].update(momentum_friction=1.0), "InvalidModel"),
    (lambda o: o["algorithm_auxiliary"]["Mpie"]["anchor"].pop(), "InvalidModel"),
    (lambda o: o.update(algorithm_auxiliary={"Unknown": {}}), "Serialization"),
])
def test_validation_cases(tmp_path, mutate, kind):
    """validate() (checkpoint.rs:158-328) and the serde layer each reject
    their own corruptions."""
    prob = _problem()
    _, ck = _oracle_checkpoint(prob, Mpie(momentum_interval=7), 2)
    obj = ck.to_json()
    ReconstructionCheckpoint.from_json(json.loads(json.dumps(obj)))
    mutate(obj)
    _write(tmp_path / "m.json", obj)
    with pytest.raises(FpmError) as e:
        ReconstructionCheckpoint.load(tmp_path / "m.json")
    assert e.value.kind == kind


def test_nan_literal_is_rejected(tmp_path):
    prob = _problem()
    _, ck = _oracle_checkpoint(prob, AlternatingProjection(), 1)
    text = json.dumps(ck.to_json()).replace('"objective": ', '"objective": NaN, "x": ', 1)
    (tmp_path / "n.json").write_text(text)
    with pytest.raises(FpmError) as e:
        ReconstructionCheckpoint.load(tmp_path / "n.json")
    assert e.value.kind == "Serialization"


def test_problem_dimension_checks():
    """validate_for_problem (checkpoint.rs:330-427): gain count, background
    length (one image or one per frame), mPIE auxiliary length."""
    prob = _problem()
    _, ck = _oracle_checkpoint(prob, Mpie(momentum_interval=7), 1)
    ck.validate_for_problem(prob)
    for field, value in (("frame_gains", [1.0] * 4), ("background", [0.0] * 65)):
        bad = ReconstructionCheckpoint.from_json(ck.to_json())
        setattr(bad, field, value)
        with pytest.raises(FpmError, match="not match the model"):
            bad.validate_for_problem(prob)
    ok = ReconstructionCheckpoint.from_json(ck.to_json())
    ok.background = [0.0] * (64 * 5)
    ok.validate_for_problem(prob)
    small = ReconstructionCheckpoint.from_json(ck.to_json())
    small.algorithm_auxiliary.velocity = small.algorithm_auxiliary.velocity[:-1]
    small.algorithm_auxiliary.anchor = small.algorithm_auxiliary.anchor[:-1]
    with pytest.raises(FpmError, match="auxiliary dimensions"):
        small.validate_for_problem(prob)


def test_checkpoint_every_names_and_cadence(tmp_path):
    """callbacks/checkpoint.rs: frequency 0 counts as 1; file names are
    checkpoint_{iteration:05}.json; the directory is made at start."""
    cb = CheckpointEvery(0, tmp_path / "sub" / "dir")
    assert cb.every == 1 and cb.due(1) and cb.due(7)
    cb.on_start()
    assert os.path.isdir(tmp_path / "sub" / "dir")
    assert os.path.basename(cb.path(42)) == "checkpoint_00042.json"
    three = CheckpointEvery(3, tmp_path)
    assert [i for i in range(1, 10) if three.due(i)] == [3, 6, 9]
    prob = _problem()
    _, ck = _oracle_checkpoint(prob, Epry(pupil_step=0.05), 2)
    cb.write(ck)
    assert ReconstructionCheckpoint.load(cb.path(2)).completed_iterations == 2
