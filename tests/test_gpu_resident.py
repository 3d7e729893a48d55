"""The ResidentBackend seam on the device (src/backend/traits.rs:77-106) and
ReconstructionState::from_object (src/reconstruction/state.rs:356-383)."""
import sys

import numpy as np
import pytest

import oracle
import refobjects as ro
from fpmhip import AlternatingProjection, MeasurementStack, Mpie, ReconstructionProblem, runtime
from fpmhip import synthetic as syn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_resident_buffers_reference_case(ctx):
    """tests/forward_model_tests.rs:33-72 on the device: capabilities, a 7 x 9
    complex buffer through forward and inverse resident FFTs back to its
    values (1e-10), and a real buffer round trip."""
    caps = runtime.capabilities(ctx)
    assert caps == {"preferred_memory": "device", "resident_buffers": True}
    shape = (7, 9)
    idx = np.arange(shape[0] * shape[1])
    original = idx / 17.0 + 1j * (idx % 3)
    b = runtime.ResidentBuffer("complex", original.size, ctx)
    info = b.info()
    assert info["location"] == "device" and info["len"] == original.size and info["device_address"] != 0
    assert np.array_equal(b.download(), np.zeros(original.size, np.complex128))  # zeroed
    b.upload(original)
    b.fft2(shape, "forward")
    b.fft2(shape, "inverse")
    assert np.abs(b.download() - original).max() <= 1e-10
    r = runtime.ResidentBuffer("real", 3, ctx)
    r.upload([1.0, 2.0, 3.0])
    assert r.download().tolist() == [1.0, 2.0, 3.0]
    b.close()
    r.close()


@pytest.mark.parametrize("shape", [(128, 128), (375, 375), (7, 9)])
def test_resident_fft_matches_oracle(ctx, shape):
    x = np.random.default_rng(3).standard_normal(shape) + 1j * np.random.default_rng(4).standard_normal(shape)
    b = runtime.ResidentBuffer("complex", x.size, ctx)
    for direction in ("forward", "inverse"):
        b.upload(x)
        b.fft2(shape, direction)
        ref = oracle.fft2(x, direction)
        e = np.linalg.norm(b.download().reshape(shape) - ref) / np.linalg.norm(ref)
        say(shape, direction, e)
        assert e <= 1e-14
    b.close()


def test_resident_buffer_errors_follow_cpu_backend(ctx):
    """cpu.rs:238-285: the other kind is InvalidParameter named by the typed
    operation; a length mismatch is InvalidShape; a shape that does not match
    the buffer is InvalidShape."""
    c = runtime.ResidentBuffer("complex", 4, ctx)
    r = runtime.ResidentBuffer("real", 4, ctx)
    with pytest.raises(runtime.FpmError) as e:
        r.upload(np.zeros(4), kind="complex")
    assert e.value.kind == "InvalidParameter" and e.value.name == "complex buffer"
    with pytest.raises(runtime.FpmError) as e:
        c.download(kind="real")
    assert e.value.kind == "InvalidParameter" and e.value.name == "real buffer"
    with pytest.raises(runtime.FpmError) as e:
        c.upload(np.zeros(5, np.complex128))
    assert e.value.kind == "InvalidShape"
    with pytest.raises(runtime.FpmError) as e:
        c.download(length=3)
    assert e.value.kind == "InvalidShape"
    with pytest.raises(runtime.FpmError) as e:
        c.fft2((3, 3))
    assert e.value.kind == "InvalidShape"
    with pytest.raises(runtime.FpmError) as e:
        r.fft2((2, 2))
    assert e.value.kind == "InvalidParameter" and e.value.name == "complex buffer"
    c.close()
    r.close()


@pytest.mark.parametrize("precision,tol", [("c128", 1e-10), ("f32", 1e-5)])
def test_initialize_from_object_then_run_matches_oracle(ctx, precision, tol):
    """from_object: the spectrum is fftshift(FFT(object) / M) of the given
    object (1e-14 against the oracle's transform), the mPIE anchor is that
    spectrum, and reconstruction continues from it as the oracle does from
    the same state (mPIE 3 sweeps, 49 LEDs, 64^2)."""
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(7), (64, 64))
    truth = syn.random_field_object(model.reconstruction_shape, 5)
    I = runtime.forward_stack(model, syn.object_spectrum(truth))
    meas = MeasurementStack(frames=I)
    start = syn.uniform_object(model.reconstruction_shape, 6)
    alg = Mpie(iterations=3, momentum_interval=20, momentum_friction=0.5, momentum_feedback=0.5)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision=precision)
    s.initialize_from_object(start)
    spec0 = oracle.shift(oracle.fft2(start, "forward"))
    e0 = np.linalg.norm(s.spectrum() - spec0) / np.linalg.norm(spec0)
    v, a, c = s.momentum()
    say(precision, "initial spectrum", e0)
    assert e0 <= (1e-14 if precision == "c128" else 1e-7)
    assert c == 0 and np.abs(v).max() == 0.0
    s.run(3)
    ref = oracle.run(model, meas, alg.to_c(), 3, resume={"spectrum": spec0, "pupil": model.pupil})
    e = np.linalg.norm(s.spectrum() - ref["object_spectrum"]) / np.linalg.norm(ref["object_spectrum"])
    say(precision, "after 3 sweeps", e)
    assert e <= tol
