"""Precision form (MIXED_F32): complex64 FFTs and updates, FP64 object
accumulation, denominators and reductions. EPRY on 225 LEDs with a true
pupil carrying Noll Z4-Z15 aberrations at 0.5 rad RMS (the reconstruction
model keeps the compiled circular pupil); amplitude and phase relative L2
against the FP64 oracle on identical inputs.

Measured errors of the precision form (object amplitude / phase / pupil):

  128^2 -> 375^2, 30 sweeps: 1.6e-6 / 2.6e-6 / 1.9e-6   (held to 1e-5)
  256^2 -> 729^2, 30 sweeps: 3.1e-5 / 4.7e-5 / 5.6e-6   (above 1e-5: strict xfail)

EPRY amplifies perturbations before it converges: in the FP64 oracle alone
one FP32 rounding of the initial spectrum (relative 8e-9) ends 170x / 260x
larger after 30 sweeps at 256^2 (tests/test_oracle_conditioning.py), and an
FP32 implementation injects such a rounding at every LED update. An FP32
emulation of the oracle (FP32 fields and FFTs, everything else FP64) gives
1.5e-5..4e-5 at 256^2 depending on where the roundings fall, and still
1.0e-5 with exact FFTs and FP32 fields only: the 256^2 error is a property of
the problem at FP32 working precision, not of one kernel. The 256^2 case is
therefore held to its measured FP32 floor, 6e-5, as a regression bound, and
the 1e-5 target is kept as a strict xfail (it fails until the error drops).
"""
import sys

import numpy as np
import pytest

import oracle
from fpmhip import Epry, MeasurementStack, ReconstructionProblem, runtime
from fpmhip import synthetic as syn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def problem(n, seed=7):
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(15), (n, n))
    obj = syn.random_field_object(model.reconstruction_shape, seed)
    true_pupil = syn.zernike_pupil(model, 0.5, seed + 100)
    I = runtime.forward_stack(model, syn.object_spectrum(obj), true_pupil)
    counts, ppp = syn.poisson_counts(I, 1e4, seed)
    model.with_frame_gains(np.full(model.frame_count, ppp))
    return model, MeasurementStack(counts=counts), true_pupil


def errors(got, ref):
    ea = np.linalg.norm(np.abs(got) - np.abs(ref)) / np.linalg.norm(np.abs(ref))
    ep = np.linalg.norm(np.angle(got * np.conj(ref))) / np.linalg.norm(np.angle(ref))
    return ea, ep


CONDITIONING_256 = ("the 256^2 EPRY problem amplifies FP32 rounding beyond the 1e-5 target: one FP32 rounding of "
                    "the initial state grows 170x / 260x in the FP64 oracle (tests/test_oracle_conditioning.py), and "
                    "an FP32 emulation of the oracle gives 1.5e-5..4e-5, still 1.0e-5 with exact FFTs")


@pytest.mark.parametrize("n,iters,bound", [
    (128, 30, 1e-5),
    pytest.param(256, 30, 1e-5, marks=pytest.mark.xfail(strict=True, reason=CONDITIONING_256)),
    (256, 30, 6e-5),  # the measured FP32 floor of this problem (see CONDITIONING_256), as a regression bound
])
def test_epry_precision_form_vs_fp64_oracle(ctx, n, iters, bound):
    model, meas, true_pupil = problem(n)
    say("precision form EPRY", n, model.reconstruction_shape, iters)
    alg = Epry(iterations=iters)
    ref = oracle.run(model, meas, alg.to_c(), iters)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision="f32")
    s.initialize()
    trace = s.run(iters)
    obj, pup = s.object(), s.pupil()
    ea, ep = errors(obj, ref["object"])
    epu = np.linalg.norm(pup - ref["recovered_pupil"]) / np.linalg.norm(ref["recovered_pupil"])
    te = np.abs(trace - ref["trace"]).max() / np.abs(ref["trace"]).max()
    say("amplitude", ea, "phase", ep, "pupil", epu, "trace", te)
    assert ea <= bound and ep <= bound and epu <= 1e-5 and te <= 1e-5


def test_epry_plain_c128_at_256_matches_oracle(ctx):
    """The FP64 path at 256^2 -> 729^2 (3 sweeps) within 1e-10."""
    model, meas, _ = problem(256, seed=9)
    alg = Epry(iterations=3)
    ref = oracle.run(model, meas, alg.to_c(), 3)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision="c128")
    s.initialize()
    s.run(3)
    for got, key in ((s.spectrum(), "object_spectrum"), (s.pupil(), "recovered_pupil"), (s.object(), "object")):
        e = np.linalg.norm(got - ref[key]) / np.linalg.norm(ref[key])
        say(key, e)
        assert e <= 1e-10
