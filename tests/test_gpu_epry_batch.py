"""EPRY in the batched form: the fused sweep with per-LED pupil recovery
(common.rs:271-300) and the device gauge after every sweep (gauge.rs:20-296)
against the FP64 oracle, on the precision-form problem (225 LEDs, true
pupil with Noll Z4-Z15 aberrations at 0.5 rad RMS, reconstruction from the
compiled circular pupil). Amplitude / phase relative L2 <= 1e-5."""
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


def problem(n, seed=7, k=15):
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(k), (n, n))
    obj = syn.random_field_object(model.reconstruction_shape, seed)
    true_pupil = syn.zernike_pupil(model, 0.5, seed + 100)
    I = runtime.forward_stack(model, syn.object_spectrum(obj), true_pupil)
    counts, ppp = syn.poisson_counts(I, 1e4, seed)
    model.with_frame_gains(np.full(model.frame_count, ppp))
    return ReconstructionProblem(MeasurementStack(counts=counts), model)


def errors(got, ref):
    ea = np.linalg.norm(np.abs(got) - np.abs(ref)) / np.linalg.norm(np.abs(ref))
    ep = np.linalg.norm(np.angle(got * np.conj(ref))) / np.linalg.norm(np.angle(ref))
    return ea, ep


@pytest.mark.parametrize("n,seed", [(64, 7), (64, 8), (128, 7), (128, 8)])
def test_batched_epry_matches_fp64_oracle(ctx, n, seed):
    """30 sweeps, the tile beside a second one (seed + 10) in the batch."""
    iters = 30
    probs = [problem(n, seed=seed), problem(n, seed=seed + 10)]
    alg = Epry(iterations=iters)
    say("batched EPRY", n, seed, probs[0].model.reconstruction_shape, iters)
    ss = [runtime.Session(p, alg, precision="f32") for p in probs]
    b = runtime.Batch(ss)
    trace = b.run(iters)
    p, s, tr = probs[0], ss[0], trace[0]
    ref = oracle.run(p.model, p.measurements, alg.to_c(), iters)
    ea, ep = errors(s.object(), ref["object"])
    epu = np.linalg.norm(s.pupil() - ref["recovered_pupil"]) / np.linalg.norm(ref["recovered_pupil"])
    te = np.abs(tr - ref["trace"]).max() / np.abs(ref["trace"]).max()
    say("amplitude", ea, "phase", ep, "pupil", epu, "trace", te)
    b.close()
    assert ea <= 1e-5 and ep <= 1e-5 and epu <= 1e-5 and te <= 1e-5


def test_batched_epry_keeps_the_gauge_and_is_identical_alone_and_in_mix(ctx):
    """After every sweep the device gauge holds the reference's invariants
    (reconstruction_tests.rs:2674): supported pupil energy equal to the
    compiled pupil's, overlap with it real and positive, spectrum DC real
    and non-negative; and a tile is bitwise the same in a batch of three as
    alone."""
    probs = [problem(64, seed=s, k=9) for s in (21, 22, 23)]
    alg = Epry(iterations=4)
    mix = [runtime.Session(p, alg, precision="f32") for p in probs]
    bm = runtime.Batch(mix)
    bm.run(4)
    for p, s in zip(probs, mix):
        P, ref, sup = s.pupil(), p.model.pupil, p.model.support.astype(bool)
        e, e0 = np.sum(np.abs(P[sup]) ** 2), np.sum(np.abs(ref[sup]) ** 2)
        ov = np.sum(np.conj(ref[sup]) * P[sup])
        spec = s.spectrum()
        dc = spec[spec.shape[0] // 2, spec.shape[1] // 2]
        say("energy", e / e0 - 1, "overlap phase", np.angle(ov), "dc phase", np.angle(dc))
        assert abs(e / e0 - 1) <= 1e-6
        assert abs(np.angle(ov)) <= 1e-6 and ov.real > 0
        assert abs(np.angle(dc)) <= 1e-12 and dc.real >= 0
    alone = runtime.Session(probs[1], alg, precision="f32")
    ba = runtime.Batch([alone])
    ba.run(4)
    assert np.array_equal(alone.spectrum(), mix[1].spectrum())
    assert np.array_equal(alone.pupil(), mix[1].pupil())
    ba.close()
    bm.close()
