"""512^2 tiles in the fused batched sweep: the field in the tile's global
scratch, passed through LDS in 16-line chunks with three radix-8 stages
per line transform, on a cooperative group of 32 workgroups (or one
workgroup running every chunk). Each tile against the FP64 oracle
(amplitude / phase relative L2 <= 1e-5), group == single workgroup bit for
bit, and a tile bitwise the same alone and in a mixed batch."""
import sys

import numpy as np
import pytest

import oracle
from fpmhip import Fpie, MeasurementStack, Mpie, ReconstructionProblem, runtime
from fpmhip import synthetic as syn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tile(n, k, seed, peak=4e3, true_pupil_rms=None):
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(k), (n, n))
    if true_pupil_rms is not None:
        model.pupil = syn.zernike_pupil(model, true_pupil_rms, seed + 100)
    obj = syn.random_field_object(model.reconstruction_shape, seed)
    I = runtime.forward_stack(model, syn.object_spectrum(obj))
    counts, ppp = syn.poisson_counts(I, peak, seed)
    model.with_frame_gains(np.full(model.frame_count, ppp))
    return model, MeasurementStack(counts=counts)


def amp_phase_err(got, ref):
    ea = np.linalg.norm(np.abs(got) - np.abs(ref)) / np.linalg.norm(np.abs(ref))
    ep = np.linalg.norm(np.angle(got * np.conj(ref))) / np.linalg.norm(np.angle(ref))
    return ea, ep


def sessions_for(tiles, alg):
    return [runtime.Session(ReconstructionProblem(m, mo), alg, precision="f32") for mo, m in tiles]


def test_batch_with_512_tile_matches_fp64_oracle(ctx):
    """mPIE with momentum events, a 512^2 tile with a Zernike pupil beside
    a 128^2 and a 64^2 tile."""
    alg = Mpie(iterations=3, momentum_interval=20, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(64, 7, 31), tile(512, 5, 32, true_pupil_rms=0.4), tile(128, 5, 33)]
    say("batch with 512", [t[0].reconstruction_shape for t in tiles])
    ss = sessions_for(tiles, alg)
    b = runtime.Batch(ss)
    trace = b.run(3)
    for (model, meas), s, tr in zip(tiles, ss, trace):
        ref = oracle.run(model, meas, alg.to_c(), 3)
        ea, ep = amp_phase_err(s.object(), ref["object"])
        te = np.abs(tr - ref["trace"]) / np.abs(ref["trace"])
        say(model.image_shape, "amp", ea, "phase", ep, "trace", te.max())
        assert ea <= 1e-5 and ep <= 1e-5 and te.max() <= 1e-5
    b.close()


def test_batch_512_masks_background_weights(ctx):
    """Masks, a per-frame background, frame gains and a zero-weight frame on
    the 512^2 path (FPIE, 2 sweeps)."""
    model, meas = tile(512, 3, 41)
    F = model.frame_count
    rng = np.random.default_rng(41)
    model.with_background(rng.uniform(0.0, 3.0, (F,) + tuple(model.image_shape)))
    masks = (rng.random((F,) + tuple(model.image_shape)) > 0.05).astype(np.uint8)
    w = np.ones(F)
    w[2] = 0.0
    w[5] = 0.5
    meas = MeasurementStack(counts=meas.counts, weights=w, masks=masks)
    alg = Fpie(iterations=2)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision="f32")
    b = runtime.Batch([s])
    trace = b.run(2)[0]
    ref = oracle.run(model, meas, alg.to_c(), 2)
    ea, ep = amp_phase_err(s.object(), ref["object"])
    te = np.abs(trace - ref["trace"]) / np.abs(ref["trace"])
    say("512 masked", "amp", ea, "phase", ep, "trace", te.max())
    assert ea <= 1e-5 and ep <= 1e-5 and te.max() <= 1e-5
    b.close()


def test_512_group_matches_single_workgroup_and_mix(ctx):
    """The group of 32 equals one workgroup running all 32 chunks bit for
    bit, and the tile is the same alone and beside other tiles."""
    alg = Mpie(iterations=2, momentum_interval=11, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(512, 3, 51), tile(128, 5, 52), tile(256, 3, 53)]
    results = {}
    for g in (8, 1):
        ss = sessions_for(tiles, alg)
        b = runtime.Batch(ss)
        b.set_schedule(group_members=g)
        tr = b.run(2)
        results[g] = ([s.spectrum() for s in ss], tr)
        say("group setting", g, "sweep ms", b.timing()[0])
        b.close()
    for a, c in zip(results[8][0], results[1][0]):
        assert np.array_equal(a, c)
    assert np.allclose(results[8][1], results[1][1], rtol=1e-12, atol=0)
    alone = sessions_for(tiles[:1], alg)
    ba = runtime.Batch(alone)
    ba.run(2)
    assert np.array_equal(alone[0].spectrum(), results[8][0][0])
    ba.close()
