"""Batched form with its defining input: per-tile aberrated (Noll Z4-Z15)
pupils. Tiles of bench.py's own 96-tile uneven batch (13x13 LEDs, 64^2 /
128^2 / 256^2, uint16 Poisson counts at a 4e3 peak), each simulated and
reconstructed with its own complex pupil, run through the fused sweep in
one batch (the 256^2 tile as a cooperative group) for the configuration's
40 FP32 mPIE sweeps, and checked against the FP64 oracle on amplitude and
phase (relative L2 <= 1e-5)."""
import os
import sys
import threading

import numpy as np
import pytest

import oracle
from fpmhip import runtime

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the measured workloads' constructions)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def amp_phase_err(got, ref):
    ea = np.linalg.norm(np.abs(got) - np.abs(ref)) / np.linalg.norm(np.abs(ref))
    ep = np.linalg.norm(np.angle(got * np.conj(ref))) / np.linalg.norm(np.angle(ref))
    return ea, ep


def oracle_runs(probs, alg, iters):
    """The oracle on each problem, one host thread each (the oracle releases
    the GIL inside its C calls)."""
    out = [None] * len(probs)

    def work(i):
        out[i] = oracle.run(probs[i].model, probs[i].measurements, alg.to_c(), iters)

    th = [threading.Thread(target=work, args=(i,)) for i in range(len(probs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    return out


def test_uneven_tiles_with_zernike_pupils_match_oracle_full_length(ctx):
    # tile 0: 256^2, tile 16: 128^2, tile 64: 64^2 of the uneven batch
    probs = bench.uneven_tiles(only={0, 16, 64})
    for p in probs:
        P = p.model.pupil[p.model.support.astype(bool)]
        assert np.abs(np.angle(P)).max() > 0.05  # the pupil is complex, not the 0/1 support
    say("config-4 tiles", [p.model.image_shape for p in probs])
    alg = bench.algorithm(40)
    sessions = [runtime.Session(p, alg, precision="f32") for p in probs]
    b = runtime.Batch(sessions)
    trace = b.run(40)
    refs = oracle_runs(probs, alg, 40)
    for p, s, tr, ref in zip(probs, sessions, trace, refs):
        ea, ep = amp_phase_err(s.object(), ref["object"])
        te = np.abs(tr - ref["trace"]).max() / np.abs(ref["trace"]).max()
        say(p.model.image_shape, "amplitude", ea, "phase", ep, "trace", te)
        assert ea <= 1e-5 and ep <= 1e-5 and te <= 1e-5
    b.close()


def test_uneven_tile_with_zernike_pupil_is_identical_alone_and_in_mix(ctx):
    """A complex-pupil tile's spectrum is bitwise the same in a mixed batch of
    all three sizes as alone (3 sweeps)."""
    probs = bench.uneven_tiles(only={1, 17, 65})
    alg = bench.algorithm(3)
    mix = [runtime.Session(p, alg, precision="f32") for p in probs]
    bm = runtime.Batch(mix)
    bm.run(3)
    for i, p in enumerate(probs):
        alone = runtime.Session(p, alg, precision="f32")
        ba = runtime.Batch([alone])
        ba.run(3)
        assert np.array_equal(alone.spectrum(), mix[i].spectrum()), p.model.image_shape
        ba.close()
    bm.close()
