"""Batched form: uneven tiles advanced by one fused sweep launch per
iteration. Each tile is checked against the FP64 oracle (amplitude and phase
relative L2) and is bitwise the same alone and in a mixed batch."""
import sys

import numpy as np
import pytest

import oracle
from fpmhip import MeasurementStack, Mpie, Fpie, AlternatingProjection, ReconstructionProblem, runtime
from fpmhip import synthetic as syn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def tile(n, k, seed, peak=4e3):
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(k), (n, n))
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


@pytest.mark.parametrize("alg", [Mpie(iterations=3, momentum_interval=30, momentum_friction=0.5,
                                      momentum_feedback=0.5),
                                 Fpie(iterations=3), AlternatingProjection(iterations=3)])
def test_batch_matches_fp64_oracle(ctx, alg):
    tiles = [tile(64, 9, 11), tile(128, 7, 12), tile(32, 5, 13)]
    say("batch", type(alg).__name__, [t[0].reconstruction_shape for t in tiles])
    ss = sessions_for(tiles, alg)
    b = runtime.Batch(ss)
    trace = b.run(3)
    for (model, meas), s, tr in zip(tiles, ss, trace):
        ref = oracle.run(model, meas, alg.to_c(), 3)
        ea, ep = amp_phase_err(s.object(), ref["object"])
        es = np.linalg.norm(s.spectrum() - ref["object_spectrum"]) / np.linalg.norm(ref["object_spectrum"])
        te = np.abs(tr - ref["trace"]) / np.abs(ref["trace"])
        say(model.image_shape, "amp", ea, "phase", ep, "spec", es, "trace", te.max())
        assert ea <= 1e-5 and ep <= 1e-5 and es <= 1e-5
        assert te.max() <= 1e-5


def test_batch_tile_is_identical_alone_and_in_mix(ctx):
    alg = Mpie(iterations=2, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(64, 9, 21), tile(128, 7, 22), tile(64, 11, 23)]
    mix = sessions_for(tiles, alg)
    runtime.Batch(mix).run(2)
    alone = sessions_for(tiles[1:2], alg)
    runtime.Batch(alone).run(2)
    assert np.array_equal(mix[1].spectrum(), alone[0].spectrum())


def test_batch_repeatable_and_graph_replays_match(ctx):
    """Two runs of the same batch (graph replays with eager work between them)
    give bitwise identical spectra."""
    alg = Mpie(iterations=3, momentum_interval=7, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(64, 9, 31), tile(32, 5, 32)]
    ss = sessions_for(tiles, alg)
    b = runtime.Batch(ss)
    b.run(3)
    first = [s.spectrum() for s in ss]
    b.run(3)
    second = [s.spectrum() for s in ss]
    for a, c in zip(first, second):
        assert np.array_equal(a, c)
    mean, mx, n = b.timing()
    assert n == 3 and mean > 0


def test_batch_vs_per_stage_mixed_session(ctx):
    """The fused sweep and the per-stage MIXED_F32 path agree to FP32 level."""
    alg = Mpie(iterations=2, momentum_friction=0.5, momentum_feedback=0.5)
    model, meas = tile(128, 15, 41, peak=1e4)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision="f32")
    runtime.Batch([s]).run(2)
    p = runtime.Session(ReconstructionProblem(meas, model), alg, precision="f32")
    p.initialize()
    p.run(2)
    e = np.linalg.norm(s.spectrum() - p.spectrum()) / np.linalg.norm(p.spectrum())
    say("fused vs per-stage", e)
    assert e <= 1e-5


def test_batch_with_256_tile_matches_fp64_oracle(ctx):
    """A 256^2 tile (field in the tile's global scratch, chunked through LDS)
    next to LDS-resident tiles: mPIE with momentum events on 13x13 LEDs
    (the uneven-batch shape), each tile against the FP64 oracle."""
    alg = Mpie(iterations=3, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(64, 9, 51), tile(256, 13, 52), tile(128, 7, 53)]
    say("batch with 256", [t[0].reconstruction_shape for t in tiles])
    ss = sessions_for(tiles, alg)
    trace = runtime.Batch(ss).run(3)
    for (model, meas), s, tr in zip(tiles, ss, trace):
        ref = oracle.run(model, meas, alg.to_c(), 3)
        ea, ep = amp_phase_err(s.object(), ref["object"])
        te = np.abs(tr - ref["trace"]) / np.abs(ref["trace"])
        say(model.image_shape, "amp", ea, "phase", ep, "trace", te.max())
        assert ea <= 1e-5 and ep <= 1e-5 and te.max() <= 1e-5


def test_batch_256_masks_background_weights(ctx):
    """The 256^2 path honours masks, a per-frame background, frame gains and
    a zero-weight frame like the oracle (FPIE, 2 sweeps)."""
    model, meas = tile(256, 5, 61)
    F = model.frame_count
    rng = np.random.default_rng(61)
    bg = rng.uniform(0.0, 3.0, (F,) + tuple(model.image_shape))
    model.with_background(bg)
    masks = (rng.random((F,) + tuple(model.image_shape)) > 0.05).astype(np.uint8)
    w = np.ones(F)
    w[3] = 0.0
    w[7] = 0.5
    meas = MeasurementStack(counts=meas.counts, weights=w, masks=masks)
    alg = Fpie(iterations=2)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision="f32")
    trace = runtime.Batch([s]).run(2)[0]
    ref = oracle.run(model, meas, alg.to_c(), 2)
    ea, ep = amp_phase_err(s.object(), ref["object"])
    te = np.abs(trace - ref["trace"]) / np.abs(ref["trace"])
    say("256 masked", "amp", ea, "phase", ep, "trace", te.max())
    assert ea <= 1e-5 and ep <= 1e-5 and te.max() <= 1e-5


def test_batch_256_tile_identical_alone_and_in_mix(ctx):
    alg = Mpie(iterations=2, momentum_interval=11, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(128, 7, 71), tile(256, 9, 72), tile(64, 5, 73), tile(256, 7, 74)]
    mix = sessions_for(tiles, alg)
    runtime.Batch(mix).run(2)
    for i in (1, 3):
        alone = sessions_for(tiles[i:i + 1], alg)
        runtime.Batch(alone).run(2)
        assert np.array_equal(mix[i].spectrum(), alone[0].spectrum())


@pytest.mark.parametrize("g", [8, 11, 0])
def test_cooperative_group_matches_single_workgroup(ctx, g):
    """A 256^2 tile split over a cooperative group (8 or 11 workgroups, or
    the automatic size, which takes the co-resident room left over: 32 here)
    gives the same spectrum, bit for bit, as one workgroup running all chunks
    (the per-pixel arithmetic and insert order are the same whatever the row
    chunks and column blocks); the objective sums differ only by summation
    order."""
    alg = Mpie(iterations=3, momentum_interval=20, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(256, 9, 91), tile(128, 7, 92), tile(256, 7, 93)]
    results = {}
    for gm in (g, 1):
        ss = sessions_for(tiles, alg)
        b = runtime.Batch(ss)
        b.set_schedule(group_members=gm)
        assert b.schedule()["group_members"] == gm
        tr = b.run(3)
        results[gm] = ([s.spectrum() for s in ss], tr)
        say("group", gm, "sweep ms", b.timing()[0])
        b.close()
    for a, c in zip(results[g][0], results[1][0]):
        assert np.array_equal(a, c)
    assert np.allclose(results[g][1], results[1][1], rtol=1e-12, atol=0)


@pytest.mark.parametrize("radix", [[4, 4, 8, 0], [8, 16, 16, 0]])
def test_radix_plans_match_oracle(ctx, radix):
    """Every first-stage radix plan of the LDS transforms stays within the
    precision-form bound of the FP64 oracle."""
    alg = Fpie(iterations=2)
    tiles = [tile(32, 5, 101), tile(64, 7, 102), tile(128, 7, 103)]
    ss = sessions_for(tiles, alg)
    b = runtime.Batch(ss)
    b.set_schedule(radix=radix)
    b.run(2)
    for (model, meas), s in zip(tiles, ss):
        ref = oracle.run(model, meas, alg.to_c(), 2)
        ea, ep = amp_phase_err(s.object(), ref["object"])
        say(radix, model.image_shape, "amp", ea, "phase", ep)
        assert ea <= 1e-5 and ep <= 1e-5


def test_sweeps_per_launch_do_not_change_results(ctx):
    """7 sweeps as launches of 1, 3 (3 + 3 + 1) and 10 sweeps each: the same
    spectra and traces, bit for bit (a launch runs whole sweeps; mPIE's
    counter and the trace carry across launches)."""
    alg = Mpie(iterations=7, momentum_interval=11, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(64, 7, 141), tile(128, 5, 142), tile(256, 5, 143)]
    out = []
    for per in (1, 3, 10):
        ss = sessions_for(tiles, alg)
        b = runtime.Batch(ss)
        b.set_schedule(sweeps_per_launch=per)
        assert b.schedule()["sweeps_per_launch"] == per
        tr = b.run(7)
        out.append(([s.spectrum() for s in ss], tr))
    for spec, tr in out[1:]:
        for a, c in zip(out[0][0], spec):
            assert np.array_equal(a, c)
        assert np.array_equal(out[0][1], tr)


def test_batch_objects_match_session_objects(ctx):
    """fpmhip_batch_objects returns each tile's state_object, bit for bit the
    per-session download, for tiles of different shapes."""
    alg = Fpie(iterations=2)
    tiles = [tile(64, 5, 151), tile(128, 5, 152), tile(32, 5, 153)]
    ss = sessions_for(tiles, alg)
    b = runtime.Batch(ss)
    b.run(2)
    objs = b.objects()
    pinned = b.objects(pinned=True)
    for s, o, q in zip(ss, objs, pinned):
        assert o.shape == tuple(s.problem.model.reconstruction_shape)
        assert np.array_equal(o, s.object())
        assert np.array_equal(q, o)


def test_batched_initialisation_matches_session_initialisation(ctx):
    """The batch initialises all tiles with one launch per pass and one FFT
    per distinct R; every tile's state (spectrum, pupil, velocity, anchor,
    counter) is bit for bit what Session.initialize gives, for uneven tiles
    (two 64^2 tiles around a 128^2 one, a 256^2 tile with masks, weights and
    a per-frame background, a 32^2 tile with a shared background)."""
    say("batched init vs session init")
    alg = Mpie(iterations=1, momentum_interval=11, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(64, 5, 161), tile(128, 7, 162), tile(64, 7, 163), tile(256, 5, 164), tile(32, 5, 165)]
    rng = np.random.default_rng(166)
    model, meas = tiles[3]
    F = model.frame_count
    model.with_background(rng.uniform(0.0, 3.0, (F,) + tuple(model.image_shape)))
    w = np.ones(F)
    w[2] = 0.0
    w[4] = 0.25
    masks = (rng.random((F,) + tuple(model.image_shape)) > 0.05).astype(np.uint8)
    tiles[3] = (model, MeasurementStack(counts=meas.counts, weights=w, masks=masks))
    tiles[4][0].with_background(np.full(tiles[4][0].image_shape, 1.5))
    ss = sessions_for(tiles, alg)
    b = runtime.Batch(ss)
    for rep in range(2):  # a second run re-initialises from scratch
        b.run(1 if rep else 0)
        if rep:
            b.run(0)
        for (mo, me), s in zip(tiles, ss):
            ref = runtime.Session(ReconstructionProblem(me, mo), alg, precision="f32")
            ref.initialize()
            assert np.array_equal(s.spectrum(), ref.spectrum()), mo.image_shape
            assert np.array_equal(s.pupil(), ref.pupil())
            v, a, c = s.momentum()
            rv, ra, rc = ref.momentum()
            assert np.array_equal(v, rv) and np.array_equal(a, ra) and c == rc == 0
            assert np.array_equal(s.object(), ref.object())


def test_device_objects_are_the_host_objects(ctx):
    """fpmhip_batch_objects_device leaves each tile's object in HBM: copied
    back with hipMemcpy it is the object fpmhip_batch_objects returns."""
    say("device-resident objects")
    import ctypes as ct
    hip = ct.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ct.c_int
    hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
    alg = Fpie(iterations=2)
    tiles = [tile(64, 5, 171), tile(128, 5, 172), tile(64, 5, 173)]
    ss = sessions_for(tiles, alg)
    b = runtime.Batch(ss)
    b.run(2)
    host = [o.copy() for o in b.objects()]
    dev = b.objects_device()
    assert len({p for p, _ in dev}) == len(dev) and all(p for p, _ in dev)
    for (ptr, shape), h in zip(dev, host):
        out = np.empty(shape, np.complex128)
        assert hip.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        assert np.array_equal(out, h)


def test_launch_order_does_not_change_results(ctx):
    alg = Mpie(iterations=2, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(64, 5, 111), tile(256, 5, 112), tile(32, 5, 113)]
    out = []
    for first in (1, 0):
        ss = sessions_for(tiles, alg)
        b = runtime.Batch(ss)
        b.set_schedule(largest_first=first)
        b.run(2)
        out.append([s.spectrum() for s in ss])
    for a, c in zip(*out):
        assert np.array_equal(a, c)


def test_autotuner_keeps_a_valid_schedule(ctx):
    """fpmhip_batch_tune times candidate schedules on the batch's own tiles,
    keeps one, and the batch still reconstructs within the FP64 bound."""
    alg = Mpie(iterations=2, momentum_friction=0.5, momentum_feedback=0.5)
    tiles = [tile(256, 7, 121), tile(128, 7, 122), tile(64, 7, 123), tile(32, 5, 124)]
    ss = sessions_for(tiles, alg)
    b = runtime.Batch(ss)
    rep = b.tune(1)
    say("tuning", rep)
    # the default, then group size (none, 8), radix of 32/64/128 and launch
    # order: 1 + 2 + 1 + 2 + 1 + 1 candidates
    assert rep["candidates"] == 8
    assert rep["best_ms"] <= rep["first_ms"]
    sch = rep["schedule"]
    assert sch["group_members"] in (0, 1, 8) and sch["largest_first"] in (0, 1)
    b.run(2)
    for (model, meas), s in zip(tiles, ss):
        ref = oracle.run(model, meas, alg.to_c(), 2)
        ea, ep = amp_phase_err(s.object(), ref["object"])
        assert ea <= 1e-5 and ep <= 1e-5


def test_invalid_schedule_is_rejected(ctx):
    ss = sessions_for([tile(64, 5, 131)], Fpie(iterations=1))
    b = runtime.Batch(ss)
    for kw in ({"group_members": 3}, {"group_members": 33}, {"radix": [0, 32, 0, 0]}, {"largest_first": 2},
               {"sweeps_per_launch": 1001}):
        with pytest.raises(runtime.FpmError) as e:
            b.set_schedule(**kw)
        assert e.value.kind == "InvalidParameter"


def test_bench_workload_full_length_precision(ctx):
    """The measured workload itself, at full length: 225 LEDs, 128^2 images
    (375^2 object), Poisson counts at 1e4 peak as uint16 (seed 7), mPIE 50
    sweeps with interval 30 (375 momentum events), friction/feedback 0.5.
    The batched FP32 result stays within the precision form's bound of the
    FP64 oracle run on the same inputs: amplitude and phase relative L2
    <= 1e-5, and every iteration objective of the trace within 1e-5."""
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(15), (128, 128))
    obj = syn.random_field_object(model.reconstruction_shape, 7)
    I = runtime.forward_stack(model, syn.object_spectrum(obj))
    counts, ppp = syn.poisson_counts(I, 1e4, 7)
    model.with_frame_gains(np.full(model.frame_count, ppp))
    meas = MeasurementStack(counts=counts)
    alg = Mpie(iterations=50, momentum_friction=0.5, momentum_feedback=0.5)
    s = runtime.Session(ReconstructionProblem(meas, model), alg, precision="f32")
    trace = runtime.Batch([s]).run(50)[0]
    say("bench workload: oracle, 50 sweeps")
    ref = oracle.run(model, meas, alg.to_c(), 50)
    ea, ep = amp_phase_err(s.object(), ref["object"])
    es = np.linalg.norm(s.spectrum() - ref["object_spectrum"]) / np.linalg.norm(ref["object_spectrum"])
    te = np.abs(trace - ref["trace"]) / np.abs(ref["trace"])
    say("50 sweeps: amp", ea, "phase", ep, "spec", es, "trace", te.max())
    assert ea <= 1e-5 and ep <= 1e-5 and es <= 1e-5
    assert te.max() <= 1e-5
