"""Serving form: continuous batching. Jobs of mixed sizes arrive while
others run; each result equals the same job run alone (bitwise) and the
FP64 oracle within 1e-5; admission honours the tile-slot limit."""
import sys
import time

import numpy as np
import pytest

import oracle
from fpmhip import MeasurementStack, Mpie, ReconstructionProblem, runtime
from fpmhip import synthetic as syn

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def job(n, k, seed, wavelength=532e-9, peak=4e3):
    model = runtime.compile_model(syn.base_optics(wavelength), syn.planar_array(k), (n, n))
    obj = syn.random_field_object(model.reconstruction_shape, seed)
    I = runtime.forward_stack(model, syn.object_spectrum(obj))
    counts, ppp = syn.poisson_counts(I, peak, seed)
    model.with_frame_gains(np.full(model.frame_count, ppp))
    return ReconstructionProblem(MeasurementStack(counts=counts), model)


ALG = Mpie(iterations=4, momentum_interval=30, momentum_friction=0.5, momentum_feedback=0.5)


def test_result_is_identical_alone_and_in_the_mix(ctx):
    jobs = [job(64, 9, 1, 450e-9), job(128, 7, 2), job(256, 9, 6, 630e-9), job(64, 11, 3, 630e-9), job(32, 5, 4)]
    say("serving mix", [p.model.reconstruction_shape for p in jobs])
    alone = []
    for p in jobs:
        srv = runtime.Server(max_tiles=5)
        jid = srv.submit(p, ALG)
        srv.wait(jid, 120)
        alone.append(srv.result(jid))
        srv.close()
    srv = runtime.Server(max_tiles=5)
    ids = []
    for i, p in enumerate(jobs):
        ids.append(srv.submit(p, ALG))
        time.sleep(0.002 * i)  # staggered arrivals: later jobs join running ticks
    for jid in ids:
        srv.wait(jid, 120)
    mixed = [srv.result(j) for j in ids]
    st = srv.stats()
    srv.close()
    say("stats", st)
    assert st["completed"] == 5 and st["failed"] == 0
    for a, b in zip(alone, mixed):
        assert np.array_equal(a[1], b[1])
        assert np.array_equal(a[2], b[2])


@pytest.mark.parametrize("n", [64, 256])
def test_served_job_matches_fp64_oracle(ctx, n):
    p = job(n, 9, 5)
    srv = runtime.Server(max_tiles=2)
    jid = srv.submit(p, ALG)
    srv.wait(jid, 120)
    obj, spec, trace, lat, q = srv.result(jid)
    srv.close()
    ref = oracle.run(p.model, p.measurements, ALG.to_c(), 4)
    e = np.linalg.norm(obj - ref["object"]) / np.linalg.norm(ref["object"])
    say("served vs oracle", e, "latency", lat, "queue", q)
    assert e <= 1e-5
    assert np.abs(trace - ref["trace"]).max() / np.abs(ref["trace"]).max() <= 1e-5
    assert lat >= q >= 0


def test_admission_respects_slots_and_reports_failures(ctx):
    srv = runtime.Server(max_tiles=2)
    ids = [srv.submit(job(32, 5, 10 + i), ALG) for i in range(5)]
    seen_max = 0
    for _ in range(2000):
        st = srv.stats()
        seen_max = max(seen_max, st["resident_tiles"])
        if st["completed"] + st["failed"] == 5:
            break
        time.sleep(0.001)
    for jid in ids:
        srv.wait(jid, 120)
        srv.result(jid)
    assert seen_max <= 2
    # an unsupported job fails cleanly with a typed error
    p = job(32, 5, 21)
    from fpmhip import Epry
    jid = srv.submit(p, Epry(iterations=2))
    srv.wait(jid, 60)
    with pytest.raises(runtime.FpmError) as e:
        srv.result(jid)
    assert e.value.kind == "Unsupported"
    srv.close()


def test_small_job_is_not_paced_by_a_large_one(ctx):
    """Sweep quotas: alone, a job runs two sweeps per tick (a tick lasts about
    two sweeps of its heaviest tile); next to a 256^2 job, a 32^2 job (about
    13x less work per sweep: 25 LED updates of a 32^2 tile against 81 of a
    256^2 group) runs all of its sweeps in one tick, finishes before the
    large job, and gives the bitwise result it gives alone."""
    alg = Mpie(iterations=8, momentum_interval=30, momentum_friction=0.5, momentum_feedback=0.5)
    big, small = job(256, 9, 81), job(32, 5, 82)
    srv = runtime.Server(max_tiles=2)
    alone_id = srv.submit(small, alg)
    srv.wait(alone_id, 120)
    alone = srv.result(alone_id)
    ticks_alone = srv.stats()["ticks"]
    srv.close()
    srv = runtime.Server(max_tiles=2)
    ib = srv.submit(big, alg)
    i_s = srv.submit(small, alg)
    srv.wait(ib, 120)
    srv.wait(i_s, 120)
    st = srv.stats()
    rs = srv.result(i_s)
    rb = srv.result(ib)
    srv.close()
    say("ticks alone", ticks_alone, "mixed", st["ticks"], "latency small", rs[3], "big", rb[3])
    ticks = -(-alg.iterations // 2)
    assert ticks_alone == ticks
    assert st["ticks"] <= ticks + 1
    assert rs[3] < rb[3]
    assert np.array_equal(alone[1], rs[1]) and np.array_equal(alone[2], rs[2])


def test_pooled_buffers_are_reused_and_results_unchanged(ctx):
    """A served job's device buffers come from the server's memory pool.
    Served a second time, after the first runs have finished, the same jobs
    allocate nothing new: every buffer is a block of the first runs, stale
    contents and all. The objects, spectra and traces are bitwise those of
    the first runs."""
    jobs = [job(128, 9, 21), job(256, 7, 22, 630e-9)]
    srv = runtime.Server(max_tiles=4)

    def serve_each():
        out = []
        for p in jobs:
            jid = srv.submit(p, ALG)
            srv.wait(jid, 120)
            out.append(srv.result(jid))
        return out, srv.stats()

    first, st1 = serve_each()
    again, st2 = serve_each()
    srv.close()
    say("pool after the first runs", st1, "after the second", st2)
    assert st2["completed"] == 4 and st2["failed"] == 0
    assert st1["pool_allocations"] > 0
    assert st2["pool_allocations"] == st1["pool_allocations"]
    assert st2["pool_reuses"] - st1["pool_reuses"] == st1["pool_allocations"] + st1["pool_reuses"]
    for x, y in zip(first, again):
        for k in range(3):
            assert np.array_equal(x[k], y[k])


def test_admission_keeps_large_tiles_as_cooperative_groups(ctx):
    """A 256^2 tile is admitted as a cooperative group of 8 workgroups: with
    every workgroup of a tick held on the device at once (one k_sweep
    workgroup per CU, 256 CUs on an MI355X), no tick of 40 queued 256^2 jobs
    launches more than 256 workgroups (32 groups; the largest tick's grid is
    whole groups), and all jobs complete. Jobs that a tick completes make room
    for the next tick's admissions while their results are downloaded, so
    more than 32 jobs can hold device state at one time."""
    probs = [job(256, 3, 200 + i) for i in range(40)]
    srv = runtime.Server(max_tiles=256)
    ids = [srv.submit(p, ALG) for p in probs]
    for jid in ids:
        srv.wait(jid, 120)
    st = srv.stats()
    first = srv.result(ids[0])
    srv.close()
    say("stats", st)
    assert st["completed"] == len(ids) and st["failed"] == 0
    assert 0 < st["peak_tick_workgroups"] <= 256 and st["peak_tick_workgroups"] % 8 == 0
    assert np.isfinite(first[1]).all()


def test_served_job_with_aberrated_pupil_matches_fp64_oracle(ctx):
    """A job whose model pupil carries Noll Z4-Z15 aberrations (0.6 rad RMS),
    simulated and reconstructed with it, served beside a plain-pupil job:
    within 1e-5 of the oracle, and bitwise the same as served alone."""
    p = job(128, 9, 8)
    p.model.pupil = syn.zernike_pupil(p.model, 0.6, 77)
    obj = syn.random_field_object(p.model.reconstruction_shape, 8)
    I = runtime.forward_stack(p.model, syn.object_spectrum(obj))
    counts, ppp = syn.poisson_counts(I, 4e3, 8)
    p.model.with_frame_gains(np.full(p.model.frame_count, ppp))
    p = ReconstructionProblem(MeasurementStack(counts=counts), p.model)
    other = job(64, 7, 9)
    srv = runtime.Server(max_tiles=2)
    jid = srv.submit(p, ALG)
    oid = srv.submit(other, ALG)
    srv.wait(jid, 120)
    srv.wait(oid, 120)
    obj_mix, spec_mix, trace, _, _ = srv.result(jid)
    srv.result(oid)
    srv.close()
    srv = runtime.Server(max_tiles=1)
    jid = srv.submit(p, ALG)
    srv.wait(jid, 120)
    _, spec_alone, _, _, _ = srv.result(jid)
    srv.close()
    ref = oracle.run(p.model, p.measurements, ALG.to_c(), 4)
    e = np.linalg.norm(obj_mix - ref["object"]) / np.linalg.norm(ref["object"])
    say("aberrated served job vs oracle", e)
    assert e <= 1e-5
    assert np.abs(trace - ref["trace"]).max() / np.abs(ref["trace"]).max() <= 1e-5
    assert np.array_equal(spec_mix, spec_alone)
