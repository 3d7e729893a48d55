"""Throughput of fpm-rs's mPIE reconstruction on one MI355X.

Workload (one "step" = one complete reconstruction of a batch of tiles:
init, 50 mPIE sweeps over 225 LEDs, result objects complete in HBM): 15x15 planar LED array, 4 mm pitch, 90 mm,
532 nm, NA 0.10, 4x, 6.5 um pixels; 128^2 images -> 375^2 ('smooth');
object = smoothed Gaussian random field (sigma 2 px, seed 7) mapped to
amplitude [0.3, 1] and phase [-pi, pi]; Poisson counts at 1e4 peak photons
stored as uint16; mPIE with the default step 0.2, stability 0.05 and
interval 30, friction/feedback 0.5 (fpm-rs's 0.9/0.9 diverges on 225 LEDs).

Prints one JSON line (rank 0). Only the cpu_baseline leg uses the oracle.
"""
import argparse
import copy
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fpm-rs_amd"))

import numpy as np  # noqa: E402

from fpmhip import Epry, MeasurementStack, Mpie, ReconstructionProblem, abi, runtime  # noqa: E402
from fpmhip import synthetic as syn  # noqa: E402

HBM_PEAK_GBS = 8000.0
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_sweep_mpie_225led_128px_256tiles.md")


def measured_traffic():
    """HBM bytes per sweep of every tile (k_sweep dispatch bytes / the sweeps
    the dispatch ran) from the committed rocprofv3 counter summary, when it
    was taken with this exact library build (else None)."""
    import hashlib

    try:
        lib = open(runtime.LIB_PATH, "rb").read()
        digest = hashlib.sha256(lib).hexdigest()
        for line in open(TRAFFIC_PROFILE):
            if line.startswith("traffic_bytes_per_sweep:"):
                parts = line.split()
                if parts[3] == digest:
                    return float(parts[1])
    except (OSError, IndexError, ValueError):
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(n=128, k=15, seed=7, peak=1e4):
    model = runtime.compile_model(syn.base_optics(), syn.planar_array(k), (n, n))
    obj = syn.random_field_object(model.reconstruction_shape, seed)
    I = runtime.forward_stack(model, syn.object_spectrum(obj))
    counts, ppp = syn.poisson_counts(I, peak, seed)
    model.with_frame_gains(np.full(model.frame_count, ppp))
    return model, MeasurementStack(counts=counts)


def workload_tiles(count, n=128, k=15, seed=7, peak=1e4, true_pupil_rms=None):
    """`count` distinct FOV tiles of the workload: one compiled geometry, and
    per tile its own object (seed + i), its own Poisson counts and its own
    frame gains (photons per pixel of its own peak). With true_pupil_rms the
    data are simulated through a Noll Z4-Z15 pupil of that RMS (seed + 100;
    the precision form), and reconstructed from the compiled pupil."""
    from concurrent.futures import ThreadPoolExecutor

    base = runtime.compile_model(syn.base_optics(), syn.planar_array(k), (n, n))
    ctx = runtime.default_context()
    dm = runtime.DeviceModel(ctx, base)
    true_pupil = None if true_pupil_rms is None else syn.zernike_pupil(base, true_pupil_rms, seed + 100)

    def counts_of(args):
        i, intensity = args
        counts, ppp = syn.poisson_counts(intensity, peak, seed + i)
        model = copy.deepcopy(base)
        model.with_frame_gains(np.full(model.frame_count, ppp))
        return model, MeasurementStack(counts=counts)

    try:
        # numpy's generators release the GIL: objects and counts are drawn on
        # host threads, the ideal intensities come from the device forward stack
        with ThreadPoolExecutor(host_threads()) as pool:
            specs = pool.map(lambda i: syn.object_spectrum(syn.random_field_object(base.reconstruction_shape, seed + i)),
                             range(count))
            stacks = ((i, runtime.forward_stack(base, sp, true_pupil, context=ctx, dmodel=dm))
                      for i, sp in enumerate(specs))
            return list(pool.map(counts_of, stacks))
    finally:
        dm.close()


def algorithm(iterations):
    return Mpie(iterations=iterations, momentum_friction=0.5, momentum_feedback=0.5)


def tile_bytes(model, iterations, interval, s_m, s_o, s_p, s_v):
    """Algorithmic HBM bytes of one tile: per LED the frame, the object
    window read and written and the pupil; per momentum event the object
    read and written (s_o each) and the velocity and anchor read and written
    at their stored size s_v; the init read of every frame and the R^2
    init/result passes."""
    N = model.image_len
    M = model.reconstruction_shape[0] * model.reconstruction_shape[1]
    F = model.frame_count
    b_led = N * (s_m + 2 * s_o + s_p)
    events = (iterations * F) // interval
    return iterations * F * b_led + events * M * (2 * s_o + 4 * s_v) + F * N * s_m + 4 * M * s_o


def sweep_bytes(model, iterations, interval, s_m, s_o, s_p, s_v):
    """tile_bytes of one sweep (the per-LED and momentum terms only)."""
    M = model.reconstruction_shape[0] * model.reconstruction_shape[1]
    fixed = model.frame_count * model.image_len * s_m + 4 * M * s_o
    return (tile_bytes(model, iterations, interval, s_m, s_o, s_p, s_v) - fixed) / iterations


def close_all(*objs):
    """Destroys device handles in the order given (dependents first)."""
    for o in objs:
        for x in (o if isinstance(o, (list, tuple)) else [o]):
            if x is not None:
                x.close()


def run_plain(model, meas, iterations, steps, warmup, precision):
    prob = ReconstructionProblem(meas, model)
    ctx = runtime.default_context()
    dm = runtime.DeviceModel(ctx, model)
    dme = runtime.DeviceMeasurements(ctx, dm, meas)
    alg = algorithm(iterations)

    def one():
        s = runtime.Session(prob, alg, precision=precision, context=ctx, dmodel=dm, dmeas=dme)
        try:
            s.initialize()
            s.run(iterations)
            return s.object()
        finally:
            s.close()

    try:
        for i in range(warmup):
            one()
            log(f"warmup tile {i + 1}/{warmup} done")
        ctx.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            one()
            log(f"tile {i + 1}/{steps} done")
        ctx.synchronize()
        return time.perf_counter() - t0
    finally:
        close_all(dme, dm)


def run_batched(tiles, iterations, steps, warmup, alg=None):
    """The workload tiles (each its own model, counts and state) advanced
    together; one step = one batch of tiles."""
    ctx = runtime.default_context()
    alg = alg or algorithm(iterations)
    dms, dmes, sessions, batch = [], [], [], None
    try:
        for model, meas in tiles:
            dms.append(runtime.DeviceModel(ctx, model))
            dmes.append(runtime.DeviceMeasurements(ctx, dms[-1], meas))
            sessions.append(runtime.Session(ReconstructionProblem(meas, model), alg, precision="f32", context=ctx,
                                            dmodel=dms[-1], dmeas=dmes[-1]))
        batch = runtime.Batch(sessions, context=ctx)
        for i in range(warmup):
            batch.run(iterations)
            log(f"warmup batch {i + 1}/{warmup} done")
        ctx.synchronize()
        t0 = time.perf_counter()
        sweep_ms = []
        for i in range(steps):
            batch.run(iterations)
            batch.objects_device()  # every tile's result object, complete in HBM
            sweep_ms.append(batch.timing()[0])
            log(f"batch {i + 1}/{steps} done")
        ctx.synchronize()
        secs = time.perf_counter() - t0
        # EPRY batches run one sweep per launch (the gauge follows each sweep)
        per_launch = 1 if isinstance(alg, Epry) else (batch.schedule()["sweeps_per_launch"] or 50)
        return secs, float(np.mean(sweep_ms)), min(per_launch, iterations), download_ms(batch)
    finally:
        close_all(batch, sessions, dmes, dms)


def download_ms(batch):
    """Extra wall time of taking the batch's result objects to page-locked
    host memory over PCIe instead of leaving them in HBM (best of 3)."""
    dev, host = [], []
    for _ in range(3):
        t = time.perf_counter()
        batch.objects_device()
        dev.append(time.perf_counter() - t)
        t = time.perf_counter()
        batch.objects(pinned=True)
        host.append(time.perf_counter() - t)
    return 1000.0 * max(min(host) - min(dev), 0.0)


def serving(requests, rate, iterations, seed=5, max_tiles=256):
    """Poisson arrivals (mean `rate` req/s) of jobs with tile size in
    {64^2, 128^2, 256^2}, 1-3 channels from {450, 532, 630} nm (each channel
    one tile with its own compiled pupil), k x k LEDs with k in 9..15; FP32
    mPIE `iterations` sweeps. Inputs are built before the arrival clock
    starts. Returns latency percentiles and throughput."""
    rng = np.random.default_rng(seed)
    wavelengths = (450e-9, 532e-9, 630e-9)
    pool = {}

    def problem(n, k, wl):
        key = (n, k, wl)
        if key not in pool:
            model = runtime.compile_model(syn.base_optics(wl), syn.planar_array(k), (n, n))
            obj = syn.random_field_object(model.reconstruction_shape, int(rng.integers(1 << 30)))
            I = runtime.forward_stack(model, syn.object_spectrum(obj))
            counts, ppp = syn.poisson_counts(I, 4e3, int(rng.integers(1 << 30)))
            model.with_frame_gains(np.full(model.frame_count, ppp))
            pool[key] = ReconstructionProblem(MeasurementStack(counts=counts), model)
        return pool[key]

    reqs = []
    t = 0.0
    for _ in range(requests):
        t += rng.exponential(1.0 / rate)
        n = int(rng.choice([64, 128, 256]))
        k = int(rng.integers(9, 16))
        ch = rng.choice(3, size=int(rng.integers(1, 4)), replace=False)
        reqs.append((t, [problem(n, k, wavelengths[c]) for c in sorted(ch)]))
    alg = Mpie(iterations=iterations, momentum_friction=0.5, momentum_feedback=0.5)
    server = runtime.Server(max_tiles=max_tiles)
    try:
        # one warm-up job so that first-use costs stay out of the measurement
        w = server.submit(reqs[0][1][0], alg)
        server.wait(w, 120)
        server.result(w)
        log(f"serving: {requests} requests, {sum(len(r[1]) for r in reqs)} tiles, rate {rate}/s")
        ids = []
        t0 = time.perf_counter()
        for (ta, probs) in reqs:
            delay = t0 + ta - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            ids.append([server.submit(p, alg) for p in probs])
        lat = []
        for group in ids:
            worst = 0.0
            for jid in group:
                server.wait(jid, 600)
                _, _, _, l, _ = server.result(jid)
                worst = max(worst, l)
            lat.append(worst)
        stats = server.stats()
        wall = time.perf_counter() - t0
    finally:
        server.close()
    tiles = sum(len(g) for g in ids)
    lat = np.array(lat) * 1000.0
    return {"requests": requests, "tiles": tiles, "rate_per_s": rate, "iterations": iterations,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "max_ms": float(lat.max()), "tiles_per_s": tiles / wall, "ticks": stats["ticks"],
            "pool_reuses": stats["pool_reuses"], "pool_allocations": stats["pool_allocations"],
            "sizes": "64^2,128^2,256^2", "channels": "1-3 of 450/532/630 nm", "leds": "9x9..15x15"}


def uneven_tiles(seed=11, only=None):
    """96 FOV tiles of uneven size: 32 x 64^2, 48 x 128^2, 16 x 256^2 images
    (-> 168^2, 336^2, 672^2 'smooth'), 13x13 LEDs, each tile with its own
    object (seed) and its own Zernike Z4-Z15 pupil with RMS ~ U(0, 0.8] rad
    (simulated with it and reconstructed with it); uint16 Poisson counts at
    a 4e3 photon peak. Tiles 0-15 are 256^2, 16-63 128^2, 64-95 64^2.
    `only`: indices of the tiles to build (the others' random draws are
    still taken, so a tile is the same whichever subset is built)."""
    rng = np.random.default_rng(seed)
    out = []
    compiled = {}
    index = 0
    for n, count in ((256, 16), (128, 48), (64, 32)):
        for _ in range(count):
            rms = float(rng.uniform(1e-3, 0.8))
            pupil_seed, obj_seed, count_seed = (int(rng.integers(1 << 30)) for _ in range(3))
            if only is None or index in only:
                if n not in compiled:
                    compiled[n] = runtime.compile_model(syn.base_optics(), syn.planar_array(13), (n, n))
                model = copy.deepcopy(compiled[n])
                model.pupil = syn.zernike_pupil(model, rms, pupil_seed)
                obj = syn.random_field_object(model.reconstruction_shape, obj_seed)
                I = runtime.forward_stack(model, syn.object_spectrum(obj))
                counts, ppp = syn.poisson_counts(I, 4e3, count_seed)
                model.with_frame_gains(np.full(model.frame_count, ppp))
                out.append(ReconstructionProblem(MeasurementStack(counts=counts), model))
            index += 1
    return out


def run_uneven(iterations=40, steps=1, warmup=1):
    """The uneven 96-tile batch, FP32 mPIE `iterations` sweeps, replayed
    from the batch's captured graph: one persistent fused launch runs the
    schedule's sweeps_per_launch (default 50) sweeps of every tile. The
    schedule (group size, radix per tile size, launch order) is the one the
    batch's autotuner picks, before the timed region. Returns tiles/s, the
    sweep-kernel figures, the tuning and the algorithmic bytes."""
    log("building the uneven 96-tile batch")
    probs = uneven_tiles()
    ctx = runtime.default_context()
    alg = algorithm(iterations)
    sessions, batch = [], None
    try:
        sessions = [runtime.Session(p, alg, precision="f32", context=ctx) for p in probs]
        batch = runtime.Batch(sessions, context=ctx)
        tuning = batch.tune(3)
        log(f"uneven batch schedule: {tuning}")
        for i in range(warmup):
            batch.run(iterations)
        ctx.synchronize()
        t0 = time.perf_counter()
        sweep_ms = []
        for i in range(steps):
            batch.run(iterations)
            batch.objects_device()  # every tile's result object, complete in HBM
            sweep_ms.append(batch.timing()[0])
            log(f"uneven batch {i + 1}/{steps} done")
        ctx.synchronize()
        secs = time.perf_counter() - t0
        dl = download_ms(batch)
    finally:
        close_all(batch, sessions, [s.dmeas for s in sessions], [s.dmodel for s in sessions])
    total = sum(tile_bytes(p.model, iterations, 30, 2, 16, 8, 8) for p in probs)
    per_sweep = sum(sweep_bytes(p.model, iterations, 30, 2, 16, 8, 8) for p in probs)
    ms = float(np.mean(sweep_ms))
    return {"tiles": len(probs), "sizes": "32x64^2, 48x128^2, 16x256^2", "leds": "13x13", "iterations": iterations,
            "tiles_per_s": steps * len(probs) / secs, "ms_per_batch": 1000.0 * secs / steps,
            "pcie_inclusive_tiles_per_s": len(probs) / (secs / steps + dl / 1000.0), "result_download_ms": dl,
            "sweep_ms": ms, "sweep_achieved_GBs": per_sweep / (ms * 1e-3) / 1e9, "tuning": tuning,
            "sweep_frac": per_sweep / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "batch_algorithmic_GB": total / 1e9}


def run_epry(tiles=256, iterations=30, steps=1, warmup=1):
    """The precision form in the batched kernel: `tiles` distinct 225-LED
    128^2 tiles simulated through a 0.5 rad RMS Noll Z4-Z15 pupil, EPRY
    (object step 0.8, pupil step 0.1, support constraint) `iterations`
    sweeps from the compiled pupil, with the device gauge after every sweep.
    Roofline bytes per tile and sweep: F N (s_m + 2 s_o + 2 s_p) (the pupil
    read and written at every LED) + the gauge's 5 M s_o + 2 N s_p."""
    log(f"building the EPRY batch ({tiles} tiles)")
    ts = workload_tiles(tiles, true_pupil_rms=0.5)
    alg = Epry(iterations=iterations)
    secs, sweep_ms, per_launch, dl = run_batched(ts, iterations, steps, warmup, alg)
    model = ts[0][0]
    N, F = model.image_len, model.frame_count
    M = model.reconstruction_shape[0] * model.reconstruction_shape[1]
    per_sweep = F * N * (2 + 2 * 16 + 2 * 8) + 5 * M * 16 + 2 * N * 8
    achieved = per_sweep * tiles / (sweep_ms * 1e-3) / 1e9
    return {"tiles": tiles, "workload": "225 LEDs 128^2 -> 375^2 EPRY 30 sweeps, Zernike 0.5 rad true pupil",
            "tiles_per_s": steps * tiles / secs, "ms_per_batch": 1000.0 * secs / steps, "sweep_ms": sweep_ms,
            "sweeps_per_launch": per_launch,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "note": "fused EPRY sweep k_sweep_epry + k_gauge per sweep; algorithmic bytes per sweep of "
                                 "every tile / mean per-sweep time (HIP events)"}}


def host_threads():
    """Host threads for the CPU baseline: the CPUs this process may run on,
    at most 16 (the host share that comes with one GPU of the box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def fft_speed_vs_pocketfft(n=128, reps=200):
    """Time of the oracle's 2-D FFT (this project's Stockham code, which the
    CPU baseline uses) over numpy's pocketfft (single thread) on one n x n
    complex128 forward transform: > 1 means the oracle's FFT is slower."""
    import oracle

    x = (np.random.default_rng(0).standard_normal((n, n)) + 1j * np.random.default_rng(1).standard_normal((n, n)))
    oracle.fft2(x)
    np.fft.fft2(x)
    t = time.perf_counter()
    for _ in range(reps):
        oracle.fft2(x)
    t_or = (time.perf_counter() - t) / reps
    t = time.perf_counter()
    for _ in range(reps):
        np.fft.fft2(x) / x.size
    t_np = (time.perf_counter() - t) / reps
    return {"n": n, "oracle_us": 1e6 * t_or, "pocketfft_us": 1e6 * t_np, "ratio": t_or / t_np}


def cpu_baseline(tiles, iterations, threads):
    """The oracle (scalar C++ FP64 restatement of the reference's serial
    per-tile path) on `threads` host threads, one whole tile (all
    `iterations` sweeps) per thread, tiles distinct."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    isa = oracle.use_widest_isa()
    alg = algorithm(iterations).to_c()

    def work(i):
        model, meas = tiles[i % len(tiles)]
        oracle.run(model, meas, alg, iterations)

    th = [threading.Thread(target=work, args=(i,)) for i in range(threads)]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    wall = time.perf_counter() - t0
    fft = fft_speed_vs_pocketfft()
    return {"value": threads / wall, "unit": "tiles/s", "cores": threads, "kind": "port",
            "sample": f"{threads} distinct workload tiles, all {iterations} mPIE sweeps each, one tile per thread on "
                      f"{threads} threads ({wall:.1f} s wall); oracle built -O3 -march=x86-64-{isa}",
            "fft_vs_pocketfft": fft}


def copy_calibration(gib=1, reps=3):
    """Counter calibration: `reps` device-to-device copies of `gib` GiB
    (one __amd_rocclr_copyBuffer dispatch each). rocprofv3's FETCH_SIZE and
    WRITE_SIZE of these dispatches, divided by the bytes each copy reads and
    writes, are the factors that turn the counters of another kernel into
    HBM bytes. Device memory comes from torch (plumbing only)."""
    import torch

    n = gib << 28  # float32 elements
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    times = []
    for _ in range(reps):
        t = time.perf_counter()
        dst.copy_(src)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    return {"metric": "device copy (counter calibration)", "bytes_read_per_copy": n * 4, "bytes_written_per_copy": n * 4,
            "copies": reps, "best_GBs": 2 * n * 4 / min(times) / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--iterations", type=int, default=50)
    ap.add_argument("--mode", default="batched",
                    choices=["batched", "plain", "mixed", "serving", "uneven", "epry", "copy"])
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--sustained-rate", type=float, default=40.0,
                    help="second serving leg below the GPU's capacity (the first leg's rate overloads it; 0: none)")
    ap.add_argument("--no-extras", action="store_true", help="batched mode without the plain and serving legs")
    ap.add_argument("--tiles", type=int, default=256)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported: tiles are independent, "
              "so more GPUs would each run their own batch of tiles, and no figure over several GPUs "
              "is measured", file=sys.stderr)
        sys.exit(2)
    if args.mode == "copy":
        print(json.dumps(copy_calibration()), flush=True)
        return
    if args.mode == "serving":
        res = serving(args.requests, args.rate, 30)
        sus = serving(args.requests, args.sustained_rate, 30) if args.sustained_rate > 0 else None
        print(json.dumps({"metric": "serving p99 latency, FP32 mPIE 30 sweeps", "value": res["p99_ms"],
                          "unit": "ms", "n_gpus": 1, "higher_is_better": False, "data": "synthetic",
                          "config": {"workload": "Poisson arrivals", **res}, "sustained": sus}), flush=True)
        return
    if args.mode == "epry":
        res = run_epry(args.tiles, 30, args.steps, args.warmup)
        print(json.dumps({"metric": "EPRY tiles/sec, precision form, 225 LEDs x 128^2, 30 sweeps",
                          "value": res["tiles_per_s"], "unit": "tiles/s", "n_gpus": 1, "higher_is_better": True,
                          "data": "synthetic", "config": res}), flush=True)
        return
    if args.mode == "uneven":
        res = run_uneven(40, args.steps, args.warmup)
        print(json.dumps({"metric": "uneven batch tiles/sec, FP32 mPIE 40 sweeps", "value": res["tiles_per_s"],
                          "unit": "tiles/s", "n_gpus": 1, "higher_is_better": True, "data": "synthetic",
                          "config": {"workload": "96 uneven tiles", **res}}), flush=True)
        return
    log("building workload")
    precision = "c128" if args.mode == "plain" else "f32"
    ntiles = args.tiles if args.mode == "batched" else 1
    t_build = time.perf_counter()
    tiles = workload_tiles(max(ntiles, 1 if args.no_cpu else host_threads()))
    model, meas = tiles[0]
    log(f"workload ready: {len(tiles)} tiles {model.reconstruction_shape} in {time.perf_counter() - t_build:.1f} s")
    s_o = 16
    s_p = 16 if precision == "c128" else 8
    s_v = 16 if precision == "c128" else 8  # velocity / anchor storage
    B = tile_bytes(model, args.iterations, 30, 2, s_o, s_p, s_v)
    traffic = None
    if args.mode == "batched":
        secs, sweep_ms, per_launch, dl = run_batched(tiles[:ntiles], args.iterations, args.steps, args.warmup)
        if args.tiles == 256:
            traffic = measured_traffic()
        tiles_per_s = args.steps * args.tiles / secs
        # dominant kernel: the fused sweep; one dispatch runs `per_launch`
        # sweeps of every tile, and bytes and time are both per sweep here
        achieved = sweep_bytes(model, args.iterations, 30, 2, s_o, s_p, s_v) * args.tiles / (sweep_ms * 1e-3) / 1e9
        note = (f"fused sweep kernel k_sweep, {per_launch} sweeps of every tile per dispatch: algorithmic bytes "
                "of one sweep of every tile (object window c128 read+write, pupil c64, uint16 frame per LED; "
                "object c128 and velocity/anchor c64 read+write per momentum event) / mean per-sweep time "
                "(HIP events around each dispatch / its sweeps); traffic is counter-measured HBM bytes per sweep")
    else:
        secs = run_plain(model, meas, args.iterations, args.steps, args.warmup, precision)
        tiles_per_s = args.steps / secs
        achieved = B * tiles_per_s / 1e9
        note = "whole-tile algorithmic bytes / wall time"
    line = {
        "metric": "mPIE tiles/sec, 225 LEDs x 128^2 tiles (375^2 object), 50 sweeps: batched FP32 vs plain FP64; p99 latency",
        "value": tiles_per_s,
        "unit": "tiles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * secs / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if precision == "c128" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "225 LEDs 128^2 -> 375^2 mPIE 50 sweeps, uint16 Poisson 1e4, distinct tiles",
                   "path": args.mode, "tiles_per_batch": ntiles},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "note": note},
    }
    if traffic:
        line["roofline"]["traffic_GBs"] = traffic / (sweep_ms * 1e-3) / 1e9
        line["roofline"]["traffic_frac"] = line["roofline"]["traffic_GBs"] / HBM_PEAK_GBS
    if args.mode == "batched":
        # results left in HBM count in `value`; the same batch with its
        # results downloaded to host memory over PCIe:
        line["result_download_ms"] = dl
        line["pcie_inclusive_tiles_per_s"] = args.tiles / (secs / args.steps + dl / 1000.0)
    if args.mode == "batched" and not args.no_extras:
        log("plain complex128 GPU leg")
        psecs = run_plain(model, meas, args.iterations, 1, 1, "c128")
        line["plain_c128_tiles_per_s"] = 1.0 / psecs
        line["speedup_vs_plain_c128"] = tiles_per_s * psecs
        # the plain path's whole-tile algorithmic bytes (object window, pupil,
        # velocity and anchor all complex128) over its wall time
        b128 = tile_bytes(model, args.iterations, 30, 2, 16, 16, 16)
        line["plain_c128_roofline"] = {"bound": "hbm", "achieved": b128 / psecs / 1e9, "peak": HBM_PEAK_GBS,
                                       "unit": "GB/s", "frac": b128 / psecs / 1e9 / HBM_PEAK_GBS,
                                       "note": "per-stage kernels, one tile at a time, each sweep replayed "
                                               "from a captured graph"}
        # the fused FP32 kernel on a batch of one tile: the per-tile part of the
        # batched / plain ratio (the rest is 256 tiles running side by side)
        log("one-tile fused leg")
        fsecs, fsweep_ms, _, _ = run_batched(tiles[:1], args.iterations, 1, 1)
        line["fused_one_tile"] = {"tiles_per_s": 1.0 / fsecs, "sweep_ms": fsweep_ms,
                                  "speedup_vs_plain_c128": psecs / fsecs,
                                  "batching_factor": tiles_per_s * fsecs}
        log("uneven batch leg")
        line["uneven_batch"] = run_uneven(40, 1, 1)
        log("EPRY batch leg")
        line["epry_batch"] = run_epry(256, 30, 1, 1)
        log("serving leg")
        line["serving"] = serving(args.requests, args.rate, 30)
        if args.sustained_rate > 0:
            log("serving leg at a sustainable rate")
            line["serving_sustained"] = serving(args.requests, args.sustained_rate, 30)
        # 80 % of the capacity the saturated leg measured (its tiles/s over its
        # tiles per request): a load the server keeps up with, near its limit
        sat = line["serving"]
        near = 0.8 * sat["tiles_per_s"] * sat["requests"] / sat["tiles"]
        log(f"serving leg at 80 % of capacity ({near:.1f} req/s)")
        line["serving_near_capacity"] = serving(args.requests, near, 30)
    if not args.no_cpu:
        threads = args.cpu_threads or host_threads()
        log(f"cpu baseline on {threads} threads")
        line["cpu_baseline"] = cpu_baseline(tiles, args.iterations, threads)
        line["speedup_vs_cpu"] = tiles_per_s / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    try:
        main()
    finally:
        runtime.close_all()
