"""The FP64 oracle pinned against the reference's own closed-form tests,
golden vectors and reference-held thresholds, and against the mathematics
(DFT definition, numpy.fft, brute force). CPU only."""
import math

import numpy as np
import pytest

import oracle
import refobjects as ro
from fpmhip import Epry, Fpie, MeasurementStack, Mpie, AlternatingProjection, abi


def dft2_bruteforce(x, sign):
    h, w = x.shape
    rr = np.exp(sign * 2j * np.pi * np.outer(np.arange(h), np.arange(h)) / h)
    cc = np.exp(sign * 2j * np.pi * np.outer(np.arange(w), np.arange(w)) / w)
    return rr @ x @ cc


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (7, 9), (8, 8), (12, 10), (16, 16), (42, 42), (11, 13)])
def test_fft_matches_dft_definition(shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    n = x.size
    fwd, inv = dft2_bruteforce(x, -1) / n, dft2_bruteforce(x, +1)
    np.testing.assert_allclose(oracle.fft2(x, "forward"), fwd, rtol=0, atol=1e-14 * np.abs(fwd).max() * 10)
    np.testing.assert_allclose(oracle.fft2(x, "inverse"), inv, rtol=0, atol=1e-14 * np.abs(inv).max() * 10)


@pytest.mark.parametrize("shape", [(128, 128), (375, 375), (729, 729), (64, 96)])
def test_fft_matches_numpy_at_path_sizes(shape):
    rng = np.random.default_rng(2)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    ref = np.fft.fft2(x) / x.size
    got = oracle.fft2(x, "forward")
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max() * 10


def test_fft_round_trip_reference_case():
    """tests/forward_model_tests.rs:13: 7x9, values index/13 + i (index % 5), 1e-10."""
    idx = np.arange(63)
    x = (idx / 13.0 + 1j * (idx % 5)).reshape(7, 9)
    back = oracle.fft2(oracle.fft2(x, "forward"), "inverse")
    np.testing.assert_allclose(back, x, rtol=0, atol=1e-10)


@pytest.mark.parametrize("shape", [(5, 7), (8, 8), (375, 3), (1, 4)])
def test_shifts_match_numpy(shape):
    x = np.arange(np.prod(shape)).reshape(shape) + 0j
    assert np.array_equal(oracle.shift(x), np.fft.fftshift(x))
    assert np.array_equal(oracle.shift(x, inverse=True), np.fft.ifftshift(x))


def test_constant_object_has_constant_central_frame():
    """tests/forward_model_tests.rs:76: constant 2 -> intensity 4.0 within 1e-10."""
    model = ro.direct_model()
    spec = ro.spectrum_of(np.full((16, 16), 2.0 + 0j))
    I = oracle.forward_stack(model, spec)
    assert np.all(np.abs(I[2] - 4.0) < 1e-10)


def test_patch_extraction_and_update_are_consistent():
    """tests/forward_model_tests.rs:104."""
    upd = np.full((8, 8), 1.0 - 0.5j)
    spec = oracle.insert_adjoint(np.zeros((16, 16), complex), (4, 4), upd, 1.0)
    assert np.array_equal(oracle.extract(spec, (4, 4), (8, 8)), upd)


def test_subpixel_patch_uses_bilinear_fourier_sampling():
    """tests/forward_model_tests.rs:117: linear ramp sampled at 4.25 + r, 4.5 + c, 1e-12."""
    r, c = np.meshgrid(np.arange(16.0), np.arange(16.0), indexing="ij")
    spec = (1 + 2 * r + 3 * c) + 1j * (-r + 0.5 * c)
    patch = oracle.extract(spec, (4, 4), (8, 8), (0.25, 0.5))
    sr, sc = np.meshgrid(4.25 + np.arange(8), 4.5 + np.arange(8), indexing="ij")
    np.testing.assert_allclose(patch, (1 + 2 * sr + 3 * sc) + 1j * (-sr + 0.5 * sc), rtol=0, atol=1e-12)


@pytest.mark.parametrize("cycles", [1, 4, 8])
def test_bilinear_error_is_cosine_attenuation(cycles):
    """tests/forward_model_tests.rs:155: half-pixel error = 1 - cos(pi nu / H), 1e-12."""
    H = 32
    rows = np.arange(H)[:, None] * np.ones((1, H))
    spec = np.exp(2j * np.pi * cycles * rows / H)
    s = oracle.extract(spec, (10, 10), (8, 8), (0.5, 0.0))
    ref_rows = 10 + np.arange(8)[:, None] + 0.5
    ref = np.exp(2j * np.pi * cycles * ref_rows / H) * np.ones((1, 8))
    err = math.sqrt(np.mean(np.abs(s - ref) ** 2))
    assert abs(err - (1 - math.cos(math.pi * cycles / H))) < 1e-12


@pytest.mark.parametrize("offset", [(-0.3, 0.4), (0.25, 0.0), (0.0, -0.7), (1.0, 0.0)])
def test_insert_is_exact_adjoint(offset):
    """tests/forward_model_tests.rs:219: <E x, u> = <x, E* u> within 1e-12."""
    idx = np.arange(256)
    spec = ((idx % 13) / 7.0 + 1j * ((idx % 17) / 11.0 - 0.5)).reshape(16, 16)
    j = np.arange(64)
    upd = ((j % 7) / 5.0 - 0.4 + 1j * (j % 11) / 9.0).reshape(8, 8)
    ex = oracle.extract(spec, (4, 4), (8, 8), offset)
    adj = oracle.insert_adjoint(np.zeros((16, 16), complex), (4, 4), upd, 1.0, offset)
    left = np.sum(np.conj(ex) * upd)
    right = np.sum(np.conj(spec) * adj)
    assert abs(left - right) < 1e-12


def test_geometry_sign_and_fractional_offsets():
    """tests/geometry_tests.rs:474: k = (1.25 dk, -0.5 dk) on 8x8 -> 16x16 gives
    column offset 0.25, row offset 0.5 (round half away from zero), start
    column right of and start row left of the central start 4. Expressed with a
    one-LED array whose position produces exactly that k."""
    opt = ro.base_optics()
    px = opt.camera_pixel_size / opt.magnification
    dk = 2 * math.pi / (8 * px)
    k0 = 2 * math.pi / opt.wavelength_vacuum_m
    kx, ky = 1.25 * dk, -0.5 * dk
    # direction d = -p/|p| with p = (x, y, -z): choose z = 90 mm
    z = 90e-3
    dx, dy = kx / k0, ky / k0
    dz = math.sqrt(1 - dx * dx - dy * dy)
    pos = (-dx / dz * z, -dy / dz * z)
    arr = ro.PlanarLedArray((1, 1), (4e-3, 4e-3), (0.0, 0.0), (pos[0], pos[1], -z))
    m = oracle.compile_model(opt, arr, (8, 8), (16, 16))
    crop, off = m.crop_start[0], m.offset[0]
    assert crop[1] > 4 and crop[0] < 4
    assert abs(off[1] - 0.25) < 1e-9 and abs(off[0] - 0.5) < 1e-9


def test_shape_policies_by_brute_force():
    """Minimum = smallest fitting multiplier (brute force over multipliers);
    Smooth = next 2/3/5/7-smooth; PowerOfTwo = next power of two
    (image_plane_fpm.rs:287-348, 1116-1198)."""
    from refobjects import PlanarLedArray
    opt = ro.base_optics()
    for k, n in [(3, 32), (15, 128), (13, 64), (9, 128)]:
        arr = PlanarLedArray((k, k), (4e-3, 4e-3), ((k - 1) / 2, (k - 1) / 2), (0, 0, -90e-3))
        mn = oracle.compile_model(opt, arr, (n, n), "minimum").reconstruction_shape[0]
        sm = oracle.compile_model(opt, arr, (n, n), "smooth").reconstruction_shape[0]
        p2 = oracle.compile_model(opt, arr, (n, n), "power_of_two").reconstruction_shape[0]
        with pytest.raises(oracle.OracleError):
            oracle.compile_model(opt, arr, (n, n), (mn - 1, mn - 1))
        oracle.compile_model(opt, arr, (n, n), (mn, mn))
        smooth = min(v for v in range(mn, 4 * mn) if _is_smooth(v))
        assert sm == smooth
        assert p2 == 1 << (mn - 1).bit_length()


def _is_smooth(v):
    for p in (2, 3, 5, 7):
        while v % p == 0:
            v //= p
    return v == 1


def test_baseline_grid_sizes():
    """Smooth grids of the measured configurations: 42, 375, 729, 168, 336, 672."""
    from refobjects import PlanarLedArray
    opt = ro.base_optics()
    cases = [(3, 32, 42), (15, 128, 375), (15, 256, 729), (13, 64, 168), (13, 128, 336), (13, 256, 672)]
    for k, n, R in cases:
        arr = PlanarLedArray((k, k), (4e-3, 4e-3), ((k - 1) / 2, (k - 1) / 2), (0, 0, -90e-3))
        assert oracle.compile_model(opt, arr, (n, n), "smooth").reconstruction_shape == (R, R)


def test_circular_pupil_closed_form():
    """Pupil::circular (pupil.rs:78-124): support |k| <= 2 pi NA / lambda on
    the centred grid; config-2 support has 4809 samples."""
    from refobjects import PlanarLedArray
    opt = ro.base_optics()
    arr = PlanarLedArray((15, 15), (4e-3, 4e-3), (7, 7), (0, 0, -90e-3))
    m = oracle.compile_model(opt, arr, (128, 128), "smooth")
    px = opt.camera_pixel_size / opt.magnification
    dk = 2 * math.pi / (128 * px)
    r, c = np.meshgrid(np.arange(128) - 64, np.arange(128) - 64, indexing="ij")
    inside = np.hypot(c * dk, r * dk) <= 2 * math.pi * opt.objective_na / opt.wavelength_vacuum_m
    assert np.array_equal(m.support.astype(bool), inside)
    assert int(inside.sum()) == 4809
    assert np.array_equal(m.pupil, inside.astype(complex))


def test_momentum_golden_vector():
    """src/algorithms/common.rs:513-535, exact."""
    o, v, a = oracle.momentum_event([3 + 2j, -1 + 0j], [1 + 0j, -2j], [1 + 1j, -2 + 0j], 0.5, 0.25)
    assert list(v) == [2.5 + 1j, 1 - 1j]
    assert list(o) == [3.625 + 2.25j, -0.75 - 0.25j]
    assert list(a) == list(o)


def test_momentum_rejects_non_finite():
    """src/algorithms/common.rs:538."""
    with pytest.raises(oracle.OracleError) as e:
        oracle.momentum_event([complex(math.inf, 0)], [0j], [0j], 0.9, 0.9)
    assert e.value.kind == "Numerical"


def test_initialization_undoes_gain_and_background():
    """tests/reconstruction_tests.rs:2713: gain 4, background 9, amplitude 2
    object -> amplitude 2 within 1e-10 after 0 iterations."""
    model = ro.direct_model().with_frame_gains(np.full(5, 4.0)).with_background(np.full(64, 9.0))
    spec = ro.spectrum_of(np.full((16, 16), 2.0 + 0j))
    I = oracle.forward_stack(model, spec)
    res = oracle.run(model, MeasurementStack(frames=I), AlternatingProjection().to_c(), 0)
    np.testing.assert_allclose(res["amplitude"], 2.0, rtol=0, atol=1e-10)


def _ideal(model, obj):
    return MeasurementStack(frames=oracle.forward_stack(model, ro.spectrum_of(obj)))


def test_ap_and_fpie_loss_decrease():
    """tests/reconstruction_tests.rs:66 and :638 (8 iterations, resolution target)."""
    model = ro.direct_model()
    meas = _ideal(model, ro.resolution_target((16, 16)))
    for alg in (AlternatingProjection(), Fpie()):
        t = oracle.run(model, meas, alg.to_c(), 8)["trace"]
        assert t[-1] < t[0]


def test_ap_matches_textbook_amplitude_replacement():
    """With a unit pupil, integer crops and a single source per frame, AP's
    step is the textbook Gerchberg-Saxton Fourier update: replace the
    detector amplitude, keep the phase, and add the spectral difference into
    the crop (object_step 1, eps -> negligible). Written here independently
    with numpy.fft."""
    model = ro.direct_model()
    obj = ro.mixed_test_pattern((16, 16))
    meas = _ideal(model, obj * 0.9 + 0.1)
    I = meas.frames
    spec = oracle.initialize(model, meas)
    ref = spec.copy()
    for it in range(2):
        for f in range(5):
            r0, c0 = (int(v) for v in model.crop_start[f])
            patch = ref[r0:r0 + 8, c0:c0 + 8].copy()
            field = np.fft.ifft2(np.fft.ifftshift(patch)) * 64
            amp = np.abs(field)
            new = np.where(amp ** 2 > 1e-10, field * np.sqrt(I[f]) / np.where(amp > 0, amp, 1), np.sqrt(I[f]))
            newspec = np.fft.fftshift(np.fft.fft2(new) / 64)
            ref[r0:r0 + 8, c0:c0 + 8] += (newspec - patch) / (1.0 + 1e-10)
    got = oracle.run(model, meas, AlternatingProjection().to_c(), 2)["object_spectrum"]
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-12


def test_mpie_zero_feedback_equals_fpie_bitwise():
    """tests/reconstruction_tests.rs:655."""
    model = ro.direct_model()
    meas = _ideal(model, ro.mixed_test_pattern((16, 16)))
    f = oracle.run(model, meas, Fpie(object_step=0.2, stability=0.05).to_c(), 4)
    m = oracle.run(model, meas, Mpie(momentum_interval=3, momentum_feedback=0.0).to_c(), 4)
    assert np.array_equal(f["object_spectrum"], m["object_spectrum"])
    assert np.array_equal(f["trace"], m["trace"])


def test_mpie_accelerates_rpie_reference_threshold():
    """tests/reconstruction_tests.rs:992: noiseless_mixed preset (3x3 LEDs,
    32 -> 64), 4x phase; mPIE(20, interval 10, 0.7/0.7) < 0.95 FPIE(20, 0.2,
    0.05) and reaches 5.4e-5 first."""
    model = oracle.compile_model(ro.base_optics(), ro.base_array(), (32, 32), (64, 64))
    base = ro.mixed_test_pattern((64, 64))
    strong = np.abs(base) * np.exp(4j * np.angle(base))
    meas = _ideal(model, strong)
    f = oracle.run(model, meas, Fpie(iterations=20, object_step=0.2, stability=0.05).to_c(), 20)["trace"]
    m = oracle.run(model, meas, Mpie(iterations=20, momentum_interval=10, momentum_friction=0.7,
                                      momentum_feedback=0.7).to_c(), 20)["trace"]
    assert m[-1] < 0.95 * f[-1]
    mi = np.argmax(m < 5.4e-5) if (m < 5.4e-5).any() else None
    fi = np.argmax(f < 5.4e-5) if (f < 5.4e-5).any() else None
    assert mi is not None and (fi is None or mi < fi)


def test_mpie_counter_semantics():
    """tests/reconstruction_tests.rs:773 (counter 3 after 2 x 5 frames, interval
    7) and :837 (zero-weight frames do not count; a multiplexed frame counts once)."""
    model = ro.direct_model()
    meas = _ideal(model, ro.mixed_test_pattern((16, 16)))
    assert oracle.run(model, meas, Mpie(momentum_interval=7).to_c(), 2)["counter"] == 3
    w = MeasurementStack(frames=meas.frames, weights=np.array([0, 0, 1, 1, 1.0]))
    assert oracle.run(model, w, Mpie(momentum_interval=3).to_c(), 1)["counter"] == 0
    mux = ro.direct_model().with_multiplexing([[(0, 0.6), (1, 0.4)], [(2, 0.3), (3, 0.2), (4, 0.5)]])
    mm = _ideal(mux, ro.mixed_test_pattern((16, 16)))
    assert oracle.run(mux, mm, Mpie(momentum_interval=2).to_c(), 1)["counter"] == 0


def test_mpie_resume_is_exact():
    """tests/reconstruction_tests.rs:773: 2 iterations, checkpoint, 2 more ==
    4 uninterrupted (1e-14)."""
    model = ro.direct_model()
    meas = _ideal(model, ro.mixed_test_pattern((16, 16)))
    alg = Mpie(momentum_interval=7).to_c()
    a = oracle.run(model, meas, alg, 2)
    b = oracle.run(model, meas, alg, 2, resume={"spectrum": a["object_spectrum"], "pupil": a["recovered_pupil"],
                                                 "velocity": a["velocity"], "anchor": a["anchor"],
                                                 "counter": a["counter"]})
    full = oracle.run(model, meas, alg, 4)
    assert np.abs(b["object_spectrum"] - full["object_spectrum"]).max() <= 1e-14


def test_epry_gauge_invariants():
    """tests/reconstruction_tests.rs:2674: after EPRY the supported pupil
    energy equals the reference (1e-11), the overlap is real positive, and the
    centred DC coefficient is real non-negative."""
    model = ro.direct_model()
    meas = _ideal(model, ro.phase_disk((16, 16), 4.5, 0.8))
    r = oracle.run(model, meas, Epry(iterations=3).to_c(), 3)
    P, Pref, sup = r["recovered_pupil"], model.pupil, model.support.astype(bool)
    assert abs(np.sum(np.abs(P[sup]) ** 2) - np.sum(np.abs(Pref[sup]) ** 2)) < 1e-11
    ov = np.sum(np.conj(Pref[sup]) * P[sup])
    assert ov.real > 0 and abs(ov.imag) <= 1e-11 * ov.real
    dc = r["object_spectrum"][8, 8]
    assert dc.real >= 0 and abs(dc.imag) <= 1e-11 * max(abs(dc), 1)


def test_gauge_preserves_predictions():
    """src/algorithms/gauge.rs:436 and :465: canonicalisation leaves the
    predicted intensities unchanged (2e-12) and removes the exact-axis ramp."""
    model = ro.direct_model()
    rng = np.random.default_rng(3)
    spec = (rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))) * np.exp(0.37j)
    y, x = np.meshgrid(np.arange(8) - 4, np.arange(8) - 4, indexing="ij")
    P = 2.5 * np.exp(1j * (-0.21 + 0.17 * y - 0.23 * x))
    before = oracle.forward_stack(model, spec, P)
    s2, p2 = oracle.canonicalize(model, spec, P)
    after = oracle.forward_stack(model, s2, p2)
    np.testing.assert_allclose(after, before, rtol=2e-12, atol=2e-12)
    np.testing.assert_allclose(p2, np.ones((8, 8)), atol=2e-13)


def test_forward_stack_matches_numpy_definition():
    """I_f = g |IFFT_unscaled(ifftshift(patch_f * P))|^2 + bg, multiplexed as
    sum_s w_s |.|^2 (forward.rs:245-292), with subpixel offsets."""
    model = ro.direct_model().with_subpixel_offsets([[0, 0], [0.3, -0.2], [0, 0], [-0.4, 0.1], [0, 0]])
    model.with_multiplexing([[(0, 0.7), (1, 0.3)], [(2, 0.4), (3, 0.6)], [(4, 1.0)]])
    model.with_frame_gains(np.array([2.0, 3.0, 0.5])).with_background(np.linspace(0, 1, 64))
    rng = np.random.default_rng(4)
    spec = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    P = np.exp(1j * rng.standard_normal((8, 8)))
    got = oracle.forward_stack(model, spec, P)
    for f, row in enumerate(model.multiplexing):
        acc = np.zeros((8, 8))
        for s, w in row:
            patch = oracle.extract(spec, model.crop_start[s], (8, 8), model.offset[s])
            acc += w * np.abs(np.fft.ifft2(np.fft.ifftshift(patch * P)) * 64) ** 2
        np.testing.assert_allclose(got[f], model.frame_gain[f] * acc + model.background.reshape(8, 8),
                                   rtol=1e-12, atol=1e-12)


def test_masked_and_zero_weight_frames():
    """tests/reconstruction_tests.rs:2782 / :2820: a zero-weight frame and
    masked pixels do not affect the result (values there are arbitrary)."""
    model = ro.direct_model()
    meas = _ideal(model, ro.mixed_test_pattern((16, 16)))
    bad = meas.frames.copy()
    bad[1] *= 7.0
    w = np.array([1, 0, 1, 1, 1.0])
    a = oracle.run(model, MeasurementStack(frames=meas.frames, weights=w), Fpie().to_c(), 3)
    b = oracle.run(model, MeasurementStack(frames=bad, weights=w), Fpie().to_c(), 3)
    assert np.array_equal(a["object_spectrum"], b["object_spectrum"])
    mask = np.ones((5, 8, 8), np.uint8)
    mask[2, 3, 3] = 0
    bad2 = meas.frames.copy()
    bad2[2, 3, 3] = 1e6
    a = oracle.run(model, MeasurementStack(frames=meas.frames, masks=mask), Fpie().to_c(), 3)
    b = oracle.run(model, MeasurementStack(frames=bad2, masks=mask), Fpie().to_c(), 3)
    assert np.array_equal(a["object_spectrum"], b["object_spectrum"])


def test_epry_amplifies_initial_perturbations():
    """Conditioning of EPRY before convergence, in FP64: on the 225-LED
    128^2 problem with a 0.5 rad Zernike pupil, a 1e-7 relative perturbation
    of the initial spectrum grows more than tenfold in the object after 10
    sweeps. Rounding of the size FP32 arithmetic makes in every LED update
    therefore shows up at the 1e-6..1e-5 level in early sweeps."""
    from fpmhip import synthetic as syn
    model = oracle.compile_model(syn.base_optics(), syn.planar_array(15), (128, 128))
    obj = syn.random_field_object(model.reconstruction_shape, 7)
    I = oracle.forward_stack(model, syn.object_spectrum(obj), syn.zernike_pupil(model, 0.5, 107))
    counts, ppp = syn.poisson_counts(I, 1e4, 7)
    model.with_frame_gains(np.full(model.frame_count, ppp))
    meas = MeasurementStack(counts=counts)
    init = oracle.initialize(model, meas)
    rng = np.random.default_rng(3)
    pert = init * (1 + 1e-7 * (rng.standard_normal(init.shape) + 1j * rng.standard_normal(init.shape)))
    alg = Epry(iterations=10).to_c()
    base = oracle.run(model, meas, alg, 10, resume={"spectrum": init, "pupil": model.pupil})["object"]
    moved = oracle.run(model, meas, alg, 10, resume={"spectrum": pert, "pupil": model.pupil})["object"]
    ea = np.linalg.norm(np.abs(moved) - np.abs(base)) / np.linalg.norm(np.abs(base))
    ep = np.linalg.norm(np.angle(moved * np.conj(base))) / np.linalg.norm(np.angle(base))
    assert ea > 1e-6 and ep > 2e-6


def _aligned_gain_error(reference, candidate):
    """compare_frame_gains (src/metrics/model.rs:130-161): relative error after
    the least-squares scale of the candidate."""
    reference, candidate = np.asarray(reference), np.asarray(candidate)
    alignment = np.dot(candidate, reference) / max(np.dot(candidate, candidate), np.finfo(float).eps)
    err = np.sum((alignment * candidate - reference) ** 2)
    return math.sqrt(err / max(np.dot(reference, reference), np.finfo(float).eps))


def test_epry_recovers_relative_frame_gain_mismatch():
    """tests/reconstruction_tests.rs:3037: true gains 0.5/1/1.5/2/0.75 on the
    direct model, constant object 2, EPRY 2 iterations with object step 1e-6,
    no pupil recovery, gain recovery at step 1: the aligned relative gain
    error falls below a tenth of that of the mean gain."""
    true_gains = np.array([0.5, 1.0, 1.5, 2.0, 0.75])
    true_model = ro.direct_model().with_frame_gains(true_gains)
    meas = MeasurementStack(frames=oracle.forward_stack(true_model, ro.spectrum_of(np.full((16, 16), 2.0 + 0j))))
    alg = Epry(iterations=2, object_step=1e-6, recover_pupil=False, recover_frame_gains=True, gain_step=1.0)
    res = oracle.run(ro.direct_model(), meas, alg.to_c(), 2)
    recovered = _aligned_gain_error(true_gains, res["frame_gains"])
    initial = math.sqrt(np.sum((true_gains.mean() - true_gains) ** 2) / np.sum(true_gains ** 2))
    assert recovered < 0.1 * initial


def test_epry_recovers_relative_per_frame_background():
    """tests/reconstruction_tests.rs:3081: per-frame backgrounds 0/1/2/3/0.5,
    EPRY 1 iteration, background recovery at step 1 within (-10, 10): the
    recovered frame means relative to their mean equal the truth's (1e-8)."""
    true_bg = np.array([0.0, 1.0, 2.0, 3.0, 0.5])
    true_model = ro.direct_model().with_background(np.repeat(true_bg, 64))
    meas = MeasurementStack(frames=oracle.forward_stack(true_model, ro.spectrum_of(np.full((16, 16), 2.0 + 0j))))
    alg = Epry(iterations=1, object_step=1e-6, recover_pupil=False, recover_background=True, background_step=1.0,
               minimum_background=-10.0, maximum_background=10.0)
    res = oracle.run(ro.direct_model(), meas, alg.to_c(), 1)
    means = res["background"].reshape(5, 64).mean(axis=1)
    assert np.allclose(means - means.mean(), true_bg - true_bg.mean(), atol=1e-8, rtol=0)
