"""The oracle's camera (CPU): Philox4x64-10 against numpy's Philox bit
generator, the Poisson and normal samplers against their exact laws, and
the deterministic detector steps against the reference's own camera tests
(tests/simulation_tests.rs)."""
import numpy as np
import pytest
from scipy import stats

import oracle
from fpmhip.simulation import CameraModel

import refobjects as ro


@pytest.mark.parametrize("key,counter", [((0, 0), (0, 0, 0, 0)), ((12345, 678), (7, 3, 1, 0)),
                                         ((2**64 - 1, 2**63 + 5), (2**64 - 2, 9, 2**40, 77))])
def test_philox_matches_numpy_bit_generator(key, counter):
    # numpy's Philox increments its 256-bit counter before each block
    g = np.random.Philox(key=np.array(key, np.uint64), counter=np.array(counter, np.uint64))
    want = g.random_raw(4)
    c = list(counter)
    c[0] = (c[0] + 1) % 2**64
    got = oracle.philox(c, key)
    assert np.array_equal(got, want)


def test_philox_counter_carry_matches_numpy():
    g = np.random.Philox(key=np.array([5, 6], np.uint64), counter=np.array([2**64 - 1, 0, 0, 0], np.uint64))
    assert np.array_equal(oracle.philox([0, 1, 0, 0], [5, 6]), g.random_raw(4))


@pytest.mark.parametrize("mean", [0.37, 3.0, 9.99, 10.0, 57.3, 1.0e4])
def test_poisson_sampler_follows_the_poisson_law(mean):
    n = 200_000
    x = oracle.camera_draws(mean, n, seed=11)
    assert np.all(x == np.floor(x)) and x.min() >= 0
    # moments: 5 standard errors
    assert abs(x.mean() - mean) < 5 * np.sqrt(mean / n)
    assert abs(x.var(ddof=1) / mean - 1) < 5 * np.sqrt(2.0 / n) + 5 * np.sqrt(1.0 / (mean * n))
    # chi-square over bins of expected count >= 50
    lo, hi = stats.poisson.ppf(1e-4, mean), stats.poisson.ppf(1 - 1e-4, mean)
    edges = np.unique(np.round(np.linspace(lo, hi + 1, 40)))
    cdf = stats.poisson.cdf(edges - 1, mean)
    probs = np.diff(np.concatenate([[0.0], cdf, [1.0]]))
    counts = np.bincount(np.searchsorted(edges, x, side="right"), minlength=probs.size)
    keep = probs * n >= 50
    obs = np.concatenate([counts[keep], [counts[~keep].sum()]])
    exp = np.concatenate([probs[keep], [probs[~keep].sum()]]) * n
    obs, exp = obs[exp > 0], exp[exp > 0]
    chi2 = ((obs - exp) ** 2 / exp).sum()
    assert stats.chi2.sf(chi2, obs.size - 1) > 1e-4


def test_normal_sampler_follows_the_standard_normal_law():
    x = oracle.camera_draws(-1.0, 200_000, seed=3)
    assert abs(x.mean()) < 5 / np.sqrt(x.size)
    assert abs(x.var(ddof=1) - 1) < 5 * np.sqrt(2.0 / x.size)
    assert stats.kstest(x, "norm").pvalue > 1e-4


def test_draws_depend_on_seed_and_pixel_only():
    a = oracle.camera_draws(25.0, 1000, seed=9)
    b = oracle.camera_draws(25.0, 400, seed=9)
    assert np.array_equal(a[:400], b)
    assert not np.array_equal(a, oracle.camera_draws(25.0, 1000, seed=10))


def test_ideal_camera_is_identity():
    # simulation_tests.rs:101-108
    assert list(oracle.camera_measure(CameraModel.ideal(), [0.0, 0.5, 2.0], 1)) == [0.0, 0.5, 2.0]


def test_camera_pixel_sensitivity_and_electronics():
    # simulation_tests.rs:111-123
    cam = (CameraModel.ideal().with_photons_per_pixel(10.0).with_pixel_sensitivity([1.0, 0.5])
           .with_dark_current_electrons(3.0).gain(2.0).with_offset_counts(5.0))
    assert list(oracle.camera_measure(cam, [2.0, 2.0], 2)) == [51.0, 31.0]


def test_camera_clips_then_quantizes_before_bad_pixels():
    # simulation_tests.rs:208-222
    cam = (CameraModel.ideal().with_photons_per_pixel(10.0).gain(2.0).with_offset_counts(1.0).with_bit_depth(4)
           .saturation(10.0).with_quantize(True).with_bad_pixels([1], 7.4))
    assert list(oracle.camera_measure(cam, [-1.0, 0.12, 0.55, 1.0], 7)) == [1.0, 7.0, 10.0, 10.0]


def test_camera_dark_current_and_bad_pixels_on_a_simulated_stack():
    # simulation_tests.rs:554-573: constant object 1 through the direct model
    model = ro.direct_model()
    I = oracle.forward_stack(model, ro.spectrum_of(np.ones((16, 16), complex)))
    cam = CameraModel.new().with_photons_per_pixel(10.0).with_dark_current_electrons(7.0).with_bad_pixels([0, 5], 123.0)
    out = oracle.camera_measure(cam, I, 0)
    for f in range(out.shape[0]):
        v = out[f].ravel()
        assert v[0] == 123.0 and v[5] == 123.0 and v[1] == 17.0


@pytest.mark.parametrize("kind", ["shot", "read", "shot+read"])
def test_camera_noise_moments(kind):
    # simulation_tests.rs:144-205 (65536 pixels, the reference's bounds)
    frame = np.ones(65_536)
    if kind == "shot":
        cam = CameraModel.ideal().with_photons_per_pixel(100.0).with_shot_noise(True)
        mean, var, tm, tv = 100.0, 100.0, 0.5, 4.0
    elif kind == "read":
        frame = np.zeros(65_536)
        cam = (CameraModel.ideal().with_dark_current_electrons(50.0).with_read_noise_electrons(4.0).gain(2.0)
               .with_offset_counts(7.0))
        mean, var, tm, tv = 107.0, 64.0, 0.35, 3.0
    else:
        cam = (CameraModel.ideal().with_photons_per_pixel(80.0).with_shot_noise(True).with_read_noise_electrons(3.0)
               .gain(1.5).with_offset_counts(11.0))
        mean, var, tm, tv = 1.5 * 80 + 11, 1.5**2 * (80 + 9), 0.6, 5.0
    out = oracle.camera_measure(cam, frame, 4)
    assert abs(out.mean() - mean) < tm
    assert abs(out.var(ddof=1) - var) < tv
