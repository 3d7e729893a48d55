"""Oracle pins taken from answers the reference's own tests hold: the loss
types' scalar values, the frame schedules' golden orders, the EPRY pupil
recovery relation of the defocus + astigmatism experiment, and closed-form
properties of Pupil::circular's aberration branch. CPU only."""
import math

import numpy as np
import pytest

import oracle
import refobjects as ro
from fpmhip import Epry, MeasurementStack, abi
from fpmhip.model import Optics

L = abi.LOSS


# ---- point_loss (src/algorithms/objective.rs:47-71) ----------------------------

def test_point_loss_matches_reference_scalar_values():
    """tests/diagnostics_tests.rs:34-58: predicted 4, measured 1."""
    got = {k: float(oracle.point_loss([4.0], [1.0], v)[0]) for k, v in L.items()}
    assert got["amplitude_mse"] == pytest.approx(1.0, abs=1e-15)
    assert got["intensity_mse"] == pytest.approx(9.0, abs=1e-15)
    assert got["huber_amplitude"] == pytest.approx(0.5, abs=1e-15)
    assert got["poisson_nll"] == pytest.approx(4.0 - math.log(4.0), abs=1e-15)


def test_point_loss_branches_by_closed_form():
    """The branches the scalar case leaves open: the Huber quadratic zone
    (|r| <= 1), clamping of negative intensities for the amplitude losses and
    the Poisson mean floor of 1e-12; intensity MSE keeps negative values."""
    pred = np.array([1.21, -2.0, 9.0, 0.0, -3.0])
    meas = np.array([1.0, 4.0, -1.0, 2.0, 1.0])
    amp = oracle.point_loss(pred, meas, L["amplitude_mse"])
    hub = oracle.point_loss(pred, meas, L["huber_amplitude"])
    poi = oracle.point_loss(pred, meas, L["poisson_nll"])
    ims = oracle.point_loss(pred, meas, L["intensity_mse"])
    # residuals sqrt(max(p,0)) - sqrt(max(y,0)): 0.1, -2, 3, -sqrt 2, -1
    np.testing.assert_allclose(amp, [0.01, 4.0, 9.0, 2.0, 1.0], rtol=1e-14)
    np.testing.assert_allclose(hub, [0.005, 1.5, 2.5, math.sqrt(2) - 0.5, 0.5], rtol=1e-14)
    np.testing.assert_allclose(ims, [0.0441, 36.0, 100.0, 4.0, 16.0], rtol=1e-13)
    floor = 1e-12
    np.testing.assert_allclose(poi, [1.21 - math.log(1.21), floor - 4.0 * math.log(floor), 9.0,
                                     floor - 2.0 * math.log(floor), floor - math.log(floor)], rtol=1e-14)


# ---- FrameSchedule (src/reconstruction/schedule.rs:35-146) ----------------------

def test_brightfield_and_spiral_orders_match_reference():
    """tests/schedule_tests.rs:15-22 on the direct model."""
    model = ro.direct_model()
    meas = MeasurementStack(frames=np.ones((5, 8, 8)))
    assert int(oracle.schedule(model, meas, "brightfield_first")[0]) == 2
    assert oracle.schedule(model, meas, "spiral_out").tolist() == [2, 4, 3, 0, 1]


def weighted_problem():
    """tests/schedule_tests.rs:73-100: frame levels 4, 9, 100, 16, 1; a 1e12
    outlier in frame 1 under a mask; frame 2 weight 0 and fully masked;
    frame 3 weight 0.5."""
    model = ro.direct_model()
    frames = np.stack([np.full((8, 8), v) for v in (4.0, 9.0, 100.0, 16.0, 1.0)])
    frames[1, 0, 0] = 1e12
    weights = np.array([1.0, 1.0, 0.0, 0.5, 1.0])
    masks = np.ones((5, 8, 8), np.uint8)
    masks[1, 0, 0] = 0
    masks[2] = 0
    return model, MeasurementStack(frames=frames, weights=weights, masks=masks)


def test_snr_weighted_order_matches_reference():
    """tests/schedule_tests.rs:37-48: [1, 0, 3, 4, 2], stable ties."""
    model, meas = weighted_problem()
    assert oracle.schedule(model, meas, "snr_weighted").tolist() == [1, 0, 3, 4, 2]
    # the masked outlier is what keeps frame 1 from dominating by far: unmasked,
    # the order of the top two frames is unchanged but for a different reason
    assert oracle.schedule(model, meas, "sequential").tolist() == [0, 1, 2, 3, 4]


# ---- Pupil::circular aberrations (src/model/pupil.rs:76-124) --------------------

N_AB = 32


def pupil_of(**kw):
    optics = ro.base_optics()
    optics.defocus_distance = kw.pop("defocus", None)
    ab = kw.pop("aberration", None)
    optics.pupil_aberration = ab
    m = oracle.compile_model(optics, ro.base_array(), (N_AB, N_AB), "smooth")
    return m.pupil, m.support.astype(bool)


def axis_samples(p, support):
    """Pupil values along +x (row c, columns c + j), +y (rows c + j, column
    c), -x and the +x+y diagonal, for j = 1.. while inside the support."""
    c = N_AB // 2
    out = {"+x": [], "+y": [], "-x": [], "diag": []}
    for j in range(1, c):
        if support[c, c + j]:
            out["+x"].append(p[c, c + j])
            out["+y"].append(p[c + j, c])
            out["-x"].append(p[c, c - j])
        if support[c + j, c + j]:
            out["diag"].append(p[c + j, c + j])
    return {k: np.array(v) for k, v in out.items()}


def test_unaberrated_pupil_is_real_unit_inside_support():
    p, s = pupil_of()
    assert np.all(p[s] == 1.0) and np.all(p[~s] == 0.0)


def test_spherical_term_is_one_at_dc_and_symmetric():
    """spherical * (6 rho^4 - 6 rho^2 + 1): exactly the coefficient at rho = 0,
    independent of the direction of k."""
    p, s = pupil_of(aberration=(0.0, 0.0, 0.3, 0.0))
    c = N_AB // 2
    assert np.angle(p[c, c]) == pytest.approx(0.3, abs=1e-15)
    a = axis_samples(p, s)
    np.testing.assert_allclose(np.angle(a["+x"]), np.angle(a["+y"]), atol=1e-15)
    np.testing.assert_allclose(np.abs(p[s]), 1.0, atol=1e-15)


def test_astigmatism_term_orientation_and_scaling():
    """astigmatism * rho^2 * cos(2 theta), theta = atan2(ky, kx): positive on
    the kx axis, the opposite on the ky axis, zero on the diagonal, quadratic
    in the radius, and at most the coefficient (rho <= 1 in the support)."""
    p, s = pupil_of(aberration=(0.2, 0.0, 0.0, 0.0))
    a = axis_samples(p, s)
    px, py = np.angle(a["+x"]), np.angle(a["+y"])
    assert np.all(px > 0)
    np.testing.assert_allclose(py, -px, atol=1e-15)
    np.testing.assert_allclose(np.angle(a["diag"]), 0.0, atol=1e-15)
    np.testing.assert_allclose(px[1::2][: len(px) // 2], 4.0 * px[: len(px) // 2][: len(px[1::2])], rtol=1e-12)
    assert px.max() <= 0.2 and px.max() > 0.8 * 0.2


def test_coma_term_is_odd_in_kx():
    """coma * (3 rho^3 - 2 rho) * cos(theta): odd under kx -> -kx, zero on the
    ky axis, negative near the centre (3 rho^3 < 2 rho for rho^2 < 2/3)."""
    p, s = pupil_of(aberration=(0.0, 0.25, 0.0, 0.0))
    a = axis_samples(p, s)
    np.testing.assert_allclose(np.angle(a["-x"]), -np.angle(a["+x"]), atol=1e-15)
    np.testing.assert_allclose(np.angle(a["+y"]), 0.0, atol=1e-15)
    assert np.angle(a["+x"][0]) < 0


def test_apodization_is_gaussian_in_rho():
    """exp(-edge_apodization * rho^2): 1 at DC, isotropic, log quadratic."""
    p, s = pupil_of(aberration=(0.0, 0.0, 0.0, 0.7))
    c = N_AB // 2
    assert abs(p[c, c]) == pytest.approx(1.0, abs=1e-15)
    a = axis_samples(p, s)
    lx = np.log(np.abs(a["+x"]))
    np.testing.assert_allclose(lx, np.log(np.abs(a["+y"])), atol=1e-15)
    n = len(lx) // 2
    np.testing.assert_allclose(lx[1::2][:n], 4.0 * lx[:n], rtol=1e-12)
    assert np.all(np.abs(np.angle(p[s])) <= 1e-15)


def test_defocus_is_the_paraxial_propagation_phase():
    """-defocus |k|^2 / (2 k_medium) with k_medium = 2 pi n / lambda and the
    pupil sampling dk = 2 pi mag / (n_pix * pixel): along kx the phase of
    sample j is -d (j dk)^2 lambda / (4 pi); a negative defocus gives a
    positive phase."""
    d = -24e-6
    p, s = pupil_of(defocus=d)
    o = ro.base_optics()
    dk = 2.0 * math.pi * o.magnification / (N_AB * o.camera_pixel_size)
    a = axis_samples(p, s)
    j = np.arange(1, len(a["+x"]) + 1)
    expect = -d * (j * dk) ** 2 * o.wavelength_vacuum_m / (4.0 * math.pi)
    np.testing.assert_allclose(np.angle(a["+x"]), expect, rtol=1e-12)
    np.testing.assert_allclose(np.angle(a["+y"]), expect, rtol=1e-12)


# ---- EPRY pupil recovery (tests/reconstruction_tests.rs:2923) -------------------

def pupil_phase_rmse(recovered, truth, support):
    """tests/reconstruction_tests.rs:3128-3149: global phase from the
    support overlap, wrapped phase differences."""
    cross = np.sum(recovered[support] * np.conj(truth[support]))
    corr = np.exp(-1j * np.angle(cross))
    d = np.mod(np.angle(recovered[support] * corr) - np.angle(truth[support]) + math.pi, 2 * math.pi) - math.pi
    return float(np.sqrt(np.mean(d * d)))


def compare_pupils_phase(reference, candidate, support):
    """compare_pupils (src/metrics/model.rs:19-72) phase_rmse: least-squares
    complex alignment of the candidate, wrapped phase error."""
    num = np.sum(np.conj(candidate[support]) * reference[support])
    den = np.sum(np.abs(candidate[support]) ** 2)
    align = num / den if den > np.finfo(float).eps else 1.0
    c = candidate[support] * align
    d = np.mod(np.angle(c) - np.angle(reference[support]) + math.pi, 2 * math.pi) - math.pi
    return float(np.sqrt(np.mean(d * d)))


def epry_experiment():
    assumed = ro.base_optics()
    true = Optics(532e-9, 0.10, 4.0, 6.5e-6, 1.0, 1.0, defocus_distance=-24e-6,
                  pupil_aberration=(0.2, 0.0, 0.0, 0.0))
    recon = oracle.compile_model(assumed, ro.base_array(), (16, 16), (32, 32))
    truth = oracle.compile_model(true, ro.base_array(), (16, 16), (32, 32))
    obj = ro.phase_disk((32, 32), 8.0, 0.9)
    frames = oracle.forward_stack(truth, ro.spectrum_of(obj))  # Simulator::new, no camera: ideal
    return recon, truth, MeasurementStack(frames=frames)


def test_epry_reduces_known_pupil_phase_error():
    recon, truth, meas = epry_experiment()
    support = truth.support.astype(bool)
    initial = pupil_phase_rmse(recon.pupil, truth.pupil, support)
    res = oracle.run(recon, meas, Epry(iterations=40, pupil_step=0.05).to_c(), 40)
    recovered = compare_pupils_phase(truth.pupil, res["recovered_pupil"], support)
    # without the pupil update the same run leaves the error where it was:
    # the decrease is the pupil gradient step's (common.rs:271-300)
    fixed = oracle.run(recon, meas, Epry(iterations=40, pupil_step=0.05, recover_pupil=False).to_c(), 40)
    unchanged = compare_pupils_phase(truth.pupil, fixed["recovered_pupil"], support)
    print("pupil phase rmse", initial, "->", recovered, "(no pupil update:", unchanged, ")")
    assert initial > 0.05
    assert recovered < initial
    assert unchanged == pytest.approx(initial, rel=1e-9)
