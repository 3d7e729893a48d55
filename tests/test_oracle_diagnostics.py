"""The oracle's diagnostics metrics (CPU) against the reference's golden
values (tests/diagnostics_tests.rs:68-122) and closed forms."""
import math

import numpy as np
import pytest

import oracle


def test_raw_frame_statistics_match_known_values():
    # diagnostics_tests.rs:68-78
    s = oracle.intensity_stats([0.0, 1.0, 2.0, 3.0], saturation=2.0)
    assert abs(s["mean"] - 1.5) < 1e-14 and abs(s["std"] - math.sqrt(1.25)) < 1e-14
    assert s["min"] == 0.0 and s["max"] == 3.0 and s["sum"] == 6.0
    assert s["saturated_pixels"] == 2 and s["zero_pixels"] == 1
    assert oracle.intensity_stats([0.0, 1.0, 2.0, 3.0])["saturated_pixels"] == 0


def test_frame_comparison_matches_known_values():
    # diagnostics_tests.rs:80-95
    d = oracle.compare_intensity([1.0, 2.0], [2.0, 0.0], saturation=2.0)
    want = {"reference_sum": 3.0, "estimate_sum": 2.0, "residual_l1": 3.0, "residual_l2": math.sqrt(5.0),
            "residual_mean": -0.5, "residual_std": 1.5, "residual_max_abs": 2.0, "normalized_l2": 1.0}
    for k, v in want.items():
        assert abs(d[k] - v) < 1e-14, k
    assert d["saturated_pixels"] == 1


def test_frame_comparison_zero_reference_and_masks():
    # diagnostics_tests.rs:97-122
    d = oracle.compare_intensity([0.0, 0.0], [1.0, 2.0])
    assert math.isfinite(d["normalized_l2"]) and d["normalized_l2"] > 0
    d = oracle.compare_intensity([1.0, 100.0], [1.0, 0.0], mask=[1, 0])
    assert d["reference_sum"] == 1.0 and d["estimate_sum"] == 1.0 and d["residual_l2"] == 0.0
    assert d["normalized_l2"] == 0.0
    with pytest.raises(oracle.OracleError) as e:
        oracle.compare_intensity([1.0], [1.0], mask=[0])
    assert e.value.name == "valid_mask"
    with pytest.raises(oracle.OracleError):
        oracle.compare_intensity([np.nan], [1.0])
    with pytest.raises(oracle.OracleError):
        oracle.intensity_stats([1.0, np.inf])


def test_frame_objective_closed_form():
    # runner.rs:361-373: mean of (sqrt p - sqrt y)^2 over unmasked pixels
    p, y = np.array([4.0, 9.0, -1.0, 16.0]), np.array([1.0, 4.0, 4.0, 0.0])
    assert oracle.frame_objective(p, y) == pytest.approx((1 + 1 + 4 + 16) / 4, abs=1e-15)
    assert oracle.frame_objective(p, y, mask=[1, 0, 1, 0]) == pytest.approx((1 + 4) / 2, abs=1e-15)
    assert oracle.frame_objective(p, y, mask=[0, 0, 0, 0]) == 0.0
