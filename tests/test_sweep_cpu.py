"""Disk-position sweep helpers on the CPU (examples/transformer_optim_utilities.jl):
the oracle's cubic spacing spline pinned by an independent construction, and
the host-side cost helpers against closed forms."""
import numpy as np
import pytest
from scipy.interpolate import BSpline

import BoostFractor as B
from oracle import sweep as OS


def _scipy_natural_oncell(y, xs):
    """cubic spline on [1/2, n+1/2] with knots 2..n-1, interpolating y_i at x = i
    and with zero second derivative at both ends, built by B-spline collocation"""
    n = len(y)
    t = np.r_[[0.5] * 4, np.arange(2, n, dtype=float), [n + 0.5] * 4]
    nb = len(t) - 4
    assert nb == n + 2
    rows = []
    for i in range(1, n + 1):
        rows.append([BSpline(t, np.eye(nb)[b], 3)(float(i)) for b in range(nb)])
    for x in (0.5, n + 0.5):
        rows.append([BSpline(t, np.eye(nb)[b], 3).derivative(2)(x) for b in range(nb)])
    A = np.array(rows)
    rhs = np.r_[y, [0.0, 0.0]]
    c = np.linalg.solve(A, rhs.astype(complex))
    return BSpline(t, c.real, 3)(xs) + 1j * BSpline(t, c.imag, 3)(xs)


@pytest.mark.parametrize("n", [2, 3, 4, 7, 12])
def test_oracle_spline_matches_independent_construction(n):
    rng = np.random.default_rng(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    xs = np.r_[0.5, n + 0.5, np.arange(1, n + 1), rng.uniform(0.5, n + 0.5, 40)]
    c = OS.spline_coefficients(y)
    got = np.array([OS.spline_eval(c, x) for x in xs])
    want = _scipy_natural_oncell(y, xs)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(y))
    # nodes
    assert np.max(np.abs(got[2:2 + n] - y)) < 1e-13 * np.max(np.abs(y))


def test_oracle_spline_reproduces_lines_and_is_linear_in_data():
    n = 9
    a, b = 0.3 - 1.1j, 2.5 + 0.4j
    y = a + b * np.arange(1, n + 1)
    c = OS.spline_coefficients(y)
    for x in np.linspace(0.5, n + 0.5, 37):
        assert abs(OS.spline_eval(c, x) - (a + b * x)) < 1e-12
    rng = np.random.default_rng(1)
    y1, y2 = rng.standard_normal(n), rng.standard_normal(n)
    c12 = OS.spline_coefficients(2 * y1 - 3 * y2)
    assert np.allclose(c12, 2 * OS.spline_coefficients(y1) - 3 * OS.spline_coefficients(y2), atol=1e-13)


def test_oracle_spline_matrix_series():
    """coefficients of a stack of matrices = per-element coefficients"""
    rng = np.random.default_rng(5)
    y = rng.standard_normal((6, 3, 2, 2)) + 1j * rng.standard_normal((6, 3, 2, 2))
    c = OS.spline_coefficients(y)
    for idx in [(0, 0, 0), (2, 1, 1), (1, 0, 1)]:
        ce = OS.spline_coefficients(y[(slice(None),) + idx])
        assert np.max(np.abs(c[(slice(None),) + idx] - ce)) < 1e-14


def test_p_norm_and_soft_box_penalty_closed_forms():
    assert B.p_norm([3.0, 3.0, 3.0], -20) == pytest.approx(3.0, rel=1e-15)
    # two points: max * ((1 + (x/max)^p)/2)^(1/p)
    want = 2.0 * ((1.0 + 0.5 ** -20) / 2.0) ** (-1.0 / 20)
    assert B.p_norm([1.0, 2.0], -20) == pytest.approx(want, rel=1e-14)
    assert OS.p_norm([1.0, 2.0], -20) == pytest.approx(want, rel=1e-14)
    # large negative p tends to the minimum
    assert B.p_norm([1.0, 5.0, 7.0], -200) == pytest.approx(1.0 * 3 ** (1 / 200), rel=1e-12)
    bounds = (-1.0, 1.0)  # soft depth 0.1: soft box [-0.9, 0.9]
    assert B.soft_box_penalty([0.1, -0.2, 0.3], bounds) == 1.0
    assert B.soft_box_penalty([0.95, -0.5], bounds) == pytest.approx(1.0 - 0.5 ** 6, rel=1e-12)
    # the sum of the shifts counts as well
    assert B.soft_box_penalty([0.5, 0.45], bounds) == pytest.approx(1.0 - 0.5 ** 6, rel=1e-12)
    assert B.soft_box_penalty([-0.95, 0.0], bounds) == pytest.approx(1.0 - 0.5 ** 6, rel=1e-12)
    for s in ([0.95, -0.5], [0.5, 0.45], [-0.2, -0.8]):
        assert B.soft_box_penalty(s, bounds) == pytest.approx(OS.soft_box_penalty(s, bounds), rel=1e-15)
