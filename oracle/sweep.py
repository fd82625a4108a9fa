"""CPU restatement of the disk-position sweep helpers
(examples/transformer_optim_utilities.jl) for the parity tests.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
(boostfractor.jl_amd/) never imports this module.

The spacing interpolation is Interpolations.jl's
BSpline(Cubic(Natural(OnCell()))) (examples/transformer_optim_utilities.jl:65-69).
Interpolations.jl is not part of the reference and no version is recorded
for the example's environment; this follows its published definition:
Natural is Line, i.e. zero second derivative at the outer cell edges
x = 1/2 and x = n + 1/2 (1-based grid index x), with one padding coefficient
on each side and the boundary cells' polynomials extended to those edges.
The system is solved here in its unreduced (n+2) x (n+2) form with a dense
solver, independently of the engine's reduced tridiagonal elimination.
Pinned in tests/test_sweep_cpu.py by an independent construction (scipy
B-spline collocation on the knots 1/2, 2, ..., n-1, n+1/2 with s'' = 0 at
both ends), reproduction of linear data and interpolation at the nodes.
"""
import numpy as np


def spline_weights(d):
    """cubic B-spline weights of the four coefficients around a cell, offset d"""
    e = 1.0 - d
    return np.array([e ** 3 / 6.0, 2.0 / 3.0 - d * d + d ** 3 / 2.0, 2.0 / 3.0 - e * e + e ** 3 / 2.0, d ** 3 / 6.0])


def spline_coefficients(y):
    """padded coefficients c_0..c_{n+1} along axis 0 of y (n >= 2 values)"""
    y = np.asarray(y)
    n = y.shape[0]
    A = np.zeros((n + 2, n + 2))
    A[0, :4] = [3.0, -7.0, 5.0, -1.0]            # s''(1/2) = 0 on the first cell's polynomial
    for i in range(1, n + 1):
        A[i, i - 1:i + 2] = [1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0]
    A[n + 1, n - 2:n + 2] = [-1.0, 5.0, -7.0, 3.0]  # s''(n + 1/2) = 0 on the last cell's polynomial
    rhs = np.zeros((n + 2,) + y.shape[1:], dtype=np.result_type(y, np.float64))
    rhs[1:n + 1] = y
    flat = rhs.reshape(n + 2, -1)
    return np.linalg.solve(A, flat).reshape(rhs.shape)


def spline_eval(c, x):
    """value at the 1-based grid coordinate x (1/2 <= x <= n + 1/2)"""
    n = c.shape[0] - 2
    i = int(min(max(np.floor(x), 1), n - 1))
    w = spline_weights(x - i)
    return sum(w[q] * c[i - 1 + q] for q in range(4))


class SpacingInterpolation:
    """construct_prop_matrix_interpolation (:58-74) over a uniform spacing grid:
    grid[r, f, s, k, j] -> coefficients along s"""

    def __init__(self, grid, spacing_grid):
        g = np.asarray(spacing_grid, dtype=np.float64)
        self.first = g[0]
        self.step = (g[-1] - g[0]) / (len(g) - 1)
        self.ns = len(g)
        self.coef = spline_coefficients(np.moveaxis(np.asarray(grid), 2, 0))  # (ns+2, r, f, k, j)

    def bounds(self):
        return self.first - 0.5 * self.step, self.first + (self.ns - 1) * self.step + 0.5 * self.step

    def __call__(self, r, f, shift):
        """itp(r, f, shift, 1, 1, :, :) with 0-based r, f"""
        lo, hi = self.bounds()
        if not lo <= shift <= hi:
            raise IndexError(f"BoundsError: {shift} outside [{lo}, {hi}]")
        x = (shift - self.first) / self.step + 1.0
        return spline_eval(self.coef[:, r, f], x)


def interpolate_prop_matrix(itp, dist_shift):
    """:79-91 -> (n_region, n_freq, n, n)"""
    n_region, n_freq = itp.coef.shape[1], itp.coef.shape[2]
    out = []
    for r in range(1, n_region + 1):
        x = 0.0 if (r + 1) % 2 == 0 or r == n_region else float(dist_shift[r // 2 - 1])
        out.append([itp(r - 1, f, x) for f in range(n_freq)])
    return np.array(out)


def p_norm(X, p):
    """:211-215"""
    X = np.asarray(X, dtype=np.float64)
    m = np.max(X)
    return m * (np.sum((X / m) ** p) / len(X)) ** (1.0 / p)


def soft_box_penalty(shift, bound_hard):
    """:150-161"""
    depth = (bound_hard[1] - bound_hard[0]) * 0.05
    lo, hi = bound_hard[0] + depth, bound_hard[1] - depth
    ext = list(shift) + [float(np.sum(shift))]
    excess = max(max(ext) - hi, -(min(ext) - lo), 0.0)
    return 1.0 - (excess / depth) ** 6
