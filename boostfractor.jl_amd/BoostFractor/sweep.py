"""Disk-position sweep around the transformer (examples/transformer_optim_utilities.jl).

calc_propagation_matrices_grid builds the propagation matrices of a booster
over a grid of gap shifts (and tilts) on the device; SpacingInterpolation keeps
the cubic B-spline of that grid along the spacing axis in device memory
(construct_prop_matrix_interpolation), and evaluates it and the transfer-matrix
chain for one or many candidate shift vectors per call.  The cost helpers
mirror the example's optimizer objective.
"""
import ctypes

import numpy as np

from . import _lib as L
from .propagators import prop_kind, propagator
from .transformer import _modes_struct, _nan_to_1e30, axion_induced_modes, transformer


def _as_grid(v):
    return np.ascontiguousarray(np.atleast_1d(np.asarray(v, dtype=np.float64)))


def calc_propagation_matrices_grid(sbdry, coords, modes, spacing_grid, frequencies, tilt_x_grid=0, tilt_y_grid=0,
                                   prop=propagator, diskR=0.15):
    """:11-52 -> complex (n_region, n_freq, n_spacing, n_tilt_x, n_tilt_y, n_mode, n_mode),
    [r, f, s, tx, ty] = the n x n propagation matrix of region r."""
    m = L.Marshalled()
    c = m.coords(coords)
    s = m.setup(sbdry, coords.ny, coords.nx)
    md = _modes_struct(m, modes)
    sg, gx, gy = (m.hold(_as_grid(v)) for v in (spacing_grid, tilt_x_grid, tilt_y_grid))
    fr = m.hold(_as_grid(frequencies))
    n, nreg = modes.n, len(sbdry.distance)
    out = np.zeros((len(sg), len(gx), len(gy), len(fr), nreg, n, n), dtype=np.complex128)
    ctx = L.context()
    L.check(L.lib().bf_propagation_matrices_grid(ctx.handle, ctypes.byref(s), ctypes.byref(c), ctypes.byref(md),
                                                 L.cptr(sg), len(sg), L.cptr(gx), len(gx), L.cptr(gy), len(gy),
                                                 L.cptr(fr), len(fr), prop_kind(prop), float(diskR), L.cptr(out)),
            ctx.handle)
    # [s, tx, ty, f, r, j, k] (column-major matrices) -> [r, f, s, tx, ty, k, j]
    return out.transpose(4, 3, 0, 1, 2, 6, 5)


class SpacingInterpolation:
    """construct_prop_matrix_interpolation (:58-74): cubic B-spline
    (Interpolations.jl BSpline(Cubic(Natural(OnCell())))) of a propagation-matrix
    grid along its spacing axis, scaled to the uniform spacing grid, resident on
    the device.  Like the reference's interpolation object it is evaluated at
    tilt index (1, 1)."""

    def __init__(self, prop_matrix_grid, spacing_grid, ctx=None):
        g = np.asarray(prop_matrix_grid)
        if g.ndim != 7:
            raise L.DimensionMismatch(f"prop_matrix_grid has {g.ndim} dimensions, expected 7")
        sp = _as_grid(spacing_grid)
        nreg, nf, ns, _, _, n, n2 = g.shape
        if len(sp) != ns or n != n2:
            raise L.DimensionMismatch(f"spacing grid of {len(sp)} points for a grid of shape {g.shape}")
        if ns < 2:
            raise L.ArgumentError("the cubic spacing interpolation needs at least 2 grid points")
        step = (sp[-1] - sp[0]) / (ns - 1)
        if not step > 0 or not np.allclose(np.diff(sp), step, rtol=1e-9, atol=0.0):
            raise L.ArgumentError("spacing_grid must be a uniform increasing range")
        self.ctx = ctx or L.context()
        self.nreg, self.nf, self.ns, self.n = nreg, nf, ns, n
        # [r, f, s, k, j] at tilt (1, 1) -> [s, f, r, j, k]
        dev = np.ascontiguousarray(g[:, :, :, 0, 0].transpose(2, 1, 0, 4, 3), dtype=np.complex128)
        h = ctypes.c_void_p()
        L.check(L.lib().bf_spacing_itp_create(self.ctx.handle, L.cptr(dev), nreg, nf, ns, n, float(sp[0]),
                                              float(step), ctypes.byref(h)), self.ctx.handle)
        self.handle = h

    def bounds(self):
        """Interpolations.bounds along the spacing axis"""
        lo, hi = ctypes.c_double(), ctypes.c_double()
        L.check(L.lib().bf_spacing_itp_bounds(self.handle, ctypes.byref(lo), ctypes.byref(hi)), self.ctx.handle)
        return lo.value, hi.value

    @property
    def ndisk(self):
        return (self.nreg - 2) // 2

    def _shifts(self, dist_shift):
        d = np.asarray(dist_shift, dtype=np.float64)
        single = d.ndim == 1
        d = np.atleast_2d(d)
        if d.shape[1] < self.ndisk:
            raise L.BoundsError(f"{d.shape[1]} shifts for {self.ndisk} gaps")
        d = np.ascontiguousarray(d[:, :self.ndisk])
        lo, hi = self.bounds()
        if not (np.all(d >= lo) and np.all(d <= hi) and lo <= 0.0 <= hi):
            # solids and the last gap are evaluated at a shift of 0 (:85)
            raise L.BoundsError(f"spacing shifts outside the interpolation bounds [{lo}, {hi}]")
        return single, d

    def matrices(self, dist_shift):
        """interpolate_prop_matrix for one shift vector (n_region, n_freq, n, n) or a
        batch (ncand, ndisk) -> (ncand, n_region, n_freq, n, n)"""
        single, d = self._shifts(dist_shift)
        out = np.zeros((len(d), self.nf, self.nreg, self.n, self.n), dtype=np.complex128)
        L.check(L.lib().bf_spacing_itp_matrices(self.handle, len(d), L.cptr(d), L.cptr(out)), self.ctx.handle)
        res = out.transpose(0, 2, 1, 4, 3)
        return res[0] if single else res

    def transformer(self, dist_shift, eps, emit, reflect=None):
        """transformer on the interpolated matrices of each candidate:
        boost (ncand, n_freq, n) [and refl], or (n_freq, n) for one shift vector"""
        single, d = self._shifts(dist_shift)
        ep = L.as_c128(eps)
        if len(ep) != self.nreg:
            raise L.DimensionMismatch(f"eps of length {len(ep)} for {self.nreg} regions")
        em = L.as_c128(np.asarray(emit).reshape(-1))
        rf = L.as_c128(np.broadcast_to(reflect, (self.n,))) if reflect is not None else None
        boost = np.zeros((len(d), self.nf, self.n), dtype=np.complex128)
        refl = np.zeros_like(boost) if reflect is not None else None
        L.check(L.lib().bf_spacing_itp_transformer(self.handle, len(d), L.cptr(d), L.cptr(ep), L.cptr(em),
                                                   L.cptr(rf), L.cptr(boost), L.cptr(refl)), self.ctx.handle)
        if single:
            boost = boost[0]
            refl = refl[0] if refl is not None else None
        return boost if reflect is None else (boost, refl)

    def close(self):
        if getattr(self, "handle", None):
            L.lib().bf_spacing_itp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def construct_prop_matrix_interpolation(prop_matrix_grid, spacing_grid):
    """:58-74"""
    return SpacingInterpolation(prop_matrix_grid, spacing_grid)


def interpolate_prop_matrix(itp, dist_shift):
    """:79-91 -> (n_region, n_freq, n, n): [r][f] is region r's matrix at frequency f"""
    return itp.matrices(dist_shift)


def _per_freq_sets(prop_matrices_set):
    P = np.asarray(prop_matrices_set)
    return [[P[r, f] for r in range(P.shape[0])] for f in range(P.shape[1])]


def calc_boostfactor_modes(sbdry, coords, modes, frequencies, prop_matrices_set, prop=propagator, diskR=0.15):
    """:99-109 -> EoutModes0 (1, n_modes, n_freq)"""
    fr = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
    boost = transformer(sbdry, coords, modes, f=fr, prop=prop, propagation_matrices=_per_freq_sets(prop_matrices_set),
                        diskR=diskR)
    return np.ascontiguousarray(np.asarray(boost).reshape(len(fr), -1).T[None])


def calc_modes(sbdry, coords, modes, frequencies, prop_matrices_set, reflect, prop=propagator, diskR=0.15):
    """:115-125 -> EoutModes0 (2, n_modes, n_freq): boost and refl"""
    fr = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
    boost, refl = transformer(sbdry, coords, modes, f=fr, prop=prop,
                              propagation_matrices=_per_freq_sets(prop_matrices_set), diskR=diskR, reflect=reflect)
    return np.stack([np.asarray(boost).reshape(len(fr), -1).T, np.asarray(refl).reshape(len(fr), -1).T])


def p_norm(X, p):
    """:211-215: magnitude * (mean((X/magnitude)^p))^(1/p), magnitude = max(X)"""
    X = np.asarray(X, dtype=np.float64)
    mag = np.max(X)
    return mag * (np.sum((X / mag) ** p) / len(X)) ** (1.0 / p)


def soft_box_penalty(shift, bound_hard):
    """:150-161"""
    depth = (bound_hard[1] - bound_hard[0]) * 0.05
    lo, hi = bound_hard[0] + depth, bound_hard[1] - depth
    shift = np.asarray(shift, dtype=np.float64)
    total = float(np.sum(shift))
    excess_pos = max(float(np.max(shift)), total) - hi
    excess_neg = -(min(float(np.min(shift)), total) - lo)
    excess = max(excess_pos, excess_neg, 0.0)
    return 1.0 - (excess / depth) ** 6


def calc_boostfactor_cost_batch(dist_shifts, itp, frequencies, sbdry, coords, modes, m_reflect, prop=propagator,
                                diskR=0.15):
    """calc_boostfactor_cost (:130-145) for many candidate shift vectors
    (ncand, ndisk) in one device call: candidates outside the interpolation
    bounds cost 1000.0; the others -p_norm(coupled power, -20) * penalty."""
    d = np.atleast_2d(np.asarray(dist_shifts, dtype=np.float64))
    lo, hi = itp.bounds()
    ok = np.all((d > lo) & (d < hi), axis=1)
    cost = np.full(len(d), 1000.0)
    if not np.any(ok):
        return cost
    _nan_to_1e30(sbdry)  # transformer mutates the setup (src/boost_fractor_transformer.jl:289)
    fr = np.atleast_1d(np.asarray(frequencies, dtype=np.float64))
    emit = axion_induced_modes(coords, modes, B=None, diskR=diskR, f=float(fr[0]))
    boost = itp.transformer(d[ok], sbdry.eps, emit)  # (nok, n_freq, n)
    m = np.asarray(m_reflect).reshape(-1)
    cpld = np.abs(np.sum(np.conj(boost) * m, axis=2)) ** 2  # (nok, n_freq)
    pen = np.array([soft_box_penalty(x, (lo, hi)) for x in d[ok]])
    cost[ok] = -np.array([p_norm(c, -20) for c in cpld]) * pen
    return cost


def calc_boostfactor_cost(dist_shift, itp, frequencies, sbdry, coords, modes, m_reflect, prop=propagator, diskR=0.15):
    """:130-145 (one candidate)"""
    return float(calc_boostfactor_cost_batch(np.asarray(dist_shift, dtype=np.float64)[None], itp, frequencies, sbdry,
                                             coords, modes, m_reflect, prop=prop, diskR=diskR)[0])


def copy_setup_boundaries(sbdry, coords=None):
    """:200-204"""
    return sbdry.copy()
