"""GPU parity of the disk-position sweep (examples/transformer_optim_utilities.jl):
the propagation-matrix grid against the oracle's propagation_matrix, the
device spline against the oracle spline, and the interpolated transformer and
optimizer cost against the oracle transformer."""
import numpy as np
import pytest

import BoostFractor as B
from BoostFractor import workloads as W
from helpers import DISKR, O, ocoords, rel
from oracle import sweep as OS

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _setup(N=32, ndisk=2, seed=1400):
    coords = W.grid_coords(N)
    sb = W.booster(coords, ndisk=ndisk, seed=seed)
    pat, kt = W.lg_modes(coords, pmax=1, lmax=1, w0=0.1)
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=2, Lmax=1, pattern_input=pat, kT_arr=kt)
    return coords, sb, modes


def _omat(sb, coords, modes, r, f, dz_shift=0.0, tx_shift=0.0, ty_shift=0.0):
    pat, kt = modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1)
    return O.propagation_matrix(sb.distance[r] + dz_shift, DISKR, sb.eps[r], sb.relative_tilt_x[r] + tx_shift,
                                sb.relative_tilt_y[r] + ty_shift, sb.relative_surfaces[r], B.wavelength(f),
                                ocoords(coords), pat, kt, O.PROP_PROPAGATOR)


def test_propagation_matrices_grid_matches_oracle():
    """gaps shifted by every spacing and tilt of the grid, solids unshifted;
    the untilted gaps take the Parseval path, the tilted ones explicit propagation"""
    coords, sb, modes = _setup()
    spacing = [-2e-4, 0.0, 3e-4]
    tgx, tgy = [0.0, 1e-4], [0.0]
    freqs = [19.9e9, 20.2e9]
    grid = B.calc_propagation_matrices_grid(sb, coords, modes, spacing, freqs, tilt_x_grid=tgx, tilt_y_grid=tgy,
                                            diskR=DISKR)
    nreg = len(sb.distance)
    assert grid.shape == (nreg, 2, 3, 2, 1, modes.n, modes.n)
    assert np.all(np.isnan(grid[0]))  # the mirror (eps NaN), as in the reference
    for fi, f in enumerate(freqs):
        for si, sp in enumerate(spacing):
            for xi, tx in enumerate(tgx):
                for r in range(1, nreg):
                    gap = r % 2 == 1
                    want = _omat(sb, coords, modes, r, f, sp if gap else 0.0, tx if gap else 0.0)
                    got = grid[r, fi, si, xi, 0]
                    assert rel(got, want) < TOL, (r, fi, si, xi, rel(got, want))


def _grid(coords, sb, modes, spacing, freqs):
    return B.calc_propagation_matrices_grid(sb, coords, modes, spacing, freqs, diskR=DISKR)


def test_spacing_interpolation_matches_oracle_spline():
    coords, sb, modes = _setup(ndisk=3)
    spacing = np.linspace(-3e-4, 3e-4, 7)
    freqs = [19.8e9, 20.0e9, 20.3e9]
    grid = _grid(coords, sb, modes, spacing, freqs)
    itp = B.construct_prop_matrix_interpolation(grid, spacing)
    oitp = OS.SpacingInterpolation(grid[:, :, :, 0, 0], spacing)
    lo, hi = itp.bounds()
    assert lo == pytest.approx(-3.5e-4, rel=1e-12) and hi == pytest.approx(3.5e-4, rel=1e-12)
    cands = [[0.0, 0.0, 0.0], [1.37e-4, -2.9e-4, 0.6e-4], [lo, hi, -3.2e-4], [3.45e-4, -3.49e-4, 1e-4]]
    got = itp.matrices(np.array(cands))
    for c, d in enumerate(cands):
        want = OS.interpolate_prop_matrix(oitp, d)
        assert got[c].shape == want.shape
        for r in range(1, len(sb.distance)):
            assert rel(got[c][r], want[r]) < 1e-12, (c, r, rel(got[c][r], want[r]))
    # one vector -> (n_region, n_freq, n, n); at grid nodes the spline returns the grid
    one = B.interpolate_prop_matrix(itp, [spacing[1], spacing[4], spacing[6]])
    for r, s in [(1, 1), (3, 4), (5, 6), (7, 3)]:
        assert rel(one[r], grid[r, :, s, 0, 0]) < 1e-12, r
    with pytest.raises(B.BoundsError):
        itp.matrices([0.0, 4e-4, 0.0])


@pytest.mark.parametrize("chunk_bytes", [None, "20000"])
def test_spacing_interpolation_transformer(engine_options, chunk_bytes):
    """chain on the interpolated matrices of several candidates (one launch, or
    several chunks) equals transformer(...; propagation_matrices=...) and the oracle"""
    if chunk_bytes:
        engine_options(batch_bytes=int(chunk_bytes))
    coords, sb, modes = _setup(ndisk=2)
    spacing = np.linspace(-2e-4, 2e-4, 5)
    freqs = [19.9e9, 20.1e9]
    itp = B.construct_prop_matrix_interpolation(_grid(coords, sb, modes, spacing, freqs), spacing)
    emit = B.axion_induced_modes(coords, modes, diskR=DISKR)
    refl_in = np.linspace(1, 2, modes.n) + 0.2j
    cands = np.array([[0.0, 0.0], [1.1e-4, -0.7e-4], [-1.9e-4, 2.1e-4]])
    boost, refl = itp.transformer(cands, sb.eps, emit, reflect=refl_in)
    eps = sb.eps.copy()
    eps[np.isnan(eps.real)] = 1e30
    pat, kt = modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1)
    for c, d in enumerate(cands):
        P = itp.matrices(d)
        want_dev = B.calc_modes(sb.copy(), coords, modes, freqs, P, refl_in, diskR=DISKR)
        assert rel(boost[c], want_dev[0].T) < TOL and rel(refl[c], want_dev[1].T) < TOL
        for fi, f in enumerate(freqs):
            os_ = O.Setup(sb.distance, eps, None, sb.relative_tilt_x, sb.relative_tilt_y, None)
            wb, wr = O.transformer(os_, ocoords(coords), pat, kt, f, O.PROP_PROPAGATOR,
                                   prop_mats=[P[r, fi] for r in range(len(sb.distance))], diskR=DISKR, emit=emit,
                                   reflect=refl_in)
            assert rel(boost[c, fi], wb) < 1e-9 and rel(refl[c, fi], wr) < 1e-9, (c, fi)


def test_boostfactor_cost_matches_oracle_composition():
    coords, sb, modes = _setup(ndisk=2)
    spacing = np.linspace(-2e-4, 2e-4, 5)
    freqs = [19.95e9, 20.0e9, 20.05e9]
    grid = _grid(coords, sb, modes, spacing, freqs)
    itp = B.construct_prop_matrix_interpolation(grid, spacing)
    oitp = OS.SpacingInterpolation(grid[:, :, :, 0, 0], spacing)
    m_reflect = np.zeros(modes.n)
    m_reflect[1] = 1.0
    emit = B.axion_induced_modes(coords, modes, diskR=DISKR)
    eps = sb.eps.copy()
    eps[np.isnan(eps.real)] = 1e30
    pat, kt = modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1)
    lo, hi = itp.bounds()
    cands = np.array([[0.0, 0.0], [1.5e-4, -0.4e-4], [2.6e-4, 0.0], [0.9e-4, 1.6e-4]])
    costs = B.calc_boostfactor_cost_batch(cands, itp, freqs, sb.copy(), coords, modes, m_reflect, diskR=DISKR)
    assert costs[2] == 1000.0  # outside the interpolation bounds
    for c, d in enumerate(cands):
        if not np.all((d > lo) & (d < hi)):
            continue
        P = OS.interpolate_prop_matrix(oitp, d)
        os_ = O.Setup(sb.distance, eps, None, sb.relative_tilt_x, sb.relative_tilt_y, None)
        Eout = np.array([O.transformer(os_, ocoords(coords), pat, kt, f, O.PROP_PROPAGATOR,
                                       prop_mats=[P[r, fi] for r in range(len(sb.distance))], diskR=DISKR, emit=emit)
                         for fi, f in enumerate(freqs)])  # (n_freq, n)
        cpld = np.abs(np.sum(np.conj(Eout) * m_reflect, axis=1)) ** 2
        want = -OS.p_norm(cpld, -20) * OS.soft_box_penalty(d, (lo, hi))
        assert abs(costs[c] / want - 1) < 1e-9, (c, costs[c], want)
        one = B.calc_boostfactor_cost(d, itp, freqs, sb.copy(), coords, modes, m_reflect, diskR=DISKR)
        assert one == pytest.approx(costs[c], rel=1e-14)
