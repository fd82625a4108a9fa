"""Coordinate system and boundary setup (src/boost_fractor.jl:16-108).

Host-side data model only: these constructors build the small descriptor
arrays (grids, distances, permittivities, reflectivities) that every engine
call takes.  Array layout: Julia's E[x, y] is a numpy array of shape (ny, nx)
(x fastest), so `arr.ravel()[i]` is Julia's linear index `arr[i+1]`; Julia's
relative_surfaces (nreg, Nx, Ny) is (nreg, ny, nx) here.
"""
import cmath
import math

import numpy as np

from .jlrange import julia_range

SPEED_OF_LIGHT = 299792458.0


def wavelength(frequency):
    """src/boost_fractor.jl:16-19"""
    return SPEED_OF_LIGHT / float(frequency)


class CoordinateSystem:
    """src/boost_fractor.jl:21-28: real-space grids X, Y and k-space grids kX, kY."""

    def __init__(self, X, Y, kX, kY):
        self.X = np.ascontiguousarray(X, dtype=np.float64)
        self.Y = np.ascontiguousarray(Y, dtype=np.float64)
        self.kX = np.ascontiguousarray(kX, dtype=np.float64)
        self.kY = np.ascontiguousarray(kY, dtype=np.float64)

    @property
    def nx(self):
        return len(self.X)

    @property
    def ny(self):
        return len(self.Y)

    def __repr__(self):
        return f"CoordinateSystem(nx={self.nx}, ny={self.ny})"


def _as_grid(v):
    if isinstance(v, tuple) and len(v) == 3:
        return julia_range(*v)
    return np.asarray(v, dtype=np.float64)


def get_kspace_coords(real_space_coords):
    """src/boost_fractor.jl:42-47: dk = 2pi/(2 max X); range -dk(N-1)/2 : dk : dk(N-1)/2."""
    X = np.asarray(real_space_coords, dtype=np.float64)
    minimum_kspace = 2 * math.pi / (float(np.max(X)) * 2)
    maximum_kspace = minimum_kspace * (len(X) - 1) / 2
    return julia_range(-maximum_kspace, minimum_kspace, maximum_kspace)


def SeedCoordinateSystem(X=(-0.5, 0.01, 0.5), Y=(-0.5, 0.01, 0.5)):
    """src/boost_fractor.jl:30-34.  X and Y are arrays or Julia range triples
    (start, step, stop) evaluated exactly as Julia's `start:step:stop`."""
    X = _as_grid(X)
    Y = _as_grid(Y)
    return CoordinateSystem(X, Y, get_kspace_coords(X), get_kspace_coords(Y))


class SetupBoundaries:
    """src/boost_fractor.jl:63-71.  Regions are numbered from the mirror side:
    distance/eps/tilts/surfaces have one entry per region, r one per boundary."""

    def __init__(self, distance, r, eps, relative_tilt_x, relative_tilt_y, relative_surfaces):
        self.distance = np.array(distance, dtype=np.float64)
        self.r = np.array(r, dtype=np.complex128)
        self.eps = np.array(eps, dtype=np.complex128)
        self.relative_tilt_x = np.array(relative_tilt_x, dtype=np.float64)
        self.relative_tilt_y = np.array(relative_tilt_y, dtype=np.float64)
        self.relative_surfaces = np.array(relative_surfaces, dtype=np.complex128)

    def copy(self):
        return SetupBoundaries(self.distance, self.r, self.eps, self.relative_tilt_x, self.relative_tilt_y,
                               self.relative_surfaces)


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch (src/boost_fractor.jl:105)."""


def _cast_real_eps(eps):
    """Real permittivities are stored as re - 0.0im (src/boost_fractor.jl:97-98) so
    that evanescent kz = sqrt(k0^2 - kt^2) takes the decaying branch."""
    eps = np.array(eps, dtype=np.complex128)
    real = eps.imag == 0.0
    v = eps.view(np.float64).reshape(-1, 2)
    v[real, 1] = -0.0
    return eps


def _csqrt(z):
    return cmath.sqrt(complex(z))


def SeedSetupBoundaries(coords, diskno=3, distance=None, epsilon=None, relative_tilt_x=None,
                        relative_tilt_y=None, relative_surfaces=None):
    """src/boost_fractor.jl:83-108."""
    n = 2 * diskno + 2
    if relative_tilt_x is None:
        relative_tilt_x = np.zeros(n)
    if relative_tilt_y is None:
        relative_tilt_y = np.zeros(n)
    if relative_surfaces is None:
        relative_surfaces = np.zeros((n, coords.ny, coords.nx), dtype=np.complex128)
    if distance is None:
        distance = [0.0] + [8e-3 if x % 2 == 1 else 1e-3 for x in range(1, 2 * diskno + 1)] + [0.0]
    if epsilon is None:
        epsilon = [math.nan] + [1.0 if x % 2 == 1 else 9.0 for x in range(1, 2 * diskno + 1)] + [1.0]
    eps = _cast_real_eps(epsilon)
    refl = [complex(1.0, 0.0)]
    for i in range(2, len(eps)):
        a, b = _csqrt(eps[i - 1]), _csqrt(eps[i])
        refl.append((a - b) / (a + b))
    relative_surfaces = np.asarray(relative_surfaces)
    if not (len(distance) == len(refl) + 1 == len(eps) == len(relative_tilt_x) == len(relative_tilt_y)
            == relative_surfaces.shape[0]):
        raise DimensionMismatch("the arrays in your SetupBoundaries objects don't fit together!")
    return SetupBoundaries(distance, np.array(refl, dtype=np.complex128), eps, relative_tilt_x, relative_tilt_y,
                           relative_surfaces)
