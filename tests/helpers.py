"""Shared inputs for the tests: the reference's own test geometries
(test/runtests.jl) and oracle adapters."""
import json
import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

import BoostFractor as B
from oracle import oracle as O

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.json")))

DISKR = 0.15


def ref_coords():
    """test/runtests.jl:33-34: X = Y = -0.5:0.01:0.5 (101 points)."""
    return B.SeedCoordinateSystem(X=(-0.5, 0.01, 0.5), Y=(-0.5, 0.01, 0.5))


def ref_dx(coords):
    """test/runtests.jl:49: dx = 0.0001*floor(10000*(X[2]-X[1]))"""
    return 0.0001 * math.floor(10000.0 * (coords.X[1] - coords.X[0]))


def ref_setup_default(coords):
    """test/runtests.jl:38: SeedSetupBoundaries(coords, diskno=1) (eps 9)."""
    return B.SeedSetupBoundaries(coords, diskno=1)


def ref_setup_eps9(coords):
    """test/runtests.jl:89-93"""
    eps = np.array([np.nan, 1, 9, 1], dtype=complex)
    return B.SeedSetupBoundaries(coords, diskno=1, distance=np.array([0.0, 0.15, 0.15 / 3, 0.0]) * 1e-2,
                                 epsilon=eps)


def deps_distance(deps):
    """test/runtests.jl:65,107,156: [0, pi*3e-2/(2pi), deps*1e-2/(2pi), 0]"""
    return np.array([0.0, math.pi * 3e-2 / (2 * math.pi), deps * 1e-2 / (2 * math.pi), 0.0])


def with_distance(sb, distance):
    s = sb.copy()
    s.distance = np.asarray(distance, dtype=np.float64)
    return s


def ocoords(c):
    return O.Coords(c.X, c.Y, c.kX, c.kY)


def osetup(sb):
    surf = sb.relative_surfaces if np.any(sb.relative_surfaces != 0) else None
    return O.Setup(sb.distance, sb.eps, sb.r, sb.relative_tilt_x, sb.relative_tilt_y, surf)


def pmap(fn, items, workers=8):
    with ThreadPoolExecutor(max_workers=workers) as ex:
        return list(ex.map(fn, items))


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)
