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


def chain_sensitivity(Ps, eps, emit, trials=3, seed=0):
    """relative change of the boost vector when every propagation matrix is
    perturbed at the 1e-15 level: the intrinsic FP64 sensitivity of the chain"""
    n = len(emit)
    N = len(eps)

    def run(rng):
        T_ = np.eye(2 * n, dtype=complex)
        ab = np.zeros(n, complex)
        for s in range(N - 1, 0, -1):
            ir, ir1 = N - s, N - s - 1
            n1, n0 = np.sqrt(eps[ir1]), np.sqrt(eps[ir])
            ab += (1 / n1 ** 2 - 1 / n0 ** 2) / 2 * (T_[n:, :n] @ emit + T_[n:, n:] @ emit)
            P = Ps[ir]
            if rng is not None:
                P = P * (1 + 1e-15 * (rng.standard_normal(P.shape) + 1j * rng.standard_normal(P.shape)))
            a, b = (n1 + n0) / (2 * n1), (n1 - n0) / (2 * n1)
            Pi = np.linalg.inv(P)
            T_ = T_ @ np.block([[a * P, b * Pi], [b * P, a * Pi]])
        return -np.linalg.solve(T_[n:, n:], ab)

    b0 = run(None)
    rng = np.random.default_rng(seed)
    return max(np.max(np.abs(run(rng) - b0)) / np.max(np.abs(b0)) for _ in range(trials))
