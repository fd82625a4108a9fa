"""Seeded synthetic workloads: booster geometries and grids.

These build the inputs of the benchmark and parity workloads (20-disk
boosters with jittered spacings, power-of-two grids spanning 1 m).  The random
numbers come from SplitMix64 (Steele, Lea, Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014) so every consumer (device
engine, CPU oracle, tests) sees bit-identical geometry.
"""
import math

import numpy as np

from .setup import CoordinateSystem, SeedSetupBoundaries, get_kspace_coords

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & _MASK

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self):
        """U[0,1) from the top 53 bits"""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def normal(self):
        """Box-Muller, one variate per call"""
        u1 = self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(1.0 - u1)) * math.cos(2.0 * math.pi * u2)


def grid_coords(N, span=1.0):
    """N x N grid over `span` metres: X = ((0:N-1) .- N/2) ./ N * span (dx = span/N)."""
    X = np.array([((j - N // 2) / N) * span for j in range(N)], dtype=np.float64)
    return CoordinateSystem(X, X.copy(), get_kspace_coords(X), get_kspace_coords(X))


def booster(coords, ndisk=20, eps_disk=24.0, thickness=1e-3, spacing=8e-3, jitter=0.3e-3, seed=1000,
            spacings=None):
    """Mirror + ndisk dielectric disks: eps = [NaN, (1, eps_disk) x ndisk, 1],
    distance = [0, s1, t, s2, t, ..., s_d, t, 0], s_i = spacing + U(-jitter, jitter)."""
    if spacings is None:
        rng = SplitMix64(seed)
        spacings = [spacing + (2.0 * rng.uniform() - 1.0) * jitter for _ in range(ndisk)]
    distance = [0.0]
    eps = [math.nan]
    for s in spacings:
        distance += [float(s), float(thickness)]
        eps += [1.0, float(eps_disk)]
    distance.append(0.0)
    eps.append(1.0)
    return SeedSetupBoundaries(coords, diskno=ndisk, distance=distance, epsilon=eps)


def beta2_dancer(E, coords, diskR):
    """Boost factor of an emitted field: sum|E|^2 dx dy / (pi R^2) (docs/src/3d_algorithms.md:22-36)."""
    dx = abs(coords.X[1] - coords.X[0]) if coords.nx > 1 else 1.0
    dy = abs(coords.Y[1] - coords.Y[0]) if coords.ny > 1 else 1.0
    return np.sum(np.abs(E) ** 2, axis=(-2, -1)) * dx * dy / (math.pi * diskR ** 2)
