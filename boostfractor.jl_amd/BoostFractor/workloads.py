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

    def uniforms(self, count):
        """`count` successive uniform() values at once (same sequence)"""
        i = np.arange(1, count + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + i * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
        self.state = (self.state + count * 0x9E3779B97F4A7C15) & _MASK
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))

    def normals(self, count):
        """`count` successive normal() values at once (same sequence)"""
        u = self.uniforms(2 * count)
        return np.sqrt(-2.0 * np.log(1.0 - u[0::2])) * np.cos(2.0 * np.pi * u[1::2])


def grid_coords(N, span=1.0):
    """N x N grid over `span` metres: X = ((0:N-1) .- N/2) ./ N * span (dx = span/N)."""
    X = np.array([((j - N // 2) / N) * span for j in range(N)], dtype=np.float64)
    return CoordinateSystem(X, X.copy(), get_kspace_coords(X), get_kspace_coords(X))


def booster(coords, ndisk=20, eps_disk=24.0, thickness=1e-3, spacing=8e-3, jitter=0.3e-3, seed=1000,
            spacings=None, flat=False):
    """Mirror + ndisk dielectric disks: eps = [NaN, (1, eps_disk) x ndisk, 1],
    distance = [0, s1, t, s2, t, ..., s_d, t, 0], s_i = spacing + U(-jitter, jitter).
    flat=True: no surface stack at all (relative_surfaces of shape (0, ny, nx),
    which the engine reads as flat), instead of (nreg, ny, nx) zeros."""
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
    if not flat:
        return SeedSetupBoundaries(coords, diskno=ndisk, distance=distance, epsilon=eps)
    sb = SeedSetupBoundaries(coords, diskno=ndisk, distance=distance, epsilon=eps,
                             relative_surfaces=np.zeros((len(distance), 0, 0), dtype=np.complex128))
    sb.relative_surfaces = np.zeros((0, coords.ny, coords.nx), dtype=np.complex128)
    return sb


def _genlaguerre(p, a, x):
    """generalised Laguerre polynomial L_p^a(x) by its three-term recurrence"""
    l0 = np.ones_like(x)
    if p == 0:
        return l0
    l1 = 1.0 + a - x
    for k in range(1, p):
        l0, l1 = l1, ((2 * k + 1 + a - x) * l1 - (k + a) * l0) / (k + 1)
    return l1


def lg_modes(coords, pmax=4, lmax=4, w0=0.1):
    """Laguerre-Gauss modes LG_{p,l}, p = 0..pmax, l = -lmax..lmax, in the
    layout of SeedModes' pattern_input (src/boost_fractor_transformer.jl:46-57,
    mode m = p + 1): patterns (pmax+1, 2 lmax+1, 1, ny, nx), each
    (sqrt(2) r/w0)^|l| L_p^|l|(2 r^2/w0^2) exp(-r^2/w0^2) exp(-i l phi)
    normalised to sum |.|^2 = 1 on the grid, and kT_arr (pmax+1, 2 lmax+1)
    with the paraxial kT^2 = 4 (2p + |l| + 1) / w0^2.  The reference has no LG
    modes; this is the workload's own convention."""
    XX, YY = np.meshgrid(coords.X, coords.Y)
    r2 = XX ** 2 + YY ** 2
    phi = np.arctan2(YY, XX)
    M, nl = pmax + 1, 2 * lmax + 1
    pat = np.zeros((M, nl, 1, coords.ny, coords.nx), dtype=np.complex128)
    kt = np.zeros((M, nl), dtype=np.complex128)
    for p in range(M):
        for l in range(-lmax, lmax + 1):
            a = abs(l)
            u = (np.sqrt(2.0 * r2) / w0) ** a * _genlaguerre(p, a, 2.0 * r2 / w0 ** 2) * np.exp(-r2 / w0 ** 2)
            f = u * np.exp(-1j * l * phi)
            pat[p, l + lmax, 0] = f / np.sqrt(np.sum(np.abs(f) ** 2))
            kt[p, l + lmax] = math.sqrt(4.0 * (2 * p + a + 1)) / w0
    return pat, kt


def rough_surface(coords, rng, sigma=10e-6, corr=1e-2, trunc=10.0, diskR=0.15):
    """Surface height map (ny, nx) after the reference's surface_roughness
    (examples/transformer_optim_utilities.jl:166-195): white N(0,1) noise
    convolved with exp(-r^2/(2 corr^2)) (circular FFT convolution in place of
    conv2), rescaled to rms `sigma` over the grid, truncated at +-trunc*sigma,
    zero outside r <= diskR."""
    ny, nx = coords.ny, coords.nx
    w = rng.normals(ny * nx).reshape(ny, nx)
    XX, YY = np.meshgrid(coords.X, coords.Y)
    g = np.exp(-(XX ** 2 + YY ** 2) / (2.0 * corr ** 2))
    s = np.real(np.fft.ifft2(np.fft.fft2(w) * np.fft.fft2(np.fft.ifftshift(g))))
    s *= sigma / np.sqrt(np.mean(s ** 2))
    s = np.clip(s, -trunc * sigma, trunc * sigma)
    s[XX ** 2 + YY ** 2 > diskR ** 2] = 0.0
    return s


def rough_booster(coords, ndisk=20, seed=1002, tilt_sigma=50e-6, sigma=10e-6, corr=1e-2, diskR=0.15):
    """The 20-disk booster with per-disk random tilts N(0, tilt_sigma) in x and
    y and Gaussian surface roughness on every disk region (all from one
    SplitMix64 stream)."""
    sb = booster(coords, ndisk=ndisk, seed=seed)
    rng = SplitMix64(seed + 1)
    n = len(sb.distance)
    tx, ty = np.zeros(n), np.zeros(n)
    surf = np.zeros((n, coords.ny, coords.nx), dtype=np.complex128)
    for i in range(2, n - 1, 2):  # disk regions: eps = eps_disk at even indices 2 .. 2 ndisk
        tx[i] = tilt_sigma * rng.normal()
        ty[i] = tilt_sigma * rng.normal()
        surf[i] = rough_surface(coords, rng, sigma=sigma, corr=corr, diskR=diskR)
    sb.relative_tilt_x = tx
    sb.relative_tilt_y = ty
    sb.relative_surfaces = surf
    return sb


# disk-position sweep: (M, L) of the Bessel mode sets, n = M (2L + 1) = 9, 15, 25
SWEEP_MODESETS = [(3, 1), (5, 1), (5, 2)]


def sweep_geometries(coords, nconf, seed=5000, first=0):
    """geometries first .. first+nconf-1 of the disk-position sweep: disk count
    uniform in 10..80, mode set uniform over SWEEP_MODESETS, spacings
    8 mm + U(-0.5, 0.5) mm, eps 24, 1 mm disks, flat and untilted (one
    SplitMix64 stream, so every slice of the sweep is the same geometries)"""
    rng = SplitMix64(seed)
    geoms, set_of = [], []
    for c in range(first + nconf):
        nd = 10 + int(rng.uniform() * 71)
        ms = int(rng.uniform() * len(SWEEP_MODESETS))
        sp = 8e-3 + (2.0 * rng.uniforms(nd) - 1.0) * 0.5e-3
        if c < first:
            continue
        geoms.append(booster(coords, ndisk=nd, spacings=list(sp), flat=True))
        set_of.append(ms)
    return geoms, np.array(set_of, dtype=np.int32)


def beta2_dancer(E, coords, diskR):
    """Boost factor of an emitted field: sum|E|^2 dx dy / (pi R^2) (docs/src/3d_algorithms.md:22-36)."""
    dx = abs(coords.X[1] - coords.X[0]) if coords.nx > 1 else 1.0
    dy = abs(coords.Y[1] - coords.Y[0]) if coords.ny > 1 else 1.0
    return np.sum(np.abs(E) ** 2, axis=(-2, -1)) * dx * dy / (math.pi * diskR ** 2)
