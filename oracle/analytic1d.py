"""CPU restatement of the 1D analytical disk model (src/Analytical1D.py) for the
parity tests.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker.  The product
(boostfractor.jl_amd/) never imports this module.

Scalar complex arithmetic per frequency point, numpy over frequencies.
Pinned in tests/test_analytic1d_cpu.py by energy conservation (|gamma| = 1 for
a lossless mirror-backed booster), the bare-mirror amplitude |ax| = 1, and the
boost factor of the transfer-matrix transformer (oracle/bf_oracle.c, one mode,
1D propagator) on the same boosters, lossy ones included.
"""
import numpy as np

c = 299792458.


def transform_surface(freq, gamma, ax, eps_i, eps_m, l, surfaceloss=0.0, ray=True):
    """src/Analytical1D.py:5-67; returns (gamma, ax) instead of mutating ax"""
    freq = np.asarray(freq, dtype=np.float64)
    si, sm = np.sqrt(eps_i), np.sqrt(eps_m)  # real or complex as given, like the reference
    r_int = np.asarray(gamma, dtype=np.complex128)
    g = 1.0 - surfaceloss
    r_s = 0.0
    if ray:
        r_s = (sm - si) / (sm + si)                                               # :8
        gam = g * r_s + g ** 2 * (1 + r_s) * (1 - r_s) * r_int / (1 + r_int * r_s * g)  # :15
    else:
        z = (1 + r_int) / (1 - r_int) * ((1 / si) / (1 / sm))                     # :22-30
        gam = ((z - 1) / (z + 1)) * (1 - surfaceloss)                             # :33
    wavel = c / (freq * sm)                                                       # :36
    gam = gam * np.exp(-1j * 4 * np.pi * l / wavel)                               # :37
    if ax is not None:
        den = eps_i * sm + eps_m * si                                             # :44
        ax_i = si * (1 - eps_m / eps_i) / den                                     # :45
        ax_m = sm * (1 - eps_i / eps_m) / den                                     # :46
        ax_i = ax_i * g * r_int * (1 - r_s) / (1 + g * r_int * r_s)               # :49
        ax = ax * g * (1 - r_s) / (1 + r_int * r_s * g)                           # :53
        ax = ax + ax_i + ax_m                                                     # :56
        ax = ax * np.exp(-1j * 2 * np.pi * l / wavel)                             # :60
    return gam, ax


def disk_system(freq, eps_disk, eps_gap, spacings, num_disk, thickness=1e-3, mirror=True, surfaceloss=None):
    """src/Analytical1D.py:82-131 with the permittivities given (lossy_eps applied by the caller)"""
    freq = np.atleast_1d(np.asarray(freq, dtype=np.float64))
    ax = np.zeros(freq.shape, dtype=np.complex128)
    gam = np.zeros(freq.shape, dtype=np.complex128)
    sl = (lambda i: 0.0) if surfaceloss is None else (lambda i: surfaceloss[i])
    if mirror:
        gam, ax = transform_surface(freq, 0.0, ax, 1.0, 1e20, 100e-3)
    gam, ax = transform_surface(freq, gam, ax, 1e20 if mirror else 1.0, eps_gap, spacings[0], sl(0))
    for i in range(num_disk):
        gam, ax = transform_surface(freq, gam, ax, eps_gap, eps_disk, thickness, sl(2 * i + 1))
        gam, ax = transform_surface(freq, gam, ax, eps_disk, eps_gap, spacings[i + 1], sl(2 * i + 2))
    return gam, ax
