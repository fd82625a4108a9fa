"""The 1D analytical disk model (src/Analytical1D.py) on the device.

Same functions and keywords as the reference module; `disk_system` evaluates
every frequency in one launch and `disk_system_batch` every (spacing
candidate, frequency) point of an optimisation sweep in one launch.
Permittivities are prepared on the host (`lossy_eps`), every transform runs
in libboostfractor_hip.so.
"""
import ctypes

import numpy as np

from . import _lib as L

c = 299792458.


def _c2(z):
    z = complex(z)
    return np.array([z.real, z.imag], dtype=np.float64)


def lossy_eps(freq, eps, tand, SvenssonDjordjevic=False):
    """src/Analytical1D.py:69-80 (host arithmetic on the permittivity)"""
    if not SvenssonDjordjevic:
        return eps - 1j * eps * tand
    FreqForEpsrTanD, HighFreqForTanD, LowFreqForTanD = 1e9, 1e12, 1e3
    Lg = np.log((HighFreqForTanD + 1j * FreqForEpsrTanD) / (LowFreqForTanD + 1j * FreqForEpsrTanD))
    a = -(eps * tand) / np.imag(Lg)
    Einf = eps - a * np.real(Lg)
    return Einf + a * np.log((HighFreqForTanD + 1j * freq) / (LowFreqForTanD + 1j * freq))


def transform_surface(freq, gamma, axion_out=None, eps_i=1., eps_m=1., l=0, surfaceloss=0, USE_RAY=True):
    """src/Analytical1D.py:5-67; `axion_out` (complex array) is updated in place, as in the reference"""
    f = np.ascontiguousarray(np.atleast_1d(np.asarray(freq, dtype=np.float64)))
    n = len(f)
    g = np.ascontiguousarray(np.broadcast_to(np.asarray(gamma, dtype=np.complex128), (n,)))
    ax = None
    if axion_out is not None:
        if not USE_RAY:
            raise L.ArgumentError("Need to use ray approx.")
        if not (isinstance(axion_out, np.ndarray) and axion_out.dtype == np.complex128 and axion_out.shape == (n,)
                and axion_out.flags.c_contiguous):
            raise L.ArgumentError("axion_out must be a contiguous complex128 array with one entry per frequency")
        ax = axion_out
    out = np.zeros(n, dtype=np.complex128)
    ctx = L.context()
    m_real = int(not np.iscomplexobj(eps_m))  # a real eps_m gives a real wavelength in the reference's numpy
    L.check(L.lib().bf_transform_surface_1d(ctx.handle, L.cptr(f), n, L.cptr(g), L.cptr(ax), L.cptr(_c2(eps_i)),
                                             L.cptr(_c2(eps_m)), m_real, float(l), float(surfaceloss),
                                             int(bool(USE_RAY)), L.cptr(out)), ctx.handle)
    return out if np.ndim(freq) else out[0]


def _run(freqs, spacings, num_disk, disk_thickness, eps_disk, eps_gap, mirror, surfaceloss):
    f = np.ascontiguousarray(np.atleast_1d(np.asarray(freqs, dtype=np.float64)))
    sp = np.ascontiguousarray(np.atleast_2d(np.asarray(spacings, dtype=np.float64)))
    if sp.shape[1] < num_disk + 1:
        raise L.DimensionMismatch(f"{sp.shape[1]} spacings for {num_disk} disks (num_disk + 1 needed)")
    sp = np.ascontiguousarray(sp[:, :num_disk + 1])
    sl = None
    if surfaceloss is not None:
        sl = np.ascontiguousarray(np.asarray(surfaceloss, dtype=np.float64).reshape(-1))
        if len(sl) < 2 * num_disk + 1:
            raise L.DimensionMismatch(f"{len(sl)} surface losses for {2 * num_disk + 1} surfaces")
    gamma = np.zeros((len(sp), len(f)), dtype=np.complex128)
    ax = np.zeros_like(gamma)
    ctx = L.context()
    L.check(L.lib().bf_disk_system_1d(ctx.handle, L.cptr(f), len(f), len(sp), int(num_disk), L.cptr(sp),
                                       float(disk_thickness), L.cptr(_c2(eps_disk)), L.cptr(_c2(eps_gap)),
                                       int(bool(mirror)), L.cptr(sl), L.cptr(gamma), L.cptr(ax)), ctx.handle)
    return gamma, ax


def disk_system(freq, tand=0, num_disk=5, non_uniform_surfaceloss=None, spacings=None, mirror=True,
                disk_thickness=0.001, disk_epsilon=9, USE_RAY=True):
    """src/Analytical1D.py:82-131 -> (gamma, axion_out), one value per frequency"""
    if not USE_RAY:
        raise L.ArgumentError("Need to use ray approx.")
    if spacings is None:
        spacings = 8e-3 * np.ones(num_disk + 1)
    eps_1 = lossy_eps(freq, disk_epsilon, 0, SvenssonDjordjevic=False)
    eps_2_tr = lossy_eps(freq, 1.0, tand, SvenssonDjordjevic=False)
    g, a = _run(freq, spacings, num_disk, disk_thickness, eps_1, eps_2_tr, mirror, non_uniform_surfaceloss)
    return (g[0], a[0]) if np.ndim(freq) else (g[0, 0], a[0, 0])


def disk_system_batch(freq, spacings, tand=0, num_disk=None, non_uniform_surfaceloss=None, mirror=True,
                      disk_thickness=0.001, disk_epsilon=9):
    """disk_system for many spacing vectors at once: spacings (ncand, num_disk + 1)
    -> gamma, axion_out of shape (ncand, nfreq)"""
    sp = np.atleast_2d(np.asarray(spacings, dtype=np.float64))
    nd = sp.shape[1] - 1 if num_disk is None else num_disk
    return _run(freq, sp, nd, disk_thickness, lossy_eps(freq, disk_epsilon, 0), lossy_eps(freq, 1.0, tand), mirror,
                non_uniform_surfaceloss)


def disk_system_phase_depths(d_air, d_disk=np.pi / 2, tand=0, num_disk=5, non_uniform_surfaceloss=None, USE_RAY=True):
    """src/Analytical1D.py:133-183: gaps and disks given as phase depths at 20 GHz
    (disk permittivity 9.4), one result per entry of d_air"""
    if not USE_RAY:
        raise L.ArgumentError("Need to use ray approx.")
    freq = 20e9
    wavel = c / freq
    d_air = np.asarray(d_air, dtype=np.float64)
    l_air = d_air * wavel / (2 * np.pi)
    l_disk = d_disk * wavel / (2 * np.pi * np.sqrt(9.4))
    flat = np.atleast_1d(l_air).reshape(-1)
    sp = np.repeat(flat[:, None], num_disk + 1, axis=1)
    if non_uniform_surfaceloss is not None:
        sp[:, 0] = 8e-3
    g, a = _run([freq], sp, num_disk, float(l_disk), lossy_eps(freq, 9.4, 0), lossy_eps(freq, 1.0, tand), True,
                non_uniform_surfaceloss)
    return g[:, 0].reshape(d_air.shape), a[:, 0].reshape(d_air.shape)
