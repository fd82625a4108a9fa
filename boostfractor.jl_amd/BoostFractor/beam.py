"""Antenna-side helpers (src/beam_gaussian.jl, src/beam_coupling.jl), on the device."""
import ctypes

import numpy as np

from . import _lib as L


def gauss_profile(coords, z=0.000000001, omega0=0.1, f=10e9):
    """src/beam_gaussian.jl:24-37: normalised Gaussian beam profile (ny, nx) at
    distance z from the waist (radius omega0) at frequency f."""
    m = L.Marshalled()
    c = m.coords(coords)
    out = np.zeros((coords.ny, coords.nx), dtype=np.complex128)
    ctx = L.context()
    L.check(L.lib().bf_gauss_profile(ctx.handle, ctypes.byref(c), float(z), float(omega0), float(f), L.cptr(out)),
            ctx.handle)
    return out


def unnormalized_match(Eout, Eantenna, dx=0.01, dy=0.01):
    """src/beam_coupling.jl:13-15: sum(Eout .* conj(Eantenna)) * dx * dy.

    Eout may carry leading axes (e.g. one field per frequency): the match is
    taken over the trailing axes of Eantenna's shape for each of them."""
    a = np.asarray(Eantenna)
    e = np.asarray(Eout)
    if e.shape[e.ndim - a.ndim:] != a.shape:
        raise L.DimensionMismatch(f"Eout {e.shape} does not end with the antenna's shape {a.shape}")
    lead = e.shape[:e.ndim - a.ndim]
    nb = int(np.prod(lead)) if lead else 1
    m = L.Marshalled()
    ec = m.hold(L.as_c128(e.reshape(nb, -1)))
    ac = m.hold(L.as_c128(a.reshape(-1)))
    out = np.zeros(nb, dtype=np.complex128)
    ctx = L.context()
    L.check(L.lib().bf_unnormalized_match(ctx.handle, nb, ctypes.c_int64(ac.size), L.cptr(ec), L.cptr(ac),
                                          float(dx), float(dy), L.cptr(out)), ctx.handle)
    return out.reshape(lead) if lead else complex(out[0])
