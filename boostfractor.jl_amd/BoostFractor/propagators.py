"""The reference's propagator plug-ins (src/boost_fractor.jl:113-227), run on the GPU.

Each function has the reference signature
    prop(E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords) -> E
and is also the token the solvers accept as `prop=`; the solvers map it to
the engine's fused device implementation.  A user-written Python propagator
has no device equivalent and is rejected with ArgumentError.

Unlike the reference (which overwrites E0 with its FFT as a side effect,
src/boost_fractor.jl:119-120), the input array is left unchanged.
"""
import ctypes

import numpy as np

from . import _lib as L


def _eps_parts(eps):
    z = complex(eps)  # keeps a signed-zero imaginary part
    return z.real, z.imag


def _run(kind, E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords):
    E = np.array(E0, dtype=np.complex128, order="C", copy=True)
    if E.shape != (coords.ny, coords.nx):
        raise L.DimensionMismatch(f"field shape {E.shape} does not match the grid {(coords.ny, coords.nx)}")
    m = L.Marshalled()
    c = m.coords(coords)
    surf = None
    if surface is not None:
        s = np.asarray(surface)
        if s.size > 1 and np.any(s != 0):
            surf = m.hold(L.as_c128(np.broadcast_to(s, E.shape)))
        elif s.size == 1 and complex(s.reshape(-1)[0]) != 0:
            surf = m.hold(L.as_c128(np.broadcast_to(s, E.shape)))
    er, ei = _eps_parts(eps)
    ctx = L.context()
    L.check(L.lib().bf_propagate(ctx.handle, kind, ctypes.byref(c), L.cptr(E), float(dz), float(diskR), er, ei,
                                 float(tilt_x), float(tilt_y), L.cptr(surf), float(lambda_)), ctx.handle)
    return E


def propagator(E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords):
    """src/boost_fractor.jl:137-149: mask -> FFT -> exp(-i kz dz) -> IFFT -> tilt -> surface."""
    return _run(L.PROP_PROPAGATOR, E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords)


def propagatorNoTilts(E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords):
    """src/boost_fractor.jl:157-179 (eps used as given: pass re-0.0im for real eps)."""
    return _run(L.PROP_NOTILTS, E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords)


def propagatorMomentumSpace(E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords):
    """src/boost_fractor.jl:187-210: E0 and the result are fftshifted k-space fields."""
    return _run(L.PROP_MOMENTUM, E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords)


def propagator1D(E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords):
    """src/boost_fractor.jl:218-227: phase exp(-i k0 dz) only."""
    return _run(L.PROP_1D, E0, dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords)


_KINDS = {propagator: L.PROP_PROPAGATOR, propagatorNoTilts: L.PROP_NOTILTS, propagator1D: L.PROP_1D,
          propagatorMomentumSpace: L.PROP_MOMENTUM}


def prop_kind(prop):
    try:
        return _KINDS[prop]
    except (KeyError, TypeError):
        raise L.ArgumentError(
            f"prop={prop!r}: only propagator, propagatorNoTilts, propagatorMomentumSpace and propagator1D "
            "run on the device engine") from None


def precision_code(precision):
    if precision in ("fp64", "float64", L.PREC_FP64, None):
        return L.PREC_FP64
    if precision in ("mixed", L.PREC_MIXED):
        return L.PREC_MIXED
    if precision in ("fp32", L.PREC_FP32):
        return L.PREC_FP32
    raise L.ArgumentError(f"unknown precision {precision!r} (use 'fp64', 'mixed' or 'fp32')")
