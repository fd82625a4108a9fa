"""Recursive Fourier propagation solvers (src/boost_fractor_recurprop.jl), on the GPU.

`f` may be a scalar (the reference signature) or a sequence of frequencies,
in which case all frequencies run as one batched device launch sequence and
the results are stacked along a leading axis.  `precision="mixed"` selects
the FP32 field/FFT path (phase arguments formed in FP64, Eout accumulated in
FP64); the default "fp64" is the parity path.
"""
import ctypes

import numpy as np

from . import _lib as L
from .propagators import precision_code, prop_kind, propagator


def _freqs(f):
    arr = np.atleast_1d(np.asarray(f, dtype=np.float64))
    return np.ascontiguousarray(arr.reshape(-1)), np.ndim(f) == 0


def dance_intro(bdry, coords, bfield=None, velocity_x=0, f=10e9, diskR=0.1):
    """src/boost_fractor_recurprop.jl:153-201 -> fields (length(distance), 2, ny, nx).

    bfield: (length(distance)+1, ny, nx) or None (the real aperture mask).  A
    nonzero velocity_x multiplies bfield by exp(-1im*Ma*(-velocity_x)*x) in place,
    as the reference does; with a real bfield (or the default mask) the
    reference raises InexactError, and so does this."""
    m = L.Marshalled()
    c = m.coords(coords)
    nd = len(bdry.distance)
    eps = m.hold(L.as_c128(bdry.eps[:nd]))
    r1 = complex(bdry.r[0])
    out = np.zeros((nd, 2, coords.ny, coords.nx), dtype=np.complex128)
    b = None
    if bfield is not None:
        arr = np.asarray(bfield)
        if arr.shape != (nd + 1, coords.ny, coords.nx):
            raise L.DimensionMismatch(f"bfield has shape {arr.shape}, expected {(nd + 1, coords.ny, coords.nx)}")
        if velocity_x != 0 and not np.iscomplexobj(arr):
            raise L.InexactError("velocity_x needs a complex bfield: the phase cannot be stored in a real array")
        b = m.hold(L.as_c128(arr).copy())
    elif velocity_x != 0:
        raise L.InexactError("velocity_x with the default (real) B field")
    ctx = L.context()
    L.check(L.lib().bf_dance_intro(ctx.handle, ctypes.byref(c), nd, L.cptr(eps), r1.real, r1.imag, float(diskR),
                                   L.cptr(b), float(velocity_x), float(f), L.cptr(out)), ctx.handle)
    if b is not None and velocity_x != 0 and isinstance(bfield, np.ndarray) and bfield.flags.writeable:
        bfield[...] = b  # the reference's `bfield .*= ...` mutates the caller's array
    return out


def dancer(amin, nmax, sbdry, coords, f=10.0e9, prop=propagator, emit=None, reflect=None, diskR=0.1,
           returnsum=True, immediatesum=True, precision="fp64", return_power=False):
    """src/boost_fractor_recurprop.jl:25-146.

    Returns Eout (ny, nx) when returnsum, else (S, ny, nx) with S = 1 or nmax+1
    (Julia's (Nx, Ny, S)).  With several frequencies a leading frequency axis is
    added.  return_power=True also returns sum |Eout|^2 per frequency."""
    kind = prop_kind(prop)
    prec = precision_code(precision)
    freqs, scalar = _freqs(f)
    nf = len(freqs)
    m = L.Marshalled()
    c = m.coords(coords)
    s = m.setup(sbdry, coords.ny, coords.nx)
    R = len(sbdry.distance) - 1
    K = (coords.ny, coords.nx)
    em = None
    if emit is not None and reflect is None:
        em = m.hold(L.as_c128(emit))
        if em.shape != (R, 2) + K:
            raise L.DimensionMismatch(f"emit has shape {em.shape}, expected {(R, 2) + K}")
    rf = None
    if reflect is not None:
        rf = m.hold(L.as_c128(np.broadcast_to(reflect, K)))
    if returnsum:
        out = np.zeros((nf,) + K, dtype=np.complex128)
    else:
        S = 1 if immediatesum else nmax + 1
        out = np.zeros((nf, S) + K, dtype=np.complex128)
    power = np.zeros(nf)
    bounces = np.zeros(nf, dtype=np.int32)
    ctx = L.context()
    L.check(L.lib().bf_dancer(ctx.handle, ctypes.byref(s), ctypes.byref(c), float(amin), int(nmax), L.cptr(freqs),
                              nf, kind, L.cptr(em), L.cptr(rf), float(diskR), int(bool(returnsum)),
                              int(bool(immediatesum)), prec, L.cptr(out), L.cptr(power), L.cptr(bounces)),
            ctx.handle)
    res = out[0] if scalar else out
    if return_power:
        return res, (power[0] if scalar else power)
    return res


def dancer_batch(amin, nmax, sbdrys, coords, f=10.0e9, prop=propagator, emit=None, reflect=None, diskR=0.1,
                 returnsum=True, immediatesum=True, precision="fp64", return_power=False):
    """dancer (src/boost_fractor_recurprop.jl:25-146) for a list of setups at the same frequencies, in one
    call (the reference sweeps geometries point by point, test/runtests.jl:101-114); the setups may differ
    in disk count.  emit / reflect: None or one entry per setup (each as dancer's, may be None).  Returns
    the list of dancer's results per setup (with return_power, a list of (Eout, power) pairs)."""
    kind = prop_kind(prop)
    prec = precision_code(precision)
    freqs, scalar = _freqs(f)
    nf = len(freqs)
    nconf = len(sbdrys)
    m = L.Marshalled()
    c = m.coords(coords)
    setups = m.setups(sbdrys, coords.ny, coords.nx)
    K = (coords.ny, coords.nx)
    emits = [None] * nconf if emit is None else list(emit)
    reflects = [None] * nconf if reflect is None else list(reflect)
    if len(emits) != nconf or len(reflects) != nconf:
        raise L.DimensionMismatch("emit and reflect need one entry per setup")
    ep = (ctypes.c_void_p * max(nconf, 1))()
    rp = (ctypes.c_void_p * max(nconf, 1))()
    shapes, offs, tot = [], [], 0
    for i, sb in enumerate(sbdrys):
        R = len(sb.distance) - 1
        if emits[i] is not None and reflects[i] is None:
            em = m.hold(L.as_c128(emits[i]))
            if em.shape != (R, 2) + K:
                raise L.DimensionMismatch(f"emit[{i}] has shape {em.shape}, expected {(R, 2) + K}")
            ep[i] = em.ctypes.data
        if reflects[i] is not None:
            rp[i] = m.hold(L.as_c128(np.broadcast_to(reflects[i], K))).ctypes.data
        shp = (nf,) + K if returnsum else (nf, 1 if immediatesum else nmax + 1) + K
        shapes.append(shp)
        offs.append(tot)
        tot += int(np.prod(shp))
    out = np.zeros(max(tot, 1), dtype=np.complex128)
    off = np.asarray(offs if offs else [0], dtype=np.int64)
    power = np.zeros((max(nconf, 1), nf))
    bounces = np.zeros((max(nconf, 1), nf), dtype=np.int32)
    ctx = L.context()
    L.check(L.lib().bf_dancer_batch(ctx.handle, nconf, setups, ctypes.byref(c), float(amin), int(nmax),
                                    L.cptr(freqs), nf, kind, ep if emit is not None else None,
                                    rp if reflect is not None else None, float(diskR), int(bool(returnsum)),
                                    int(bool(immediatesum)), prec, L.cptr(off), L.cptr(out), L.cptr(power),
                                    L.cptr(bounces)), ctx.handle)
    res = []
    for i in range(nconf):
        e = out[offs[i]:offs[i] + int(np.prod(shapes[i]))].reshape(shapes[i])
        e = e[0] if scalar else e
        res.append((e, power[i][0] if scalar else power[i]) if return_power else e)
    return res


def cheerleader(amin, nmax, bdry, coords, f=10.0e9, prop=propagator, emit=None, reflect=None, diskR=0.1,
                returnboth=False, precision="fp64"):
    """src/boost_fractor_recurprop.jl:226-363.  Returns the field leaving on the
    right, or (left, right) when returnboth (the reference's order)."""
    kind = prop_kind(prop)
    prec = precision_code(precision)
    freqs, scalar = _freqs(f)
    nf = len(freqs)
    m = L.Marshalled()
    c = m.coords(coords)
    s = m.setup(bdry, coords.ny, coords.nx)
    Nd = len(bdry.distance)
    K = (coords.ny, coords.nx)
    em = None
    if emit is not None and reflect is None:
        em = m.hold(L.as_c128(emit))
        if em.shape != (Nd, 2) + K:
            raise L.DimensionMismatch(f"emit has shape {em.shape}, expected {(Nd, 2) + K}")
    rf = m.hold(L.as_c128(np.broadcast_to(reflect, K))) if reflect is not None else None
    right = np.zeros((nf,) + K, dtype=np.complex128)
    left = np.zeros((nf,) + K, dtype=np.complex128) if returnboth else None
    ctx = L.context()
    L.check(L.lib().bf_cheerleader(ctx.handle, ctypes.byref(s), ctypes.byref(c), float(amin), int(nmax),
                                   L.cptr(freqs), nf, kind, L.cptr(em), L.cptr(rf), float(diskR), prec,
                                   L.cptr(right), L.cptr(left)), ctx.handle)
    if scalar:
        right = right[0]
        left = left[0] if left is not None else None
    return (left, right) if returnboth else right


class DancerPlan:
    """Device-resident frequency sweep of `dancer` (default emission, summed output).

    Build once, `run()` many times; inputs stay in HBM between runs."""

    def __init__(self, amin, nmax, sbdry, coords, freqs, prop=propagator, diskR=0.1, precision="fp64",
                 freq_batch=0, ctx=None):
        self.ctx = ctx or L.context()
        self.freqs = np.ascontiguousarray(np.atleast_1d(freqs), dtype=np.float64)
        self.K = (coords.ny, coords.nx)
        m = L.Marshalled()
        c = m.coords(coords)
        s = m.setup(sbdry, coords.ny, coords.nx)
        h = ctypes.c_void_p()
        L.check(L.lib().bf_dancer_plan_create(self.ctx.handle, ctypes.byref(s), ctypes.byref(c), float(amin),
                                              int(nmax), L.cptr(self.freqs), len(self.freqs), prop_kind(prop),
                                              float(diskR), precision_code(precision), int(freq_batch),
                                              ctypes.byref(h)), self.ctx.handle)
        self.handle = h

    def run(self):
        L.check(L.lib().bf_dancer_plan_run(self.handle), self.ctx.handle)

    def fetch(self, fields=True):
        power = np.zeros(len(self.freqs))
        E = np.zeros((len(self.freqs),) + self.K, dtype=np.complex128) if fields else None
        L.check(L.lib().bf_dancer_plan_fetch(self.handle, L.cptr(E), L.cptr(power)), self.ctx.handle)
        return E, power

    def timing(self, enable=True):
        L.check(L.lib().bf_dancer_plan_timing(self.handle, int(bool(enable))), self.ctx.handle)

    def kernel_times(self):
        ms = np.zeros(2)
        n = np.zeros(2, dtype=np.int64)
        L.check(L.lib().bf_dancer_plan_kernel_times(self.handle, L.cptr(ms), L.cptr(n)), self.ctx.handle)
        return {"row_pass": (ms[0], int(n[0])), "col_pass": (ms[1], int(n[1]))}

    def info(self):
        """layout chosen by the library: frequencies per batch and the row band"""
        fb, lo, hi = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        L.check(L.lib().bf_dancer_plan_info(self.handle, ctypes.byref(fb), ctypes.byref(lo), ctypes.byref(hi)),
                self.ctx.handle)
        return {"freq_batch": fb.value, "band": (lo.value, hi.value)}

    @property
    def freq_batch(self):
        return self.info()["freq_batch"]

    def close(self):
        if self.handle:
            L.lib().bf_dancer_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
