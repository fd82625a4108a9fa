"""ctypes binding of libboostfractor_hip.so (include/boostfractor_hip.h).

There is no CPU fallback: if the HIP library is missing or no device is
visible, every engine call raises.
"""
import contextlib
import ctypes
import os
import threading

import numpy as np

from .setup import DimensionMismatch

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("BF_LIBRARY", os.path.join(PKG_DIR, "lib", "libboostfractor_hip.so"))

BF_OK, BF_ERR_DIM, BF_ERR_ARG, BF_ERR_UNSUPPORTED, BF_ERR_HIP, BF_ERR_NOMEM = 0, 1, 2, 3, 4, 5
PROP_PROPAGATOR, PROP_NOTILTS, PROP_1D, PROP_MOMENTUM = 0, 1, 2, 3
PREC_FP64, PREC_MIXED, PREC_FP32 = 0, 1, 2


class ArgumentError(ValueError):
    """Julia's ArgumentError / error(...)"""


class UnsupportedError(NotImplementedError):
    pass


class InexactError(ValueError):
    """Julia's InexactError (a complex value stored into a Float64 array)"""


class BoundsError(IndexError):
    """Julia's BoundsError"""


class HipError(RuntimeError):
    pass


_I32 = ctypes.c_int32
_D = ctypes.c_double
_P = ctypes.c_void_p


class Coords(ctypes.Structure):
    _fields_ = [("nx", _I32), ("ny", _I32), ("X", _P), ("Y", _P), ("kX", _P), ("kY", _P)]


class Setup(ctypes.Structure):
    _fields_ = [("nreg", _I32), ("distance", _P), ("r", _P), ("eps", _P), ("tilt_x", _P), ("tilt_y", _P),
                ("surfaces", _P)]


class ModesS(ctypes.Structure):
    _fields_ = [("n", _I32), ("ncomp", _I32), ("patterns", _P), ("mode_kt", _P)]


# name -> (restype, argtypes)
_SIGS = {
    "bf_version": (ctypes.c_int, []),
    "bf_last_error": (ctypes.c_char_p, [_P]),
    "bf_device_count": (ctypes.c_int, [ctypes.POINTER(_I32)]),
    "bf_ctx_create": (ctypes.c_int, [_I32, ctypes.POINTER(_P)]),
    "bf_ctx_destroy": (ctypes.c_int, [_P]),
    "bf_ctx_stream": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "bf_ctx_synchronize": (ctypes.c_int, [_P]),
    "bf_propagate": (ctypes.c_int, [_P, _I32, ctypes.POINTER(Coords), _P, _D, _D, _D, _D, _D, _D, _P, _D]),
    "bf_dance_intro": (ctypes.c_int, [_P, ctypes.POINTER(Coords), _I32, _P, _D, _D, _D, _P, _D, _D, _P]),
    "bf_gauss_profile": (ctypes.c_int, [_P, ctypes.POINTER(Coords), _D, _D, _D, _P]),
    "bf_unnormalized_match": (ctypes.c_int, [_P, _I32, ctypes.c_int64, _P, _P, _D, _D, _P]),
    "bf_dancer": (ctypes.c_int, [_P, ctypes.POINTER(Setup), ctypes.POINTER(Coords), _D, _I32, _P, _I32, _I32, _P, _P,
                                 _D, _I32, _I32, _I32, _P, _P, _P]),
    "bf_dancer_batch": (ctypes.c_int, [_P, _I32, ctypes.POINTER(Setup), ctypes.POINTER(Coords), _D, _I32, _P, _I32,
                                       _I32, _P, _P, _D, _I32, _I32, _I32, _P, _P, _P, _P]),
    "bf_cheerleader": (ctypes.c_int, [_P, ctypes.POINTER(Setup), ctypes.POINTER(Coords), _D, _I32, _P, _I32, _I32,
                                      _P, _P, _D, _I32, _P, _P]),
    "bf_dancer_plan_create": (ctypes.c_int, [_P, ctypes.POINTER(Setup), ctypes.POINTER(Coords), _D, _I32, _P, _I32,
                                             _I32, _D, _I32, _I32, ctypes.POINTER(_P)]),
    "bf_dancer_plan_run": (ctypes.c_int, [_P]),
    "bf_dancer_plan_fetch": (ctypes.c_int, [_P, _P, _P]),
    "bf_dancer_plan_timing": (ctypes.c_int, [_P, _I32]),
    "bf_dancer_plan_kernel_times": (ctypes.c_int, [_P, _P, _P]),
    "bf_dancer_plan_info": (ctypes.c_int, [_P, _P, _P, _P]),
    "bf_ctx_timing": (ctypes.c_int, [_P, ctypes.c_int32]),
    "bf_ctx_kernel_time": (ctypes.c_int, [_P, ctypes.c_char_p, _P, _P, _P]),
    "bf_ctx_set_option": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int64]),
    "bf_ctx_get_option": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]),
    "bf_ctx_set_tuning_file": (ctypes.c_int, [_P, ctypes.c_char_p]),
    "bf_ctx_tuning_report": (ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_double)]),
    "bf_dancer_plan_destroy": (ctypes.c_int, [_P]),
    "bf_stream_copy_probe": (ctypes.c_int, [_P, ctypes.c_int64, _I32, ctypes.POINTER(_D)]),
    "bf_seed_modes": (ctypes.c_int, [_P, ctypes.POINTER(Coords), _I32, _I32, _D, _P, _P]),
    "bf_field2modes": (ctypes.c_int, [_P, ctypes.POINTER(Coords), ctypes.POINTER(ModesS), _P, _D, _D, _D, _P]),
    "bf_modes2field": (ctypes.c_int, [_P, ctypes.POINTER(Coords), ctypes.POINTER(ModesS), _P, _P]),
    "bf_propagation_matrices": (ctypes.c_int, [_P, ctypes.POINTER(Setup), ctypes.POINTER(Coords),
                                               ctypes.POINTER(ModesS), _P, _I32, _I32, _D, _I32, _P]),
    "bf_transformer": (ctypes.c_int, [_P, ctypes.POINTER(Setup), ctypes.POINTER(Coords), ctypes.POINTER(ModesS), _P,
                                      _I32, _I32, _P, _D, _P, _P, _I32, _P, _P]),
    "bf_transformer_plan_create": (ctypes.c_int, [_P, ctypes.POINTER(Setup), ctypes.POINTER(Coords),
                                                  ctypes.POINTER(ModesS), _P, _I32, _I32, _D, _P, _P,
                                                  ctypes.POINTER(_P)]),
    "bf_transformer_plan_run": (ctypes.c_int, [_P]),
    "bf_transformer_plan_fetch": (ctypes.c_int, [_P, _P, _P]),
    "bf_transformer_plan_destroy": (ctypes.c_int, [_P]),
    "bf_disk_system_1d": (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _P, _D, _P, _P, _I32, _P, _P, _P]),
    "bf_transform_surface_1d": (ctypes.c_int, [_P, _P, _I32, _P, _P, _P, _P, _I32, _D, _D, _I32, _P]),
    "bf_transformer_batch": (ctypes.c_int, [_P, _I32, ctypes.POINTER(Setup), ctypes.POINTER(Coords), _I32,
                                            ctypes.POINTER(ModesS), ctypes.POINTER(ctypes.c_void_p), _P, _P, _I32,
                                            _I32, _D, _I32, _P, _P]),
    "bf_transformer_batch_plan_create": (ctypes.c_int, [_P, ctypes.POINTER(Coords), _I32, ctypes.POINTER(ModesS),
                                                        ctypes.POINTER(ctypes.c_void_p), _P, _I32, _I32, _D, _I32,
                                                        ctypes.POINTER(_P)]),
    "bf_transformer_batch_plan_create_grids": (ctypes.c_int, [_P, _I32, ctypes.POINTER(Coords),
                                                              ctypes.POINTER(ModesS), ctypes.POINTER(ctypes.c_void_p),
                                                              _P, _I32, _I32, _D, _I32, ctypes.POINTER(_P)]),
    "bf_transformer_batch_plan_run": (ctypes.c_int, [_P, _I32, ctypes.POINTER(Setup), _P, _P]),
    "bf_transformer_batch_plan_fetch": (ctypes.c_int, [_P, _P]),
    "bf_transformer_batch_plan_destroy": (ctypes.c_int, [_P]),
    "bf_propagation_matrices_grid": (ctypes.c_int, [_P, ctypes.POINTER(Setup), ctypes.POINTER(Coords),
                                                    ctypes.POINTER(ModesS), _P, _I32, _P, _I32, _P, _I32, _P, _I32,
                                                    _I32, _D, _P]),
    "bf_spacing_itp_create": (ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _D, _D, ctypes.POINTER(_P)]),
    "bf_spacing_itp_bounds": (ctypes.c_int, [_P, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "bf_spacing_itp_matrices": (ctypes.c_int, [_P, _I32, _P, _P]),
    "bf_spacing_itp_transformer": (ctypes.c_int, [_P, _I32, _P, _P, _P, _P, _P, _P]),
    "bf_spacing_itp_destroy": (ctypes.c_int, [_P]),
    "bf_transformer_trace_back": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(Setup), ctypes.POINTER(Coords),
                                                 ctypes.POINTER(ModesS), _D, _I32, _P, _D, _I32, _P, _P]),
}

_lib = None
_lock = threading.Lock()
_tls = threading.local()


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise ImportError(
                        f"libboostfractor_hip.so not found at {LIB_PATH}: build it with "
                        "`make -C boostfractor.jl_amd` (or __graft_entry__.build()); there is no CPU fallback")
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in _SIGS.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def check(rc, ctx=None):
    if rc == BF_OK:
        return
    msg = lib().bf_last_error(ctx)
    msg = msg.decode() if msg else ""
    if rc == BF_ERR_DIM:
        raise DimensionMismatch(msg)
    if rc == BF_ERR_ARG:
        raise ArgumentError(msg)
    if rc == BF_ERR_UNSUPPORTED:
        raise UnsupportedError(msg)
    if rc == BF_ERR_NOMEM:
        raise MemoryError(msg)
    raise HipError(msg)


class Context:
    """One engine context (device + HIP stream).  One per thread by default."""

    def __init__(self, device=None):
        if device is None:
            device = int(os.environ.get("BF_DEVICE", "0"))
        h = _P()
        check(lib().bf_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().bf_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stream(self):
        s = _P()
        check(lib().bf_ctx_stream(self.handle, ctypes.byref(s)), self.handle)
        return s.value

    def synchronize(self):
        check(lib().bf_ctx_synchronize(self.handle), self.handle)

    def timing(self, enable=True):
        """start (clearing the counters) or stop per-kernel-class timing"""
        check(lib().bf_ctx_timing(self.handle, int(bool(enable))), self.handle)

    def kernel_time(self, kernel):
        """(milliseconds, launches, algorithmic flops) accumulated for a kernel class"""
        ms, fl = _D(), _D()
        n = ctypes.c_int64()
        check(lib().bf_ctx_kernel_time(self.handle, kernel.encode(), ctypes.byref(ms), ctypes.byref(n),
                                       ctypes.byref(fl)), self.handle)
        return ms.value, n.value, fl.value

    def set_option(self, name, value):
        """set one engine option of this context (bf_ctx_set_option)"""
        check(lib().bf_ctx_set_option(self.handle, name.encode(), int(value)), self.handle)

    def get_option(self, name):
        v = ctypes.c_int64()
        check(lib().bf_ctx_get_option(self.handle, name.encode(), ctypes.byref(v)), self.handle)
        return v.value

    def set_tuning_file(self, path):
        """load tuned launch shapes from `path` and append new ones to it"""
        check(lib().bf_ctx_set_tuning_file(self.handle, None if path is None else os.fsencode(path)), self.handle)

    def tuning_report(self):
        """({shape key: [values]}, milliseconds spent tuning) of this context"""
        buf = ctypes.create_string_buffer(1 << 16)
        ms = ctypes.c_double()
        check(lib().bf_ctx_tuning_report(self.handle, buf, len(buf), ctypes.byref(ms)), self.handle)
        out = {}
        for line in buf.value.decode().splitlines():
            k, _, v = line.partition("\t")
            out[k] = [int(x) for x in v.split()]
        return out, ms.value

    @contextlib.contextmanager
    def options(self, **kw):
        """engine options set for the body of a `with` block, restored afterwards"""
        old = {k: self.get_option(k) for k in kw}
        try:
            for k, v in kw.items():
                self.set_option(k, v)
            yield self
        finally:
            for k, v in old.items():
                self.set_option(k, v)

    def stream_copy_probe(self, nbytes, reps=20):
        """mean ms per device-to-device copy of nbytes with the engine's copy kernel"""
        ms = _D()
        check(lib().bf_stream_copy_probe(self.handle, int(nbytes), int(reps), ctypes.byref(ms)), self.handle)
        return ms.value


def context():
    """The calling thread's context (created on first use)."""
    c = getattr(_tls, "ctx", None)
    if c is None:
        c = Context()
        _tls.ctx = c
    return c


def device_count():
    n = _I32()
    rc = lib().bf_device_count(ctypes.byref(n))
    return n.value if rc == BF_OK else 0


# ---- marshalling helpers ---------------------------------------------------
def cptr(a):
    return None if a is None else a.ctypes.data_as(_P)


def as_c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Marshalled:
    """Holds the numpy buffers a ctypes struct points into."""

    def __init__(self):
        self.keep = []

    def hold(self, a):
        self.keep.append(a)
        return a

    def coords(self, c):
        X, Y, kX, kY = (self.hold(as_f64(v)) for v in (c.X, c.Y, c.kX, c.kY))
        return Coords(len(X), len(Y), cptr(X), cptr(Y), cptr(kX), cptr(kY))

    def setup(self, b, ny, nx):
        dist = self.hold(as_f64(b.distance))
        eps = self.hold(as_c128(b.eps))
        r = self.hold(as_c128(b.r))
        tx = self.hold(as_f64(np.broadcast_to(b.relative_tilt_x, dist.shape)))
        ty = self.hold(as_f64(np.broadcast_to(b.relative_tilt_y, dist.shape)))
        surf = None
        s = np.asarray(b.relative_surfaces)
        # any nonzero entry (region by region: stops at the first rough region, no temporaries)
        if s.size and any(np.asarray(s[r]).view(np.float64 if s.dtype == np.complex128 else s.dtype).any()
                          for r in range(s.shape[0])):
            if s.shape != (len(dist), ny, nx):
                raise DimensionMismatch(
                    f"relative_surfaces has shape {s.shape}, expected {(len(dist), ny, nx)}")
            surf = self.hold(as_c128(s))
        return Setup(len(dist), cptr(dist), cptr(r), cptr(eps), cptr(tx), cptr(ty), cptr(surf))

    def setups(self, bdrys, ny, nx, nan_to_1e30=False):
        """ctypes array of Setup for many geometries: each field of all of them in one buffer, the structs
        filled at once in a numpy array laid out like Setup (the same structs as setup() per geometry).
        ny, nx: the grid of every geometry, or sequences with each geometry's grid (surface map shape).
        nan_to_1e30: NaN permittivities become 1e30, in the caller's setups too (transformer's mutation,
        src/boost_fractor_transformer.jl:289), found with one scan of the joined buffer"""
        n = len(bdrys)
        arr = (Setup * max(n, 1))()
        if n == 0:
            return arr
        # one pass over the geometries gathers every field (the joins and the structs are numpy-wide)
        dists, rs, epss, txs, tys, rough = [], [], [], [], [], []
        f64, c128 = np.float64, np.complex128
        for i, b in enumerate(bdrys):
            d = np.asarray(b.distance, dtype=f64).reshape(-1)
            l = len(d)
            dists.append(d)
            rs.append(np.asarray(b.r, dtype=c128).reshape(-1))
            epss.append(np.asarray(b.eps, dtype=c128).reshape(-1))
            for v, out in ((b.relative_tilt_x, txs), (b.relative_tilt_y, tys)):
                # a tilt array of the geometry's region count, or a value broadcast to it
                a = np.asarray(v, dtype=f64)
                out.append(a if a.shape == (l,) else np.broadcast_to(a, (l,)))
            srf = b.relative_surfaces
            if getattr(srf, "size", 1):
                rough.append(i)
        lens = np.fromiter((len(d) for d in dists), dtype=np.int64, count=n)
        rlens = np.fromiter((len(x) for x in rs), dtype=np.int64, count=n)
        dist = self.hold(np.concatenate(dists))
        eps = self.hold(np.concatenate(epss))
        if len(eps) != len(dist):
            raise DimensionMismatch("eps and distance differ in length")
        if nan_to_1e30:
            bad = np.flatnonzero(np.isnan(eps))
            if len(bad):
                eps[bad] = 1e30
                ends = np.cumsum(lens)
                owner = np.searchsorted(ends, bad, side="right")
                for q, c in zip(bad.tolist(), owner.tolist()):
                    bdrys[c].eps[q - (ends[c] - lens[c])] = 1e30
        r = self.hold(np.concatenate(rs))
        tx = self.hold(np.concatenate(txs))
        ty = self.hold(np.concatenate(tys))
        off = np.concatenate(([0], np.cumsum(lens)[:-1]))
        roff = np.concatenate(([0], np.cumsum(rlens)[:-1]))
        f = Setup
        rec = np.frombuffer(arr, dtype=np.dtype({
            "names": ["nreg", "distance", "r", "eps", "tilt_x", "tilt_y", "surfaces"],
            "formats": ["<i4", "<u8", "<u8", "<u8", "<u8", "<u8", "<u8"],
            "offsets": [f.nreg.offset, f.distance.offset, f.r.offset, f.eps.offset, f.tilt_x.offset,
                        f.tilt_y.offset, f.surfaces.offset],
            "itemsize": ctypes.sizeof(Setup)}), count=n)
        rec["nreg"] = lens
        rec["distance"] = dist.ctypes.data + 8 * off
        rec["r"] = (r.ctypes.data + 16 * roff) if len(r) else 0
        rec["eps"] = eps.ctypes.data + 16 * off
        rec["tilt_x"] = tx.ctypes.data + 8 * off
        rec["tilt_y"] = ty.ctypes.data + 8 * off
        for i in rough:  # geometries with a surface stack: passed if any map is nonzero
            s = np.asarray(bdrys[i].relative_surfaces)
            if s.size and any(np.asarray(s[q]).view(np.float64 if s.dtype == np.complex128 else s.dtype).any()
                              for q in range(s.shape[0])):
                want = (int(lens[i]), int(ny[i] if np.ndim(ny) else ny), int(nx[i] if np.ndim(nx) else nx))
                if s.shape != want:
                    raise DimensionMismatch(f"relative_surfaces has shape {s.shape}, expected {want}")
                rec["surfaces"][i] = self.hold(as_c128(s)).ctypes.data
        return arr

    def modes(self, m):
        pat = self.hold(as_c128(m.patterns_flat()))
        kt = self.hold(as_c128(np.asarray(m.mode_kt).reshape(-1)))
        n = m.n
        return ModesS(n, m.ncomp, cptr(pat), cptr(kt))
