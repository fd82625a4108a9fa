"""ctypes wrapper of the FP64 CPU oracle (oracle/bf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker.  The product (boostfractor.jl_amd/)
never imports this module.

Array conventions: complex128 numpy arrays; a field is shaped (ny, nx) so that
the C-order flat index equals Julia's column-major linear index of E[x, y].
Stacked objects keep Julia's leading axes first: dancer fields (R, 2, ny, nx),
surfaces (nreg, ny, nx), mode patterns (n, ny, nx) with n = (m-1)(2L+1)+l+L.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libbf_oracle.so")

PROP_PROPAGATOR, PROP_NOTILTS, PROP_1D, PROP_MOMENTUM = 0, 1, 2, 3

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        _lib = ctypes.CDLL(LIBPATH)
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.complex128)


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


_D = ctypes.c_double
_I = ctypes.c_int


def wavelength(f):
    lib().bfo_wavelength.restype = _D
    return lib().bfo_wavelength(_D(f))


def seed_reflectivities(eps):
    """Returns (eps with real entries cast to re-0.0im, r) as SeedSetupBoundaries does."""
    eps = _c(eps).copy()
    r = np.zeros(len(eps) - 1, dtype=np.complex128)
    lib().bfo_seed_reflectivities(_I(len(eps)), _p(eps), _p(r))
    return eps, r


class Coords:
    def __init__(self, X, Y, kX, kY):
        self.X, self.Y, self.kX, self.kY = _d(X), _d(Y), _d(kX), _d(kY)
        self.nx, self.ny = len(self.X), len(self.Y)

    def args(self):
        return (_I(self.nx), _I(self.ny), _p(self.X), _p(self.Y), _p(self.kX), _p(self.kY))


class Setup:
    def __init__(self, distance, eps, r=None, tilt_x=None, tilt_y=None, surfaces=None):
        self.distance = _d(distance)
        n = len(self.distance)
        self.eps = _c(eps)
        self.r = _c(r) if r is not None else None
        self.tilt_x = _d(tilt_x if tilt_x is not None else np.zeros(n))
        self.tilt_y = _d(tilt_y if tilt_y is not None else np.zeros(n))
        self.surfaces = _c(surfaces)
        self.nreg = n

    def args(self, with_r=True):
        a = [_I(self.nreg), _p(self.distance)]
        if with_r:
            a.append(_p(self.r))
        a += [_p(self.eps), _p(self.tilt_x), _p(self.tilt_y), _p(self.surfaces)]
        return a


def propagate(kind, E, coords, dz, diskR, eps, tilt_x=0.0, tilt_y=0.0, surface=None, lam=None):
    E = _c(E).copy()
    eps = complex(eps)
    lib().bfo_propagate(_I(kind), _p(E), _I(coords.nx), _I(coords.ny), _p(coords.X), _p(coords.Y),
                        _p(coords.kX), _p(coords.kY), _D(dz), _D(diskR), _D(eps.real), _D(eps.imag),
                        _D(tilt_x), _D(tilt_y), _p(_c(surface)), _D(lam))
    return E


def dance_intro(eps, r1, coords, diskR):
    eps = _c(eps)
    out = np.zeros((len(eps), 2, coords.ny, coords.nx), dtype=np.complex128)
    lib().bfo_dance_intro(_I(len(eps)), _p(eps), _D(complex(r1).real), _D(complex(r1).imag), _I(coords.nx),
                          _I(coords.ny), _p(coords.X), _p(coords.Y), _D(diskR), _p(out))
    return out


def dance_intro_ex(eps, r1, coords, diskR, bfield=None, velocity_x=0.0, f=10e9):
    """dance_intro with bfield (nd+1, ny, nx) and velocity; returns (fields, bfield after the
    reference's in-place multiplication), or raises ValueError where the reference throws
    InexactError (velocity with the default Float64 mask)."""
    eps = _c(eps)
    nd = len(eps)
    out = np.zeros((nd, 2, coords.ny, coords.nx), dtype=np.complex128)
    bout = np.zeros((nd + 1, coords.ny, coords.nx), dtype=np.complex128) if bfield is not None else None
    rc = lib().bfo_dance_intro_ex(_I(nd), _p(eps), _D(complex(r1).real), _D(complex(r1).imag), _I(coords.nx),
                                  _I(coords.ny), _p(coords.X), _p(coords.Y), _D(diskR), _p(_c(bfield)),
                                  _D(velocity_x), _D(f), _p(bout), _p(out))
    if rc != 0:
        raise ValueError("InexactError: velocity_x with the default Float64 B field")
    return out, bout


def gauss_profile(coords, z=1e-9, omega0=0.1, f=10e9):
    out = np.zeros((coords.ny, coords.nx), dtype=np.complex128)
    lib().bfo_gauss_profile(_I(coords.nx), _I(coords.ny), _p(coords.X), _p(coords.Y), _D(z), _D(omega0), _D(f),
                            _p(out))
    return out


def unnormalized_match(Eout, Eant, dx=0.01, dy=0.01):
    a, b = _c(Eout), _c(Eant)
    out = np.zeros(2)
    lib().bfo_unnormalized_match(ctypes.c_int64(a.size), _p(a), _p(b), _D(dx), _D(dy), _p(out))
    return complex(out[0], out[1])


def dancer(setup, coords, amin, nmax, f, kind=PROP_PROPAGATOR, emit=None, reflect=None, diskR=0.1,
           returnsum=True, immediatesum=True):
    K = (coords.ny, coords.nx)
    if returnsum:
        out = np.zeros(K, dtype=np.complex128)
    else:
        out = np.zeros(((1 if immediatesum else nmax + 1),) + K, dtype=np.complex128)
    emit = _c(emit)
    reflect = _c(reflect)
    lib().bfo_dancer(*setup.args(), *coords.args(), _D(amin), _I(nmax), _D(f), _I(kind), _p(emit), _p(reflect),
                     _D(diskR), _I(int(returnsum)), _I(int(immediatesum)), _p(out))
    return out


def dancer_sweep(setup, coords, amin, nmax, freqs, kind=PROP_PROPAGATOR, diskR=0.1, nthreads=1):
    freqs = _d(freqs)
    out = np.zeros((len(freqs), coords.ny, coords.nx), dtype=np.complex128)
    lib().bfo_dancer_sweep(*setup.args(), *coords.args(), _D(amin), _I(nmax), _p(freqs), _I(len(freqs)),
                           _I(kind), _D(diskR), _I(nthreads), _p(out))
    return out


def cheerleader(setup, coords, amin, nmax, f, kind=PROP_PROPAGATOR, emit=None, reflect=None, diskR=0.1,
                returnboth=False):
    K = (coords.ny, coords.nx)
    right = np.zeros(K, dtype=np.complex128)
    left = np.zeros(K, dtype=np.complex128) if returnboth else None
    lib().bfo_cheerleader(*setup.args(), *coords.args(), _D(amin), _I(nmax), _D(f), _I(kind), _p(_c(emit)),
                          _p(_c(reflect)), _D(diskR), _p(right), _p(left))
    return (left, right) if returnboth else right


def besselj_zero(nu, m):
    lib().bfo_besselj_zero.restype = _D
    return lib().bfo_besselj_zero(_I(nu), _I(m))


def seed_modes(coords, M, L, diskR):
    n = M * (2 * L + 1)
    pat = np.zeros((n, coords.ny, coords.nx), dtype=np.complex128)
    kt = np.zeros(n, dtype=np.complex128)
    lib().bfo_seed_modes(_I(coords.nx), _I(coords.ny), _p(coords.X), _p(coords.Y), _I(M), _I(L), _D(diskR),
                         _p(pat), _p(kt))
    return pat, kt


def field2modes(coords, patterns, B=None, velocity_x=0.0, diskR=0.15, f=20e9):
    patterns = _c(patterns)
    n = patterns.shape[0]
    out = np.zeros(n, dtype=np.complex128)
    lib().bfo_field2modes(_I(coords.nx), _I(coords.ny), _p(coords.X), _p(coords.Y), _I(n), _p(patterns),
                          _p(_c(B)), _D(velocity_x), _D(diskR), _D(f), _p(out))
    return out


def modes2field(patterns, coeffs):
    patterns = _c(patterns)
    n = patterns.shape[0]
    out = np.zeros(patterns.shape[1:], dtype=np.complex128)
    lib().bfo_modes2field(_I(int(np.prod(patterns.shape[1:]))), _I(n), _p(patterns), _p(_c(coeffs)), _p(out))
    return out


def propagation_matrix(dz, diskR, eps, tilt_x, tilt_y, surface, lam, coords, patterns, kt,
                       kind=PROP_PROPAGATOR, onlydiagonal=False):
    patterns = _c(patterns)
    n = patterns.shape[0]
    out = np.zeros((n, n), dtype=np.complex128, order="F")
    eps = complex(eps)
    lib().bfo_propagation_matrix(_D(dz), _D(diskR), _D(eps.real), _D(eps.imag), _D(tilt_x), _D(tilt_y),
                                 _p(_c(surface)), _D(lam), *coords.args(), _I(n), _p(patterns), _p(_c(kt)),
                                 _I(kind), _I(int(onlydiagonal)), out.ctypes.data_as(ctypes.c_void_p))
    return out


def _fortran_stack(mats):
    """list of n x n matrices -> contiguous buffer of column-major matrices"""
    return np.ascontiguousarray(np.stack([np.asarray(m, dtype=np.complex128).T for m in mats]))


def transformer(setup, coords, patterns, kt, f, kind=PROP_PROPAGATOR, prop_mats=None, diskR=0.15, emit=None,
                reflect=None):
    patterns = _c(patterns)
    n = patterns.shape[0]
    boost = np.zeros(n, dtype=np.complex128)
    refl = np.zeros(n, dtype=np.complex128) if reflect is not None else None
    pm = _fortran_stack(prop_mats) if prop_mats is not None else None
    lib().bfo_transformer(*setup.args(with_r=False), *coords.args(), _I(n), _p(patterns), _p(_c(kt)), _D(f),
                          _I(kind), _p(pm), _D(diskR), _p(_c(emit)), _p(_c(reflect)), _p(boost), _p(refl))
    return (boost, refl) if reflect is not None else boost


def transformer_sweep(setup, coords, patterns, kt, freqs, emit, kind=PROP_PROPAGATOR, diskR=0.15, nthreads=1):
    patterns = _c(patterns)
    n = patterns.shape[0]
    freqs = _d(freqs)
    boost = np.zeros((len(freqs), n), dtype=np.complex128)
    lib().bfo_transformer_sweep(*setup.args(with_r=False), *coords.args(), _I(n), _p(patterns), _p(_c(kt)),
                                _p(freqs), _I(len(freqs)), _I(kind), _D(diskR), _p(_c(emit)), _I(nthreads),
                                _p(boost))
    return boost


def transformer_trace_back(reflected, inp, setup, coords, patterns, kt, f, kind=PROP_PROPAGATOR, prop_mats=None,
                           diskR=0.15, includes_axion=False, emit=None):
    patterns = _c(patterns)
    n = patterns.shape[0]
    out = np.zeros((setup.nreg, 2, n), dtype=np.complex128)
    pm = _fortran_stack(prop_mats) if prop_mats is not None else None
    lib().bfo_transformer_trace_back(_p(_c(np.broadcast_to(reflected, (n,)))), _p(_c(np.broadcast_to(inp, (n,)))),
                                     *setup.args(with_r=False), *coords.args(), _I(n), _p(patterns), _p(_c(kt)),
                                     _D(f), _I(kind), _p(pm), _D(diskR), _I(int(includes_axion)), _p(_c(emit)),
                                     _p(out))
    return out
