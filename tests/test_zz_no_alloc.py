"""The allocation-free chain (bf_transformer_chain_noalloc) against the existing
entry point it varies, bf_transformer with precomputed propagation matrices
(transformer(...; propagation_matrices=P), src/boost_fractor_transformer.jl:285-359).

Both run the same inverse and chain kernels on the same matrices, so boost
and reflectivity agree bit for bit.  Two calls in a row, the second smaller
(fewer regions, modes and frequencies): the shared scratch array keeps the
first call's contents and must not leak into the second.  Device buffers come
from the HIP runtime the library has loaded (hipMalloc / hipMemcpy through
ctypes), in the test only: the entry point itself allocates nothing.
"""
import ctypes

import numpy as np
import pytest

import BoostFractor as B
from BoostFractor import _lib as L
from BoostFractor import workloads as W
from helpers import DISKR

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        self.h = ctypes.CDLL("libamdhip64.so")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipSetDevice.argtypes = [ctypes.c_int]
        self.bufs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipSetDevice(L.context().device) == 0
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def get(self, p, shape):
        out = np.zeros(shape, dtype=np.complex128)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _lib():
    fn = L.lib().bf_transformer_chain_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    return fn


def _case(N, ndisk, M, Lm, freqs, seed):
    coords = W.grid_coords(N)
    sb = W.booster(coords, ndisk=ndisk, seed=seed)
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=M, Lmax=Lm, diskR=DISKR)
    emit = B.axion_induced_modes(coords, modes, diskR=DISKR)
    refl_in = np.linspace(1.0, 2.0, modes.n) + 0.25j
    mats = B.calc_propagation_matrices(sb, coords, modes, f=freqs, diskR=DISKR)
    return coords, sb, modes, emit, refl_in, mats


def _run_both(N, ndisk, M, Lm, freqs, seed, hip):
    coords, sb, modes, emit, refl_in, mats = _case(N, ndisk, M, Lm, freqs, seed)
    n, nreg, nf = modes.n, len(sb.distance), len(freqs)
    print(f"no-alloc chain: {ndisk} disks, n={n}, {nf} freqs", flush=True)
    want_b, want_r = B.transformer(sb.copy(), coords, modes, f=freqs, propagation_matrices=mats, diskR=DISKR,
                                   emit=emit, reflect=refl_in)
    # the library's layout: [nfreq][nreg][n][n], each matrix column-major
    pm = np.ascontiguousarray(np.stack([np.stack([np.asarray(m).T for m in mf]) for mf in mats]))
    eps = np.ascontiguousarray(sb.eps, dtype=np.complex128)  # NaN mirror, taken as 1e30 by the entry point
    assert np.isnan(eps[0].real)
    dP, de, dr = hip.put(pm), hip.put(np.asarray(emit, dtype=np.complex128)), hip.put(refl_in.astype(np.complex128))
    db, dref = hip.alloc(16 * nf * n), hip.alloc(16 * nf * n)
    ctx = L.context()
    rc = _lib()(ctx.handle, nreg, eps.ctypes.data, n, nf, dP, de, dr, db, dref)
    L.check(rc, ctx.handle)
    got_b, got_r = hip.get(db, (nf, n)), hip.get(dref, (nf, n))
    assert np.any(got_b != 0)
    assert np.array_equal(got_b, want_b), np.max(np.abs(got_b - want_b))
    assert np.array_equal(got_r, want_r), np.max(np.abs(got_r - want_r))
    # without reflect: boost only
    rc = _lib()(ctx.handle, nreg, eps.ctypes.data, n, nf, dP, de, None, db, None)
    L.check(rc, ctx.handle)
    assert np.array_equal(hip.get(db, (nf, n)), want_b)


def test_noalloc_chain_matches_transformer_two_calls():
    hip = _Hip()
    try:
        _run_both(48, 3, 3, 1, [19.7e9, 19.95e9, 20.1e9, 20.3e9, 20.45e9], 1201, hip)  # n = 9, 8 regions
        _run_both(32, 1, 2, 0, [19.9e9, 20.2e9], 1202, hip)                           # n = 2, 4 regions
    finally:
        hip.free()


def test_noalloc_chain_rejects_oversized_calls():
    """beyond the scratch array, more than 48 modes or 512 regions: an error before any device work"""
    ctx = L.context()
    eps = np.ones(8, dtype=np.complex128)
    fake = ctypes.c_void_p(16)  # never dereferenced: the checks run on the host first
    fn = _lib()
    # 8 regions x 48^2 x 16 B = 295 KB per frequency: 120 frequencies exceed 32 MiB
    assert fn(ctx.handle, 8, eps.ctypes.data, 48, 120, fake, fake, None, fake, None) == L.BF_ERR_UNSUPPORTED
    assert b"32 MiB" in L.lib().bf_last_error(ctx.handle)
    assert fn(ctx.handle, 8, eps.ctypes.data, 49, 1, fake, fake, None, fake, None) == L.BF_ERR_UNSUPPORTED
    big = np.ones(513, dtype=np.complex128)
    assert fn(ctx.handle, 513, big.ctypes.data, 4, 1, fake, fake, None, fake, None) == L.BF_ERR_UNSUPPORTED
    assert fn(ctx.handle, 8, eps.ctypes.data, 4, 1, fake, fake, fake, fake, None) == L.BF_ERR_ARG
    assert fn(ctx.handle, 1, eps.ctypes.data, 4, 1, fake, fake, None, fake, None) == L.BF_ERR_DIM
