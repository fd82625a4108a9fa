"""Mode-matching transformer (src/boost_fractor_transformer.jl), on the GPU.

Array conventions (Julia column-major -> numpy C order, spatial axes last):
  mode_patterns (M, 2L+1, Nx, Ny, C)  ->  (M, 2L+1, C, ny, nx)
  mode_kt       (M, 2L+1)             ->  (M, 2L+1)
  modes2field   (Nx, Ny, C)           ->  (C, ny, nx)
  fields_regions (Nreg, 2, n)         ->  (Nreg, 2, n)   (same element indexing)
Mode vectors use the flat index (m-1)(2L+1) + l + L (m = 1..M, l = -L..L).

`f` may be a scalar (reference signature) or a sequence: the frequencies then
run batched on the device and results gain a leading frequency axis.
"""
import ctypes

import numpy as np

from . import _lib as L
from .propagators import precision_code, prop_kind, propagator
from .setup import SPEED_OF_LIGHT, SetupBoundaries, wavelength


class ErrorException(RuntimeError):
    """Julia's error(...)"""


class Modes:
    """src/boost_fractor_transformer.jl:31-38"""

    def __init__(self, M, L_, mode_patterns, mode_kt, id_=None, zeromatrix=None):
        self.M = int(M)
        self.L = int(L_)
        self.mode_patterns = np.asarray(mode_patterns, dtype=np.complex128)
        self.mode_kt = np.asarray(mode_kt, dtype=np.complex128)
        n = self.M * (2 * self.L + 1)
        self.id = np.eye(n, dtype=np.complex128) if id_ is None else id_
        self.zeromatrix = np.zeros((n, n), dtype=np.complex128) if zeromatrix is None else zeromatrix

    @property
    def n(self):
        return self.M * (2 * self.L + 1)

    @property
    def ncomp(self):
        return self.mode_patterns.shape[2]

    def patterns_flat(self):
        """[n][C][ny][nx] with n = (m-1)(2L+1) + l + L"""
        p = self.mode_patterns
        return p.reshape((self.n,) + p.shape[2:])


def e_field_dimensions(modes):
    """src/boost_fractor_transformer.jl:23: 1 (scalar field) or 3."""
    return modes.mode_patterns.shape[2]


def index(modes, k):
    """src/boost_fractor_transformer.jl:76-78 (0-based slice of sub-block k = 1, 2)."""
    n = modes.n
    return slice((k - 1) * n, k * n)


def _seed_bessel(coords, M, L_, diskR):
    n = M * (2 * L_ + 1)
    pat = np.zeros((n, coords.ny, coords.nx), dtype=np.complex128)
    kt = np.zeros(n, dtype=np.complex128)
    m = L.Marshalled()
    c = m.coords(coords)
    ctx = L.context()
    L.check(L.lib().bf_seed_modes(ctx.handle, ctypes.byref(c), int(M), int(L_), float(diskR), L.cptr(pat),
                                  L.cptr(kt)), ctx.handle)
    return pat, kt


def SeedModes(coords, ThreeDim=False, Mmax=1, Lmax=0, diskR=0.15, pattern_input=None, kT_arr=None):
    """src/boost_fractor_transformer.jl:46-69"""
    if ThreeDim:
        M, Lm = int(Mmax), int(Lmax)
        if pattern_input is None:
            pat, kt = _seed_bessel(coords, M, Lm, diskR)
            patterns = pat.reshape(M, 2 * Lm + 1, 1, coords.ny, coords.nx)
            mode_kt = kt.reshape(M, 2 * Lm + 1)
        else:
            if kT_arr is None:
                raise ErrorException("you have to define kTs as well when putting in an input pattern")
            pi = np.asarray(pattern_input, dtype=np.complex128)
            patterns = np.ascontiguousarray(pi[:M, :2 * Lm + 1])
            mode_kt = np.asarray(kT_arr, dtype=np.complex128)[:M, :2 * Lm + 1].copy()
    else:
        M, Lm = 1, 0
        patterns = np.ones((1, 1, 1, coords.ny, coords.nx), dtype=np.complex128)
        mode_kt = np.zeros((1, 1), dtype=np.complex128)
    return Modes(M, Lm, patterns, mode_kt)


def mode(m, l, L_, coords, diskR=0.15, k0=2 * np.pi / 0.03, pattern_input=None, kT_arr=None):
    """src/boost_fractor_transformer.jl:88-107 -> (kt, pattern (C, ny, nx))."""
    if pattern_input is None:
        pat, kt = _seed_bessel(coords, m, abs(l), diskR)
        idx = (m - 1) * (2 * abs(l) + 1) + l + abs(l)
        return kt[idx], pat[idx][None]
    if kT_arr is None:
        raise ErrorException("you have to define kTs as well when putting in an input pattern")
    return complex(np.asarray(kT_arr)[m - 1, l + L_]), np.asarray(pattern_input)[m - 1, l + L_]


def _modes_struct(m, modes):
    return m.modes(modes)


def axion_induced_modes(coords, modes, B=None, velocity_x=0, diskR=0.15, f=20e9):
    """src/boost_fractor_transformer.jl:115-151 -> mode vector (n,)."""
    C = modes.ncomp
    if B is None and C == 3:
        B = np.zeros((3, coords.ny, coords.nx), dtype=np.complex128)
        B[1] = 1.0
    m = L.Marshalled()
    c = m.coords(coords)
    md = _modes_struct(m, modes)
    b = None
    if B is not None:
        b = np.asarray(B, dtype=np.complex128)
        if b.ndim == 2:
            b = np.broadcast_to(b, (C,) + b.shape)
        b = m.hold(L.as_c128(b))
        if b.shape != (C, coords.ny, coords.nx):
            raise L.DimensionMismatch(f"B has shape {b.shape}, expected {(C, coords.ny, coords.nx)}")
    out = np.zeros(modes.n, dtype=np.complex128)
    ctx = L.context()
    L.check(L.lib().bf_field2modes(ctx.handle, ctypes.byref(c), ctypes.byref(md), L.cptr(b), float(velocity_x),
                                   float(diskR), float(f), L.cptr(out)), ctx.handle)
    return out


def field2modes(pattern, coords, modes, diskR=0.15):
    """src/boost_fractor_transformer.jl:156 (the caller's array is not modified)."""
    return axion_induced_modes(coords, modes, B=pattern, diskR=diskR)


def modes2field(mode_coeffs, coords, modes):
    """src/boost_fractor_transformer.jl:161-167 -> field (C, ny, nx)."""
    m = L.Marshalled()
    c = m.coords(coords)
    md = _modes_struct(m, modes)
    coeffs = m.hold(L.as_c128(np.asarray(mode_coeffs).reshape(-1)))
    if len(coeffs) != modes.n:
        raise L.DimensionMismatch(f"{len(coeffs)} coefficients for {modes.n} modes")
    out = np.zeros((modes.ncomp, coords.ny, coords.nx), dtype=np.complex128)
    ctx = L.context()
    L.check(L.lib().bf_modes2field(ctx.handle, ctypes.byref(c), ctypes.byref(md), L.cptr(coeffs), L.cptr(out)),
            ctx.handle)
    return out


def _fortran_stack(mats, n):
    return np.ascontiguousarray(np.stack([np.asarray(x, dtype=np.complex128).reshape(n, n).T for x in mats]))


def _run_matrices(bdry, coords, modes, freqs, prop, diskR, precision):
    m = L.Marshalled()
    c = m.coords(coords)
    s = m.setup(bdry, coords.ny, coords.nx)
    md = _modes_struct(m, modes)
    fr = m.hold(L.as_f64(freqs))
    n, nreg = modes.n, len(bdry.distance)
    out = np.zeros((len(fr), nreg, n, n), dtype=np.complex128)  # column-major matrices
    ctx = L.context()
    L.check(L.lib().bf_propagation_matrices(ctx.handle, ctypes.byref(s), ctypes.byref(c), ctypes.byref(md),
                                            L.cptr(fr), len(fr), prop_kind(prop), float(diskR),
                                            precision_code(precision), L.cptr(out)), ctx.handle)
    return np.swapaxes(out, -1, -2)  # -> row-major indexing P[f][reg][i, j]


def propagation_matrix(dz, diskR, eps, tilt_x, tilt_y, surface, lambda_, coords, modes, is_air=None,
                       onlydiagonal=False, prop=propagator):
    """src/boost_fractor_transformer.jl:174-227 (one region)."""
    eps = complex(eps)
    if is_air is not None and bool(is_air) != (eps.real == 1.0):
        raise L.UnsupportedError("is_air must follow real(eps) == 1 on the device engine")
    surf = np.zeros((1, coords.ny, coords.nx), dtype=np.complex128)
    if surface is not None:
        surf[0] = np.broadcast_to(np.asarray(surface), (coords.ny, coords.nx))
    b = SetupBoundaries([dz], np.zeros(0), [eps], [tilt_x], [tilt_y], surf)
    P = _run_matrices(b, coords, modes, [SPEED_OF_LIGHT / float(lambda_)], prop, diskR, "fp64")[0, 0]
    if onlydiagonal:
        P = np.diag(np.diag(P))
    return P


def calc_propagation_matrices(bdry, coords, modes, f=10.0e9, prop=propagator, diskR=0.15, precision="fp64"):
    """src/boost_fractor_transformer.jl:274-280: list of Nreg n x n matrices (list of lists for several f)."""
    freqs = np.atleast_1d(np.asarray(f, dtype=np.float64))
    P = _run_matrices(bdry, coords, modes, freqs, prop, diskR, precision)
    res = [[P[q, r] for r in range(P.shape[1])] for q in range(P.shape[0])]
    return res[0] if np.ndim(f) == 0 else res


def _nan_to_1e30(bdry):
    bad = np.isnan(bdry.eps)  # complex: either part NaN
    if bad.any():
        bdry.eps[bad] = 1e30  # the reference mutates its argument (src/boost_fractor_transformer.jl:289)


def transformer(bdry, coords, modes, f=10.0e9, velocity_x=0, prop=propagator, propagation_matrices=None, diskR=0.15,
                emit=None, reflect=None, precision="fp64"):
    """src/boost_fractor_transformer.jl:285-359 -> boost (n,) or (boost, refl); batched over f."""
    _nan_to_1e30(bdry)
    freqs = np.atleast_1d(np.asarray(f, dtype=np.float64))
    scalar = np.ndim(f) == 0
    n, nreg = modes.n, len(bdry.distance)
    if emit is None and velocity_x != 0 and len(freqs) > 1:
        # the default emission depends on f through the velocity phase: one call per frequency
        res = [transformer(bdry, coords, modes, f=float(q), velocity_x=velocity_x, prop=prop,
                           propagation_matrices=(propagation_matrices[i] if propagation_matrices else None),
                           diskR=diskR, reflect=reflect, precision=precision) for i, q in enumerate(freqs)]
        if reflect is None:
            return np.stack(res)
        return np.stack([r[0] for r in res]), np.stack([r[1] for r in res])
    if emit is None:
        emit = axion_induced_modes(coords, modes, B=None, velocity_x=velocity_x, diskR=diskR, f=float(freqs[0]))
    m = L.Marshalled()
    c = m.coords(coords)
    s = m.setup(bdry, coords.ny, coords.nx)
    md = _modes_struct(m, modes)
    fr = m.hold(L.as_f64(freqs))
    em = m.hold(L.as_c128(np.asarray(emit).reshape(-1)))
    rf = m.hold(L.as_c128(np.broadcast_to(reflect, (n,)))) if reflect is not None else None
    pm = None
    if propagation_matrices is not None and len(propagation_matrices) > 0:
        mats = propagation_matrices
        if scalar or not isinstance(mats[0], (list, tuple)):
            mats = [mats] * len(fr)
        pm = m.hold(np.ascontiguousarray(np.stack([_fortran_stack(x, n) for x in mats])))
        if pm.shape != (len(fr), nreg, n, n):
            raise L.DimensionMismatch(f"propagation_matrices has shape {pm.shape}, expected {(len(fr), nreg, n, n)}")
    boost = np.zeros((len(fr), n), dtype=np.complex128)
    refl = np.zeros((len(fr), n), dtype=np.complex128) if reflect is not None else None
    ctx = L.context()
    L.check(L.lib().bf_transformer(ctx.handle, ctypes.byref(s), ctypes.byref(c), ctypes.byref(md), L.cptr(fr),
                                   len(fr), prop_kind(prop), L.cptr(pm), float(diskR), L.cptr(em), L.cptr(rf),
                                   precision_code(precision), L.cptr(boost), L.cptr(refl)), ctx.handle)
    if scalar:
        boost = boost[0]
        refl = refl[0] if refl is not None else None
    return boost if reflect is None else (boost, refl)


class TransformerPlan:
    """Device-resident frequency sweep of `transformer` over one geometry
    (src/boost_fractor_transformer.jl:285-359): the setup, the modes and
    emit/reflect are uploaded once; every `run()` recomputes the sweep from them
    on the device; `fetch()` returns boost (nfreq, n) [and refl].  NaN
    permittivities are read as 1e30 (the caller's setup is not modified)."""

    def __init__(self, bdry, coords, modes, f, prop=propagator, diskR=0.15, emit=None, reflect=None, ctx=None):
        self.ctx = ctx or L.context()
        self.freqs = np.ascontiguousarray(np.atleast_1d(np.asarray(f, dtype=np.float64)))
        self.n = modes.n
        if emit is None:
            emit = axion_induced_modes(coords, modes, B=None, diskR=diskR, f=float(self.freqs[0]))
        m = L.Marshalled()
        c = m.coords(coords)
        s = m.setup(bdry, coords.ny, coords.nx)
        md = _modes_struct(m, modes)
        em = m.hold(L.as_c128(np.asarray(emit).reshape(-1)))
        if len(em) != self.n:
            raise L.DimensionMismatch(f"emit of length {len(em)} for {self.n} modes")
        rf = m.hold(L.as_c128(np.broadcast_to(reflect, (self.n,)))) if reflect is not None else None
        self.has_reflect = reflect is not None
        h = ctypes.c_void_p()
        L.check(L.lib().bf_transformer_plan_create(self.ctx.handle, ctypes.byref(s), ctypes.byref(c),
                                                   ctypes.byref(md), L.cptr(self.freqs), len(self.freqs),
                                                   prop_kind(prop), float(diskR), L.cptr(em), L.cptr(rf),
                                                   ctypes.byref(h)), self.ctx.handle)
        self.handle = h

    def run(self):
        L.check(L.lib().bf_transformer_plan_run(self.handle), self.ctx.handle)

    def fetch(self):
        boost = np.zeros((len(self.freqs), self.n), dtype=np.complex128)
        refl = np.zeros_like(boost) if self.has_reflect else None
        L.check(L.lib().bf_transformer_plan_fetch(self.handle, L.cptr(boost), L.cptr(refl)), self.ctx.handle)
        return boost if refl is None else (boost, refl)

    def close(self):
        if getattr(self, "handle", None):
            L.lib().bf_transformer_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _batch_layout(nconf, modesets, mode_set_of, nf):
    """mode_set_of as int32 (default all 0), checked; per-geometry output offsets and sizes"""
    nsets = len(modesets)
    if mode_set_of is None:
        mode_set_of = np.zeros(nconf, dtype=np.int32)
    mode_set_of = np.ascontiguousarray(mode_set_of, dtype=np.int32)
    if mode_set_of.shape != (nconf,):
        raise L.DimensionMismatch(f"mode_set_of has shape {mode_set_of.shape}, expected {(nconf,)}")
    if np.any(mode_set_of < 0) or np.any(mode_set_of >= nsets):
        raise L.ArgumentError("mode_set_of entries must index modesets")
    sizes = np.array([modesets[s].n * nf for s in mode_set_of], dtype=np.int64)
    offs = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64))
    return mode_set_of, offs, sizes


def _batch_sources(m, coords, modesets, emits, diskR, f0):
    """marshalled mode sets and one axion source per set (default: axion_induced_modes, uniform field);
    coords: the grid of every set, or a list with each set's grid"""
    nsets = len(modesets)
    grids = coords if isinstance(coords, (list, tuple)) else [coords] * nsets
    for g, md in zip(grids, modesets):
        if tuple(md.mode_patterns.shape[-2:]) != (g.ny, g.nx):
            raise L.DimensionMismatch(f"mode patterns of shape {md.mode_patterns.shape[-2:]} on a "
                                      f"{g.ny}x{g.nx} grid")
    if emits is None:
        emits = [axion_induced_modes(g, md, B=None, diskR=diskR, f=f0) for g, md in zip(grids, modesets)]
    if len(emits) != nsets:
        raise L.DimensionMismatch(f"{len(emits)} emits for {nsets} mode sets")
    mds = (L.ModesS * nsets)(*[_modes_struct(m, md) for md in modesets])
    em = []
    for md, e in zip(modesets, emits):
        a = m.hold(L.as_c128(np.asarray(e).reshape(-1)))
        if len(a) != md.n:
            raise L.DimensionMismatch(f"emit of length {len(a)} for a mode set of {md.n} modes")
        em.append(a.ctypes.data)
    return mds, (ctypes.c_void_p * nsets)(*em)


class TransformerBatchPlan:
    """Device-resident disk-position sweep (transformer_batch) for an optimiser
    that evaluates many candidate geometries against fixed modes
    (examples/transformer_optim_utilities.jl:105,121,202): the grid, the mode sets,
    their axion sources (emits) and the frequencies are uploaded once; each
    `run(bdrys, mode_set_of)` takes the candidate setups from the host and
    computes every boost on the device; `fetch()` returns one (nfreq, n) array per
    geometry of the last run.  Like transformer, NaN permittivities of each setup
    are replaced by 1e30 in place.

    coords: one coordinate system for every mode set, or a list with one per mode
    set (grids of different sizes in one plan; a geometry's surface maps are
    sampled on its mode set's grid)."""

    def __init__(self, coords, modesets, f, prop=propagator, diskR=0.15, emits=None, precision="fp64", ctx=None):
        self.ctx = ctx or L.context()
        self.modesets = list(modesets) if isinstance(modesets, (list, tuple)) else [modesets]
        self.freqs = np.ascontiguousarray(np.atleast_1d(np.asarray(f, dtype=np.float64)))
        nsets = len(self.modesets)
        per_set = isinstance(coords, (list, tuple))
        if per_set and len(coords) != nsets:
            raise L.DimensionMismatch(f"{len(coords)} grids for {nsets} mode sets")
        self.coords = coords
        self._grids = list(coords) if per_set else [coords] * nsets
        m = L.Marshalled()
        mds, emp = _batch_sources(m, coords, self.modesets, emits, diskR, float(self.freqs[0]))
        h = ctypes.c_void_p()
        args = (mds, emp, L.cptr(self.freqs), len(self.freqs), prop_kind(prop), float(diskR),
                precision_code(precision), ctypes.byref(h))
        if per_set:
            cs = (L.Coords * nsets)(*[m.coords(g) for g in coords])
            rc = L.lib().bf_transformer_batch_plan_create_grids(self.ctx.handle, nsets, cs, *args)
        else:
            c = m.coords(coords)
            rc = L.lib().bf_transformer_batch_plan_create(self.ctx.handle, ctypes.byref(c), nsets, *args)
        L.check(rc, self.ctx.handle)
        self.handle = h
        self._layout = None

    def run(self, bdrys, mode_set_of=None):
        nconf = len(bdrys)
        mode_set_of, offs, sizes = _batch_layout(nconf, self.modesets, mode_set_of, len(self.freqs))
        m = L.Marshalled()
        ny = np.array([g.ny for g in self._grids])[mode_set_of]
        nx = np.array([g.nx for g in self._grids])[mode_set_of]
        setups = m.setups(bdrys, ny, nx, nan_to_1e30=True)
        L.check(L.lib().bf_transformer_batch_plan_run(self.handle, nconf, setups, L.cptr(mode_set_of), L.cptr(offs)),
                self.ctx.handle)
        self._layout = (offs, sizes)

    def fetch(self):
        if self._layout is None:
            raise L.ArgumentError("fetch() before run()")
        offs, sizes = self._layout
        out = np.zeros(int(sizes.sum()), dtype=np.complex128)
        L.check(L.lib().bf_transformer_batch_plan_fetch(self.handle, L.cptr(out)), self.ctx.handle)
        return [out[o:o + sz].reshape(len(self.freqs), -1) for o, sz in zip(offs, sizes)]

    def close(self):
        if getattr(self, "handle", None):
            L.lib().bf_transformer_batch_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def transformer_batch(bdrys, coords, modesets, f=10.0e9, mode_set_of=None, prop=propagator, diskR=0.15, emits=None,
                      precision="fp64"):
    """The disk-position sweep of examples/transformer_optim_utilities.jl:105,121,202
    (calc_propagation_matrices, then transformer(...; propagation_matrices=...), per
    candidate geometry) in one device call.

    bdrys: list of SetupBoundaries (region counts may differ); modesets: a Modes or
    a list of Modes; geometry c uses modesets[mode_set_of[c]] (default 0).  emits:
    one axion mode vector per mode set (default: axion_induced_modes with a uniform
    field).  Returns a list with one (nfreq, n) boost array per geometry (n of its
    mode set), equal to transformer(bdrys[c], coords, modes, f=freqs, ...).
    Like transformer, NaN permittivities of each setup are replaced by 1e30 in place."""
    if not isinstance(modesets, (list, tuple)):
        modesets = [modesets]
    nconf, nsets = len(bdrys), len(modesets)
    freqs = np.atleast_1d(np.asarray(f, dtype=np.float64))
    mode_set_of, offs, sizes = _batch_layout(nconf, modesets, mode_set_of, len(freqs))
    m = L.Marshalled()
    c = m.coords(coords)
    mds, emp = _batch_sources(m, coords, modesets, emits, diskR, float(freqs[0]))
    setups = m.setups(bdrys, coords.ny, coords.nx, nan_to_1e30=True)
    fr = m.hold(L.as_f64(freqs))
    nf = len(fr)
    out = np.zeros(int(sizes.sum()), dtype=np.complex128)
    ctx = L.context()
    L.check(L.lib().bf_transformer_batch(ctx.handle, nconf, setups, ctypes.byref(c), nsets, mds, emp,
                                         L.cptr(mode_set_of), L.cptr(fr), nf, prop_kind(prop), float(diskR),
                                         precision_code(precision), L.cptr(offs), L.cptr(out)), ctx.handle)
    return [out[o:o + sz].reshape(nf, -1) for o, sz in zip(offs, sizes)]


def transformer_trace_back(reflected_beam, input_beam, bdry, coords, modes, inlcudes_axion=False, f=10.0e9,
                           velocity_x=0, prop=propagator, propagation_matrices=None, diskR=0.15, emit=None):
    """src/boost_fractor_transformer.jl:372-421 -> fields_regions (Nreg, 2, n)."""
    n, nreg = modes.n, len(bdry.distance)
    if emit is None:
        emit = axion_induced_modes(coords, modes, B=np.ones((coords.ny, coords.nx)), diskR=diskR)
    m = L.Marshalled()
    c = m.coords(coords)
    s = m.setup(bdry, coords.ny, coords.nx)
    md = _modes_struct(m, modes)
    refl = m.hold(L.as_c128(np.broadcast_to(reflected_beam, (n,))))
    inp = m.hold(L.as_c128(np.broadcast_to(input_beam, (n,))))
    em = m.hold(L.as_c128(np.asarray(emit).reshape(-1)))
    pm = None
    if propagation_matrices is not None:
        pm = m.hold(_fortran_stack(propagation_matrices, n))
    out = np.zeros((nreg, 2, n), dtype=np.complex128)
    ctx = L.context()
    L.check(L.lib().bf_transformer_trace_back(ctx.handle, L.cptr(refl), L.cptr(inp), ctypes.byref(s),
                                              ctypes.byref(c), ctypes.byref(md), float(f), prop_kind(prop),
                                              L.cptr(pm), float(diskR), int(bool(inlcudes_axion)), L.cptr(em),
                                              L.cptr(out)), ctx.handle)
    return out
