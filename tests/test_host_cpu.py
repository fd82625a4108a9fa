"""Host-side tests (no device): grids, setup, the C-ABI library's exports and
its argument validation, the Python mirror's marshalling and error mapping."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import BoostFractor as B
from BoostFractor import _lib as L
from BoostFractor import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "boostfractor_hip.h")


def test_julia_ranges():
    X = B.julia_range(-0.5, 0.01, 0.5)
    assert len(X) == 101 and X[0] == -0.5 and X[50] == 0.0 and X[-1] == 0.5
    assert X[1] == -0.49 and X[37] == -0.13  # exact decimal elements
    X = B.julia_range(-0.4445, 0.007, 0.4445)
    assert len(X) == 128 and X[0] == -0.4445 and X[-1] == 0.4445
    k = B.get_kspace_coords(X)
    assert len(k) == 128 and k[0] == -448.79895051282756 and k[-1] == 448.79895051282756
    assert k[1] == -441.73125050475153  # as printed by examples/Beam Shapes.ipynb cell 13
    c = B.SeedCoordinateSystem()
    assert c.nx == 101 and c.kX[50] == 0.0 and abs(c.kX[1] - c.kX[0] - 2 * math.pi) < 1e-12


def test_kspace_even_grid_dc_offset():
    """dk = 2pi/(2 max X); an even grid's fftshifted DC bin carries dk/2 (not 0)"""
    c = W.grid_coords(256)
    dk = 2 * math.pi / (2 * c.X.max())
    assert abs(c.kX[128] - dk / 2) < 1e-9 and abs(c.kX[1] - c.kX[0] - dk) < 1e-9


def test_seed_setup_boundaries():
    c = B.SeedCoordinateSystem()
    sb = B.SeedSetupBoundaries(c, diskno=1)
    assert np.isnan(sb.eps[0].real)
    assert list(sb.distance) == [0.0, 8e-3, 1e-3, 0.0]
    # real eps stored as re - 0.0im
    assert all(math.copysign(1.0, e.imag) == -1.0 for e in sb.eps)
    assert np.allclose(sb.r, [1.0, -0.5, 0.5])
    with pytest.raises(B.DimensionMismatch):
        B.SeedSetupBoundaries(c, diskno=1, distance=[0.0, 1.0, 0.0])


def test_booster_generator_is_seeded():
    c = W.grid_coords(32)
    a = W.booster(c, ndisk=20, seed=1000)
    b = W.booster(c, ndisk=20, seed=1000)
    assert np.array_equal(a.distance, b.distance) and len(a.distance) == 42
    s = a.distance[1:-1:2]
    assert np.all(np.abs(s - 8e-3) <= 0.3e-3) and np.all(a.distance[2:-1:2] == 1e-3)
    assert a.eps[2] == 24 and a.eps[1] == 1
    assert W.SplitMix64(0).next_u64() == 0xE220A8397B1DCDAF  # published SplitMix64 first output for seed 0


def _header_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(bf_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_header_symbol():
    lib = L.lib()
    syms = _header_symbols()
    assert len(syms) >= 24
    for s in syms:
        assert hasattr(lib, s), s
    assert lib.bf_version() >= 10000


def _coords_struct(nx=8, ny=8):
    m = L.Marshalled()
    c = W.grid_coords(nx)
    return m, m.coords(c)


def test_cabi_validation_without_device():
    """argument and dimension checks run on the host before any device work"""
    lib = L.lib()
    m, c = _coords_struct()
    E = np.zeros((8, 8), dtype=complex)
    # unknown propagator kind -> ArgumentError
    rc = lib.bf_propagate(None, 7, ctypes.byref(c), L.cptr(E), 0.0, 0.1, 1.0, 0.0, 0.0, 0.0, None, 0.03)
    assert rc == L.BF_ERR_ARG and b"propagator" in lib.bf_last_error(None)
    # grid too large
    big = L.Coords(2048, 8, c.X, c.Y, c.kX, c.kY)
    rc = lib.bf_propagate(None, 0, ctypes.byref(big), L.cptr(E), 0.0, 0.1, 1.0, 0.0, 0.0, 0.0, None, 0.03)
    assert rc == L.BF_ERR_UNSUPPORTED
    # valid arguments but no context -> ArgumentError, raised only after the checks
    rc = lib.bf_propagate(None, 0, ctypes.byref(c), L.cptr(E), 0.0, 0.1, 1.0, 0.0, 0.0, 0.0, None, 0.03)
    assert rc == L.BF_ERR_ARG and b"ctx" in lib.bf_last_error(None)
    # too few regions -> DimensionMismatch
    sb = W.booster(W.grid_coords(8), ndisk=1)
    s = m.setup(sb, 8, 8)
    s.nreg = 1
    f = np.array([20e9])
    out = np.zeros((1, 8, 8), dtype=complex)
    rc = lib.bf_dancer(None, ctypes.byref(s), ctypes.byref(c), 0.0, 10, L.cptr(f), 1, 0, None, None, 0.15, 1, 1, 0,
                       L.cptr(out), None, None)
    assert rc == L.BF_ERR_DIM
    with pytest.raises(B.DimensionMismatch):
        L.check(rc)
    # bad precision / negative frequency
    s.nreg = len(sb.distance)
    rc = lib.bf_dancer(None, ctypes.byref(s), ctypes.byref(c), 0.0, 10, L.cptr(f), 1, 0, None, None, 0.15, 1, 1, 9,
                       L.cptr(out), None, None)
    assert rc == L.BF_ERR_ARG
    fneg = np.array([-1.0])
    rc = lib.bf_dancer(None, ctypes.byref(s), ctypes.byref(c), 0.0, 10, L.cptr(fneg), 1, 0, None, None, 0.15, 1, 1, 0,
                       L.cptr(out), None, None)
    assert rc == L.BF_ERR_ARG
    # includes_axion with several modes is a DimensionMismatch in the reference
    pat = np.ones((3, 8, 8), dtype=complex)
    kt = np.zeros(3, dtype=complex)
    md = L.ModesS(3, 1, L.cptr(pat), L.cptr(kt))
    v = np.zeros(3, dtype=complex)
    fields = np.zeros((len(sb.distance), 2, 3), dtype=complex)
    rc = lib.bf_transformer_trace_back(None, L.cptr(v), L.cptr(v), ctypes.byref(s), ctypes.byref(c),
                                       ctypes.byref(md), 20e9, 0, None, 0.15, 1, L.cptr(v), L.cptr(fields))
    assert rc == L.BF_ERR_DIM
    # dance_intro: an axion velocity needs a complex B field (the reference throws InexactError)
    eps = np.ones(4, dtype=complex)
    out = np.zeros((4, 2, 8, 8), dtype=complex)
    rc = lib.bf_dance_intro(None, ctypes.byref(c), 4, L.cptr(eps), 1.0, 0.0, 0.15, None, 1e-3, 20e9, L.cptr(out))
    assert rc == L.BF_ERR_ARG and b"InexactError" in lib.bf_last_error(None)
    with pytest.raises(B.InexactError):
        B.dance_intro(sb, W.grid_coords(8), velocity_x=1e-3)
    # cheerleader: nmax below -1 is rejected like the dancer's (no device work)
    rc = lib.bf_cheerleader(None, ctypes.byref(s), ctypes.byref(c), 0.0, -2, L.cptr(f), 1, 0, None, None, 0.15, 0,
                            L.cptr(out), None)
    assert rc == L.BF_ERR_ARG and b"nmax" in lib.bf_last_error(None)
    rc = lib.bf_cheerleader(None, ctypes.byref(s), ctypes.byref(c), 0.0, -(2 ** 31), L.cptr(f), 1, 0, None, None,
                            0.15, 0, L.cptr(out), None)
    assert rc == L.BF_ERR_ARG
    # engine options need a context and a known name
    assert lib.bf_ctx_set_option(None, b"prune", 0) == L.BF_ERR_ARG
    v = ctypes.c_int64()
    assert lib.bf_ctx_get_option(None, b"prune", ctypes.byref(v)) == L.BF_ERR_ARG
    assert lib.bf_ctx_set_tuning_file(None, b"/tmp/x") == L.BF_ERR_ARG
    assert lib.bf_ctx_tuning_report(None, None, 0, None) == L.BF_ERR_ARG
    # gauss_profile / unnormalized_match argument checks
    g = np.zeros((8, 8), dtype=complex)
    assert lib.bf_gauss_profile(None, ctypes.byref(c), 1e-9, -0.1, 10e9, L.cptr(g)) == L.BF_ERR_ARG
    assert lib.bf_unnormalized_match(None, 0, 64, L.cptr(g), L.cptr(g), 0.01, 0.01, L.cptr(g)) == L.BF_ERR_ARG


def test_library_reads_no_environment():
    """engine switches are per-context options (bf_ctx_set_option), never
    process-environment variables: no source of the library calls getenv"""
    csrc = os.path.join(ROOT, "boostfractor.jl_amd", "csrc")
    for f in os.listdir(csrc):
        assert "getenv" not in open(os.path.join(csrc, f)).read(), f


def test_mirror_rejects_user_propagators():
    with pytest.raises(B.ArgumentError):
        from BoostFractor.propagators import prop_kind
        prop_kind(lambda *a: a[0])


def test_no_cpu_fallback_when_library_missing(monkeypatch, tmp_path):
    """the mirror raises instead of computing anything on the host"""
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(ImportError):
        L.lib()


def test_oracle_not_imported_by_product():
    """the product never imports, links or loads the CPU oracle"""
    pkg = os.path.join(ROOT, "boostfractor.jl_amd")
    pat = re.compile(r"(import\s+oracle|from\s+oracle|libbf_oracle|bf_oracle\.c|bfo_)")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".cpp", "Makefile")) or f == "Makefile":
                assert not pat.search(open(os.path.join(dirpath, f)).read()), f


def test_batch_plan_validation_without_device():
    """the resident sweep plan's entry points reject bad arguments on the host"""
    lib = L.lib()
    m, c = _coords_struct()
    fr = np.array([20e9])
    h = ctypes.c_void_p()
    # no mode set
    rc = lib.bf_transformer_batch_plan_create(None, ctypes.byref(c), 0, None, None, L.cptr(fr), 1, 0, 0.15, 0,
                                              ctypes.byref(h))
    assert rc == L.BF_ERR_ARG and b"nsets" in lib.bf_last_error(None)
    assert not h.value
    # NULL output handle
    rc = lib.bf_transformer_batch_plan_create(None, ctypes.byref(c), 1, None, None, L.cptr(fr), 1, 0, 0.15, 0, None)
    assert rc == L.BF_ERR_ARG
    # a grid per mode set: no mode set, no grids, a grid of zero size
    h = ctypes.c_void_p(1)
    rc = lib.bf_transformer_batch_plan_create_grids(None, 0, ctypes.byref(c), None, None, L.cptr(fr), 1, 0, 0.15, 0,
                                                    ctypes.byref(h))
    assert rc == L.BF_ERR_ARG and b"nsets" in lib.bf_last_error(None)
    assert not h.value
    rc = lib.bf_transformer_batch_plan_create_grids(None, 1, None, None, None, L.cptr(fr), 1, 0, 0.15, 0,
                                                    ctypes.byref(h))
    assert rc == L.BF_ERR_ARG and b"coords" in lib.bf_last_error(None)
    bad = (L.Coords * 2)(c, L.Coords(0, 4, c.X, c.Y, c.kX, c.kY))
    rc = lib.bf_transformer_batch_plan_create_grids(None, 2, bad, None, None, L.cptr(fr), 1, 0, 0.15, 0,
                                                    ctypes.byref(h))
    assert rc == L.BF_ERR_DIM
    # NULL plan
    assert lib.bf_transformer_batch_plan_run(None, 0, None, None, None) == L.BF_ERR_ARG
    assert lib.bf_transformer_batch_plan_fetch(None, None) == L.BF_ERR_ARG
    assert lib.bf_transformer_batch_plan_destroy(None) == L.BF_OK


def test_batched_setup_marshalling_matches_per_geometry():
    """Marshalled.setups (one buffer per field for a batch of geometries) gives the same structs as
    Marshalled.setup per geometry: region counts, and every array read back through its pointer;
    a rough geometry keeps its surface array, flat ones pass none"""
    coords = W.grid_coords(16)
    geoms = [W.booster(coords, ndisk=3, seed=1), W.booster(coords, ndisk=7, seed=2, flat=True),
             W.rough_booster(coords, ndisk=2, seed=3), W.booster(coords, ndisk=1, seed=4)]
    geoms[3].relative_tilt_x = 1e-4  # a scalar tilt, broadcast over the regions
    m = L.Marshalled()
    batch = m.setups(geoms, coords.ny, coords.nx)
    for i, g in enumerate(geoms):
        m1 = L.Marshalled()  # holds the arrays `one` points into
        one = m1.setup(g, coords.ny, coords.nx)
        got = batch[i]
        assert got.nreg == one.nreg
        nreg = one.nreg

        def read(p, count, dt):
            if not p:
                return None
            buf = (ctypes.c_char * (count * np.dtype(dt).itemsize)).from_address(p)
            return np.frombuffer(bytes(buf), dtype=dt)

        for field, dt, cnt in (("distance", np.float64, nreg), ("eps", np.complex128, nreg),
                               ("tilt_x", np.float64, nreg), ("tilt_y", np.float64, nreg),
                               ("r", np.complex128, len(np.asarray(g.r).reshape(-1)))):
            a, b = read(getattr(got, field), cnt, dt), read(getattr(one, field), cnt, dt)
            assert np.array_equal(a, b, equal_nan=True), (i, field)
        assert bool(got.surfaces) == bool(one.surfaces), i
        if one.surfaces:
            cnt = nreg * coords.ny * coords.nx
            assert np.array_equal(read(got.surfaces, cnt, np.complex128), read(one.surfaces, cnt, np.complex128))
    # grids per geometry (plans with a grid per mode set): each rough geometry's maps checked against its own
    small = W.grid_coords(12)
    mixed = [W.rough_booster(small, ndisk=2, seed=5), geoms[2]]
    batch = L.Marshalled().setups(mixed, np.array([12, 16]), np.array([12, 16]))
    assert batch[0].surfaces and batch[1].surfaces
    with pytest.raises(L.DimensionMismatch):
        L.Marshalled().setups(mixed, np.array([16, 16]), np.array([16, 16]))


def test_batched_marshalling_replaces_nan_in_place():
    """Marshalled.setups(nan_to_1e30=True) passes 1e30 for NaN permittivities and writes 1e30 into the
    caller's eps arrays (the reference's in-place bdry.eps[isnan] .= 1e30), leaving other entries alone"""
    coords = W.grid_coords(16)
    geoms = [W.booster(coords, ndisk=k, seed=10 + k) for k in (1, 3, 2)]
    geoms[1].eps[3] = complex(np.nan, 0.0)
    before = [g.eps.copy() for g in geoms]
    m = L.Marshalled()
    batch = m.setups(geoms, coords.ny, coords.nx, nan_to_1e30=True)
    for i, g in enumerate(geoms):
        want = before[i].copy()
        want[np.isnan(want)] = 1e30
        assert np.array_equal(g.eps, want), i
        got = np.frombuffer(bytes((ctypes.c_char * (16 * batch[i].nreg)).from_address(batch[i].eps)),
                            dtype=np.complex128)
        assert np.array_equal(got, want), i


def test_dancer_batch_validation_without_device():
    """bf_dancer_batch checks its arguments on the host: negative count, missing setups or offsets"""
    lib = L.lib()
    m, c = _coords_struct()
    sbs = [W.booster(W.grid_coords(8), ndisk=k, seed=k) for k in (1, 2)]
    setups = m.setups(sbs, 8, 8)
    f = np.array([20e9])
    out = np.zeros(200, dtype=complex)
    off = np.zeros(2, dtype=np.int64)
    args = lambda n, st, offp: (None, n, st, ctypes.byref(c), 0.0, 10, L.cptr(f), 1, 0, None, None, 0.1, 1, 1, 0,
                                offp, L.cptr(out), None, None)
    assert lib.bf_dancer_batch(*args(-1, setups, L.cptr(off))) == L.BF_ERR_ARG
    assert lib.bf_dancer_batch(*args(2, None, L.cptr(off))) == L.BF_ERR_ARG
    assert lib.bf_dancer_batch(*args(2, setups, None)) == L.BF_ERR_ARG
    # valid arguments but no context -> ArgumentError, raised only after the checks
    assert lib.bf_dancer_batch(*args(2, setups, L.cptr(off))) == L.BF_ERR_ARG
    assert b"ctx" in lib.bf_last_error(None)
