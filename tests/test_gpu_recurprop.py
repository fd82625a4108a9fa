"""GPU parity of the propagators and the recursive Fourier solvers.

FP64 path: results must match the CPU oracle (and the reference goldens) to
1e-10 relative (max |diff| / max |oracle| for fields, per-point relative for
boost factors).  Mixed precision: relative beta^2 error <= 1e-4 against FP64.
"""
import math

import numpy as np
import pytest

import BoostFractor as B
from BoostFractor import workloads as W
from helpers import (DISKR, GOLDEN, O, deps_distance, ocoords, osetup, pmap, ref_coords, ref_dx,
                     ref_setup_default, ref_setup_eps9, rel, with_distance)

pytestmark = pytest.mark.gpu

TOL = 1e-10
KINDS = [(B.propagator, O.PROP_PROPAGATOR), (B.propagatorNoTilts, O.PROP_NOTILTS),
         (B.propagator1D, O.PROP_1D), (B.propagatorMomentumSpace, O.PROP_MOMENTUM)]


def _field(coords, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((coords.ny, coords.nx)) + 1j * rng.standard_normal((coords.ny, coords.nx))


def _grid(nx, ny):
    X = np.array([((j - nx // 2) / nx) for j in range(nx)]) if nx > 1 else np.array([1e-7])
    Y = np.array([((j - ny // 2) / ny) for j in range(ny)]) if ny > 1 else np.array([1e-7])
    return B.CoordinateSystem(X, Y, B.get_kspace_coords(X), B.get_kspace_coords(Y))


@pytest.mark.parametrize("nx,ny", [(64, 64), (101, 101), (128, 32), (100, 64), (1, 1)])
@pytest.mark.parametrize("kind", range(4))
def test_propagators_match_oracle(nx, ny, kind):
    coords = _grid(nx, ny) if (nx, ny) != (101, 101) else ref_coords()
    oc = ocoords(coords)
    prop, okind = KINDS[kind]
    lam = B.wavelength(20e9)
    rng = np.random.default_rng(7)
    surf = 1e-5 * (rng.standard_normal((coords.ny, coords.nx)) + 0.1j * rng.standard_normal((coords.ny, coords.nx)))
    cases = [(complex(1.0, -0.0), 8e-3, 0.0, 0.0, None), (complex(24.0, -0.0), 1e-3, 2e-4, -1e-4, surf),
             (complex(1.5, -0.02), 5e-3, 1e-4, 3e-4, surf), (complex(1.0, -0.0), 0.0, 0.0, 0.0, None)]
    for i, (eps, dz, tx, ty, s) in enumerate(cases):
        E0 = _field(coords, 11 + i)
        got = prop(E0, dz, 0.3, eps, tx, ty, s, lam, coords)
        want = O.propagate(okind, E0, oc, dz, 0.3, eps, tx, ty, s, lam)
        assert rel(got, want) < TOL, (nx, ny, kind, i, rel(got, want))


def test_propagator_dz0_is_mask():
    """dz = 0, no tilt: the FFT round trip leaves the aperture-masked field."""
    coords = W.grid_coords(128)
    E0 = _field(coords, 3)
    got = B.propagator(E0, 0.0, 0.2, complex(1.0, -0.0), 0.0, 0.0, None, 0.015, coords)
    XX, YY = np.meshgrid(coords.X, coords.Y)
    assert np.max(np.abs(got - np.where(XX ** 2 + YY ** 2 < 0.2 ** 2, E0, 0))) < 1e-13


def test_dancer_1d_golden():
    """test/runtests.jl:214 (dancer with propagator1D, nmax=200)."""
    coords = ref_coords()
    sb = ref_setup_eps9(coords)
    dx = ref_dx(coords)
    got = []
    for d in np.arange(0, 3 * math.pi, 1.0):
        E = B.dancer(0, 200, with_distance(sb, deps_distance(d)), coords, prop=B.propagator1D, f=10e9, diskR=DISKR)
        got.append(np.sum(np.abs(E) ** 2) * dx * dx / (math.pi * DISKR ** 2))
    g = np.array(GOLDEN["dancer_1d_deps_0_1_3pi"]["beta2"])
    assert np.max(np.abs(np.array(got) / g - 1)) < TOL


def test_dancer_3d_golden():
    """test/runtests.jl:215 (3D dancer, nmax=200, Deps = 0:0.1:3pi), a spread of points."""
    coords = ref_coords()
    sb = ref_setup_eps9(coords)
    dx = ref_dx(coords)
    Deps = np.arange(0, 3 * math.pi, 0.1)
    g = np.array(GOLDEN["dancer_3d_deps_0_0.1_3pi"]["beta2"])
    for i in [0, 17, 47, 80, 94]:
        E = B.dancer(0, 200, with_distance(sb, deps_distance(Deps[i])), coords, f=10e9, diskR=DISKR)
        b = np.sum(np.abs(E) ** 2) * dx * dx / (math.pi * DISKR ** 2)
        assert abs(b / g[i] - 1) < TOL, (i, b, g[i])


def test_cheerleader_golden():
    """test/runtests.jl:211 (3D cheerleader, nmax=100), all 10 points."""
    coords = ref_coords()
    sb = ref_setup_default(coords)
    dx = ref_dx(coords)
    got = []
    for d in np.arange(0, 3 * math.pi, 1.0):
        E = B.cheerleader(0, 100, with_distance(sb, deps_distance(d)), coords, prop=B.propagator, f=10e9,
                          diskR=DISKR)
        got.append(np.sum(np.abs(E) ** 2) * dx * dx / (math.pi * DISKR ** 2))
    g = np.array(GOLDEN["cheerleader_3d_deps_0_1_3pi"]["beta2"])
    assert np.max(np.abs(np.array(got) / g - 1)) < TOL


def test_cheerleader_fine_golden():
    """test/runtests.jl:212 (commented out; Deps = 0:0.1:3pi), a spread of points
    that are not on the :211 grid."""
    coords = ref_coords()
    sb = ref_setup_default(coords)
    dx = ref_dx(coords)
    Deps = np.arange(0, 3 * math.pi, 0.1)
    g = np.array(GOLDEN["cheerleader_3d_deps_0_0.1_3pi"]["beta2"])
    for i in [3, 17, 26, 48, 63, 79, 94]:
        E = B.cheerleader(0, 100, with_distance(sb, deps_distance(Deps[i])), coords, prop=B.propagator, f=10e9,
                          diskR=DISKR)
        b = np.sum(np.abs(E) ** 2) * dx * dx / (math.pi * DISKR ** 2)
        assert abs(b / g[i] - 1) < TOL, (i, b, g[i])


def _small_booster(N=64, ndisk=3, seed=1001, tilt=True, span=1.0):
    coords = W.grid_coords(N, span)
    sb = W.booster(coords, ndisk=ndisk, seed=seed)
    if tilt:
        rng = W.SplitMix64(seed + 7)
        n = len(sb.distance)
        sb.relative_tilt_x = np.array([2e-4 * rng.normal() for _ in range(n)])
        sb.relative_tilt_y = np.array([2e-4 * rng.normal() for _ in range(n)])
        surf = np.zeros((n, N, N), dtype=complex)
        XX, YY = np.meshgrid(coords.X, coords.Y)
        for i in range(2, n - 1, 2):
            surf[i] = 5e-6 * np.cos(20 * XX + i) * np.sin(13 * YY)
        sb.relative_surfaces = surf
    return coords, sb


@pytest.mark.parametrize("prop,okind", [(B.propagator, O.PROP_PROPAGATOR), (B.propagatorNoTilts, O.PROP_NOTILTS),
                                        (B.propagator1D, O.PROP_1D),
                                        (B.propagatorMomentumSpace, O.PROP_MOMENTUM)])
def test_dancer_batched_matches_oracle(prop, okind):
    coords, sb = _small_booster()
    freqs = [19.7e9, 20.0e9, 20.3e9]
    got = B.dancer(0, 25, sb, coords, f=freqs, prop=prop, diskR=DISKR)
    os_, oc = osetup(sb), ocoords(coords)
    want = pmap(lambda f: O.dancer(os_, oc, 0, 25, f, okind, diskR=DISKR), freqs)
    for q in range(len(freqs)):
        assert rel(got[q], want[q]) < TOL, (q, rel(got[q], want[q]))


def test_dancer_history_and_reverse_sum():
    coords, sb = _small_booster(N=32, ndisk=2, tilt=False)
    os_, oc = osetup(sb), ocoords(coords)
    hist = B.dancer(0, 12, sb, coords, f=20e9, diskR=DISKR, returnsum=False, immediatesum=False)
    want = O.dancer(os_, oc, 0, 12, 20e9, diskR=DISKR, returnsum=False, immediatesum=False)
    assert hist.shape == (13, 32, 32)
    assert rel(hist, want) < TOL
    summed = B.dancer(0, 12, sb, coords, f=20e9, diskR=DISKR, returnsum=True, immediatesum=False)
    assert rel(summed, O.dancer(os_, oc, 0, 12, 20e9, diskR=DISKR, returnsum=True, immediatesum=False)) < TOL
    one = B.dancer(0, 12, sb, coords, f=20e9, diskR=DISKR, returnsum=False, immediatesum=True)
    assert one.shape == (1, 32, 32)
    assert rel(one[0], summed) < 1e-12


def test_dancer_amin_early_stop():
    coords, sb = _small_booster(N=32, ndisk=2, tilt=False)
    os_, oc = osetup(sb), ocoords(coords)
    for amin in [0.3, 0.05, 2.0]:
        got = B.dancer(amin, 40, sb, coords, f=20e9, diskR=DISKR)
        want = O.dancer(os_, oc, amin, 40, 20e9, diskR=DISKR)
        assert rel(got, want) < TOL if np.any(want) else np.all(got == 0), amin


def test_dancer_emit_and_reflect():
    coords, sb = _small_booster(N=32, ndisk=2, tilt=False)
    os_, oc = osetup(sb), ocoords(coords)
    R = len(sb.distance) - 1
    rng = np.random.default_rng(5)
    emit = rng.standard_normal((R, 2, 32, 32)) + 1j * rng.standard_normal((R, 2, 32, 32))
    got = B.dancer(0, 10, sb, coords, f=20e9, emit=emit, diskR=DISKR)
    assert rel(got, O.dancer(os_, oc, 0, 10, 20e9, emit=emit, diskR=DISKR)) < TOL
    refl = rng.standard_normal((32, 32)) + 0j
    got = B.dancer(0, 10, sb, coords, f=20e9, reflect=refl, emit=emit, diskR=DISKR)
    assert rel(got, O.dancer(os_, oc, 0, 10, 20e9, reflect=refl, diskR=DISKR)) < TOL


@pytest.mark.parametrize("prop,okind", [(B.propagator, O.PROP_PROPAGATOR), (B.propagator1D, O.PROP_1D),
                                        (B.propagatorMomentumSpace, O.PROP_MOMENTUM)])
def test_cheerleader_matches_oracle(prop, okind):
    coords, sb = _small_booster(ndisk=2)
    freqs = [19.9e9, 20.2e9]
    left, right = B.cheerleader(0, 15, sb, coords, f=freqs, prop=prop, diskR=DISKR, returnboth=True)
    os_, oc = osetup(sb), ocoords(coords)
    for q, f in enumerate(freqs):
        wl, wr = O.cheerleader(os_, oc, 0, 15, f, okind, diskR=DISKR, returnboth=True)
        assert rel(right[q], wr) < TOL and rel(left[q], wl) < TOL, q


def test_cheerleader_amin_more_than_64_frequencies():
    """amin with 70 frequencies in one batch: the stop criterion covers every
    frequency slot, including those at index 64 and above, against the oracle;
    the criterion really stops those points early (their result differs from
    the run to nmax)"""
    coords, sb = _small_booster(N=32, ndisk=2, tilt=False)
    os_, oc = osetup(sb), ocoords(coords)
    freqs = np.linspace(19.5e9, 20.5e9, 70)
    got = B.cheerleader(1.5, 40, sb, coords, f=freqs, diskR=DISKR)
    full = B.cheerleader(0, 40, sb, coords, f=freqs[64:], diskR=DISKR)
    for q in (0, 63, 64, 66, 69):
        want = O.cheerleader(os_, oc, 1.5, 40, freqs[q], diskR=DISKR)
        assert rel(got[q], want) < TOL, (q, rel(got[q], want))
        if q >= 64:
            assert rel(got[q], full[q - 64]) > 0.1, q  # stopped early (the oracle stops these near 0.7 rel)


def test_cheerleader_amin_reflect():
    coords, sb = _small_booster(N=32, ndisk=2, tilt=False)
    os_, oc = osetup(sb), ocoords(coords)
    got = B.cheerleader(0.05, 40, sb, coords, f=20e9, diskR=DISKR)
    assert rel(got, O.cheerleader(os_, oc, 0.05, 40, 20e9, diskR=DISKR)) < TOL
    refl = np.ones((32, 32), dtype=complex)
    got = B.cheerleader(0, 10, sb, coords, f=20e9, reflect=refl, diskR=DISKR)
    assert rel(got, O.cheerleader(os_, oc, 0, 10, 20e9, reflect=refl, diskR=DISKR)) < TOL


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("N,ndisk,nmax", [(64, 3, 20), (128, 2, 9), (64, 1, 7), (64, 2, 0)])
def test_cheerleader_wavefront_matches_sweep(prec, N, ndisk, nmax):
    """Steps of several iterations sharing launches (wavefront order) against the
    step-by-step sweep: the same kernels on the same fields in an order that
    keeps every dependency, so both outputs agree bit for bit (Nd = 4 and
    nmax = 0 included)."""
    coords, sb = _small_booster(N=N, ndisk=ndisk, seed=1011 + N + ndisk)
    out = {}
    for wz in ("0", "1"):
        with B.context().options(chwave=int(wz)):
            out[wz] = B.cheerleader(0, nmax, sb, coords, f=[19.9e9, 20.05e9, 20.2e9], diskR=DISKR, precision=prec,
                                    returnboth=True)
    assert np.any(out["1"][1] != 0)
    assert np.array_equal(out["1"][0], out["0"][0]) and np.array_equal(out["1"][1], out["0"][1])


def test_cheerleader_wavefront_several_batches():
    """260 frequencies run as two batches (256 + 4): the wavefront lists of the
    smaller last batch, against the step-by-step sweep bit for bit and two
    points (one per batch) against the oracle."""
    coords, sb = _small_booster(N=32, ndisk=1, seed=1013)
    freqs = np.linspace(19.5e9, 20.5e9, 260)
    out = {}
    for wz in ("0", "1"):
        with B.context().options(chwave=int(wz)):
            out[wz] = B.cheerleader(0, 6, sb, coords, f=freqs, diskR=DISKR)
    assert np.array_equal(out["1"], out["0"])
    os_, oc = osetup(sb), ocoords(coords)
    for q in (3, 258):
        want = O.cheerleader(os_, oc, 0, 6, freqs[q], diskR=DISKR)
        assert rel(out["1"][q], want) < TOL, (q, rel(out["1"][q], want))


@pytest.mark.parametrize("prec,tol", [("fp64", TOL), ("mixed", 1e-5)])
def test_cheerleader_emit_matches_oracle(prec, tol):
    """User-supplied emission fields (Nd x 2 x grid) instead of dance_intro, on a
    power-of-two grid where the propagations run on the aperture band and the
    pair steps store the next masked row transform; both outputs."""
    coords, sb = _small_booster(N=64, ndisk=2, seed=1010)
    Nd = len(sb.distance)
    rng = np.random.default_rng(11)
    emit = rng.standard_normal((Nd, 2, 64, 64)) + 1j * rng.standard_normal((Nd, 2, 64, 64))
    left, right = B.cheerleader(0, 12, sb, coords, f=20.1e9, emit=emit, diskR=DISKR, precision=prec,
                                returnboth=True)
    wl, wr = O.cheerleader(osetup(sb), ocoords(coords), 0, 12, 20.1e9, emit=emit, diskR=DISKR, returnboth=True)
    assert np.any(wr != 0) and np.any(wl != 0)
    assert rel(right, wr) < tol and rel(left, wl) < tol, (rel(right, wr), rel(left, wl))


def test_dancer_mixed_precision_beta2():
    """20-disk booster, 128^2, FP32 fields/FFT: beta^2 within 1e-4 of the FP64 engine."""
    coords = W.grid_coords(128)
    sb = W.booster(coords, ndisk=20, seed=1000)
    freqs = np.linspace(19.5e9, 20.5e9, 6)
    e64 = B.dancer(0, 300, sb, coords, f=freqs, diskR=DISKR)
    e32 = B.dancer(0, 300, sb, coords, f=freqs, diskR=DISKR, precision="mixed")
    b64 = W.beta2_dancer(e64, coords, DISKR)
    b32 = W.beta2_dancer(e32, coords, DISKR)
    assert np.max(np.abs(b32 / b64 - 1)) < 1e-4, np.abs(b32 / b64 - 1)


def test_dance_intro_matches_oracle():
    coords = ref_coords()
    sb = ref_setup_eps9(coords)
    # untrimmed setup: the NaN mirror permittivity propagates exactly as in the reference
    got = B.dance_intro(sb, coords, diskR=DISKR)
    want = O.dance_intro(sb.eps, sb.r[0], ocoords(coords), DISKR)
    np.testing.assert_allclose(got, want, rtol=1e-14, atol=0, equal_nan=True)
    # the trimmed view dancer() uses (src/boost_fractor_recurprop.jl:28-34)
    t = sb.copy()
    t.distance, t.eps = sb.distance[1:], sb.eps[1:]
    got = B.dance_intro(t, coords, diskR=DISKR)
    want = O.dance_intro(sb.eps[1:], sb.r[0], ocoords(coords), DISKR)
    assert np.all(np.isfinite(got)) and rel(got, want) < 1e-14


def test_dancer_plan_matches_one_shot():
    coords = W.grid_coords(64)
    sb = W.booster(coords, ndisk=4, seed=1003)
    freqs = np.linspace(19.5e9, 20.5e9, 5)
    E1 = B.dancer(0, 30, sb, coords, f=freqs, diskR=DISKR)
    plan = B.DancerPlan(0, 30, sb, coords, freqs, diskR=DISKR, freq_batch=2)
    plan.run()
    plan.run()  # idempotent: every run restarts from the axion source
    E2, power = plan.fetch()
    assert rel(E2, E1) < 1e-13
    assert np.allclose(power, np.sum(np.abs(E1) ** 2, axis=(1, 2)), rtol=1e-12)
    plan.close()


@pytest.mark.parametrize("N,prec", [(64, "fp64"), (64, "mixed"), (64, "fp32"), (48, "fp64"), (40, "mixed")])
def test_dancer_row_band_is_exact(N, prec):
    """Skipping the rows the aperture zeroes changes nothing: the fused bounce with
    the row band equals the bounce over every row, bit for bit (register and LDS
    FFT paths, all precisions, summed and history outputs)."""
    coords = W.grid_coords(N)
    sb = W.booster(coords, ndisk=3, seed=1002)
    freqs = [19.8e9, 20.1e9]
    out = {}
    for band in ("0", "1"):
        with B.context().options(band=int(band)):
            e = B.dancer(0, 20, sb, coords, f=freqs, diskR=DISKR, precision=prec)
            h = B.dancer(0, 9, sb, coords, f=20e9, diskR=DISKR, precision=prec, returnsum=False, immediatesum=False)
        out[band] = (e, h)
    assert np.any(out["1"][0] != 0)
    # the band really is narrower than the grid here
    XX, YY = np.meshgrid(coords.X, coords.Y)
    assert np.count_nonzero(np.any(XX ** 2 + YY ** 2 < DISKR ** 2, axis=1)) < N // 2
    assert np.array_equal(out["0"][0], out["1"][0])
    assert np.array_equal(out["0"][1], out["1"][1])


# (N, span): the aperture's rows and columns fall in elements [5, 10] of a
# thread's 16 (N/16 apart) on a 1 m grid, in [3, 12] on 0.5 m and in [4, 11]
# on 0.6 m, so the column and row passes run their band-pruned transforms.
@pytest.mark.parametrize("N,span,prec", [(128, 1.0, "fp64"), (128, 1.0, "mixed"), (128, 1.0, "fp32"),
                                         (128, 0.5, "fp64"), (128, 0.6, "mixed"), (256, 1.0, "mixed"),
                                         (256, 1.0, "fp64")])
def test_dancer_pruned_transforms_match_unpruned(N, span, prec):
    """Band-pruned register transforms (butterflies of zero elements skipped,
    elements that are not stored not formed) against the unpruned ones: summed and
    history outputs, tilted and rough disks (tables applied in the row pass) and
    the field maximum of the amin criterion.  The arithmetic is the same up to FMA
    contraction (a product fused where the pruned code uses it once), so the
    results agree to rounding: 1e-12 of the field maximum in FP64, 1e-5 with
    complex64 storage (mixed), 1e-4 in FP32 arithmetic."""
    coords, sb = _small_booster(N=N, ndisk=3, seed=1003, span=span)
    freqs = [19.8e9, 20.1e9]
    out = {}
    for prune in ("0", "1"):
        with B.context().options(prune=int(prune)):
            e = B.dancer(0, 20, sb, coords, f=freqs, diskR=DISKR, precision=prec)
            h = B.dancer(0, 6, sb, coords, f=20e9, diskR=DISKR, precision=prec, returnsum=False, immediatesum=False)
            a = B.dancer(0.05, 40, sb, coords, f=20e9, diskR=DISKR, precision=prec)
        out[prune] = (e, h, a)
    tol = {"fp64": 1e-12, "mixed": 1e-5, "fp32": 1e-4}[prec]
    assert np.any(out["1"][0] != 0)
    for q in range(3):
        assert rel(out["1"][q], out["0"][q]) < tol, (q, rel(out["1"][q], out["0"][q]))


@pytest.mark.parametrize("N", [128, 256])
def test_dancer_pruned_matches_oracle(N):
    """The pruned path (tilted and rough disks; 256 x 256 is the benchmark's grid)
    against the oracle."""
    coords, sb = _small_booster(N=N, ndisk=2, seed=1004)
    freqs = [19.9e9, 20.2e9]
    got = B.dancer(0, 15, sb, coords, f=freqs, diskR=DISKR)
    os_, oc = osetup(sb), ocoords(coords)
    want = pmap(lambda f: O.dancer(os_, oc, 0, 15, f, O.PROP_PROPAGATOR, diskR=DISKR), freqs)
    for q in range(len(freqs)):
        assert rel(got[q], want[q]) < TOL, (q, rel(got[q], want[q]))


@pytest.mark.parametrize("N,prec", [(64, "mixed"), (128, "mixed"), (128, "fp32"), (256, "mixed")])
def test_dancer_quarter_tables_match_full(N, prec):
    """Complex64-field paths read quarter k-space tables ([N/2+1]^2 entries, a
    mirrored bin reads its mirror's phase, equal up to the rounding of the k
    range): against full tables within the storage rounding (1e-5 of the field
    maximum with FP64 arithmetic, 1e-4 in FP32), tilted and rough disks."""
    coords, sb = _small_booster(N=N, ndisk=3, seed=1005)
    freqs = [19.8e9, 20.1e9]
    out = {}
    for q in ("0", "1"):
        with B.context().options(qtab=int(q)):
            out[q] = B.dancer(0, 20, sb, coords, f=freqs, diskR=DISKR, precision=prec)
    assert np.any(out["1"] != 0)
    tol = {"mixed": 1e-5, "fp32": 1e-4}[prec]
    assert rel(out["1"], out["0"]) < tol, rel(out["1"], out["0"])


@pytest.mark.parametrize("N,prec,amin", [(64, "fp64", 0.0), (128, "mixed", 0.0), (64, "mixed", 0.05)])
def test_dancer_plan_graph_replays_match_eager(N, prec, amin):
    """Plans replay each frequency batch's bounce loop from a HIP graph captured by
    the first run: three runs in a row, with eager work and synchronisations
    between them, each equal bit for bit to the plan launched bounce by bounce;
    5 frequencies in batches of 2 (three graphs, the last batch partial); the
    instrumented (timing) run takes the launch-by-launch path and agrees too."""
    coords, sb = _small_booster(N=N, ndisk=3, seed=1006)
    freqs = [19.7e9, 19.9e9, 20.0e9, 20.2e9, 20.4e9]
    with B.context().options(graph=0):
        eager = B.DancerPlan(amin, 30, sb, coords, freqs, diskR=DISKR, precision=prec, freq_batch=2)
        eager.run()
        want_E, want_p = eager.fetch()
        eager.close()
    assert B.context().get_option("graph") == 1
    plan = B.DancerPlan(amin, 30, sb, coords, freqs, diskR=DISKR, precision=prec, freq_batch=2)
    for rep in range(3):
        plan.run()
        E, p = plan.fetch()
        assert np.any(E != 0)
        assert np.array_equal(E, want_E), rep
        assert np.array_equal(p, want_p), rep
        B.dancer(0, 4, sb, coords, f=20e9, diskR=DISKR)  # eager work on the same stream in between
        B.context().synchronize()
    plan.timing(True)
    plan.run()
    plan.timing(False)
    E, p = plan.fetch()
    assert np.array_equal(E, want_E)
    plan.close()


def test_cheerleader_band_matches_oracle():
    """The cheerleader's propagations on the aperture band (masked row FFT and
    column loads on band rows only, band-pruned column transforms at 128 x 128),
    tilted and rough disks, both outputs, against the oracle."""
    coords, sb = _small_booster(N=128, ndisk=2, seed=1007)
    freqs = [19.9e9, 20.2e9]
    left, right = B.cheerleader(0, 12, sb, coords, f=freqs, diskR=DISKR, returnboth=True)
    os_, oc = osetup(sb), ocoords(coords)
    for q, f in enumerate(freqs):
        wl, wr = O.cheerleader(os_, oc, 0, 12, f, O.PROP_PROPAGATOR, diskR=DISKR, returnboth=True)
        assert rel(right[q], wr) < TOL and rel(left[q], wl) < TOL, (q, rel(right[q], wr), rel(left[q], wl))


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cheerleader_band_and_quarter_tables_match_full(prec):
    """Band-limited propagations (and, with complex64 fields, quarter phase
    tables) against every row and full tables: 1e-12 of the field maximum in
    FP64, 1e-5 with complex64 fields."""
    coords, sb = _small_booster(N=128, ndisk=3, seed=1008)
    out = {}
    for on in ("0", "1"):
        with B.context().options(band=int(on), qtab=int(on)):
            out[on] = B.cheerleader(0, 20, sb, coords, f=[19.8e9, 20.1e9], diskR=DISKR, precision=prec)
    assert np.any(out["1"] != 0)
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert rel(out["1"], out["0"]) < tol, rel(out["1"], out["0"])


@pytest.mark.parametrize("N", [64, 128])
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_cheerleader_fused_pair_rows_match_separate(prec, N):
    """The pair steps' inverse row pass fused with the boundary update against the
    separate passes (tilted and rough disks, both outputs, amin criterion; with
    amin = 0 the fused pair steps keep the aperture band's rows only and store
    the next propagation's masked row transform, option chxf=0 stores real space):
    1e-12 of the field maximum in FP64, 1e-5 with complex64 fields."""
    coords, sb = _small_booster(N=N, ndisk=3, seed=1009)
    out = {}
    for fz, xf in (("0", "1"), ("1", "0"), ("1", "1")):
        with B.context().options(chfuse=int(fz), chxf=int(xf)):
            lr = B.cheerleader(0, 20, sb, coords, f=[19.8e9, 20.1e9], diskR=DISKR, precision=prec, returnboth=True)
            am = B.cheerleader(0.05, 40, sb, coords, f=20e9, diskR=DISKR, precision=prec)
        out[fz + xf] = (lr[0], lr[1], am)
    tol = 1e-12 if prec == "fp64" else 1e-5
    assert np.any(out["11"][1] != 0) and np.any(out["11"][2] != 0)  # (the mirror side's output is zero)
    for key in ("10", "11"):
        for q in range(3):
            assert rel(out[key][q], out["01"][q]) < tol, (key, q, rel(out[key][q], out["01"][q]))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dancer_batch_matches_per_geometry(precision):
    """dancer_batch over boosters of 1, 3 and 2 disks (uneven region counts) at the same frequencies equals
    dancer per geometry bit for bit (summed output, the whole history, and a reflect case), and the
    oracle at 1e-10 on one geometry (FP64)"""
    coords = W.grid_coords(64)
    sbs = [W.booster(coords, ndisk=k, seed=1200 + k) for k in (1, 3, 2)]
    freqs = [19.7e9, 20.3e9]
    got = B.dancer_batch(0, 30, sbs, coords, f=freqs, diskR=0.15, precision=precision)
    for sb, g in zip(sbs, got):
        want = B.dancer(0, 30, sb, coords, f=freqs, diskR=0.15, precision=precision)
        assert np.array_equal(g, want)
    hist = B.dancer_batch(0, 12, sbs, coords, f=freqs, diskR=0.15, returnsum=False, immediatesum=False,
                          precision=precision)
    for sb, g in zip(sbs, hist):
        want = B.dancer(0, 12, sb, coords, f=freqs, diskR=0.15, returnsum=False, immediatesum=False,
                        precision=precision)
        assert g.shape == want.shape and np.array_equal(g, want)
    refl = [None, np.exp(1j * np.linspace(0, 1, 64 * 64)).reshape(64, 64), None]
    rb = B.dancer_batch(0, 20, sbs, coords, f=freqs, diskR=0.15, reflect=refl, precision=precision)
    for sb, r, g in zip(sbs, refl, rb):
        want = B.dancer(0, 20, sb, coords, f=freqs, diskR=0.15, reflect=r, precision=precision)
        assert np.array_equal(g, want)
    if precision == "fp64":
        oc = O.Coords(coords.X, coords.Y, coords.kX, coords.kY)
        os_ = O.Setup(sbs[1].distance, sbs[1].eps, sbs[1].r)
        w = O.dancer(os_, oc, 0, 30, freqs[0], O.PROP_PROPAGATOR, diskR=0.15)
        assert np.max(np.abs(got[1][0] - w)) / np.max(np.abs(w)) < TOL
