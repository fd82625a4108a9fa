"""GPU parity of the mode-matching transformer (src/boost_fractor_transformer.jl).

FP64 path at 1e-10 relative against the reference goldens (test/runtests.jl,
examples/Beam Shapes.ipynb) and the CPU oracle on richer cases (several
modes and azimuthal orders, tilts, rough surfaces, all propagator kinds,
reflectivity, precomputed matrices, trace-back).
"""
import math

import numpy as np
import pytest

import BoostFractor as B
from BoostFractor import workloads as W
from helpers import (DISKR, GOLDEN, O, chain_sensitivity, deps_distance, ocoords, pmap, ref_coords, ref_setup_eps9,
                     rel, with_distance)

pytestmark = pytest.mark.gpu
TOL = 1e-10


def gauss_profile(coords, z=1e-9, omega0=0.1, f=10e9):
    """input recipe of src/beam_gaussian.jl:24-37 (numpy), used to build golden inputs"""
    XX, YY = np.meshgrid(coords.X, coords.Y)
    radius = np.sqrt(XX ** 2 + YY ** 2)
    lam = B.wavelength(f)
    k = 2 * math.pi / lam
    zR = math.pi * omega0 ** 2 / lam
    om = omega0 * math.sqrt(1 + (z / zR) ** 2)
    Rz = z * (1 + (zR / z) ** 2)
    psi = math.atan(z / zR)
    return omega0 / om * np.exp(-radius ** 2 / om ** 2) * np.exp(-1j * (k * z + k * radius ** 2 / (2 * Rz) - psi))


def test_transformer_golden():
    """test/runtests.jl:217: 3D transformer, M=10, L=0, Deps = 0:1:3pi."""
    coords = ref_coords()
    sb = ref_setup_eps9(coords)
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=10, Lmax=0, diskR=DISKR)
    got = []
    for d in np.arange(0, 3 * math.pi, 1.0):
        s = with_distance(sb, deps_distance(d))
        boost = B.transformer(s, coords, modes, prop=B.propagator, f=10e9, diskR=DISKR)
        got.append(np.sum(np.abs(boost) ** 2))
    g = np.array(GOLDEN["transformer_3d_deps_0_1_3pi"]["beta2"])
    assert np.max(np.abs(np.array(got) / g - 1)) < TOL


def test_transformer_golden_batched_geometry_and_frequency():
    """the same golden through the batched frequency axis (one geometry, f = [10 GHz] * 2)."""
    coords = ref_coords()
    sb = with_distance(ref_setup_eps9(coords), deps_distance(2.0))
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=10, Lmax=0, diskR=DISKR)
    boost = B.transformer(sb, coords, modes, f=[10e9, 10e9], diskR=DISKR)
    g = GOLDEN["transformer_3d_deps_0_1_3pi"]["beta2"][2]
    assert np.all(np.abs(np.sum(np.abs(boost) ** 2, axis=1) / g - 1) < TOL)


def test_modes_golden():
    """test/runtests.jl:166-171"""
    coords = ref_coords()
    sb = ref_setup_eps9(coords)
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=10, Lmax=0, diskR=DISKR)
    coeffs = B.field2modes(None, coords, modes, diskR=DISKR)
    field = B.modes2field(coeffs, coords, modes)
    g = GOLDEN["modes2field_field5000"]
    assert abs(field.ravel()[4999] - (g["re"] + 1j * g["im"])) < TOL * abs(g["re"])
    sb.eps[0] = 1e30  # state after the transformer sweep of the reference test (it mutates bdry.eps)
    propmat = B.calc_propagation_matrices(sb, coords, modes)
    g = GOLDEN["propmat1_4_6"]
    # a Gram entry of nearly orthogonal patterns: absolute agreement relative to the unit diagonal
    assert abs(propmat[0][3, 5] - (g["re"] + 1j * g["im"])) < 1e-14
    assert abs(propmat[0][3, 3] - 1.0) < 1e-13


def test_trace_back_golden():
    """test/runtests.jl:174-207,221 (1D, with the Deps=9 geometry left in the global by the loop)."""
    c1 = B.SeedCoordinateSystem(X=[1e-7], Y=[1e-7])
    sb = ref_setup_eps9(c1)
    modes = B.SeedModes(c1, ThreeDim=False, Mmax=1, Lmax=0, diskR=DISKR)
    res = []
    for d in np.arange(0, 3 * math.pi, 1.0):
        sb.distance = deps_distance(d)
        res.append(B.transformer(sb, c1, modes, prop=B.propagator1D, f=10e9, diskR=DISKR))
    full = B.transformer_trace_back(res[4], 0, sb, c1, modes, prop=B.propagator1D, f=10e9)
    g = np.array(GOLDEN["trace_back_full_fields"]["re"]) + 1j * np.array(GOLDEN["trace_back_full_fields"]["im"])
    assert np.max(np.abs(full[:, :, 0] - g)) < 1e-12


def test_beam_shapes_goldens():
    """examples/Beam Shapes.ipynb: 128^2 grid (dx=0.007), uniform beam (M=7) and Gaussian beam (M=4)."""
    coords = B.SeedCoordinateSystem(X=(-0.4445, 0.007, 0.4445), Y=(-0.4445, 0.007, 0.4445))
    modes7 = B.SeedModes(coords, ThreeDim=True, Mmax=7, Lmax=0, diskR=0.15)
    XX, YY = np.meshgrid(coords.X, coords.Y)
    uniform = np.where(XX ** 2 + YY ** 2 > 0.15 ** 2, 0.0, 1.0)
    m = B.field2modes(uniform, coords, modes7, diskR=0.15)
    g = np.array(GOLDEN["beam_shapes_uniform_field2modes_M7"]["re"])
    assert np.max(np.abs(m - g)) < TOL
    modes4 = B.SeedModes(coords, ThreeDim=True, Mmax=4, Lmax=0, diskR=0.15)
    gs = gauss_profile(coords, z=1e-9, omega0=0.097, f=22e9)
    m = B.field2modes(gs, coords, modes4)
    g = np.array(GOLDEN["beam_shapes_gauss_field2modes_M4"]["re"]) + \
        1j * np.array(GOLDEN["beam_shapes_gauss_field2modes_M4"]["im"])
    assert np.max(np.abs(m - g)) < 1e-9 * np.max(np.abs(g))


def test_seed_modes_match_oracle():
    coords = W.grid_coords(64)
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=3, Lmax=2, diskR=0.15)
    pat, kt = O.seed_modes(ocoords(coords), 3, 2, 0.15)
    assert rel(modes.mode_kt.reshape(-1), kt) < 1e-14
    assert rel(modes.patterns_flat()[:, 0], pat) < 1e-12


def _lg_like_modes(coords, M, Lm, seed=3, kspace=False):
    """user patterns through pattern_input/kT_arr (smooth, non-orthogonal, complex).
    kspace=True gives their fftshifted spectra: the momentum-space propagator
    expects its input in that form (a real-space pattern fed to it lands at the
    grid edge after the inverse FFT and is masked away to roundoff)."""
    rng = np.random.default_rng(seed)
    XX, YY = np.meshgrid(coords.X, coords.Y)
    r2 = XX ** 2 + YY ** 2
    phi = np.arctan2(YY, XX)
    pat = np.zeros((M, 2 * Lm + 1, 1, coords.ny, coords.nx), dtype=complex)
    kt = np.zeros((M, 2 * Lm + 1), dtype=complex)
    for m in range(M):
        for li, l in enumerate(range(-Lm, Lm + 1)):
            w = 0.08 + 0.02 * m
            p = (np.sqrt(r2) / w) ** abs(l) * np.exp(-r2 / w ** 2) * np.exp(-1j * l * phi) * (1 + 0.1 * m * r2 / w ** 2)
            p *= np.exp(1j * rng.uniform(0, 0.1))
            if kspace:
                p = np.fft.fftshift(np.fft.fft2(p))
            pat[m, li, 0] = p / np.sqrt(np.sum(np.abs(p) ** 2))
            kt[m, li] = 2.0 * math.sqrt(2 * m + abs(l) + 1) / w
    return B.SeedModes(coords, ThreeDim=True, Mmax=M, Lmax=Lm, pattern_input=pat, kT_arr=kt)


def _rough_booster(N=48, ndisk=2, seed=1100):
    coords = W.grid_coords(N)
    sb = W.booster(coords, ndisk=ndisk, seed=seed)
    n = len(sb.distance)
    rng = W.SplitMix64(seed + 1)
    sb.relative_tilt_x = np.array([3e-4 * rng.normal() for _ in range(n)])
    sb.relative_tilt_y = np.array([3e-4 * rng.normal() for _ in range(n)])
    XX, YY = np.meshgrid(coords.X, coords.Y)
    surf = np.zeros((n, N, N), dtype=complex)
    for i in range(1, n - 1):
        surf[i] = 4e-6 * np.cos(17 * XX + i) * np.sin(11 * YY - i)
    sb.relative_surfaces = surf
    return coords, sb


def _flat_booster(N=48, ndisk=2, seed=1100):
    coords = W.grid_coords(N)
    return coords, W.booster(coords, ndisk=ndisk, seed=seed)


def _otransformer(sb, coords, modes, f, okind, emit, reflect=None):
    surf = sb.relative_surfaces if np.any(sb.relative_surfaces != 0) else None
    eps = sb.eps.copy()
    eps[np.isnan(eps.real)] = 1e30
    os_ = O.Setup(sb.distance, eps, None, sb.relative_tilt_x, sb.relative_tilt_y, surf)
    return O.transformer(os_, ocoords(coords), modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1), f, okind,
                         diskR=DISKR, emit=emit, reflect=reflect)


def _oracle_mats(sb, coords, modes, f, okind, eps=None):
    eps = sb.eps if eps is None else eps
    lam = B.wavelength(f)
    pat, kt = modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1)
    return [O.propagation_matrix(sb.distance[r], DISKR, eps[r], sb.relative_tilt_x[r], sb.relative_tilt_y[r],
                                 sb.relative_surfaces[r], lam, ocoords(coords), pat, kt, okind)
            for r in range(len(sb.distance))]


KINDS = [(B.propagator, O.PROP_PROPAGATOR), (B.propagatorNoTilts, O.PROP_NOTILTS), (B.propagator1D, O.PROP_1D),
         (B.propagatorMomentumSpace, O.PROP_MOMENTUM)]


@pytest.mark.parametrize("builder", [_flat_booster, _rough_booster])
@pytest.mark.parametrize("prop,okind", KINDS)
def test_propagation_matrices_per_kind(builder, prop, okind):
    """every region's matrix (no inversion involved) at 1e-10"""
    coords, sb = builder()
    modes = _lg_like_modes(coords, 2, 1, kspace=(okind == O.PROP_MOMENTUM))
    got = B.calc_propagation_matrices(sb, coords, modes, f=[19.8e9, 20.25e9], prop=prop, diskR=DISKR)
    for q, f in enumerate([19.8e9, 20.25e9]):
        want = _oracle_mats(sb, coords, modes, f, okind)
        for r in range(1, len(sb.distance)):
            assert rel(got[q][r], want[r]) < TOL, (q, r, rel(got[q][r], want[r]))


@pytest.mark.parametrize("builder", [_flat_booster, _rough_booster])
@pytest.mark.parametrize("prop,okind", KINDS)
def test_transformer_matches_oracle(builder, prop, okind):
    """end to end, with the tolerance scaled by the problem's own FP64 sensitivity
    (these synthetic multimode boosters have cond(T22) up to ~1e7)"""
    coords, sb = builder()
    modes = _lg_like_modes(coords, 2, 1, kspace=(okind == O.PROP_MOMENTUM))
    emit = B.axion_induced_modes(coords, modes, diskR=DISKR)
    refl_in = np.linspace(1, 2, modes.n) + 0.3j
    freqs = [19.8e9, 20.25e9]
    boost, refl = B.transformer(sb.copy(), coords, modes, f=freqs, prop=prop, diskR=DISKR, emit=emit,
                                reflect=refl_in)
    want = pmap(lambda f: _otransformer(sb, coords, modes, f, okind, emit, refl_in), freqs)
    eps = sb.eps.copy()
    eps[np.isnan(eps.real)] = 1e30
    for q, f in enumerate(freqs):
        sens = chain_sensitivity(_oracle_mats(sb, coords, modes, f, okind, eps), eps, emit)
        tol = max(TOL, 10 * sens)
        assert tol < 1e-7
        assert rel(boost[q], want[q][0]) < tol, (q, rel(boost[q], want[q][0]), sens)
        assert rel(refl[q], want[q][1]) < tol, (q, rel(refl[q], want[q][1]), sens)


def test_transformer_many_modes_matches_oracle():
    """45 orthonormal Bessel modes (M=5, L=4): several MFMA tiles, off-diagonal tiles, padding rows"""
    coords, sb = _flat_booster(N=128, ndisk=1)
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=5, Lmax=4, diskR=DISKR)
    emit = B.axion_induced_modes(coords, modes, diskR=DISKR)
    got = B.transformer(sb.copy(), coords, modes, f=20e9, diskR=DISKR, emit=emit)
    mats = B.calc_propagation_matrices(sb, coords, modes, f=20e9, diskR=DISKR)
    want_m = _oracle_mats(sb, coords, modes, 20e9, O.PROP_PROPAGATOR)
    for r in range(1, len(sb.distance)):
        assert rel(mats[r], want_m[r]) < TOL, r
    want = _otransformer(sb, coords, modes, 20e9, O.PROP_PROPAGATOR, emit)
    eps = sb.eps.copy()
    eps[np.isnan(eps.real)] = 1e30
    sens = chain_sensitivity(_oracle_mats(sb, coords, modes, 20e9, O.PROP_PROPAGATOR, eps), eps, emit)
    assert rel(got, want) < max(TOL, 10 * sens), (rel(got, want), sens)


def test_propagation_matrices_match_oracle():
    coords, sb = _rough_booster(N=32)
    modes = _lg_like_modes(coords, 2, 1)
    mats = B.calc_propagation_matrices(sb, coords, modes, f=20e9, diskR=DISKR)
    lam = B.wavelength(20e9)
    pat, kt = modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1)
    for r in range(1, len(sb.distance)):
        want = O.propagation_matrix(sb.distance[r], DISKR, sb.eps[r], sb.relative_tilt_x[r], sb.relative_tilt_y[r],
                                    sb.relative_surfaces[r], lam, ocoords(coords), pat, kt)
        assert rel(mats[r], want) < TOL, r
    # the NaN mirror region propagates NaN as in the reference
    assert np.all(np.isnan(mats[0]))
    one = B.propagation_matrix(sb.distance[2], DISKR, sb.eps[2], sb.relative_tilt_x[2], sb.relative_tilt_y[2],
                               sb.relative_surfaces[2], lam, coords, modes, onlydiagonal=True)
    assert np.allclose(np.diag(one), np.diag(mats[2]), rtol=1e-12) and np.count_nonzero(one - np.diag(np.diag(one))) == 0


def test_transformer_precomputed_matrices_and_eps_mutation():
    coords, sb = _flat_booster(N=32)
    modes = _lg_like_modes(coords, 2, 0)
    s1 = sb.copy()
    direct = B.transformer(s1, coords, modes, f=20e9, diskR=DISKR)
    assert s1.eps[0] == 1e30  # bdry.eps[isnan] .= 1e30
    mats = B.calc_propagation_matrices(s1, coords, modes, f=20e9, diskR=DISKR)
    via = B.transformer(s1, coords, modes, f=20e9, diskR=DISKR, propagation_matrices=mats)
    assert rel(via, direct) < 1e-13


def test_trace_back_multimode_matches_oracle():
    coords, sb = _flat_booster(N=32)
    modes = _lg_like_modes(coords, 2, 1)
    sb.eps[0] = 1e30
    rng = np.random.default_rng(2)
    reflected = rng.standard_normal(modes.n) + 1j * rng.standard_normal(modes.n)
    inp = rng.standard_normal(modes.n) + 0j
    got = B.transformer_trace_back(reflected, inp, sb, coords, modes, f=20e9)
    os_ = O.Setup(sb.distance, sb.eps, None)
    want = O.transformer_trace_back(reflected, inp, os_, ocoords(coords), modes.patterns_flat()[:, 0],
                                    modes.mode_kt.reshape(-1), 20e9)
    assert rel(got, want) < TOL


def test_transformer_reference_vs_cheerleader_near_resonance():
    """1 disk in front of a mirror: the transformer with many modes approaches the 3D
    cheerleader's boost factor (examples/Single Disk.ipynb, cell 33) -- a coarse cross-check
    between the two independent device solvers."""
    coords = ref_coords()
    sb = with_distance(ref_setup_eps9(coords), deps_distance(2.0))
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=10, Lmax=0, diskR=DISKR)
    bt = np.sum(np.abs(B.transformer(sb.copy(), coords, modes, f=10e9, diskR=DISKR)) ** 2)
    E = B.cheerleader(0, 100, sb, coords, f=10e9, diskR=DISKR)
    bc = np.sum(np.abs(E) ** 2) * 1e-4 / (math.pi * DISKR ** 2)
    assert abs(bt / bc - 1) < 0.05


@pytest.mark.parametrize("M,Lm", [(7, 3), (13, 2), (9, 4)])
def test_propagation_matrices_many_modes(M, Lm):
    """n = 49, 65 and 81 modes (4x4 to 6x6 output tiles of 16 per problem: the
    ungrouped Gram kernel, in several tile batches for the larger two); flat
    air (Parseval), tilted/rough disks and an explicitly propagated air region,
    all against the oracle."""
    coords, sb = _rough_booster(N=32)
    sb.relative_tilt_x[1] = 2e-4  # one tilted air region: propagated explicitly
    modes = _lg_like_modes(coords, M, Lm)
    got = B.calc_propagation_matrices(sb, coords, modes, f=20.1e9, diskR=DISKR)
    want = _oracle_mats(sb, coords, modes, 20.1e9, O.PROP_PROPAGATOR)
    for r in range(1, len(sb.distance)):
        assert rel(got[r], want[r]) < TOL, (r, rel(got[r], want[r]))


@pytest.mark.parametrize("M,Lm", [(3, 1), (5, 2), (5, 4)])
def test_propagation_matrices_grouped_batches(M, Lm):
    """n = 9, 25, 45 modes (1x1, 2x2, 3x3 output tiles: the grouped, double-
    buffered Gram kernel), several frequencies: problems sharing A and B are
    grouped (up to four per block) with a partial last group; every matrix
    against the oracle."""
    coords, sb = _rough_booster(N=32, ndisk=3)
    modes = _lg_like_modes(coords, M, Lm)
    freqs = [19.7e9, 20.0e9, 20.3e9]
    got = B.calc_propagation_matrices(sb, coords, modes, f=freqs, diskR=DISKR)
    for q, f in enumerate(freqs):
        want = _oracle_mats(sb, coords, modes, f, O.PROP_PROPAGATOR)
        for r in range(1, len(sb.distance)):
            assert rel(got[q][r], want[r]) < TOL, (q, r, rel(got[q][r], want[r]))


def _batch_geometries(N=32):
    """uneven sweep: 1..4 disks, flat, tilted and rough, two mode sets"""
    coords = W.grid_coords(N)
    geoms = []
    for c, nd in enumerate([1, 3, 2, 4, 1, 2]):
        sb = W.booster(coords, ndisk=nd, seed=1200 + c, jitter=0.5e-3)
        if c == 2:
            _, rb = _rough_booster(N=N, ndisk=nd, seed=1300)
            sb = rb
        geoms.append(sb)
    return coords, geoms


@pytest.mark.parametrize("chunk_bytes", [None, "60000"])
def test_transformer_batch_matches_single_geometry(engine_options, chunk_bytes):
    """transformer_batch over geometries with uneven region counts and two mode
    sets (n = 6 and 9) equals transformer per geometry; with a small matrix
    budget the batch runs in several chunks per mode set"""
    if chunk_bytes:
        engine_options(batch_bytes=int(chunk_bytes))
    coords, geoms = _batch_geometries()
    sets = [_lg_like_modes(coords, 2, 1), _lg_like_modes(coords, 3, 1, seed=5)]
    set_of = [0, 1, 1, 0, 1, 0]
    emits = [B.axion_induced_modes(coords, md, diskR=DISKR) for md in sets]
    freqs = [19.6e9, 20.0e9, 20.4e9]
    got = B.transformer_batch([g.copy() for g in geoms], coords, sets, f=freqs, mode_set_of=set_of, diskR=DISKR,
                              emits=emits)
    assert len(got) == len(geoms)
    for c, sb in enumerate(geoms):
        md = sets[set_of[c]]
        assert got[c].shape == (len(freqs), md.n)
        want = B.transformer(sb.copy(), coords, md, f=freqs, diskR=DISKR, emit=emits[set_of[c]])
        eps = sb.eps.copy()
        eps[np.isnan(eps.real)] = 1e30
        for q, f in enumerate(freqs):
            err = rel(got[c][q], want[q])
            if err > TOL:
                sens = chain_sensitivity(_oracle_mats(sb, coords, md, f, O.PROP_PROPAGATOR, eps), eps,
                                         emits[set_of[c]])
                assert err < 10 * sens, (c, q, err, sens)


@pytest.mark.parametrize("chunk_bytes", [None, "60000"])
def test_transformer_batch_plan_matches_batch(engine_options, chunk_bytes):
    """TransformerBatchPlan (mode sets, sources and frequencies resident) equals
    transformer_batch bit for bit, on two runs in a row, the second over fewer
    geometries in another order and mode-set assignment; fetch() returns the last
    run's geometries only; with a small matrix budget every run is chunked"""
    if chunk_bytes:
        engine_options(batch_bytes=int(chunk_bytes))
    coords, geoms = _batch_geometries()
    sets = [_lg_like_modes(coords, 2, 1), _lg_like_modes(coords, 3, 1, seed=5)]
    emits = [B.axion_induced_modes(coords, md, diskR=DISKR) for md in sets]
    freqs = [19.6e9, 20.0e9, 20.4e9]
    plan = B.TransformerBatchPlan(coords, sets, freqs, diskR=DISKR, emits=emits)
    for pick, set_of in [(list(range(6)), [0, 1, 1, 0, 1, 0]), ([4, 1, 5], [1, 0, 1])]:
        sel = [geoms[i].copy() for i in pick]
        want = B.transformer_batch([g.copy() for g in sel], coords, sets, f=freqs, mode_set_of=set_of,
                                   diskR=DISKR, emits=emits)
        plan.run(sel, mode_set_of=set_of)
        got = plan.fetch()
        assert len(got) == len(pick)
        for c in range(len(pick)):
            assert got[c].shape == want[c].shape
            assert np.array_equal(got[c], want[c]), c
    plan.close()


def test_transformer_batch_matches_oracle():
    """two geometries of a batch against the CPU oracle (matrices and chain)"""
    coords, geoms = _batch_geometries()
    md = _lg_like_modes(coords, 2, 1)
    emit = B.axion_induced_modes(coords, md, diskR=DISKR)
    pick = [geoms[1], geoms[2]]
    got = B.transformer_batch([g.copy() for g in pick], coords, md, f=[20.1e9], diskR=DISKR, emits=[emit])
    for c, sb in enumerate(pick):
        want = _otransformer(sb, coords, md, 20.1e9, O.PROP_PROPAGATOR, emit)
        eps = sb.eps.copy()
        eps[np.isnan(eps.real)] = 1e30
        sens = chain_sensitivity(_oracle_mats(sb, coords, md, 20.1e9, O.PROP_PROPAGATOR, eps), eps, emit)
        assert rel(got[c][0], want) < max(TOL, 10 * sens), (c, rel(got[c][0], want), sens)


@pytest.mark.parametrize("lossy", [False, True])
def test_propagation_matrices_disk_frequency_expansion(engine_options, lossy):
    """tilted/rough disks over a 24-frequency band take the frequency expansion
    (moments of the height map): every region against the oracle at 1e-10 and
    against per-frequency weight tables; a lossy disk has a complex k0"""
    coords, sb = _rough_booster(N=32, ndisk=2)
    if lossy:
        sb.eps[2] = 24.0 - 0.5j
    modes = _lg_like_modes(coords, 2, 1)
    freqs = np.linspace(19.5e9, 20.5e9, 24)
    got = B.calc_propagation_matrices(sb, coords, modes, f=freqs, diskR=DISKR)
    engine_options(disk_expansion=0)
    direct = B.calc_propagation_matrices(sb, coords, modes, f=freqs, diskR=DISKR)
    for q in range(len(freqs)):
        for r in range(1, len(sb.distance)):
            assert rel(got[q][r], direct[q][r]) < 1e-13, (q, r, rel(got[q][r], direct[q][r]))
    for q in (0, 9, 23):
        want = _oracle_mats(sb, coords, modes, freqs[q], O.PROP_PROPAGATOR)
        for r in range(1, len(sb.distance)):
            assert rel(got[q][r], want[r]) < TOL, (q, r, rel(got[q][r], want[r]))


def test_transformer_plan_matches_transformer():
    """the resident sweep gives transformer's results run after run (rough booster:
    Parseval air, frequency-series disks, reflectivity)"""
    coords, sb = _rough_booster(N=32, ndisk=3)
    modes = _lg_like_modes(coords, 2, 1)
    emit = B.axion_induced_modes(coords, modes, diskR=DISKR)
    refl_in = np.linspace(1, 2, modes.n) + 0.1j
    freqs = np.linspace(19.6e9, 20.4e9, 9)
    want_b, want_r = B.transformer(sb.copy(), coords, modes, f=freqs, diskR=DISKR, emit=emit, reflect=refl_in)
    eps0 = sb.eps.copy()
    plan = B.TransformerPlan(sb, coords, modes, freqs, diskR=DISKR, emit=emit, reflect=refl_in)
    assert np.array_equal(sb.eps, eps0, equal_nan=True)  # the plan leaves the setup alone
    for _ in range(2):
        plan.run()
        b, r = plan.fetch()
        assert rel(b, want_b) < 1e-12 and rel(r, want_r) < 1e-12
    plan.close()


def test_transformer_more_than_48_modes_uses_global_chain():
    """n = 49 modes: the inverses and the chain take the global-memory kernels
    (the LDS ones hold n <= 48); boost and reflectivity against the oracle"""
    coords, sb = _flat_booster(N=32, ndisk=1)
    modes = _lg_like_modes(coords, 7, 3)
    assert modes.n == 49
    emit = B.axion_induced_modes(coords, modes, diskR=DISKR)
    refl_in = np.linspace(1, 2, modes.n) + 0.3j
    boost, refl = B.transformer(sb.copy(), coords, modes, f=20.0e9, diskR=DISKR, emit=emit, reflect=refl_in)
    want_b, want_r = _otransformer(sb, coords, modes, 20.0e9, O.PROP_PROPAGATOR, emit, refl_in)
    eps = sb.eps.copy()
    eps[np.isnan(eps.real)] = 1e30
    sens = chain_sensitivity(_oracle_mats(sb, coords, modes, 20.0e9, O.PROP_PROPAGATOR, eps), eps, emit)
    tol = max(TOL, 10 * sens)
    assert rel(boost, want_b) < tol and rel(refl, want_r) < tol, (rel(boost, want_b), rel(refl, want_r), sens)


def test_parseval_groups_match_per_bin_weights(engine_options):
    """flat air regions: the k-bin group GEMM against per-bin weight tables in the
    Gram kernel (option parseval_groups=0) and against the oracle"""
    coords, sb = _flat_booster(N=48, ndisk=2)
    modes = _lg_like_modes(coords, 3, 1)
    freqs = [19.8e9, 20.2e9]
    got = B.calc_propagation_matrices(sb, coords, modes, f=freqs, diskR=DISKR)
    engine_options(parseval_groups=0)
    per_bin = B.calc_propagation_matrices(sb, coords, modes, f=freqs, diskR=DISKR)
    for q, f in enumerate(freqs):
        want = _oracle_mats(sb, coords, modes, f, O.PROP_PROPAGATOR)
        for r in range(1, len(sb.distance)):
            assert rel(got[q][r], per_bin[q][r]) < 1e-12, (q, r)
            assert rel(per_bin[q][r], want[r]) < TOL and rel(got[q][r], want[r]) < TOL, (q, r)


def test_parseval_series_matches_direct(engine_options):
    """air gaps of one frequency whose spacings lie in one cluster take the Taylor
    series in the spacing (a few coefficient matrices per cluster, a Horner sum
    per gap): a 64-disk booster (64 gaps of 8 mm +- 0.3 mm per frequency and the
    last region at dz = 0, which stays direct) at 64 x 64 against the direct
    product (option parseval_series=0) to 1e-13 per matrix, and against the
    oracle on some gaps at 1e-10"""
    coords = W.grid_coords(64)
    sb = W.booster(coords, ndisk=64, seed=1710)
    modes = _lg_like_modes(coords, 3, 1)
    freqs = [19.8e9, 20.2e9]
    ctx = B.context()
    ctx.timing(True)
    got = B.calc_propagation_matrices(sb, coords, modes, f=freqs, diskR=DISKR)
    series_launches = ctx.kernel_time("pseries")[1]
    ctx.timing(False)
    assert series_launches >= 2  # coefficient and evaluation launches ran
    engine_options(parseval_series=0)
    direct = B.calc_propagation_matrices(sb, coords, modes, f=freqs, diskR=DISKR)
    nreg = len(sb.distance)
    lam = [B.wavelength(f) for f in freqs]
    pat, kt = modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1)
    for q in range(len(freqs)):
        for r in range(1, nreg):
            assert rel(got[q][r], direct[q][r]) < 1e-13, (q, r, rel(got[q][r], direct[q][r]))
        for r in (1, 3, 63, nreg - 3, nreg - 1):
            want = O.propagation_matrix(sb.distance[r], DISKR, sb.eps[r], 0.0, 0.0, None, lam[q], ocoords(coords),
                                        pat, kt, O.PROP_PROPAGATOR)
            assert rel(got[q][r], want) < TOL, (q, r, rel(got[q][r], want))


@pytest.mark.parametrize("flat", [False, True])
def test_transformer_batch_plan_matches_single_geometry_256(flat):
    """The sweep benchmark's shape (256 x 256, Bessel modes n = 9 and 25, boosters
    of 12 and 24 disks with spacings 8 mm +- 0.5 mm): the resident batch plan
    against transformer per geometry, per point within 1e-10 of the boost vector.
    flat=True (no surface arrays, the benchmark's form) takes the plan's flat
    chunks: gap matrices and their inverses from the spacing series, disks from
    G and inv(G); transformer forms every matrix by the direct product and every
    inverse by the LU."""
    coords = W.grid_coords(256)
    rng = W.SplitMix64(5200)
    geoms = []
    for nd in (12, 24):
        sp = 8e-3 + (2.0 * rng.uniforms(nd) - 1.0) * 0.5e-3
        geoms.append(W.booster(coords, ndisk=nd, spacings=list(sp), flat=flat))
    sets = [B.SeedModes(coords, ThreeDim=True, Mmax=3, Lmax=1, diskR=DISKR),
            B.SeedModes(coords, ThreeDim=True, Mmax=5, Lmax=2, diskR=DISKR)]
    set_of = [1, 0]
    emits = [B.axion_induced_modes(coords, md, diskR=DISKR) for md in sets]
    freqs = [19.3e9, 20.0e9, 20.7e9]
    plan = B.TransformerBatchPlan(coords, sets, freqs, diskR=DISKR, emits=emits)
    plan.run([g.copy() for g in geoms], mode_set_of=set_of)
    got = plan.fetch()
    plan.close()
    for c, sb in enumerate(geoms):
        md = sets[set_of[c]]
        want = B.transformer(sb.copy(), coords, md, f=freqs, diskR=DISKR, emit=emits[set_of[c]])
        for q in range(len(freqs)):
            assert rel(got[c][q], want[q]) < 1e-10, (c, q, rel(got[c][q], want[q]))


def test_batch_plan_flat_chunks_match_general_path(engine_options):
    """chunks of flat, untilted geometries take the frequency-expanded path (gap
    matrices and their inverses from the spacing series, disks from inv(G));
    with the series off the same plan takes the general path (direct products,
    LU inverses): boosts agree to 1e-10, every geometry and frequency, with one
    chunk and with many, and with the gaps' inverses by the LU.  A run whose
    inverse series fails its convergence check is redone with the LU inverses
    (forced here by series_inverse=2): bitwise the series_inverse=0 result.
    The chains of the geometries (3 to 45 disks) start longest first; in job
    order (chain_order=0) the boosts are bitwise the same."""
    coords = W.grid_coords(64)
    rng = W.SplitMix64(1750)
    geoms = []
    for nd in (3, 30, 12, 45, 7, 20):
        sp = 8e-3 + (2.0 * rng.uniforms(nd) - 1.0) * 0.5e-3
        geoms.append(W.booster(coords, ndisk=nd, spacings=list(sp), flat=True))
    sets = [B.SeedModes(coords, ThreeDim=True, Mmax=3, Lmax=1, diskR=DISKR),
            B.SeedModes(coords, ThreeDim=True, Mmax=5, Lmax=1, diskR=DISKR)]
    set_of = [0, 1, 1, 0, 1, 0]
    emits = [B.axion_induced_modes(coords, md, diskR=DISKR) for md in sets]
    freqs = np.linspace(19.5e9, 20.5e9, 20)
    out = {}
    for name, kw in (("flat", {}), ("flat_chunked", dict(batch_bytes=2_000_000)), ("general", dict(parseval_series=0)),
                     ("flat_lu", dict(series_inverse=0)), ("flat_rerun", dict(series_inverse=2)),
                     ("flat_job_order", dict(chain_order=0)), ("general_job_order", dict(parseval_series=0,
                                                                                          chain_order=0))):
        engine_options(**kw)
        plan = B.TransformerBatchPlan(coords, sets, freqs, diskR=DISKR, emits=emits)
        plan.run([g.copy() for g in geoms], mode_set_of=set_of)
        out[name] = plan.fetch()
        plan.close()
        engine_options(parseval_series=1, batch_bytes=0, series_inverse=1, chain_order=1)
    for c in range(len(geoms)):
        assert rel(out["flat"][c], out["general"][c]) < 1e-10, (c, rel(out["flat"][c], out["general"][c]))
        assert rel(out["flat_chunked"][c], out["general"][c]) < 1e-10, c
        assert rel(out["flat_lu"][c], out["general"][c]) < 1e-10, c
        assert np.array_equal(out["flat_rerun"][c], out["flat_lu"][c]), c
        assert np.array_equal(out["flat_job_order"][c], out["flat"][c]), c
        assert np.array_equal(out["general_job_order"][c], out["general"][c]), c


def test_parseval_product_over_upper_triangles_matches_full_product(engine_options):
    """Bessel patterns vanish outside the aperture, so the masked and unmasked
    pattern transforms agree and every group Gram H_g is Hermitian: the
    Parseval product then forms the upper triangles only (n (n + 1) / 2 of the
    n^2 entries per problem, read off the executed flops) and mirrors them.
    Boosts agree with the full product (pgemm_herm=0) to 1e-12 on the series
    path (flat chunks) and on the direct path (parseval_series=0).  Patterns
    that do not vanish outside the aperture keep the full product."""
    coords = W.grid_coords(64)
    rng = W.SplitMix64(1770)
    geoms = []
    for nd in (4, 11, 19):
        sp = 8e-3 + (2.0 * rng.uniforms(nd) - 1.0) * 0.5e-3
        geoms.append(W.booster(coords, ndisk=nd, spacings=list(sp), flat=True))
    freqs = np.linspace(19.5e9, 20.5e9, 12)
    ctx = B.context()

    def run(modes, **kw):
        engine_options(**kw)
        plan = B.TransformerBatchPlan(coords, [modes], freqs, diskR=DISKR)
        ctx.timing(True)
        plan.run([g.copy() for g in geoms])
        out = plan.fetch()
        flops = ctx.kernel_time("pgemm")[2]
        ctx.timing(False)
        plan.close()
        engine_options(pgemm_herm=1, parseval_series=1)
        return out, flops

    bessel = B.SeedModes(coords, ThreeDim=True, Mmax=3, Lmax=1, diskR=DISKR)
    n = bessel.n
    for kw in ({}, dict(parseval_series=0)):
        herm, fh = run(bessel, **kw)
        full, ff = run(bessel, pgemm_herm=0, **kw)
        assert fh > 0 and abs(fh / ff - (n + 1) / (2 * n)) < 1e-12, (kw, fh / ff)
        for c in range(len(geoms)):
            assert rel(herm[c], full[c]) < 1e-12, (kw, c, rel(herm[c], full[c]))
    lg = _lg_like_modes(coords, 3, 1)
    _, fh = run(lg)
    _, ff = run(lg, pgemm_herm=0)
    assert fh == ff


@pytest.mark.parametrize("mmax,lmax,n", [(3, 1, 9), (5, 1, 15), (5, 2, 25)])
def test_inverse_series_on_matrix_cores_matches_valu(engine_options, mmax, lmax, n):
    """the gaps' inverses from the inverse spacing series: the matrix-core kernel
    (8 < n <= 32; here n = 9, 15 and 25) against the FP64 VALU kernel (sinv_mfma=0),
    and both against the LU inverses (series_inverse=0): boosts to 1e-10"""
    coords = W.grid_coords(64)
    rng = W.SplitMix64(1780)
    geoms = []
    for nd in (6, 23, 14):
        sp = 8e-3 + (2.0 * rng.uniforms(nd) - 1.0) * 0.5e-3
        geoms.append(W.booster(coords, ndisk=nd, spacings=list(sp), flat=True))
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=mmax, Lmax=lmax, diskR=DISKR)
    assert modes.n == n
    freqs = np.linspace(19.5e9, 20.5e9, 16)
    out = {}
    for name, kw in (("mfma", {}), ("valu", dict(sinv_mfma=0)), ("lu", dict(series_inverse=0))):
        engine_options(**kw)
        plan = B.TransformerBatchPlan(coords, [modes], freqs, diskR=DISKR)
        plan.run([g.copy() for g in geoms])
        out[name] = plan.fetch()
        plan.close()
        engine_options(sinv_mfma=1, series_inverse=1)
    for c in range(len(geoms)):
        assert rel(out["mfma"][c], out["valu"][c]) < 1e-10, (c, rel(out["mfma"][c], out["valu"][c]))
        assert rel(out["mfma"][c], out["lu"][c]) < 1e-10, (c, rel(out["mfma"][c], out["lu"][c]))


def test_batch_plan_grid_per_mode_set_matches_plans_per_grid():
    """a plan whose mode sets live on grids of different sizes (64x64 and 48x40,
    bf_transformer_batch_plan_create_grids), run once over geometries of both
    (uneven disk counts, flat ones and a tilted rough one sampled on its set's
    grid), gives each geometry the boosts of a plan on its grid alone"""
    ga = W.grid_coords(64)
    gb = B.SeedCoordinateSystem(X=(-0.24, 0.01, 0.23), Y=(-0.2, 0.01, 0.19))
    assert (gb.nx, gb.ny) == (48, 40)
    sets = [B.SeedModes(ga, ThreeDim=True, Mmax=3, Lmax=1, diskR=DISKR),
            B.SeedModes(gb, ThreeDim=True, Mmax=4, Lmax=0, diskR=DISKR)]
    emits = [B.axion_induced_modes(g, md, diskR=DISKR) for g, md in zip((ga, gb), sets)]
    freqs = np.linspace(19.5e9, 20.5e9, 12)
    rng = W.SplitMix64(1760)
    geoms, set_of = [], []
    for nd, k in ((5, 0), (17, 1), (3, 1), (26, 0), (9, 1), (12, 0)):
        sp = 8e-3 + (2.0 * rng.uniforms(nd) - 1.0) * 0.5e-3
        geoms.append(W.booster(ga if k == 0 else gb, ndisk=nd, spacings=list(sp), flat=True))
        set_of.append(k)
    rough = W.booster(gb, ndisk=3, seed=1761)
    nreg = len(rough.distance)
    rough.relative_tilt_x = np.array([2e-4 * rng.normal() for _ in range(nreg)])
    XX, YY = np.meshgrid(gb.X, gb.Y)
    surf = np.zeros((nreg, gb.ny, gb.nx), dtype=complex)
    for i in range(1, nreg - 1):
        surf[i] = 3e-6 * np.cos(13 * XX + i) * np.sin(7 * YY - i)
    rough.relative_surfaces = surf
    geoms.append(rough)
    set_of.append(1)
    plan = B.TransformerBatchPlan([ga, gb], sets, freqs, diskR=DISKR, emits=emits)
    plan.run([g.copy() for g in geoms], mode_set_of=set_of)
    both = plan.fetch()
    plan.close()
    for k, g in enumerate((ga, gb)):
        idx = [c for c in range(len(geoms)) if set_of[c] == k]
        alone = B.TransformerBatchPlan(g, [sets[k]], freqs, diskR=DISKR, emits=[emits[k]])
        alone.run([geoms[c].copy() for c in idx])
        got = alone.fetch()
        alone.close()
        for c, a in zip(idx, got):
            assert both[c].shape == (len(freqs), sets[k].n)
            assert rel(both[c], a) < 1e-12, (c, rel(both[c], a))


# one n per LDS chain kernel shape: the fixed-n kernels (9, 15, 25, 45) and the generic ones with 1, 2, 3, 4,
# 5, 6, 7, 9 and 10 columns per thread; 16, 32, 35, 42 and 45 have no spare column of the thread grid
# (the axion sum takes its own loop), the others form it as column n of the products
CHAIN_NS = [1, 4, 9, 15, 16, 20, 24, 25, 30, 32, 33, 35, 36, 40, 42, 44, 45, 48]


@pytest.mark.parametrize("n", CHAIN_NS)
def test_chain_every_shape_matches_oracle(n):
    """boost and reflectivity from given propagation matrices (random, well conditioned) for every mode
    count class of the chain kernels, against the oracle's chain"""
    coords, sb = _flat_booster(N=16, ndisk=3)
    modes = _lg_like_modes(coords, n, 0)
    assert modes.n == n
    rng = np.random.default_rng(500 + n)
    nreg = len(sb.distance)
    mats = [np.eye(n, dtype=complex)]
    for _ in range(1, nreg):
        ph = np.exp(2j * np.pi * rng.uniform(size=n))
        X = np.eye(n) + 0.2 / np.sqrt(n) * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        mats.append(ph[:, None] * X)
    emit = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    refl_in = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    boost, refl = B.transformer(sb.copy(), coords, modes, f=20e9, diskR=DISKR, emit=emit, reflect=refl_in,
                                propagation_matrices=mats)
    eps = sb.eps.copy()
    eps[np.isnan(eps.real)] = 1e30
    os_ = O.Setup(sb.distance, eps, None, sb.relative_tilt_x, sb.relative_tilt_y, None)
    want_b, want_r = O.transformer(os_, ocoords(coords), modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1),
                                   20e9, O.PROP_PROPAGATOR, prop_mats=mats, diskR=DISKR, emit=emit, reflect=refl_in)
    sens = chain_sensitivity(mats, eps, emit)
    tol = max(TOL, 10 * sens)
    assert tol < 1e-7
    assert rel(boost, want_b) < tol and rel(refl, want_r) < tol, (rel(boost, want_b), rel(refl, want_r), sens)


@pytest.mark.parametrize("n", [9, 15, 25, 45])
def test_chain_tiled_matches_lds_chain(engine_options, n):
    """the register-tiled chain kernels (n = 9, 15, 25, 45) against the LDS chain kernel (option
    chain_tile=0) on the same matrices: every sum runs in the same order, so 9, 15 and 25 agree bitwise;
    for 45 (no spare column: the LDS kernel sums the axion term as one dot product of U + V) to 1e-13"""
    coords, sb = _flat_booster(N=16, ndisk=4)
    modes = _lg_like_modes(coords, n, 0)
    rng = np.random.default_rng(900 + n)
    mats = [np.eye(n, dtype=complex)]
    for _ in range(1, len(sb.distance)):
        ph = np.exp(2j * np.pi * rng.uniform(size=n))
        X = np.eye(n) + 0.2 / np.sqrt(n) * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        mats.append(ph[:, None] * X)
    emit = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    refl_in = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out = {}
    for tile in (1, 0):
        engine_options(chain_tile=tile, chain_mfma=0)
        out[tile] = B.transformer(sb.copy(), coords, modes, f=20e9, diskR=DISKR, emit=emit, reflect=refl_in,
                                  propagation_matrices=mats)
    for a, b in zip(out[1], out[0]):
        if n == 45:
            assert rel(a, b) < 1e-13, rel(a, b)
        else:
            assert np.array_equal(a, b), rel(a, b)


@pytest.mark.parametrize("form", [2, 3, 4])
@pytest.mark.parametrize("n", [1, 7, 9, 15, 16, 17, 25, 31, 32])
def test_chain_mfma_matches_valu_chain(engine_options, n, form):
    """the matrix-core chains forced for every n <= 32 (operands padded to 16 or 32; both products in one
    wave with four real MFMAs per complex product, option chain_mfma=2, or three, 3; the products in
    separate waves, 4) against the VALU chain kernels (chain_mfma=0) on the same matrices, boost and
    reflectivity to 1e-12 (each MFMA sums its four k in an order of its own); 16 and 32 fill the padded
    size (the axion sum by its own dot product)"""
    coords, sb = _flat_booster(N=16, ndisk=4)
    modes = _lg_like_modes(coords, n, 0)
    rng = np.random.default_rng(950 + n)
    mats = [np.eye(n, dtype=complex)]
    for _ in range(1, len(sb.distance)):
        ph = np.exp(2j * np.pi * rng.uniform(size=n))
        X = np.eye(n) + 0.2 / np.sqrt(n) * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        mats.append(ph[:, None] * X)
    emit = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    refl_in = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    out = {}
    for mf in (form, 0):
        engine_options(chain_mfma=mf)
        out[mf] = B.transformer(sb.copy(), coords, modes, f=20e9, diskR=DISKR, emit=emit, reflect=refl_in,
                                propagation_matrices=mats)
    for a, b in zip(out[form], out[0]):
        assert rel(a, b) < 1e-12, rel(a, b)
