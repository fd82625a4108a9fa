"""GPU parity on the benchmark configurations' own workloads.

One test (or a pair) per configuration, on its geometry, grid, modes and
frequency band, through the same entry points the benchmark times (dancer
plans, propagation matrices, the resident sweep plan), against the CPU
oracle.  Where the oracle's cost forces a cut (bounces or frequencies), the
docstring states it; the full-size runs are then checked through
size-independent properties (mixed precision against FP64 over every point).

FP64: 1e-10 relative (fields: max |diff| / max |oracle|; boost vectors per
frequency).  Mixed precision: relative beta^2 error <= 1e-4 against FP64.
"""
import numpy as np
import pytest

import BoostFractor as B
from BoostFractor import workloads as W
from helpers import DISKR, O, chain_sensitivity, ocoords, osetup, pmap, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10
MIXED_TOL = 1e-4


def _beta2(E, coords):
    return W.beta2_dancer(E, coords, DISKR)


def _plan(sb, coords, freqs, nmax, prec):
    plan = B.DancerPlan(0, nmax, sb, coords, freqs, diskR=DISKR, precision=prec)
    plan.run()
    E, _ = plan.fetch()
    plan.close()
    return E


def test_config1_single_disk_in_full():
    """configuration 1 in full: one disk (eps 24, 1 mm, R = 0.15 m) 8 mm in front of
    the mirror, 128 x 128 over 1 m, 20 GHz, nmax = 500 (501 bounces), FP64 dancer"""
    coords = W.grid_coords(128)
    sb = W.booster(coords, ndisk=1, spacings=[8e-3])
    print("config 1: dancer 128x128, 1 disk, nmax=500, 20 GHz", flush=True)
    got = B.dancer(0, 500, sb, coords, f=20e9, diskR=DISKR)
    want = O.dancer(osetup(sb), ocoords(coords), 0, 500, 20e9, O.PROP_PROPAGATOR, diskR=DISKR)
    assert rel(got, want) < TOL, rel(got, want)
    assert abs(_beta2(got, coords) / _beta2(want, coords) - 1) < TOL
    plan = _plan(sb, coords, [20e9], 500, "fp64")
    assert rel(plan[0], want) < TOL


def test_config2_dancer_fp64_and_mixed_vs_oracle():
    """configuration 2's booster (20 disks, spacings from seed 1000) on its grid
    (256 x 256) through the benchmark's plan over all 50 frequencies; the oracle
    checks frequencies 0 and 37 with nmax cut to 20 (21 bounces, ~16 s of oracle
    per point): FP64 fields at 1e-10, mixed precision beta^2 at 1e-4"""
    coords = W.grid_coords(256)
    sb = W.booster(coords, ndisk=20, seed=1000)
    freqs = np.linspace(19.5e9, 20.5e9, 50)
    nmax = 20
    print("config 2: dancer 256x256, 20 disks, 50 freqs, nmax=20 (fp64, mixed)", flush=True)
    E64 = _plan(sb, coords, freqs, nmax, "fp64")
    E32 = _plan(sb, coords, freqs, nmax, "mixed")
    os_, oc = osetup(sb), ocoords(coords)
    pick = [0, 37]
    want = pmap(lambda q: O.dancer(os_, oc, 0, nmax, freqs[q], O.PROP_PROPAGATOR, diskR=DISKR), pick)
    for q, w in zip(pick, want):
        assert rel(E64[q], w) < TOL, (q, rel(E64[q], w))
        assert abs(_beta2(E32[q], coords) / _beta2(w, coords) - 1) < MIXED_TOL
    err = np.abs(_beta2(E32, coords) / _beta2(E64, coords) - 1)
    assert np.max(err) < MIXED_TOL, np.max(err)


def test_config2_full_size_mixed_vs_fp64():
    """configuration 2 at full size: 50 frequencies, nmax = 2000; every point's
    mixed-precision beta^2 within 1e-4 of FP64 (the FP64 path is held to the
    oracle on the same booster and grid above)"""
    coords = W.grid_coords(256)
    sb = W.booster(coords, ndisk=20, seed=1000)
    freqs = np.linspace(19.5e9, 20.5e9, 50)
    print("config 2: full size, 50 freqs, nmax=2000, mixed vs fp64", flush=True)
    b64 = _beta2(_plan(sb, coords, freqs, 2000, "fp64"), coords)
    b32 = _beta2(_plan(sb, coords, freqs, 2000, "mixed"), coords)
    err = np.abs(b32 / b64 - 1)
    assert np.all(np.isfinite(b64)) and np.min(b64) > 0
    assert np.max(err) < MIXED_TOL, np.max(err)


def _rough_512():
    coords = W.grid_coords(512)
    return coords, W.rough_booster(coords, ndisk=20, seed=1002)


def _oracle_region_mats(sb, coords, pat, kt, f, eps):
    lam = B.wavelength(f)
    oc = ocoords(coords)
    regions = list(range(1, len(sb.distance)))

    surf = np.asarray(sb.relative_surfaces)
    tx = np.broadcast_to(sb.relative_tilt_x, (len(sb.distance),))
    ty = np.broadcast_to(sb.relative_tilt_y, (len(sb.distance),))

    def one(r):
        s = surf[r] if surf.shape[0] else None
        return O.propagation_matrix(sb.distance[r], DISKR, eps[r], tx[r], ty[r], s, lam, oc, pat, kt)

    mats = pmap(one, regions, workers=16)
    return [np.eye(pat.shape[0], dtype=complex)] + mats


def test_config3_lg_propagation_matrices_vs_oracle():
    """configuration 3's matrices: the 20-disk booster with per-disk tilts
    N(0, 50 urad) and 10 um roughness, 512 x 512, 45 Laguerre-Gauss modes,
    computed for all 200 frequencies (air regions by Parseval k-bin groups, the
    rough disks by the frequency series, as the benchmark runs them); every one
    of the 41 regions at frequency 57 against the oracle at 1e-10.  The matrix
    entries carry the parity figure: with these modes the chain's boost vector
    is dominated by rounding (condition ~1e7), see the Bessel-mode chain below."""
    coords, sb = _rough_512()
    pat, kt = W.lg_modes(coords)
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=5, Lmax=4, pattern_input=pat, kT_arr=kt)
    freqs = np.linspace(19.5e9, 20.5e9, 200)
    q = 57
    print("config 3: propagation matrices 512x512, 45 LG modes, 200 freqs", flush=True)
    got = B.calc_propagation_matrices(sb, coords, modes, f=freqs, diskR=DISKR)
    eps = sb.eps.copy()
    want = _oracle_region_mats(sb, coords, modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1), freqs[q], eps)
    for r in range(1, len(sb.distance)):
        assert rel(got[q][r], want[r]) < TOL, (r, rel(got[q][r], want[r]))


def test_config3_booster_chain_with_bessel_modes():
    """configuration 3's booster (tilted, rough, 512 x 512) end to end with 45
    orthonormal Bessel modes (M = 5, L = 4), where the chain is well conditioned:
    boost vector at frequency 20 GHz against the oracle transformer at 1e-10
    (tolerance scaled by the chain's measured FP64 sensitivity if that is larger)"""
    coords, sb = _rough_512()
    modes = B.SeedModes(coords, ThreeDim=True, Mmax=5, Lmax=4, diskR=DISKR)
    emit = B.axion_induced_modes(coords, modes, diskR=DISKR)
    f = 20e9
    print("config 3: transformer 512x512, 45 Bessel modes, rough booster", flush=True)
    got = B.transformer(sb.copy(), coords, modes, f=[19.9e9, f], diskR=DISKR, emit=emit)[1]
    eps = sb.eps.copy()
    eps[np.isnan(eps.real)] = 1e30
    pat, kt = modes.patterns_flat()[:, 0], modes.mode_kt.reshape(-1)
    mats = _oracle_region_mats(sb, coords, pat, kt, f, eps)
    surf = sb.relative_surfaces
    want = O.transformer(O.Setup(sb.distance, eps, None, sb.relative_tilt_x, sb.relative_tilt_y, surf),
                         ocoords(coords), pat, kt, f, O.PROP_PROPAGATOR, prop_mats=mats, diskR=DISKR, emit=emit)
    sens = chain_sensitivity(mats, eps, emit)
    assert sens < 1e-11, sens
    assert rel(got, want) < max(TOL, 10 * sens), (rel(got, want), sens)


def test_config4_dancer512_vs_oracle():
    """configuration 4's booster at 512 x 512 (the three-pass register FFT and
    the 512-point row pass) through the benchmark's plans over all 100
    frequencies; the oracle checks frequencies 0 and 61 with nmax cut to 7
    (8 bounces, ~26 s of oracle per point): FP64 fields at 1e-10, mixed
    precision beta^2 at 1e-4"""
    coords = W.grid_coords(512)
    sb = W.booster(coords, ndisk=20, seed=1000)
    freqs = np.linspace(19.5e9, 20.5e9, 100)
    nmax = 7
    print("config 4: dancer 512x512, 20 disks, 100 freqs, nmax=7 (fp64, mixed)", flush=True)
    E64 = _plan(sb, coords, freqs, nmax, "fp64")
    E32 = _plan(sb, coords, freqs, nmax, "mixed")
    os_, oc = osetup(sb), ocoords(coords)
    pick = [0, 61]
    want = pmap(lambda q: O.dancer(os_, oc, 0, nmax, freqs[q], O.PROP_PROPAGATOR, diskR=DISKR), pick)
    for q, w in zip(pick, want):
        assert rel(E64[q], w) < TOL, (q, rel(E64[q], w))
        assert abs(_beta2(E32[q], coords) / _beta2(w, coords) - 1) < MIXED_TOL


def test_config4_full_nmax_mixed_vs_fp64():
    """configuration 4 at its full bounce count (nmax = 3000) on 4 of its 100
    frequencies: mixed-precision beta^2 within 1e-4 of FP64"""
    coords = W.grid_coords(512)
    sb = W.booster(coords, ndisk=20, seed=1000)
    freqs = np.linspace(19.5e9, 20.5e9, 100)[[0, 33, 66, 99]]
    print("config 4: 512x512, 4 freqs, nmax=3000, mixed vs fp64", flush=True)
    b64 = _beta2(_plan(sb, coords, freqs, 3000, "fp64"), coords)
    b32 = _beta2(_plan(sb, coords, freqs, 3000, "mixed"), coords)
    err = np.abs(b32 / b64 - 1)
    assert np.max(err) < MIXED_TOL, np.max(err)


# the sweep's geometries #395 (10 disks, 9 modes), #426 (45 disks, 15 modes), #132 (80 disks, 25 modes)
SWEEP_PICK = [395, 426, 132]


def test_config5_sweep_geometries_vs_oracle():
    """configuration 5: three of its 4096 seeded geometries (10, 45 and 80 disks
    with 9, 15 and 25 Bessel modes), 256 x 256, through the resident sweep plan
    over all 100 frequencies as the benchmark runs it; boost vectors at
    frequencies 0 and 61 against the oracle transformer at 1e-10"""
    coords = W.grid_coords(256)
    geoms, set_of = [], []
    for c in SWEEP_PICK:
        g, s = W.sweep_geometries(coords, 1, first=c)
        geoms += g
        set_of.append(int(s[0]))
    nd = [(len(g.distance) - 2) // 2 for g in geoms]
    assert nd == [10, 45, 80] and set_of == [0, 1, 2]
    sets = [B.SeedModes(coords, ThreeDim=True, Mmax=M, Lmax=L_, diskR=DISKR) for M, L_ in W.SWEEP_MODESETS]
    emits = [B.axion_induced_modes(coords, md, diskR=DISKR) for md in sets]
    freqs = np.linspace(19e9, 21e9, 100)
    print("config 5: sweep plan, 3 geometries (10/45/80 disks), 256x256, 100 freqs", flush=True)
    plan = B.TransformerBatchPlan(coords, sets, freqs, diskR=DISKR, emits=emits)
    plan.run([g.copy() for g in geoms], mode_set_of=set_of)
    got = plan.fetch()
    plan.close()
    for g, s, out in zip(geoms, set_of, got):
        md = sets[s]
        pat, kt = md.patterns_flat()[:, 0], md.mode_kt.reshape(-1)
        eps = g.eps.copy()
        eps[np.isnan(eps.real)] = 1e30
        for q in (0, 61):
            mats = _oracle_region_mats(g, coords, pat, kt, freqs[q], eps)
            want = O.transformer(O.Setup(g.distance, eps), ocoords(coords), pat, kt, freqs[q], O.PROP_PROPAGATOR,
                                 prop_mats=mats, diskR=DISKR, emit=emits[s])
            err = rel(out[q], want)
            if err > TOL:
                sens = chain_sensitivity(mats, eps, emits[s])
                assert err < 10 * sens and sens < 1e-9, (len(g.distance), q, err, sens)
