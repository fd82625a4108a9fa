"""The CPU oracle pinned against the reference's own golden values (CPU only).

Each check below would fail for a plausible slip in the restatement: the
signed-zero kz branch (evanescent growth -> NaN), the fftshift offset, the
nmax+1 bounce count, the strict/inclusive mask inequalities, the eps
mutation of the transformer, the Deps=9 trace-back geometry.
"""
import math

import numpy as np
import pytest

import BoostFractor as B
from helpers import (DISKR, GOLDEN, O, deps_distance, ocoords, osetup, pmap, ref_coords, ref_dx, ref_setup_default,
                     ref_setup_eps9, with_distance)

TOL = 1e-12


@pytest.fixture(scope="module")
def rc():
    return ref_coords()


def test_cheerleader_3d_golden(rc):
    """test/runtests.jl:211"""
    oc, dx = ocoords(rc), ref_dx(rc)
    sb = ref_setup_default(rc)
    res = pmap(lambda d: O.cheerleader(osetup(with_distance(sb, deps_distance(d))), oc, 0, 100, 10e9,
                                       O.PROP_PROPAGATOR, diskR=DISKR), list(np.arange(0, 3 * math.pi, 1.0)))
    b = np.array([np.sum(np.abs(E) ** 2) * dx * dx / (math.pi * DISKR ** 2) for E in res])
    g = np.array(GOLDEN["cheerleader_3d_deps_0_1_3pi"]["beta2"])
    assert np.max(np.abs(b / g - 1)) < TOL
    # the commented-out fine sweep (:212) holds the same values every 10th point
    g2 = np.array(GOLDEN["cheerleader_3d_deps_0_0.1_3pi"]["beta2"])
    assert np.max(np.abs(g2[::10] / g - 1)) < 1e-14


def test_dancer_1d_golden_and_closed_form(rc):
    """test/runtests.jl:214, and the mirror-only invariant: beta^2 = area ratio (709 dx^2/(pi R^2))."""
    oc, dx = ocoords(rc), ref_dx(rc)
    sb = ref_setup_eps9(rc)
    b = []
    for d in np.arange(0, 3 * math.pi, 1.0):
        E = O.dancer(osetup(with_distance(sb, deps_distance(d))), oc, 0, 200, 10e9, O.PROP_1D, diskR=DISKR)
        b.append(np.sum(np.abs(E) ** 2) * dx * dx / (math.pi * DISKR ** 2))
    g = np.array(GOLDEN["dancer_1d_deps_0_1_3pi"]["beta2"])
    assert np.max(np.abs(np.array(b) / g - 1)) < TOL
    XX, YY = np.meshgrid(rc.X, rc.Y)
    inside = np.count_nonzero(XX ** 2 + YY ** 2 <= DISKR ** 2)
    assert inside == 709
    # Deps = 0 is a zero-thickness disk: the mirror alone, boost factor = area fraction emitting
    mirror = dx * dx * inside / (math.pi * DISKR ** 2)
    assert abs(g[0] / mirror - 1) < 1e-2


def test_dancer_3d_golden(rc):
    """test/runtests.jl:215 (commented out): 3D dancer nmax=200 on Deps=0:0.1:3pi (spread of points)."""
    oc, dx = ocoords(rc), ref_dx(rc)
    sb = ref_setup_eps9(rc)
    Deps = np.arange(0, 3 * math.pi, 0.1)
    idx = [0, 17, 47, 80]
    res = pmap(lambda i: O.dancer(osetup(with_distance(sb, deps_distance(Deps[i]))), oc, 0, 200, 10e9,
                                  O.PROP_PROPAGATOR, diskR=DISKR), idx)
    g = np.array(GOLDEN["dancer_3d_deps_0_0.1_3pi"]["beta2"])
    for i, E in zip(idx, res):
        assert abs(np.sum(np.abs(E) ** 2) * dx * dx / (math.pi * DISKR ** 2) / g[i] - 1) < TOL


def test_transformer_modes_traceback_goldens(rc):
    """test/runtests.jl:166-171, 217, 221"""
    oc = ocoords(rc)
    sb = ref_setup_eps9(rc)
    pat, kt = O.seed_modes(oc, 10, 0, DISKR)
    emit = O.field2modes(oc, pat, None, 0.0, DISKR, 10e9)
    field = O.modes2field(pat, emit)
    g = GOLDEN["modes2field_field5000"]
    assert abs(field.ravel()[4999] - g["re"]) < 1e-14
    eps = sb.eps.copy()
    eps[0] = 1e30
    b = pmap(lambda d: np.sum(np.abs(O.transformer(O.Setup(deps_distance(d), eps), oc, pat, kt, 10e9,
                                                   emit=emit, diskR=DISKR)) ** 2),
             list(np.arange(0, 3 * math.pi, 1.0)))
    g = np.array(GOLDEN["transformer_3d_deps_0_1_3pi"]["beta2"])
    assert np.max(np.abs(np.array(b) / g - 1)) < TOL
    pm = O.propagation_matrix(0.0, 0.15, 1e30, 0.0, 0.0, None, B.wavelength(10e9), oc, pat, kt)
    assert abs(pm[3, 5] - GOLDEN["propmat1_4_6"]["re"]) < 1e-14
    # trace back: 1D coords, the reference traces through the Deps=9 geometry left by its loop
    c1 = B.SeedCoordinateSystem(X=[1e-7], Y=[1e-7])
    o1 = ocoords(c1)
    p1, k1 = np.ones((1, 1, 1), dtype=complex), np.zeros(1, dtype=complex)
    e1 = O.field2modes(o1, p1, None, 0.0, 0.15, 10e9)
    Deps = np.arange(0, 3 * math.pi, 1.0)
    rtb = [O.transformer(O.Setup(deps_distance(d), eps), o1, p1, k1, 10e9, O.PROP_1D, emit=e1, diskR=DISKR)[0]
           for d in Deps]
    ff = O.transformer_trace_back(rtb[4], 0.0, O.Setup(deps_distance(Deps[-1]), eps), o1, p1, k1, 10e9, O.PROP_1D)
    gg = np.array(GOLDEN["trace_back_full_fields"]["re"]) + 1j * np.array(GOLDEN["trace_back_full_fields"]["im"])
    assert np.max(np.abs(ff[:, :, 0] - gg)) < 1e-12
    # with the Deps = 4 geometry (no leaked global) the golden does NOT match
    ff4 = O.transformer_trace_back(rtb[4], 0.0, O.Setup(deps_distance(Deps[4]), eps), o1, p1, k1, 10e9, O.PROP_1D)
    assert np.max(np.abs(ff4[:, :, 0] - gg)) > 1e-3


def test_beam_shapes_goldens():
    """examples/Beam Shapes.ipynb cell 23 (uniform beam on 7 Bessel modes)."""
    c = B.SeedCoordinateSystem(X=(-0.4445, 0.007, 0.4445), Y=(-0.4445, 0.007, 0.4445))
    # k-space grid printed by the notebook (cell 13)
    assert c.kX[0] == -448.79895051282756 and c.kX[-1] == 448.79895051282756
    oc = ocoords(c)
    pat, kt = O.seed_modes(oc, 7, 0, 0.15)
    XX, YY = np.meshgrid(c.X, c.Y)
    ub = np.where(XX ** 2 + YY ** 2 > 0.15 ** 2, 0.0, 1.0)
    m = O.field2modes(oc, pat, ub, 0.0, 0.15, 20e9)
    g = np.array(GOLDEN["beam_shapes_uniform_field2modes_M7"]["re"])
    assert np.max(np.abs(m - g)) < 1e-14


def test_bessel_zeros_against_tabulated_values():
    """zeros of J_0, J_1, J_4 from standard tables (Abramowitz & Stegun 9.5)"""
    tab = {(0, 1): 2.404825557695773, (0, 2): 5.520078110286311, (1, 1): 3.831705970207512,
           (1, 3): 10.17346813506272, (4, 1): 7.588342434503805}
    for (nu, m), v in tab.items():
        assert abs(O.besselj_zero(nu, m) - v) < 1e-13


def test_propagator_closed_forms(rc):
    """dz = 0 leaves the masked field (FFT round trip); a plane wave at DC in the
    aperture-free 1D limit picks up exp(-i k0 dz); evanescent components decay."""
    oc = ocoords(rc)
    rng = np.random.default_rng(0)
    E = rng.standard_normal((101, 101)) + 1j * rng.standard_normal((101, 101))
    out = O.propagate(O.PROP_PROPAGATOR, E, oc, 0.0, 0.2, 1.0, 0, 0, None, 0.03)
    XX, YY = np.meshgrid(rc.X, rc.Y)
    assert np.max(np.abs(out - np.where(XX ** 2 + YY ** 2 < 0.2 ** 2, E, 0))) < 1e-13
    lam, dz = 0.03, 0.0123
    one = O.propagate(O.PROP_1D, np.ones((1, 1)), O.Coords([1e-7], [1e-7], [0.0], [0.0]), dz, 0.15, 4.0, 0, 0,
                      None, lam)
    assert abs(one[0, 0] - np.exp(-1j * 2 * math.pi / lam * 2.0 * dz)) < 1e-14
    # |exp(-i kz dz)| <= 1 for every k: the propagated energy never grows
    out = O.propagate(O.PROP_NOTILTS, E, oc, 0.05, 1.0, complex(1.0, -0.0), 0, 0, None, 0.03)
    assert np.sum(np.abs(out) ** 2) <= np.sum(np.abs(E) ** 2) * (1 + 1e-12)
    # with +0.0 imaginary eps (no signed-zero cast) the evanescent modes grow: the trap the cast avoids
    bad = O.propagate(O.PROP_NOTILTS, E, oc, 0.05, 1.0, complex(1.0, 0.0), 0, 0, None, 0.03)
    assert np.sum(np.abs(bad) ** 2) > 10 * np.sum(np.abs(E) ** 2)


def test_dft_against_brute_force():
    """the oracle's general-N FFT against a brute-force 2D DFT (12 x 10 grid)"""
    rng = np.random.default_rng(1)
    E = rng.standard_normal((10, 12)) + 1j * rng.standard_normal((10, 12))
    X = np.linspace(-0.5, 0.5, 12)
    Y = np.linspace(-0.5, 0.5, 10)
    kX = B.get_kspace_coords(X)
    kY = B.get_kspace_coords(Y)
    oc = O.Coords(X, Y, kX, kY)
    # dz = 0 with a mask covering everything: FFT -> x1 -> IFFT, the identity
    out = O.propagate(O.PROP_NOTILTS, E, oc, 0.0, 10.0, 1.0, 0, 0, None, 0.03)
    assert np.max(np.abs(out - E)) < 1e-13
    # momentum-space propagator at dz=0 without aperture/tilt: IFFT then FFT with shifts = identity
    out = O.propagate(O.PROP_MOMENTUM, E, oc, 0.0, 10.0, 1.0, 0, 0, None, 0.03)
    assert np.max(np.abs(out - E)) < 1e-13
    # one propagation against numpy's FFT with the explicit fftshifted kz phase
    dz, lam = 0.01, 0.02
    k0 = 2 * math.pi / lam
    KX, KY = np.meshgrid(kX, kY)
    arg = k0 * k0 - KX ** 2 - KY ** 2
    kz = np.where(arg >= 0, np.sqrt(np.abs(arg)) + 0j, -1j * np.sqrt(np.abs(arg)))
    ph = np.exp(-1j * kz * dz)
    want = np.fft.ifft2(np.fft.ifftshift(np.fft.fftshift(np.fft.fft2(E)) * ph))
    got = O.propagate(O.PROP_NOTILTS, E, oc, dz, 10.0, complex(1.0, -0.0), 0, 0, None, lam)
    assert np.max(np.abs(got - want)) < 1e-12


def test_gauss_profile_golden_and_closed_forms():
    """test/runtests.jl:46-47: gauss_profile(coords; z=1e-9, omega0=0.097, f=22e9)[4000] on the
    default 101^2 grid; near the waist |profile| = exp(-r^2/omega0^2) and the beam's power
    integral is pi omega0^2 / 2."""
    c = B.SeedCoordinateSystem()
    oc = ocoords(c)
    g = O.gauss_profile(oc, z=1e-9, omega0=0.097, f=22e9)
    gd = GOLDEN["gauss_profile_4000"]
    assert g.ravel()[3999] == complex(gd["re"], gd["im"])  # Julia linear index 4000, x fastest
    XX, YY = np.meshgrid(c.X, c.Y)
    assert np.max(np.abs(np.abs(g) - np.exp(-(XX ** 2 + YY ** 2) / 0.097 ** 2))) < 1e-15
    assert abs(np.sum(np.abs(g) ** 2) * 1e-4 / (math.pi * 0.097 ** 2 / 2) - 1) < 1e-3
    # far from the waist the amplitude drops as omega0/omega(z): omega(z)^2 = omega0^2 (1 + (z/zR)^2)
    z, lam = 0.5, 299792458.0 / 22e9
    zR = math.pi * 0.097 ** 2 / lam
    gz = O.gauss_profile(oc, z=z, omega0=0.097, f=22e9)
    assert abs(np.abs(gz[50, 50]) - 1 / math.sqrt(1 + (z / zR) ** 2)) < 1e-14


def test_unnormalized_match_closed_forms():
    """src/beam_coupling.jl:13-15: self match = sum |E|^2 dx dy (real); match(a, b) = conj(match(b, a))"""
    rng = np.random.default_rng(4)
    a = rng.standard_normal((37, 41)) + 1j * rng.standard_normal((37, 41))
    b = rng.standard_normal((37, 41)) + 1j * rng.standard_normal((37, 41))
    s = O.unnormalized_match(a, a, 0.02, 0.03)
    assert s.imag == 0.0 and abs(s.real / (np.sum(np.abs(a) ** 2) * 0.02 * 0.03) - 1) < 1e-14
    assert O.unnormalized_match(a, b) == np.conj(O.unnormalized_match(b, a))
    assert abs(O.unnormalized_match(a, b) - np.sum(a * np.conj(b)) * 1e-4) < 1e-13


def test_dance_intro_bfield_and_velocity():
    """src/boost_fractor_recurprop.jl:157-166: the default B field is the real aperture mask;
    a velocity multiplies the B field by exp(1im*Ma*v*x) in place; with the real default the
    reference's in-place multiplication throws InexactError."""
    c = B.SeedCoordinateSystem()
    oc = ocoords(c)
    sb = B.SeedSetupBoundaries(c, diskno=2)
    nd = len(sb.eps)
    XX, YY = np.meshgrid(c.X, c.Y)
    mask = np.where(XX ** 2 + YY ** 2 > 0.15 ** 2, 0.0, 1.0).astype(complex)
    f0 = O.dance_intro(sb.eps, sb.r[0], oc, 0.15)
    f1, _ = O.dance_intro_ex(sb.eps, sb.r[0], oc, 0.15, bfield=np.broadcast_to(mask, (nd + 1,) + mask.shape).copy())
    assert np.array_equal(f0, f1, equal_nan=True)
    bf = (np.broadcast_to(mask, (nd + 1,) + mask.shape) * (1 + 0.5j)).copy()
    f2, b2 = O.dance_intro_ex(sb.eps, sb.r[0], oc, 0.15, bfield=bf, velocity_x=1e-3, f=20e9)
    ph = np.exp(1j * (2 * math.pi / (299792458.0 / 20e9)) * 1e-3 * XX)
    assert np.max(np.abs(b2 - bf * ph[None])) < 1e-15
    ok = np.isfinite(f0[:, 0])
    assert np.max(np.abs((f2[:, 0] - f0[:, 0] * (1 + 0.5j) * ph[None])[ok])) < 1e-15
    with pytest.raises(ValueError):
        O.dance_intro_ex(sb.eps, sb.r[0], oc, 0.15, velocity_x=1e-3)
