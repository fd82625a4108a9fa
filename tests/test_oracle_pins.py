"""Pins of the oracle's option paths that the reference's goldens leave open
(CPU only).  Each check is fixed by mathematics independent of the oracle's
code, chosen so that a plausible slip fails it:

* tilt and surface phases (src/boost_fractor.jl:145-147): at dz = 0 with no
  aperture the propagator is the pointwise phase exp(-i k0 (tx x + ty y)) *
  exp(-i k0 s), checked on a non-square grid (an x/y swap fails it);
* propagatorMomentumSpace (:187-210): identities that tie it to numpy's FFT
  and to the golden-pinned propagatorNoTilts: phase before mask (a mask
  applied in the wrong domain fails), tilt in real space after the mask;
* the reflect and left-output paths of cheerleader / dancer / transformer
  (recurprop.jl:53-63,241-244,358-362; transformer.jl:356-357): a 1D stack
  against the effective-reflection (Parratt/Airy) recursion built from the
  Fresnel coefficients, with energy conservation for lossless stacks;
* linearity in emit and reflect; the dancer's history output and its reverse
  summation (recurprop.jl:46-50,141-145); the amin stop (:71,128-130)
  against a brute-force 1D bounce written as matrix products;
* dancer (Jacobi) against cheerleader (Gauss-Seidel) at convergence on a
  tilted, rough 3D booster (docs/src/3d_algorithms.md:67): the dancer's
  output is the cheerleader's behind the aperture mask;
* the axion-velocity phase of axion_induced_modes (transformer.jl:127-132):
  a pattern carrying the same phase projects to sqrt(#disk pixels) exactly.
"""
import math

import numpy as np
import pytest

import BoostFractor as B
from helpers import DISKR, O, ocoords, osetup, rel

C0 = 299792458.0


def _grid(nx, ny, span=1.0):
    X = (np.arange(nx) - nx // 2) * (span / nx)
    Y = (np.arange(ny) - ny // 2) * (span / ny)
    return O.Coords(X, Y, B.get_kspace_coords(X), B.get_kspace_coords(Y))


def _field(ny, nx, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((ny, nx)) + 1j * rng.standard_normal((ny, nx))


def _k0(lam, eps):
    e = complex(eps)
    if e.imag == 0.0:
        e = complex(e.real, -0.0)
    return 2 * math.pi / lam * np.sqrt(e)


def _fwd(E):
    """Julia's fftshift(fft(E)) for E[x, y] stored as (ny, nx)"""
    return np.fft.fftshift(np.fft.fft2(E))


def _inv(Eh):
    return np.fft.ifft2(np.fft.ifftshift(Eh))


# ---- tilt and surface -------------------------------------------------------
@pytest.mark.parametrize("eps", [complex(1.0, -0.0), complex(24.0, -0.0), complex(2.5, -0.03)])
def test_tilt_surface_phase_closed_form(eps):
    """dz = 0, aperture wider than the grid: propagator = E * exp(-i k0 (tx x + ty y)) * exp(-i k0 s)
    on a 24 x 40 grid (x and y of different lengths and spans)"""
    oc = _grid(24, 40, span=0.8)
    E = _field(40, 24, 1)
    rng = np.random.default_rng(2)
    s = 1e-5 * (rng.standard_normal((40, 24)) + 0.3j * rng.standard_normal((40, 24)))
    lam, tx, ty = C0 / 20e9, 3e-4, -7e-4
    k0 = _k0(lam, eps)
    want = E * np.exp(-1j * k0 * (tx * oc.X[None, :] + ty * oc.Y[:, None])) * np.exp(-1j * k0 * s)
    got = O.propagate(O.PROP_PROPAGATOR, E, oc, 0.0, 10.0, eps, tx, ty, s, lam)
    assert rel(got, want) < 1e-13
    # the same tilt with x and y exchanged is a different field: the check sees orientation
    swapped = E * np.exp(-1j * k0 * (ty * oc.X[None, :] + tx * oc.Y[:, None])) * np.exp(-1j * k0 * s)
    assert rel(got, swapped) > 1e-3
    # propagatorNoTilts ignores tilt and surface
    got = O.propagate(O.PROP_NOTILTS, E, oc, 0.0, 10.0, eps, tx, ty, s, lam)
    assert rel(got, E) < 1e-13


def test_tilt_surface_after_diffraction():
    """dz != 0, finite aperture: propagator = (propagatorNoTilts result) * tilt * surface phase,
    the phases applied after the propagation (src/boost_fractor.jl:143-147)"""
    oc = _grid(32, 48, span=1.0)
    E = _field(48, 32, 3)
    s = 2e-5 * _field(48, 32, 4)
    lam, eps = C0 / 19.5e9, complex(1.0, -0.0)
    k0 = _k0(lam, eps)
    nt = O.propagate(O.PROP_NOTILTS, E, oc, 0.013, 0.3, eps, 0, 0, None, lam)
    ph = np.exp(-1j * k0 * (2e-4 * oc.X[None, :] + 5e-4 * oc.Y[:, None])) * np.exp(-1j * k0 * s)
    got = O.propagate(O.PROP_PROPAGATOR, E, oc, 0.013, 0.3, eps, 2e-4, 5e-4, s, lam)
    assert rel(got, nt * ph) < 1e-13


# ---- propagatorMomentumSpace ------------------------------------------------
def test_momentum_space_identities():
    """E held in k-space: M(Eh) = fftshift(fft(mask * tilt * ifft(ifftshift(Eh * phase)))).
    (1) no aperture, dz != 0: M(fftshift(fft E)) = fftshift(fft(propagatorNoTilts(E)));
    (2) finite aperture, dz = 0: M(fftshift(fft E)) = fftshift(fft(mask E)) (mask strict <);
    (3) finite aperture, dz != 0: the phase acts before the mask, so
        M(fftshift(fft E)) = fftshift(fft(mask * propagatorNoTilts_noaperture(E)));
    (4) tilt: applied in real space after the mask, before the forward transform."""
    oc = _grid(20, 36, span=1.0)
    E = _field(36, 20, 5)
    lam, eps = C0 / 20e9, complex(1.0, -0.0)
    XX, YY = np.meshgrid(oc.X, oc.Y)
    R = 0.31
    mask = (XX ** 2 + YY ** 2 < R ** 2).astype(float)
    Eh = _fwd(E)
    big = 10.0
    # (1)
    got = O.propagate(O.PROP_MOMENTUM, Eh, oc, 0.013, big, eps, 0, 0, None, lam)
    want = _fwd(O.propagate(O.PROP_NOTILTS, E, oc, 0.013, big, eps, 0, 0, None, lam))
    assert rel(got, want) < 1e-12
    # (2)
    got = O.propagate(O.PROP_MOMENTUM, Eh, oc, 0.0, R, eps, 0, 0, None, lam)
    assert rel(got, _fwd(mask * E)) < 1e-12
    # (3)
    got = O.propagate(O.PROP_MOMENTUM, Eh, oc, 0.013, R, eps, 0, 0, None, lam)
    want = _fwd(mask * O.propagate(O.PROP_NOTILTS, E, oc, 0.013, big, eps, 0, 0, None, lam))
    assert rel(got, want) < 1e-12
    wrong = _fwd(O.propagate(O.PROP_NOTILTS, E, oc, 0.013, R, eps, 0, 0, None, lam))  # mask before the phase
    assert rel(got, wrong) > 1e-3
    # (4)
    k0 = _k0(lam, eps)
    tilt = np.exp(-1j * k0 * (4e-4 * XX - 6e-4 * YY))
    got = O.propagate(O.PROP_MOMENTUM, Eh, oc, 0.0, R, eps, 4e-4, -6e-4, None, lam)
    assert rel(got, _fwd(tilt * mask * E)) < 1e-12


# ---- reflect paths against the effective-reflection recursion ---------------
def _parratt(eps, dist, lam, r):
    """A wave incident from the last region onto a stack (regions 1..N, boundary
    i between regions i and i+1 with reflectivity r_i for a right-moving wave,
    -r_i for a left-moving one, transmissivities 1 + r_i and 1 - r_i; region j
    adds the phase exp(-i k0 sqrt(eps_j) d_j) per crossing).  Returns the total
    reflection (in the last region, at its boundary) and the transmission into
    region 1 (at boundary 1), by the effective-reflection recursion
        G_2 = -r_1,  T_2 = 1 - r_1,
        G_{j+1} = -r_j + (1 - r_j)(1 + r_j) p_j^2 G_j / (1 - r_j p_j^2 G_j),
        T_{j+1} = (1 - r_j) p_j T_j / (1 - r_j p_j^2 G_j)."""
    N = len(eps)
    G, T = -r[0], 1 - r[0]
    for j in range(2, N):  # region j (1-based) between boundaries j-1 and j
        p = np.exp(-1j * _k0(lam, eps[j - 1]) * dist[j - 1])
        rj = r[j - 1]
        den = 1 - rj * p * p * G
        G, T = -rj + (1 - rj) * (1 + rj) * p * p * G / den, (1 - rj) * p * T / den
    return G, T


def _stack(eps, dist):
    """oracle Setup with r from the permittivities (also at boundary 1: a transmissive stack)"""
    eps = np.array(eps, dtype=complex)
    eps[eps.imag == 0] = eps[eps.imag == 0].real - 0.0j
    sq = np.sqrt(eps)
    r = (sq[:-1] - sq[1:]) / (sq[:-1] + sq[1:])
    return O.Setup(np.array(dist, dtype=float), eps, r), r


def _one_pixel():
    z = np.array([1e-7])
    return O.Coords(z, z, np.array([0.0]), np.array([0.0]))


STACKS = [
    # transmissive slab, lossless and lossy; two disks; a lossy gap
    ([1, 24, 1], [0, 1e-3, 0]),
    ([1, 24 - 0.4j, 1], [0, 1e-3, 0]),
    ([1, 9, 1, 24, 1], [0, 1.3e-3, 7.9e-3, 1e-3, 0]),
    ([2.0, 1, 9.4, 1 - 1e-3j, 1], [0, 8.1e-3, 1e-3, 6e-3, 0]),
]


@pytest.mark.parametrize("eps,dist", STACKS)
def test_cheerleader_reflect_and_transmission_closed_form(eps, dist):
    """cheerleader(reflect=E_in; returnboth=true): right output = G E_in, left output
    = T E_in (1D propagator, one pixel), converged; lossless stacks conserve
    energy: |G|^2 + (n_1/n_N) |T|^2 = 1"""
    st, r = _stack(eps, dist)
    oc = _one_pixel()
    f = 20.3e9
    lam = C0 / f
    Ein = np.array([[0.7 - 0.2j]])
    left, right = O.cheerleader(st, oc, 0, 400, f, O.PROP_1D, reflect=Ein, diskR=DISKR, returnboth=True)
    G, T = _parratt(st.eps, st.distance, lam, r)
    assert abs(right[0, 0] - G * Ein[0, 0]) < 1e-12
    assert abs(left[0, 0] - T * Ein[0, 0]) < 1e-12
    if all(np.imag(eps) == 0):
        n1, nN = math.sqrt(eps[0].real if isinstance(eps[0], complex) else eps[0]), 1.0
        assert abs(abs(G) ** 2 + n1 / nN * abs(T) ** 2 - 1) < 1e-12


@pytest.mark.parametrize("eps_disk", [24.0, 24 - 0.3j])
def test_mirror_reflection_all_solvers(eps_disk):
    """a mirror-backed 2-disk booster (r_1 = 1): cheerleader and dancer reflect
    outputs, and the transformer's refl (one mode, eps 1e30 mirror), against the
    recursion; a lossless booster reflects everything (|G| = 1)"""
    eps = np.array([np.nan, 1, eps_disk, 1, eps_disk, 1], dtype=complex)
    dist = np.array([0, 7.9e-3, 1e-3, 8.2e-3, 1e-3, 0])
    eps[eps.imag == 0] = eps[eps.imag == 0].real - 0.0j
    _, r = O.seed_reflectivities(eps)
    st = O.Setup(dist, eps, r)
    oc = _one_pixel()
    f = 19.8e9
    lam = C0 / f
    G, _ = _parratt(eps, dist, lam, r)
    if np.imag(eps_disk) == 0:
        assert abs(abs(G) - 1) < 1e-13
    Ein = np.array([[1.0 + 0.5j]])
    ch = O.cheerleader(st, oc, 0, 2000, f, O.PROP_1D, reflect=Ein, diskR=DISKR)
    assert abs(ch[0, 0] / Ein[0, 0] - G) < 1e-10
    da = O.dancer(st, oc, 0, 4000, f, O.PROP_1D, reflect=Ein, diskR=DISKR)
    assert abs(da[0, 0] / Ein[0, 0] - G) < 1e-10
    et = eps.copy()
    et[0] = 1e30
    Gt, _ = _parratt(et, dist, lam, np.r_[(1e15 - 1) / (1e15 + 1), r[1:]])
    b, refl = O.transformer(O.Setup(dist, et), oc, np.ones((1, 1, 1), dtype=complex), np.zeros(1, dtype=complex), f,
                            O.PROP_1D, diskR=DISKR, emit=np.ones(1, dtype=complex), reflect=np.ones(1, dtype=complex))
    assert abs(refl[0] - Gt) < 1e-10
    if np.imag(eps_disk) == 0:
        assert abs(abs(refl[0]) - 1) < 1e-12


# ---- linearity, history, amin ------------------------------------------------
def _small_3d(N=24, ndisk=2, seed=7):
    """mirror-backed booster with tilted, rough inner regions (the mirror region
    and the outer region, distance 0, are flat and untilted)"""
    X = (np.arange(N) - N // 2) / N
    c = B.CoordinateSystem(X, X, B.get_kspace_coords(X), B.get_kspace_coords(X))
    rng = np.random.default_rng(seed)
    nreg = 2 * ndisk + 2
    eps = [np.nan] + [1.0 if i % 2 == 0 else 24.0 for i in range(2 * ndisk)] + [1.0]
    dist = [0.0] + [8e-3 + 2e-4 * rng.standard_normal() if i % 2 == 0 else 1e-3 for i in range(2 * ndisk)] + [0.0]
    inner = np.r_[0.0, np.ones(nreg - 2), 0.0]
    sb = B.SeedSetupBoundaries(c, diskno=ndisk, distance=dist, epsilon=eps,
                               relative_tilt_x=1e-4 * rng.standard_normal(nreg) * inner,
                               relative_tilt_y=1e-4 * rng.standard_normal(nreg) * inner,
                               relative_surfaces=1e-5 * rng.standard_normal((nreg, N, N)) * inner[:, None, None]
                               + 0j)
    return c, sb


def test_solvers_are_linear_in_emit_and_reflect():
    c, sb = _small_3d()
    os_, oc = osetup(sb), ocoords(c)
    R = len(sb.distance) - 1
    rng = np.random.default_rng(11)
    e1, e2 = (rng.standard_normal((R, 2, 24, 24)) + 1j * rng.standard_normal((R, 2, 24, 24)) for _ in range(2))
    a, b = 0.6 - 1.1j, -0.3 + 0.2j
    d = [O.dancer(os_, oc, 0, 12, 20e9, emit=e, diskR=DISKR) for e in (e1, e2, a * e1 + b * e2)]
    assert rel(d[2], a * d[0] + b * d[1]) < 1e-12
    E1, E2 = _field(24, 24, 12), _field(24, 24, 13)
    d = [O.dancer(os_, oc, 0, 12, 20e9, reflect=e, diskR=DISKR) for e in (E1, E2, a * E1 + b * E2)]
    assert rel(d[2], a * d[0] + b * d[1]) < 1e-12
    ce1, ce2 = (rng.standard_normal((R + 1, 2, 24, 24)) + 0j for _ in range(2))
    for kw in [dict(emit=x) for x in (ce1, ce2, a * ce1 + b * ce2)], [dict(reflect=x) for x in (E1, E2, a * E1 + b * E2)]:
        lr = [O.cheerleader(os_, oc, 0, 6, 20e9, diskR=DISKR, returnboth=True, **k) for k in kw]
        for side in range(2):
            assert rel(lr[2][side], a * lr[0][side] + b * lr[1][side]) < 1e-12


def test_dancer_history_and_reverse_summation():
    """returnsum=false, immediatesum=false returns one slice per bounce (nmax+1);
    returnsum=true with immediatesum=false sums them in reverse bounce order,
    bit for bit a sequential sum from the last slice; immediatesum=true
    accumulates in bounce order"""
    c, sb = _small_3d(seed=8)
    os_, oc = osetup(sb), ocoords(c)
    nmax = 9
    h = O.dancer(os_, oc, 0, nmax, 20.1e9, diskR=DISKR, returnsum=False, immediatesum=False)
    assert h.shape == (nmax + 1, 24, 24) and np.all(np.any(h != 0, axis=(1, 2))[1:])
    s = np.zeros((24, 24), dtype=complex)
    for j in reversed(range(nmax + 1)):
        s = s + h[j]
    rev = O.dancer(os_, oc, 0, nmax, 20.1e9, diskR=DISKR, returnsum=True, immediatesum=False)
    assert np.array_equal(rev, s)
    fwd = np.zeros((24, 24), dtype=complex)
    for j in range(nmax + 1):
        fwd = fwd + h[j]
    imm = O.dancer(os_, oc, 0, nmax, 20.1e9, diskR=DISKR)
    assert np.array_equal(imm, fwd)
    one = O.dancer(os_, oc, 0, nmax, 20.1e9, diskR=DISKR, returnsum=False, immediatesum=True)
    assert one.shape == (1, 24, 24) and np.array_equal(one[0], imm)


def _bounce_1d(eps, dist, r, lam):
    """the dancer's Jacobi bounce on one pixel with the 1D propagator as matrix
    products: state = (right_1, left_1, ..., right_R, left_R) after the mirror is
    trimmed (regions 2..N of the setup, r appended with 0); one bounce is
    state -> S (P state), the output of the bounce is right_R of P state"""
    R = len(eps) - 1
    rr = list(r) + [0.0]
    ph = np.array([np.exp(-1j * _k0(lam, eps[i + 1]) * dist[i + 1]) for i in range(R)])
    P = np.diag(np.repeat(ph, 2))
    S = np.zeros((2 * R, 2 * R), dtype=complex)
    S[0, 1] = -rr[0]  # boundary 1 (mirror): left-moving of region 1 reflects into its right-moving
    for i in range(1, R):  # boundary between regions i and i+1 (0-based i-1, i)
        rb = rr[i]
        S[2 * (i - 1) + 1, 2 * (i - 1)] = rb       # left-moving in i-1 after: r * right_{i-1}
        S[2 * (i - 1) + 1, 2 * i + 1] = 1 - rb     #  + (1 - r) left_i
        S[2 * i, 2 * i + 1] = -rb                  # right-moving in i after: -r left_i
        S[2 * i, 2 * (i - 1)] = 1 + rb             #  + (1 + r) right_{i-1}
    # the new left-moving field of region R is r_{R+1} * right_R = 0
    return P, S, R


def test_dancer_amin_stop_against_bounce_matrices():
    """amin > 1 stops before the first bounce (a starts at 1): zero output; with
    0 < amin <= 1 the output is the bounce-matrix sum up to and including the
    first bounce after which max |fields| < amin (recurprop.jl:71,128-130)"""
    eps = np.array([np.nan, 1, 24 - 0.5j, 1, 24 - 0.5j, 1], dtype=complex)  # lossy disks: the maximum decays
    dist = np.array([0, 7.9e-3, 1e-3, 8.2e-3, 1e-3, 0])
    eps, r = O.seed_reflectivities(eps)
    st = O.Setup(dist, eps, r)
    oc = _one_pixel()
    f = 19.0e9
    lam = C0 / f
    P, S, R = _bounce_1d(eps, dist, r, lam)
    # initial state from dance_intro (pinned by the 1D golden), then bounces
    fi = O.dance_intro(eps[1:], r[0], oc, DISKR)
    s = np.array([fi[i, q, 0, 0] for i in range(R) for q in range(2)])
    outs, maxes = [], []
    for n in range(200):
        sp = P @ s
        outs.append(sp[2 * (R - 1)])  # (1 + 0) * right_R after propagation
        s = S @ sp
        maxes.append(np.max(np.abs(s)))
    assert abs(O.dancer(st, oc, 1.5, 199, f, O.PROP_1D, diskR=DISKR)[0, 0]) == 0.0
    full = O.dancer(st, oc, 0, 199, f, O.PROP_1D, diskR=DISKR)[0, 0]
    assert abs(full - sum(outs)) < 1e-12
    # the maximum is not monotone (0.796, 0.899, 0.899, 0.801, ...): the first crossing counts
    ks = []
    for amin in (0.85, 0.79, 0.62, 0.45, 0.2, 0.05):
        k = next(n for n, m in enumerate(maxes) if m < amin)
        ks.append(k)
        got = O.dancer(st, oc, amin, 199, f, O.PROP_1D, diskR=DISKR)[0, 0]
        assert abs(got - sum(outs[:k + 1])) < 1e-13, amin
        assert abs(got - sum(outs[:k + 2])) > 1e-9  # one bounce more would differ
    assert ks[0] == 0 and len(set(ks)) == len(ks)


def test_amin_3d_stops_on_a_history_prefix():
    """3D: the amin result equals a prefix sum of the history (the stop lands after
    a whole bounce), later for smaller amin; cheerleader stops no later than nmax"""
    c, sb = _small_3d(seed=9)
    os_, oc = osetup(sb), ocoords(c)
    h = O.dancer(os_, oc, 0, 80, 20e9, diskR=DISKR, returnsum=False, immediatesum=False)
    pref = np.cumsum(h, axis=0)
    ks = []
    for amin in (0.9, 0.8, 0.7, 0.5):
        got = O.dancer(os_, oc, amin, 80, 20e9, diskR=DISKR)
        k = int(np.argmin([rel(got, p) for p in pref]))
        assert rel(got, pref[k]) < 1e-13
        ks.append(k)
    assert ks == sorted(ks) and len(set(ks)) == len(ks)
    assert np.all(O.dancer(os_, oc, 1.0000001, 80, 20e9, diskR=DISKR) == 0)


# ---- axion velocity phase ----------------------------------------------------
@pytest.mark.parametrize("v", [0.05, -0.08])
def test_field2modes_velocity_phase(v):
    """axion_induced_modes with velocity_x (transformer.jl:127-132): the source is
    B exp(i Ma v x) on the disk, normalised to power 1.  A pattern equal to that
    phase on the disk projects to sqrt(#disk pixels) exactly; the conjugate
    phase (a sign slip) or the phase along y (an axis slip) do not.  The
    velocities are large (a phase of several radians across the disk) so that
    the slips are visible."""
    oc = _grid(48, 40, span=0.4)
    f = 20e9
    Ma = 2 * math.pi * f / C0
    XX, YY = np.meshgrid(oc.X, oc.Y)
    mask = (np.sqrt(XX ** 2 + YY ** 2) <= DISKR).astype(float)
    pats = np.stack([mask * np.exp(1j * Ma * v * XX), mask * np.exp(-1j * Ma * v * XX),
                     mask * np.exp(1j * Ma * v * YY)])
    c = O.field2modes(oc, pats, None, v, DISKR, f)
    npix = mask.sum()
    assert abs(c[0] - math.sqrt(npix)) < 1e-12 * math.sqrt(npix)
    assert abs(c[1]) < 0.5 * math.sqrt(npix) and abs(c[2]) < 0.5 * math.sqrt(npix)
    # v = 0: the plain disk
    c0 = O.field2modes(oc, mask[None].astype(complex), None, 0.0, DISKR, f)
    assert abs(c0[0] - math.sqrt(npix)) < 1e-12 * math.sqrt(npix)


# ---- dancer against cheerleader ----------------------------------------------
def test_dancer_and_cheerleader_agree_at_convergence():
    """tilted, rough 3D booster: the Jacobi dancer with 2k bounces and the
    Gauss-Seidel cheerleader with k iterations converge to the same emitted
    field; the dancer propagates its last region (distance 0), so its output is
    the cheerleader's behind the strict aperture mask (docs/src/3d_algorithms.md:67)"""
    c, sb = _small_3d(N=32, seed=10)
    os_, oc = osetup(sb), ocoords(c)
    XX, YY = np.meshgrid(c.X, c.Y)
    mask = XX ** 2 + YY ** 2 < DISKR ** 2
    ch = O.cheerleader(os_, oc, 0, 600, 19.7e9, diskR=DISKR)
    da = O.dancer(os_, oc, 0, 1200, 19.7e9, diskR=DISKR)
    assert rel(da, mask * ch) < 1e-8
    # the unmasked cheerleader output differs outside the aperture
    assert rel(da, ch) > 1e-5
