"""GPU parity of the axion-source options and the antenna helpers:
dance_intro with a B field and an axion velocity (src/boost_fractor_recurprop.jl:153-201),
gauss_profile (src/beam_gaussian.jl:24-37), unnormalized_match (src/beam_coupling.jl:13-15)."""
import math
import types

import numpy as np
import pytest

import BoostFractor as B
from BoostFractor import workloads as W
from helpers import DISKR, GOLDEN, O, ocoords, osetup, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10


def test_gauss_profile_matches_oracle_and_golden():
    c = B.SeedCoordinateSystem()
    g = B.gauss_profile(c, z=1e-9, omega0=0.097, f=22e9)
    gd = GOLDEN["gauss_profile_4000"]
    assert abs(g.ravel()[3999] - complex(gd["re"], gd["im"])) < 1e-15
    for z, w0, f in [(1e-9, 0.1, 10e9), (0.3, 0.05, 20e9), (-0.2, 0.12, 18e9)]:
        c2 = W.grid_coords(64)
        assert rel(B.gauss_profile(c2, z=z, omega0=w0, f=f), O.gauss_profile(ocoords(c2), z, w0, f)) < 1e-13


def test_unnormalized_match_matches_oracle():
    rng = np.random.default_rng(7)
    ant = rng.standard_normal((48, 40)) + 1j * rng.standard_normal((48, 40))
    E = rng.standard_normal((5, 48, 40)) + 1j * rng.standard_normal((5, 48, 40))
    got = B.unnormalized_match(E, ant, dx=0.02, dy=0.025)
    assert got.shape == (5,)
    for q in range(5):
        want = O.unnormalized_match(E[q], ant, 0.02, 0.025)
        assert abs(got[q] - want) < 1e-12 * abs(want)
    one = B.unnormalized_match(E[0], ant, dx=0.02, dy=0.025)
    assert abs(one - got[0]) < 1e-15 * abs(got[0])
    with pytest.raises(B.DimensionMismatch):
        B.unnormalized_match(E, ant[:, :39])


def test_dance_intro_bfield_velocity_matches_oracle():
    c = W.grid_coords(32)
    sb = W.booster(c, ndisk=2, seed=1200)
    nd = len(sb.distance)
    rng = np.random.default_rng(8)
    bf = rng.standard_normal((nd + 1, 32, 32)) + 1j * rng.standard_normal((nd + 1, 32, 32))
    # B field, no velocity: the caller's array is left alone
    keep = bf.copy()
    got = B.dance_intro(sb, c, bfield=bf, diskR=DISKR)
    want, _ = O.dance_intro_ex(sb.eps, sb.r[0], ocoords(c), DISKR, bfield=keep)
    assert np.array_equal(bf, keep)
    ok = np.isfinite(want)
    assert np.array_equal(np.isfinite(got), ok) and rel(got[ok], want[ok]) < TOL
    # with a velocity the B field is multiplied in place, as in the reference
    got = B.dance_intro(sb, c, bfield=bf, velocity_x=2e-3, f=20.3e9, diskR=DISKR)
    want, bwant = O.dance_intro_ex(sb.eps, sb.r[0], ocoords(c), DISKR, bfield=keep, velocity_x=2e-3, f=20.3e9)
    assert rel(bf, bwant) < 1e-14
    assert rel(got[ok], want[ok]) < TOL
    # a real B field cannot hold the phase
    with pytest.raises(B.InexactError):
        B.dance_intro(sb, c, bfield=np.ones((nd + 1, 32, 32)), velocity_x=1e-3)
    # the default B field is the aperture mask
    d0 = B.dance_intro(sb, c, diskR=DISKR)
    w0 = O.dance_intro(sb.eps, sb.r[0], ocoords(c), DISKR)
    assert np.array_equal(d0, w0, equal_nan=True)


def test_dancer_with_velocity_source():
    """the velocity-shifted source as `emit`, built on the setup the dancer itself
    passes to dance_intro: mirror region dropped, r[1] kept (recurprop.jl:28-34)"""
    c = W.grid_coords(32)
    sb = W.booster(c, ndisk=2, seed=1201)
    trimmed = types.SimpleNamespace(distance=sb.distance[1:], eps=sb.eps[1:], r=sb.r)
    nd = len(trimmed.distance)
    XX, YY = np.meshgrid(c.X, c.Y)
    mask = np.where(XX ** 2 + YY ** 2 > DISKR ** 2, 0.0, 1.0).astype(complex)
    bf = np.broadcast_to(mask, (nd + 1,) + mask.shape).copy()
    emit = B.dance_intro(trimmed, c, bfield=bf, velocity_x=1e-3, f=20e9, diskR=DISKR)
    assert np.all(np.isfinite(emit))
    got = B.dancer(0, 15, sb, c, f=20e9, emit=emit, diskR=DISKR)
    want = O.dancer(osetup(sb), ocoords(c), 0, 15, 20e9, emit=emit, diskR=DISKR)
    assert rel(got, want) < TOL
    plain = B.dancer(0, 15, sb, c, f=20e9, diskR=DISKR)
    assert rel(got, plain) > 1e-6  # the velocity changes the emitted field
