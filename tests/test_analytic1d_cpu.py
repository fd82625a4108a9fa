"""The oracle of the 1D analytical model (src/Analytical1D.py) pinned on the CPU:
energy conservation, the bare mirror, the equivalence of the ray and impedance
forms at lossless surfaces, and the transfer-matrix transformer of the oracle
(one mode, 1D propagator) on the same boosters."""
import numpy as np
import pytest

import BoostFractor as B
from oracle import analytic1d as A
from oracle import oracle as O


def _transformer_1d(eps_disk, eps_gap, spacings, nd, thick, f):
    """|boost|^2 of the oracle transformer on the same booster (mirror eps 1e20)"""
    dist, eps = [0.0], [1e20]
    for i in range(nd + 1):
        dist.append(spacings[i])
        eps.append(eps_gap)
        if i < nd:
            dist.append(thick)
            eps.append(eps_disk)
    eps, dist = np.array(eps, dtype=complex), np.array(dist)
    n = len(dist)
    st = O.Setup(dist, eps, None, np.zeros(n), np.zeros(n), None)
    X = np.array([0.0])
    co = O.Coords(X, X, np.array([0.0]), np.array([0.0]))
    b = O.transformer(st, co, np.ones((1, 1, 1), dtype=complex), np.zeros(1, dtype=complex), f, O.PROP_1D,
                      diskR=0.15, emit=np.ones(1, dtype=complex))
    return abs(b[0]) ** 2


@pytest.mark.parametrize("nd,eps_disk,eps_gap", [(0, 24, 1), (1, 24, 1), (3, 24 - 0.05j, 1 - 1e-4j), (5, 9.4, 1)])
def test_oracle_boost_matches_transformer(nd, eps_disk, eps_gap):
    """an independent algorithm for the same 1D physics; the 1D model's metal
    mirror (a 0.1 m slab of eps 1e20 with vacuum behind it) differs from the
    transformer's at the 1e-10 level, amplified by the resonance"""
    f = np.linspace(19.7e9, 20.3e9, 5)
    sp = np.array([8e-3 + 1e-4 * ((i * 7) % 5) for i in range(nd + 1)])
    _, a = A.disk_system(f, complex(eps_disk), complex(eps_gap), sp, nd)
    bt = np.array([_transformer_1d(complex(eps_disk), complex(eps_gap), sp, nd, 1e-3, q) for q in f])
    assert np.max(np.abs(np.abs(a) ** 2 / bt - 1)) < 1e-7


def test_oracle_lossless_and_bare_mirror():
    f = np.linspace(18e9, 22e9, 41)
    rng = np.random.default_rng(3)
    for nd in (0, 2, 7):
        sp = 8e-3 + 1e-3 * rng.uniform(-1, 1, nd + 1)
        g, a = A.disk_system(f, 24.0 + 0j, 1.0 + 0j, sp, nd)
        assert np.max(np.abs(np.abs(g) - 1)) < 1e-12  # everything is reflected
        if nd == 0:
            assert np.max(np.abs(np.abs(a) - 1)) < 1e-9  # a bare mirror emits unit amplitude


def test_oracle_ray_and_impedance_forms_agree_without_loss():
    f = np.linspace(19e9, 21e9, 7)
    g0 = np.exp(1j * np.linspace(0, 3, 7)) * 0.7
    for ei, em in ((1.0, 24.0), (24.0, 1.0), (1.0, 9.4 - 0.01j)):
        gr, _ = A.transform_surface(f, g0, None, ei, em, 3e-3, 0.0, ray=True)
        gz, _ = A.transform_surface(f, g0, None, ei, em, 3e-3, 0.0, ray=False)
        assert np.max(np.abs(gr - gz)) < 1e-13


def test_lossy_eps_closed_forms():
    assert B.Analytical1D.lossy_eps(20e9, 24.0, 1e-3) == pytest.approx(24.0 - 0.024j, abs=1e-15)
    # Svensson/Djordjevic: eps at the reference frequency 1 GHz reproduces eps (1 - i tand) to first order
    e = B.Analytical1D.lossy_eps(1e9, 9.0, 1e-3, SvenssonDjordjevic=True)
    assert abs(e.real - 9.0) < 1e-9 and abs(e.imag / 9.0 + 1e-3) < 1e-5
