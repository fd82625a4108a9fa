"""GPU parity of the 1D analytical model (src/Analytical1D.py) against the oracle.

The model's metal mirror propagates 0.1 m inside eps = 1e20, a phase of
~1e12 rad: the device forms it with numpy's operation order and reduces it by
2 pi exactly before sincos, so results agree with numpy's to ~1e-13 (the
device's own reduction alone would leave ~1e-4 rad there)."""
import numpy as np
import pytest

import BoostFractor as B
from helpers import rel
from oracle import analytic1d as A

pytestmark = pytest.mark.gpu
A1 = B.Analytical1D
TOL = 1e-12


@pytest.mark.parametrize("nd,tand,mirror,loss", [(0, 0, True, False), (5, 0, True, False), (20, 1e-4, True, False),
                                                 (4, 2e-3, False, False), (6, 0, True, True)])
def test_disk_system_matches_oracle(nd, tand, mirror, loss):
    f = np.linspace(19.5e9, 20.5e9, 101)
    rng = np.random.default_rng(nd + 10)
    sp = 7e-3 + 1e-3 * rng.uniform(0, 1, nd + 1)
    sl = rng.uniform(0, 0.02, 2 * nd + 1) if loss else None
    g, a = A1.disk_system(f, tand=tand, num_disk=nd, spacings=sp, mirror=mirror, disk_epsilon=24,
                          non_uniform_surfaceloss=sl)
    wg, wa = A.disk_system(f, A1.lossy_eps(f, 24, 0), A1.lossy_eps(f, 1.0, tand), sp, nd, mirror=mirror,
                           surfaceloss=sl)
    assert rel(g, wg) < TOL and rel(a, wa) < TOL, (rel(g, wg), rel(a, wa))


def test_disk_system_batch_and_scalar_frequency():
    f = np.linspace(21.9e9, 22.1e9, 33)
    rng = np.random.default_rng(7)
    sp = 7e-3 + 2e-4 * rng.standard_normal((9, 21))
    g, a = A1.disk_system_batch(f, sp, num_disk=20, disk_epsilon=24)
    assert g.shape == (9, 33)
    for c in range(9):
        wg, wa = A.disk_system(f, 24 + 0j, 1 + 0j, sp[c], 20)
        assert rel(g[c], wg) < TOL and rel(a[c], wa) < TOL
    g0, a0 = A1.disk_system(22e9, num_disk=20, spacings=sp[0], disk_epsilon=24)
    wg, wa = A.disk_system([22e9], 24 + 0j, 1 + 0j, sp[0], 20)
    assert np.ndim(g0) == 0 and abs(g0 - wg[0]) < TOL and abs(a0 - wa[0]) < TOL * abs(wa[0])


def test_transform_surface_in_place_axion_and_forms():
    f = np.linspace(19e9, 21e9, 17)
    g0 = 0.8 * np.exp(1j * np.linspace(0, 2, 17))
    ax = np.linspace(0.1, 0.5, 17) + 0.2j
    ax0 = ax.copy()
    got = A1.transform_surface(f, g0, axion_out=ax, eps_i=1.0, eps_m=24.0 - 0.1j, l=1e-3, surfaceloss=0.01)
    wg, wa = A.transform_surface(f, g0, ax0, 1.0, 24.0 - 0.1j, 1e-3, 0.01)
    assert rel(got, wg) < 1e-13 and rel(ax, wa) < 1e-13  # axion_out updated in place
    gz = A1.transform_surface(f, g0, eps_i=24.0, eps_m=1.0, l=8e-3, surfaceloss=0.02, USE_RAY=False)
    wz, _ = A.transform_surface(f, g0, None, 24.0, 1.0, 8e-3, 0.02, ray=False)
    assert rel(gz, wz) < 1e-13
    with pytest.raises(B.ArgumentError):
        A1.transform_surface(f, g0, axion_out=ax, USE_RAY=False)


def test_phase_depths_matches_oracle():
    d_air = np.linspace(2.5, 3.5, 11)
    g, a = A1.disk_system_phase_depths(d_air, num_disk=5)
    wavel = A.c / 20e9
    for i, d in enumerate(d_air):
        l_air = d * wavel / (2 * np.pi)
        l_disk = (np.pi / 2) * wavel / (2 * np.pi * np.sqrt(9.4))
        wg, wa = A.disk_system([20e9], 9.4 + 0j, 1 + 0j, [l_air] * 6, 5, thickness=l_disk)
        assert abs(g[i] - wg[0]) < TOL and abs(a[i] - wa[0]) < TOL * max(1.0, abs(wa[0]))
