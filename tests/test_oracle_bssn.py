"""Pins the FP64 BSSN oracle (the scheme has no reference implementation).

* A Schwarzschild black hole in Kerr-Schild coordinates is stationary: with
  its own lapse and shift, d_t of chi, gt_ij, K, At_ij and Gt^i vanish, and
  so do the Hamiltonian and momentum constraints.  Every term of the RHS
  (curvature, Lie derivatives, lapse terms) is non-zero there, so a dropped
  term, a wrong sign or a transposed index fails this test.
* The harmonic gauge wave (Alcubierre et al. 2004) is an exact solution with
  zero shift: its analytic d_t must be reproduced, lapse included.
* Flat space is a fixed point.
Inputs are exact (forward-mode derivatives, tests/bssn_exact.py).
"""
import numpy as np
import pytest

from tests import bssn_exact as be

PHYS = list(range(17))  # chi, gt, K, At, Gt


@pytest.mark.parametrize("point", [(1.1, 0.7, -0.5), (-2.0, 0.3, 1.2), (0.4, -3.1, 2.2), (5.0, 4.0, -3.0)])
@pytest.mark.parametrize("given_christoffel", [False, True])
def test_kerr_schild_is_stationary(orc, point, given_christoffel):
    u, d1, dd, G = be.kerr_schild(point)
    rhs, hm = orc.bssn_point(u, d1, dd, G if given_christoffel else None, (2.0, 0.75, 1e-6, 1, 0))
    scale = np.max(np.abs(d1)) + 1
    assert np.max(np.abs(rhs[PHYS])) < 1e-12 * scale * 10, rhs[PHYS]
    assert np.max(np.abs(hm)) < 1e-11 * scale, hm


def test_kerr_schild_detects_perturbations(orc):
    """Sensitivity: perturbing one input derivative breaks stationarity."""
    u, d1, dd, G = be.kerr_schild((1.1, 0.7, -0.5))
    base, _ = orc.bssn_point(u, d1, dd, None)
    for v, k in [(0, 1), (3, 2), (7, 0), (12, 1), (15, 2), (17, 0), (19, 1)]:
        d = d1.copy()
        d[v, k] += 1e-3
        rhs, _ = orc.bssn_point(u, d, dd, None)
        assert np.max(np.abs(rhs[PHYS] - base[PHYS])) > 1e-5, (v, k)
    for h in range(11):
        e = dd.copy()
        e[h, (h * 5) % 6] += 1e-3
        rhs, _ = orc.bssn_point(u, d1, e, None)
        assert np.max(np.abs(rhs[PHYS] - base[PHYS])) > 1e-6, h


@pytest.mark.parametrize("point", [(0.1, 0.2, 0.3), (0.37, -1.0, 0.5), (0.8, 0.0, 0.0)])
def test_harmonic_gauge_wave(orc, point):
    u, d1, dd, G, dt = be.gauge_wave(point, A=0.1, d=1.0)
    rhs, hm = orc.bssn_point(u, d1, dd, None, (2.0, 0.75, 1e-6, 0, 1))
    idx = PHYS + [17]
    assert np.allclose(rhs[idx], dt[idx], rtol=0, atol=1e-12 * np.max(np.abs(dt))), (rhs[idx], dt[idx])
    assert np.max(np.abs(hm)) < 1e-11


def test_flat_space_fixed_point(orc):
    u = np.zeros(24)
    u[[0, 1, 4, 6, 17]] = 1.0
    rhs, hm = orc.bssn_point(u, np.zeros((24, 3)), np.zeros((11, 6)))
    assert np.all(rhs == 0) and np.all(hm == 0)


def test_one_plus_log_lapse(orc):
    """d_t alpha = b.d alpha - 2 alpha K (1+log) on Kerr-Schild data, where
    alpha is not stationary under 1+log: compare with the harmonic value
    scaled, which the gauge wave pins."""
    u, d1, dd, G = be.kerr_schild((1.1, 0.7, -0.5))
    r_log, _ = orc.bssn_point(u, d1, dd, None, (2.0, 0.75, 1e-6, 1, 0))
    r_har, _ = orc.bssn_point(u, d1, dd, None, (2.0, 0.75, 1e-6, 1, 1))
    adv = u[18:21] @ d1[17]
    alpha, K = u[17], u[7]
    assert np.isclose(r_log[17] - adv, (r_har[17] - adv) * 2 / alpha, rtol=1e-13)
    assert K != 0
