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


# ------------------------------------------------------------------ the
# Gamma-driver shift and the constraint normalisation.  The Kerr-Schild and
# gauge-wave pins above leave the shift equations (fields 18-23) and the
# absolute scale of H and M open; each check below fails for a plausible
# mistake there, and asserts that it would (the altered closed form differs
# from the pinned one by far more than the tolerance).

ETA, SHIFT = 2.0, 0.75


def flat_state():
    u = np.zeros(24)
    u[[0, 1, 4, 6, 17]] = 1.0
    return u


@pytest.mark.parametrize("eta,shift,advect", [(2.0, 0.75, 1), (0.7, 0.5, 0), (3.0, 1.25, 1)])
def test_gamma_driver_flat_space(orc, eta, shift, advect):
    """Flat space, zero shift, uniform B: d_t beta^i = shift_coeff B^i and
    d_t B^i = -eta B^i exactly (d_t Gt^i = 0 there)."""
    u = flat_state()
    B = np.array([0.3, -0.2, 0.1])
    u[21:24] = B
    rhs, _ = orc.bssn_point(u, np.zeros((24, 3)), np.zeros((11, 6)), None, (eta, shift, 1e-6, advect, 0))
    assert np.array_equal(rhs[18:21], shift * B)
    assert np.array_equal(rhs[21:24], -eta * B)
    assert np.all(rhs[:18] == 0)
    assert np.min(np.abs(2 * eta * B)) > 0.1  # a flipped eta would be off by 2 eta |B|


@pytest.mark.parametrize("point", [(0.1, 0.2, 0.3), (0.37, -1.0, 0.5), (0.8, 0.0, 0.0)])
def test_gamma_driver_on_gauge_wave(orc, point):
    """Gauge wave (an exact solution with zero shift), shift frozen in
    position only (no advection): d_t B^i = d_t Gt^i - eta B^i with d_t Gt^i
    the exact one (-d_x Gt^i), and d_t beta^i = 3/4 B^i."""
    u, d1, dd, G, dt = be.gauge_wave(point, A=0.1, d=1.0)
    B = np.array([0.05, -0.02, 0.03])
    u = u.copy()
    u[21:24] = B
    rhs, _ = orc.bssn_point(u, d1, dd, None, (ETA, SHIFT, 1e-6, 0, 1))
    want = dt[14:17] - ETA * B
    scale = np.max(np.abs(dt[14:17]))
    assert scale > 0.1  # d_t Gt^x does not vanish: the coupling is tested
    assert np.allclose(rhs[21:24], want, rtol=0, atol=1e-12 * scale), (rhs[21:24], want)
    assert np.allclose(rhs[18:21], SHIFT * B, rtol=0, atol=1e-15)


def _linear_shift_case(seed):
    rng = np.random.default_rng(seed)
    L = rng.uniform(-0.5, 0.5, (3, 3))   # d_k beta^i = L[i, k]
    N = rng.uniform(-0.5, 0.5, (3, 3))   # d_k B^i
    Mg = rng.uniform(-0.5, 0.5, (3, 3))  # d_k Gt^i
    beta = rng.uniform(-0.8, 0.8, 3)
    B = rng.uniform(-0.4, 0.4, 3)
    Gam = rng.uniform(-0.3, 0.3, 3)
    u = flat_state()
    u[14:17], u[18:21], u[21:24] = Gam, beta, B
    d1 = np.zeros((24, 3))
    d1[14:17], d1[18:21], d1[21:24] = Mg, L, N
    return u, d1, L, N, Mg, beta, B, Gam


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gamma_driver_advection_closed_form(orc, seed):
    """Flat metric, unit lapse, beta, B and Gt linear in x (their second
    derivatives vanish).  The textbook system then reduces to
      d_t Gt^i = beta.d Gt^i - Gt^k d_k beta^i + 2/3 Gt^i d_k beta^k
      d_t beta^i = beta.d beta^i + 3/4 B^i
      d_t B^i = beta.d B^i + d_t Gt^i - beta.d Gt^i - eta B^i
    evaluated here in closed form.  Dropping the -beta.d Gt^i of the B
    equation, or the advection of beta or B, changes the result by more
    than 1e-2."""
    u, d1, L, N, Mg, beta, B, Gam = _linear_shift_case(seed)
    rhs, _ = orc.bssn_point(u, d1, np.zeros((11, 6)), None, (ETA, SHIFT, 1e-6, 1, 0))
    trL = np.trace(L)
    dtG = Mg @ beta - L @ Gam + (2 / 3) * Gam * trL
    want_beta = L @ beta + SHIFT * B
    want_B = N @ beta + dtG - Mg @ beta - ETA * B
    assert np.allclose(rhs[14:17], dtG, rtol=0, atol=1e-14)
    assert np.allclose(rhs[18:21], want_beta, rtol=0, atol=1e-14)
    assert np.allclose(rhs[21:24], want_B, rtol=0, atol=1e-14)
    # the test's power against the plausible mistakes
    assert np.max(np.abs(Mg @ beta)) > 1e-2      # -beta.d Gt dropped
    assert np.max(np.abs(L @ beta)) > 1e-2       # beta not advected
    assert np.max(np.abs(N @ beta)) > 1e-2       # B not advected
    # without advection the advective terms go, d_t Gt^i stays whole
    rhs0, _ = orc.bssn_point(u, d1, np.zeros((11, 6)), None, (ETA, SHIFT, 1e-6, 0, 0))
    assert np.allclose(rhs0[18:21], SHIFT * B, rtol=0, atol=1e-15)
    assert np.allclose(rhs0[21:24], dtG - ETA * B, rtol=0, atol=1e-14)


@pytest.mark.parametrize("point", [(0.3, -0.4, 0.2), (1.0, 0.5, -0.7), (-0.6, 0.9, 0.4)])
def test_hamiltonian_scale_conformally_flat(orc, point):
    """gamma = psi^4 delta, K = 0: H is the Ricci scalar -8 psi^-5 lap(psi),
    non-zero, so an overall factor on H fails."""
    gamma, K, R = be.conformally_flat(point)
    u, d1, dd, G = be.bssn_of(gamma, K)
    _, hm = orc.bssn_point(u, d1, dd, None)
    assert abs(R) > 0.05
    assert np.isclose(hm[0], R, rtol=1e-12, atol=0), (hm[0], R)
    assert np.max(np.abs(hm[1:])) < 1e-13


@pytest.mark.parametrize("seed,point", [(0, (0.3, -0.4, 0.2)), (1, (1.0, 0.5, -0.7)), (2, (-0.6, 0.9, 0.4)),
                                        (3, (0.1, 0.1, 0.1))])
def test_constraints_match_adm(orc, seed, point):
    """Generic ADM data (a polynomial metric near flat space and a linear
    extrinsic curvature, neither symmetric in any axis) converted to BSSN
    variables: the oracle's H and M_i equal the ADM Hamiltonian R + K^2 -
    K_ij K^ij and momentum constraint D_j K^j_i - D_i K computed from the
    physical metric alone (tests/bssn_exact.py).  Both are O(0.1), so a
    scale error, a dropped chi term or a wrong Christoffel index fails."""
    gamma, K = be.generic_adm(point, seed)
    H, M = be.adm_constraints(gamma, K)
    u, d1, dd, G = be.bssn_of(gamma, K)
    _, hm = orc.bssn_point(u, d1, dd, None)
    assert abs(H) > 1e-2 and np.max(np.abs(M)) > 1e-2
    assert np.isclose(hm[0], H, rtol=1e-11, atol=1e-13), (hm[0], H)
    assert np.allclose(hm[1:], M, rtol=1e-11, atol=1e-13), (hm[1:], M)
    # the same with the Christoffel words given rather than computed
    _, hm2 = orc.bssn_point(u, d1, dd, G)
    assert np.allclose(hm2, hm, rtol=1e-12, atol=1e-14)
