"""The FP64 scalar-wave oracle pinned by what the mathematics fixes.

* Time integrator: for a single Fourier mode the semi-discrete system is
  y' = A y with A = [[0, 1], [mu, 0]] (mu the stencil's eigenvalue on that
  mode), so n RK4 steps multiply (u, v) by the closed-form matrix
  (sum_{j<=4} (dt A)^j / j!)^n.  A wrong stage weight fails this.
* Stencils: the error against the continuum solution cos(|k| t) cos(k.x)
  falls at order 2 (q = 1) and 4 (q = 2) under refinement at fixed dt/h.  A
  wrong coefficient breaks the order.
* stage_from_words chained over the four stages equals one rk4_uniform step.
"""
import math

import numpy as np
import pytest

from oracle import swave as W


def _mode(N, kvec):
    h = 1.0 / N
    x = (np.arange(N) + 0.5) * h
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    return np.cos(2 * np.pi * (kvec[0] * xx + kvec[1] * y + kvec[2] * z)), h


@pytest.mark.parametrize("q", [1, 2, 3])
def test_rk4_matches_closed_form_amplification(q):
    N, kvec, steps = 16, (1, 2, 3), 37
    u0, h = _mode(N, kvec)
    dt = 0.25 * h
    lap = W.laplacian(u0, h, q)
    mu = float(np.sum(lap * u0) / np.sum(u0 * u0))
    assert np.allclose(lap, mu * u0, atol=1e-9 * abs(mu))  # a Fourier mode is an eigenvector
    A = np.array([[0.0, 1.0], [mu, 0.0]]) * dt
    M = sum(np.linalg.matrix_power(A, j) / math.factorial(j) for j in range(5))
    Mn = np.linalg.matrix_power(M, steps)
    u, v, _ = W.rk4_uniform(u0, np.zeros_like(u0), h, dt, steps, q)
    assert np.max(np.abs(u - Mn[0, 0] * u0)) < 1e-11
    assert np.max(np.abs(v - Mn[1, 0] * u0)) < 1e-9 * abs(mu) ** 0.5


@pytest.mark.parametrize("q,order", [(1, 2), (2, 4)])
def test_convergence_order(q, order):
    kvec, T = (1, 1, 0), 0.25
    errs = []
    for N in (16, 32):
        u0, h = _mode(N, kvec)
        dt = 0.125 * h
        steps = int(round(T / dt))
        u, _, _ = W.rk4_uniform(u0, np.zeros_like(u0), h, dt, steps, q)
        w = 2 * np.pi * math.sqrt(sum(k * k for k in kvec))
        errs.append(np.max(np.abs(u - math.cos(w * steps * dt) * u0)))
    rate = math.log2(errs[0] / errs[1])
    assert abs(rate - order) < 0.3, (errs, rate)


def test_energy_of_a_mode():
    """E = 1/2 |grad u|^2 integrated: for cos(2 pi k.x) on the unit cube the
    continuum value is (1/2)(2 pi)^2 |k|^2 <sin^2> = pi^2 |k|^2; the order-4
    gradient is within O(h^4) of it."""
    kvec = (1, 0, 1)
    u0, h = _mode(32, kvec)
    e = W.energy(u0, np.zeros_like(u0), h, 2)
    assert abs(e - np.pi ** 2 * 2) < 1e-3 * e


def test_stages_chain_to_one_step():
    rng = np.random.default_rng(1)
    N = 8
    h, dt, q = 1.0 / N, 0.25 / N, 2
    u = rng.standard_normal((N, N, N))
    v = rng.standard_normal((N, N, N))
    ref_u, ref_v, _ = W.rk4_uniform(u, v, h, dt, 1, q)
    old = (u.ravel(), v.ravel())
    state, acc = old, (np.zeros(u.size), np.zeros(u.size))
    for s in range(4):
        lap = W.laplacian(state[0].reshape(u.shape), h, q).ravel()
        acc, new = W.stage_from_words(s, state, old, acc, lap, dt)
        state = new
    assert np.allclose(state[0], ref_u.ravel(), atol=1e-12)
    assert np.allclose(state[1], ref_v.ravel(), atol=1e-12)
