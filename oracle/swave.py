"""FP64 restatement of the scalar-wave RK4 scheme (the checker).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the smoke entry point as the
checker.  The product never imports it.

The scheme: u_tt = lap u written as u_t = v, v_t = lap u; the Laplacian is
the sum of the three centred second differences of order 2q whose
coefficients are the reference's (AMR.h:1144-1150, the same table as the FD
kernel); classical RK4 with the stage update of the RK4 driver:
  k1 = f(y), k2 = f(y + dt/2 k1), k3 = f(y + dt/2 k2), k4 = f(y + dt k3),
  y + dt/6 (k1 + 2 k2 + 2 k3 + k4).
`rk4_uniform` runs it on one periodic uniform grid in FP64 (no codec), which
is what the octree run reduces to when every leaf has one depth.
`stage_from_words` is one stage in FP64 on the GPU's own derivative words.
"""
from __future__ import annotations

import numpy as np

# centred second-difference weights by half-width q (AMR.h:1148-1150), as exact rationals
D2 = {1: [1.0, -2.0, 1.0],
      2: [-1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12],
      3: [1 / 90, -3 / 20, 3 / 2, -49 / 18, 3 / 2, -3 / 20, 1 / 90]}
D1 = {1: [-0.5, 0.0, 0.5],
      2: [1 / 12, -2 / 3, 0.0, 2 / 3, -1 / 12],
      3: [-1 / 60, 3 / 20, -3 / 4, 0.0, 3 / 4, -3 / 20, 1 / 60]}


def _stencil(u, w, axis):
    q = (len(w) - 1) // 2
    out = np.zeros_like(u)
    for i, c in enumerate(w):
        if c != 0.0:
            out += c * np.roll(u, q - i, axis=axis)  # roll by -(i - q): u[x + i - q]
    return out


def laplacian(u, h, q):
    return sum(_stencil(u, D2[q], a) for a in range(3)) / (h * h)


def gradient(u, h, q):
    return [_stencil(u, D1[q], a) / h for a in range(3)]


def energy(u, v, h, q):
    """E = 1/2 sum (v^2 + |grad u|^2) h^3 with the order-2q centred gradient."""
    g = gradient(u, h, q)
    return 0.5 * float(np.sum(v * v + g[0] ** 2 + g[1] ** 2 + g[2] ** 2)) * h ** 3


def scheme_energy(u, v, h, q):
    """1/2 sum (v^2 - u L u) h^3 with L the order-2q Laplacian: the quadratic
    form the semi-discrete system conserves (L is symmetric negative
    semi-definite on the periodic grid), which RK4 at a stable dt can only
    decrease."""
    return 0.5 * float(np.sum(v * v - u * laplacian(u, h, q))) * h ** 3


def rk4_uniform(u, v, h, dt, steps, q, energy_every=0):
    """`steps` RK4 steps on a periodic cube (arrays indexed [z, y, x]).
    Returns (u, v, energies) with the energy before the first step and after
    every `energy_every` steps when energy_every > 0."""
    u = np.array(u, np.float64)
    v = np.array(v, np.float64)
    es = [energy(u, v, h, q)] if energy_every else []

    def f(a, b):
        return b, laplacian(a, h, q)

    for n in range(steps):
        k1 = f(u, v)
        k2 = f(u + 0.5 * dt * k1[0], v + 0.5 * dt * k1[1])
        k3 = f(u + 0.5 * dt * k2[0], v + 0.5 * dt * k2[1])
        k4 = f(u + dt * k3[0], v + dt * k3[1])
        u = u + dt / 6 * (k1[0] + 2 * k2[0] + 2 * k3[0] + k4[0])
        v = v + dt / 6 * (k1[1] + 2 * k2[1] + 2 * k3[1] + k4[1])
        if energy_every and (n + 1) % energy_every == 0:
            es.append(energy(u, v, h, q))
    return u, v, es


def stage_from_words(stage, state, old, acc, lap, dt):
    """One RK4 stage on interior values in FP64: state/old/acc are (u, v)
    pairs of 1-D arrays, lap the Laplacian unpacked from the d2diag word.
    Returns (acc_out, new_out)."""
    k = (np.asarray(state[1], np.float64), np.asarray(lap, np.float64))
    acc_out, new_out = [], []
    for c in range(2):
        yn = np.asarray(state[c] if stage == 0 else old[c], np.float64)
        a = np.asarray(acc[c], np.float64)
        if stage == 0:
            acc_out.append(k[c].copy())
            new_out.append(yn + 0.5 * dt * k[c])
        elif stage in (1, 2):
            acc_out.append(a + 2 * k[c])
            new_out.append(yn + (0.5 * dt if stage == 1 else dt) * k[c])
        else:
            acc_out.append(a)
            new_out.append(yn + dt / 6 * (a + k[c]))
    return acc_out, new_out
