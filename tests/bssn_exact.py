"""Exact BSSN data for the oracle pins, by forward-mode differentiation.

A ``Jet`` carries a value, its gradient and its Hessian at one point, so
derivatives are exact to round-off (no finite differences).  Builders:

* Schwarzschild in Kerr-Schild coordinates: stationary, so with its own
  lapse and shift every evolution equation except the gauge ones has
  d_t = 0, and the constraints vanish;
* the harmonic gauge wave of Alcubierre et al. (2004), ds^2 = -H dt^2 +
  H dx^2 + dy^2 + dz^2, H = 1 - A sin(2 pi (x - t) / d), with analytic time
  derivatives (t enters through the phase only, so d_t = -d_x).

Each returns the 24 state values, first derivatives, the Hessians of the
11 fields that need one, the Christoffel symbols Gt^i_jk (word order
xx..zz), and for the gauge wave the exact d_t of every field.
"""
import math

import numpy as np

SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
HESS_VARS = [0, 1, 2, 3, 4, 5, 6, 17, 18, 19, 20]


class Jet:
    """Value, gradient (3) and Hessian (3x3, or None once a derivative has
    been taken)."""

    __slots__ = ("v", "g", "h")

    def __init__(self, v, g=None, h=None):
        self.v = float(v)
        self.g = np.zeros(3) if g is None else np.asarray(g, dtype=np.float64)
        self.h = h

    @staticmethod
    def const(c):
        return Jet(c, np.zeros(3), np.zeros((3, 3)))

    @staticmethod
    def coord(i, x):
        g = np.zeros(3)
        g[i] = 1.0
        return Jet(x, g, np.zeros((3, 3)))

    def _lift(o):
        return o if isinstance(o, Jet) else Jet.const(o)

    def __add__(a, b):
        b = Jet._lift(b)
        return Jet(a.v + b.v, a.g + b.g, None if a.h is None or b.h is None else a.h + b.h)

    __radd__ = __add__

    def __neg__(a):
        return Jet(-a.v, -a.g, None if a.h is None else -a.h)

    def __sub__(a, b):
        return a + (-Jet._lift(b))

    def __rsub__(a, b):
        return Jet._lift(b) + (-a)

    def __mul__(a, b):
        b = Jet._lift(b)
        h = None
        if a.h is not None and b.h is not None:
            h = a.v * b.h + b.v * a.h + np.outer(a.g, b.g) + np.outer(b.g, a.g)
        return Jet(a.v * b.v, a.v * b.g + b.v * a.g, h)

    __rmul__ = __mul__

    def unary(a, f, f1, f2):
        h = None if a.h is None else f1 * a.h + f2 * np.outer(a.g, a.g)
        return Jet(f, f1 * a.g, h)

    def inv(a):
        return a.unary(1 / a.v, -1 / a.v ** 2, 2 / a.v ** 3)

    def __truediv__(a, b):
        return a * Jet._lift(b).inv()

    def __rtruediv__(a, b):
        return Jet._lift(b) * a.inv()

    def __pow__(a, p):
        return a.unary(a.v ** p, p * a.v ** (p - 1), p * (p - 1) * a.v ** (p - 2))

    def sqrt(a):
        return a ** 0.5

    def sin(a):
        return a.unary(math.sin(a.v), math.cos(a.v), -math.sin(a.v))

    def d(a, k):
        """d_k, one order lower."""
        if a.h is None:
            raise ValueError("derivative of a first-order jet")
        return Jet(a.g[k], a.h[k], None)


def mat_inv(m):
    a, b, c = m[0]
    d, e, f = m[1]
    g, h, i = m[2]
    det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g)
    cof = [[e * i - f * h, c * h - b * i, b * f - c * e],
           [f * g - d * i, a * i - c * g, c * d - a * f],
           [d * h - e * g, b * g - a * h, a * e - b * d]]
    return [[cof[r][s] / det for s in range(3)] for r in range(3)], det


def _bssn_from_adm(gamma, K, alpha, beta_up):
    """gamma, alpha, beta: second-order jets; K: first-order jets."""
    gi, det = mat_inv(gamma)
    chi = det ** (-1.0 / 3.0)
    gt = [[chi * gamma[i][j] for j in range(3)] for i in range(3)]
    Ktr = sum((gi[i][j] * K[i][j] for i in range(3) for j in range(3)), Jet(0.0))
    At = [[chi * (K[i][j] - gamma[i][j] * Ktr / 3.0) for j in range(3)] for i in range(3)]
    gtu, _ = mat_inv(gt)
    Gl = [[[0.5 * (gt[i][j].d(k) + gt[i][k].d(j) - gt[j][k].d(i)) for k in range(3)] for j in range(3)]
          for i in range(3)]
    G = [[[sum((gtu[i][l] * Gl[l][j][k] for l in range(3)), Jet(0.0)) for k in range(3)] for j in range(3)]
         for i in range(3)]
    Gam = [sum((gtu[j][k] * G[i][j][k] for j in range(3) for k in range(3)), Jet(0.0)) for i in range(3)]
    zero = Jet.const(0.0)
    state = [chi] + [gt[a][b] for a, b in SYM] + [Ktr] + [At[a][b] for a, b in SYM] + Gam + [alpha] + \
        list(beta_up) + [zero, zero, zero]
    u = np.array([s.v for s in state])
    d1 = np.array([s.g for s in state])
    dd = np.array([[state[v].h[a, b] for a, b in SYM] for v in HESS_VARS])
    Gw = np.array([[G[i][a][b].v for i in range(3)] for a, b in SYM])
    return u, d1, dd, Gw


def kerr_schild(point, M=1.0):
    x = [Jet.coord(i, point[i]) for i in range(3)]
    r = (x[0] * x[0] + x[1] * x[1] + x[2] * x[2]).sqrt()
    H = M / r
    l = [x[i] / r for i in range(3)]
    delta = lambda i, j: 1.0 if i == j else 0.0
    gamma = [[delta(i, j) + 2 * H * l[i] * l[j] for j in range(3)] for i in range(3)]
    alpha = (1 + 2 * H) ** -0.5
    beta_dn = [2 * H * l[i] for i in range(3)]
    gi = [[delta(i, j) - 2 * H * l[i] * l[j] / (1 + 2 * H) for j in range(3)] for i in range(3)]
    beta_up = [sum((gi[i][j] * beta_dn[j] for j in range(3)), Jet(0.0, None, np.zeros((3, 3)))) for i in range(3)]
    # K_ij = (D_i b_j + D_j b_i) / (2 alpha), Gamma^k_ij of the physical metric
    dg = [[[gamma[i][j].d(k) for k in range(3)] for j in range(3)] for i in range(3)]
    Gam = [[[sum((gi[k][m] * (dg[m][i][j] + dg[m][j][i] - dg[i][j][m]) for m in range(3)), Jet(0.0)) * 0.5
             for j in range(3)] for i in range(3)] for k in range(3)]
    Db = [[beta_dn[j].d(i) - sum((Gam[k][i][j] * beta_dn[k] for k in range(3)), Jet(0.0)) for j in range(3)]
          for i in range(3)]
    K = [[(Db[i][j] + Db[j][i]) / (2 * alpha) for j in range(3)] for i in range(3)]
    return _bssn_from_adm(gamma, K, alpha, beta_up)


def gauge_wave(point, A=0.1, d=1.0):
    """State at t = 0 and its exact d_t (= -d_x for every field)."""
    x = [Jet.coord(i, point[i]) for i in range(3)]
    Hf = 1 - A * (2 * math.pi / d * x[0]).sin()
    one, zero = Jet.const(1.0), Jet.const(0.0)
    gamma = [[Hf, zero, zero], [zero, one, zero], [zero, zero, one]]
    alpha = Hf.sqrt()
    # K_xx = -d_t gamma_xx / (2 alpha) with d_t H = -d_x H
    K = [[Hf.d(0) / (2 * alpha) if (i, j) == (0, 0) else Jet(0.0) for j in range(3)] for i in range(3)]
    u, d1, dd, Gw = _bssn_from_adm(gamma, K, alpha, [zero, zero, zero])
    return u, d1, dd, Gw, -d1[:, 0]
