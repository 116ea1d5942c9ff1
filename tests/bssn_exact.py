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


def adm_constraints(gamma, K):
    """Hamiltonian and momentum constraints of ADM data, computed from the
    ADM definitions without any BSSN variable:
      Gamma^k_ij = gamma^kl (d_i gamma_lj + d_j gamma_li - d_l gamma_ij) / 2
      R_ij = d_k Gamma^k_ij - d_j Gamma^k_ik + Gamma^k_kl Gamma^l_ij
             - Gamma^k_jl Gamma^l_ik
      H = R + K^2 - K_ij K^ij,   M_i = D_j K^j_i - D_i K
    gamma: second-order jets; K: first-order (or better) jets.  Returns
    (H, M[3]) as floats."""
    gi, _ = mat_inv(gamma)
    G = [[[0.5 * sum((gi[k][l] * (gamma[l][j].d(i) + gamma[l][i].d(j) - gamma[i][j].d(l)) for l in range(3)),
                     Jet(0.0))
           for j in range(3)] for i in range(3)] for k in range(3)]
    R = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            r = 0.0
            for k in range(3):
                r += G[k][i][j].g[k] - G[k][i][k].g[j]
                for l in range(3):
                    r += G[k][k][l].v * G[l][i][j].v - G[k][j][l].v * G[l][i][k].v
            R[i, j] = r
    giv = np.array([[gi[i][j].v for j in range(3)] for i in range(3)])
    Kv = np.array([[K[i][j].v for j in range(3)] for i in range(3)])
    Gv = np.array([[[G[k][i][j].v for j in range(3)] for i in range(3)] for k in range(3)])
    dK = np.array([[K[i][j].g for j in range(3)] for i in range(3)])  # dK[i, j, k] = d_k K_ij
    Rs = np.sum(giv * R)
    Ktr = np.sum(giv * Kv)
    Ku = giv @ Kv @ giv
    H = Rs + Ktr ** 2 - np.sum(Kv * Ku)
    # D_k K_ij = d_k K_ij - Gamma^l_ki K_lj - Gamma^l_kj K_il
    DK = dK.copy()
    for i in range(3):
        for j in range(3):
            for k in range(3):
                DK[i, j, k] -= np.sum(Gv[:, k, i] * Kv[:, j]) + np.sum(Gv[:, k, j] * Kv[i, :])
    dKtr = np.array([np.sum(giv * DK[:, :, k]) for k in range(3)])  # D_k K (gamma is covariantly constant)
    M = np.array([np.sum(giv * DK[:, i, :].T) - dKtr[i] for i in range(3)])
    return H, M


def generic_adm(point, seed=0):
    """A non-symmetric smooth spatial metric and extrinsic curvature near
    flat space: gamma_ij = delta_ij + quadratic polynomials, K_ij = linear
    polynomials, with seeded coefficients.  Returns (gamma, K) jets."""
    rng = np.random.default_rng(seed)
    x = [Jet.coord(i, point[i]) for i in range(3)]
    mono = [Jet.const(1.0)] + x + [x[a] * x[b] for a in range(3) for b in range(a, 3)]
    gamma = [[None] * 3 for _ in range(3)]
    K = [[None] * 3 for _ in range(3)]
    for a, b in SYM:
        cg = rng.uniform(-0.08, 0.08, len(mono))
        ck = rng.uniform(-0.3, 0.3, 4)
        gval = sum((c * m for c, m in zip(cg, mono)), Jet.const(1.0 if a == b else 0.0))
        kval = sum((c * m for c, m in zip(ck, mono[:4])), Jet.const(0.0))
        gamma[a][b] = gamma[b][a] = gval
        K[a][b] = K[b][a] = kval
    return gamma, K


def conformally_flat(point, a=0.05):
    """gamma_ij = psi^4 delta_ij, psi = 1 + a r^2, K_ij = 0: the Hamiltonian
    constraint is the Ricci scalar R = -8 psi^-5 lap(psi) = -48 a psi^-5."""
    x = [Jet.coord(i, point[i]) for i in range(3)]
    psi = 1 + a * (x[0] * x[0] + x[1] * x[1] + x[2] * x[2])
    p4 = psi ** 4
    zero = Jet.const(0.0)
    gamma = [[p4 if i == j else zero for j in range(3)] for i in range(3)]
    K = [[Jet(0.0) for _ in range(3)] for _ in range(3)]
    return gamma, K, -48.0 * a * psi.v ** -5


def bssn_of(gamma, K, alpha=None, beta_up=None):
    """BSSN point data (u, d1, dd, G) of ADM data (unit lapse, zero shift by
    default)."""
    zero = Jet.const(0.0)
    alpha = Jet.const(1.0) if alpha is None else alpha
    beta_up = [zero, zero, zero] if beta_up is None else beta_up
    return _bssn_from_adm(gamma, K, alpha, beta_up)


def gauge_wave_Bx(w, t, A, wavelength, eta, nodes=400):
    """The Gamma-driver's B^x on the harmonic gauge wave with the shift held
    at zero (no advection): dB/dt = d_t Gt^x - eta B, B(0) = 0, so
      B(t) = int_0^t e^{-eta (t - s)} d_s Gt^x(w, s) ds
           = g(w - t) - e^{-eta t} g(w) - eta int_0^t e^{-eta (t - s)} g(w - s) ds
    with Gt^x = g(w - t) = (2/3) H^(-5/3) d_w H, H = 1 - A sin(2 pi (w - t) /
    wavelength).  The last integral by Gauss-Legendre quadrature (the
    integrand is smooth; 400 nodes resolve it to round-off for t <= 64).
    w: array of positions along the wave axis."""
    k = 2 * np.pi / wavelength

    def g(u):
        H = 1 - A * np.sin(k * u)
        return (2 / 3) * H ** (-5 / 3) * (-A * k * np.cos(k * u))

    w = np.asarray(w, np.float64)
    if t == 0:
        return np.zeros_like(w)
    x, wq = np.polynomial.legendre.leggauss(nodes)
    s = (x + 1) / 2 * t
    integral = (np.exp(-eta * (t - s))[None, :] * g(w[:, None] - s[None, :])) @ (wq * t / 2)
    return g(w - t) - np.exp(-eta * t) * g(w) - eta * integral
