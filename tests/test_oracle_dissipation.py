"""Pins the oracle's upwinded advection and Kreiss-Oliger dissipation (an
extension for the moving-puncture runs; the reference has neither):

* the float restatement of the FD's output (orc_fd_adv, which the GPU FD
  must equal bit for bit): the sixth difference vanishes on polynomials of
  degree <= 5 and takes the mode (-1)^(x+y+z) to -192 f, so the dissipation
  is -3 sigma/h f there; the lopsided first derivatives are exact on
  polynomials of degree <= 4; a unit impulse shows which five offsets the
  stencil reads for each sign of the shift and with which weights;
* the RHS with the advection given (orc_bssn_point_adv) equals the centred
  RHS when what it is given is the centred advection (the B equation's
  d_0 Gt form included), and the dissipation of a field enters only its own
  equation (and B's, through d_t Gt, without advect_shift).
"""
import numpy as np
import pytest

from tests import fields as F
from tests import meshes

S, P, OUTER = 16, 3, 4.0
T = S + 2 * P


def mesh(orc):
    ops, n = meshes.uniform(1)
    t = orc.OracleTree(n, True)
    for p_, c in ops:
        t.add_node(p_, c)
    e = t.export()
    return e, n, [s for s in F.used_slots(e) if F.depth_of(e, s) == 1]


def sample(e, n, fn):
    return F.sample(e["positions"], n, S, P, OUTER, fn).astype(np.float32)


def interior(n, a):
    return a.reshape(n, S, S, S)


def test_ko_vanishes_on_quintics(orc):
    e, n, tiles = mesh(orc)
    h = OUTER / S / 2  # depth-1 cell size
    # degree 5 in each variable, in cell units (values stay O(1) across a tile)
    f = sample(e, n, lambda x, y, z: ((x / h) ** 5 - 3 * (y / h) ** 4 * (z / h) + (x / h) ** 2 * (z / h) ** 3) * 1e-5)
    zero = np.zeros_like(f)
    out = orc.fd_adv(f, [zero, zero, zero], e["positions"], tiles, 0, 1, S, P, OUTER / S, 0.5, False, n)
    scale = np.max(np.abs(f)) * 64  # sum of |weights| times the largest value read
    out = interior(n, out)[tiles]
    assert np.max(np.abs(out)) < 2e-6 * scale * (0.5 / h / 64), np.max(np.abs(out))


def test_ko_damps_the_nyquist_mode(orc):
    """f = (-1)^(i+j+k) on the cell indices: each sixth difference is -64 f,
    so out = -192 f (sigma / 64) (1 / h) = -3 sigma f / h, exactly in float
    (integers, then one rounded product)."""
    e, n, tiles = mesh(orc)
    idx = np.indices((T, T, T)).sum(0)
    f = np.tile(np.where(idx % 2 == 0, 1.0, -1.0).reshape(-1), n).astype(np.float32)
    zero = np.zeros_like(f)
    sigma = 0.25
    out = orc.fd_adv(f, [zero, zero, zero], e["positions"], tiles, 0, 1, S, P, OUTER / S, sigma, False, n)
    dx1 = np.float32(2) / np.float32(OUTER / S)
    kos = np.float32(np.float32(sigma) * np.float32(1 / 64)) * dx1
    centre = f.reshape(n, T, T, T)[:, P:P + S, P:P + S, P:P + S]
    want = (np.float32(-192) * centre * kos).astype(np.float32)
    got = interior(n, out)
    assert np.array_equal(got[tiles], want[tiles])
    assert np.allclose(got[tiles], -3 * sigma * dx1 * centre[tiles], rtol=1e-6)


@pytest.mark.parametrize("seed", [0, 1])
def test_lopsided_exact_on_quartics(orc, seed):
    """beta . grad f with the shift's sign varying from cell to cell: exact
    for a quartic f up to float rounding (sigma = 0)."""
    e, n, tiles = mesh(orc)
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, 6)
    f64 = lambda x, y, z: c[0] * x ** 4 + c[1] * x * y ** 3 + c[2] * z ** 2 * y ** 2 + c[3] * y + c[4] * z ** 3 + c[5]
    grad = lambda x, y, z: (4 * c[0] * x ** 3 + c[1] * y ** 3,
                            3 * c[1] * x * y ** 2 + 2 * c[2] * z ** 2 * y + c[3],
                            2 * c[2] * z * y ** 2 + 3 * c[4] * z ** 2)
    f = sample(e, n, f64)
    shift = [sample(e, n, lambda x, y, z, k=k: np.sin(3 * x + 2 * k) * np.cos(2 * y - k) + 0.1 * z) for k in range(3)]
    out = interior(n, orc.fd_adv(f, shift, e["positions"], tiles, 0, 1, S, P, OUTER / S, 0.0, True, n))
    for slot in tiles:
        cx, cy, cz = F.cell_coords(e["positions"], slot, S, P, OUTER)
        sl = (slice(P, P + S),) * 3
        X, Y, Z = cx[sl], cy[sl], cz[sl]
        g = grad(X, Y, Z)
        b = [s.reshape(n, T, T, T)[slot][sl].astype(np.float64) for s in shift]
        want = b[0] * g[0] + b[1] * g[1] + b[2] * g[2]
        scale = np.max(np.abs(f)) * 3 * (OUTER / S / 2) ** -1 * 3
        assert np.max(np.abs(out[slot] - want)) < 1e-5 * scale, slot


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_lopsided_reads_the_upwind_side(orc, sign):
    """A unit impulse at one cell, shift (sign, 0, 0), sigma = 0: the output
    is non-zero exactly at the cells whose stencil reads the impulse, with
    the stencil's weight times 1/h times the shift -- offsets -1..3 for a
    positive shift, -3..1 for a negative one, along x only (the y and z
    shift components are 0, so their terms vanish)."""
    e, n, tiles = mesh(orc)
    f = np.zeros(n * T ** 3, np.float32)
    slot, (z0, y0, x0) = tiles[3], (7, 8, 8)
    f[slot * T ** 3 + ((z0 + P) * T + y0 + P) * T + x0 + P] = 1.0
    zero = np.zeros_like(f)
    bx = np.full_like(f, sign)
    out = interior(n, orc.fd_adv(f, [bx, zero, zero], e["positions"], tiles, 0, 1, S, P, OUTER / S, 0.0, True, n))
    nz = np.argwhere(out[slot] != 0)
    dx1 = np.float32(2) / np.float32(OUTER / S)
    weights = ([-3, -10, 18, -6, 1], range(-1, 4)) if sign > 0 else ([-1, 6, -18, 10, 3], range(-3, 2))
    want = {}
    for w, m in zip(*weights):
        want[(z0, y0, x0 - m)] = np.float32(sign) * (np.float32(w / 12) * dx1)  # cell x reads x + m
    assert {tuple(map(int, p)) for p in nz} == set(want), nz
    for k, v in want.items():
        assert out[slot][k] == v, (k, out[slot][k], v)


def _kerr_point():
    from tests import bssn_exact as be
    u, d1, dd, G = be.kerr_schild((1.1, 0.7, -0.5))
    u = u.copy()
    u[21:24] = [0.05, -0.02, 0.03]  # a non-zero B
    d1 = d1.copy()
    d1[21:24] = np.random.default_rng(3).uniform(-0.1, 0.1, (3, 3))
    return u, d1, dd


@pytest.mark.parametrize("advect", [0, 1])
def test_given_centred_advection_reproduces_the_centred_rhs(orc, advect):
    u, d1, dd = _kerr_point()
    params = (2.0, 0.75, 1e-6, advect, 0)
    want, _ = orc.bssn_point(u, d1, dd, None, params)
    adv = d1 @ u[18:21]  # beta . grad of every field
    if not advect:
        adv[18:24] = 0.0  # the shift is then not advected
    got = orc.bssn_point_adv(u, d1, dd, adv, params)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-14 * np.max(np.abs(want)))


@pytest.mark.parametrize("advect", [0, 1])
def test_dissipation_enters_its_own_equation(orc, advect):
    """Adding a dissipation term q to field v's advection input adds q to
    v's RHS only -- except Gt^i, whose d_t feeds B^i's equation without
    advect_shift (with it, d_0 Gt drops Gt's own advection and dissipation)."""
    u, d1, dd = _kerr_point()
    params = (2.0, 0.75, 1e-6, advect, 0)
    adv = d1 @ u[18:21]
    if not advect:
        adv[18:24] = 0.0
    base = orc.bssn_point_adv(u, d1, dd, adv, params)
    for v in range(24):
        a = adv.copy()
        a[v] += 0.125
        d = orc.bssn_point_adv(u, d1, dd, a, params) - base
        want = np.zeros(24)
        want[v] = 0.125
        if 14 <= v <= 16 and not advect:
            want[v + 7] = 0.125
        assert np.allclose(d, want, atol=1e-12), (v, d)
