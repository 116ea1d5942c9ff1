"""Pins the oracle's stencils, interpolants and pointwise kernels by
exactness on polynomials, closed forms and library routines.

FD (AMR.h:1144-1219): the order-2q centred first derivative is exact on
polynomials of degree <= 2q, the second derivative on degree <= 2q+1; the
mixed-partial formulas, interior and edge (AMR.h:1221-1266), are exact on
quadratics.  Exact values that lie on the compressed_float3 grid decode
exactly (the dither amplitude is below half a step); others decode within one
grid step of their lane's shared exponent.
"""
import numpy as np
import pytest

from tests import fields, meshes

S, P = 16, 3
OUTER = 16.0  # outer_dx = 1, so inverse spacings are powers of two


def build(orc, recipe="graded", periodic=False):
    ops, n = {"graded": meshes.graded(), "uniform1": meshes.uniform(1), "uniform2": meshes.uniform(2),
              "nested3": meshes.nested(3)}[recipe]
    t = orc.OracleTree(n, periodic)
    meshes.apply(t, ops)
    return t, t.export(), n


def grid_tol(dec):
    """One grid step of each word's shared exponent."""
    m = np.abs(dec).max(axis=-1, keepdims=True)
    e = np.clip(np.floor(np.log2(np.maximum(m, 1e-30))) + 127 - 112, 0, 31)
    return np.exp2(e - 22) * 1.0001


def dec_words(orc, w, n, S_):
    return orc.cf3_unpack(w).reshape(n, S_, S_, S_, 3).astype(np.float64)


@pytest.mark.parametrize("q", [1, 2, 3])
def test_fd_polynomial_exactness(orc, q):
    t, e, n = build(orc)
    tiles = fields.used_slots(e)
    rng = np.random.default_rng(q)
    c = rng.integers(-3, 4, size=(3, 2 * q + 2)).astype(np.float64) / 8

    def f(x, y, z):  # degree 2q in each axis, plus an (ignored) degree-2q+1 part in d2
        return sum(c[0, k] * x ** k + c[1, k] * y ** k + c[2, k] * z ** k for k in range(2 * q + 1))

    def fp(v, a):
        return sum(k * c[a, k] * v ** (k - 1) for k in range(1, 2 * q + 1))

    def fpp(v, a):
        return sum(k * (k - 1) * c[a, k] * v ** (k - 2) for k in range(2, 2 * q + 1))

    field = fields.sample(e["positions"], n, S, P, OUTER, f)
    d1, d2 = orc.differentiate(field, e["positions"], tiles, 0, e["max_depth"], 1234, S, P, q, OUTER / S, n)
    g, h = dec_words(orc, d1, n, S), dec_words(orc, d2, n, S)
    for s in tiles:
        x, y, z = (fields.interior(a[None], 1, S, P)[0] for a in fields.cell_coords(e["positions"], s, S, P, OUTER))
        want1 = np.stack([fp(x, 0), fp(y, 1), fp(z, 2)], -1)
        want2 = np.stack([fpp(x, 0), fpp(y, 1), fpp(z, 2)], -1)
        assert np.all(np.abs(g[s] - want1) <= grid_tol(want1)), s
        assert np.all(np.abs(h[s] - want2) <= grid_tol(want2)), s


def test_fd_linear_bit_exact(orc):
    """A linear field's gradient lies on the grid: the words decode exactly."""
    t, e, n = build(orc)
    tiles = fields.used_slots(e)
    field = fields.sample(e["positions"], n, S, P, OUTER, lambda x, y, z: 3 * x + 5 * y - 7 * z + 1)
    for q in (1, 2, 3):
        d1, d2 = orc.differentiate(field, e["positions"], tiles, 0, e["max_depth"], 77, S, P, q, OUTER / S, n)
        g = dec_words(orc, d1, n, S)[tiles]
        assert np.array_equal(g, np.broadcast_to([3.0, 5.0, -7.0], g.shape))
        # second derivatives of a linear field are float round-off only
        assert np.max(np.abs(dec_words(orc, d2, n, S)[tiles])) < 1e-3


def test_fd_gradient_only_kernel_matches_d1_of_fused(orc):
    """__differentiate(field, dif) packs d1 with the first three draws; the
    d1+d2 kernel packs d2 first, so d1 words differ but decode alike."""
    t, e, n = build(orc)
    tiles = fields.used_slots(e)
    rng = np.random.default_rng(5)
    field = rng.standard_normal(n * (S + 2 * P) ** 3).astype(np.float32)
    a, _ = orc.differentiate(field, e["positions"], tiles, 0, e["max_depth"], 9, S, P, 3, OUTER / S, n, False)
    b, _ = orc.differentiate(field, e["positions"], tiles, 0, e["max_depth"], 9, S, P, 3, OUTER / S, n, True)
    da, db = dec_words(orc, a, n, S), dec_words(orc, b, n, S)
    assert np.all(np.abs(da - db) <= 2 * grid_tol(da))
    assert not np.array_equal(a, b)


def test_mixed_partials_exact_on_quadratics(orc):
    t, e, n = build(orc)
    tiles = fields.used_slots(e)
    field = fields.sample(e["positions"], n, S, P, OUTER,
                          lambda x, y, z: x * y + 2 * x * z - 3 * y * z + x * x - 2 * z * z + 4 * y + 0.5)
    off = orc.mixed_partials(field, e["positions"], tiles, 0, e["max_depth"], 31, S, P, OUTER / S, n)
    o = dec_words(orc, off, n, S)[tiles]
    assert np.array_equal(o, np.broadcast_to([1.0, 2.0, -3.0], o.shape))


def test_mixed_partials_avoid_edge_and_corner_ghosts(orc):
    """NaN in every edge and corner ghost cell must not reach any word
    (AMR.h:1248-1261: the 7-point formula at tile edges)."""
    t, e, n = build(orc, "uniform1")
    tiles = fields.used_slots(e)
    T = S + 2 * P
    f = fields.sample(e["positions"], n, S, P, OUTER, lambda x, y, z: np.sin(x) * np.cos(y) + z * x).reshape(n, T, T, T)
    idx = np.arange(T)
    ghost = (idx < P) | (idx >= P + S)
    gz, gy, gx = np.meshgrid(ghost, ghost, ghost, indexing="ij")
    f[:, (gz.astype(int) + gy + gx) >= 2] = np.nan
    f = f.reshape(-1)
    off = orc.mixed_partials(f, e["positions"], tiles, 0, e["max_depth"], 31, S, P, OUTER / S, n)
    d1, d2 = orc.differentiate(f, e["positions"], tiles, 0, e["max_depth"], 3, S, P, 3, OUTER / S, n)
    for w in (off, d1, d2):
        assert np.all(np.isfinite(orc.cf3_unpack(w.reshape(n, -1)[tiles].reshape(-1))))


def test_inactive_depths_untouched(orc):
    """With the counter at 1 only the deepest level is active."""
    t, e, n = build(orc)
    tiles = fields.used_slots(e)
    rng = np.random.default_rng(0)
    field = rng.standard_normal(n * (S + 2 * P) ** 3).astype(np.float32)
    d1, _ = orc.differentiate(field, e["positions"], tiles, 1, e["max_depth"], 5, S, P, 3, OUTER / S, n)
    w = d1.reshape(n, -1)
    for s in tiles:
        active = fields.depth_of(e, s) == e["max_depth"]
        assert bool(np.any(w[s])) == active


# ---------------------------------------------------------------- ghost fill
def fill_all(orc, old, new, older, e, tiles_by_depth, substep, mode, copy_only_new=True):
    for d in sorted(tiles_by_depth):
        if d == 0:
            continue
        orc.copy_boundaries(old, new, older, e["boundaries"], tiles_by_depth[d], substep, e["max_depth"],
                            copy_only_new, mode, S, P)


def by_depth(e):
    out = {}
    for s in fields.used_slots(e):
        out.setdefault(fields.depth_of(e, s), []).append(s)
    return out


def face_ghost_mask():
    T = S + 2 * P
    idx = np.arange(T)
    inner = (idx >= P) & (idx < P + S)
    gz, gy, gx = np.meshgrid(inner, inner, inner, indexing="ij")
    return (gz.astype(int) + gy + gx) == 2  # exactly one axis in the ghost layers


def test_ghost_equal_depth_is_exact_copy(orc):
    """At equal depth the fractional offset is 0 and a ghost cell is a copy
    of the neighbour's interior cell -- except on the last interior row of each
    axis, where the reference clamps the read index to T-P-2 (AMR.h:747-750)
    and evaluates the 10-point quadratic at offset 1.  That is exact only for
    fields the quadratic reproduces (see the polynomial test), not a copy."""
    t, e, n = build(orc, "uniform2", periodic=True)
    T = S + 2 * P
    rng = np.random.default_rng(1)
    new = (rng.random(n * T ** 3) + 0.5).astype(np.float32)
    old = (rng.random(n * T ** 3) + 0.5).astype(np.float32)
    want = new.copy()
    fill_all(orc, old, new, None, e, by_depth(e), 0, 0)
    nb = new.reshape(n, T, T, T)
    bnd = e["boundaries"].reshape(-1, 6)
    src = want.reshape(n, T, T, T)
    for s in range(9, n):
        # +x face ghost layer l copies interior layer l of the +x neighbour
        tgt = int(bnd[s, 0]) & 0xFFFFFF
        got, ref = nb[s, P:P + S, P:P + S, P + S:], src[tgt, P:P + S, P:P + S, P:2 * P]
        assert np.array_equal(got[:S - 1, :S - 1], ref[:S - 1, :S - 1])
        tgt = int(bnd[s, 5]) & 0xFFFFFF
        got, ref = nb[s, :P, P:P + S, P:P + S], src[tgt, S:S + P, P:P + S, P:P + S]
        # the -z ghost layer next to the interior reads source z = S+P-1, also clamped
        assert np.array_equal(got[:P - 1, :S - 1, :S - 1], ref[:P - 1, :S - 1, :S - 1])


@pytest.mark.parametrize("S_,P_", [(16, 3), (16, 2), (28, 2)])
def test_ghost_equal_copy_is_the_neighbour_cell(orc, S_, P_):
    """The RK4 driver's equal-depth rule (equal_copy): every cell of every
    P-layer face slab across an equal-depth face is the neighbour's interior
    cell S away, bit for bit -- also at S = 28, where 1/S is inexact in fp32
    and the mapped index is a whole number only within rounding."""
    ops, n = meshes.uniform(2)
    t = orc.OracleTree(n, True)
    meshes.apply(t, ops)
    e = t.export()
    T = S_ + 2 * P_
    new = np.random.default_rng(S_ + P_).random(n * T ** 3).astype(np.float32)
    src = new.copy().reshape(n, T, T, T)
    tiles = [s for s in fields.used_slots(e) if fields.depth_of(e, s) > 0]
    orc.copy_boundaries(new.copy(), new, None, e["boundaries"], tiles, 0, e["max_depth"], True, 0, S_, P_,
                        equal_copy=True)
    got_all = new.reshape(n, T, T, T)
    inner = slice(P_, P_ + S_)
    checked = 0
    for s in tiles:
        for f in range(6):
            nd = int(e["boundaries"][s * 6 + f]) & 0xFFFFFFFF
            if (nd >> 24) & 15:
                continue  # a coarse face
            tgt = nd & 0xFFFFFF
            gi, si = [inner] * 3, [inner] * 3  # x, y, z
            gi[f % 3] = slice(P_ + S_, T) if f < 3 else slice(0, P_)
            si[f % 3] = slice(P_, 2 * P_) if f < 3 else slice(S_, S_ + P_)
            got = got_all[s, gi[2], gi[1], gi[0]]
            ref = src[tgt, si[2], si[1], si[0]]
            assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), (S_, s, f)
            checked += 1
    assert checked == 6 * len(tiles)


@pytest.mark.parametrize("mode,fn", [
    (0, lambda x, y, z: 1 + x - 2 * y + 0.5 * z + 0.1 * x * x - 0.2 * y * z + 0.05 * z * z + 0.3 * x * y),
    (1, lambda x, y, z: 1 + x - 2 * y + 0.5 * z + 0.1 * x * y * z),
    (2, lambda x, y, z: 1 + x - 2 * y + 0.5 * z + 0.01 * x ** 3 - 0.02 * y * y * z + 0.03 * x * y * z),
])
def test_ghost_coarse_fine_polynomial_exactness(orc, mode, fn):
    """Face ghosts of every tile, from equal or coarser targets, reproduce a
    polynomial the interpolant is exact on."""
    t, e, n = build(orc, "graded")
    new = fields.sample(e["positions"], n, S, P, OUTER, fn)
    old = new.copy()
    want = fields.sample(e["positions"], n, S, P, OUTER, fn, dtype=np.float64)
    mask = np.zeros((n, S + 2 * P, S + 2 * P, S + 2 * P), bool)
    mask[1:] = face_ghost_mask()
    # everything the fill must write gets a sentinel (finite: equal-depth
    # reads touch a neighbour's ghost layer with weight exactly 0)
    new[mask.reshape(-1)] = 1e30
    fill_all(orc, old, new, None, e, by_depth(e), 0, mode)
    got = new.reshape(mask.shape)[mask]
    ref = want.reshape(mask.shape)[mask]
    assert np.max(np.abs(got - ref)) < 2e-5 * max(1.0, np.max(np.abs(ref)))


def test_ghost_time_interpolation(orc):
    """Linear in time between old and new at (k, k+1)/2^dd; quadratic through
    older/old/new in the tricubic mode (exact on t^2)."""
    t, e, n = build(orc, "graded")
    T = S + 2 * P
    fine = [s for s in fields.used_slots(e) if fields.depth_of(e, s) == 3]
    bnd = e["boundaries"].reshape(-1, 6)
    for mode in (0, 2):
        for sub in range(8):
            a = lambda tt: 2.0 + tt + (0.25 * tt * tt if mode == 2 else 0.0)
            older = np.full(n * T ** 3, a(-1.0), np.float32)
            old = np.full(n * T ** 3, a(0.0), np.float32)
            new = np.full(n * T ** 3, a(1.0), np.float32)
            orc.copy_boundaries(old, new, older if mode == 2 else None, e["boundaries"], fine, sub, e["max_depth"],
                                True, mode, S, P)
            nb = new.reshape(n, T, T, T)
            for s in fine:
                for face in range(6):
                    dd = (int(bnd[s, face]) >> 24) & 15
                    k = (sub >> (e["max_depth"] - 3)) & ((1 << dd) - 1)
                    ty = (k + 1) / 2 ** dd
                    want = a(ty) if dd else a(1.0)
                    v = nb[s, P:P + S, P:P + S, P + S:] if face == 0 else nb[s, :P, P:P + S, P:P + S]
                    if face in (0, 5):
                        assert np.allclose(v, want, rtol=2e-6, atol=0), (mode, sub, s, face)


def test_copy_new_to_old_restriction(orc):
    t, e, n = build(orc, "graded")
    T = S + 2 * P
    rng = np.random.default_rng(3)
    new = rng.standard_normal(n * T ** 3).astype(np.float32)
    old = np.zeros_like(new)
    tiles = fields.used_slots(e)
    orc.copy_new_to_old(old, new, e["positions"], e["children"], tiles, 0, e["max_depth"], S, P)
    nb, ob = new.reshape(n, T, T, T), old.reshape(n, T, T, T)
    ch = e["children"].reshape(-1, 8)
    for s in tiles:
        want = nb[s].copy()
        for o in range(8):
            c = ch[s, o]
            if c == -1:
                continue
            ox, oy, oz = o & 1, (o >> 1) & 1, (o >> 2) & 1
            blk = nb[c, P:P + S, P:P + S, P:P + S].reshape(S // 2, 2, S // 2, 2, S // 2, 2)
            # sum in the kernel's order: x, then y, then z pairs
            acc = np.zeros((S // 2,) * 3, np.float32)
            for dz in (0, 1):
                for dy in (0, 1):
                    for dx in (0, 1):
                        acc = acc + blk[:, dz, :, dy, :, dx]
            sl = (slice(P + oz * S // 2, P + (oz + 1) * S // 2), slice(P + oy * S // 2, P + (oy + 1) * S // 2),
                  slice(P + ox * S // 2, P + (ox + 1) * S // 2))
            want[sl] = acc * np.float32(0.125)
        assert np.allclose(ob[s], want, rtol=1e-6, atol=1e-6)
        assert np.array_equal(ob[s][want == nb[s]], nb[s][want == nb[s]])


@pytest.mark.parametrize("mode", [0, 1])
def test_copy_to_child_exact(orc, mode):
    t, e, n = build(orc, "uniform1")
    fn = (lambda x, y, z: 1 + x - y + 0.3 * z + 0.1 * x * y - 0.05 * z * z + 0.02 * x * x) if mode == 0 else \
        (lambda x, y, z: 1 + x - y + 0.3 * z + 0.1 * x * y * z)
    buf = fields.sample(e["positions"], n, S, P, OUTER, fn, slots=[0])
    want = fields.sample(e["positions"], n, S, P, OUTER, fn, dtype=np.float64)
    T = S + 2 * P
    for child in range(1, 9):
        orc.copy_to_child(buf, child - 1, child, 0, mode, S, P)
    got = buf.reshape(n, T, T, T)[1:]
    assert np.max(np.abs(got - want.reshape(n, T, T, T)[1:])) < 2e-5 * np.max(np.abs(want))


# ----------------------------------------------------------- pointwise
def test_wave_stages_one_step(orc):
    """u = a + b x: every FD is exact and b = 3/8 lands mid-binade on the
    grid (a value just under a power of two would saturate its mantissa);
    lap is round-off, so predictor u* = u - dt_d u b and corrector
    0.5 (u* + u - dt_d u b)."""
    t, e, n = build(orc, "uniform1")
    tiles = fields.used_slots(e)
    T = S + 2 * P
    dt = np.float32(0.4)
    u = fields.sample(e["positions"], n, S, P, OUTER, lambda x, y, z: 2.0 + 0.375 * x)
    d1, d2 = orc.differentiate(u, e["positions"], tiles, 0, 1, 1, S, P, 3, OUTER / S, n)
    new = u.copy()
    orc.wave_stage(u, d1, d2, new, e["positions"], tiles, 0, 1, False, S, P, dt)
    ui = fields.interior(u, n, S, P).astype(np.float64)
    for s in tiles:
        dtd = 0.4 / 2 ** fields.depth_of(e, s)
        want = ui[s] - dtd * ui[s] * 0.375
        assert np.allclose(fields.interior(new, n, S, P)[s], want, rtol=2e-7, atol=1e-6)
    new2 = new.copy()
    orc.wave_stage(u, d1, d2, new2, e["positions"], tiles, 0, 1, True, S, P, dt)
    for s in tiles:
        dtd = 0.4 / 2 ** fields.depth_of(e, s)
        want = 0.5 * (fields.interior(new, n, S, P)[s] + ui[s] - dtd * ui[s] * 0.375)
        assert np.allclose(fields.interior(new2, n, S, P)[s], want, rtol=2e-7, atol=1e-6)


def random_spd(rng, shape):
    a = rng.standard_normal(shape + (3, 3)) * 0.2
    return np.eye(3) + a @ np.swapaxes(a, -1, -2)


SYM = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def test_deriv_trace_matches_inverse(orc):
    t, e, n = build(orc, "uniform1")
    tiles = fields.used_slots(e)
    T = S + 2 * P
    rng = np.random.default_rng(7)
    g = random_spd(rng, (n, T, T, T))
    metric = [np.ascontiguousarray(g[..., a, b].astype(np.float32).reshape(-1)) for a, b in SYM]
    hd = orc.cf3_pack_array(rng.standard_normal((n * S ** 3, 3)))
    ho = orc.cf3_pack_array(rng.standard_normal((n * S ** 3, 3)))
    out = orc.deriv_trace(hd, ho, metric, e["positions"], tiles, 0, 1, S, P, n)
    H = np.zeros((n * S ** 3, 3, 3))
    d, o = orc.cf3_unpack(hd), orc.cf3_unpack(ho)
    H[:, 0, 0], H[:, 1, 1], H[:, 2, 2] = d[:, 0], d[:, 1], d[:, 2]
    H[:, 0, 1] = H[:, 1, 0] = o[:, 0]
    H[:, 0, 2] = H[:, 2, 0] = o[:, 1]
    H[:, 1, 2] = H[:, 2, 1] = o[:, 2]
    gi = np.linalg.inv(np.stack([np.stack([metric[SYM.index(tuple(sorted((a, b))))] for b in range(3)], -1)
                                 for a in range(3)], -2).astype(np.float64))
    gi = gi.reshape(n, T, T, T, 3, 3)[:, P:P + S, P:P + S, P:P + S].reshape(-1, 3, 3)
    want = np.einsum("nij,nij->n", gi, H)
    assert np.allclose(out, want, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("compat", [0, 1])
def test_christoffel_matches_definition(orc, compat):
    """Gamma^l_jk = g^lm (d_j g_mk + d_k g_mj - d_m g_jk)/2 from the decoded
    derivative words; compat contracts with g_lm instead (BSSN.h:82-83)."""
    t, e, n = build(orc, "uniform1")
    tiles = fields.used_slots(e)
    T = S + 2 * P
    rng = np.random.default_rng(11)
    g = random_spd(rng, (n, T, T, T))
    metric = [np.ascontiguousarray(g[..., a, b].astype(np.float32).reshape(-1)) for a, b in SYM]
    dg = [orc.cf3_pack_array(rng.standard_normal((n * S ** 3, 3)) * 0.3) for _ in range(6)]
    out = orc.christoffel(metric, dg, compat, e["positions"], tiles, 0, 1, 4242, S, P, n)
    D = np.zeros((n * S ** 3, 3, 3, 3))  # D[c, a, b] = d_c g_ab
    for k, (a, b) in enumerate(SYM):
        v = orc.cf3_unpack(dg[k]).astype(np.float64)
        D[:, :, a, b] = v
        D[:, :, b, a] = v
    low = 0.5 * (np.einsum("njmk->nmjk", D) + np.einsum("nkmj->nmjk", D) - D)  # Gamma_{m jk}
    gm = np.stack([np.stack([metric[SYM.index(tuple(sorted((a, b))))] for b in range(3)], -1)
                   for a in range(3)], -2).astype(np.float64)
    gm = gm.reshape(n, T, T, T, 3, 3)[:, P:P + S, P:P + S, P:P + S].reshape(-1, 3, 3)
    C = gm if compat else np.linalg.inv(gm)
    want = np.einsum("nlm,nmjk->nljk", C, low)
    for k, (a, b) in enumerate(SYM):
        got = orc.cf3_unpack(out[k]).astype(np.float64)
        w = want[:, :, a, b]
        tol = grid_tol(w) + 1e-5 * np.abs(w).max(axis=-1, keepdims=True)
        assert np.all(np.abs(got - w) <= tol), (compat, a, b)
