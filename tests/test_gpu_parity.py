"""HIP kernels (through the C ABI) against the oracle, bit for bit.

Every kernel that restates a reference kernel must produce the oracle's
exact bits on the same seeded inputs: FD words (all orders, every supported
tile geometry, every launch shape), ghost fills (three interpolation modes,
every counter value), restriction, prolongation, SimpleWave stages,
deriv_trace and Christoffel words, and a whole SimpleWave run through the
native driver.
"""
import ctypes as C

import numpy as np
import pytest

from tests import fields, meshes

pytestmark = pytest.mark.gpu

GEOMS = [(16, 3, 6), (16, 3, 4), (16, 3, 2), (16, 2, 4), (16, 2, 2), (28, 2, 4)]


def dev(torch, a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    if a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a.copy()).cuda()


def host(t, dtype):
    a = t.cpu().numpy()
    return a.view(dtype) if a.dtype != dtype else a


def setup_mesh(am, orc, recipe, periodic=False):
    ops, n = {"graded": meshes.graded(), "uniform1": meshes.uniform(1), "uniform2": meshes.uniform(2),
              "nested3": meshes.nested(3)}[recipe]
    tree = am.Tree(n, periodic)
    otree = orc.OracleTree(n, periodic)
    meshes.apply(tree, ops)
    meshes.apply(otree, ops)
    return tree, otree.export(), n


def mesh_for(tree, depth, substep):
    m = tree.mesh(depth)
    m.substep = substep
    return m


def sync(torch):
    torch.cuda.synchronize()


@pytest.mark.parametrize("S,P,order", GEOMS)
@pytest.mark.parametrize("recipe", ["graded", "uniform2"])
@pytest.mark.parametrize("substep", [0, 1, 2])
def test_fd_bit_exact(am, orc, torch_cuda, S, P, order, recipe, substep):
    torch = torch_cuda
    print(f"fd S={S} P={P} order={order} {recipe} substep={substep}", flush=True)
    g = am.geometry(S, P, order, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, recipe)
    T = S + 2 * P
    rng = np.random.default_rng(S * 100 + order + substep)
    field = rng.standard_normal(n * T ** 3).astype(np.float32)
    seed_d, seed_m = int(rng.integers(-2 ** 31, 2 ** 31)), int(rng.integers(-2 ** 31, 2 ** 31))
    tiles = fields.used_slots(e)
    r1, r2 = orc.differentiate(field, e["positions"], tiles, substep, e["max_depth"], seed_d, S, P, order // 2,
                               g.outer_dx, n)
    r1only, _ = orc.differentiate(field, e["positions"], tiles, substep, e["max_depth"], seed_d, S, P, order // 2,
                                  g.outer_dx, n, with_d2=False)
    rm = orc.mixed_partials(field, e["positions"], tiles, substep, e["max_depth"], seed_m, S, P, g.outer_dx, n)

    f = dev(torch, field)
    m = mesh_for(tree, -1, substep)
    L = am.lib()
    st = am.stream_handle()
    W = n * S ** 3
    for zb, pf in ((0, 1), (S // 2, 1), (S // 4, 1), (0, 2), (0, 3), (S // 2, 3), (S // 4, 2)):
        d1, d2, off = (torch.zeros(W, dtype=torch.int32, device="cuda") for _ in range(3))
        b = am.FdBatch()
        b.n_vars = 1
        b.z_block = zb
        b.prefetch_planes = pf
        b.field[0], b.d1[0], b.d2diag[0], b.d2off[0] = f.data_ptr(), d1.data_ptr(), d2.data_ptr(), off.data_ptr()
        b.seed_d[0], b.seed_m[0] = seed_d, seed_m
        am.check(L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st), "am_fd_batched")
        sync(torch)
        assert np.array_equal(host(d1, np.uint32), r1), zb
        assert np.array_equal(host(d2, np.uint32), r2), zb
        assert np.array_equal(host(off, np.uint32), rm), zb
    # single-purpose entry points
    d1 = torch.zeros(W, dtype=torch.int32, device="cuda")
    am.check(L.am_fd_d1(f.data_ptr(), d1.data_ptr(), C.byref(m), C.byref(g), seed_d, st), "am_fd_d1")
    off = torch.zeros(W, dtype=torch.int32, device="cuda")
    am.check(L.am_fd_mixed(f.data_ptr(), off.data_ptr(), C.byref(m), C.byref(g), seed_m, st), "am_fd_mixed")
    sync(torch)
    assert np.array_equal(host(d1, np.uint32), r1only)
    assert np.array_equal(host(off, np.uint32), rm)


def extreme_field(rng, n_slots, T):
    """Per tile, normal values scaled by 1, 1e-6 (words whose exponent clamps
    at 0), 1e-39 (subnormal values and products) or 0, with NaN, +-inf,
    values near the float maximum and signed zeros sprinkled in: stencil
    sums then cover NaN (from inf - inf or a NaN tap), inf, words whose
    exponent clamps at 31 next to tiny components, exponent 0 and exact
    zeros of either sign."""
    f = rng.standard_normal((n_slots, T ** 3)).astype(np.float32)
    f *= np.array([1.0, 1e-6, 1e-39, 0.0], np.float32)[rng.integers(0, 4, n_slots)][:, None]
    f = f.reshape(-1)
    pick = rng.random(f.size)
    specials = np.array([np.nan, np.inf, -np.inf, 3e38, -3e38, 1e30, -2e-40, 3e-41, 0.0, -0.0, 1e-30, 7e-39],
                        np.float32)
    sel = pick < 0.02
    f[sel] = specials[rng.integers(0, len(specials), sel.sum())]
    return f


@pytest.mark.parametrize("S,P,order", [(16, 3, 4), (28, 2, 4), (16, 3, 6)])
def test_fd_words_on_extreme_values(am, orc, torch_cuda, S, P, order):
    """The dithered pack over the whole float range: FD words of a field
    holding NaN, inf, huge, subnormal and signed-zero values equal the
    oracle's bit for bit (the saturating clamp, the exponent rule at both
    ends, the sign of every component)."""
    torch = torch_cuda
    print(f"fd extreme values S={S} P={P} order={order}", flush=True)
    g = am.geometry(S, P, order, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, "graded")
    T = S + 2 * P
    rng = np.random.default_rng(S * 31 + order)
    field = extreme_field(rng, n, T)
    seed_d, seed_m = int(rng.integers(-2 ** 31, 2 ** 31)), int(rng.integers(-2 ** 31, 2 ** 31))
    tiles = fields.used_slots(e)
    with np.errstate(all="ignore"):
        r1, r2 = orc.differentiate(field, e["positions"], tiles, 0, e["max_depth"], seed_d, S, P, order // 2,
                                   g.outer_dx, n)
        rm = orc.mixed_partials(field, e["positions"], tiles, 0, e["max_depth"], seed_m, S, P, g.outer_dx, n)
    # the sample does reach both exponent ends and the NaN / inf lanes
    ex = r1 >> 27
    assert (ex == 31).any() and (ex == 0).any() and ((r1 & 255) == 255).any()
    f = dev(torch, field)
    m = mesh_for(tree, -1, 0)
    W = n * S ** 3
    d1, d2, off = (torch.zeros(W, dtype=torch.int32, device="cuda") for _ in range(3))
    b = am.FdBatch()
    b.n_vars = 1
    b.field[0], b.d1[0], b.d2diag[0], b.d2off[0] = f.data_ptr(), d1.data_ptr(), d2.data_ptr(), off.data_ptr()
    b.seed_d[0], b.seed_m[0] = seed_d, seed_m
    am.check(am.lib().am_fd_batched(C.byref(b), C.byref(m), C.byref(g), am.stream_handle()), "am_fd_batched")
    sync(torch)
    for got, want, what in ((d1, r1, "d1"), (d2, r2, "d2diag"), (off, rm, "d2off")):
        g_ = host(got, np.uint32)
        bad = np.flatnonzero(g_ != want)
        assert bad.size == 0, (what, bad[:5], g_[bad[:5]], want[bad[:5]])


@pytest.mark.parametrize("S,P,order", GEOMS)
@pytest.mark.parametrize("recipe,periodic", [("graded", False), ("uniform2", True), ("nested3", False)])
def test_fd_direct_faces(am, orc, torch_cuda, S, P, order, recipe, periodic):
    """direct_faces: ghost cells across equal-depth faces of the chosen axes
    come from the neighbours' interiors.  With every equal-depth ghost layer
    of those axes poisoned (NaN), the words equal the oracle's on the field
    whose ghost layers hold the neighbours' copies, for every axis mask
    tried, launch shape and two counter values."""
    torch = torch_cuda
    print(f"fd direct S={S} P={P} order={order} {recipe} periodic={periodic}", flush=True)
    g = am.geometry(S, P, order, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, recipe, periodic)
    T = S + 2 * P
    rng = np.random.default_rng(S * 10 + P + order)
    field = rng.standard_normal(n * T ** 3).astype(np.float32)
    tiles = [s for s in fields.used_slots(e) if fields.depth_of(e, s) > 0]
    filled = field.copy()
    orc.copy_boundaries(field.copy(), filled, None, e["boundaries"], tiles, 0, e["max_depth"], True, 0, S, P,
                        equal_copy=True)
    eq_all = fields.outer_push_layers(e, n, tiles, S, P, 0)
    assert eq_all.any()
    good = field.copy()
    good[eq_all] = filled[eq_all]
    L = am.lib()
    st = am.stream_handle()
    W = n * S ** 3
    for axes in (7, 4, 3, 1):
        bad = good.copy()
        bad[fields.outer_push_layers(e, n, tiles, S, P, 0, axes)] = np.nan
        fb = dev(torch, bad)
        for substep in (0, 1):
            seed_d, seed_m = int(rng.integers(-2 ** 31, 2 ** 31)), int(rng.integers(-2 ** 31, 2 ** 31))
            r1, r2 = orc.differentiate(good, e["positions"], fields.used_slots(e), substep, e["max_depth"], seed_d,
                                       S, P, order // 2, g.outer_dx, n)
            rm = orc.mixed_partials(good, e["positions"], fields.used_slots(e), substep, e["max_depth"], seed_m, S,
                                    P, g.outer_dx, n)
            m = mesh_for(tree, -1, substep)
            for zb, pf in ((0, 1), (S // 2, 2), (S // 4, 3)):
                d1, d2, off = (torch.zeros(W, dtype=torch.int32, device="cuda") for _ in range(3))
                b = am.FdBatch()
                b.n_vars = 1
                b.z_block = zb
                b.prefetch_planes = pf
                b.direct_faces = axes
                b.field[0], b.d1[0], b.d2diag[0], b.d2off[0] = (fb.data_ptr(), d1.data_ptr(), d2.data_ptr(),
                                                                off.data_ptr())
                b.seed_d[0], b.seed_m[0] = seed_d, seed_m
                am.check(L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st), "am_fd_batched")
                sync(torch)
                assert np.array_equal(host(d1, np.uint32), r1), (axes, substep, zb)
                assert np.array_equal(host(d2, np.uint32), r2), (axes, substep, zb)
                assert np.array_equal(host(off, np.uint32), rm), (axes, substep, zb)
        # without direct reads the poisoned layers reach the words
        b.direct_faces = 0
        am.check(L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st), "am_fd_batched")
        sync(torch)
        assert not np.array_equal(host(d1, np.uint32), r1), axes
    # an axis mask outside 0..7, or direct reads without the boundary table
    b.direct_faces = 8
    assert L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st) == -1
    m.boundaries = None
    b.direct_faces = 1
    assert L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st) == -1


def test_fd_batched_many_variables(am, orc, torch_cuda):
    """24 fields in one launch, mixed outputs per variable, equal the
    one-at-a-time oracle results."""
    torch = torch_cuda
    S, P = 16, 3
    g = am.geometry(S, P, 4, 64.0, 0.25)
    tree, e, n = setup_mesh(am, orc, "uniform2")
    T = S + 2 * P
    rng = np.random.default_rng(7)
    tiles = fields.used_slots(e)
    b = am.FdBatch()
    b.n_vars = 24
    keep, want = [], []
    for v in range(24):
        fld = rng.standard_normal(n * T ** 3).astype(np.float32)
        sd, sm = int(rng.integers(-2 ** 31, 2 ** 31)), int(rng.integers(-2 ** 31, 2 ** 31))
        hess = v % 2 == 0
        r1, r2 = orc.differentiate(fld, e["positions"], tiles, 0, e["max_depth"], sd, S, P, 2, g.outer_dx, n, hess)
        rm = orc.mixed_partials(fld, e["positions"], tiles, 0, e["max_depth"], sm, S, P, g.outer_dx, n) if hess else None
        f = dev(torch, fld)
        outs = [torch.zeros(n * S ** 3, dtype=torch.int32, device="cuda") for _ in range(3 if hess else 1)]
        b.field[v] = f.data_ptr()
        b.d1[v] = outs[0].data_ptr()
        b.d2diag[v] = outs[1].data_ptr() if hess else None
        b.d2off[v] = outs[2].data_ptr() if hess else None
        b.seed_d[v], b.seed_m[v] = sd, sm
        keep += [f] + outs
        want.append((outs, [r1, r2, rm] if hess else [r1]))
    m = mesh_for(tree, -1, 0)
    am.check(am.lib().am_fd_batched(C.byref(b), C.byref(m), C.byref(g), am.stream_handle()), "batched")
    sync(torch)
    for outs, refs in want:
        for o, r in zip(outs, refs):
            assert np.array_equal(host(o, np.uint32), r)


@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("recipe,periodic", [("graded", False), ("nested3", False), ("uniform2", True)])
@pytest.mark.parametrize("copy_only_new", [True, False])
def test_ghost_fill_bit_exact(am, orc, torch_cuda, mode, recipe, periodic, copy_only_new):
    torch = torch_cuda
    S, P = 16, 3
    print(f"ghost mode={mode} {recipe} periodic={periodic} con={copy_only_new}", flush=True)
    g = am.geometry(S, P, 6, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, recipe, periodic)
    T = S + 2 * P
    rng = np.random.default_rng(mode * 10 + len(recipe))
    depths = {}
    for s in fields.used_slots(e):
        depths.setdefault(fields.depth_of(e, s), []).append(s)
    for substep in range(1 << e["max_depth"]):
        old = rng.standard_normal(n * T ** 3).astype(np.float32)
        new = rng.standard_normal(n * T ** 3).astype(np.float32)
        older = rng.standard_normal(n * T ** 3).astype(np.float32)
        do, dn, dol = dev(torch, old), dev(torch, new), dev(torch, older)
        for d in sorted(depths):
            if d == 0:
                continue
            orc.copy_boundaries(old, new, older if mode == 2 else None, e["boundaries"], depths[d], substep,
                                e["max_depth"], copy_only_new, mode, S, P)
            m = mesh_for(tree, d, substep)
            am.check(am.lib().am_ghost_fill(do.data_ptr(), dn.data_ptr(), dol.data_ptr() if mode == 2 else None,
                                            C.byref(m), C.byref(g), 1 if copy_only_new else 0, mode,
                                            am.stream_handle()), "am_ghost_fill")
        sync(torch)
        assert np.array_equal(host(dn, np.float32).view(np.uint32), new.view(np.uint32)), substep
        assert np.array_equal(host(do, np.float32).view(np.uint32), old.view(np.uint32)), substep


@pytest.mark.parametrize("S,P", [(28, 2), (16, 3)])
@pytest.mark.parametrize("mode,equal", [(0, 1), (1, 1), (2, 1), (1, 0), (2, 0)])
@pytest.mark.parametrize("recipe,periodic", [("graded", False), ("uniform2", True)])
def test_ghost_fill_stage_equal_modes(am, orc, torch_cuda, S, P, mode, equal, recipe, periodic):
    """The RK4 driver's stage fill (am_ghost_fill_stage_batched, 8 fields per
    launch) with equal-depth faces copied (equal=1, AM_EQUAL_COPY) or
    interpolated by the formula (equal=0, AM_EQUAL_FORMULA) equals the
    oracle's copy_boundaries bit for bit at stage fractions 1/2 and 1, on
    the reference's default tile (S=28, P=2, where 1/S is inexact and the
    mapped read positions are a rounding away from whole cells) and on
    S=16.  The quadratic formula (mode 0) is tested with the copy rule only:
    at S=28 its equal-depth reads reach the target's ghost cells, which the
    same pass writes."""
    torch = torch_cuda
    print(f"stage fill S={S} P={P} mode={mode} equal={equal} {recipe}", flush=True)
    g = am.geometry(S, P, 2 * P, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, recipe, periodic)
    T = S + 2 * P
    nv = 8
    rng = np.random.default_rng(S * 100 + mode * 10 + equal)
    depths = {}
    for s in fields.used_slots(e):
        depths.setdefault(fields.depth_of(e, s), []).append(s)
    for substep in range(min(1 << e["max_depth"], 4)):
        for frac in (0.5, 1.0):
            old = rng.standard_normal((nv, n * T ** 3)).astype(np.float32)
            new = rng.standard_normal((nv, n * T ** 3)).astype(np.float32)
            older = rng.standard_normal((nv, n * T ** 3)).astype(np.float32)
            do, dn, dl = dev(torch, old), dev(torch, new), dev(torch, older)
            fb = am.FieldBatch()
            fb.n_vars = nv
            for v in range(nv):
                fb.old_[v] = do[v].data_ptr()
                fb.new_[v] = dn[v].data_ptr()
                fb.older[v] = dl[v].data_ptr() if mode == 2 else None
            for d in sorted(depths):
                if d == 0:
                    continue
                for v in range(nv):
                    orc.copy_boundaries(old[v], new[v], older[v] if mode == 2 else None, e["boundaries"], depths[d],
                                        substep, e["max_depth"], True, mode, S, P, frac, equal_copy=equal == 1)
                m = mesh_for(tree, d, substep)
                am.check(am.lib().am_ghost_fill_stage_batched(C.byref(fb), C.byref(m), C.byref(g), C.c_float(frac),
                                                              mode, equal, am.stream_handle()), "stage fill")
            sync(torch)
            got = host(dn, np.float32).view(np.uint32)
            assert np.array_equal(got, new.view(np.uint32)), (substep, frac)


@pytest.mark.parametrize("S,P", [(28, 2), (16, 3)])
@pytest.mark.parametrize("mode", [0, 1, 2])
@pytest.mark.parametrize("layout", [0, 1])
def test_ghost_fill_stage_faces_equals_full(am, orc, torch_cuda, S, P, mode, layout):
    """am_ghost_fill_stage_faces over the list of a depth's faces whose
    target is coarser (the RK4 driver's call), with equal-depth and outer
    faces mixed in, writes what am_ghost_fill_stage_batched writes with
    AM_EQUAL_SKIP over all the depth's tiles, bit for bit, on the graded
    mesh in both tile layouts; the list argument is checked."""
    torch = torch_cuda
    print(f"face-list fill S={S} P={P} mode={mode} layout={layout}", flush=True)
    g = am.geometry(S, P, 2 * P, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, "graded", False)
    T = S + 2 * P
    W = n * int(am.lib().am_tile_stride(C.byref(g), layout))
    nv = 5
    rng = np.random.default_rng(S + mode + 10 * layout)
    depths = {}
    for s in fields.used_slots(e):
        depths.setdefault(fields.depth_of(e, s), []).append(s)
    lists = {}
    for d, ts in depths.items():
        coarse, other = [], []
        for t in ts:
            for f in range(6):
                nd = int(e["boundaries"][t * 6 + f]) & 0xFFFFFFFF
                is_coarse = not (nd & 0x800000) and nd >> 28 not in (0, 15) and (nd >> 24) & 15
                (coarse if is_coarse else other).append(t * 6 + f)
        lists[d] = (coarse, other)
    assert sum(len(c) for c, _ in lists.values()) > 0
    for substep in range(4):
        src = [rng.standard_normal((nv, W)).astype(np.float32) for _ in range(3)]
        dev_a = [dev(torch, x) for x in src]
        dev_b = [dev(torch, x) for x in src]
        fbs = []
        for dv in (dev_a, dev_b):
            fb = am.FieldBatch()
            fb.n_vars = nv
            fb.tile_layout = layout
            for v in range(nv):
                fb.old_[v], fb.new_[v] = dv[0][v].data_ptr(), dv[1][v].data_ptr()
                fb.older[v] = dv[2][v].data_ptr() if mode == 2 else None
            fbs.append(fb)
        for d in sorted(depths):
            if d == 0:
                continue
            m = mesh_for(tree, d, substep)
            am.check(am.lib().am_ghost_fill_stage_batched(C.byref(fbs[0]), C.byref(m), C.byref(g), C.c_float(0.5),
                                                          mode, am.AM_EQUAL_SKIP, am.stream_handle()), "full")
            coarse, other = lists[d]
            fl = np.array(coarse + other[::3], np.int32)
            rng.shuffle(fl)
            dfl = torch.from_numpy(fl).cuda()
            am.check(am.lib().am_ghost_fill_stage_faces(C.byref(fbs[1]), C.byref(m), dfl.data_ptr(), len(fl),
                                                        C.byref(g), C.c_float(0.5), mode, am.AM_EQUAL_SKIP,
                                                        am.stream_handle()), "faces")
            sync(torch)
        a, b = host(dev_a[1], np.float32), host(dev_b[1], np.float32)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), substep
        assert not np.array_equal(a.view(np.uint32), src[1].view(np.uint32))
    m = mesh_for(tree, 1, 0)
    assert am.lib().am_ghost_fill_stage_faces(C.byref(fbs[1]), C.byref(m), None, 3, C.byref(g), C.c_float(0.5),
                                              mode, am.AM_EQUAL_SKIP, am.stream_handle()) == -1
    assert am.lib().am_ghost_fill_stage_faces(C.byref(fbs[1]), C.byref(m), None, 0, C.byref(g), C.c_float(0.5),
                                              mode, am.AM_EQUAL_SKIP, am.stream_handle()) == 0


@pytest.mark.parametrize("S,P", [(16, 3), (16, 2), (28, 2)])
def test_copy_new_to_old_bit_exact(am, orc, torch_cuda, S, P):
    torch = torch_cuda
    g = am.geometry(S, P, 2 * P, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, "graded")
    T = S + 2 * P
    rng = np.random.default_rng(S + P)
    for substep in range(4):
        new = rng.standard_normal(n * T ** 3).astype(np.float32)
        old = rng.standard_normal(n * T ** 3).astype(np.float32)
        do, dn = dev(torch, old), dev(torch, new)
        orc.copy_new_to_old(old, new, e["positions"], e["children"], fields.used_slots(e), substep, e["max_depth"],
                            S, P)
        m = mesh_for(tree, -1, substep)
        am.check(am.lib().am_copy_new_to_old(do.data_ptr(), dn.data_ptr(), C.byref(m), C.byref(g),
                                             am.stream_handle()), "copy_new_to_old")
        sync(torch)
        assert np.array_equal(host(do, np.float32).view(np.uint32), old.view(np.uint32))


@pytest.mark.parametrize("S,P", [(16, 3), (28, 2)])
def test_copy_new_to_old_interior(am, orc, torch_cuda, S, P):
    """The driver's copy-back: tiles flagged in keep_ghosts copy like the
    reference (every cell, restriction, older <- old first); on the others
    every interior cell is copied the same way, and each ghost cell either
    keeps its value or receives the reference copy's (the interior planes
    are copied as contiguous vector runs, which take the x-ghost cells
    along).  new then equals old on the interior of every active tile
    (restricted cells copied back into new) and is unchanged elsewhere."""
    torch = torch_cuda
    g = am.geometry(S, P, 2 * P, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, "graded")
    T = S + 2 * P
    rng = np.random.default_rng(S * P)
    keep = (rng.random(n) < 0.4).astype(np.uint8)
    ghost = np.ones((n, T, T, T), bool)
    ghost[:, P:P + S, P:P + S, P:P + S] = False
    lean = np.repeat(keep == 0, T ** 3) & ghost.reshape(-1)
    for substep in range(4):
        new = rng.standard_normal(n * T ** 3).astype(np.float32)
        old = rng.standard_normal(n * T ** 3).astype(np.float32)
        older = rng.standard_normal(n * T ** 3).astype(np.float32)
        do, dn, dd = dev(torch, old), dev(torch, new), dev(torch, older)
        want_old, want_older = old.copy(), older.copy()
        active = [s for s in fields.used_slots(e)
                  if (substep & ((1 << (e["max_depth"] - fields.depth_of(e, s))) - 1)) == 0]
        idx = np.concatenate([np.arange(s * T ** 3, (s + 1) * T ** 3) for s in active])
        want_older[idx] = old[idx]
        orc.copy_new_to_old(want_old, new, e["positions"], e["children"], fields.used_slots(e), substep,
                            e["max_depth"], S, P)
        want_new = new.copy()
        inner = np.zeros((n, T, T, T), bool)
        inner[active, P:P + S, P:P + S, P:P + S] = True
        inner = inner.reshape(-1)
        want_new[inner] = want_old[inner]
        fb = am.FieldBatch()
        fb.n_vars = 1
        fb.old_[0], fb.new_[0], fb.older[0] = do.data_ptr(), dn.data_ptr(), dd.data_ptr()
        kg = torch.tensor(keep, device="cuda")
        m = mesh_for(tree, -1, substep)
        am.check(am.lib().am_copy_new_to_old_interior(C.byref(fb), C.byref(m), kg.data_ptr(), C.byref(g),
                                                      am.stream_handle()), "copy_new_to_old_interior")
        sync(torch)
        got_old, got_older = host(do, np.float32).view(np.uint32), host(dd, np.float32).view(np.uint32)
        assert np.array_equal(host(dn, np.float32).view(np.uint32), want_new.view(np.uint32)), substep
        full = ~lean
        assert np.array_equal(got_old[full], want_old.view(np.uint32)[full]), substep
        assert np.array_equal(got_older[full], want_older.view(np.uint32)[full]), substep
        assert np.all((got_old[lean] == old.view(np.uint32)[lean]) | (got_old[lean] == want_old.view(np.uint32)[lean]))
        assert np.all((got_older[lean] == older.view(np.uint32)[lean]) |
                      (got_older[lean] == want_older.view(np.uint32)[lean]))


@pytest.mark.parametrize("mode", [0, 1])
def test_copy_to_child_bit_exact(am, orc, torch_cuda, mode):
    torch = torch_cuda
    S, P = 16, 3
    g = am.geometry(S, P, 6, 100.0, 0.4)
    T = S + 2 * P
    n = 9
    rng = np.random.default_rng(mode)
    buf = rng.standard_normal(n * T ** 3).astype(np.float32)
    d = dev(torch, buf)
    for c in range(8):
        orc.copy_to_child(buf, c, c + 1, 0, mode, S, P)
        am.check(am.lib().am_copy_to_child(d.data_ptr(), c, c + 1, 0, C.byref(g), mode, am.stream_handle()), "ctc")
    sync(torch)
    assert np.array_equal(host(d, np.float32).view(np.uint32), buf.view(np.uint32))


@pytest.mark.parametrize("mode", [0, 1])
def test_copy_to_child_batched_bit_exact(am, orc, torch_cuda, mode):
    """One launch prolonging every old/new/older buffer of a 5-field batch
    equals the oracle applied buffer by buffer."""
    torch = torch_cuda
    S, P = 16, 3
    g = am.geometry(S, P, 4, 64.0, 0.25)
    T = S + 2 * P
    n, nv = 9, 5
    rng = np.random.default_rng(10 + mode)
    bufs = {(k, v): rng.standard_normal(n * T ** 3).astype(np.float32) for k in range(3) for v in range(nv)}
    devs = {key: dev(torch, b) for key, b in bufs.items()}
    fb = am.FieldBatch()
    fb.n_vars = nv
    for v in range(nv):
        fb.old_[v] = devs[(0, v)].data_ptr()
        fb.new_[v] = devs[(1, v)].data_ptr()
        fb.older[v] = devs[(2, v)].data_ptr() if v % 2 == 0 else None
    child, parent, off = 5, 0, 4
    am.check(am.lib().am_copy_to_child_batched(C.byref(fb), off, child, parent, C.byref(g), mode,
                                               am.stream_handle()), "ctc batched")
    sync(torch)
    for (k, v), b in bufs.items():
        ref = b.copy()
        if k < 2 or v % 2 == 0:
            orc.copy_to_child(ref, off, child, parent, mode, S, P)
        assert np.array_equal(host(devs[(k, v)], np.float32).view(np.uint32), ref.view(np.uint32)), (k, v)


def test_pointwise_bit_exact(am, orc, torch_cuda):
    """SimpleWave stages, deriv_trace and both Christoffel modes."""
    torch = torch_cuda
    S, P = 16, 3
    g = am.geometry(S, P, 6, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, "graded")
    T = S + 2 * P
    rng = np.random.default_rng(3)
    tiles = fields.used_slots(e)
    L, st = am.lib(), am.stream_handle()
    for substep in range(3):
        m = mesh_for(tree, -1, substep)
        u = rng.standard_normal(n * T ** 3).astype(np.float32)
        d1 = orc.cf3_pack_array(rng.standard_normal((n * S ** 3, 3)))
        d2 = orc.cf3_pack_array(rng.standard_normal((n * S ** 3, 3)))
        for corr in (False, True):
            new = rng.standard_normal(n * T ** 3).astype(np.float32)
            dnew = dev(torch, new)
            orc.wave_stage(u, d1, d2, new, e["positions"], tiles, substep, e["max_depth"], corr, S, P, g.outer_dt)
            fn = L.am_wave_corrector if corr else L.am_wave_predictor
            du, dd1, dd2 = dev(torch, u), dev(torch, d1), dev(torch, d2)
            am.check(fn(du.data_ptr(), dd1.data_ptr(), dd2.data_ptr(), dnew.data_ptr(), C.byref(m), C.byref(g), st),
                     "wave")
            sync(torch)
            assert np.array_equal(host(dnew, np.float32).view(np.uint32), new.view(np.uint32))
        a = rng.standard_normal((n * T ** 3, 3, 3)) * 0.2
        met = np.eye(3) + a @ np.swapaxes(a, -1, -2)
        metric = [np.ascontiguousarray(met[:, i, j].astype(np.float32)) for i, j in
                  [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]]
        dmet = [dev(torch, x) for x in metric]
        mt = (C.c_void_p * 6)(*[x.data_ptr() for x in dmet])
        hd, ho = d1, d2
        want = orc.deriv_trace(hd, ho, metric, e["positions"], tiles, substep, e["max_depth"], S, P, n)
        out = torch.zeros(n * S ** 3, dtype=torch.float32, device="cuda")
        dhd, dho = dev(torch, hd), dev(torch, ho)
        am.check(L.am_deriv_trace(dhd.data_ptr(), dho.data_ptr(), mt, out.data_ptr(), C.byref(m), C.byref(g), st),
                 "trace")
        sync(torch)
        assert np.array_equal(host(out, np.float32).view(np.uint32), want.view(np.uint32))
        dg = [orc.cf3_pack_array(rng.standard_normal((n * S ** 3, 3)) * 0.3) for _ in range(6)]
        ddg = [dev(torch, x) for x in dg]
        dgt = (C.c_void_p * 6)(*[x.data_ptr() for x in ddg])
        for compat in (0, 1):
            seed = int(rng.integers(-2 ** 31, 2 ** 31))
            want = orc.christoffel(metric, dg, compat, e["positions"], tiles, substep, e["max_depth"], seed, S, P, n)
            outs = [torch.zeros(n * S ** 3, dtype=torch.int32, device="cuda") for _ in range(6)]
            ot = (C.c_void_p * 6)(*[x.data_ptr() for x in outs])
            am.check(L.am_christoffel(mt, dgt, ot, compat, C.byref(m), C.byref(g), seed, st), "christoffel")
            sync(torch)
            for o, w in zip(outs, want):
                assert np.array_equal(host(o, np.uint32), w)


def oracle_wave_timestep(orc, e, n, old, new, rng, S, P, g, tiles_by_depth):
    """wave_AMR_data::timestep (SimpleWave.h:268-278) driven on the oracle,
    with the reference's seed draws: one per differentiate call."""
    tiles = sorted(s for v in tiles_by_depth.values() for s in v)
    maxd = e["max_depth"]
    timer = 0
    while True:
        orc.copy_new_to_old(old, new, e["positions"], e["children"], tiles, timer, maxd, S, P)
        d1, d2 = orc.differentiate(old, e["positions"], tiles, timer, maxd, orc.prng_ints(rng[0], 1)[0], S, P,
                                   g.fd_order // 2, g.outer_dx, n)
        rng[0] = orc.prng_ints(rng[0], 1)[0]
        orc.wave_stage(old, d1, d2, new, e["positions"], tiles, timer, maxd, False, S, P, g.outer_dt)
        for d in sorted(tiles_by_depth):
            if d > 0:
                orc.copy_boundaries(old, new, None, e["boundaries"], tiles_by_depth[d], timer, maxd, True, 0, S, P)
        d1, d2 = orc.differentiate(new, e["positions"], tiles, timer, maxd, orc.prng_ints(rng[0], 1)[0], S, P,
                                   g.fd_order // 2, g.outer_dx, n)
        rng[0] = orc.prng_ints(rng[0], 1)[0]
        orc.wave_stage(old, d1, d2, new, e["positions"], tiles, timer, maxd, True, S, P, g.outer_dt)
        timer = (timer + 1) & ((1 << maxd) - 1)
        if timer == 0:
            return


def test_simplewave_driver_bit_exact(am, orc, torch_cuda):
    """The native AMR<wave_AMR_data> (static graded mesh: modify_nodes off)
    over two full timesteps equals the oracle driven through the same phases
    and seeds."""
    S, P = 16, 3
    g = am.geometry(S, P, 6, 100.0, 0.4)
    ops, n = meshes.graded()
    sim = am.WaveSim(g, n)
    sim.set_regrid(False)
    otree = orc.OracleTree(n)
    T = S + 2 * P
    # initial data on the root, prolonged into children as add_node does
    root = orc.OracleTree(1).export()
    blob = fields.sample(root["positions"], 1, S, P, 100.0,
                         lambda x, y, z: 10 / ((x - 25) ** 2 + (y - .5) ** 2 + z * z + 10)
                         - 10 / ((x + 25) ** 2 + (y + .5) ** 2 + z * z + 10))
    init = np.zeros(n * T ** 3, np.float32)
    init[:T ** 3] = blob
    sim.set_field(0, init)
    sim.set_field(1, init)
    old, new = init.copy(), init.copy()
    for p, c in ops:
        s = sim.add_node(p, c)
        assert otree.add_node(p, c) == s
        orc.copy_to_child(old, c, s, p, 0, S, P)
        orc.copy_to_child(new, c, s, p, 0, S, P)
    e = otree.export()
    by_depth = {}
    for s in fields.used_slots(e):
        by_depth.setdefault(fields.depth_of(e, s), []).append(s)
    assert np.array_equal(sim.get_field(0).view(np.uint32), old.view(np.uint32))
    sim.copy_bounds()
    for d in sorted(by_depth):
        if d > 0:
            orc.copy_boundaries(old, new, None, e["boundaries"], by_depth[d], 0, e["max_depth"], True, 0, S, P)
    rng = [12894712]
    for step in range(2):
        sim.timestep()
        oracle_wave_timestep(orc, e, n, old, new, rng, S, P, g, by_depth)
        assert sim.rng_state == rng[0]
        assert np.array_equal(sim.get_field(1).view(np.uint32), new.view(np.uint32)), step
        assert np.array_equal(sim.get_field(0).view(np.uint32), old.view(np.uint32)), step
