"""Kernels on interior-first tiles (am_tile_layout 1, the RK4 drivers'
storage) against the oracle on the same data in the reference's padded
layout, bit for bit: the layout conversion itself, the FD stencils (every
geometry and z-block, own ghost layers or direct y / z faces), the stage ghost
fill (three interpolations, copy and formula rules, S = 16 and 28), the
driver's copy-back with restriction, and prolongation into a child.  Arrays
move between the layouts on the host with adaptivemesh_amd's
interior_first_index, which tests/test_tile_layout.py pins to the
library's own map."""
import ctypes as C

import numpy as np
import pytest

from tests import fields, meshes
from tests.test_gpu_parity import GEOMS, dev, host, mesh_for, setup_mesh, sync

pytestmark = pytest.mark.gpu


def to_if(am, a, S, P, n):
    """Padded host array of n slots -> interior-first host array."""
    T = S + 2 * P
    g = am.geometry(S, P, 2, 64.0, 0.25)
    TS = int(am.lib().am_tile_stride(C.byref(g), 1))
    out = np.full((n, TS), np.nan, a.dtype)
    out[:, am.interior_first_index(S, P)] = a.reshape(n, T ** 3)
    return out.reshape(-1)


def from_if(am, b, S, P, n):
    T = S + 2 * P
    return b.reshape(n, -1)[:, am.interior_first_index(S, P)].reshape(n * T ** 3)


@pytest.mark.parametrize("S,P", [(16, 3), (16, 2), (28, 2)])
def test_convert_round_trip(am, torch_cuda, S, P):
    torch = torch_cuda
    print(f"layout convert S={S} P={P}", flush=True)
    T = S + 2 * P
    n = 11
    g = am.geometry(S, P, 2, 64.0, 0.25)
    a = np.random.default_rng(S + P).standard_normal(n * T ** 3).astype(np.float32)
    TS = int(am.lib().am_tile_stride(C.byref(g), 1))
    pa = dev(torch, a)
    ia = torch.zeros(n * TS, dtype=torch.float32, device="cuda")
    back = torch.zeros_like(pa)
    L, st = am.lib(), am.stream_handle()
    am.check(L.am_tile_layout_convert(pa.data_ptr(), ia.data_ptr(), n, C.byref(g), 0, 1, st), "to interior-first")
    am.check(L.am_tile_layout_convert(ia.data_ptr(), back.data_ptr(), n, C.byref(g), 1, 0, st), "to padded")
    sync(torch)
    want = to_if(am, a, S, P, n).reshape(n, TS)[:, :T ** 3]
    assert np.array_equal(host(ia, np.float32).reshape(n, TS)[:, :T ** 3].view(np.uint32), want.view(np.uint32))
    assert np.array_equal(host(back, np.float32).view(np.uint32), a.view(np.uint32))


@pytest.mark.parametrize("S,P,order", GEOMS)
@pytest.mark.parametrize("recipe,periodic", [("graded", False), ("uniform2", True)])
def test_fd_interior_first(am, orc, torch_cuda, S, P, order, recipe, periodic):
    """FD words of an interior-first field equal the oracle's on the padded
    field, for every z-block and two counter values; with direct faces
    (masks 2 y, 4 z, 6 both) the equal-depth ghost layers on those axes are
    poisoned (NaN) and the words equal the oracle's on the field whose ghost
    layers hold the neighbours' copies."""
    torch = torch_cuda
    print(f"fd interior-first S={S} P={P} order={order} {recipe}", flush=True)
    g = am.geometry(S, P, order, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, recipe, periodic)
    T = S + 2 * P
    rng = np.random.default_rng(S * 7 + P + order)
    field = rng.standard_normal(n * T ** 3).astype(np.float32)
    tiles = [s for s in fields.used_slots(e) if fields.depth_of(e, s) > 0]
    filled = field.copy()
    orc.copy_boundaries(field.copy(), filled, None, e["boundaries"], tiles, 0, e["max_depth"], True, 0, S, P,
                        equal_copy=True)
    eq_yz = fields.outer_push_layers(e, n, tiles, S, P, 0, 6)
    good = field.copy()
    good[eq_yz] = filled[eq_yz]
    L, st = am.lib(), am.stream_handle()
    W = n * S ** 3
    for direct in (0, 2, 4, 6):
        src = good.copy()
        if direct:
            poison = fields.outer_push_layers(e, n, tiles, S, P, 0, direct)
            assert poison.any()
            src[poison] = np.nan
        f = dev(torch, to_if(am, src, S, P, n))
        for substep in (0, 1):
            seed_d, seed_m = int(rng.integers(-2 ** 31, 2 ** 31)), int(rng.integers(-2 ** 31, 2 ** 31))
            r1, r2 = orc.differentiate(good, e["positions"], fields.used_slots(e), substep, e["max_depth"], seed_d,
                                       S, P, order // 2, g.outer_dx, n)
            rm = orc.mixed_partials(good, e["positions"], fields.used_slots(e), substep, e["max_depth"], seed_m, S,
                                    P, g.outer_dx, n)
            m = mesh_for(tree, -1, substep)
            for zb in (0, S // 2, S // 4):
                d1, d2, off = (torch.zeros(W, dtype=torch.int32, device="cuda") for _ in range(3))
                b = am.FdBatch()
                b.n_vars = 1
                b.z_block = zb
                b.tile_layout = 1
                b.direct_faces = direct
                b.field[0], b.d1[0], b.d2diag[0], b.d2off[0] = (f.data_ptr(), d1.data_ptr(), d2.data_ptr(),
                                                                off.data_ptr())
                b.seed_d[0], b.seed_m[0] = seed_d, seed_m
                am.check(L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st), "am_fd_batched")
                sync(torch)
                assert np.array_equal(host(d1, np.uint32), r1), (direct, substep, zb)
                assert np.array_equal(host(d2, np.uint32), r2), (direct, substep, zb)
                assert np.array_equal(host(off, np.uint32), rm), (direct, substep, zb)
    # the interior-first FD takes direct masks 0, 2, 4 and 6 and the one-plane prefetch only
    for bad_direct, bad_pf in ((1, 1), (7, 1), (3, 1), (5, 1), (0, 2), (6, 2)):
        b.direct_faces, b.prefetch_planes = bad_direct, bad_pf
        assert L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st) == -1


@pytest.mark.parametrize("S,P", [(28, 2), (16, 3)])
@pytest.mark.parametrize("mode,equal", [(0, 1), (1, 1), (2, 1), (1, 0), (2, 0), (2, 2)])
@pytest.mark.parametrize("recipe,periodic", [("graded", False), ("uniform2", True)])
def test_stage_fill_interior_first(am, orc, torch_cuda, S, P, mode, equal, recipe, periodic):
    """am_ghost_fill_stage_batched on interior-first fields (6 per launch)
    equals the oracle's copy_boundaries on the padded fields bit for bit,
    at stage fractions 1/2 and 1 (equal=2 skips equal-depth faces: the
    oracle's result there is replaced by the input).  Cells the fill does
    not write (edges, corners, interiors, unfilled faces) keep their
    values."""
    torch = torch_cuda
    print(f"stage fill interior-first S={S} P={P} mode={mode} equal={equal} {recipe}", flush=True)
    g = am.geometry(S, P, 2 * P, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, recipe, periodic)
    T = S + 2 * P
    nv = 6
    rng = np.random.default_rng(S * 31 + mode * 7 + equal)
    depths = {}
    for s in fields.used_slots(e):
        depths.setdefault(fields.depth_of(e, s), []).append(s)
    for substep in range(min(1 << e["max_depth"], 4)):
        for frac in (0.5, 1.0):
            old = rng.standard_normal((nv, n * T ** 3)).astype(np.float32)
            new = rng.standard_normal((nv, n * T ** 3)).astype(np.float32)
            older = rng.standard_normal((nv, n * T ** 3)).astype(np.float32)
            conv = lambda a: np.stack([to_if(am, a[v], S, P, n) for v in range(nv)])
            do, dn, dl = dev(torch, conv(old)), dev(torch, conv(new)), dev(torch, conv(older))
            fb = am.FieldBatch()
            fb.n_vars = nv
            fb.tile_layout = 1
            for v in range(nv):
                fb.old_[v], fb.new_[v] = do[v].data_ptr(), dn[v].data_ptr()
                fb.older[v] = dl[v].data_ptr() if mode == 2 else None
            for d in sorted(depths):
                if d == 0:
                    continue
                for v in range(nv):
                    before = new[v].copy()
                    orc.copy_boundaries(old[v], new[v], older[v] if mode == 2 else None, e["boundaries"], depths[d],
                                        substep, e["max_depth"], True, mode, S, P, frac, equal_copy=equal == 1)
                    if equal == 2:  # equal-depth faces left alone
                        keep = fields.outer_push_layers(e, n, depths[d], S, P, 0)
                        new[v][keep] = before[keep]
                m = mesh_for(tree, d, substep)
                am.check(am.lib().am_ghost_fill_stage_batched(C.byref(fb), C.byref(m), C.byref(g), C.c_float(frac),
                                                              mode, equal, am.stream_handle()), "stage fill")
            sync(torch)
            got = host(dn, np.float32)
            for v in range(nv):
                assert np.array_equal(from_if(am, got[v], S, P, n).view(np.uint32), new[v].view(np.uint32)), \
                    (substep, frac, v)


@pytest.mark.parametrize("S,P", [(16, 3), (28, 2)])
def test_copy_back_interior_first(am, orc, torch_cuda, S, P):
    """The driver's copy-back on interior-first tiles: on a tile flagged in
    keep_ghosts every cell is the reference copy (older <- old first, the
    interior restricted from the children); on the others the interior
    cells are, and the ghost cells of old and older keep their values; new
    then equals old on the interior of every active tile."""
    torch = torch_cuda
    g = am.geometry(S, P, 2 * P, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, "graded")
    T = S + 2 * P
    rng = np.random.default_rng(S * P + 1)
    keep = (rng.random(n) < 0.4).astype(np.uint8)
    ghost = np.ones((n, T, T, T), bool)
    ghost[:, P:P + S, P:P + S, P:P + S] = False
    lean = np.repeat(keep == 0, T ** 3) & ghost.reshape(-1)
    for substep in range(4):
        new = rng.standard_normal(n * T ** 3).astype(np.float32)
        old = rng.standard_normal(n * T ** 3).astype(np.float32)
        older = rng.standard_normal(n * T ** 3).astype(np.float32)
        do, dn, dd = (dev(torch, to_if(am, a, S, P, n)) for a in (old, new, older))
        want_old, want_older = old.copy(), older.copy()
        active = [s for s in fields.used_slots(e)
                  if (substep & ((1 << (e["max_depth"] - fields.depth_of(e, s))) - 1)) == 0]
        idx = np.concatenate([np.arange(s * T ** 3, (s + 1) * T ** 3) for s in active])
        want_older[idx] = old[idx]
        orc.copy_new_to_old(want_old, new, e["positions"], e["children"], fields.used_slots(e), substep,
                            e["max_depth"], S, P)
        want_old[lean] = old[lean]
        want_older[lean] = older[lean]
        want_new = new.copy()
        inner = np.zeros((n, T, T, T), bool)
        inner[active, P:P + S, P:P + S, P:P + S] = True
        inner = inner.reshape(-1)
        want_new[inner] = want_old[inner]
        fb = am.FieldBatch()
        fb.n_vars = 1
        fb.tile_layout = 1
        fb.old_[0], fb.new_[0], fb.older[0] = do.data_ptr(), dn.data_ptr(), dd.data_ptr()
        kg = torch.tensor(keep, device="cuda")
        m = mesh_for(tree, -1, substep)
        am.check(am.lib().am_copy_new_to_old_interior(C.byref(fb), C.byref(m), kg.data_ptr(), C.byref(g),
                                                      am.stream_handle()), "copy_new_to_old_interior")
        sync(torch)
        for d_, want in ((dn, want_new), (do, want_old), (dd, want_older)):
            got = from_if(am, host(d_, np.float32), S, P, n)
            assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), substep


@pytest.mark.parametrize("mode", [0, 1])
def test_copy_to_child_interior_first(am, orc, torch_cuda, mode):
    """Prolongation of a 3-field batch into a new child on interior-first
    tiles equals the oracle on padded tiles, every cell of the child."""
    torch = torch_cuda
    S, P = 16, 3
    g = am.geometry(S, P, 4, 64.0, 0.25)
    T = S + 2 * P
    n, nv = 9, 3
    rng = np.random.default_rng(20 + mode)
    bufs = [rng.standard_normal(n * T ** 3).astype(np.float32) for _ in range(nv)]
    devs = [dev(torch, to_if(am, b, S, P, n)) for b in bufs]
    fb = am.FieldBatch()
    fb.n_vars = nv
    fb.tile_layout = 1
    for v in range(nv):
        fb.new_[v] = devs[v].data_ptr()
    child, parent, off = 3, 0, 6
    am.check(am.lib().am_copy_to_child_batched(C.byref(fb), off, child, parent, C.byref(g), mode,
                                               am.stream_handle()), "ctc batched")
    sync(torch)
    for v in range(nv):
        ref = bufs[v].copy()
        orc.copy_to_child(ref, off, child, parent, mode, S, P)
        got = from_if(am, host(devs[v], np.float32), S, P, n)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), v
