"""Device-side compaction of the octree's launch lists (am_tile_lists).

Each list is checked against its definition, evaluated by brute force over
the exported tree: every slot up to the last used one; the slots of each
depth; the leaves (a used slot with no child); per depth >= 1 the faces
(slot * 6 + face) whose boundary target is a coarser tile; the slots those
faces read.  Entries must come in ascending slot order, as a sequential loop
over the slots writes them (the order the kernels' results never depend on,
but the recorded launches and the host's counts do).  Trees are random
add/remove sequences, sized so the scan spans one block and many (256 slots
per block), and the three standard meshes.
"""
import ctypes as C

import numpy as np
import pytest

from tests import meshes

pytestmark = pytest.mark.gpu

DEPTH_LIMIT = 8


def random_tree(am, n, seed, periodic, steps):
    rng = np.random.default_rng(seed)
    t = am.Tree(n, periodic=periodic)
    for _ in range(steps):
        e = t.export()
        used = [i for i in range(n) if e["positions"][i] != 0xFFFFFFFF]
        if rng.random() < 0.85 or len(used) == 1:
            p = int(rng.choice(used))
            if (int(e["positions"][p]) >> 27) >= 6:
                continue
            t.add_node(p, int(rng.integers(0, 8)))
        else:
            t.remove_node(int(rng.choice(used[1:])))
    return t


def expected(e, max_slots):
    pos, ch, bnd = e["positions"], e["children"].reshape(-1, 8), e["boundaries"]
    n = e["final_used"] + 1
    depth = np.where(pos == 0xFFFFFFFF, -1, (pos >> 27).astype(np.int64))
    tiles = [np.arange(n)]
    for d in range(DEPTH_LIMIT + 1):
        tiles.append(np.array([i for i in range(n) if depth[i] == d], np.int64))
    tiles.append(np.array([i for i in range(n) if depth[i] >= 0 and (ch[i] == -1).all()], np.int64))
    faces = [[] for _ in range(DEPTH_LIMIT + 1)]
    source = np.zeros(max_slots, np.uint8)
    for i in range(n):
        if depth[i] < 1:
            continue
        for f in range(6):
            nd = int(bnd[i * 6 + f]) & 0xFFFFFFFF
            tgt = nd & 0xFFFFFF
            if tgt & 0x800000 or ((nd >> 24) & 15) == 0:
                continue
            faces[depth[i]].append(i * 6 + f)
            source[tgt] = 1
    return n, tiles, [np.array(f, np.int64) for f in faces], source


def run_lists(am, torch, e, max_slots):
    L = am.lib()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    pos, ch = dev(e["positions"].view(np.int32)), dev(e["children"])
    bnd = dev(e["boundaries"].view(np.int64))
    n = e["final_used"] + 1
    tiles = torch.full((3 * max_slots,), -7, dtype=torch.int32, device="cuda")
    faces = torch.full((6 * max_slots,), -7, dtype=torch.int32, device="cuda")
    source = torch.full((max_slots,), 9, dtype=torch.uint8, device="cuda")
    ws_n = L.am_tile_lists_workspace(max_slots, DEPTH_LIMIT)
    assert ws_n == (max_slots + 255) // 256 * (2 * DEPTH_LIMIT + 4)
    ws = torch.zeros(max(ws_n, 1), dtype=torch.int32, device="cuda")
    counts = torch.zeros(4 * DEPTH_LIMIT + 8, dtype=torch.int32, device="cuda")
    am.check(L.am_tile_lists(pos.data_ptr(), ch.data_ptr(), bnd.data_ptr(), n, max_slots, DEPTH_LIMIT,
                             tiles.data_ptr(), faces.data_ptr(), source.data_ptr(), ws.data_ptr(), counts.data_ptr(),
                             am.stream_handle()), "am_tile_lists")
    torch.cuda.synchronize()
    return tiles.cpu().numpy(), faces.cpu().numpy(), source.cpu().numpy(), counts.cpu().numpy()


def check_lists(am, torch, e, max_slots, what):
    n, want_t, want_f, want_s = expected(e, max_slots)
    tiles, faces, source, counts = run_lists(am, torch, e, max_slots)
    NT, NF = DEPTH_LIMIT + 3, DEPTH_LIMIT + 1
    off = 0
    for k in range(NT):
        assert counts[k] == off and counts[NT + k] == len(want_t[k]), (what, "tile list", k)
        assert np.array_equal(tiles[off:off + len(want_t[k])], want_t[k]), (what, "tile list", k)
        off += len(want_t[k])
    assert (tiles[off:] == -7).all(), what  # nothing written past the lists
    off = 0
    for d in range(NF):
        assert counts[2 * NT + d] == off and counts[2 * NT + NF + d] == len(want_f[d]), (what, "face list", d)
        assert np.array_equal(faces[off:off + len(want_f[d])], want_f[d]), (what, "face list", d)
        off += len(want_f[d])
    assert (faces[off:] == -7).all(), what
    assert np.array_equal(source, want_s), what  # every slot < max_slots cleared or marked


@pytest.mark.parametrize("n,seed,periodic,steps", [(40, 0, False, 60), (300, 1, True, 400), (1200, 2, False, 1500),
                                                   (1200, 3, True, 900)])
def test_lists_random_trees(am, torch_cuda, n, seed, periodic, steps):
    print(f"tile lists random tree n={n} seed={seed} periodic={periodic}", flush=True)
    t = random_tree(am, n, seed, periodic, steps)
    check_lists(am, torch_cuda, t.export(), n, (n, seed))


@pytest.mark.parametrize("recipe", ["uniform3", "nested6", "graded"])
def test_lists_meshes_and_driver_views(am, torch_cuda, recipe):
    """The standard meshes; the tree's own device views (am_tree_mesh, which
    the drivers launch over) are the same lists."""
    torch = torch_cuda
    print(f"tile lists {recipe}", flush=True)
    ops, n = {"uniform3": meshes.uniform(3), "nested6": meshes.nested(6), "graded": meshes.graded()}[recipe]
    t = am.Tree(n + 5, periodic=recipe == "uniform3")
    meshes.apply(t, ops)
    e = t.export()
    check_lists(am, torch, e, n + 5, recipe)
    _, want_t, _, _ = expected(e, n + 5)
    for d in [-1] + list(range(DEPTH_LIMIT + 1)):
        m = t.mesh(d)
        want = want_t[0] if d < 0 else want_t[d + 1]
        assert m.n_tiles == len(want), (recipe, d)
        if m.n_tiles:
            got = np.empty(m.n_tiles, np.int32)
            am.check(am.lib().am_memcpy_d2h(got.ctypes.data, C.c_void_p(m.tiles), 4 * m.n_tiles, None), "d2h")
            assert np.array_equal(got, want), (recipe, d)
