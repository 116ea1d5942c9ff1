"""Octree encodings, neighbour search, timer and slot management.

The oracle (a C restatement of AMR.h:22-294) is pinned by the survey's known
boundary table, by closed forms of the position encoding and the sub-cycling
counter, and by a brute-force geometric neighbour search; the product's C++
octree (include/adaptivemesh/AMR.hpp through the C ABI) must then equal the
oracle array for array.
"""
import itertools

import numpy as np
import pytest

from tests import meshes

KEYS = ("parent", "children", "positions", "boundaries", "lifetime", "scalars")


def decode_boundary(orc, b):
    nd = int(b) & 0xFFFFFFFF
    tgt = nd & 0xFFFFFF
    if tgt & (1 << 23):
        tgt -= 1 << 24
    depth = nd >> 28
    return tgt, (nd >> 24) & 15, (-1 if depth == 15 else depth), orc.cf3_unpack(int(b) >> 32)[0]


def centre(pos):
    """Tile centre in units of outer_size and its width (AMR.h:33-43)."""
    d = pos >> 27
    c = np.zeros(3)
    for i in range(d):
        o = (pos >> (3 * i)) & 7
        c += np.array([1 if o & 1 else -1, 1 if o & 2 else -1, 1 if o & 4 else -1]) * 0.25 / 2 ** i
    return c, 1.0 / 2 ** d


def test_survey_boundary_table(orc):
    t = orc.OracleTree(9)
    meshes.apply(t, meshes.uniform(1)[0])
    b = t.export()["boundaries"].reshape(-1, 6)[1]
    got = [decode_boundary(orc, x) for x in b]
    assert [g[0] for g in got] == [2, 3, 5, 0, 0, 0]
    assert [g[1] for g in got] == [0, 0, 0, 1, 1, 1]
    for g, r in zip(got[:3], ([-1, 0, 0], [0, -1, 0], [0, 0, -1])):
        assert g[3].tolist() == r
    for g in got[3:]:
        assert g[3].tolist() == [-0.25, -0.25, -0.25]


def test_position_encoding_closed_form(orc):
    t = orc.OracleTree(600)
    ops, n = meshes.uniform(3)
    slots = meshes.apply(t, ops)
    e = t.export()
    for (p, c), s in zip(ops, slots):
        pp = int(e["positions"][p])
        d = pp >> 27
        assert int(e["positions"][s]) == ((d + 1) << 27) | (pp & ((1 << 27) - 1)) | (c << (3 * d))
        assert e["parent"][s] == p and e["children"][8 * p + c] == s


def brute_neighbour(pos, n_used, face, periodic):
    """Deepest existing tile no deeper than `node` that contains a point just
    beyond the face centre (0 = the root, standing in for the outer
    boundary), and the periodic shift applied to that point."""
    axis, sgn = face % 3, (1 if face < 3 else -1)
    out, shifts = [], []
    for node in range(n_used):
        c, w = centre(int(pos[node]))
        p = c.copy()
        p[axis] += sgn * w * 0.75
        shift = np.zeros(3)
        if np.any(np.abs(p) > 0.5):
            if not periodic:
                out.append(0)
                shifts.append(shift)
                continue
            q = (p + 0.5) % 1.0 - 0.5
            shift, p = q - p, q
        shifts.append(shift)
        best, bd = 0, -1
        for j in range(n_used):
            pj = int(pos[j])
            if pj == 0xFFFFFFFF:
                continue
            cj, wj = centre(pj)
            dj = pj >> 27
            if dj <= (int(pos[node]) >> 27) and np.all(np.abs(p - cj) < wj / 2) and dj > bd:
                best, bd = j, dj
        out.append(best)
    return out, shifts


@pytest.mark.parametrize("recipe", ["uniform2", "nested4", "graded"])
@pytest.mark.parametrize("periodic", [False, True])
def test_neighbour_search_brute_force(orc, recipe, periodic):
    ops, n = {"uniform2": meshes.uniform(2), "nested4": meshes.nested(4), "graded": meshes.graded()}[recipe]
    t = orc.OracleTree(n, periodic)
    meshes.apply(t, ops)
    e = t.export()
    bnd = e["boundaries"].reshape(-1, 6)
    for face in range(6):
        want, shifts = brute_neighbour(e["positions"], n, face, periodic)
        for node in range(1, n):
            tgt, dd, depth, rel = decode_boundary(orc, bnd[node, face])
            assert tgt == want[node], (node, face)
            pn, pt = int(e["positions"][node]), int(e["positions"][tgt])
            assert depth == pn >> 27 and dd == (pn >> 27) - (pt >> 27)
            # rel_pos = 2^d_target (c_this - c_target) in units of outer_size,
            # with the target image that lies across the face
            cn, _ = centre(pn)
            ct, _ = centre(pt)
            r = (cn - (ct - shifts[node])) * 2 ** (pt >> 27)
            if tgt != 0 or periodic:
                assert np.sign(r[face % 3]) == (-1 if face < 3 else 1)
            assert np.allclose(rel, r, atol=0), (node, face, rel, r)
        # root has no boundary
        assert decode_boundary(orc, bnd[0, face])[0] == -1


def test_equal_depth_symmetry(orc):
    t = orc.OracleTree(585, True)
    meshes.apply(t, meshes.uniform(3)[0])
    bnd = t.export()["boundaries"].reshape(-1, 6)
    leaves = range(73, 585)
    for node in leaves:
        for f in range(6):
            tgt, dd, _, rel = decode_boundary(orc, bnd[node, f])
            assert dd == 0
            back = decode_boundary(orc, bnd[tgt, (f + 3) % 6])[0]
            assert back == node
            assert sorted(np.abs(rel).tolist()) == [0.0, 0.0, 1.0]


@pytest.mark.parametrize("max_depth", [1, 2, 3, 4, 5])
def test_timer_closed_form(orc, max_depth):
    """Counter runs 0..2^max-1; depth d is active when t % 2^(max-d) == 0;
    lifetime counts the activations (AMR.h:180-188, 640-644)."""
    t = orc.OracleTree(64)
    parent = 0
    for d in range(max_depth):
        parent = t.add_node(parent, 0)
    seen = []
    for _ in range(2 * 2 ** max_depth):
        seen.append(t.export()["timer"])
        t.increment_timer()
    assert seen == list(range(2 ** max_depth)) * 2
    life = t.export()["lifetime"]
    for d in range(max_depth + 1):
        assert life[d] == 2 * 2 ** d


def test_timer_rescale_on_depth_change(orc):
    """Adding a deeper tile shifts the counter left; removing it shifts right."""
    t = orc.OracleTree(32)
    a = t.add_node(0, 0)
    t.increment_timer()
    assert t.export()["timer"] == 1
    b = t.add_node(a, 3)
    assert t.export()["timer"] == 2 and t.export()["max_depth"] == 2
    t.remove_node(b)
    assert t.export()["timer"] == 1 and t.export()["max_depth"] == 1


def test_refusals_leave_tree_unchanged(orc, am):
    for T in (orc.OracleTree(3), am.Tree(3)):
        assert T.add_node(0, 0) == 1
        assert T.add_node(0, 1) == 2
        before = T.export()
        assert T.add_node(0, 2) < 0  # no free slot (AMR.h:210)
        after = T.export()
        for k in KEYS:
            assert np.array_equal(before[k], after[k])
    t = am.Tree(4)
    assert t.add_node(0, 0) == 1
    assert t.add_node(0, 0) == -5  # occupied
    t = am.Tree(16)
    p = 0
    for d in range(8):
        p = t.add_node(p, 0)
    assert t.add_node(p, 0) == -4  # depth limit (AMR.h:211)


@pytest.mark.parametrize("seed", range(6))
def test_product_tree_equals_oracle_random(orc, am, seed):
    """Random add/remove sequences: the product octree equals the oracle
    after every operation."""
    rng = np.random.default_rng(seed)
    n = 80
    a, b = am.Tree(n, periodic=bool(seed & 1)), orc.OracleTree(n, periodic=bool(seed & 1))
    for step in range(120):
        eb = b.export()
        used = [i for i in range(n) if eb["positions"][i] != 0xFFFFFFFF]
        if rng.random() < 0.8 or len(used) == 1:
            p = int(rng.choice(used))
            if (int(eb["positions"][p]) >> 27) >= 6:
                continue
            c = int(rng.integers(0, 8))
            ra, rb = a.add_node(p, c), b.add_node(p, c)
            assert (ra < 0) == (rb < 0) and (ra < 0 or ra == rb)
        else:
            v = int(rng.choice(used[1:]))
            a.remove_node(v)
            b.remove_node(v)
        if rng.random() < 0.3:
            a.increment_timer()
            b.increment_timer()
        ea, eb = a.export(), b.export()
        for k in KEYS:
            assert np.array_equal(ea[k], eb[k]), (seed, step, k)


@pytest.mark.parametrize("recipe", ["uniform3", "nested6", "graded"])
def test_product_tree_equals_oracle_meshes(orc, am, recipe):
    ops, n = {"uniform3": meshes.uniform(3), "nested6": meshes.nested(6), "graded": meshes.graded()}[recipe]
    a, b = am.Tree(n + 3), orc.OracleTree(n + 3)
    assert meshes.apply(a, ops) == meshes.apply(b, ops)
    for _ in range(5):
        a.increment_timer()
        b.increment_timer()
    ea, eb = a.export(), b.export()
    for k in KEYS:
        assert np.array_equal(ea[k], eb[k]), k
    for packed, md in itertools.product([0, 1, 0o7777, 0o1234567, 12345678], [0, 1, 3, 9, 0xFFFFFFFF]):
        assert a.find_node_idx(packed, md) == b.find_node_idx(packed, md)
