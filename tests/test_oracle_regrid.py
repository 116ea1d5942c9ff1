"""The regridding oracle (SimpleWave.h:91-248) pinned by hand-made cases.

* criterion: single known words (3-4-0 and 1-2-2 triangles, exactly
  representable) placed at one cell of chosen octants give |v| dx / 2^d and
  |v| dx^2 / 4^d at exactly that octant (o = x + 2y + 4z), zero words give 0
  and inactive or unused slots give NaN;
* pressures: a criterion table whose averages are worked out by hand
  (clamping at 2, NaN slots left out of the sums and of the count);
* triage: a tree and table chosen so that each factor of the two weights
  decides at least one node -- dropping the lifetime clamp, one urgency
  factor, the square root of the depth ratio, the (1 - urgency) divisor, or
  swapping the two averages changes the removed or added set.
"""
import numpy as np
import pytest

from oracle import regrid as rg

S, P = 16, 3
T = S + 2 * P


def tree_with(orc, n, ops):
    t = orc.OracleTree(n)
    for p, c in ops:
        t.add_node(p, c)
    return t


def put(orc, words, slot, x, y, z, v):
    w = orc.cf3_pack(*v)
    assert np.array_equal(orc.cf3_unpack(w)[0], np.asarray(v, np.float32))  # exactly representable
    words[slot * S ** 3 + (z * S + y) * S + x] = w


def test_criterion_known_words(orc):
    # root, children 0 and 3 (slots 1, 2), child 5 (slot 3) removed again
    t = tree_with(orc, 4, [(0, 0), (0, 3), (0, 5)])
    t.remove_node(3)
    e = t.export()
    assert e["max_depth"] == 1 and e["final_used"] == 2
    d1 = np.zeros(4 * S ** 3, np.uint32)
    d2 = np.zeros(4 * S ** 3, np.uint32)
    assert np.array_equal(orc.cf3_unpack(np.uint32(0))[0], np.zeros(3, np.float32))
    put(orc, d1, 1, S - 1, 0, 0, (3., 4., 0.))       # octant 1: |v| = 5
    put(orc, d2, 1, 0, S - 1, 0, (0., 0., 2.))       # octant 2: |v| = 2
    put(orc, d1, 1, 0, 0, S - 1, (1., 2., 2.))       # octant 4: |v| = 3
    put(orc, d2, 1, 0, 0, S - 1, (0., 0., 1.))       # octant 4, smaller than the d1 term
    put(orc, d2, 0, S // 2, S // 2, S // 2, (0., 3., 4.))  # root octant 7: |v| = 5
    put(orc, d1, 0, 0, S - 1, S - 1, (0., 0., 1.))   # root octant 6: |v| = 1
    dx = 0.5
    c0 = rg.octant_criterion(orc, d1, d2, e["positions"], 4, 0, 1, S, dx).reshape(4, 8)
    want_root = [0, 0, 0, 0, 0, 0, 1 * dx, 5 * dx * dx]
    want_1 = [0, 5 * dx / 2, 2 * dx * dx / 4, 0, 3 * dx / 2, 0, 0, 0]
    assert np.array_equal(c0[0], np.asarray(want_root, np.float32))
    assert np.array_equal(c0[1], np.asarray(want_1, np.float32))
    assert np.array_equal(c0[2], np.zeros(8, np.float32))
    assert np.isnan(c0[3]).all()  # unused slot
    c1 = rg.octant_criterion(orc, d1, d2, e["positions"], 4, 1, 1, S, dx).reshape(4, 8)
    assert np.isnan(c1[0]).all()  # depth 0 inactive at odd sub-steps
    assert np.array_equal(c1[1], c0[1]) and np.isnan(c1[3]).all()


def test_criterion_depth_scaling(orc):
    """At depth 2 the two terms scale by 1/4 and 1/16."""
    t = tree_with(orc, 3, [(0, 0), (1, 7)])
    e = t.export()
    d1 = np.zeros(3 * S ** 3, np.uint32)
    d2 = np.zeros(3 * S ** 3, np.uint32)
    put(orc, d1, 2, 3, 3, 3, (0., 4., 0.))
    put(orc, d2, 2, S - 2, S - 3, S - 4, (0., 0., 8.))
    c = rg.octant_criterion(orc, d1, d2, e["positions"], 3, 0, 2, S, 2.0).reshape(3, 8)
    assert c[2, 0] == np.float32(4 * 2.0 / 4) and c[2, 7] == np.float32(8 * 4.0 / 16)
    assert np.count_nonzero(c[2]) == 2


def triage_table():
    """Per-slot octant values: slot 4 inactive (NaN); maxima above 2 clamp."""
    m = {0: 1., 1: 1., 2: 1., 3: 1., 5: .10, 6: .20, 7: .15, 8: 1., 9: .20, 10: .03}
    c = np.full(11 * 8, np.nan, np.float32)
    for s, v in m.items():
        c[s * 8:(s + 1) * 8] = v
    c[1:8] = 0.         # root: one octant at 1
    c[1 * 8 + 1] = 9.   # slot 1, octant 1: already refined (slot 10)
    c[2 * 8 + 3] = 5.   # slot 2, octant 3: above the add weight
    c[3 * 8 + 6] = 3.   # slot 3, octant 6: just below it
    return c


def triage_tree(orc):
    """Root, its 8 children (slots 1-8), slot 9 under slot 1; 48 sub-steps;
    slot 10 under slot 1; 3 more sub-steps.  Lifetimes: depth 1 -> 26,
    slot 9 -> 51, slot 10 -> 3."""
    t = tree_with(orc, 20, [(0, c) for c in range(8)] + [(1, 0)])
    for _ in range(48):
        t.increment_timer()
    assert t.add_node(1, 1) == 10
    for _ in range(3):
        t.increment_timer()
    e = t.export()
    assert list(e["lifetime"][1:11]) == [26] * 8 + [51, 3]
    return t


def test_pressures_by_hand():
    node, avg1, avg2 = rg.pressures(triage_table(), 11)
    assert np.array_equal(node[[0, 1, 2, 3, 5]], np.asarray([1, 9, 5, 3, .1], np.float32)) and np.isnan(node[4])
    # per-slot maxima clamped to 2: 1+2+2+2+.1+.2+.15+1+.2+.03 = 8.68 over 10 slots
    assert avg1 == pytest.approx(0.868, rel=1e-6)
    # octants clamped to 2: 1 + 3*9 + .8 + 1.6 + 1.2 + 8 + 1.6 + .24 = 41.44, /8, /10
    assert avg2 == pytest.approx(0.518, rel=1e-6)


def test_triage_by_hand(orc):
    """urgency 10/20; removal weights: depth 1 .25*.868*sqrt(.5) = .1534,
    slot 9 .25*.868 = .217, slot 10 .1*.217 = .0217; add weight
    .518*3/.5 = 3.108.  Slots 5 (.10), 7 (.15) and 9 (.20) go; 6 (.20) and
    10 (.03) stay; slot 2 octant 3 (5.0) gets a child in the first free
    slot, 5, prolonged from slot 2; slot 3 octant 6 (3.0) and the refined
    slot 1 octant 1 get none."""
    t = triage_tree(orc)
    old = np.zeros(20 * T ** 3, np.float32)
    old[2 * T ** 3:3 * T ** 3] = np.arange(T ** 3, dtype=np.float32)
    new = old * 2
    avg1, avg2, removed, added = rg.modify_nodes(orc, t, triage_table(), old, new, S, P)
    assert (removed, added) == ([5, 7, 9], [5])
    e = t.export()
    assert e["parent"][5] == 2 and e["children"][2 * 8 + 3] == 5 and e["lifetime"][5] == 0
    assert e["children"][1 * 8 + 0] == -1 and int(e["positions"][7]) == 0xFFFFFFFF
    want = np.zeros(20 * T ** 3, np.float32)
    want[2 * T ** 3:3 * T ** 3] = old[2 * T ** 3:3 * T ** 3]
    orc.copy_to_child(want, 3, 5, 2, 0, S, P)
    assert np.array_equal(old, want) and np.array_equal(new, want * 2)


def test_single_slot_override(orc):
    """With only the root in use both averages are .15, so an octant
    refines above .45 (urgency 0)."""
    for v, kids in ((.5, 1), (.4, 0)):
        t = orc.OracleTree(4)
        c = np.zeros(8, np.float32)
        c[6] = v
        old = np.zeros(4 * T ** 3, np.float32)
        avg1, avg2, removed, added = rg.modify_nodes(orc, t, c, old, old.copy(), S, P)
        assert (avg1, avg2) == (np.float32(.15), np.float32(.15))
        assert removed == [] and len(added) == kids
        if kids:
            assert t.export()["children"][6] == added[0]
