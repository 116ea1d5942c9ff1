"""CPU oracle of the SimpleWave plugin's regridding (test infrastructure only;
nothing in the product imports it).

* octant_criterion: __amr_criterion_1 and __amr_criterion_2
  (SimpleWave.h:91-144) -- per slot and octant o = x + 2y + 4z, the largest
  over the octant's cells of max(|d1| * outer_dx / 2^d, |d2| * outer_dx *
  outer_dx / 4^d) from the decoded words, in float32 with the reference's
  operation order (length = sqrt(x*x + y*y + z*z)); NaN for an unused slot
  or one not active at the sub-step.
* pressures: the host half of amr_criterion (SimpleWave.h:146-187) -- the
  per-slot maxima and the two average pressures, float32 sums in slot
  order.
* modify_nodes: the triage (SimpleWave.h:228-248) on an OracleTree, with
  node_ctor's prolongation of old and new into every added child
  (SimpleWave.h:219-225).  A parent removed earlier in the same call gets no
  child (the product's add_node refuses it; the reference would build one
  from an empty position).

Pinned by tests/test_oracle_regrid.py with hand-made cases: exactly
representable words at single cells, averages worked by hand, and a triage
table on which each factor of the two weights decides at least one tile.
"""
from __future__ import annotations

import numpy as np

F = np.float32


def _depth(p) -> int:
    p = int(p)
    return -1 if p == 0xFFFFFFFF else p >> 27


def _active(substep: int, depth: int, max_depth: int) -> bool:
    return (substep & ((1 << (max_depth - depth)) - 1)) == 0


def octant_criterion(orc, d1w, d2w, positions, n_slots, substep, max_depth, S, outer_dx):
    """(n_slots * 8,) float32."""
    H = S // 2
    out = np.full(n_slots * 8, np.nan, np.float32)
    dx = F(outer_dx)
    for s in range(n_slots):
        d = _depth(positions[s])
        if d == -1 or not _active(substep, d, max_depth):
            continue
        sl = slice(s * S ** 3, (s + 1) * S ** 3)
        v1 = orc.cf3_unpack(d1w[sl]).astype(np.float32)
        v2 = orc.cf3_unpack(d2w[sl]).astype(np.float32)
        l1 = np.sqrt(v1[:, 0] * v1[:, 0] + v1[:, 1] * v1[:, 1] + v1[:, 2] * v1[:, 2])
        l2 = np.sqrt(v2[:, 0] * v2[:, 0] + v2[:, 1] * v2[:, 1] + v2[:, 2] * v2[:, 2])
        c = np.maximum(l1 * dx / F(1 << d), l2 * dx * dx / F(1 << (2 * d))).reshape(S, S, S)  # [z, y, x]
        for o in range(8):
            ox, oy, oz = (o & 1) * H, ((o >> 1) & 1) * H, (o >> 2) * H
            out[s * 8 + o] = max(F(0), c[oz:oz + H, oy:oy + H, ox:ox + H].max())
    return out


def _clamp(f, a, b):
    return F(max(F(a), min(F(f), F(b))))


def pressures(oct_crit, n_used):
    """(node maxima (n_used,), avg1, avg2) as amr_criterion returns them."""
    p1, p2, active = F(0), F(0), 0
    for i in range(n_used):
        for j in range(8):
            v = oct_crit[i * 8 + j]
            if not np.isnan(v):
                p2 = F(p2 + _clamp(v, 0, 2))
    node = np.empty(n_used, np.float32)
    for i in range(n_used):
        t = oct_crit[i * 8]
        for j in range(1, 8):
            b = oct_crit[i * 8 + j]
            t = t if np.isnan(b) else (b if np.isnan(t) else max(t, b))  # fmaxf
        node[i] = t
        if not np.isnan(t):
            p1 = F(p1 + _clamp(t, 0, 2))
            active += 1
    return node, F(p1 / F(active)), F(F(p2 * F(.125)) / F(active))


def modify_nodes(orc, tree, oct_crit, old, new, S, P):
    """The triage on `tree` (an OracleTree); returns (avg1, avg2) after the
    single-slot override, the removed and the added slots."""
    e = tree.export()
    s = e["final_used"] + 1
    node, avg1, avg2 = pressures(oct_crit, s)
    if s == 1:
        avg1 = avg2 = F(.15)
    urgency = F(F(s - 1) / F(tree.max_slots))
    removed, added = [], []
    for i in range(1, s):
        e = tree.export()
        d = _depth(e["positions"][i])
        ratio = F(F(d) / F(e["max_depth"])) if e["max_depth"] else F(np.nan)
        if not ratio > 0:
            continue
        w = F(F(F(F(_clamp(F(F(e["lifetime"][i]) * F(.1)) - F(.2), 0, 1) * urgency) * urgency) * avg1)
              * np.sqrt(ratio))
        if node[i] < w:
            tree.remove_node(i)
            removed.append(i)
    weight = F(F(avg2 * F(3)) / F(F(1) - urgency))
    e = tree.export()
    n8 = (e["final_used"] + 1) * 8
    for i in range(n8):
        e = tree.export()
        if oct_crit[i] > weight and e["children"][i] == -1:
            parent, off = i >> 3, i & 7
            if _depth(e["positions"][parent]) == -1:
                continue
            slot = tree.add_node(parent, off)
            if slot >= 0:
                orc.copy_to_child(old, off, slot, parent, 0, S, P)
                orc.copy_to_child(new, off, slot, parent, 0, S, P)
                added.append(slot)
    return avg1, avg2, removed, added
