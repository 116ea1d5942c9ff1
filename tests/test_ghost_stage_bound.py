"""The tricubic coarse-face fill (k_ghost_tricubic) stages, per face, the
block of target nodes its stencils read in a fixed LDS array.  This checks
on the CPU that every block fits: for each tile geometry, depth difference
0..8, face, and relative position of the fine tile in its target (every
multiple of 2^-ddiff in [-1, 1] per axis), the clamped stencil bases are
computed in float32 with the kernel's operations, and the block of
nodes (first base - 1 .. last base + 2 per axis) times the staged time
levels (3 one level or more coarser, 2 at equal depth) is compared with the
capacity: max(2 fields x 3 levels x ET^2 EN, 2 x (S + 4)^2 (P + 4)) floats,
ET = (S + 1) / 2 + 4, EN = P / 2 + 5."""
import numpy as np
import pytest


def bases(S, P, cnt, wlo, rel_scale, rl, root):
    T = S + 2 * P
    f = np.float32
    w = (np.arange(cnt) + wlo).astype(np.float32)
    rp = (w - f(P) + f(.5)) * (f(1) / f(S)) - f(.5)
    rp = rp * f(rel_scale) + f(rl)
    rp = (rp + f(.5)) * f(S) + f(P) - f(.5)
    lo_c = 1.0 if root else float(P + 1)
    hi_c = T - 3 if root else T - P - 3
    ri = np.minimum(np.maximum(rp, f(lo_c)).astype(np.uint32), hi_c)
    assert np.all(np.diff(ri.astype(np.int64)) >= 0)  # the bases rise with the index
    return ri


@pytest.mark.parametrize("S,P", [(16, 3), (16, 2), (28, 2)])
def test_staged_block_fits(S, P):
    ET, EN = (S + 1) // 2 + 4, P // 2 + 5
    cap = max(2 * 3 * ET * ET * EN, 2 * (S + 4) ** 2 * (P + 4))
    worst = 0
    for dd in range(9):
        rel = 1.0 / 2 ** dd
        offs = [k * rel for k in range(-2 ** dd, 2 ** dd + 1)] if dd <= 4 else \
            [k * rel for k in (-2 ** dd, -2 ** dd + 1, -1, 0, 1, 2 ** dd - 1, 2 ** dd)]
        nlev = 3 if dd else 2
        for d in range(6):
            na = d % 3
            cnt = [P if k == na else S for k in range(3)]
            wlo = [P if k != na else (P + S if d < 3 else 0) for k in range(3)]
            for root in (False, True):
                ext = []
                for k in range(3):
                    e = 0
                    for rl in offs:
                        ri = bases(S, P, cnt[k], wlo[k], rel, rl, root)
                        e = max(e, int(ri[-1]) - int(ri[0]) + 4)
                    ext.append(e)
                need = nlev * ext[0] * ext[1] * ext[2]
                worst = max(worst, need)
                assert need <= cap, (S, P, dd, d, root, ext)
    print(f"S={S} P={P}: largest staged block {worst} of {cap} floats")


def test_staging_index_division():
    """The staging loop splits an element index i by d (fields x levels,
    block, row, plane sizes) as floor((i + 1/2) * fl(1/d)) in float32: equal
    to i // d for every i and d up to the largest capacity (12288 floats)."""
    n = 12288
    i = np.arange(n, dtype=np.int64)
    fi = i.astype(np.float32) + np.float32(.5)
    for d in range(1, n + 1):
        inv = np.float32(1) / np.float32(d)
        q = (fi * inv).astype(np.int64)  # float32 product, truncated as (int)
        assert np.array_equal(q, i // d), d
