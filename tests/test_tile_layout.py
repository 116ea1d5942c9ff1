"""The interior-first tile layout (am_tile_layout): the library's cell map
(am_tile_cell_offsets, the same function the kernels use) equals the numpy
restatement, is a bijection onto the slot, and puts hand-worked cells where
the header's definition says; the slot stride is T^3 rounded up to 32."""
import ctypes as C

import numpy as np
import pytest


@pytest.mark.parametrize("S,P", [(16, 3), (16, 2), (28, 2)])
def test_cell_map(am, S, P):
    g = am.geometry(S, P, 2 * min(P, 2), 64.0, 0.25)
    T = S + 2 * P
    lib = am.lib()
    got = np.empty(T ** 3, np.uint32)
    am.check(lib.am_tile_cell_offsets(C.byref(g), 1, got.ctypes.data), "offsets")
    assert np.array_equal(got.astype(np.int64), am.interior_first_index(S, P))
    assert np.array_equal(np.sort(got), np.arange(T ** 3))
    pad = np.empty(T ** 3, np.uint32)
    am.check(lib.am_tile_cell_offsets(C.byref(g), 0, pad.ctypes.data), "offsets")
    assert np.array_equal(pad, np.arange(T ** 3))
    stride = lib.am_tile_stride(C.byref(g), 1)
    assert stride % 32 == 0 and T ** 3 <= stride < T ** 3 + 32
    assert lib.am_tile_stride(C.byref(g), 0) == T ** 3
    at = lambda x, y, z: int(got[(z * T + y) * T + x])
    S2, S3 = S * S, S ** 3
    # interior: row-major from the first interior cell
    assert at(P, P, P) == 0 and at(P + 1, P, P) == 1 and at(P, P + 1, P) == S and at(P, P, P + 1) == S2
    assert at(P + S - 1, P + S - 1, P + S - 1) == S3 - 1
    # face slabs: +x layer 0 at (a, b) = (y, z); -x is face 3; layers count outward
    assert at(P + S, P, P) == S3
    assert at(P + S + 1, P + 2, P + 5) == S3 + ((0 * P + 1) * S + 5) * S + 2
    assert at(P - 1, P + 3, P + 4) == S3 + ((3 * P + 0) * S + 4) * S + 3
    assert at(0, P + 3, P + 4) == S3 + ((3 * P + P - 1) * S + 4) * S + 3
    assert at(P + 7, P + S, P + 2) == S3 + ((1 * P + 0) * S + 2) * S + 7      # +y: (a, b) = (x, z)
    assert at(P + 7, P + 9, P - 2) == S3 + ((5 * P + 1) * S + 9) * S + 7     # -z: (a, b) = (x, y)
    # every face slab row (fixed layer and b, a = 0..S-1) is S consecutive offsets
    for f in range(6):
        k, high = f % 3, f < 3
        for l in range(P):
            g_ = P + S + l if high else P - 1 - l
            for b in (0, S // 2, S - 1):
                cells = []
                for a in range(S):
                    c = [0, 0, 0]
                    c[k] = g_
                    ax_a, ax_b = [i for i in range(3) if i != k]
                    c[ax_a], c[ax_b] = P + a, P + b
                    cells.append(at(*c))
                assert cells == list(range(cells[0], cells[0] + S)), (f, l, b)
                assert cells[0] == S3 + ((f * P + l) * S + b) * S
    E0 = S3 + 6 * P * S2
    assert at(0, 0, P) >= E0 and at(0, 0, 0) >= E0 + 3 * S * (2 * P) ** 2


def test_bad_arguments(am):
    g = am.geometry(16, 3, 4, 64.0, 0.25)
    lib = am.lib()
    assert lib.am_tile_stride(C.byref(g), 2) == -1
    out = np.empty(8, np.uint32)
    assert lib.am_tile_cell_offsets(C.byref(g), 5, out.ctypes.data) == -1
    assert lib.am_tile_layout_convert(16, 32, 1, C.byref(g), 0, 0, None) == -1
