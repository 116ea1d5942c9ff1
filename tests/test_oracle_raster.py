"""The slice rasteriser and RGBA8 encoder of the oracle (AMR.h:1313-1353,
image_process.h:14-21), pinned by cases worked out by hand.

* Interpolation and descent: on every tile the field is one global
  quadratic q(X, Y, Z) sampled where the rasteriser puts the cells (cell i
  of a tile of width w at its lower edge + (i - P) w / S, i.e. half a cell
  below the cell centres the initial data uses, as in the reference), so
  the 10-point form is exact and every pixel is q at its point + 1/2, on a
  graded mesh at several slices.  A point exactly on the upper face of a
  lower octant (t = 0 at some level) is taken to the lower face of that
  octant, since fracf(2 * 0) - 1/2 = -1/2 (AMR.h:1324): those pixels sample
  the octant's far edge, checked on its own.
* Boxes: on the root's 8 children at 64x64 pixels a pixel is red exactly
  when x mod 32 or y mod 32 is in {0, 1, 2, 30, 31} (|t| >= 1/2 - 4/64 in
  tile units), 1180 pixels.
* Encoding: known words (grey 1/2 -> 0xFF7F7F7F, clamping, truncation,
  magenta for NaN) for 4-float pixels, floats and compressed_float3 words.
"""
import numpy as np
import pytest

from tests import meshes

S, P = 16, 3
T = S + 2 * P


def q(X, Y, Z):
    return (0.2 + 0.3 * X - 0.25 * Y + 0.4 * Z + 0.5 * X * Y - 0.3 * X * Z + 0.2 * Y * Z + 0.35 * X * X
            - 0.15 * Y * Y + 0.1 * Z * Z)


def raster_grid_field(e, n):
    out = np.zeros((n, T, T, T), np.float64)
    r = (np.arange(T) - P) / S
    for s, pos in enumerate(e["positions"]):
        pos = int(pos)
        if pos == 0xFFFFFFFF:
            continue
        d = pos >> 27
        lo, w = np.full(3, -0.5), 1.0
        for i in range(d):
            o = (pos >> (3 * i)) & 7
            w /= 2
            lo += np.array([o & 1, (o >> 1) & 1, (o >> 2) & 1], dtype=float) * w
        Z, Y, X = np.meshgrid(lo[2] + r * w, lo[1] + r * w, lo[0] + r * w, indexing="ij")
        out[s] = q(X, Y, Z)
    return out.astype(np.float32).reshape(-1)


@pytest.mark.parametrize("z_slice", [1e-9, -0.3, 0.41, -0.5])
def test_quadratic_field_is_reproduced(orc, z_slice):
    t = orc.OracleTree(25)
    ops, n = meshes.graded()
    meshes.apply(t, ops)
    e = t.export()
    f = raster_grid_field(e, n)
    W, H = 96, 80
    img = orc.rasterise(f, e["children"], W, H, S, P, z_slice, boxes=False)
    X = np.arange(W, dtype=np.float32) / np.float32(W) - np.float32(.5)
    Y = np.arange(H, dtype=np.float32) / np.float32(H) - np.float32(.5)
    want = q(X[None, :].astype(np.float64), Y[:, None].astype(np.float64), float(np.float32(z_slice))) + 0.5
    on_face = lambda c: (c * 8) == np.round(c * 8)  # multiples of 1/8: tile faces down to depth 3
    inside = ~on_face(Y)[:, None] & ~on_face(X)[None, :]
    assert inside.sum() > 0.8 * W * H
    assert np.max(np.abs(img[..., 0] - want)[inside]) < 2e-5
    assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 0], img[..., 2])
    assert np.all(img[..., 3] == 1.0)


def test_point_on_an_octant_face_goes_to_its_far_edge(orc):
    """Root + 8 children, 64x64: column 32 is X = 0, the upper face of the
    lower octants, and samples X = -1/2; row 32 likewise samples Y = -1/2."""
    t = orc.OracleTree(9)
    for c in range(8):
        t.add_node(0, c)
    e = t.export()
    f = raster_grid_field(e, 9)
    img = orc.rasterise(f, e["children"], 64, 64, S, P, 0.2, boxes=False)[..., 0]
    X = np.arange(64) / 64 - 0.5
    assert np.max(np.abs(img[5, 32] - (q(-0.5, X[5], 0.2) + 0.5))) < 2e-5
    assert np.max(np.abs(img[32, 7] - (q(X[7], -0.5, 0.2) + 0.5))) < 2e-5
    assert np.max(np.abs(img[5, 31] - (q(X[31], X[5], 0.2) + 0.5))) < 2e-5


def test_tile_edge_boxes(orc):
    t = orc.OracleTree(9)
    for c in range(8):
        t.add_node(0, c)
    e = t.export()
    f = np.zeros(9 * T ** 3, np.float32)
    img = orc.rasterise(f, e["children"], 64, 64, S, P, 1e-9, boxes=True)
    red = (img[..., 0] == 1) & (img[..., 1] == 0) & (img[..., 2] == 0)
    band = np.isin(np.arange(64) % 32, [0, 1, 2, 30, 31])
    assert np.array_equal(red, band[None, :] | band[:, None])
    assert red.sum() == 1180
    assert np.all(img[~red][:, 0] == 0.5)  # zero field: grey 1/2


def test_encode_known_words(orc):
    r = np.array([[.5, .5, .5, 1], [np.nan, 0, 0, 1], [2, -1, .999, 1], [0, 0, 0, 0], [1, 1, 1, 1],
                  [.003, .004, .9999, .5], [0, 0, 0, np.nan]], np.float32)
    w = orc.encode_rgba8(r)
    assert [int(x) for x in w] == [0xFF7F7F7F, 0xFFFF00FF, 0xFFFE00FF, 0, 0xFFFFFFFF,
                                   (127 << 24) | (254 << 16) | (1 << 8), 0xFFFF00FF]


def test_encode_grey_and_word_known_answers(orc):
    g = orc.encode_rgba8_f32(np.array([.5, np.nan, -1, 2, .2], np.float32))
    assert [int(x) for x in g] == [0xFF7F7F7F, 0xFFFF00FF, 0xFF000000, 0xFFFFFFFF, 0xFF333333]
    w = [orc.cf3_pack(1., .5, 0.), orc.cf3_pack(-1., 2., .25), orc.cf3_pack(0., 0., 0.)]
    c = orc.encode_rgba8_cf3(np.array(w, np.uint32))
    assert [int(x) for x in c] == [0xFF007FFF, 0xFF3FFF00, 0xFF000000]
