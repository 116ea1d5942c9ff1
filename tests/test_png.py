"""am_write_png (the lodepng_encode32_file step of save_image,
image_process.h:55-62) on the CPU: the file is a valid PNG by the
specification -- signature, IHDR for 8-bit RGBA, chunk CRCs, one zlib
stream whose Adler-32 checks -- and decodes (Python's zlib) to exactly the
pixels written, R in the low byte."""
import struct
import zlib

import numpy as np
import pytest

import adaptivemesh_amd as am


def chunks(b):
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    at, out = 8, []
    while at < len(b):
        n, = struct.unpack(">I", b[at:at + 4])
        typ, data = b[at + 4:at + 8], b[at + 8:at + 8 + n]
        crc, = struct.unpack(">I", b[at + 8 + n:at + 12 + n])
        assert crc == zlib.crc32(typ + data), typ
        out.append((typ, data))
        at += 12 + n
    return out


@pytest.mark.parametrize("w,h", [(1, 1), (200, 100), (1024, 70)])
def test_png_round_trip(tmp_path, w, h):
    rng = np.random.default_rng(w * h)
    px = rng.integers(0, 2 ** 32, (h, w), dtype=np.uint32)
    path = tmp_path / "img.png"
    am.write_png(path, px)
    cs = chunks(path.read_bytes())
    assert [c[0] for c in cs] == [b"IHDR", b"IDAT", b"IEND"]
    assert struct.unpack(">IIBBBBB", cs[0][1]) == (w, h, 8, 6, 0, 0, 0)
    raw = zlib.decompress(cs[1][1])  # checks the Adler-32 too
    rows = np.frombuffer(raw, np.uint8).reshape(h, 4 * w + 1)
    assert np.all(rows[:, 0] == 0)
    got = rows[:, 1:].copy().view("<u4")
    assert np.array_equal(got, px)


def test_png_bad_arguments(tmp_path):
    with pytest.raises(am.AMError) as ex:
        am.write_png(tmp_path / "no" / "such" / "dir.png", np.zeros((2, 2), np.uint32))
    assert ex.value.code == -10
    with pytest.raises(am.AMError) as ex:
        am.write_png(tmp_path / "x.png", np.zeros((0, 3), np.uint32))
    assert ex.value.code == -1
