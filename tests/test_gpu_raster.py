"""Slice rasteriser and RGBA8 encoder on the GPU against the oracle, bit for
bit (AMR.h:1313-1353, image_process.h:10-54): random fields on graded and
uniform meshes, image sizes that are and are not multiples of the 16x16
workgroup, boxes on and off, slices at the root's faces; the drivers'
raster calls and a PNG written from them."""
import ctypes as C
import zlib

import numpy as np
import pytest

from tests import meshes
from tests.test_png import chunks

pytestmark = pytest.mark.gpu


def dev(torch, a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a.copy()).cuda()


@pytest.mark.parametrize("S,P", [(16, 3), (28, 2)])
@pytest.mark.parametrize("recipe", ["graded", "uniform2"])
def test_rasterise_bit_exact(am, orc, torch_cuda, S, P, recipe):
    torch = torch_cuda
    print(f"raster S={S} P={P} {recipe}", flush=True)
    g = am.geometry(S, P, 2 * P, 100.0, 0.4)
    ops, n = {"graded": meshes.graded(), "uniform2": meshes.uniform(2)}[recipe]
    tree, otree = am.Tree(n), orc.OracleTree(n)
    meshes.apply(tree, ops)
    meshes.apply(otree, ops)
    e = otree.export()
    T = S + 2 * P
    f = np.random.default_rng(S + n).standard_normal(n * T ** 3).astype(np.float32) * 0.3
    df = dev(torch, f)
    m = tree.mesh(-1)
    for (w, h, z, boxes) in ((96, 80, 1e-9, True), (100, 75, -0.5, False), (256, 256, 0.49999997, True),
                             (33, 517, 0.1234, False)):
        raster = torch.empty(w * h * 4, dtype=torch.float32, device="cuda")
        px = torch.empty(w * h, dtype=torch.int32, device="cuda")
        am.check(am.lib().am_rasterise_slice(df.data_ptr(), C.byref(m), C.byref(g), w, h, z, 1 if boxes else 0,
                                             raster.data_ptr(), am.stream_handle()), "am_rasterise_slice")
        am.check(am.lib().am_encode_rgba8(raster.data_ptr(), px.data_ptr(), w * h, am.stream_handle()), "encode")
        torch.cuda.synchronize()
        want = orc.rasterise(f, e["children"], w, h, S, P, z, boxes)
        got = raster.cpu().numpy().reshape(h, w, 4)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (w, h, z)
        assert np.array_equal(px.cpu().numpy().view(np.uint32).reshape(h, w), orc.encode_rgba8(want))


def test_encode_nan_and_clamp(am, orc, torch_cuda):
    torch = torch_cuda
    r = np.array([[.5, .5, .5, 1], [np.nan, 0, 0, 1], [2, -1, .999, 1], [0, 0, 0, np.nan], [.004, 0, 1, .5]],
                 np.float32)
    dr = dev(torch, r.reshape(-1))
    px = torch.empty(5, dtype=torch.int32, device="cuda")
    am.check(am.lib().am_encode_rgba8(dr.data_ptr(), px.data_ptr(), 5, am.stream_handle()), "encode")
    torch.cuda.synchronize()
    assert np.array_equal(px.cpu().numpy().view(np.uint32), orc.encode_rgba8(r))


def test_rasterise_rejects_bad_arguments(am, torch_cuda):
    torch = torch_cuda
    g = am.geometry(16, 3, 6, 100.0, 0.4)
    tree = am.Tree(2)
    m = tree.mesh(-1)
    f = torch.zeros(2 * 22 ** 3, dtype=torch.float32, device="cuda")
    r = torch.empty(64 * 4, dtype=torch.float32, device="cuda")
    L, st = am.lib(), am.stream_handle()
    for z in (0.5, -0.50001, float("nan")):
        assert L.am_rasterise_slice(f.data_ptr(), C.byref(m), C.byref(g), 8, 8, z, 1, r.data_ptr(), st) == -1
    assert L.am_rasterise_slice(f.data_ptr(), C.byref(m), C.byref(g), 0, 8, 0.0, 1, r.data_ptr(), st) == -1
    assert L.am_rasterise_slice(f.data_ptr(), C.byref(m), C.byref(g), 8, 8, 0.0, 1, r.data_ptr() + 4, st) == -6


def test_driver_raster_and_png(am, orc, torch_cuda, tmp_path):
    S, P = 16, 3
    T = S + 2 * P
    g = am.geometry(S, P, 6, 100.0, 0.4)
    ops, n = meshes.graded()
    sim = am.WaveSim(g, n)
    f = np.random.default_rng(3).standard_normal(n * T ** 3).astype(np.float32) * 0.2
    sim.set_field(1, f)
    meshes.apply(sim, ops)  # prolongs the root's field into the children
    e = sim.export()
    field = sim.get_field(1)
    px = sim.raster(1, 200, 120, 0.05)
    want = orc.encode_rgba8(orc.rasterise(field, e["children"], 200, 120, S, P, 0.05, True))
    assert np.array_equal(px, want)
    path = tmp_path / "slice.png"
    sim.save_image(path, 1, 200, 120, 0.05)
    cs = chunks(path.read_bytes())
    raw = np.frombuffer(zlib.decompress(cs[1][1]), np.uint8).reshape(120, 801)
    assert np.array_equal(raw[:, 1:].copy().view("<u4"), want)

    bs = am.BssnSim(am.geometry(S, P, 4, 16.0, 0.25), n, am.bssn_options())
    meshes.apply(bs, ops)
    chi = np.random.default_rng(4).standard_normal(n * T ** 3).astype(np.float32)
    bs.set_field(0, 0, chi)
    got = bs.raster(0, which=0, width=64, height=48, z_slice=-0.2, boxes=False)
    assert np.array_equal(got, orc.encode_rgba8(orc.rasterise(chi, bs.export()["children"], 64, 48, S, P, -0.2,
                                                              False)))


def test_encode_float_and_word_images(am, orc, torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(11)
    img = rng.uniform(-0.5, 1.5, 4099).astype(np.float32)
    img[::97] = np.nan
    words = orc.cf3_pack_array(rng.uniform(-0.5, 1.5, (4099, 3)))
    for arr, fn, want in ((img, "am_encode_rgba8_f32", orc.encode_rgba8_f32(img)),
                          (words, "am_encode_rgba8_cf3", orc.encode_rgba8_cf3(words))):
        d = dev(torch, arr)
        px = torch.empty(arr.size, dtype=torch.int32, device="cuda")
        am.check(getattr(am.lib(), fn)(d.data_ptr(), px.data_ptr(), arr.size, am.stream_handle()), fn)
        torch.cuda.synchronize()
        assert np.array_equal(px.cpu().numpy().view(np.uint32), want), fn
