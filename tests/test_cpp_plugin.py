"""The drop-in C++ headers (include/adaptivemesh/*.hpp) from user programs:
tests/cpp/wave_plugin_demo.cpp and bssn_plugin_demo.cpp are the reference's
wave_test and BSSN_test (kernel.cu:45-111) written against them.  On the CPU
both must compile with a plain C++17 compiler and no warnings (the headers
hold no device code).  On the GPU the wave program's fields, seed and every
saved PNG must equal the same run driven through the Python bindings, and
the BSSN program (the reference-shaped BSSN_AMR_data plugin: metric
derivatives, traces and Christoffel words over a 6-level nested mesh, two
timesteps) must equal the oracle driven through the same phases and seeds,
bit for bit.  tests/cpp/stream_pool_demo.cpp drives component bundles through
the wrappers' stream-pool forms (AMR.h:902-933, 1020-1051, 1083-1115,
1304-1310) and through their one-buffer forms on the default stream; the two
runs must leave every buffer and the seed the same."""
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

from tests import fields
from tests.test_png import chunks

ROOT = Path(__file__).resolve().parents[1]
DEMOS = {"wave": ROOT / "tests" / "cpp" / "wave_plugin_demo.cpp",
         "bssn": ROOT / "tests" / "cpp" / "bssn_plugin_demo.cpp",
         "pool": ROOT / "tests" / "cpp" / "stream_pool_demo.cpp"}


def build(tmp_path, werror=False, demo="wave"):
    exe = tmp_path / f"{demo}_plugin_demo"
    lib = ROOT / "adaptivemesh_amd"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", str(ROOT / "include"), str(DEMOS[demo]), "-L",
           str(lib), "-ladaptivemesh_amd", f"-Wl,-rpath,{lib}", "-o", str(exe)]
    if werror:
        cmd.insert(1, "-Werror")
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    return exe


@pytest.mark.parametrize("demo", sorted(DEMOS))
def test_headers_compile_as_plain_cpp(am, tmp_path, demo):
    assert build(tmp_path, werror=True, demo=demo).exists()


def png_pixels(path):
    cs = chunks(Path(path).read_bytes())
    w, h = int.from_bytes(cs[0][1][:4], "big"), int.from_bytes(cs[0][1][4:8], "big")
    raw = np.frombuffer(zlib.decompress(cs[1][1]), np.uint8).reshape(h, 4 * w + 1)
    return raw[:, 1:].copy().view("<u4")


@pytest.mark.gpu
def test_demo_matches_python_driver(am, orc, torch_cuda, tmp_path):
    S, P, n, steps, W, H = 16, 3, 128, 3, 96, 64
    out = tmp_path / "out"
    out.mkdir()
    exe = build(tmp_path)
    print("running the C++ demo", flush=True)
    r = subprocess.run([str(exe), str(out), str(steps), str(W), str(H)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    seed, used = int(r.stdout.split()[1]), int(r.stdout.split()[3])

    g = am.geometry(S, P, 6, 100.0, 0.4)
    sim = am.WaveSim(g, n)
    sim.init_temp(1.0)
    for c in range(8):
        sim.add_node(0, c)
    sim.copy_bounds()
    assert np.array_equal(png_pixels(out / "wave_0.png"), sim.raster(1, W, H))
    for i in range(1, steps + 1):
        sim.timestep()
        assert np.array_equal(png_pixels(out / f"wave_{i}.png"), sim.raster(1, W, H)), i
    assert sim.rng_state == seed and sim.export()["scalars"][2] + 1 == used
    for which, name in ((0, "old.bin"), (1, "new.bin")):
        got = np.frombuffer((out / name).read_bytes(), np.float32)
        assert np.array_equal(got.view(np.uint32), sim.get_field(which).view(np.uint32)), name


def _merge(dst, fresh, e, tiles, timer, n_inner):
    for s in tiles:
        if (timer & ((1 << (e["max_depth"] - (int(e["positions"][s]) >> 27))) - 1)) == 0:
            dst[s * n_inner:(s + 1) * n_inner] = fresh[s * n_inner:(s + 1) * n_inner]


def oracle_bssn_plugin(orc, S, P, g, steps):
    """BSSN_AMR_data (BSSN.h:104-171) on the oracle: the BSSN_test mesh and
    initial data, `steps` timesteps of copy_back -> derivatives of old ->
    copy_bounds -> derivatives of new, 13 seed draws per derivative phase
    (per component differentiate and mixed partials, then the Christoffel
    words, BSSN.h:38-53, 93-102)."""
    n = 128
    t = orc.OracleTree(n)
    children = []
    for c in range(8):
        children.append(t.add_node(0, c))
    for _ in range(1, 6):
        for c in range(8):
            children[c] = t.add_node(children[c], (~c) & 7)
    e = t.export()
    tiles = fields.used_slots(e)
    maxd = e["max_depth"]
    old, new = [], []
    for c, mul in enumerate((1.0, 0.0, 0.0, 1.0, 0.0, 1.0)):
        o, w = orc.init_temp(e["positions"], tiles, n, S, P, g.outer_size, g.outer_dx, mul)
        old.append(o)
        new.append(w)
    n3 = S ** 3
    dk = [np.zeros(n * n3, np.uint32) for _ in range(6)]
    cG = [np.zeros(n * n3, np.uint32) for _ in range(6)]
    cR = [np.zeros(n * n3, np.float32) for _ in range(6)]
    rng = [12894712]

    def draw():
        rng[0] = orc.prng_ints(rng[0], 1)[0]
        return rng[0]

    def derivatives(metric, timer):
        for c in range(6):
            d1, d2 = orc.differentiate(metric[c], e["positions"], tiles, timer, maxd, draw(), S, P, g.fd_order // 2,
                                       g.outer_dx, n)
            off = orc.mixed_partials(metric[c], e["positions"], tiles, timer, maxd, draw(), S, P, g.outer_dx, n)
            tr = orc.deriv_trace(d2, off, metric, e["positions"], tiles, timer, maxd, S, P, n)
            _merge(dk[c], d1, e, tiles, timer, n3)
            _merge(cR[c], tr, e, tiles, timer, n3)
        ch = orc.christoffel(metric, dk, 1, e["positions"], tiles, timer, maxd, draw(), S, P, n)
        for c in range(6):
            _merge(cG[c], ch[c], e, tiles, timer, n3)

    by_depth = {}
    for s in tiles:
        by_depth.setdefault(fields.depth_of(e, s), []).append(s)
    for _ in range(steps):
        timer = 0
        while True:
            for c in range(6):
                orc.copy_new_to_old(old[c], new[c], e["positions"], e["children"], tiles, timer, maxd, S, P)
            derivatives(old, timer)
            for d in sorted(by_depth):
                if d > 0:
                    for c in range(6):
                        orc.copy_boundaries(old[c], new[c], None, e["boundaries"], by_depth[d], timer, maxd, True, 0,
                                            S, P)
            derivatives(new, timer)
            timer = (timer + 1) & ((1 << maxd) - 1)
            if timer == 0:
                break
    return dict(new=new, cG=cG, cR=cR, seed=rng[0], used=len(tiles), e=e)


@pytest.mark.gpu
def test_bssn_demo_matches_oracle(am, orc, torch_cuda, tmp_path):
    S, P = 16, 3
    T = S + 2 * P
    out = tmp_path / "out"
    out.mkdir()
    exe = build(tmp_path, demo="bssn")
    print("running the C++ BSSN demo", flush=True)
    r = subprocess.run([str(exe), str(out)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    last = r.stdout.strip().splitlines()[-1].split()
    seed, used = int(last[1]), int(last[3])
    assert "Node index: 0" in r.stdout  # to_string_debug
    g = am.geometry(S, P, 6, 100.0, 0.4)
    print("oracle run of the same program", flush=True)
    want = oracle_bssn_plugin(orc, S, P, g, 2)
    assert (seed, used) == (want["seed"], want["used"])
    new = np.frombuffer((out / "new.bin").read_bytes(), np.uint32).reshape(6, used * T ** 3)
    cG = np.frombuffer((out / "cG.bin").read_bytes(), np.uint32).reshape(6, used * S ** 3)
    cR = np.frombuffer((out / "cR.bin").read_bytes(), np.uint32).reshape(6, used * S ** 3)
    for c in range(6):
        assert np.array_equal(new[c], want["new"][c][:used * T ** 3].view(np.uint32)), c
        assert np.array_equal(cG[c], want["cG"][c][:used * S ** 3]), c
        assert np.array_equal(cR[c], want["cR"][c][:used * S ** 3].view(np.uint32)), c
    for c, nm in enumerate(("xx", "xy", "xz", "yy", "yz", "zz")):
        px = png_pixels(out / f"test_{nm}.png")
        full = np.zeros(128 * S ** 3, np.uint32)
        full[:used * S ** 3] = want["cG"][c][:used * S ** 3]
        assert np.array_equal(px.reshape(-1)[:used * S ** 3], orc.encode_rgba8_cf3(full)[:used * S ** 3]), nm


@pytest.mark.gpu
def test_stream_pool_forms_match_serial_forms(am, torch_cuda, tmp_path):
    exe = build(tmp_path, demo="pool")
    lines = {}
    for mode in ("serial", "pool"):
        print(f"running the stream-pool demo ({mode})", flush=True)
        r = subprocess.run([str(exe), mode], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        lines[mode] = r.stdout.strip().splitlines()
    assert lines["pool"] == lines["serial"]
    rows = {ln.split()[0]: ln.split()[1:] for ln in lines["pool"]}
    assert len(rows) == 3 * 2 + 9 * 2 + 6 * 2 + 3 + 1
    for name in ("gradient", "hessian.diag", "hessian.off_diag", "metric.xx.new", "k.zz.old", "shift.x.new"):
        assert int(rows[name][1]) > 0, name  # nonzero words: the phases ran
    assert int(rows["seed"][0]) != 12894712 and int(rows["seed"][2]) == 15
