"""The drop-in C++ headers (include/adaptivemesh/*.hpp) from a user program:
tests/cpp/wave_plugin_demo.cpp is the reference's wave_test loop
(kernel.cu:45-64) written against them.  On the CPU it must compile with a
plain C++17 compiler and no warnings (the headers hold no device code); on
the GPU its fields, seed and every saved PNG must equal the same run driven
through the Python bindings, bit for bit."""
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

from tests.test_png import chunks

ROOT = Path(__file__).resolve().parents[1]
DEMO = ROOT / "tests" / "cpp" / "wave_plugin_demo.cpp"


def build(tmp_path, werror=False):
    exe = tmp_path / "wave_plugin_demo"
    lib = ROOT / "adaptivemesh_amd"
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-I", str(ROOT / "include"), str(DEMO), "-L", str(lib),
           "-ladaptivemesh_amd", f"-Wl,-rpath,{lib}", "-o", str(exe)]
    if werror:
        cmd.insert(1, "-Werror")
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=300)
    return exe


def test_headers_compile_as_plain_cpp(am, tmp_path):
    assert build(tmp_path, werror=True).exists()


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
