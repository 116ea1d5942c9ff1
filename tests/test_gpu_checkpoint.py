"""Checkpoint and resume of the drivers on the GPU, bit for bit.

For each driver a run is saved part-way, continued, and compared with a
fresh driver that loads the file and runs the same number of steps: the
octree (every exported array and counter), the seed and every field must be
identical.  The SimpleWave run regrids (its saved derivative words decide
the refinement criterion of tiles not differentiated since), the BSSN run
uses tricubic ghosts (its `older` level is saved) and is saved in the middle
of a coarse step.  A load that does not fit the driver leaves it unchanged.
"""
import numpy as np
import pytest

from adaptivemesh_amd import checkpoint as ck
from adaptivemesh_amd import workloads as WL
from tests import fields, meshes

pytestmark = pytest.mark.gpu

TREE_KEYS = ("parent", "children", "positions", "boundaries", "lifetime", "scalars")


def same_tree(a, b):
    return all(np.array_equal(a[k], b[k]) for k in TREE_KEYS)


def wave_sim(am, orc, n=40):
    S, P = 16, 3
    T = S + 2 * P
    g = am.geometry(S, P, 6, 100.0, 0.4)
    sim = am.WaveSim(g, n)
    root = orc.OracleTree(1).export()
    blob = fields.sample(root["positions"], 1, S, P, 100.0,
                         lambda x, y, z: 10 / ((x - 25) ** 2 + (y - .5) ** 2 + z * z + 10)
                         - 10 / ((x + 25) ** 2 + (y + .5) ** 2 + z * z + 10))
    init = np.zeros(n * T ** 3, np.float32)
    init[:T ** 3] = blob
    sim.set_field(0, init)
    sim.set_field(1, init)
    ops, _ = meshes.graded()
    for p, c in ops:
        sim.add_node(p, c)
    sim.copy_bounds()
    return sim, g


def test_wave_resume_bit_exact(am, orc, torch_cuda, tmp_path):
    print("wave checkpoint", flush=True)
    sim, g = wave_sim(am, orc)
    sim.timestep()
    path = tmp_path / "wave.ckpt"
    sim.save(path)
    saved = (sim.export(), sim.get_field(0), sim.get_field(1), sim.rng_state)
    h, secs = ck.read(path)
    used = saved[0]["scalars"][2] + 1
    assert h["timer"] == saved[0]["scalars"][3] and h["seed"] == saved[3] and h["flags"] == 1
    assert np.array_equal(secs["positions"], saved[0]["positions"])
    assert np.array_equal(secs["new"].ravel(), saved[2][:used * g.T ** 3])
    assert am.inspect_checkpoint(path).n_sections == 9
    for _ in range(2):
        sim.timestep()
    want = (sim.export(), sim.get_field(0), sim.get_field(1), sim.rng_state, sim.pressure())

    fresh, _ = wave_sim(am, orc)
    fresh.load(path)
    got = fresh.export()
    assert same_tree(got, saved[0]) and fresh.rng_state == saved[3]
    assert np.array_equal(fresh.get_field(1)[:used * g.T ** 3].view(np.uint32),
                          saved[2][:used * g.T ** 3].view(np.uint32))
    for _ in range(2):
        fresh.timestep()
    assert same_tree(fresh.export(), want[0])
    assert np.array_equal(fresh.get_field(0).view(np.uint32), want[1].view(np.uint32))
    assert np.array_equal(fresh.get_field(1).view(np.uint32), want[2].view(np.uint32))
    assert fresh.rng_state == want[3] and fresh.pressure() == want[4]
    assert not same_tree(saved[0], want[0])  # the mesh did change after the save


def bssn_sim(am, n_ops, graph=False):
    S, P, outer = 16, 3, 32.0
    g = am.geometry(S, P, 4, outer, 0.25)
    ops, n = n_ops
    sim = am.BssnSim(g, n, am.bssn_options(interp=am.AM_INTERP_TRICUBIC))
    meshes.apply(sim, ops)
    e = sim.export()
    f = WL.gauge_wave_fields(e["positions"], S, P, outer, amplitude=0.05)
    for v in range(24):
        sim.set_field(0, v, f[v])
        sim.set_field(1, v, f[v])
    sim.fill_ghosts()
    sim.use_graph(graph)
    return sim, g, n


def bssn_state(sim, n, T):
    used = sim.export()["scalars"][2] + 1
    return [sim.get_field(w, v)[:used * T ** 3].copy() for w in (0, 1) for v in range(24)]


@pytest.mark.parametrize("graph", [False, True])
def test_bssn_resume_mid_coarse_step_bit_exact(am, torch_cuda, tmp_path, graph):
    print(f"bssn checkpoint graph={graph}", flush=True)
    mesh = meshes.nested(2)
    sim, g, n = bssn_sim(am, mesh)
    sim.timestep()
    sim.substep()
    sim.sync()
    timer = sim.export()["scalars"][3]
    assert timer != 0
    path = tmp_path / "bssn.ckpt"
    sim.save(path)
    h = am.inspect_checkpoint(path)
    assert (h.kind, h.n_vars, h.n_sections, h.timer) == (2, 24, 8, timer)
    for _ in range(3):
        sim.substep()
    sim.timestep()
    sim.sync()
    want_tree, want_rng, want = sim.export(), sim.rng_state, bssn_state(sim, n, g.T)

    fresh, _, _ = bssn_sim(am, mesh, graph=graph)
    fresh.load(path)
    assert fresh.export()["scalars"][3] == timer
    for _ in range(3):
        fresh.substep()
    fresh.timestep()
    fresh.sync()
    assert same_tree(fresh.export(), want_tree) and fresh.rng_state == want_rng
    for a, b in zip(bssn_state(fresh, n, g.T), want):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_swave_resume_bit_exact(am, torch_cuda, tmp_path):
    print("swave checkpoint", flush=True)
    S, P = 16, 3
    g = am.geometry(S, P, 4, 16.0, 0.25)
    ops, n = meshes.graded()

    def make():
        sim = am.SwaveSim(g, n, am.swave_options(periodic=False))
        meshes.apply(sim, ops)
        e = sim.export()
        u = fields.sample(e["positions"], n, S, P, 16.0, lambda x, y, z: np.exp(-(x * x + y * y + z * z) / 8.0))
        sim.set_field(0, 0, u)
        sim.set_field(1, 0, u)
        sim.fill_ghosts()
        return sim

    sim = make()
    sim.timestep()
    path = tmp_path / "swave.ckpt"
    sim.save(path)
    sim.timestep()
    sim.sync()
    want = [sim.get_field(w, v) for w in (0, 1) for v in range(2)]
    fresh = make()
    fresh.load(path)
    fresh.timestep()
    fresh.sync()
    for a, b in zip([fresh.get_field(w, v) for w in (0, 1) for v in range(2)], want):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert fresh.rng_state == sim.rng_state


def test_load_that_does_not_fit_changes_nothing(am, orc, torch_cuda, tmp_path):
    sim, g = wave_sim(am, orc, n=40)
    sim.timestep()
    path = tmp_path / "w.ckpt"
    sim.save(path)
    other, _ = wave_sim(am, orc, n=41)  # other slot count
    before = (other.export(), other.get_field(1), other.rng_state)
    with pytest.raises(am.AMError) as ex:
        other.load(path)
    assert ex.value.code == -11
    off, _ = wave_sim(am, orc, n=40)
    off.set_regrid(False)  # other options
    with pytest.raises(am.AMError):
        off.load(path)
    damaged = tmp_path / "d.ckpt"
    damaged.write_bytes(path.read_bytes()[:-4])
    with pytest.raises(am.AMError):
        other.load(damaged)
    assert same_tree(other.export(), before[0]) and other.rng_state == before[2]
    assert np.array_equal(other.get_field(1).view(np.uint32), before[1].view(np.uint32))
    other.timestep()  # still runs
