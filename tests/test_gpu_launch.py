"""Launch machinery of the RK4 driver: none of it may change a single bit.

* A recorded coarse step (hipGraph) replayed three times gives the same
  fields, accumulator, seeds and zone-cycle count as eager launches, checked
  after every replay with eager work (downloads) in between; with tricubic
  ghosts (three time levels) on a graded non-periodic mesh and with the
  scalar-wave System.
* Persistent launch shapes (capped grids, workgroups looping over a level's
  work items) and other FD z-blocks give bit-identical results.
* autotune() (whole coarse steps from a snapshot) leaves the state as it
  found it: the next timestep equals an untuned simulation's, counter,
  seeds and tile lifetimes included.
* The overlap of coarser depths on a second stream (on by default) gives
  the one-stream results bit for bit, eager and replayed, on uniform levels
  with their ancestors and on a graded mesh whose finest depth reads
  coarser faces (its first coarse-face fill waits for the coarser depths),
  with tricubic three-level fills.
"""
import numpy as np
import pytest

from adaptivemesh_amd import workloads as WL
from tests import fields as F
from tests import meshes

pytestmark = pytest.mark.gpu

S, P, ORDER = 16, 3, 4
T = S + 2 * P


def smooth(e, n, outer, seed):
    rng = np.random.default_rng(seed)
    out = []
    for v in range(24):
        k = rng.integers(1, 3, size=3)
        ph = rng.uniform(0, 2 * np.pi, 3)
        a = 0.03 * rng.uniform(0.5, 1.0)
        out.append(F.sample(e["positions"], n, S, P, outer,
                            lambda x, y, z, v=v, k=k, ph=ph, a=a: WL.FLAT[v]
                            + a * np.sin(2 * np.pi * k[0] * x / outer + ph[0]) * np.cos(2 * np.pi * k[1] * y / outer
                                                                                       + ph[1])
                            + 0.5 * a * np.sin(2 * np.pi * k[2] * z / outer + ph[2])))
    return out


def bssn_pair(am, orc, interp, periodic, recipe="graded", outer=16.0):
    ops, n = {"graded": meshes.graded(), "uniform1": meshes.uniform(1), "uniform2": meshes.uniform(2)}[recipe]
    g = am.geometry(S, P, ORDER, outer, 0.25)
    sims = [am.BssnSim(g, n, am.bssn_options(periodic=periodic, interp=interp)) for _ in range(2)]
    ot = orc.OracleTree(n, periodic)
    for p_, c in ops:
        ot.add_node(p_, c)
        for s in sims:
            s.add_node(p_, c)
    e = ot.export()
    st = smooth(e, n, outer, 3)
    for s in sims:
        for v in range(24):
            s.set_field(0, v, st[v])
            s.set_field(1, v, st[v])
        s.fill_ghosts()
    return sims, n


def same_state(a, b, nv, which=(0, 1, 2)):
    for w in which:
        for v in range(nv):
            x, y = a.get_field(w, v), b.get_field(w, v)
            if not np.array_equal(x.view(np.uint32), y.view(np.uint32)):
                return (w, v, int(np.sum(x.view(np.uint32) != y.view(np.uint32))))
    return None


@pytest.mark.parametrize("interp,periodic", [(0, True), (2, False)])
def test_graph_replay_equals_eager_bssn(am, orc, torch_cuda, interp, periodic):
    print(f"graph replay vs eager, BSSN graded, interp={interp}", flush=True)
    (eager, graph), n = bssn_pair(am, orc, interp, periodic)
    graph.use_graph(True)
    for step in range(3):
        eager.timestep()
        graph.timestep()
        eager.sync()
        graph.sync()
        diff = same_state(eager, graph, 24, (0, 1, 2, 3) if interp == 2 else (0, 1, 2))
        assert diff is None, (step, diff)
        assert eager.rng_state == graph.rng_state, step
        assert eager.zone_cycles == graph.zone_cycles, step
        assert eager.counter() == graph.counter() == (0, 3)


def test_graph_replay_equals_eager_swave(am, orc, torch_cuda):
    ops, n = meshes.uniform(1)
    g = am.geometry(S, P, ORDER, 1.0, 0.25)
    sims = [am.SwaveSim(g, n, am.swave_options(periodic=True)) for _ in range(2)]
    for s in sims:
        meshes.apply(s, ops)
    ot = orc.OracleTree(n, True)
    meshes.apply(ot, ops)
    e = ot.export()
    u0 = F.sample(e["positions"], n, S, P, 1.0, lambda x, y, z: np.exp(-(x * x + y * y + z * z) / 0.01))
    for s in sims:
        for w in (0, 1):
            s.set_field(w, 0, u0)
            s.set_field(w, 1, np.zeros_like(u0))
        s.fill_ghosts()
    sims[1].use_graph(True)
    for step in range(3):
        for s in sims:
            s.timestep()
            s.sync()
        assert same_state(sims[0], sims[1], 2) is None, step
        assert sims[0].rng_state == sims[1].rng_state


@pytest.mark.parametrize("launch", [(8, 0, 0, 1), (4, 64, 32, 1), (16, 1, 1, 1), (0, 512, 7, 1), (16, 0, 0, 1),
                                    (4, 0, 0, 1)])
def test_launch_shapes_bit_identical(am, orc, torch_cuda, launch):
    print(f"launch shape {launch}", flush=True)
    (ref, alt), n = bssn_pair(am, orc, 0, True)
    alt.set_launch(*launch[:3])
    alt.set_fd_prefetch(launch[3])
    assert alt.get_launch() == launch
    ref.timestep()
    alt.timestep()
    ref.sync()
    alt.sync()
    assert same_state(ref, alt, 24) is None


def test_autotune_restores_state(am, orc, torch_cuda):
    (ref, tuned), n = bssn_pair(am, orc, 2, False)
    ref.timestep()
    tuned.timestep()
    rep = tuned.autotune(reps=2)
    print(f"autotune: {rep}", flush=True)
    assert rep["z_block"] in (4, 8, 16) and rep["fd_prefetch"] == 1 and len(rep["trials"]) == 13
    # the mask it keeps need not be the default: the steps below still equal the reference run's
    assert rep["direct_faces"] in (0, 4, 6) and tuned.direct_faces == rep["direct_faces"]
    assert [t[1] for t in rep["trials"] if t[0] == "direct_faces"] == [4, 6, 0]
    assert [t[1] for t in rep["trials"] if t[0] == "overlap_depths"] == [1, 0]
    assert tuned.overlap_depths == rep["overlap_depths"]
    assert tuned.get_launch() == (rep["z_block"], rep["fd_grid_limit"], rep["rhs_grid_limit"], rep["fd_prefetch"])
    tuned.sync()
    assert same_state(ref, tuned, 24, (0, 1, 2, 3)) is None
    assert ref.rng_state == tuned.rng_state and ref.counter() == tuned.counter()
    assert np.array_equal(ref.export()["lifetime"], tuned.export()["lifetime"])
    ref.timestep()
    tuned.timestep()
    ref.sync()
    tuned.sync()
    assert same_state(ref, tuned, 24, (0, 1, 2, 3)) is None


@pytest.mark.parametrize("recipe,periodic,interp", [("uniform1", True, 0), ("uniform2", True, 0),
                                                    ("graded", False, 2)])
def test_overlap_depths_bit_identical(am, orc, torch_cuda, recipe, periodic, interp):
    print(f"overlap of depths: {recipe} periodic={periodic} interp={interp}", flush=True)
    (one, two), n = bssn_pair(am, orc, interp, periodic, recipe)
    assert two.overlap_depths  # the default
    one.set_overlap_depths(False)
    assert not one.overlap_depths
    for graph in (False, True):
        one.use_graph(graph)
        two.use_graph(graph)
        for step in range(2):
            one.timestep()
            two.timestep()
            one.sync()
            two.sync()
            assert same_state(one, two, 24) is None, (graph, step)
            assert one.rng_state == two.rng_state and one.counter() == two.counter()
    with pytest.raises(am.AMError):  # the switch is 0 or 1
        am.check(one._f("set_overlap_depths")(one.h, 2), "set_overlap_depths")
