"""Regridding of the SimpleWave plugin (SimpleWave.h:91-248) on the GPU
against the oracle, bit for bit.

* am_amr_criterion: octant maxima from random derivative words on a graded
  mesh with an unused slot inside the used range, at every sub-step of a
  coarse step (inactive depths and the unused slot give NaN);
* the native AMR<wave_AMR_data> with modify_nodes on, three timesteps from a
  two-bump initial field: the tree (positions, parents, children, lifetimes,
  counters), both fields, the seed state and the two pressures equal the
  oracle's after every timestep, on a run that adds tiles, removes them and
  reuses freed slots.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import regrid as rg
from tests import fields, meshes

pytestmark = pytest.mark.gpu


def dev(torch, a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(a.copy()).cuda()


def same_bits(a, b):
    a, b = np.asarray(a, np.float32), np.asarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return np.array_equal(na, nb) and np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


@pytest.mark.parametrize("S,P", [(16, 3), (28, 2)])
def test_amr_criterion_bit_exact(am, orc, torch_cuda, S, P):
    torch = torch_cuda
    print(f"amr_criterion S={S} P={P}", flush=True)
    g = am.geometry(S, P, 2 * P, 100.0, 0.4)
    ops, n = meshes.graded()
    tree, otree = am.Tree(n), orc.OracleTree(n)
    meshes.apply(tree, ops)
    meshes.apply(otree, ops)
    assert tree.remove_node(12) == 0 and otree.remove_node(12) == 0
    e = otree.export()
    assert int(e["positions"][12]) == 0xFFFFFFFF and e["final_used"] == n - 1
    rng = np.random.default_rng(S)
    vec = lambda: (rng.standard_normal((n * S ** 3, 3)) * 10.0 ** rng.uniform(-3, 3, (n * S ** 3, 1)))
    d1w, d2w = orc.cf3_pack_array(vec()), orc.cf3_pack_array(vec())
    d1, d2 = dev(torch, d1w), dev(torch, d2w)
    out = torch.full((n * 8,), -7.0, dtype=torch.float32, device="cuda")
    for substep in range(1 << e["max_depth"]):
        m = tree.mesh(-1)
        m.substep = substep
        am.check(am.lib().am_amr_criterion(d1.data_ptr(), d2.data_ptr(), C.byref(m), C.byref(g), out.data_ptr(),
                                           am.stream_handle()), "am_amr_criterion")
        torch.cuda.synchronize()
        want = rg.octant_criterion(orc, d1w, d2w, e["positions"], n, substep, e["max_depth"], S, g.outer_dx)
        got = out.cpu().numpy()
        assert same_bits(got, want), substep
        assert np.isnan(got[12 * 8:13 * 8]).all()
        assert np.count_nonzero(~np.isnan(got)) == 8 * sum(
            1 for s in fields.used_slots(e)
            if (substep & ((1 << (e["max_depth"] - fields.depth_of(e, s))) - 1)) == 0)


def test_amr_criterion_rejects_bad_arguments(am, torch_cuda):
    g = am.geometry(16, 3, 6, 100.0, 0.4)
    tree = am.Tree(2)
    m = tree.mesh(-1)
    assert am.lib().am_amr_criterion(None, None, C.byref(m), C.byref(g), None, am.stream_handle()) != 0


def _merge(words, fresh, e, tiles, timer, S):
    """The kernels write the words of the tiles active at `timer` only; the
    rest keep what an earlier sub-step left."""
    for s in tiles:
        if (timer & ((1 << (e["max_depth"] - fields.depth_of(e, s))) - 1)) == 0:
            words[s * S ** 3:(s + 1) * S ** 3] = fresh[s * S ** 3:(s + 1) * S ** 3]


def oracle_regrid_timestep(orc, tree, n, old, new, w1, w2, rng, S, P, g, log):
    """wave_AMR_data::timestep (SimpleWave.h:268-278) with modify_nodes on,
    driven on the oracle: the octree, its timer and lifetimes live in the
    oracle tree; the derivative words persist between sub-steps."""
    while True:
        e = tree.export()
        tiles = fields.used_slots(e)
        orc.copy_new_to_old(old, new, e["positions"], e["children"], tiles, e["timer"], e["max_depth"], S, P)
        crit = rg.octant_criterion(orc, w1, w2, e["positions"], e["final_used"] + 1, e["timer"], e["max_depth"], S,
                                   g.outer_dx)
        avg1, avg2, removed, added = rg.modify_nodes(orc, tree, crit, old, new, S, P)
        log.append((removed, added))
        e = tree.export()
        tiles, timer, maxd = fields.used_slots(e), e["timer"], e["max_depth"]
        by_depth = {}
        for s in tiles:
            by_depth.setdefault(fields.depth_of(e, s), []).append(s)
        for src, corrector in ((old, False), (new, True)):
            if corrector:
                for d in sorted(by_depth):
                    if d > 0:
                        orc.copy_boundaries(old, new, None, e["boundaries"], by_depth[d], timer, maxd, True, 0, S, P)
            d1, d2 = orc.differentiate(src, e["positions"], tiles, timer, maxd, orc.prng_ints(rng[0], 1)[0], S, P,
                                       g.fd_order // 2, g.outer_dx, n)
            rng[0] = orc.prng_ints(rng[0], 1)[0]
            _merge(w1, d1, e, tiles, timer, S)
            _merge(w2, d2, e, tiles, timer, S)
            orc.wave_stage(old, d1, d2, new, e["positions"], tiles, timer, maxd, corrector, S, P, g.outer_dt)
        tree.increment_timer()
        if tree.export()["timer"] == 0:
            return avg1, avg2


def test_simplewave_regrid_bit_exact(am, orc, torch_cuda):
    S, P = 16, 3
    T = S + 2 * P
    g = am.geometry(S, P, 6, 100.0, 0.4)
    ops, _ = meshes.graded()
    n = 40
    sim = am.WaveSim(g, n)
    otree = orc.OracleTree(n)
    root = orc.OracleTree(1).export()
    blob = fields.sample(root["positions"], 1, S, P, 100.0,
                         lambda x, y, z: 10 / ((x - 25) ** 2 + (y - .5) ** 2 + z * z + 10)
                         - 10 / ((x + 25) ** 2 + (y + .5) ** 2 + z * z + 10))
    old = np.zeros(n * T ** 3, np.float32)
    old[:T ** 3] = blob
    sim.set_field(0, old)
    sim.set_field(1, old)
    new = old.copy()
    for p, c in ops:
        s = sim.add_node(p, c)
        assert otree.add_node(p, c) == s
        orc.copy_to_child(old, c, s, p, 0, S, P)
        orc.copy_to_child(new, c, s, p, 0, S, P)
    sim.copy_bounds()
    e = otree.export()
    for d in range(1, e["max_depth"] + 1):
        tiles = [s for s in fields.used_slots(e) if fields.depth_of(e, s) == d]
        orc.copy_boundaries(old, new, None, e["boundaries"], tiles, 0, e["max_depth"], True, 0, S, P)
    w1 = np.zeros(n * S ** 3, np.uint32)
    w2 = np.zeros(n * S ** 3, np.uint32)
    rng = [12894712]
    log = []
    for step in range(3):
        print(f"regrid timestep {step}", flush=True)
        sim.timestep()
        avg = oracle_regrid_timestep(orc, otree, n, old, new, w1, w2, rng, S, P, g, log)
        got, want = sim.export(), otree.export()
        for k in ("positions", "parent", "children", "lifetime", "scalars"):
            assert np.array_equal(got[k], want[k]), (step, k)
        assert sim.rng_state == rng[0]
        assert sim.pressure() == (float(avg[0]), float(avg[1])), step
        assert np.array_equal(sim.get_field(1).view(np.uint32), new.view(np.uint32)), step
        assert np.array_equal(sim.get_field(0).view(np.uint32), old.view(np.uint32)), step
    removed = {s for r, _ in log for s in r}
    added = [s for _, a in log for s in a]
    assert removed and added and removed & set(added), log  # removals, additions and a reused slot
