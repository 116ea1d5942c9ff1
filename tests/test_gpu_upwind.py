"""Upwinded advection with Kreiss-Oliger dissipation on the GPU against the
oracle (tests/test_oracle_dissipation.py pins the oracle):

* the FD kernel's advection + dissipation output equals the oracle's float
  restatement (orc_fd_adv) bit for bit, on padded and interior-first tiles,
  with direct y and z faces (their ghost layers poisoned), every z-block, two
  counter values, with and without the advection term; its words stay the
  oracle's;
* the BSSN driver with set_dissipation: per RK4 stage the words and the
  advection + dissipation arrays are bit-exact with the oracle on the GPU's
  stage input, the fp32 RHS + update match the FP64 oracle given the same
  words and arrays within 5e-5 of the largest |k| per field, and the push
  refreshes every ghost layer (P of them) as exact copies.
"""
import ctypes as C

import numpy as np
import pytest

from tests import fields as F
from tests import meshes
from tests.test_gpu_bssn import (CFL, HESS, OUTER, ORDER, STAGE_FRAC, P, S, T, download, interior_mask, sim_words,
                                 smooth_state, tile_mask)
from tests.test_gpu_layout import to_if
from tests.test_gpu_parity import dev, host, mesh_for, setup_mesh, sync

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout,direct", [(0, 0), (0, 4), (1, 0), (1, 4), (1, 2), (1, 6)])
@pytest.mark.parametrize("recipe,periodic", [("graded", False), ("uniform2", True)])
def test_fd_adv_bit_exact(am, orc, torch_cuda, layout, direct, recipe, periodic):
    torch = torch_cuda
    print(f"fd adv layout={layout} direct={direct} {recipe}", flush=True)
    g = am.geometry(S, P, ORDER, 100.0, 0.4)
    tree, e, n = setup_mesh(am, orc, recipe, periodic)
    rng = np.random.default_rng(layout * 10 + direct + len(recipe))
    nv = 4
    flds = [rng.standard_normal(n * T ** 3).astype(np.float32) for _ in range(nv)]
    shift = [(0.5 * rng.standard_normal(n * T ** 3)).astype(np.float32) for _ in range(3)]
    tiles = [s for s in F.used_slots(e) if F.depth_of(e, s) > 0]
    eq_yz = F.outer_push_layers(e, n, tiles, S, P, 0, 6)
    good = []
    for f in flds:
        filled = f.copy()
        orc.copy_boundaries(f.copy(), filled, None, e["boundaries"], tiles, 0, e["max_depth"], True, 0, S, P,
                            equal_copy=True)
        x = f.copy()
        x[eq_yz] = filled[eq_yz]
        good.append(x)
    src = [x.copy() for x in good]
    if direct:  # the equal-depth layers on the direct axes poisoned: they must come from the neighbours
        poison = F.outer_push_layers(e, n, tiles, S, P, 0, direct)
        assert poison.any()
        for x in src:
            x[poison] = np.nan
    conv = (lambda a: to_if(am, a, S, P, n)) if layout else (lambda a: a)
    dfl = [dev(torch, conv(x)) for x in src]
    dsh = [dev(torch, conv(x)) for x in shift]
    L, st = am.lib(), am.stream_handle()
    W = n * S ** 3
    for substep in (0, 1):
        for sigma in (0.0, 0.37):
            seeds = [int(s_) for s_ in rng.integers(-2 ** 31, 2 ** 31, size=nv)]
            m = mesh_for(tree, -1, substep)
            for zb in (0, S // 2, S // 4):
                d1s = [torch.zeros(W, dtype=torch.int32, device="cuda") for _ in range(nv)]
                advs = [torch.full((W,), float("nan"), dtype=torch.float32, device="cuda") for _ in range(nv)]
                b = am.FdBatch()
                b.n_vars = nv
                b.z_block = zb
                b.tile_layout = layout
                b.direct_faces = direct
                for v in range(nv):
                    b.field[v], b.d1[v], b.adv[v] = dfl[v].data_ptr(), d1s[v].data_ptr(), advs[v].data_ptr()
                    b.seed_d[v] = seeds[v]
                for k in range(3):
                    b.shift[k] = dsh[k].data_ptr()
                b.advect_mask = 0b0101  # fields 0 and 2 advected, 1 and 3 dissipation only
                b.ko_sigma = sigma
                am.check(L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st), "am_fd_batched adv")
                sync(torch)
                for v in range(nv):
                    want = orc.fd_adv(good[v], shift, e["positions"], F.used_slots(e), substep, e["max_depth"], S, P,
                                      g.outer_dx, sigma, v % 2 == 0, n)
                    wmask = np.zeros((n, S ** 3), bool)
                    for s_ in F.used_slots(e):
                        if (substep & ((1 << (e["max_depth"] - F.depth_of(e, s_))) - 1)) == 0:
                            wmask[s_] = True
                    got = host(advs[v], np.float32)
                    assert np.array_equal(got[wmask.reshape(-1)].view(np.uint32),
                                          want[wmask.reshape(-1)].view(np.uint32)), (substep, sigma, zb, v)
                    w1, _ = orc.differentiate(good[v], e["positions"], F.used_slots(e), substep, e["max_depth"],
                                              seeds[v], S, P, ORDER // 2, g.outer_dx, n, with_d2=False)
                    assert np.array_equal(host(d1s[v], np.uint32), w1), (substep, zb, v)
    # the upwind output is refused where it is not built
    b.prefetch_planes = 2
    assert L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st) == -1
    b.prefetch_planes = 1
    for bad_direct in (1, 3, 7):  # x faces are never direct
        b.direct_faces = bad_direct
        assert L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st) == -1
    b.direct_faces = direct
    b.prefetch_planes = 1
    b.shift[2] = None
    assert L.am_fd_batched(C.byref(b), C.byref(m), C.byref(g), st) == -1


@pytest.mark.parametrize("recipe,periodic,depth,direct,advect", [
    ("uniform1", True, 1, 4, 1), ("graded", False, 2, 4, 1), ("graded", False, 3, 0, 1), ("uniform1", True, 1, 0, 0),
    ("uniform1", True, 1, 6, 1), ("graded", False, 3, 6, 1)])
def test_bssn_upwind_stages_match_oracle(am, orc, torch_cuda, recipe, periodic, depth, direct, advect):
    print(f"bssn upwind stages {recipe} depth={depth} direct={direct} advect={advect}", flush=True)
    sigma = 0.3
    ops, n = {"uniform1": meshes.uniform(1), "graded": meshes.graded()}[recipe]
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    sim = am.BssnSim(g, n, am.bssn_options(periodic=periodic, advect_shift=advect))
    sim.set_dissipation(sigma, True)
    assert sim.dissipation == (pytest.approx(sigma), True)
    assert sim.push_layers == P
    sim.set_direct_faces(direct)
    ot = orc.OracleTree(n, periodic)
    for p_, c in ops:
        assert sim.add_node(p_, c) == ot.add_node(p_, c)
    e = ot.export()
    state0 = smooth_state(e, n, seed=depth + 10 * advect)
    for v in range(24):
        sim.set_field(0, v, state0[v])
        sim.set_field(1, v, state0[v])
    tiles = [s for s in F.used_slots(e) if F.depth_of(e, s) == depth]
    outer_dx = float(np.float32(OUTER) / np.float32(S))
    dtd = np.array([float(np.float32(np.float32(outer_dx) * np.float32(CFL)) / np.float32(2 ** d)) for d in range(9)])
    params = np.array((2.0, 0.75, 1e-4, advect, 0), np.float64)
    imask = interior_mask(n, tiles)
    wmask = tile_mask(n, tiles)
    W = n * S ** 3
    old = [x.copy() for x in state0]
    for stage in range(4):
        state_in = old if stage == 0 else [sim.get_field(1, v) for v in range(24)]
        state_fd = F.fd_input(orc, e, n, tiles, state_in, old, stage, direct, S, P, STAGE_FRAC[stage])
        acc_in = [sim.get_field(2, v).astype(np.float64) for v in range(24)]
        seed = sim.rng_state
        sim.stage_depth(depth, stage)
        sim.sync()
        d1, d2d, d2o = sim_words(am, sim, n)
        advg = [download(am, sim.device_ptr(8), W, np.float32, v * W) for v in range(24)]
        s_ = seed
        h = 0
        for v in range(24):
            sd = orc.prng_ints(s_, 1)[0]
            s_ = sd
            hess = v in HESS
            w1, w2 = orc.differentiate(state_fd[v], e["positions"], tiles, 0, e["max_depth"], sd, S, P, ORDER // 2,
                                       outer_dx, n, with_d2=hess)
            assert np.array_equal(d1[v][wmask], w1[wmask]), (stage, v)
            if hess:
                sm = orc.prng_ints(s_, 1)[0]
                s_ = sm
                wo = orc.mixed_partials(state_fd[v], e["positions"], tiles, 0, e["max_depth"], sm, S, P, outer_dx, n)
                assert np.array_equal(d2d[h][wmask], w2[wmask]), (stage, v)
                assert np.array_equal(d2o[h][wmask], wo[wmask]), (stage, v)
                h += 1
            wa = orc.fd_adv(state_fd[v], state_fd[18:21], e["positions"], tiles, 0, e["max_depth"], S, P, outer_dx,
                            sigma, v < 18 or advect, n)
            assert np.array_equal(advg[v][wmask].view(np.uint32), wa[wmask].view(np.uint32)), (stage, v)
        assert sim.rng_state == s_
        acc_o = [np.zeros(n * T ** 3) for _ in range(24)]
        new_o = [np.zeros(n * T ** 3) for _ in range(24)]
        sin32 = [np.ascontiguousarray(x, np.float32) for x in state_in]
        tl = np.array(tiles, np.int32)  # alive through the call
        orc.lib().orc_bssn_stage_from_words_adv(orc.ptrs(sin32), orc.ptrs(old), orc.ptrs(acc_in), orc.ptrs(acc_o),
                                                orc.ptrs(new_o), orc.ptrs(d1), orc.ptrs(d2d), orc.ptrs(d2o),
                                                orc.ptrs(advg), e["positions"].ctypes.data,
                                                tl.ctypes.data, len(tiles), 0,
                                                e["max_depth"], stage, dtd.ctypes.data, params.ctypes.data, S, P)
        new_g = [sim.get_field(1, v) for v in range(24)]
        acc_g = [sim.get_field(2, v) for v in range(24)]
        worst = 0.0
        for v in range(24):
            if stage < 3:
                k_o = (acc_o[v] - acc_in[v])[imask] / (1 if stage == 0 else 2)
                k_g = (acc_g[v].astype(np.float64) - acc_in[v])[imask] / (1 if stage == 0 else 2)
                scale = max(np.max(np.abs(k_o)), 1e-3)
                err = np.max(np.abs(k_g - k_o)) / scale
                worst = max(worst, err)
                assert err < 5e-5, (stage, v, err)
            dn = np.max(np.abs(new_g[v][imask] - new_o[v][imask]))
            assert dn < 1e-6 * max(1.0, np.max(np.abs(new_o[v][imask]))), (stage, v, dn)
        print(f"  stage {stage}: worst relative k error {worst:.2e}", flush=True)
        if depth > 0:
            ref = [x.copy() for x in new_g]
            keep = F.unpushed_layers(e, n, tiles, S, P, sim.push_layers, direct, stage)
            for v in range(24):
                orc.copy_boundaries(old[v].copy(), ref[v], None, e["boundaries"], tiles, 0, e["max_depth"], True, 0,
                                    S, P, STAGE_FRAC[stage], equal_copy=True)
                ref[v][keep] = state_in[v][keep]
                assert np.array_equal(ref[v].view(np.uint32), new_g[v].view(np.uint32)), (stage, v)


def test_dissipation_setting_refusals(am, torch_cuda):
    g = am.geometry(S, P, ORDER, OUTER, CFL)
    sim = am.BssnSim(g, 9, am.bssn_options(periodic=True))
    with pytest.raises(am.AMError):
        sim.set_dissipation(0.2, False)  # dissipation comes with the upwinded advection
    with pytest.raises(am.AMError):
        sim.set_dissipation(-0.1, True)
    g2 = am.geometry(S, 2, 4, OUTER, CFL)
    sim2 = am.BssnSim(g2, 9, am.bssn_options(periodic=True))
    with pytest.raises(am.AMError):
        sim2.set_dissipation(0.2, True)  # P = 2: the stencils reach 3 cells
    sim.set_dissipation(0.0, False)
    assert sim.dissipation == (0.0, False) and sim.push_layers == ORDER // 2
    # the upwind FD reads direct y and z faces as the plain one does: upwinding keeps the mask
    sim.set_direct_faces(6)
    sim.set_dissipation(0.2, True)
    assert sim.direct_faces == 6
    sim.set_direct_faces(2)
    assert sim.direct_faces == 2
    with pytest.raises(am.AMError):
        sim.set_direct_faces(3)
